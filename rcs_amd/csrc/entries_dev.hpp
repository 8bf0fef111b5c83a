// Device-side MoM entry functions shared by the block-entry and ACA kernels:
// RWG EFIE (src/EFIEMatrix.cpp:247-783) and Laplace (src/ElectrostaticsMatrix.cpp:140-340).
#pragma once
#include "entries.hpp"

namespace rcs {

template <typename R>
struct V3 {
  R x, y, z;
  __device__ V3() {}
  __device__ V3(R a, R b, R c) : x(a), y(b), z(c) {}
};
template <typename R> __device__ inline V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> __device__ inline V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> __device__ inline V3<R> operator*(V3<R> a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <typename R> __device__ inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> __device__ inline V3<R> cross(V3<R> a, V3<R> b) {
  return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename R> __device__ inline R nrm(V3<R> a) { return sqrt(dot(a, a)); }

template <typename R> __device__ inline V3<R> node(const DevGeometry& g, int i) { return V3<R>((R)g.nx[i], (R)g.ny[i], (R)g.nz[i]); }
template <typename R> __device__ inline V3<R> centroid(const DevGeometry& g, int t) { return V3<R>((R)g.cx[t], (R)g.cy[t], (R)g.cz[t]); }

__device__ inline void dsincos(double x, double* s, double* c) { sincos(x, s, c); }
__device__ inline void dsincos(float x, float* s, float* c) { sincosf(x, s, c); }

// Triangle rule, E:284-286 (constants as truncated in the reference).
static __constant__ double kQ7A[7] = {0.3333333, 0.0597158, 0.470142, 0.470142, 0.7974269, 0.1012865, 0.1012865};
static __constant__ double kQ7B[7] = {0.3333333, 0.470142, 0.0597158, 0.470142, 0.1012865, 0.7974269, 0.1012865};
static __constant__ double kQ7W[7] = {0.225, 0.13239415278851, 0.13239415278851, 0.13239415278851,
                               0.12593918054483, 0.12593918054483, 0.12593918054483};

// Closed-form integral of 1/R over the triangle (a, b, c) seen from tc (E:478-577).
template <typename R>
__device__ R analytic_laplace(V3<R> tc, V3<R> a, V3<R> b, V3<R> c) {
  V3<R> n = cross(b - a, c - a);
  n = n * (R(-1) / nrm(n));
  R d = dot(tc - a, n);
  V3<R> cp = tc - n * d;
  d = fabs(d);
  R s = 0;
  V3<R> P[4] = {a, b, c, a};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    V3<R> p1 = P[i], p2 = P[i + 1];
    R len = nrm(p1 - p2);
    R Rp = nrm(p1 - tc), Rm = nrm(p2 - tc);
    V3<R> Pp = p1 - cp, Pm = p2 - cp;
    V3<R> ul = (p1 - p2) * (R(1) / len);
    R Lp = dot(Pp, ul), Lm = dot(Pm, ul);
    V3<R> uu = cross(ul, n);
    R P0 = fabs(dot(Pp, uu));
    if (P0 < R(1.0e-10) * len) continue;
    R R0 = sqrt(P0 * P0 + d * d);
    V3<R> up = Pp - ul * Lp;
    up = up * (R(1) / nrm(up));
    R T1 = P0 * log((Rp + Lp) / (Rm + Lm));
    if (P0 < R(1e-6) * len && R0 < R(1e-4) * len) T1 = 0;
    R T2 = atan(P0 * Lp / (R0 * R0 + d * Rp));
    R T3 = atan(P0 * Lm / (R0 * R0 + d * Rm));
    s += dot(up, uu) * (T1 - d * (T2 - T3));
  }
  return s;
}

// Closed-form integral of (r' - p)/R over the triangle, p = projection of tc on its plane (E:579-668).
template <typename R>
__device__ V3<R> analytic_rho_laplace(V3<R> tc, V3<R> a, V3<R> b, V3<R> c) {
  V3<R> n = cross(b - a, c - a);
  n = n * (R(-1) / nrm(n));
  R d = dot(tc - a, n);
  V3<R> cp = tc - n * d;
  d = fabs(d);
  V3<R> s(0, 0, 0);
  V3<R> P[4] = {a, b, c, a};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    V3<R> p1 = P[i], p2 = P[i + 1];
    R Rp = nrm(p1 - tc), Rm = nrm(p2 - tc);
    V3<R> Pp = p1 - cp, Pm = p2 - cp;
    R len = nrm(p1 - p2);
    V3<R> ul = (p1 - p2) * (R(1) / len);
    R Lp = dot(Pp, ul), Lm = dot(Pm, ul);
    V3<R> uu = cross(ul, n);
    R P0 = fabs(dot(Pp, uu));
    R sq = P0 * P0 + d * d;
    R T1 = ((Rm + Lm) >= R(1.0e-10)) ? sq * log((Rp + Lp) / (Rm + Lm)) : R(0);
    R T2 = Lp * Rp - Lm * Rm;
    s = s + uu * (R(0.5) * (T1 + T2));
  }
  return s;
}

// Source triangle integrals seen from test centroid tc with source free vertex fv:
// scalar part (E:670-685) and vector part (E:386-446).
template <typename R>
__device__ void source_integrals(const DevGeometry& g, V3<R> tc, int s, int fv, cpx<R>& S, cpx<R> V[3]) {
  const R kap = (R)g.kappa;
  const R c4pi = R(-1.0 / (4.0 * M_PI));
  const int i1 = g.v1[s], i2 = g.v2[s], i3 = g.v3[s];
  V3<R> a = node<R>(g, i1), b = node<R>(g, i2), c = node<R>(g, i3);
  V3<R> sc = centroid<R>(g, s);
  const R A = (R)g.area[s], LE = (R)g.le[s];
  const R dist = nrm(tc - sc);
  V3<R> r3 = node<R>(g, fv);
  const bool nearz = dist < R(4) * LE, midz = dist < R(10) * LE;

  // vector non-singular part, 7-point rule over (v2, v3, v1)
  cpx<R> ns[3] = {cpx<R>(0, 0), cpx<R>(0, 0), cpx<R>(0, 0)};
#pragma unroll 1
  for (int q = 0; q < 7; ++q) {
    V3<R> pt = b + (c - b) * (R)kQ7A[q] + (a - b) * (R)kQ7B[q];
    R Rr = nrm(tc - pt);
    V3<R> rho = pt - r3;
    cpx<R> f;
    if (Rr < R(1e-12)) {
      f = cpx<R>(0, c4pi * (-kap));
    } else {
      R sn, cs;
      dsincos(kap * Rr, &sn, &cs);
      R inv = c4pi / Rr;
      f = cpx<R>((cs - R(1)) * inv, -sn * inv);
    }
    R w = (R)kQ7W[q];
    ns[0] = ns[0] + f * (w * rho.x);
    ns[1] = ns[1] + f * (w * rho.y);
    ns[2] = ns[2] + f * (w * rho.z);
  }
  V3<R> sg;
  cpx<R> sc_int;
  if (nearz) {
    // free vertex first, then the other two in patch order (E:478-494)
    V3<R> pa = r3, pb, pc;
    if (i1 == fv) { pb = b; pc = c; }
    else if (i2 == fv) { pb = a; pc = c; }
    else { pb = a; pc = b; }
    R lap = analytic_laplace(tc, pa, pb, pc);
    V3<R> rl = analytic_rho_laplace(tc, pa, pb, pc);
    // projection on the plane through the origin with the source normal (E:457-476)
    V3<R> nn = cross(b - a, c - a);
    nn = nn * (R(-1) / nrm(nn));
    V3<R> proj = tc - nn * dot(tc, nn);
    sg = (rl + (proj - r3) * lap) * c4pi;
    cpx<R> sum(0, 0);
#pragma unroll 1
    for (int q = 0; q < 7; ++q) {
      V3<R> pt = a + (b - a) * (R)kQ7A[q] + (c - a) * (R)kQ7B[q];
      R Rr = nrm(pt - tc);
      R w = (R)kQ7W[q];
      if (Rr < R(1e-12)) {
        sum = sum + cpx<R>(0, w * kap * R(1.0 / (4.0 * M_PI)));
      } else {
        R sn, cs;
        dsincos(kap * Rr, &sn, &cs);
        R inv = c4pi / Rr;
        sum = sum + cpx<R>((cs - R(1)) * inv * w, -sn * inv * w);
      }
    }
    sc_int = cpx<R>(c4pi * lap, 0) + sum * A;
  } else if (midz) {
    V3<R> s3(0, 0, 0);
#pragma unroll 1
    for (int q = 0; q < 7; ++q) {
      V3<R> pt = b + (c - b) * (R)kQ7A[q] + (a - b) * (R)kQ7B[q];
      R Rr = nrm(tc - pt);
      s3 = s3 + (pt - r3) * ((R)kQ7W[q] * c4pi / Rr);
    }
    sg = s3 * A;
    cpx<R> sum(0, 0);
#pragma unroll 1
    for (int q = 0; q < 7; ++q) {
      V3<R> pt = a + (b - a) * (R)kQ7A[q] + (c - a) * (R)kQ7B[q];
      R Rr = nrm(tc - pt);
      R sn, cs;
      dsincos(kap * Rr, &sn, &cs);
      R inv = c4pi / Rr * (R)kQ7W[q];
      sum = sum + cpx<R>(cs * inv, -sn * inv);
    }
    sc_int = sum * A;
  } else {
    R Rr = nrm(tc - sc);
    sg = (sc - r3) * (c4pi / Rr * A);
    R sn, cs;
    dsincos(kap * Rr, &sn, &cs);
    R inv = c4pi / Rr * A;
    sc_int = cpx<R>(cs * inv, -sn * inv);
  }
  S = sc_int;
  V[0] = ns[0] * A + cpx<R>(sg.x, 0);
  V[1] = ns[1] * A + cpx<R>(sg.y, 0);
  V[2] = ns[2] * A + cpx<R>(sg.z, 0);
}

// EFIEMatrix::getMatrixEntry, E:698-783.
template <typename R>
__device__ cpx<R> efie_entry(const DevGeometry& g, int i, int j) {
  const int ttri[2] = {g.tp[i], g.tm[i]};
  const int tfree[2] = {g.fp[i], g.fm[i]};
  const int stri[2] = {g.tp[j], g.tm[j]};
  const int sfree[2] = {g.fp[j], g.fm[j]};
  const R Lt = (R)g.len[i], Ls = (R)g.len[j];
  const R w = (R)g.omega;
  cpx<R> e(0, 0);
#pragma unroll 1
  for (int a = 0; a < 2; ++a) {
    const int t = ttri[a];
    V3<R> tc = centroid<R>(g, t);
    const R At = (R)g.area[t];
    V3<R> rho = tc - node<R>(g, tfree[a]);
#pragma unroll 1
    for (int b = 0; b < 2; ++b) {
      const int s = stri[b];
      const R As = (R)g.area[s];
      cpx<R> S, V[3];
      source_integrals<R>(g, tc, s, sfree[b], S, V);
      const R sgn = (a == b) ? R(-1) : R(1);
      // phi constant sgn*Lt*Ls/(i w eps At As) = -i sgn Lt Ls/(w eps At As)
      R phic = sgn * Lt * Ls / (w * (R)g.eps * At * As);
      cpx<R> phiI = S * At;
      e = e + cpx<R>(phiI.im * phic, -phiI.re * phic);
      // A constant sgn*i w mu Lt Ls/(4 At As)
      R ac = sgn * w * (R)g.mu * Lt * Ls / (R(4) * At * As);
      cpx<R> aI = (V[0] * rho.x + V[1] * rho.y + V[2] * rho.z) * At;
      e = e + cpx<R>(-aI.im * ac, aI.re * ac);
    }
  }
  return e;
}

// ElectrostaticsMatrix::getMatrixEntry, src/ElectrostaticsMatrix.cpp:293-340.
template <typename R>
__device__ cpx<R> laplace_entry(const DevGeometry& g, int i, int j) {
  V3<R> tc = centroid<R>(g, i);
  V3<R> a = node<R>(g, g.v1[j]), b = node<R>(g, g.v2[j]), c = node<R>(g, g.v3[j]);
  V3<R> sc = centroid<R>(g, j);
  const R A = (R)g.area[j], LE = (R)g.le[j];
  R d = nrm(tc - sc);
  R v;
  if (d < R(4) * LE) {
    v = analytic_laplace(tc, a, b, c);
  } else if (d < R(10) * LE) {
    v = 0;
    for (int q = 0; q < 7; ++q) {
      V3<R> pt = a + (b - a) * (R)kQ7A[q] + (c - a) * (R)kQ7B[q];
      v += A * (R)kQ7W[q] * (R(1) / nrm(tc - pt));
    }
  } else {
    v = A * (R(1) / d);
  }
  return cpx<R>((R)g.C0 * v, 0);
}

template <typename R>
__device__ inline cpx<R> kernel_entry(const DevGeometry& g, int kind, int i, int j) {
  return (kind == KIND_EFIE) ? efie_entry<R>(g, i, j) : laplace_entry<R>(g, i, j);
}

}  // namespace rcs
