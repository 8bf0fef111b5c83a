"""Mie series for a perfectly conducting sphere (Bohren & Huffman, ch. 4 and 5)."""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def pec_backscatter(ka, a, nmax=None):
    """Monostatic RCS (m^2) of a PEC sphere of radius a at size parameter ka:
    sigma = (pi a^2 / x^2) |sum_n (2n+1) (-1)^n (a_n - b_n)|^2 with the PEC limits
    a_n = psi_n'(x)/xi_n'(x), b_n = psi_n(x)/xi_n(x), psi = x j_n, xi = x h_n^(1)."""
    x = float(ka)
    if nmax is None:
        nmax = int(x + 4 * x ** (1 / 3) + 10)
    s = 0j
    for n in range(1, nmax + 1):
        jn, jnp = spherical_jn(n, x), spherical_jn(n, x, derivative=True)
        yn, ynp = spherical_yn(n, x), spherical_yn(n, x, derivative=True)
        hn, hnp = jn + 1j * yn, jnp + 1j * ynp
        psi, dpsi = x * jn, jn + x * jnp
        xi, dxi = x * hn, hn + x * hnp
        an, bn = dpsi / dxi, psi / xi
        s += (2 * n + 1) * (-1) ** n * (an - bn)
    return np.pi * a * a / (x * x) * abs(s) ** 2
