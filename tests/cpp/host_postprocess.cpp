// The reference driver's mesh and post-processing flow (test/sphere.cpp:82-151, 276-316)
// through include/rcs_amd.hpp on meshes with host tables only, so that it runs without a GPU:
// RWG midpoints through edge2PatchMap and nodeList, the plane-wave right-hand side,
// getJAtCentroids and getRCS_OnePointQuadrature for fixed coefficients, and the
// ElectrostaticsMatrix's getRhs / getCapacitance / getCentroids. Writes text files that
// tests/test_mesh_and_abi.py compares with the oracle.
//
// usage: host_postprocess nodeList patchList out_prefix
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>

#include "rcs_amd.hpp"

using namespace rcs;

namespace {
void write_complex(const std::string& path, const std::vector<cdouble>& v) {
  std::ofstream o(path);
  o << std::setprecision(17);
  for (const cdouble& c : v) o << c.real() << " " << c.imag() << "\n";
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const std::string out = argv[3];
  inputsToKernelClass inputsToKernel;
  inputsToKernel.nodeListFilename = argv[1];
  inputsToKernel.patchListFilename = argv[2];
  try {
    EFIEMatrix* mykernel = new EFIEMatrix(inputsToKernel, 1.0e9, /*hostOnly=*/true);
    const int N = mykernel->N;
    {
      std::ofstream o(out + "_mid.txt");
      o << std::setprecision(17);
      for (int i = 0; i < N; i++) {  // S:143-152
        auto itrEdge = std::next(mykernel->edge2PatchMap.begin(), i);
        std::pair<int, int> edge = itrEdge->first;
        Node mid = (mykernel->nodeList[edge.first] + mykernel->nodeList[edge.second]) / 2.0;
        o << mid(0) << " " << mid(1) << " " << mid(2) << "\n";
      }
    }
    Node Einc;
    Einc(0) = 1.0;
    Vec rhs;
    mykernel->getRhs(Einc, rhs);
    write_complex(out + "_rhs.txt", rhs);
    Vec coefficients(N);
    for (int i = 0; i < N; ++i) coefficients[i] = cdouble(0.001 * (i + 1), 0.5 - 0.0007 * i);
    Mat JAtCentroids;
    mykernel->getJAtCentroids(coefficients, JAtCentroids);
    write_complex(out + "_J.txt", JAtCentroids.v);  // column-major n_patch x 3
    {
      std::ofstream o(out + "_rcs.txt");
      o << std::setprecision(17);
      for (int k = 0; k < 19; ++k) {
        const double theta = 3.14159265358979323846 * k / 18.0;
        o << theta << " " << mykernel->getRCS_OnePointQuadrature(theta, 0.0, JAtCentroids) << "\n";
      }
    }
    {
      std::ofstream o(out + "_info.txt");
      o << N << " " << mykernel->conductorIndices.size() << " " << mykernel->patchList.size() << "\n";
    }
    delete mykernel;

    ElectrostaticsMatrix* lap = new ElectrostaticsMatrix(inputsToKernel, /*hostOnly=*/true);
    std::vector<double> b, sigma(lap->N, 1.0), centroids;
    lap->getRhs(0, b);
    double ones = 0;
    for (double v : b) ones += v;
    const double C = lap->getCapacitance(sigma);  // with unit density: the total area
    lap->getCentroids(centroids);
    {
      std::ofstream o(out + "_lap.txt");
      o << std::setprecision(17) << lap->N << " " << ones << " " << C << "\n";
      for (int i = 0; i < lap->N; ++i)
        o << centroids[3 * i] << " " << centroids[3 * i + 1] << " " << centroids[3 * i + 2] << "\n";
    }
    delete lap;
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
