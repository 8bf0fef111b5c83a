// The reference driver's flow (test/sphere.cpp:80-262) written against the drop-in C++
// interface include/rcs_amd.hpp: mesh files -> EFIEMatrix -> RWG midpoints and bounding
// box -> inputsToSolverClass -> AIFMM -> factorize -> getRhs(Einc = x) -> solve and
// backSubstitute1/2/3 -> getPhi -> RCS. Also exercises LowRank::ACA_only_nodes2, RRQR and kernel::getMatrix.
//
// usage: driver_sphere nodeList patchList freq tol_pow leaf_size out_prefix
// writes <out_prefix>_phi.txt (N lines "re im", input order), <out_prefix>_box.txt (box
// order), <out_prefix>_rcs.txt (theta rcs) and <out_prefix>_aca.txt / _rrqr.txt.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "rcs_amd.hpp"

static void write_vec(const std::string& f, const rcs::Vec& v) {
  std::ofstream o(f);
  o << std::setprecision(17);
  for (const auto& z : v) o << z.real() << " " << z.imag() << "\n";
}

int main(int argc, char** argv) {
  if (argc < 7) {
    std::cerr << "usage: driver_sphere nodeList patchList freq tol_pow leaf_size out_prefix\n";
    return 2;
  }
  try {
    rcs::inputsToKernelClass inputsToKernel;
    inputsToKernel.nodeListFilename = argv[1];
    inputsToKernel.patchListFilename = argv[2];
    const double freq = std::atof(argv[3]);
    const int TOL_POW = std::atoi(argv[4]), leafSize = std::atoi(argv[5]);
    const std::string out = argv[6];

    rcs::EFIEMatrix* mykernel = new rcs::EFIEMatrix(inputsToKernel, freq);
    const int N = mykernel->N;
    std::vector<double> mid = mykernel->locations();  // RWG edge midpoints, N x 3
    rcs::inputsToSolverClass inputsToSolver;
    inputsToSolver.xmin = inputsToSolver.ymin = inputsToSolver.zmin = 1e300;
    inputsToSolver.xmax = inputsToSolver.ymax = inputsToSolver.zmax = -1e300;
    for (int i = 0; i < N; ++i) {
      inputsToSolver.xmin = std::min(inputsToSolver.xmin, mid[3 * i]);
      inputsToSolver.xmax = std::max(inputsToSolver.xmax, mid[3 * i]);
      inputsToSolver.ymin = std::min(inputsToSolver.ymin, mid[3 * i + 1]);
      inputsToSolver.ymax = std::max(inputsToSolver.ymax, mid[3 * i + 1]);
      inputsToSolver.zmin = std::min(inputsToSolver.zmin, mid[3 * i + 2]);
      inputsToSolver.zmax = std::max(inputsToSolver.zmax, mid[3 * i + 2]);
    }
    inputsToSolver.leafSize = leafSize;
    inputsToSolver.tol_pow = TOL_POW;
    inputsToSolver.fillin_tol = TOL_POW;
    inputsToSolver.fillin_rank = 0;
    inputsToSolver.locations = mid;
    std::cout << "N: " << N << std::endl;

    rcs::Node Einc;
    Einc.x = 1.0;
    rcs::Vec rhs;
    mykernel->getRhs(Einc, rhs);

    rcs::AIFMM<rcs::EFIEMatrix>* aifmm = new rcs::AIFMM<rcs::EFIEMatrix>(inputsToSolver, mykernel);
    aifmm->factorize();
    rcs::Vec xbox = aifmm->solve(rhs);  // box order (src/AIFMM.cpp:102-110)
    // the reference driver's own solve sequence (test/sphere.cpp:253-264)
    aifmm->backSubstitute1(rhs);
    aifmm->backSubstitute2();
    aifmm->backSubstitute3();
    rcs::Vec phi;
    aifmm->getPhi(phi);
    write_vec(out + "_phi.txt", phi);
    write_vec(out + "_box.txt", xbox);

    std::vector<double> theta(181), phiang(181, 0.0);
    for (int i = 0; i < 181; ++i) theta[i] = 3.14159265358979323846 * i / 180.0;
    std::vector<double> sigma = mykernel->getRCS(phi, theta, phiang);
    {
      std::ofstream o(out + "_rcs.txt");
      o << std::setprecision(17);
      for (int i = 0; i < 181; ++i) o << theta[i] << " " << sigma[i] << "\n";
    }

    // LowRank::ACA_only_nodes2 on the first and last thirds of the unknowns
    std::vector<int> rows, cols;
    for (int i = 0; i < N / 3; ++i) rows.push_back(i);
    for (int i = 2 * N / 3; i < N; ++i) cols.push_back(i);
    rcs::LowRank<rcs::EFIEMatrix> lr(mykernel, 1e-6, rows, cols);
    std::vector<int> rb, cb;
    int rank = 0;
    rcs::Mat Ac, Ar;
    lr.ACA_only_nodes2(rb, cb, rank, Ac, Ar);
    {
      std::ofstream o(out + "_aca.txt");
      o << rank << " " << Ac.rows() << " " << Ac.cols() << " " << Ar.rows() << " " << Ar.cols() << "\n";
      for (int i = 0; i < rank; ++i) o << rb[i] << " " << cb[i] << "\n";
    }

    // RRQR of the same far block with a partition of 3 columns
    rcs::Mat B = mykernel->getMatrix(rows, cols);
    rcs::RRQR qr(B, TOL_POW, TOL_POW, 3, 0.0, 0, 1000);
    qr.colPivHouseholderQRWithPartition();
    {
      std::ofstream o(out + "_rrqr.txt");
      o << std::setprecision(17) << qr.rank << " " << qr.maxPivot << "\n";
    }
    delete aifmm;
    delete mykernel;
  } catch (const std::exception& e) {
    std::cerr << "driver_sphere: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
