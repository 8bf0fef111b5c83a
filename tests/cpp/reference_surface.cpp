// Compile-only check that include/rcs_amd.hpp offers the members the reference's driver
// (test/sphere.cpp:80-320) and its tree use on EFIEMatrix, ElectrostaticsMatrix, LowRank,
// RRQR and AIFMM, with the reference's names and argument order. Eigen types are replaced
// by rcs::Mat / rcs::Vec / std::vector<double>. Built with -fsyntax-only by
// tests/test_mesh_and_abi.py; tests/cpp/driver_sphere.cpp is the version that runs.
#include <iostream>
#include <iterator>

#include "rcs_amd.hpp"

using namespace rcs;

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  inputsToKernelClass inputsToKernel;
  inputsToKernel.nodeListFilename = argv[1];
  inputsToKernel.patchListFilename = argv[2];

  // S:82-83, 142-151: N, conductorIndices, RWG midpoints through edge2PatchMap and nodeList
  EFIEMatrix* mykernel = new EFIEMatrix(inputsToKernel, 1.0e9);
  int N = mykernel->N;
  int numConductors = mykernel->conductorIndices.size();
  std::vector<double> midPointOfRWGEdgeMatrix(3 * (size_t)N);
  for (int i = 0; i < N; i++) {
    auto itrEdge = std::next(mykernel->edge2PatchMap.begin(), i);
    std::pair<int, int> edge = itrEdge->first;
    Node node1 = mykernel->nodeList[edge.first];
    Node node2 = mykernel->nodeList[edge.second];
    Node mid = (node1 + node2) / 2.0;
    for (int d = 0; d < 3; ++d) midPointOfRWGEdgeMatrix[3 * i + d] = mid(d);
  }
  double L = mykernel->edge2PatchMap.begin()->second.length;  // E:712, 724
  int tp = mykernel->edge2PatchMap.begin()->second.indicesOfAssociatedPatches[0];

  // S:159-174
  inputsToSolverClass inputsToSolver;
  inputsToSolver.xmin = -1; inputsToSolver.xmax = 1;
  inputsToSolver.ymin = -1; inputsToSolver.ymax = 1;
  inputsToSolver.zmin = -1; inputsToSolver.zmax = 1;
  inputsToSolver.leafSize = 64;
  inputsToSolver.tol_pow = 6;
  inputsToSolver.locations = midPointOfRWGEdgeMatrix;
  inputsToSolver.fillin_rank = 0;
  inputsToSolver.fillin_tol = 6;

  // S:178-194: plane wave, dense LU reference solution
  Node Einc;
  Einc(0) = 1.0;
  Einc(1) = 0.0;
  Einc(2) = 0.0;
  Vec rhs;
  mykernel->getRhs(Einc, rhs);
  Vec coefficientsLU;
  mykernel->getCoefficients(rhs, coefficientsLU);

  // S:205-264: AIFMM, factorize, the three-part solve, getPhi
  AIFMM<EFIEMatrix>* aifmm = new AIFMM<EFIEMatrix>(inputsToSolver, mykernel);
  int avgLeafSize = aifmm->A->getAvgLeafSize();          // S:217
  int avgRank = aifmm->A->getAvgMultipoles();            // S:223
  double mem1 = 2 * aifmm->A->findMemory();              // S:224
  aifmm->factorize();
  double mem2 = 2 * aifmm->A->findMemoryAfterFactorize();  // S:244
  aifmm->backSubstitute1(rhs);
  aifmm->backSubstitute2();
  aifmm->backSubstitute3();
  Vec coefficients;
  aifmm->getPhi(coefficients);
  Vec xbox = aifmm->solve(rhs);
  aifmm->backSubstitute(rhs);

  // S:276-316: currents at the centroids and the monostatic RCS
  Mat JAtCentroids;
  mykernel->getJAtCentroids(coefficientsLU, JAtCentroids);
  double rcs = mykernel->getRCS_OnePointQuadrature(0.0, 0.0, JAtCentroids);

  // kernel (include/kernel.hpp:13-47)
  std::vector<int> rows{0, 1}, cols{2, 3};
  cdouble e = mykernel->getMatrixEntry(0, 1);
  Vec r1 = mykernel->getRow(0, cols);
  Vec c1 = mykernel->getCol(2, rows);
  Vec r2 = mykernel->getRow(N, 0);
  Vec c2 = mykernel->getCol(N, 0);
  Mat m1 = mykernel->getMatrix(rows, cols);
  Mat m2 = mykernel->getMatrix(0, 0, 2, 2);
  Mat m3;
  mykernel->getMatrix(m3, 2);
  Patch p = mykernel->patchList[0];

  // LowRank<K> (include/ACA.hpp:16,26) and RRQR (include/rrqr.hpp:5-27)
  LowRank<EFIEMatrix>* LR = new LowRank<EFIEMatrix>(mykernel, 1e-6, rows, cols);
  std::vector<int> row_bases, col_bases;
  int computed_rank;
  Mat Ac, Ar;
  LR->ACA_only_nodes2(row_bases, col_bases, computed_rank, Ac, Ar);
  delete LR;
  RRQR* rrqr = new RRQR(m1, 10, 10, 0, 0.0, 0, 5);
  rrqr->colPivHouseholderQRWithPartition();
  Mat Q = rrqr->Q_trunc;
  double maxPivot = rrqr->maxPivot;
  double err = rrqr->errorInQR();
  delete rrqr;

  // ElectrostaticsMatrix (include/ElectrostaticsMatrix.hpp:8-40)
  ElectrostaticsMatrix* lap = new ElectrostaticsMatrix(inputsToKernel);
  std::vector<double> b, sigma, centroids;
  lap->getRhs(0, b);
  lap->getchargeDensityUsingLU(b, sigma);
  double C = lap->getCapacitance(sigma);
  lap->getCentroids(centroids);
  double area = lap->patchList[0].area;

  std::cout << numConductors << L << tp << rcs << e << r1.size() << c1.size() << r2.size() << c2.size() << m2.rows()
            << m3.cols() << p.area << Q.cols() << maxPivot << err << C << area << xbox.size() << coefficients.size() << avgLeafSize << avgRank
            << mem1 << mem2
            << std::endl;
  delete lap;
  delete aifmm;
  delete mykernel;
  return 0;
}
