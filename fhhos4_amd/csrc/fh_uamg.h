// Uncondensed aggregation AMG (the reference's UncondensedAMG,
// src/Solver/Multigrid/UncondensedAMG/*.h): a multigrid hierarchy for the
// condensed face system built from the uncondensed hybrid blocks A_TT, A_TF,
// A_FF.  Element aggregation on the A_TF couplings, interface collapsing
// into coarse faces, reconstruct-smoothed-trace-or-inject prolongations,
// chained over passes until the face count drops by the coarsening factor.
// Host set-up (C++); the levels it returns run on the GPU multigrid.
#pragma once
#include <string>
#include <vector>

#include "fh_common.h"

namespace fh {

// Scalar CSR on the host, row-major with ascending columns, as the
// reference's SparseMatrix (Eigen RowMajor, src/Utils/Types.h:11-22).  Every
// stored entry counts, zero or not.
struct Csr {
    int nr = 0, nc = 0;
    std::vector<int> rp{0}, ci;
    std::vector<double> v;
    int64_t nnz() const { return (int64_t)ci.size(); }
};

Csr csr_copy(int nr, int nc, const int* rp, const int* ci, const double* v);
Csr csr_transpose(const Csr& A);
// C = A B, every structural product entry kept; C(i,j) accumulated over the
// entries of row i of A in order, then row k of B in order (Eigen's
// conservative sparse product).
Csr csr_mul(const Csr& A, const Csr& B);
// The full matrix of selfadjointView<Lower>(): the lower triangle of A and
// its mirror.
Csr csr_sym_lower(const Csr& A);

struct UamgArgs {
    int cell_bs = 1;                  // HHO->nCellUnknowns
    int face_bs = 1;                  // HHO->nFaceUnknowns
    double theta = 0.25;              // strong coupling threshold (SolverFactory.h:215)
    double coarsening_factor = 3.8;   // main.cpp:1862-1865
    int max_coarse = 1000;            // MatrixMaxSizeForCoarsestLevel
    int max_levels = 0;               // 0: automatic
};

// One coarsening pass of a level (for inspection and parity checks).
struct UamgPass {
    int level = 0, pass = 0;
    int n_elems = 0, n_faces = 0, n_coarse_elems = 0, n_coarse_faces = 0;
    std::vector<int> elem_agg;    // coarse element of every element
    std::vector<int> face_coarse; // coarse face of every face, -1 for removed faces
};

struct UamgHierarchy {
    // operator of every level (level 0 = the condensed operator as given)
    std::vector<Csr> A;
    // prolongation from level l+1 to level l (l < levels - 1)
    std::vector<Csr> P;
    std::vector<UamgPass> passes;
};

// Builds the hierarchy (Multigrid::Setup with UncondensedLevel::CoarsenMesh,
// Multigrid.h:88-313, UncondensedLevel.h:74-185).
UamgHierarchy uamg_build(const Csr& A, const Csr& ATT, const Csr& ATF, const Csr& AFF, const UamgArgs& args);

}  // namespace fh
