// Matrix construction and host-side set-up algorithms (colouring,
// aggregation, Galerkin operators) of the solve path.
#pragma once
#include <functional>

#include "fh_common.h"

namespace fh {

// Builds a device BCSR (FP64 values) from host block arrays.  keep_host
// keeps the structure and FP64 values on the host for set-up algorithms.
std::unique_ptr<Bcsr> bcsr_from_blocks(Ctx& c, int nbr, int nbc, int bs_r, int bs_c, const int* rowptr,
                                       const int* colidx, const double* vals, bool keep_host);

// Builds a device BCSR with square bs x bs blocks from a scalar CSR (rows
// and columns multiples of bs).  Every block containing a stored entry is
// stored dense; missing entries are 0.
std::unique_ptr<Bcsr> bcsr_from_csr(Ctx& c, int n, int m, int bs, const int* rp, const int* ci, const double* v,
                                    bool keep_host);

// Copy of A with values stored in `p` (device conversion).
std::unique_ptr<Bcsr> bcsr_convert(Ctx& c, const Bcsr& A, Prec p);

// Block transpose (host set-up, FP64): A^T as BCSR.
std::unique_ptr<Bcsr> bcsr_transpose(Ctx& c, const Bcsr& A);

// Device copy of the index of the diagonal block of every block row.
std::vector<int> diag_blocks(const Bcsr& A);

// Greedy distance-1 colouring of the block graph in block-row order.
int greedy_colouring(const Bcsr& A, std::vector<int>& colour);

// --------------------------------------------------------------- AMG set-up
// One pass of pairwise aggregation on block rows of A (host FP64 values):
// algebraic mesh with block-trace couplings and strong neighbours, priority
// ordering and pairing.  Returns the number of aggregates; agg[i] per block
// row; members[J] lists the block rows of aggregate J in the order they
// were added.  coarsest is set when the coarsening ratio is <= 1.01.
int pairwise_aggregate(const Bcsr& A, double theta, std::vector<int>& agg, std::vector<std::vector<int>>& members,
                       bool& coarsest);

// Element of an algebraic mesh for pairwise aggregation
// (AlgebraicElement / HybridAlgebraicElement): neighbours with their
// coupling, sorted by ascending coupling; strong neighbours in that order;
// count = NElementsIAmStrongNeighbourOf.
struct AggElem {
    std::vector<std::pair<int, double>> nbr;
    std::vector<int> strong;
    int count = 0;
    int agg = -1;
};

// PairwiseAggregation::Perform (PairwiseAggregation.h:9-130) on built
// elements: priority ordering (breadth-first from the element with the
// smallest count, unnumbered neighbours by ascending count) and pairing
// with the available strong neighbour of smallest count.
int pairwise_pairing(std::vector<AggElem>& E, std::vector<int>& agg, std::vector<std::vector<int>>& members,
                     bool& coarsest);

// Galerkin operator of a piecewise-constant (block identity) prolongation:
// Ac(I,J) = sum_{k in I, l in J} A_kl, entries |a| <= 1e-15 dropped, host
// result (structure + FP64 values), device copy uploaded.
std::unique_ptr<Bcsr> galerkin_aggregate(Ctx& c, const Bcsr& A, const std::vector<int>& agg, int nagg);

// Galerkin operator P^T A P of a block prolongation P (host FP64 values of A
// and P; Level::ComputeGalerkinOperator, Level.h:133-137).
std::unique_ptr<Bcsr> galerkin_rap(Ctx& c, const Bcsr& A, const Bcsr& P);

// Galerkin operator of the p-injection keeping the first nfc modes of every
// block (hierarchical bases): the leading nfc x nfc sub-block of every block.
std::unique_ptr<Bcsr> bcsr_mode_submatrix(Ctx& c, const Bcsr& A, int nfc);

// Dense inverse of a small SPD matrix given as BCSR (host values), via
// Cholesky.  Throws on a non-positive pivot.
std::vector<double> dense_spd_inverse(const Bcsr& A);

}  // namespace fh
