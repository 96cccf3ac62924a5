// Device-side description of an HHO problem for the batched local-operator
// and static-condensation kernel (fh_hho_kernels.hip).
#pragma once
#include "fh_common.h"

namespace fh {

struct HhoDev {
    int dim, kind, nvc, nfc, nvf;  // cell kind 0 simplex, 1 box
    int nc, nr, nfb, nq, ncell;    // cell / reconstruction / face basis sizes, cell rule size
    const double* V;               // vertices, 3 per vertex
    const int* cv;                 // cell vertices (nvc per cell)
    const int* cf;                 // cell faces (nfc per cell)
    const int* fv;                 // face vertices (nvf per face)
    const double* K;               // diffusion tensor, 9 per cell
    const int* cft;                // face table of every (cell, local face)
    const double* Shat;            // 9 x nr x nr reference stiffness tensors
    const double* cellPsi;         // nq x nr reconstruction basis at the cell rule
    const double* cqp;             // nq x 3 reference points
    const double* cqw;             // nq weights
    const double* Phi;             // per table: 3 x nr x nfb
    const double* Psi;             // per table: 3 x nr x nc
    const double* MFT;             // per table: nfb x nr
    int src;                       // 0 pi-sine, 1 4pi-sine, 2 random modes
    int nmodes;
    const double* modes;           // nmodes x 4: a, p0, p1, p2
    const int* blk;                // nfc x nfc per cell: block index of (face f, face g), -1 if not stored
    const int* frow;               // nfc per cell: block row of the face, -1 for a Dirichlet face
    double* vals;                  // block values (nf x nf row-major), accumulated
    double* b;                     // right-hand side, accumulated
    // element layout (optional, elem_vals null: none): the condensed cell
    // matrix by cyclic diagonals (entry (r, s) with t = (s - r) mod nF <=
    // nF / 2 at t nF + r, nF = nfc * nfb unknowns), elem_stride values per
    // cell, cell c at position elem_pos[c]
    double* elem_vals;
    const int* elem_pos;
    int elem_stride;
    // post-processing mode (post = 1): instead of the scatter, recover the
    // cell unknowns from the face solution xsol (nb x nfb), reconstruct the
    // k+1 polynomial and write the cell's integrals of (u - p)^2 and u^2
    // into cell_err[2c], cell_err[2c+1]
    int post;
    const double* xsol;
    double* cell_err;
    // general cells (kind 2, bilinear / trilinear maps): Gauss rules on
    // [-1, 1] (ng1 points for the operators, nl1 for the L2 error: nodes
    // then weights), reconstruction exponents (nr x 3, graded; the first nc
    // are the cell basis) and face exponents (nfb x 2) of the tensor
    // Legendre bases
    int ng1, nl1;
    const double* g1;
    const double* l1;
    const int* rexp;
    const int* fexp;
};

// LDS bytes per cell (one workgroup per cell).
size_t hho_local_lds_bytes(const HhoDev& d);
// Batched local operators + static condensation + scatter (vals and b must
// be zeroed), or the post-processing mode.  Synchronises the stream.
void launch_hho_local(Ctx& c, const HhoDev& d);

}  // namespace fh
