// Batched static condensation of caller-supplied element matrices and the
// recovery of the cell unknowns (device arrays; the cell-to-face numbers
// on the host).
#pragma once
#include <memory>

#include "fh_common.h"

namespace fh {

// Face system of ncells cells with nfc faces of nf unknowns and nc cell
// unknowns each.  cell_faces (host, ncells x nfc): system face number or
// -1; local (device, ncells x nl x nl, nl = nc + nfc nf): each cell's local
// matrix, row-major.  BT (ncells x nc), BF (nfaces x nf) may be null (zero);
// with b_out (nfaces x nf) non-null the right-hand side is formed too.
std::unique_ptr<Bcsr> condense_batched(Ctx& c, int ncells, int nfc, int nc, int nf, int nfaces,
                                       const int* cell_faces, const double* local, const double* BT,
                                       const double* BF, double* b_out);

// x_T = A_TT^-1 (B_T - A_TF x_F) for every cell (xT: ncells x nc).
void cell_recover_batched(Ctx& c, int ncells, int nfc, int nc, int nf, int nfaces, const int* cell_faces,
                          const double* local, const double* BT, const double* xF, double* xT);

}  // namespace fh
