// Exported linear-system format: Matrix Market text files as the reference
// writes them (ExportModule::ExportMatrix / ExportVector,
// src/Utils/ExportModule.h:68-87, through Eigen's saveMarket /
// saveMarketVector: header line, "rows cols nnz", then 1-based "i j v"
// entries in row-major order, values printed "%.64e").
#pragma once
#include <string>
#include <vector>

#include "fh_common.h"

namespace fh {

// Writes the scalar operator of A.  Stored entries: the scalar pattern A was
// built from (stored-entry mask) or, without a mask, every non-zero block
// entry.  Values come from the host copy, else from the FP64 device copy.
void market_write_matrix(Ctx& c, const Bcsr& A, const std::string& path);

// Reads a "coordinate real general" file into scalar CSR (rows sorted,
// columns ascending, duplicates summed as Eigen's setFromTriplets does).
void market_read_csr(const std::string& path, int& nrows, int& ncols, std::vector<int>& rowptr,
                     std::vector<int>& colidx, std::vector<double>& vals);

// Writes a scalar CSR matrix (columns ascending in every row).
void market_write_csr(int nrows, int ncols, const int* rp, const int* ci, const double* v, const std::string& path);

void market_write_vector(int64_t n, const double* v, const std::string& path);
// Reads an "array real general" n x 1 file.
std::vector<double> market_read_vector(const std::string& path);

}  // namespace fh
