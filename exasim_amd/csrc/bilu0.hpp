// bilu0.hpp -- host-side structure of the superelement block-ILU(0) preconditioner.
#pragma once

#include "common.hpp"

#include <array>
#include <vector>

namespace exa {

struct Bilu0Struct {
  int nse = 0, nese = 0, nfse = 0, nnz = 0;
  std::vector<int> face;              // [nse, nfse]: face[n + nse*i] = global face of slot i
  std::vector<int> row_ptr, col_ind;  // CRS pattern of one superelement (nfse rows)
  std::vector<int> diag;              // CRS index of the diagonal block of row i
  // elimination lists of the factorisation (row i: entries eptr[i]..eptr[i+1]-1;
  // entry k updates upd_jl/upd_il[uptr[k]..uptr[k+1]-1])
  std::vector<int> eptr, elim_ji, uptr, upd_jl, upd_il;
  std::vector<int> upd_ji;            // the L block (elim_ji) each update multiplies
  // lower / upper entries of each row for the triangular sweeps
  std::vector<int> lptr, lidx, lcol, uuptr, uidx, ucol;
  // each global face lies in one or two superelements (slot = n + nse*i)
  std::vector<int> slot1, slot2;
};

std::vector<int> grid_partition(const std::vector<int>& cart, int& nse, int& nese);
Bilu0Struct build_bilu0(const std::vector<int>& cart, const Conn& conn, int nfe, int ne);

}  // namespace exa
