// bilu0.cpp -- host-side structure of the superelement block-ILU(0) preconditioner
// (preconditioner = 2): superelement partition, face ordering, CRS pattern and
// the elimination index lists.  Restates backend/Common/cpuimpl.h:
// gridpartition2d/3d (:1332-1496), crs_faceordering (:1266-1330) with mke2f,
// mkfelem, mkf2eelem, facereorder, sortrecurrence, sortinteriorfaces, crs_array,
// mkface (:440-1264), and crs_indexingilu0 (:1498-1600).  The ordering is
// computed on the first superelement and applied to all of them (they are
// congruent on the structured grids cartgridpart describes).
#include "bilu0.hpp"

#include <algorithm>
#include <numeric>

namespace exa {

namespace {

std::vector<int> stable_perm(const std::vector<int>& v) {
  std::vector<int> p(v.size());
  std::iota(p.begin(), p.end(), 0);
  std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return v[a] < v[b]; });
  return p;
}

// e2f[e*nfe + l] = face
std::vector<int> make_e2f(const std::vector<std::array<int, 4>>& f2e, int nfe, int ne) {
  std::vector<int> e2f(static_cast<size_t>(nfe) * ne, -1);
  for (size_t i = 0; i < f2e.size(); ++i) {
    e2f[f2e[i][0] * nfe + f2e[i][1]] = static_cast<int>(i);
    if (f2e[i][2] >= 0) e2f[f2e[i][2] * nfe + f2e[i][3]] = static_cast<int>(i);
  }
  return e2f;
}

// f2f[i*nbf + k]: the other faces of the element(s) of face i (mkf2f, cpuimpl.h:460-495)
std::vector<int> make_f2f(const std::vector<std::array<int, 4>>& f2e, const std::vector<int>& e2f, int nfe) {
  int nbf = 2 * (nfe - 1);
  std::vector<int> f2f(f2e.size() * nbf, -1);
  for (size_t i = 0; i < f2e.size(); ++i) {
    int k = 0;
    for (int l = 0; l < nfe; ++l)
      if (l != f2e[i][1]) f2f[i * nbf + k++] = e2f[f2e[i][0] * nfe + l];
    if (f2e[i][2] >= 0)
      for (int l = 0; l < nfe; ++l)
        if (l != f2e[i][3]) f2f[i * nbf + k++] = e2f[f2e[i][2] * nfe + l];
  }
  return f2f;
}

}  // namespace

std::vector<int> grid_partition(const std::vector<int>& cart, int& nse, int& nese) {
  if (cart.empty()) throw Error(2, "preconditioner 2 needs cartgridpart in mesh.bin");
  int nd = cart[0];
  if (nd != 2 && nd != 3) throw Error(2, "cartgridpart: bad dimension");
  if (static_cast<int>(cart.size()) < 1 + 2 * nd) throw Error(2, "cartgridpart too short");
  std::vector<int> N(cart.begin() + 1, cart.begin() + 1 + nd), M(cart.begin() + 1 + nd, cart.begin() + 1 + 2 * nd);
  int opt = static_cast<int>(cart.size()) > 1 + 2 * nd ? cart[1 + 2 * nd] : 0;
  for (int d = 0; d < nd; ++d)
    if (M[d] <= 0 || N[d] % M[d]) throw Error(2, "grid dimensions must be divisible by the block sizes");
  std::vector<int> nb(nd);
  for (int d = 0; d < nd; ++d) nb[d] = N[d] / M[d];
  std::vector<std::vector<int>> blocks;
  if (nd == 2) {
    auto gid = [&](int i, int j) { return opt == 0 ? i + N[0] * j : j + N[1] * i; };
    auto add = [&](int jj, int ii) {
      std::vector<int> b;
      for (int j = 0; j < M[1]; ++j)
        for (int i = 0; i < M[0]; ++i) b.push_back(gid(ii * M[0] + i, jj * M[1] + j));
      blocks.push_back(b);
    };
    if (opt == 0) {
      for (int jj = 0; jj < nb[1]; ++jj)
        for (int ii = 0; ii < nb[0]; ++ii) add(jj, ii);
    } else {
      for (int ii = 0; ii < nb[0]; ++ii)
        for (int jj = 0; jj < nb[1]; ++jj) add(jj, ii);
    }
  } else {
    auto gid = [&](int i, int j, int k) {
      return opt == 0 ? i + N[0] * j + N[0] * N[1] * k : k + N[2] * j + N[2] * N[1] * i;
    };
    auto add = [&](int kk, int jj, int ii) {
      std::vector<int> b;
      for (int k = 0; k < M[2]; ++k)
        for (int j = 0; j < M[1]; ++j)
          for (int i = 0; i < M[0]; ++i) b.push_back(gid(ii * M[0] + i, jj * M[1] + j, kk * M[2] + k));
      blocks.push_back(b);
    };
    if (opt == 0) {
      for (int kk = 0; kk < nb[2]; ++kk)
        for (int jj = 0; jj < nb[1]; ++jj)
          for (int ii = 0; ii < nb[0]; ++ii) add(kk, jj, ii);
    } else {
      for (int ii = 0; ii < nb[0]; ++ii)
        for (int jj = 0; jj < nb[1]; ++jj)
          for (int kk = 0; kk < nb[2]; ++kk) add(kk, jj, ii);
    }
  }
  nse = static_cast<int>(blocks.size());
  nese = static_cast<int>(blocks[0].size());
  std::vector<int> elem(static_cast<size_t>(nse) * nese);  // elem[n + nse*i]
  for (int n = 0; n < nse; ++n)
    for (int i = 0; i < nese; ++i) elem[n + static_cast<size_t>(nse) * i] = blocks[n][i];
  return elem;
}

Bilu0Struct build_bilu0(const std::vector<int>& cart, const Conn& conn, int nfe, int ne) {
  Bilu0Struct s;
  std::vector<int> elem = grid_partition(cart, s.nse, s.nese);
  for (int v : elem)
    if (v < 0 || v >= ne) throw Error(2, "cartgridpart does not match the mesh");
  std::vector<std::array<int, 4>> f2e(conn.nf);
  int emax = 0;
  for (int f = 0; f < conn.nf; ++f) {
    for (int k = 0; k < 4; ++k) f2e[f][k] = conn.f2e[4 * f + k];
    emax = std::max(emax, std::max(f2e[f][0], f2e[f][2]));
  }
  std::vector<int> e2f = make_e2f(f2e, nfe, emax + 1);
  // faces of superelement 0 (mkfelem) and their local f2e (mkf2eelem)
  std::vector<int> el0(s.nese);
  for (int i = 0; i < s.nese; ++i) el0[i] = elem[0 + static_cast<size_t>(s.nse) * i];
  std::vector<int> felem;
  for (int e : el0)
    for (int j = 0; j < nfe; ++j) {
      int f = e2f[e * nfe + j];
      if (std::find(felem.begin(), felem.end(), f) == felem.end()) felem.push_back(f);
    }
  auto local = [&](int e) {
    for (int i = 0; i < s.nese; ++i)
      if (el0[i] == e) return i;
    return -1;
  };
  std::vector<std::array<int, 4>> loc;
  for (int f : felem) {
    int a1 = local(f2e[f][0]), a2 = f2e[f][2] >= 0 ? local(f2e[f][2]) : -1;
    if (a1 >= 0) loc.push_back({a1, f2e[f][1], a2, a2 >= 0 ? f2e[f][3] : -1});
    else loc.push_back({a2, f2e[f][3], -1, -1});
  }
  const int nf = static_cast<int>(loc.size());
  // facereorder: superelement-boundary faces grouped by recurrence of their element
  std::vector<int> ind0, first;
  for (int i = 0; i < nf; ++i)
    if (loc[i][3] == -1) {
      ind0.push_back(i);
      first.push_back(loc[i][0]);
    }
  std::vector<int> cnt(first.size());
  int maxrec = 0;
  for (size_t i = 0; i < first.size(); ++i) {
    cnt[i] = static_cast<int>(std::count(first.begin(), first.end(), first[i]));
    maxrec = std::max(maxrec, cnt[i]);
  }
  std::vector<int> order;
  for (int k = 1; k <= maxrec; ++k) {
    std::vector<int> grp, vals;
    for (size_t i = 0; i < first.size(); ++i)
      if (cnt[i] == k) {
        grp.push_back(static_cast<int>(i));
        vals.push_back(first[i]);
      }
    for (int p : stable_perm(vals)) order.push_back(ind0[grp[p]]);
  }
  std::vector<char> isb(nf, 0);
  for (int i : ind0) isb[i] = 1;
  for (int i = 0; i < nf; ++i)
    if (!isb[i]) order.push_back(i);
  std::vector<std::array<int, 4>> tmp(nf);
  for (int i = 0; i < nf; ++i) tmp[i] = loc[order[i]];
  loc = tmp;
  // sortinteriorfaces
  const int nbf = 2 * (nfe - 1);
  int nel = 0;
  for (auto& q : loc) nel = std::max(nel, std::max(q[0], q[2]) + 1);
  int nfs = static_cast<int>(ind0.size());
  while (nfs < nf) {
    std::vector<int> le2f = make_e2f(loc, nfe, nel);
    std::vector<int> f2f = make_f2f(loc, le2f, nfe);
    std::vector<int> counts;
    for (int i = nfs; i < nf; ++i) {
      int c = 0;
      for (int k = 0; k < nbf; ++k)
        if (f2f[static_cast<size_t>(i) * nbf + k] >= nfs) ++c;
      counts.push_back(c);
    }
    std::vector<int> p = stable_perm(counts);
    std::vector<std::array<int, 4>> t2(nf - nfs);
    for (int q = 0; q < nf - nfs; ++q) t2[q] = loc[nfs + p[q]];
    for (int q = 0; q < nf - nfs; ++q) loc[nfs + q] = t2[q];
    int inc = 0;
    for (int c : counts)
      if (c < nbf) ++inc;
    if (inc == 0) throw Error(2, "superelement face ordering does not progress");
    nfs += inc;
  }
  // crs_array
  std::vector<int> le2f = make_e2f(loc, nfe, nel);
  std::vector<int> f2f = make_f2f(loc, le2f, nfe);
  s.row_ptr.assign(1, 0);
  for (int i = 0; i < nf; ++i) {
    std::vector<int> nb;
    for (int k = 0; k < nbf; ++k)
      if (f2f[static_cast<size_t>(i) * nbf + k] >= 0) nb.push_back(f2f[static_cast<size_t>(i) * nbf + k]);
    std::sort(nb.begin(), nb.end());
    s.col_ind.push_back(i);
    s.col_ind.insert(s.col_ind.end(), nb.begin(), nb.end());
    s.row_ptr.push_back(static_cast<int>(s.col_ind.size()));
  }
  s.nfse = nf;
  s.nnz = s.row_ptr[nf];
  // mkface: face[n + nse*i]
  s.face.assign(static_cast<size_t>(s.nse) * nf, -1);
  for (int n = 0; n < s.nse; ++n)
    for (int i = 0; i < nf; ++i)
      s.face[n + static_cast<size_t>(s.nse) * i] =
          e2f[elem[n + static_cast<size_t>(s.nse) * loc[i][0]] * nfe + loc[i][1]];
  for (int v : s.face)
    if (v < 0) throw Error(2, "superelement face map is incomplete");
  // crs_indexingilu0 as flat lists
  s.diag.resize(nf);
  s.eptr.assign(1, 0);
  s.uptr.assign(1, 0);
  s.lptr.assign(1, 0);
  s.uuptr.assign(1, 0);
  auto find_in_row = [&](int r, int c) {
    for (int q = s.row_ptr[r]; q < s.row_ptr[r + 1]; ++q)
      if (s.col_ind[q] == c) return q;
    return -1;
  };
  for (int i = 0; i < nf; ++i) {
    s.diag[i] = find_in_row(i, i);
    if (s.diag[i] < 0) throw Error(2, "ILU0: missing diagonal block");
    for (int p = s.row_ptr[i]; p < s.row_ptr[i + 1]; ++p) {
      int j = s.col_ind[p];
      if (j <= i) continue;
      int idx_ji = find_in_row(j, i);
      if (idx_ji < 0) continue;
      s.elim_ji.push_back(idx_ji);
      for (int pp = s.row_ptr[i]; pp < s.row_ptr[i + 1]; ++pp) {
        int ell = s.col_ind[pp];
        if (ell <= i) continue;
        int idx_jl = find_in_row(j, ell);
        if (idx_jl < 0) continue;
        s.upd_jl.push_back(idx_jl);
        s.upd_il.push_back(pp);
        s.upd_ji.push_back(idx_ji);
      }
      s.uptr.push_back(static_cast<int>(s.upd_jl.size()));
    }
    s.eptr.push_back(static_cast<int>(s.elim_ji.size()));
    for (int p = s.row_ptr[i]; p < s.row_ptr[i + 1]; ++p) {
      if (s.col_ind[p] < i) {
        s.lidx.push_back(p);
        s.lcol.push_back(s.col_ind[p]);
      }
    }
    s.lptr.push_back(static_cast<int>(s.lidx.size()));
    for (int p = s.row_ptr[i]; p < s.row_ptr[i + 1]; ++p) {
      if (s.col_ind[p] > i) {
        s.uidx.push_back(p);
        s.ucol.push_back(s.col_ind[p]);
      }
    }
    s.uuptr.push_back(static_cast<int>(s.uidx.size()));
  }
  // each global face appears in one or two superelements: slots for the deterministic scatter
  s.slot1.assign(conn.nf, -1);
  s.slot2.assign(conn.nf, -1);
  for (int i = 0; i < nf; ++i)
    for (int n = 0; n < s.nse; ++n) {
      int f = s.face[n + static_cast<size_t>(s.nse) * i];
      int slot = n + s.nse * i;
      if (s.slot1[f] < 0) s.slot1[f] = slot;
      else if (s.slot2[f] < 0) s.slot2[f] = slot;
      else throw Error(2, "a face lies in more than two superelements");
    }
  for (int f = 0; f < conn.nf; ++f)
    if (s.slot1[f] < 0) throw Error(2, "cartgridpart does not cover every face");
  return s;
}

}  // namespace exa
