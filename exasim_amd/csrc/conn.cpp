// conn.cpp -- element/face connectivity for one process, restating buildConn
// (backend/Discretization/connectivity.cpp:1212-1385) and the helpers it calls:
// mkf2e_hash (:182-231), connectivity (:426-497), fix_f2t_consistency (:301-335),
// build_connectivity (:337-424), permindex (:1128-1205), divide_interval (:948-968).
//
// Face numbering: interior faces in the order they are completed while scanning
// elements then local faces, then boundary faces stably sorted by boundary
// condition.  Within one condition boundary faces are taken in order of first
// visit (the reference takes them in std::unordered_map iteration order).
#include "common.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <unordered_map>

namespace exa {

namespace {

// connectivity.cpp:1010-1106 (getelemface), 0-based
std::vector<std::vector<int>> local_faces(int nd, int elemtype) {
  if (nd == 2 && elemtype == 1) return {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  if (nd == 3 && elemtype == 1)
    return {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {1, 2, 6, 5}, {3, 0, 4, 7}};
  if (nd == 2 && elemtype == 0) return {{1, 2}, {2, 0}, {0, 1}};
  if (nd == 3 && elemtype == 0) return {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
  throw Error(5, "unsupported element type");
}

struct Key {
  std::array<int, 4> v{{-1, -1, -1, -1}};
  bool operator==(const Key& o) const { return v == o.v; }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int x : k.v) {
      h ^= static_cast<uint64_t>(static_cast<uint32_t>(x));
      h *= 1099511628211ull;
    }
    return static_cast<size_t>(h);
  }
};

// permindex (connectivity.cpp:1128-1205): for each orientation k, perm[i + npf*k] = j
// with xpf[i] == T_k(xpf[j]).
std::vector<int> permindex(const MasterData& m, int& ncols) {
  int npf = m.npf;
  if (m.nd == 2) {
    ncols = 1;
    std::vector<int> p(npf);
    for (int i = 0; i < npf; ++i) p[i] = npf - 1 - i;
    return p;
  }
  const double* s = m.xpf.data();
  const double* t = m.xpf.data() + npf;
  ncols = (m.elemtype == 1) ? 4 : 3;
  std::vector<int> p(static_cast<size_t>(npf) * ncols, -1);
  for (int k = 0; k < ncols; ++k) {
    for (int j = 0; j < npf; ++j) {
      double a, b;
      if (m.elemtype == 1) {
        if (k == 0) { a = t[j]; b = s[j]; }
        else if (k == 1) { a = s[j]; b = 1 - t[j]; }
        else if (k == 2) { a = 1 - t[j]; b = 1 - s[j]; }
        else { a = 1 - s[j]; b = t[j]; }
      } else {
        if (k == 0) { a = t[j]; b = s[j]; }
        else if (k == 1) { a = 1 - s[j] - t[j]; b = t[j]; }
        else { a = s[j]; b = 1 - s[j] - t[j]; }
      }
      for (int i = 0; i < npf; ++i)
        if (std::fabs(a - s[i]) < 1e-8 && std::fabs(b - t[i]) < 1e-8) {
          p[i + npf * k] = j;
          break;
        }
    }
  }
  for (int v : p)
    if (v < 0) throw Error(2, "face node permutation not found");
  return p;
}

}  // namespace

Conn build_conn(const MasterData& master, const MeshData& mesh) {
  const int nd = master.nd, elemtype = master.elemtype;
  const int npe = master.npe, npf = master.npf;
  const int ne = mesh.ne, nfe = mesh.nfe, nve = mesh.nve;
  auto faces = local_faces(nd, elemtype);
  if (static_cast<int>(faces.size()) != nfe) throw Error(2, "nfe does not match the element type");
  if (static_cast<int>(mesh.perm.size()) != npf * nfe) throw Error(2, "perm size mismatch");
  const int nvf = static_cast<int>(faces[0].size());

  // mkf2e_hash: pair faces by their sorted vertex ids
  struct Seen { int e, l; long seq; };
  std::unordered_map<Key, Seen, KeyHash> seen;
  seen.reserve(static_cast<size_t>(ne) * nfe);
  std::vector<std::array<int, 4>> interior;
  long seq = 0;
  for (int e = 0; e < ne; ++e) {
    for (int l = 0; l < nfe; ++l) {
      Key k;
      for (int v = 0; v < nvf; ++v) k.v[v] = mesh.t[faces[l][v] + nve * e];
      std::sort(k.v.begin(), k.v.begin() + nvf);
      auto it = seen.find(k);
      if (it == seen.end()) {
        seen.emplace(k, Seen{e, l, seq++});
      } else {
        interior.push_back({it->second.e, it->second.l, e, l});
        seen.erase(it);
      }
    }
  }
  std::vector<Seen> bnd;
  bnd.reserve(seen.size());
  for (auto& kv : seen) bnd.push_back(kv.second);
  std::sort(bnd.begin(), bnd.end(), [](const Seen& a, const Seen& b) { return a.seq < b.seq; });
  std::stable_sort(bnd.begin(), bnd.end(), [&](const Seen& a, const Seen& b) {
    return mesh.bf[a.l + nfe * a.e] < mesh.bf[b.l + nfe * b.e];
  });

  Conn c;
  c.nif = static_cast<int>(interior.size());
  c.nf = c.nif + static_cast<int>(bnd.size());
  c.f2e.assign(4 * static_cast<size_t>(c.nf), -1);
  for (int f = 0; f < c.nif; ++f) {
    auto q = interior[f];
    // fix_f2t_consistency: the element with the smaller elempart id comes first
    if (mesh.elempart[q[2]] < mesh.elempart[q[0]]) q = {q[2], q[3], q[0], q[1]};
    for (int k = 0; k < 4; ++k) c.f2e[4 * f + k] = q[k];
  }
  for (size_t i = 0; i < bnd.size(); ++i) {
    int f = c.nif + static_cast<int>(i);
    c.f2e[4 * f + 0] = bnd[i].e;
    c.f2e[4 * f + 1] = bnd[i].l;
  }

  // build_connectivity
  int ncols = 0;
  std::vector<int> permind = permindex(master, ncols);
  c.elemcon.assign(static_cast<size_t>(npf) * nfe * ne, -1);
  c.facecon.assign(static_cast<size_t>(2) * npf * c.nf, -1);
  const int* perm = mesh.perm.data();
  for (int f = 0; f < c.nf; ++f) {
    int e1 = c.f2e[4 * f], l1 = c.f2e[4 * f + 1], e2 = c.f2e[4 * f + 2], l2 = c.f2e[4 * f + 3];
    for (int j = 0; j < npf; ++j) {
      c.facecon[2 * j + 2 * npf * f] = e1 * npe + perm[j + npf * l1];
      c.elemcon[j + npf * (l1 + nfe * e1)] = f * npf + j;
    }
    if (e2 < 0) {
      for (int j = 0; j < npf; ++j) c.facecon[1 + 2 * j + 2 * npf * f] = c.facecon[2 * j + 2 * npf * f];
      continue;
    }
    int k = 0;
    if (nd == 3) {
      int f1[4], f2[4];
      for (int v = 0; v < nvf; ++v) {
        f1[v] = mesh.t[faces[l1][v] + nve * e1];
        f2[v] = mesh.t[faces[l2][v] + nve * e2];
      }
      k = -1;
      if (elemtype == 0) {
        if (f1[0] == f2[0] && f1[1] == f2[2] && f1[2] == f2[1]) k = 0;
        else if (f1[0] == f2[1] && f1[1] == f2[0] && f1[2] == f2[2]) k = 1;
        else if (f1[0] == f2[2] && f1[1] == f2[1] && f1[2] == f2[0]) k = 2;
      } else {
        if (f1[0] == f2[0] && f1[1] == f2[3] && f1[2] == f2[2] && f1[3] == f2[1]) k = 0;
        else if (f1[0] == f2[3] && f1[1] == f2[2] && f1[2] == f2[1] && f1[3] == f2[0]) k = 1;
        else if (f1[0] == f2[2] && f1[1] == f2[1] && f1[2] == f2[0] && f1[3] == f2[3]) k = 2;
        else if (f1[0] == f2[1] && f1[1] == f2[0] && f1[2] == f2[3] && f1[3] == f2[2]) k = 3;
      }
      if (k < 0) throw Error(2, "Mesh connectivity is wrong");
    }
    for (int j = 0; j < npf; ++j) {
      int pj = permind[j + npf * k];
      c.facecon[1 + 2 * j + 2 * npf * f] = e2 * npe + perm[pj + npf * l2];
      c.elemcon[j + npf * (l2 + nfe * e2)] = f * npf + pj;
    }
  }
  for (int v : c.elemcon)
    if (v < 0) throw Error(2, "Mesh connectivity is incomplete");

  // divide_interval(ne, nebmax)
  int nebmax = (nd == 3) ? 1024 : 4096;
  int nblk = (ne + nebmax - 1) / nebmax;
  int base = ne / nblk, rem = ne % nblk, s0 = 0;
  for (int b = 0; b < nblk; ++b) {
    int len = base + (b < rem ? 1 : 0);
    c.eblks.push_back({s0, s0 + len});
    s0 += len;
  }
  return c;
}

}  // namespace exa
