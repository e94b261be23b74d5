// common.hpp -- shared host-side definitions of libexahip.so.
//
// Arrays follow the reference's conventions (column-major, first index fastest,
// int32 indices, float64 data).  Errors are C++ exceptions inside the library
// and are turned into EXA_ERR_* codes at the C ABI (api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace exa {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

#define EXA_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::exa::Error(3, std::string("HIP error '") + hipGetErrorString(e_) +      \
                                 "' at " __FILE__ ":" + std::to_string(__LINE__));     \
  } while (0)

// Device buffer owning its allocation.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    EXA_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  template <class Stager>
  void upload(const std::vector<T>& h, Stager& st, hipStream_t s) {
    alloc(h.size());
    if (n) st.h2d(p, h.data(), n * sizeof(T), s);
  }
  void zero(hipStream_t s) {
    if (n) EXA_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Host <-> device copies go through one pinned staging buffer in fixed-size
// chunks and have completed when the call returns: caller memory (pageable,
// possibly a temporary) is never handed to an asynchronous copy.
struct Stager {
  char* buf = nullptr;
  size_t cap = 0;
  Stager() = default;
  Stager(const Stager&) = delete;
  Stager& operator=(const Stager&) = delete;
  ~Stager() {
    if (buf) (void)hipHostFree(buf);
  }
  void init(size_t bytes) {
    if (buf) return;
    EXA_HIP(hipHostMalloc(reinterpret_cast<void**>(&buf), bytes, hipHostMallocDefault));
    cap = bytes;
  }
  void h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    for (size_t off = 0; off < bytes; off += cap) {
      size_t c = std::min(cap, bytes - off);
      std::memcpy(buf, static_cast<const char*>(src) + off, c);
      EXA_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, buf, c, hipMemcpyHostToDevice, s));
      EXA_HIP(hipStreamSynchronize(s));
    }
  }
  void d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
    for (size_t off = 0; off < bytes; off += cap) {
      size_t c = std::min(cap, bytes - off);
      EXA_HIP(hipMemcpyAsync(buf, static_cast<const char*>(src) + off, c, hipMemcpyDeviceToHost, s));
      EXA_HIP(hipStreamSynchronize(s));
      std::memcpy(static_cast<char*>(dst) + off, buf, c);
    }
  }
};

// ---- datain content (readbinaryfiles.cpp) ----
struct AppData {
  std::vector<double> ndims, flag, problem, uinf, dt, factor, physicsparam, solversparam, tau;
};

struct MasterData {
  int nd = 0, elemtype = 0, nodetype = 0, porder = 0, pgauss = 0;
  int npe = 0, npf = 0, nge = 0, ngf = 0;
  std::vector<double> shapegt, shapegw, shapfgt, shapfgw, xpe, xpf, gwe, gwf;
};

struct MeshData {
  int nd = 0, ne = 0, nfe = 0, nve = 0;
  std::vector<int> elempart, perm, bf, cartgridpart, t, bcs;
};

struct SolData {
  std::vector<double> xdg, udg, odg, wdg, uh;
  int udg_ncomp = 0, nco = 0;
};

// ---- connectivity (connectivity.cpp buildConn) ----
struct Conn {
  int nf = 0, nif = 0;
  std::vector<int> f2e;      // [4, nf]: e1, l1, e2, l2
  std::vector<int> elemcon;  // [npf, nfe, ne]: global face-node index f*npf + j
  std::vector<int> facecon;  // [2, npf, nf]
  std::vector<std::pair<int, int>> eblks;  // [start, end) element blocks
};

AppData read_app(const std::string& path);
MasterData read_master(const std::string& path);
MeshData read_mesh(const std::string& path);
SolData read_sol(const std::string& path, int npe, int nc, int ncu, int ne);
Conn build_conn(const MasterData& master, const MeshData& mesh);
void write_output(const std::string& path, int a0, int a1, int a2, const std::vector<double>& data);

}  // namespace exa
