// kernels.hip -- HIP kernels of the HDG Newton-GMRES path for gfx950 (MI355X).
//
// Layouts (column-major, first index fastest, as in the reference):
//   master tables  shapegt[nge, npe, nd+1], shapegw[npe, nge, nd+1],
//                  shapfgt[ngf, npf, nd],   shapfgw[npf, ngf, nd]
//   udg[npe, nc, ne], uh[ncu, npf, nf], perm[npf, nfe], elemcon[npf, nfe, ne]
//   element geometry egeo, face geometry fgeo: structure of arrays over the
//   quadrature points (field-major), so consecutive threads touch consecutive points.
//   Per-element matrices in the condensation layouts of uequation.cpp:559-775:
//   u-DOF r = i + npe*c, trace DOF t = c + ncu*(j + npf*l).
//
// Every reduction has a fixed order and no atomics are used, so results are
// bitwise reproducible from run to run.
#include "kernels.hpp"
#include "models.hpp"
#include "common.hpp"
#include "generated/registry.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <type_traits>

namespace exa {

namespace {

inline unsigned nblocks(long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

// MFMA operand vectors (the batched GEMMs themselves are in gemm.hip)
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// One 16x16x4 step.  FP64: register i of lane l holds D[l/16 + 4i][l%16];
// FP32 (inputs rounded to FP32, FP32 accumulation): D[4(l/16) + i][l%16].
__device__ inline dbl4 mfma_step(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline flt4 mfma_step(float a, float b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Right-looking blocked LU of the augmented systems [D | RHS] (n x ncols, ld n),
// one panel of BLU_NB columns at a time: k_lu_panel factors the panel of every
// matrix in LDS (partial pivoting), swaps the rows of the columns right of the
// panel and solves L11 U12 = A12; the trailing update A22 -= L21 U12 and the
// back substitution's updates run on the MFMA GEMM.
// ---------------------------------------------------------------------------
constexpr int BLU_MIN_N = 64;  // smaller blocks use the single-kernel k_lu_solve

__device__ inline void argmax_wave(double& v, int& i) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(v, off);
    int oi = __shfl_xor(i, off);
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
  }
}

template <typename T, int BLU_NB>
__global__ __launch_bounds__(256) void k_lu_panel(T* Aall, int n, int ncols, long stride, int k0,
                                                  int* info, int* pivout) {
  T* A = Aall + stride * blockIdx.x;
  const int nb = min(BLU_NB, n - k0);
  const int rows = n - k0;
  extern __shared__ double lu_sm[];
  T* P = reinterpret_cast<T*>(lu_sm);  // rows x nb, ld rows
  __shared__ double wv[4];
  __shared__ int wi[4];
  __shared__ int piv[BLU_NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int idx = tid; idx < rows * nb; idx += 256) {
    int r = idx % rows, kk = idx / rows;
    P[idx] = A[k0 + r + static_cast<long>(n) * (k0 + kk)];
  }
  __syncthreads();
  int singular = 0;
  for (int kk = 0; kk < nb; ++kk) {
    double best = -1.0;
    int bi = kk;
    for (int r = kk + tid; r < rows; r += 256) {
      double v = fabs(P[r + rows * kk]);
      if (v > best) { best = v; bi = r; }
    }
    argmax_wave(best, bi);
    if (lane == 0) { wv[wave] = best; wi[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      double b = wv[0];
      int p = wi[0];
      for (int w = 1; w < 4; ++w)
        if (wv[w] > b || (wv[w] == b && wi[w] < p)) { b = wv[w]; p = wi[w]; }
      piv[kk] = p;
    }
    __syncthreads();
    const int p = piv[kk];
    if (p != kk && tid < nb) {
      T t = P[kk + rows * tid];
      P[kk + rows * tid] = P[p + rows * tid];
      P[p + rows * tid] = t;
    }
    __syncthreads();
    const T pv = P[kk + rows * kk];
    if (pv == 0.0) singular = 1;
    const T inv = T(1) / pv;
    for (int r = kk + 1 + tid; r < rows; r += 256) P[r + rows * kk] *= inv;
    __syncthreads();
    const int rr = rows - kk - 1, cc = nb - kk - 1;
    for (int idx = tid; idx < rr * cc; idx += 256) {
      int r = kk + 1 + idx % rr, c = kk + 1 + idx / rr;
      P[r + rows * c] = fma(-P[r + rows * kk], P[kk + rows * c], P[r + rows * c]);
    }
    __syncthreads();
  }
  for (int idx = tid; idx < rows * nb; idx += 256) {
    int r = idx % rows, kk = idx / rows;
    A[k0 + r + static_cast<long>(n) * (k0 + kk)] = P[idx];
  }
  // columns right of the panel: row swaps, then U12 = L11^{-1} A12
  for (int j = k0 + nb + tid; j < ncols; j += 256) {
    T* colj = A + static_cast<long>(n) * j + k0;
    T x[BLU_NB];
#pragma unroll
    for (int kk = 0; kk < BLU_NB; ++kk) x[kk] = (kk < nb) ? colj[kk] : 0.0;
    // pivots are applied in order; p < BLU_NB means both rows are in the register
    // block (a last panel narrower than BLU_NB has rows == nb)
#pragma unroll
    for (int kk = 0; kk < BLU_NB; ++kk) {
      if (kk < nb) {
        const int p = piv[kk];
        if (p != kk) {
          if (p < BLU_NB) {
#pragma unroll
            for (int q = kk + 1; q < BLU_NB; ++q)
              if (q == p) {
                const T t = x[q];
                x[q] = x[kk];
                x[kk] = t;
              }
          } else {
            const T t = colj[p];
            colj[p] = x[kk];
            x[kk] = t;
          }
        }
      }
    }
#pragma unroll
    for (int kk = 1; kk < BLU_NB; ++kk)
#pragma unroll
      for (int t = 0; t < kk; ++t)
        if (kk < nb) x[kk] = fma(-P[kk + rows * t], x[t], x[kk]);
#pragma unroll
    for (int kk = 0; kk < BLU_NB; ++kk)
      if (kk < nb) colj[kk] = x[kk];
  }
  if (tid == 0 && info) info[blockIdx.x] = (k0 == 0 ? 0 : info[blockIdx.x]) | singular;
  // panel-relative pivots of this panel (row k0+kk was exchanged with row k0+piv[kk])
  if (pivout && tid < nb) pivout[static_cast<long>(blockIdx.x) * n + k0 + tid] = piv[tid];
}

// The panel factorisation of k_lu_panel for panels of at most 256 rows, with the panel in
// registers: thread t holds row t of the rows x nb panel.  Per column: wave argmax of |a[kk]|
// (ties to the smaller row, as k_lu_panel) and a 4-entry combine, the pivot row and row kk
// exchanged through a double-buffered LDS row, then each lower row scales its multiplier and
// updates its own entries -- two barriers per column and no LDS traffic for the rank-1 update.
// The operations on every entry are those of k_lu_panel, so the factors are the same.  The
// columns right of the panel are left to k_lu_u12; the pivots go to piv_ws[blockIdx.x][nb]
// (and to pivout as panel-relative rows).
template <typename T, int BLU_NB>
__global__ __launch_bounds__(256) void k_lu_panel_reg(T* Aall, int n, long stride, int k0, int* info,
                                                      int* pivout, int* piv_ws) {
  T* A = Aall + stride * blockIdx.x;
  const int nb = min(BLU_NB, n - k0);
  const int rows = n - k0;  // <= 256
  __shared__ double wv[2][4];
  __shared__ int wi[2][4];
  __shared__ T prow[2][BLU_NB];
  __shared__ T xrow[2][BLU_NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool mine = tid < rows;
  T a[BLU_NB];
#pragma unroll
  for (int c = 0; c < BLU_NB; ++c)
    a[c] = (mine && c < nb) ? A[k0 + tid + static_cast<long>(n) * (k0 + c)] : 0.0;
  int singular = 0;
#pragma clang loop unroll(full)
  for (int kk = 0; kk < BLU_NB; ++kk) {
    if (kk < nb) {
      const int b = kk & 1;
      double best = (mine && tid >= kk) ? fabs(a[kk]) : -1.0;
      int bi = (mine && tid >= kk) ? tid : kk;
      argmax_wave(best, bi);
      if (lane == 0) {
        wv[b][wave] = best;
        wi[b][wave] = bi;
      }
      __syncthreads();
      double bv = wv[b][0];
      int p = wi[b][0];
      for (int w = 1; w < 4; ++w)
        if (wv[b][w] > bv || (wv[b][w] == bv && wi[b][w] < p)) { bv = wv[b][w]; p = wi[b][w]; }
      if (tid == 0) {
        piv_ws[static_cast<long>(blockIdx.x) * LU_PIV_WS + kk] = p;
        if (pivout) pivout[static_cast<long>(blockIdx.x) * n + k0 + kk] = p;
      }
      // the pivot row and row kk trade places
      if (tid == p)
#pragma unroll
        for (int c = 0; c < BLU_NB; ++c) prow[b][c] = a[c];
      if (tid == kk && p != kk)
#pragma unroll
        for (int c = 0; c < BLU_NB; ++c) xrow[b][c] = a[c];
      __syncthreads();
      if (tid == kk) {
#pragma unroll
        for (int c = 0; c < BLU_NB; ++c) a[c] = prow[b][c];
      } else if (tid == p) {
#pragma unroll
        for (int c = 0; c < BLU_NB; ++c) a[c] = xrow[b][c];
      }
      const T pv = prow[b][kk];
      if (pv == 0.0) singular = 1;
      if (mine && tid > kk) {
        const T l = a[kk] * (T(1) / pv);
        a[kk] = l;
#pragma unroll
        for (int c = kk + 1; c < BLU_NB; ++c)
          if (c < nb) a[c] = fma(-l, prow[b][c], a[c]);
      }
    }
  }
  if (mine)
#pragma unroll
    for (int c = 0; c < BLU_NB; ++c)
      if (c < nb) A[k0 + tid + static_cast<long>(n) * (k0 + c)] = a[c];
  if (tid == 0 && info) info[blockIdx.x] = (k0 == 0 ? 0 : info[blockIdx.x]) | singular;
}

// The row exchanges of a panel factored by k_lu_panel_reg, applied to the columns right of the
// panel.  Thread 0 turns the panel's swaps (row kk <-> row piv[kk], in order) into their net
// permutation of the rows they touch; every column (one per lane, 64 per workgroup) then reads
// all the moved values into registers before writing any, so a column costs one round of
// independent loads instead of a dependent load/store round trip per exchange.
template <typename T, int BLU_NB>
__global__ __launch_bounds__(64) void k_lu_swap(T* Aall, int n, int ncols, long stride, int k0, const int* piv_ws) {
  T* A = Aall + stride * blockIdx.x;
  const int nb = min(BLU_NB, n - k0);
  __shared__ int rows[2 * BLU_NB];  // touched rows (panel-relative), then the moved ones
  __shared__ int held[2 * BLU_NB];  // the original row whose value rows[i] holds
  __shared__ int dst[2 * BLU_NB], srcr[2 * BLU_NB];
  __shared__ int cnt;
  if (threadIdx.x == 0) {
    int nr = nb;
    for (int kk = 0; kk < nb; ++kk) {
      rows[kk] = kk;
      held[kk] = kk;
    }
    for (int kk = 0; kk < nb; ++kk) {
      const int p = piv_ws[static_cast<long>(blockIdx.x) * LU_PIV_WS + kk];
      if (p == kk) continue;
      int pos = p;
      if (p >= nb) {
        pos = -1;
        for (int i = nb; i < nr; ++i)
          if (rows[i] == p) pos = i;
        if (pos < 0) {
          rows[nr] = p;
          held[nr] = p;
          pos = nr++;
        }
      }
      const int t = held[kk];
      held[kk] = held[pos];
      held[pos] = t;
    }
    int c = 0;
    for (int i = 0; i < nr; ++i)
      if (held[i] != rows[i]) {
        dst[c] = rows[i];
        srcr[c] = held[i];
        ++c;
      }
    cnt = c;
  }
  __syncthreads();
  const int j = k0 + nb + blockIdx.y * 64 + threadIdx.x;
  if (j >= ncols) return;
  T* colj = A + static_cast<long>(n) * j + k0;
  const int nm = cnt;
  T v[2 * BLU_NB];
#pragma unroll
  for (int i = 0; i < 2 * BLU_NB; ++i)
    if (i < nm) v[i] = colj[srcr[i]];
#pragma unroll
  for (int i = 0; i < 2 * BLU_NB; ++i)
    if (i < nm) colj[dst[i]] = v[i];
}

// Columns right of a panel factored by k_lu_panel_reg, after k_lu_swap: U12 = L11^{-1} A12 (one
// column per thread; blockIdx.y selects a chunk of 256 columns).  Each thread reads its column's
// nb contiguous entries; a variant with one wave per 64 columns staged through LDS with coalesced
// loads was measured slower (0.47 against 0.33 ms per launch, averaged over the FP64 LUs of the
// 32^3, p = 4 mesh's set-up and assembly; k_lu_back 0.31 against 0.26 ms): the LDS tile cut
// occupancy.
template <typename T, int BLU_NB>
__global__ __launch_bounds__(256) void k_lu_u12(T* Aall, int n, int ncols, long stride, int k0,
                                                const int* piv_ws) {
  T* A = Aall + stride * blockIdx.x;
  const int nb = min(BLU_NB, n - k0);
  // L11 with a run-time leading dimension: constant LDS addresses in the unrolled solve let the
  // compiler hoist every load into registers (and spill)
  __shared__ T L11[BLU_NB * BLU_NB];
  const int ldl = nb;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int r = idx % nb, c = idx / nb;
    L11[r + ldl * c] = A[k0 + r + static_cast<long>(n) * (k0 + c)];
  }
  __syncthreads();
  const int j = k0 + nb + blockIdx.y * 256 + tid;
  if (j >= ncols) return;
  T* colj = A + static_cast<long>(n) * j + k0;
  T x[BLU_NB];
#pragma unroll
  for (int kk = 0; kk < BLU_NB; ++kk) x[kk] = (kk < nb) ? colj[kk] : 0.0;
#pragma unroll
  for (int kk = 1; kk < BLU_NB; ++kk)
#pragma unroll
    for (int t = 0; t < kk; ++t)
      if (kk < nb) x[kk] = fma(-L11[kk + ldl * t], x[t], x[kk]);
#pragma unroll
  for (int kk = 0; kk < BLU_NB; ++kk)
    if (kk < nb) colj[kk] = x[kk];
}

// Inverse of the unit lower triangle L11 (LOWER) or of the upper triangle U11 of the panel at
// k0, one workgroup per matrix: thread c computes column c by substitution (static register
// indices, predicated on c), written to W[blockIdx.x] (nb x nb, ld BLU_NB).  U12 = L11^{-1} A12
// and X = U11^{-1} X then run as in-place GEMMs on the matrix cores: each output tile covers all
// nb rows of its own columns, so it reads its B columns completely before writing them.
template <typename T, int BLU_NB, bool LOWER>
__global__ __launch_bounds__(64) void k_lu_tri_inv(const T* Aall, int n, long stride, int k0, T* W) {
  const T* A = Aall + stride * blockIdx.x;
  const int nb = min(BLU_NB, n - k0);
  __shared__ T Ts[BLU_NB][BLU_NB + 1];
  for (int idx = threadIdx.x; idx < nb * nb; idx += 64) {
    const int r = idx % nb, c = idx / nb;
    Ts[r][c] = A[k0 + r + static_cast<long>(n) * (k0 + c)];
  }
  __syncthreads();
  const int c = threadIdx.x;
  if (c >= nb) return;
  T y[BLU_NB];
  if (LOWER) {
#pragma unroll
    for (int r = 0; r < BLU_NB; ++r) {
      if (r <= c || r >= nb) {
        y[r] = (r == c) ? T(1) : T(0);
      } else {
        T s = T(0);
#pragma unroll
        for (int t = 0; t < r; ++t)
          if (t >= c) s = fma(Ts[r][t], y[t], s);
        y[r] = -s;
      }
    }
  } else {
#pragma unroll
    for (int r = BLU_NB - 1; r >= 0; --r) {
      if (r > c || r >= nb) {
        y[r] = T(0);
      } else if (r == c) {
        y[r] = T(1) / Ts[r][r];
      } else {
        T s = T(0);
#pragma unroll
        for (int t = r + 1; t < BLU_NB; ++t)
          if (t <= c) s = fma(Ts[r][t], y[t], s);
        y[r] = -s / Ts[r][r];
      }
    }
  }
  T* Wm = W + static_cast<long>(blockIdx.x) * BLU_NB * BLU_NB;
#pragma unroll
  for (int r = 0; r < BLU_NB; ++r)
    if (r < nb) Wm[r + BLU_NB * c] = y[r];
}

// X[k0:k0+nb, c] = U11^{-1} X[k0:k0+nb, c] for the right-hand-side columns c >= n
template <typename T, int BLU_NB>
__global__ __launch_bounds__(256) void k_lu_back(T* Aall, int n, int ncols, long stride, int k0) {
  T* A = Aall + stride * blockIdx.x;
  const int nb = min(BLU_NB, n - k0);
  __shared__ T U[BLU_NB][BLU_NB + 1];
  __shared__ T dinv[BLU_NB];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    int r = idx % nb, c = idx / nb;
    U[r][c] = A[k0 + r + static_cast<long>(n) * (k0 + c)];
  }
  __syncthreads();
  if (tid < nb) dinv[tid] = T(1) / U[tid][tid];
  __syncthreads();
  // one right-hand-side column per thread (blockIdx.y selects a chunk of 256)
  const int j = n + blockIdx.y * 256 + tid;
  if (j >= ncols) return;
  T* colj = A + static_cast<long>(n) * j + k0;
  T x[BLU_NB];
#pragma unroll
  for (int kk = 0; kk < BLU_NB; ++kk) x[kk] = (kk < nb) ? colj[kk] : 0.0;
#pragma unroll
  for (int kk = BLU_NB - 1; kk >= 0; --kk) {
    if (kk < nb) {
      x[kk] *= dinv[kk];
#pragma unroll
      for (int t = 0; t < kk; ++t) x[t] = fma(-U[t][kk], x[kk], x[t]);
    }
  }
#pragma unroll
  for (int kk = 0; kk < BLU_NB; ++kk)
    if (kk < nb) colj[kk] = x[kk];
}

// ---------------------------------------------------------------------------
// Blocked LU with partial pivoting + back substitution, one workgroup per matrix.
// ---------------------------------------------------------------------------
constexpr int LU_NB = 16;
constexpr int LU_CH = 64;

__global__ __launch_bounds__(256) void k_lu_solve(double* Aall, int n, int ncols, long stride,
                                                  int* info) {
  double* A = Aall + stride * blockIdx.x;
  extern __shared__ double P[];  // n x LU_NB panel (ld n)
  __shared__ double Us[LU_NB][LU_CH + 1];
  __shared__ double redv[256];
  __shared__ int redi[256];
  __shared__ int piv[LU_NB];
  const int tid = threadIdx.x;
  int singular = 0;
  for (int k0 = 0; k0 < n; k0 += LU_NB) {
    const int nb = min(LU_NB, n - k0);
    const int rows = n - k0;
    for (int idx = tid; idx < rows * nb; idx += 256) {
      int r = k0 + idx % rows, kk = idx / rows;
      P[r + n * kk] = A[r + static_cast<long>(n) * (k0 + kk)];
    }
    __syncthreads();
    for (int kk = 0; kk < nb; ++kk) {
      const int col = k0 + kk;
      double best = -1.0;
      int bi = col;
      for (int r = col + tid; r < n; r += 256) {
        double v = fabs(P[r + n * kk]);
        if (v > best) { best = v; bi = r; }
      }
      redv[tid] = best;
      redi[tid] = bi;
      __syncthreads();
      for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
          double o = redv[tid + s];
          int oi = redi[tid + s];
          if (o > redv[tid] || (o == redv[tid] && oi < redi[tid])) { redv[tid] = o; redi[tid] = oi; }
        }
        __syncthreads();
      }
      const int p = redi[0];
      if (tid == 0) piv[kk] = p;
      if (p != col && tid < nb) {
        double t = P[col + n * tid];
        P[col + n * tid] = P[p + n * tid];
        P[p + n * tid] = t;
      }
      __syncthreads();
      const double pv = P[col + n * kk];
      if (pv == 0.0) singular = 1;
      for (int r = col + 1 + tid; r < n; r += 256) P[r + n * kk] /= pv;
      __syncthreads();
      const int rr = n - col - 1, cc = nb - kk - 1;
      for (int idx = tid; idx < rr * cc; idx += 256) {
        int r = col + 1 + idx % rr, c = kk + 1 + idx / rr;
        P[r + n * c] -= P[r + n * kk] * P[col + n * c];
      }
      __syncthreads();
    }
    for (int idx = tid; idx < rows * nb; idx += 256) {
      int r = k0 + idx % rows, kk = idx / rows;
      A[r + static_cast<long>(n) * (k0 + kk)] = P[r + n * kk];
    }
    for (int j = tid; j < ncols; j += 256) {
      if (j >= k0 && j < k0 + nb) continue;
      double* colj = A + static_cast<long>(n) * j;
      for (int kk = 0; kk < nb; ++kk) {
        int p = piv[kk], r = k0 + kk;
        if (p != r) { double t = colj[r]; colj[r] = colj[p]; colj[p] = t; }
      }
    }
    __syncthreads();
    for (int j = k0 + nb + tid; j < ncols; j += 256) {
      double* colj = A + static_cast<long>(n) * j;
      double x[LU_NB];
#pragma unroll
      for (int kk = 0; kk < LU_NB; ++kk) x[kk] = (kk < nb) ? colj[k0 + kk] : 0.0;
#pragma unroll
      for (int kk = 1; kk < LU_NB; ++kk)
#pragma unroll
        for (int t = 0; t < kk; ++t)
          if (kk < nb) x[kk] -= P[k0 + kk + n * t] * x[t];
#pragma unroll
      for (int kk = 0; kk < LU_NB; ++kk)
        if (kk < nb) colj[k0 + kk] = x[kk];
    }
    __syncthreads();
    const int R = n - k0 - nb;
    if (R > 0) {
      for (int j0 = k0 + nb; j0 < ncols; j0 += LU_CH) {
        const int cw = min(LU_CH, ncols - j0);
        for (int idx = tid; idx < nb * cw; idx += 256) {
          int t = idx % nb, c = idx / nb;
          Us[t][c] = A[k0 + t + static_cast<long>(n) * (j0 + c)];
        }
        __syncthreads();
        for (int idx = tid; idx < R * cw; idx += 256) {
          int r = k0 + nb + idx % R, c = idx / R;
          double* a = A + r + static_cast<long>(n) * (j0 + c);
          double s = *a;
          for (int t = 0; t < nb; ++t) s -= P[r + n * t] * Us[t][c];
          *a = s;
        }
        __syncthreads();
      }
    }
  }
  // back substitution U X = Y for the right-hand-side columns
  const int last = ((n - 1) / LU_NB) * LU_NB;
  for (int k0 = last; k0 >= 0; k0 -= LU_NB) {
    const int nb = min(LU_NB, n - k0);
    const int rows = k0 + nb;
    for (int idx = tid; idx < rows * nb; idx += 256) {
      int r = idx % rows, kk = idx / rows;
      P[r + n * kk] = A[r + static_cast<long>(n) * (k0 + kk)];
    }
    __syncthreads();
    for (int j = n + tid; j < ncols; j += 256) {
      double* colj = A + static_cast<long>(n) * j;
      double x[LU_NB];
#pragma unroll
      for (int kk = 0; kk < LU_NB; ++kk) x[kk] = (kk < nb) ? colj[k0 + kk] : 0.0;
#pragma unroll
      for (int kk = LU_NB - 1; kk >= 0; --kk) {
        if (kk < nb) {
          double s = x[kk];
#pragma unroll
          for (int t = kk + 1; t < LU_NB; ++t)
            if (t < nb) s -= P[k0 + kk + n * t] * x[t];
          x[kk] = s / P[k0 + kk + n * kk];
        }
      }
#pragma unroll
      for (int kk = 0; kk < LU_NB; ++kk)
        if (kk < nb) colj[k0 + kk] = x[kk];
    }
    __syncthreads();
    if (k0 > 0) {
      for (int j0 = n; j0 < ncols; j0 += LU_CH) {
        const int cw = min(LU_CH, ncols - j0);
        for (int idx = tid; idx < nb * cw; idx += 256) {
          int t = idx % nb, c = idx / nb;
          Us[t][c] = A[k0 + t + static_cast<long>(n) * (j0 + c)];
        }
        __syncthreads();
        for (int idx = tid; idx < k0 * cw; idx += 256) {
          int r = idx % k0, c = idx / k0;
          double* a = A + r + static_cast<long>(n) * (j0 + c);
          double s = *a;
          for (int t = 0; t < nb; ++t) s -= P[r + n * t] * Us[t][c];
          *a = s;
        }
        __syncthreads();
      }
    }
  }
  if (tid == 0 && info) info[blockIdx.x] = singular;
}

// ---------------------------------------------------------------------------
// Geometry and setup
// ---------------------------------------------------------------------------
__device__ inline void adjugate(int nd, const double J[3][3], double* jac, double Xx[9]) {
  if (nd == 2) {
    *jac = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    Xx[0 + 2 * 0] = J[1][1];
    Xx[1 + 2 * 0] = -J[1][0];
    Xx[0 + 2 * 1] = -J[0][1];
    Xx[1 + 2 * 1] = J[0][0];
    return;
  }
  for (int j = 0; j < 3; ++j)
    for (int m = 0; m < 3; ++m)
      Xx[j + 3 * m] = J[(m + 1) % 3][(j + 1) % 3] * J[(m + 2) % 3][(j + 2) % 3] -
                      J[(m + 1) % 3][(j + 2) % 3] * J[(m + 2) % 3][(j + 1) % 3];
  *jac = J[0][0] * Xx[0 + 3 * 0] + J[0][1] * Xx[1 + 3 * 0] + J[0][2] * Xx[2 + 3 * 0];
}

// ElemGeomBlock (geometry.cpp:44-85, kokkosimpl.h:2329-2360): egeo fields
// xg[d] (d < nd), Xx[j + nd*m] at nd + j + nd*m, jac at nd + nd*nd.
__global__ void k_elem_geom(Dims d, const double* shapegt, const double* xdg, double* egeo) {
  const long P = static_cast<long>(d.ne) * d.nge;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const int e = static_cast<int>(p / d.nge), g = static_cast<int>(p % d.nge);
  const int nd = d.nd, npe = d.npe, nge = d.nge;
  double xg[3] = {0, 0, 0}, J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int i = 0; i < npe; ++i) {
    const double phi = shapegt[g + nge * i];
    for (int b = 0; b < nd; ++b) {
      const double x = xdg[i + npe * (b + nd * static_cast<long>(e))];
      xg[b] += phi * x;
      for (int a = 0; a < nd; ++a) J[a][b] += shapegt[g + nge * (i + npe * (a + 1))] * x;
    }
  }
  double jac, Xx[9];
  adjugate(nd, J, &jac, Xx);
  for (int b = 0; b < nd; ++b) egeo[b * P + p] = xg[b];
  for (int k = 0; k < nd * nd; ++k) egeo[(nd + k) * P + p] = Xx[k];
  egeo[(nd + nd * nd) * P + p] = jac;
}

// ElemFaceGeomBlock (geometry.cpp:168-205, kokkosimpl.h:2386-2416): fgeo fields
// xg[d], nl[d] at nd + d, jac at 2*nd, at point p = (e*nfe + l)*ngf + g.
__global__ void k_face_geom(Dims d, const double* shapfgt, const int* perm, const double* xdg,
                            double* fgeo) {
  const long P = static_cast<long>(d.ne) * d.nfe * d.ngf;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const int g = static_cast<int>(p % d.ngf);
  const long el = p / d.ngf;
  const int l = static_cast<int>(el % d.nfe);
  const long e = el / d.nfe;
  const int nd = d.nd, npe = d.npe, npf = d.npf, ngf = d.ngf;
  double xg[3] = {0, 0, 0}, T0[3] = {0, 0, 0}, T1[3] = {0, 0, 0};
  for (int j = 0; j < npf; ++j) {
    const int node = perm[j + npf * l];
    const double phi = shapfgt[g + ngf * j];
    const double d0 = shapfgt[g + ngf * (j + npf * 1)];
    const double d1 = (nd == 3) ? shapfgt[g + ngf * (j + npf * 2)] : 0.0;
    for (int b = 0; b < nd; ++b) {
      const double x = xdg[node + npe * (b + nd * e)];
      xg[b] += phi * x;
      T0[b] += d0 * x;
      T1[b] += d1 * x;
    }
  }
  double nl[3], jac;
  if (nd == 2) {
    jac = sqrt(T0[0] * T0[0] + T0[1] * T0[1]);
    nl[0] = T0[1] / jac;
    nl[1] = -T0[0] / jac;
  } else {
    nl[0] = T0[1] * T1[2] - T0[2] * T1[1];
    nl[1] = T0[2] * T1[0] - T0[0] * T1[2];
    nl[2] = T0[0] * T1[1] - T0[1] * T1[0];
    jac = sqrt(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]);
    nl[0] /= jac;
    nl[1] /= jac;
    nl[2] /= jac;
  }
  for (int b = 0; b < nd; ++b) {
    fgeo[b * P + p] = xg[b];
    fgeo[(nd + b) * P + p] = nl[b];
  }
  fgeo[(2 * nd) * P + p] = jac;
}

__global__ void k_interp_elem(Dims d, int ncomp, const double* shapegt, const double* src, double* dst) {
  const long P = static_cast<long>(d.ne) * d.nge;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const long e = p / d.nge;
  const int g = static_cast<int>(p % d.nge);
  for (int c = 0; c < ncomp; ++c) {
    double s = 0.0;
    for (int i = 0; i < d.npe; ++i) s += shapegt[g + d.nge * i] * src[i + d.npe * (c + ncomp * e)];
    dst[c * P + p] = s;
  }
}

__global__ void k_interp_face(Dims d, int ncomp, const double* shapfgt, const int* perm,
                              const double* src, double* dst) {
  const long P = static_cast<long>(d.ne) * d.nfe * d.ngf;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const int g = static_cast<int>(p % d.ngf);
  const long el = p / d.ngf;
  const int l = static_cast<int>(el % d.nfe);
  const long e = el / d.nfe;
  for (int c = 0; c < ncomp; ++c) {
    double s = 0.0;
    for (int j = 0; j < d.npf; ++j)
      s += shapfgt[g + d.ngf * j] * src[perm[j + d.npf * l] + d.npe * (c + ncomp * e)];
    dst[c * P + p] = s;
  }
}

// qEquationElem (qequation.cpp:66-165): Mass[r, j] = sum w phi_r phi_j jac and
// C_d[r, i] = sum_g sum_m w dphi_r/dxi_m phi_i Xx[d, m], written into the
// augmented matrix [Mass | C_1 .. C_nd | E_1 .. E_nd] (ld npe).
__global__ void k_qsetup_mc(Dims d, int e0, int eb, const double* shapegt, const double* shapegw,
                            const double* egeo, double* aug, double* Mass) {
  const int npe = d.npe, nge = d.nge, nd = d.nd;
  const long ncq = static_cast<long>(npe) + npe * nd + static_cast<long>(d.ndf) * nd;
  const long per = static_cast<long>(npe) * (npe + npe * nd);
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * eb) return;
  const int e = static_cast<int>(q / per);
  const long rc = q % per;
  const int r = static_cast<int>(rc % npe), col = static_cast<int>(rc / npe);
  const long P = static_cast<long>(d.ne) * nge;
  const long pe = static_cast<long>(e0 + e) * nge;
  double s = 0.0;
  if (col < npe) {
    for (int g = 0; g < nge; ++g)
      s += shapegw[r + npe * g] * shapegt[g + nge * col] * egeo[(nd + nd * nd) * P + pe + g];
  } else {
    const int cc = col - npe, dd = cc / npe, i = cc % npe;
    for (int g = 0; g < nge; ++g) {
      double t = 0.0;
      for (int m = 0; m < nd; ++m)
        t += shapegw[r + npe * (g + nge * (m + 1))] * egeo[(nd + dd + nd * m) * P + pe + g];
      s += t * shapegt[g + nge * i];
    }
  }
  aug[e * (npe * ncq) + r + static_cast<long>(npe) * col] = s;
  // the mass matrix is kept (it is factored in place in aug)
  if (col < npe) Mass[static_cast<long>(e0 + e) * npe * npe + r + static_cast<long>(npe) * col] = s;
}

// Left factors of the volume q-equation matrices (qEquationElem, qequation.cpp:66-160), per
// element e of the block: Aq_e[b] (npe x nge, b = 0..nd) with
//   Aq_e[0][r, g]      = phiw_0[r, g] jac_g                      (Mass = Aq_e[0] Phi^T)
//   Aq_e[1 + d][r, g]  = sum_m phiw_{m+1}[r, g] Xx[d, m]_g      (C_d  = Aq_e[1 + d] Phi^T)
// so that the products are one batched GEMM against the shared shape table Phi^T.
__global__ void k_qsetup_a(Dims d, int e0, int eb, const double* shapegw, const double* egeo, double* Aq) {
  const int npe = d.npe, nge = d.nge, nd = d.nd;
  const long per = static_cast<long>(npe) * nge;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * eb) return;
  const long e = q / per;
  const long rg = q - e * per;
  const int g = static_cast<int>(rg / npe);
  const long P = static_cast<long>(d.ne) * nge;
  const long pe = (e0 + e) * nge + g;
  double* A = Aq + e * per * (nd + 1) + rg;
  A[0] = shapegw[rg] * egeo[(nd + nd * nd) * P + pe];
  for (int dd = 0; dd < nd; ++dd) {
    double t = 0.0;
    for (int m = 0; m < nd; ++m) t += shapegw[rg + per * (m + 1)] * egeo[(nd + dd + nd * m) * P + pe];
    A[per * (1 + dd)] = t;
  }
}

// Mass of elements e0..e0+eb-1 from the first npe columns of their augmented q-equation systems
__global__ void k_qmass(Dims d, int e0, int eb, const double* aug, double* Mass) {
  const int npe = d.npe;
  const long per = static_cast<long>(npe) * npe;
  const long ncq = static_cast<long>(npe) + npe * d.nd + static_cast<long>(d.ndf) * d.nd;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * eb) return;
  const long e = q / per;
  const long rc = q - e * per;
  Mass[(e0 + e) * per + rc] = aug[e * npe * ncq + rc];
}

// qEquationElemFaceBlock (qequation.cpp:258-340): E_d[perm[k,l], i + npf*l] = sum w phi_k phi_i n_d jac.
__global__ void k_qsetup_e(Dims d, int e0, int eb, const double* shapfgt, const double* shapfgw,
                           const int* perm, const double* fgeo, double* aug) {
  const int npe = d.npe, npf = d.npf, ngf = d.ngf, nd = d.nd, nfe = d.nfe, ndf = d.ndf;
  const long ncq = static_cast<long>(npe) + npe * nd + static_cast<long>(ndf) * nd;
  const long per = static_cast<long>(nfe) * npf * npf * nd;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * eb) return;
  const int e = static_cast<int>(q / per);
  long r = q % per;
  const int k = static_cast<int>(r % npf); r /= npf;
  const int i = static_cast<int>(r % npf); r /= npf;
  const int l = static_cast<int>(r % nfe);
  const int dd = static_cast<int>(r / nfe);
  const long P = static_cast<long>(d.ne) * nfe * ngf;
  const long pf = (static_cast<long>(e0 + e) * nfe + l) * ngf;
  double s = 0.0;
  for (int g = 0; g < ngf; ++g)
    s += shapfgw[k + npf * g] * shapfgt[g + ngf * i] * fgeo[(nd + dd) * P + pf + g] *
         fgeo[(2 * nd) * P + pf + g];
  const long col = npe + static_cast<long>(npe) * nd + (i + npf * l) + static_cast<long>(ndf) * dd;
  aug[e * (npe * ncq) + perm[k + npf * l] + npe * col] = s;
}

// MinvC stacked MCs[(k + npe*d) + npe*nd*j], MinvE stacked MEs[(k + npe*d) + npe*nd*t].
__global__ void k_qpack(Dims d, int e0, int eb, const double* aug, double* MCs, double* MEs) {
  const int npe = d.npe, nd = d.nd, ndf = d.ndf;
  const long ncq = static_cast<long>(npe) + npe * nd + static_cast<long>(ndf) * nd;
  const long per = static_cast<long>(npe) * nd * (npe + ndf);
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * eb) return;
  const int e = static_cast<int>(q / per);
  long r = q % per;
  const int kd = static_cast<int>(r % (npe * nd));
  const int col = static_cast<int>(r / (npe * nd));
  const int k = kd % npe, dd = kd / npe;
  const double* X = aug + e * (npe * ncq);
  const long E = e0 + e;
  if (col < npe) {
    MCs[E * npe * nd * npe + kd + static_cast<long>(npe) * nd * col] =
        X[k + static_cast<long>(npe) * (npe + col + npe * dd)];
  } else {
    const int t = col - npe;
    MEs[E * npe * nd * ndf + kd + static_cast<long>(npe) * nd * t] =
        X[k + static_cast<long>(npe) * (npe + npe * nd + t + static_cast<long>(ndf) * dd)];
  }
}

// hdgGetQ (qequation.cpp:372-487): q_d = MinvC_d u - MinvE_d uh, stored in udg.
// One workgroup per element, one thread per row kd of M^-1 C / M^-1 E, all NCU components: the
// element's u and its gathered uh are staged in LDS, and the two matrices are read once, along
// their rows (the sums for each component run in the same order as per component).
template <int NCU>
__global__ __launch_bounds__(1024) void k_getq(Dims d, const double* MCs, const double* MEs, const int* elemcon,
                                               const double* uh, double* udg) {
  extern __shared__ double gq_sm[];
  const int npe = d.npe, nd = d.nd, nc = d.nc, ndf = d.ndf;
  const int per = npe * nd;
  const long e = blockIdx.x;
  const int tid = threadIdx.x;
  double* us = gq_sm;               // u[j + npe*c]
  double* uhs = gq_sm + npe * NCU;  // uh of the element's trace nodes, [t*NCU + c]
  const double* u = udg + static_cast<long>(npe) * nc * e;
  const int* ec = elemcon + static_cast<long>(ndf) * e;
  for (int idx = tid; idx < npe * NCU; idx += blockDim.x) us[idx] = u[idx];
  for (int idx = tid; idx < ndf * NCU; idx += blockDim.x)
    uhs[idx] = uh[NCU * static_cast<long>(ec[idx / NCU]) + idx % NCU];
  __syncthreads();
  const double* mc = MCs + e * npe * nd * npe;
  const double* me = MEs + e * npe * nd * ndf;
  for (int kd = tid; kd < per; kd += blockDim.x) {
    const int k = kd % npe, dd = kd / npe;
    double s[NCU];
#pragma unroll
    for (int c = 0; c < NCU; ++c) s[c] = 0.0;
#pragma unroll 4
    for (int j = 0; j < npe; ++j) {
      const double mv = mc[kd + static_cast<long>(npe) * nd * j];
#pragma unroll
      for (int c = 0; c < NCU; ++c) s[c] += mv * us[j + npe * c];
    }
#pragma unroll 4
    for (int t = 0; t < ndf; ++t) {
      const double mv = me[kd + static_cast<long>(npe) * nd * t];
#pragma unroll
      for (int c = 0; c < NCU; ++c) s[c] -= mv * uhs[t * NCU + c];
    }
#pragma unroll
    for (int c = 0; c < NCU; ++c) udg[k + npe * (NCU + c + NCU * dd) + static_cast<long>(npe) * nc * e] = s[c];
  }
}

// the same for any ncu, one component per thread
__global__ void k_getq_any(Dims d, const double* MCs, const double* MEs, const int* elemcon,
                           const double* uh, double* udg) {
  const int npe = d.npe, nd = d.nd, ncu = d.ncu, nc = d.nc, ndf = d.ndf;
  const long per = static_cast<long>(npe) * nd * ncu;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * d.ne) return;
  const long e = q / per;
  long r = q % per;
  const int kd = static_cast<int>(r % (npe * nd));
  const int c = static_cast<int>(r / (npe * nd));
  const int k = kd % npe, dd = kd / npe;
  const double* mc = MCs + e * npe * nd * npe;
  const double* me = MEs + e * npe * nd * ndf;
  const double* u = udg + npe * (c + nc * e);
  double s = 0.0;
  for (int j = 0; j < npe; ++j) s += mc[kd + static_cast<long>(npe) * nd * j] * u[j];
  for (int t = 0; t < ndf; ++t)
    s -= me[kd + static_cast<long>(npe) * nd * t] * uh[c + ncu * static_cast<long>(elemcon[t + ndf * e])];
  udg[k + npe * (ncu + c + ncu * dd) + npe * nc * e] = s;
}

// ---------------------------------------------------------------------------
// models with a time-derivative coefficient (Tdfunc) define tdfunc(x, uq, v, p, t, T)
template <class M, class = void>
struct has_tdfunc : std::false_type {};
template <class M>
struct has_tdfunc<M, std::void_t<decltype(&M::tdfunc)>> : std::true_type {};

// Model-library models (MODEL_PLUGIN): the library's functions have run over every point of the
// block before the point kernels, into a.pbuf (plugin_points); the point kernels read their values
// and Jacobians from there.  Staging buffer layout, P points of the block:
//   volume: inputs x [ND], uq [NC], v [nco]; outputs f [NCU ND], f_uq [NCU ND NC], s [NCU], s_uq [NCU NC]
//   faces:  inputs x [ND], uq [NC], uq with u -> uhat [NC], uhat [NCU], n [ND], v [nco];
//           outputs f(uhat, q) [NCU ND], its f_uq [NCU ND NC], then per boundary condition
//           ib = 1..nbc: fb [NCU], fb_uq [NCU NC], fb_uh [NCU NCU]
// each entry k of a group at (group offset + k) * P + p.
template <int ND_, int NCU_>
struct PluginModel {
  static constexpr int ND = ND_, NCU = NCU_, NC = NCU_ * (ND_ + 1);
  static constexpr bool SOA = true;
  __host__ __device__ static long vol_in(int nco) { return ND + NC + nco; }
  __host__ __device__ static long vol_out() { return NCU * ND + NCU * ND * NC + NCU + NCU * NC; }
  __host__ __device__ static long face_in(int nco) { return 2 * ND + 2 * NC + NCU + nco; }
  __host__ __device__ static long face_bc() { return NCU + NCU * NC + NCU * NCU; }
  __host__ __device__ static long face_out(int nbc) { return NCU * ND + NCU * ND * NC + nbc * face_bc(); }
};
template <class M, class = void>
struct is_soa : std::false_type {};
template <class M>
struct is_soa<M, std::void_t<decltype(M::SOA)>> : std::true_type {};

// Models with a w equation (NCW > 0: generated models with a Sourcew function) define flux_w,
// source_w, fbou_w and sourcew.
template <class M, class = void>
struct model_ncw : std::integral_constant<int, 0> {};
template <class M>
struct model_ncw<M, std::void_t<decltype(M::NCW)>> : std::integral_constant<int, M::NCW> {};

// B <- J^{-1} B for an N x N matrix J (column-major, overwritten) and NR right-hand sides
// (B[i + N c]): Gaussian elimination with partial pivoting, in registers.
template <int N, int NR>
__device__ __forceinline__ void small_solve(double* J, double* B) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int pr = k;
#pragma unroll
    for (int i = k + 1; i < N; ++i)
      if (fabs(J[i + N * k]) > fabs(J[pr + N * k])) pr = i;
    if (pr != k) {
#pragma unroll
      for (int j = 0; j < N; ++j) {
        const double t = J[k + N * j];
        J[k + N * j] = J[pr + N * j];
        J[pr + N * j] = t;
      }
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const double t = B[k + N * c];
        B[k + N * c] = B[pr + N * c];
        B[pr + N * c] = t;
      }
    }
    const double inv = 1.0 / J[k + N * k];
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      const double l = J[i + N * k] * inv;
#pragma unroll
      for (int j = k + 1; j < N; ++j) J[i + N * j] -= l * J[k + N * j];
#pragma unroll
      for (int c = 0; c < NR; ++c) B[i + N * c] -= l * B[k + N * c];
    }
  }
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    const double inv = 1.0 / J[k + N * k];
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      double t = B[k + N * c];
#pragma unroll
      for (int j = k + 1; j < N; ++j) t -= J[k + N * j] * B[j + N * c];
      B[k + N * c] = t * inv;
    }
  }
}

// The w equation at one point (wEquation, wequation.cpp:108-224): Newton on
//   R(w) = Sourcew(x, uq, v, w) + alpha wsrc - (alpha dtfactor + beta) w = 0
// from the guess in w, steps dw = J^{-1} R with J = (alpha dtfactor + beta) I - Sourcew_w, until
// max |dw| < 1e-8 (at most 20 steps), then w_uq = J^{-1} Sourcew_uq ([j + NCW c]).  A point that
// does not converge gets NaN, which the residual norm reports.
template <class M>
__device__ void w_point(const AsmArgs& a, const double* x, const double* uq, const double* v, const double* prm,
                        double tm, const double* ws, double* w, double* w_uq) {
  constexpr int NCW = model_ncw<M>::value, NC = M::NC;
  const double scal = a.dae_alpha * a.dtfactor + a.dae_beta;
  double sw[NCW], sw_uq[NCW * NC], sw_w[NCW * NCW], J[NCW * NCW];
  bool conv = false;
  for (int it = 0; it < 20 && !conv; ++it) {
    M::sourcew(x, uq, v, w, prm, tm, sw, sw_uq, sw_w);
    double r[NCW];
#pragma unroll
    for (int i = 0; i < NCW; ++i) {
      r[i] = sw[i] + a.dae_alpha * ws[i] - scal * w[i];
#pragma unroll
      for (int j = 0; j < NCW; ++j) J[i + NCW * j] = (i == j ? scal : 0.0) - sw_w[i + NCW * j];
    }
    small_solve<NCW, 1>(J, r);
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < NCW; ++i) {
      w[i] += r[i];
      mx = fmax(mx, fabs(r[i]));
    }
    conv = mx < 1e-8;
  }
  M::sourcew(x, uq, v, w, prm, tm, sw, sw_uq, sw_w);
#pragma unroll
  for (int i = 0; i < NCW; ++i)
#pragma unroll
    for (int j = 0; j < NCW; ++j) J[i + NCW * j] = (i == j ? scal : 0.0) - sw_w[i + NCW * j];
#pragma unroll
  for (int k = 0; k < NCW * NC; ++k) w_uq[k] = sw_uq[k];
  small_solve<NCW, NC>(J, w_uq);
  if (!conv)
#pragma unroll
    for (int i = 0; i < NCW; ++i) w[i] = __builtin_nan("");
}

// physicsparam and tau of the instance element E belongs to (one instance: a.einst is null)
__device__ __forceinline__ const double* inst_param(const AsmArgs& a, long E) {
  return a.einst ? a.param + static_cast<long>(a.einst[E]) * a.nparam : a.param;
}
__device__ __forceinline__ const double* inst_tau(const AsmArgs& a, long E) {
  return a.einst ? a.tau + static_cast<long>(a.einst[E]) * a.ntau : a.tau;
}
// the model time at element E: the instance's own time when the instances of a batch are at
// different times (a.tinst, one entry per instance), else the common a.time
__device__ __forceinline__ double inst_time(const AsmArgs& a, long E) {
  return a.tinst ? a.tinst[a.einst[E]] : a.time;
}

// Element volume terms: uEquationElemBlock (uequation.cpp:45-203)
// ---------------------------------------------------------------------------
// rv[(dp*NCU + k)*P + p]       : s*jac (dp = 0), sum_j f[k, j] Xx[j, m] (dp = m+1)
// rJ[((dp*NCU + k)*NC + c)*P + p]: the same for the Jacobians (ApplyXxJac, kokkosimpl.h:2105-2158)
template <class M, bool JAC>
__global__ __launch_bounds__(256) void k_vol_points(Dims d, AsmArgs a) {
  constexpr int ND = M::ND, NCU = M::NCU, NC = M::NC;
  const long P = static_cast<long>(a.eb) * d.nge;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const long e = p / d.nge;
  const int g = static_cast<int>(p % d.nge);
  const long E = a.e0 + e;
  const long PG = static_cast<long>(d.ne) * d.nge;
  const long pg = E * d.nge + g;
  double uq[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) uq[c] = 0.0;
  for (int i = 0; i < d.npe; ++i) {
    const double phi = a.shapegt[g + d.nge * i];
#pragma unroll
    for (int c = 0; c < NC; ++c) uq[c] += phi * a.udg[i + d.npe * (c + NC * E)];
  }
  double x[ND], Xx[ND * ND], v[8];
#pragma unroll
  for (int b = 0; b < ND; ++b) x[b] = a.egeo[b * PG + pg];
#pragma unroll
  for (int k = 0; k < ND * ND; ++k) Xx[k] = a.egeo[(ND + k) * PG + pg];
  const double jac = a.egeo[(ND + ND * ND) * PG + pg];
  for (int c = 0; c < d.nco && c < 8; ++c) v[c] = a.og[c * PG + pg];
  double f[NCU * ND], s[NCU];
  double f_uq[JAC ? NCU * ND * NC : 1], s_uq[JAC ? NCU * NC : 1];
  const double* prm = inst_param(a, E);
  const double tm = inst_time(a, E);
  if constexpr (model_ncw<M>::value > 0) {
    // w at the point from its equation, and the chain rule through w_uq (uequation.cpp:105-165)
    constexpr int NCW = model_ncw<M>::value;
    double w[NCW], ws[NCW], w_uq[NCW * NC];
#pragma unroll
    for (int c = 0; c < NCW; ++c) w[c] = ws[c] = 0.0;
    for (int i = 0; i < d.npe; ++i) {
      const double phi = a.shapegt[g + d.nge * i];
#pragma unroll
      for (int c = 0; c < NCW; ++c) {
        w[c] += phi * a.wdg[i + d.npe * (c + NCW * E)];
        ws[c] += phi * a.wsrc[i + d.npe * (c + NCW * E)];
      }
    }
    w_point<M>(a, x, uq, v, prm, tm, ws, w, w_uq);
    double f_w[JAC ? NCU * ND * NCW : 1], s_w[JAC ? NCU * NCW : 1];
    M::template flux_w<JAC>(x, uq, v, w, prm, tm, f, f_uq, f_w);
    M::template source_w<JAC>(x, uq, v, w, prm, tm, s, s_uq, s_w);
    if (JAC) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int i = 0; i < NCU * ND; ++i) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < NCW; ++j) t += f_w[i + NCU * ND * j] * w_uq[j + NCW * c];
          f_uq[i + NCU * ND * c] += t;
        }
#pragma unroll
        for (int i = 0; i < NCU; ++i) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < NCW; ++j) t += s_w[i + NCU * j] * w_uq[j + NCW * c];
          s_uq[i + NCU * c] += t;
        }
      }
    }
  } else if constexpr (is_soa<M>::value) {  // the model library's values at this point
    const double* o = a.pbuf + M::vol_in(d.nco) * P + p;
#pragma unroll
    for (int i = 0; i < NCU * ND; ++i) f[i] = o[i * P];
    o += NCU * ND * P;
    if (JAC)
#pragma unroll
      for (int i = 0; i < NCU * ND * NC; ++i) f_uq[i] = o[i * P];
    o += NCU * ND * NC * P;
#pragma unroll
    for (int i = 0; i < NCU; ++i) s[i] = o[i * P];
    o += NCU * P;
    if (JAC)
#pragma unroll
      for (int i = 0; i < NCU * NC; ++i) s_uq[i] = o[i * P];
  } else {
    M::template flux<JAC>(x, uq, v, prm, tm, f, f_uq);
    M::template source<JAC>(x, uq, v, prm, tm, s, s_uq);
  }
  if (a.tdep) {  // uequation.cpp:126-151: s += (sdg - dtfactor*u) T, s_u -= dtfactor T
    double T[NCU];
#pragma unroll
    for (int k = 0; k < NCU; ++k) T[k] = 1.0;
    if constexpr (has_tdfunc<M>::value)
      if (a.tdfunc) M::tdfunc(x, uq, v, prm, tm, T);
    double sg[NCU];
#pragma unroll
    for (int k = 0; k < NCU; ++k) sg[k] = 0.0;
    for (int i = 0; i < d.npe; ++i) {
      const double phi = a.shapegt[g + d.nge * i];
#pragma unroll
      for (int k = 0; k < NCU; ++k) sg[k] += phi * a.sdg[i + d.npe * (k + NCU * E)];
    }
#pragma unroll
    for (int k = 0; k < NCU; ++k) {
      s[k] += (sg[k] - a.dtfactor * uq[k]) * T[k];
      if (JAC) s_uq[k + NCU * k] -= a.dtfactor * T[k];
    }
  }
#pragma unroll
  for (int k = 0; k < NCU; ++k) {
    a.rv[(0 * NCU + k) * P + p] = s[k] * jac;
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < ND; ++j) t += f[k + NCU * j] * Xx[j + ND * m];
      a.rv[((m + 1) * NCU + k) * P + p] = t;
    }
  }
  if (JAC) {
#pragma unroll
    for (int k = 0; k < NCU; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        a.rJ[((0 * NCU + k) * NC + c) * P + p] = s_uq[k + NCU * c] * jac;
#pragma unroll
        for (int m = 0; m < ND; ++m) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < ND; ++j) t += f_uq[(k + NCU * j) + NCU * ND * c] * Xx[j + ND * m];
          a.rJ[(((m + 1) * NCU + k) * NC + c) * P + p] = t;
        }
      }
  }
}

// Ru[i + npe*k] = sum_dp sum_g shapegw[i, g, dp] * rv  (Gauss2Node with shapegw); one thread per
// (i, k, element), for point sets too large for the LDS of k_vol_res
template <typename T>
__global__ void k_vol_res_direct(Dims d, AsmArgs a, T* out, long ldo) {
  const int npe = d.npe, nge = d.nge, ncu = d.ncu, nd = d.nd;
  const long per = static_cast<long>(npe) * ncu;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * a.eb) return;
  const long e = q / per;
  const int r = static_cast<int>(q % per);
  const int i = r % npe, k = r / npe;
  const long P = static_cast<long>(a.eb) * nge;
  double s = 0.0;
  for (int dp = 0; dp <= nd; ++dp) {
    const double* rv = a.rv + (dp * ncu + k) * P + e * nge;
    const double* w = a.shapegw + static_cast<long>(npe) * nge * dp;
    for (int g = 0; g < nge; ++g) s += w[i + npe * g] * rv[g];
  }
  out[e * ldo + r] = static_cast<T>(s);
}

// One workgroup per element: the element's point values rv (nd+1) ncu nge are staged in LDS once;
// thread (i, k) sums over dp and the points in the same order as before (dp outer, g inner).
template <typename T>
__global__ __launch_bounds__(256) void k_vol_res(Dims d, AsmArgs a, T* out, long ldo) {
  extern __shared__ double vr_sm[];
  const int npe = d.npe, nge = d.nge, ncu = d.ncu, nd = d.nd;
  const int per = npe * ncu;
  const long e = blockIdx.x;
  const int tid = threadIdx.x;
  const long P = static_cast<long>(a.eb) * nge;
  const int L = (nd + 1) * ncu;
  // vr_sm[(dp*ncu + k)*nge + g] = rv[(dp*ncu + k)*P + e*nge + g]
  for (int idx = tid; idx < L * nge; idx += blockDim.x) {
    const int j = idx / nge, g = idx - j * nge;
    vr_sm[idx] = a.rv[j * P + e * nge + g];
  }
  __syncthreads();
  for (int r = tid; r < per; r += blockDim.x) {
    const int i = r % npe, k = r / npe;
    double sum = 0.0;
    for (int dp = 0; dp <= nd; ++dp) {
      const double* rv = vr_sm + (dp * ncu + k) * nge;
      const double* w = a.shapegw + static_cast<long>(npe) * nge * dp;
#pragma unroll 4
      for (int g = 0; g < nge; ++g) sum += w[i + npe * g] * rv[g];
    }
    out[e * ldo + r] = static_cast<T>(sum);
  }
}

// W[(e*NC + c)*NCU + k][i + npe*g] = sum_dp shapegw[i, g, dp] * rJ[dp, k, c]
// One workgroup per (element, gp consecutive points): the points' rJ values (nd+1) ncu nc each are
// staged in LDS once, then one thread per (i, g) forms all ncu*nc (k, c) blocks from LDS (the sum
// over dp runs in the same order for every block); the stores run along i (coalesced).
template <typename T>
__global__ __launch_bounds__(256) void k_vol_W(Dims d, AsmArgs a, T* W, int gp) {
  constexpr int MAXD = 4;  // nd + 1 <= 4
  extern __shared__ double vw_sm[];
  const int npe = d.npe, nge = d.nge, ncu = d.ncu, nc = d.nc, nd = d.nd;
  const int blk = npe * nge, L = (nd + 1) * ncu * nc;
  const int ngrp = (nge + gp - 1) / gp;
  const long e = blockIdx.x / ngrp;
  const int g0 = static_cast<int>(blockIdx.x % ngrp) * gp;
  const int ng = min(gp, nge - g0);
  const long P = static_cast<long>(a.eb) * nge;
  const int tid = threadIdx.x;
  // vw_sm[gl * L + (dp*ncu + k)*nc + c] = rJ[((dp*ncu + k)*nc + c)*P + e*nge + g0 + gl]
  for (int idx = tid; idx < L * ng; idx += 256) {
    const int gl = idx % ng, j = idx / ng;
    vw_sm[gl * L + j] = a.rJ[j * P + e * nge + g0 + gl];
  }
  __syncthreads();
  T* We = W + e * static_cast<long>(nc) * ncu * blk;
  for (int idx = tid; idx < npe * ng; idx += 256) {
    const int i = idx % npe, gl = idx / npe, g = g0 + gl;
    double w[MAXD];
#pragma unroll
    for (int dp = 0; dp < MAXD; ++dp) w[dp] = dp <= nd ? a.shapegw[i + npe * (g + nge * dp)] : 0.0;
    const double* r = vw_sm + gl * L;
    const int ig = i + npe * g;
    for (int kc = 0; kc < ncu * nc; ++kc) {
      const int k = kc % ncu, c = kc / ncu;
      double sum = 0.0;
#pragma unroll
      for (int dp = 0; dp < MAXD; ++dp)
        if (dp <= nd) sum += w[dp] * r[(dp * ncu + k) * nc + c];
      We[static_cast<long>(c * ncu + k) * blk + ig] = static_cast<T>(sum);
    }
  }
}

// ---------------------------------------------------------------------------
// Face terms: uEquationElemFaceBlock (uequation.cpp:206-557) with FhatDriver
// (Model/KokkosDrivers.cpp:874-934) and FbouDriver for boundary faces.
// fpt fields (JAC): fhI[NCU], fhI_uq[NCU*NC], fhI_uh[NCU*NCU], then the same
// with boundary faces replaced by FbouHdg ("B"); residual-only: fhI[NCU], fhB[NCU].
// ---------------------------------------------------------------------------
// Traced state at face point g of local face l of element E: uh, uq = (u, q) from the
// element, x, the normal nl, the face Jacobian jac and the auxiliary field v.
template <class M>
__device__ __forceinline__ void face_state(const Dims& d, const AsmArgs& a, long E, int l, int g,
                                           double* uh, double* uq, double* x, double* nl, double* v,
                                           double& jac) {
  constexpr int ND = M::ND, NCU = M::NCU, NC = M::NC;
  const long PF = static_cast<long>(d.ne) * d.nfe * d.ngf;
  const long pf = (E * d.nfe + l) * d.ngf + g;
#pragma unroll
  for (int c = 0; c < NCU; ++c) uh[c] = 0.0;
#pragma unroll
  for (int c = 0; c < NC; ++c) uq[c] = 0.0;
  for (int j = 0; j < d.npf; ++j) {
    const double phi = a.shapfgt[g + d.ngf * j];
    const long gn = a.elemcon[j + d.npf * (l + d.nfe * E)];
    const int node = a.perm[j + d.npf * l];
#pragma unroll
    for (int c = 0; c < NCU; ++c) uh[c] += phi * a.uh[c + NCU * gn];
#pragma unroll
    for (int c = 0; c < NC; ++c) uq[c] += phi * a.udg[node + d.npe * (c + NC * E)];
  }
#pragma unroll
  for (int b = 0; b < ND; ++b) {
    x[b] = a.fgeo[b * PF + pf];
    nl[b] = a.fgeo[(ND + b) * PF + pf];
  }
  jac = a.fgeo[(2 * ND) * PF + pf];
  for (int c = 0; c < d.nco && c < 8; ++c) v[c] = a.ogf[c * PF + pf];
}

// nodal w and wsrc of element E interpolated to point g of its face l
template <class M>
__device__ __forceinline__ void face_w(const Dims& d, const AsmArgs& a, long E, int l, int g, double* w,
                                       double* ws) {
  constexpr int NCW = model_ncw<M>::value;
#pragma unroll
  for (int c = 0; c < NCW; ++c) w[c] = ws[c] = 0.0;
  for (int j = 0; j < d.npf; ++j) {
    const double phi = a.shapfgt[g + d.ngf * j];
    const int node = a.perm[j + d.npf * l];
#pragma unroll
    for (int c = 0; c < NCW; ++c) {
      w[c] += phi * a.wdg[node + d.npe * (c + NCW * E)];
      ws[c] += phi * a.wsrc[node + d.npe * (c + NCW * E)];
    }
  }
}

template <class M, bool JAC>
__global__ __launch_bounds__(256) void k_face_points(Dims d, AsmArgs a) {
  constexpr int ND = M::ND, NCU = M::NCU, NC = M::NC;
  constexpr int NF1 = JAC ? (NCU + NCU * NC + NCU * NCU) : NCU;
  const long P = static_cast<long>(a.eb) * d.nfe * d.ngf;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const int g = static_cast<int>(p % d.ngf);
  const long el = p / d.ngf;
  const int l = static_cast<int>(el % d.nfe);
  const long E = a.e0 + el / d.nfe;
  double uh[NCU], uq[NC], usave[NCU], x[ND], nl[ND], v[8], jac;
  face_state<M>(d, a, E, l, g, uh, uq, x, nl, v, jac);
  const double tau0 = inst_tau(a, E)[0];
  // interior formula: flux at (uhat, q), dotted with n, plus tau*(u - uhat)
#pragma unroll
  for (int c = 0; c < NCU; ++c) {
    usave[c] = uq[c];
    uq[c] = uh[c];
  }
  double f[NCU * ND], f_uq[JAC ? NCU * ND * NC : 1];
  if constexpr (model_ncw<M>::value > 0) {
    // w at the face point with (uhat, q) in the (u, q) slots (uequation.cpp:262-280): the u
    // columns of w_uq are derivatives in uhat, as are those of f_uq here
    constexpr int NCW = model_ncw<M>::value;
    double w[NCW], ws[NCW], w_uq[NCW * NC];
    face_w<M>(d, a, E, l, g, w, ws);
    const double* prm = inst_param(a, E);
    const double tm = inst_time(a, E);
    w_point<M>(a, x, uq, v, prm, tm, ws, w, w_uq);
    double f_w[JAC ? NCU * ND * NCW : 1];
    M::template flux_w<JAC>(x, uq, v, w, prm, tm, f, f_uq, f_w);
    if (JAC)
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < NCU * ND; ++i) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < NCW; ++j) t += f_w[i + NCU * ND * j] * w_uq[j + NCW * c];
          f_uq[i + NCU * ND * c] += t;
        }
  } else if constexpr (is_soa<M>::value) {  // the model library's f(uhat, q) at this point
    const double* o = a.pbuf + M::face_in(d.nco) * P + p;
#pragma unroll
    for (int i = 0; i < NCU * ND; ++i) f[i] = o[i * P];
    if (JAC)
#pragma unroll
      for (int i = 0; i < NCU * ND * NC; ++i) f_uq[i] = o[(NCU * ND + i) * P];
  } else {
    M::template flux<JAC>(x, uq, v, inst_param(a, E), inst_time(a, E), f, f_uq);
  }
#pragma unroll
  for (int c = 0; c < NCU; ++c) uq[c] = usave[c];
#pragma unroll
  for (int k = 0; k < NCU; ++k) {
    double fh = 0.0;
#pragma unroll
    for (int b = 0; b < ND; ++b) fh += f[k + NCU * b] * nl[b];
    fh += tau0 * (usave[k] - uh[k]);
    a.fpt[k * P + p] = fh * jac;
  }
  if (JAC) {
#pragma unroll
    for (int k = 0; k < NCU; ++k)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < ND; ++b) t += f_uq[(k + NCU * b) + NCU * ND * c] * nl[b];
        if (c < NCU) {
          a.fpt[(NCU + NCU * NC + k + NCU * c) * P + p] = (t - (c == k ? tau0 : 0.0)) * jac;
          a.fpt[(NCU + k + NCU * c) * P + p] = (c == k ? tau0 : 0.0) * jac;
        } else {
          a.fpt[(NCU + k + NCU * c) * P + p] = t * jac;
        }
      }
  }
  // boundary faces get their "B" fields from k_face_bou_points
  if (a.bf[l + d.nfe * E] == 0) {
#pragma unroll
    for (int i = 0; i < NF1; ++i) a.fpt[(NF1 + i) * P + p] = a.fpt[i * P + p];
  }
}

// "B" fields of the boundary face points of the block: FbouHdg (HdgFbou, the
// FbouDriver of Model/KokkosDrivers.cpp) at every point of the faces
// bface[bf0 .. bf0 + nbf) (face ids E*nfe + l, ascending, all inside the block).
// A kernel of its own, so the model's boundary code (often far larger than the
// flux, e.g. characteristic far-field conditions) does not weigh on the interior
// face kernel, with 64-lane blocks so it may use up to 512 VGPRs.
template <class M, bool JAC>
__global__ __launch_bounds__(64) void k_face_bou_points(Dims d, AsmArgs a) {
  constexpr int ND = M::ND, NCU = M::NCU, NC = M::NC;
  constexpr int NF1 = JAC ? (NCU + NCU * NC + NCU * NCU) : NCU;
  const long P = static_cast<long>(a.eb) * d.nfe * d.ngf;
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<long>(a.nbf) * d.ngf) return;
  const int g = static_cast<int>(t % d.ngf);
  const long face = a.bface[a.bf0 + t / d.ngf];
  const long E = face / d.nfe;
  const int l = static_cast<int>(face % d.nfe);
  const long p = ((E - a.e0) * d.nfe + l) * d.ngf + g;
  double uh[NCU], uq[NC], x[ND], nl[ND], v[8], jac;
  face_state<M>(d, a, E, l, g, uh, uq, x, nl, v, jac);
  const int ib = a.bf[l + d.nfe * E];
  double fb[NCU], fb_uq[JAC ? NCU * NC : 1], fb_uh[JAC ? NCU * NCU : 1];
  if constexpr (model_ncw<M>::value > 0) {
    // the face point's w from (uhat, q), as in k_face_points; FbouHdg's derivative through w goes
    // to uhat (u columns of w_uq) and to q
    constexpr int NCW = model_ncw<M>::value;
    double w[NCW], ws[NCW], w_uq[NCW * NC], uqh[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) uqh[c] = c < NCU ? uh[c] : uq[c];
    face_w<M>(d, a, E, l, g, w, ws);
    const double* prm = inst_param(a, E);
    const double tm = inst_time(a, E);
    w_point<M>(a, x, uqh, v, prm, tm, ws, w, w_uq);
    double fb_w[JAC ? NCU * NCW : 1];
    M::template fbou_w<JAC>(ib, x, uq, v, w, uh, nl, inst_tau(a, E), prm, tm, fb, fb_uq, fb_uh, fb_w);
    if (JAC)
#pragma unroll
      for (int r = 0; r < NCU; ++r)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          double t = 0.0;
#pragma unroll
          for (int j = 0; j < NCW; ++j) t += fb_w[r + NCU * j] * w_uq[j + NCW * c];
          if (c < NCU) fb_uh[r + NCU * c] += t;
          else fb_uq[r + NCU * c] += t;
        }
  } else if constexpr (is_soa<M>::value) {  // the model library's condition ib at this point
    const int nbc = a.plugin_nbc;
    const int b = (ib >= 1 && ib <= nbc) ? ib - 1 : 0;
    const double* o = a.pbuf + (M::face_in(d.nco) + NCU * ND + NCU * ND * NC + b * M::face_bc()) * P + p;
#pragma unroll
    for (int i = 0; i < NCU; ++i) fb[i] = o[i * P];
    if (JAC) {
#pragma unroll
      for (int i = 0; i < NCU * NC; ++i) fb_uq[i] = o[(NCU + i) * P];
#pragma unroll
      for (int i = 0; i < NCU * NCU; ++i) fb_uh[i] = o[(NCU + NCU * NC + i) * P];
    }
  } else {
    M::template fbou<JAC>(ib, x, uq, v, uh, nl, inst_tau(a, E), inst_param(a, E), inst_time(a, E), fb, fb_uq,
                          fb_uh);
  }
#pragma unroll
  for (int k = 0; k < NCU; ++k) a.fpt[(NF1 + k) * P + p] = fb[k] * jac;
  if (JAC) {
#pragma unroll
    for (int i = 0; i < NCU * NC; ++i) a.fpt[(NF1 + NCU + i) * P + p] = fb_uq[i] * jac;
#pragma unroll
    for (int i = 0; i < NCU * NCU; ++i) a.fpt[(NF1 + NCU + NCU * NC + i) * P + p] = fb_uh[i] * jac;
  }
}

// Residual parts of face l: Ru[perm[a,l] + npe*k] -= int phi_a fhI; Rh = -int phi_a fhB
// (assembleRu, kokkosimpl.h:1650-1659; uequation.cpp:426-427).
template <typename T>
__global__ void k_face_res(Dims d, AsmArgs a, int l, int nf1, T* ru, long ldru, double* Rh) {
  const int npf = d.npf, ngf = d.ngf, ncu = d.ncu;
  const long per = static_cast<long>(npf) * ncu;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * a.eb) return;
  const long e = q / per;
  const int r = static_cast<int>(q % per);
  const int aa = r % npf, k = r / npf;
  const long P = static_cast<long>(a.eb) * d.nfe * ngf;
  const long pb = (e * d.nfe + l) * ngf;
  double sI = 0.0, sB = 0.0;
  for (int g = 0; g < ngf; ++g) {
    const double w = a.shapfgw[aa + npf * g];
    sI += w * a.fpt[k * P + pb + g];
    sB += w * a.fpt[(nf1 + k) * P + pb + g];
  }
  T& dst = ru[e * ldru + a.perm[aa + npf * l] + d.npe * k];
  dst = static_cast<T>(static_cast<double>(dst) - sI);
  Rh[(a.e0 + e) * d.m + k + ncu * (aa + npf * l)] = -sB;
}

// Jacobian parts of face l: D/B += int phi_a phi_b fhI_uq, F += int phi_a phi_b fhI_uh,
// K/G = int phi_a phi_b fhB_uq, H = int phi_a phi_b fhB_uh
// (assembleMatrixBD/F/GK/H, kokkosimpl.h:1661-1730).
template <typename T>
__global__ void k_face_jac(Dims d, AsmArgs a, int l, T* aug, T* Bfull, T* Kt, T* Gfull, T* H) {
  const int npf = d.npf, ngf = d.ngf, ncu = d.ncu, nc = d.nc, npe = d.npe, nd = d.nd;
  const int n = d.n, m = d.m;
  const int ncol = nc + ncu;
  const long per = static_cast<long>(npf) * npf * ncu * ncol;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * a.eb) return;
  const long e = q / per;
  long r = q % per;
  const int aa = static_cast<int>(r % npf); r /= npf;
  const int bb = static_cast<int>(r % npf); r /= npf;
  const int k = static_cast<int>(r % ncu);
  const int c = static_cast<int>(r / ncu);
  const long P = static_cast<long>(a.eb) * d.nfe * ngf;
  const long pb = (e * d.nfe + l) * ngf;
  const int nf1 = ncu + ncu * nc + ncu * ncu;
  const int fI = (c < nc) ? (ncu + k + ncu * c) : (ncu + ncu * nc + k + ncu * (c - nc));
  double sI = 0.0, sB = 0.0;
  for (int g = 0; g < ngf; ++g) {
    const double w = a.shapfgw[aa + npf * g] * a.shapfgt[g + ngf * bb];
    sI += w * a.fpt[fI * P + pb + g];
    sB += w * a.fpt[(nf1 + fI) * P + pb + g];
  }
  const int pa = a.perm[aa + npf * l], pbn = a.perm[bb + npf * l];
  const long augs = a.augs;
  const long trow = k + ncu * (aa + npf * l);
  if (c < ncu) {
    aug[e * augs + pa + npe * k + static_cast<long>(n) * (pbn + npe * c)] += static_cast<T>(sI);
    Kt[e * m * static_cast<long>(n) + trow + static_cast<long>(m) * (pbn + npe * c)] += static_cast<T>(sB);
  } else if (c < nc) {
    const int cq = c - ncu, cp = cq % ncu, dd = cq / ncu;
    const long col = pbn + npe * dd + static_cast<long>(npe) * nd * cp;
    Bfull[e * n * static_cast<long>(n) * nd + pa + npe * k + n * col] += static_cast<T>(sI);
    Gfull[e * m * static_cast<long>(n) * nd + trow + m * col] += static_cast<T>(sB);
  } else {
    const int c2 = c - nc;
    aug[e * augs + pa + npe * k + static_cast<long>(n) * (n + c2 + ncu * (bb + npf * l))] += static_cast<T>(sI);
    H[(a.e0 + e) * static_cast<long>(m) * m + trow + static_cast<long>(m) * (c2 + ncu * (bb + npf * l))] =
        static_cast<T>(sB);
  }
}

// The Jacobian parts of k_face_jac for all faces of an element, one workgroup per element.
// The face matrices of the 2 ncu (nc + ncu) face-point fields f (the I and the B parts)
//   M_f[aa, bb] = sum_g shapfgw[aa, g] shapfgt[g, bb] fpt_f[g]
// are the product F^T Q^T, F[g, f] = fpt_f[g] and Q[p, g] = shapfgw[aa, g] shapfgt[g, bb]
// (p = aa + npf bb), formed on the matrix cores: F^T (fields x points) from LDS as the A
// operand, each lane's Q entry computed from the shape tables in LDS as the B operand.  In
// the accumulators 16 lanes hold 16 consecutive pairs of one field, so their updates of
// D/B/F (aug, Bfull) and K/G/H land in a few rows of one column, as in k_face_jac.  The
// faces run in order l = 0 .. nfe-1 with a barrier between them: entries that two faces
// share (D and B at the nodes of a common edge) are updated face by face, in that order.
constexpr int FJ_THREADS = 512;

template <typename T>
__global__ __launch_bounds__(FJ_THREADS) void k_face_jac_elem(Dims d, AsmArgs a, T* aug, T* Bfull, T* Kt, T* Gfull,
                                                              T* H) {
  extern __shared__ double fj_sm[];
  const int npf = d.npf, ngf = d.ngf, ncu = d.ncu, nc = d.nc, npe = d.npe, nd = d.nd, nfe = d.nfe;
  const int n = d.n, m = d.m;
  const int nfi = ncu * (nc + ncu), nfl = 2 * nfi;
  const int kp = (ngf + 3) & ~3, nfp = (nfl + 15) & ~15;
  double* Ws = fj_sm;              // shapfgw, npf x ngf
  double* Ts = Ws + npf * ngf;     // shapfgt, ngf x npf
  double* Fs = Ts + ngf * npf;     // Fs[g * nfp + f], zero beyond ngf points / nfl fields
  int* perm = reinterpret_cast<int*>(Fs + kp * nfp);
  const long e = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  constexpr int NW = FJ_THREADS / 64;
  for (int i = tid; i < npf * ngf; i += FJ_THREADS) {
    Ws[i] = a.shapfgw[i];
    Ts[i] = a.shapfgt[i];
  }
  for (int i = tid; i < npf * nfe; i += FJ_THREADS) perm[i] = a.perm[i];
  const long P = static_cast<long>(a.eb) * nfe * ngf;
  const int npair = npf * npf, ptiles = (npair + 15) / 16, ftiles = nfp / 16;
  const long augs = a.augs;
  T* augE = aug + e * augs;
  T* BE = Bfull + e * static_cast<long>(n) * n * nd;
  T* KE = Kt + e * static_cast<long>(m) * n;
  T* GE = Gfull + e * static_cast<long>(m) * n * nd;
  T* HE = H + (a.e0 + e) * static_cast<long>(m) * m;
  for (int l = 0; l < nfe; ++l) {
    const long pb = (e * nfe + l) * ngf;
    __syncthreads();  // the previous face's reads of Fs and its updates are done
    for (int i = tid; i < kp * nfp; i += FJ_THREADS) {
      const int g = i / nfp, f = i % nfp;
      // I fields are fpt fields ncu .. ncu+nfi-1, B fields 2 ncu + nfi .. 2 ncu + 2 nfi - 1
      Fs[i] = (g < ngf && f < nfl) ? a.fpt[(f < nfi ? ncu + f : 2 * ncu + f) * P + pb + g] : 0.0;
    }
    __syncthreads();
    for (int pt = wave; pt < ptiles; pt += NW) {
      const int p = pt * 16 + li;
      const bool pok = p < npair;
      const int aa = pok ? p % npf : 0, bb = pok ? p / npf : 0;
      const int pa = perm[aa + npf * l], pbn = perm[bb + npf * l];
      for (int f0 = 0; f0 < ftiles; f0 += 4) {
        dbl4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = dbl4{0, 0, 0, 0};
        for (int k0 = 0; k0 < kp; k0 += 4) {
          const int g = k0 + lk;
          const double q = (pok && g < ngf) ? Ws[aa + npf * g] * Ts[g + ngf * bb] : 0.0;
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (f0 + t < ftiles) acc[t] = mfma_step(Fs[g * nfp + (f0 + t) * 16 + li], q, acc[t]);
        }
        if (!pok) continue;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          if (f0 + t >= ftiles) break;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int f = (f0 + t) * 16 + lk + 4 * i;
            if (f >= nfl) continue;
            const double v = acc[t][i];
            const T vt = static_cast<T>(v);
            const bool bpart = f >= nfi;
            const int fl = bpart ? f - nfi : f;
            const int k = fl % ncu, c = fl / ncu;
            const long trow = k + ncu * (aa + npf * l);
            if (c < ncu) {
              if (bpart) KE[trow + static_cast<long>(m) * (pbn + npe * c)] += vt;
              else augE[pa + npe * k + static_cast<long>(n) * (pbn + npe * c)] += vt;
            } else if (c < nc) {
              const int cq = c - ncu, cp = cq % ncu, dd = cq / ncu;
              const long col = pbn + npe * dd + static_cast<long>(npe) * nd * cp;
              if (bpart) GE[trow + m * col] += vt;
              else BE[pa + npe * k + n * col] += vt;
            } else {
              const long col = (c - nc) + ncu * (bb + npf * l);
              if (bpart) HE[trow + static_cast<long>(m) * col] = vt;
              else augE[pa + npe * k + static_cast<long>(n) * (n + col)] += vt;
            }
          }
        }
      }
    }
  }
}

// LDS bytes of k_face_jac_elem
size_t face_jac_elem_lds(const Dims& d) {
  const int nfl = 2 * d.ncu * (d.nc + d.ncu);
  const long kp = (d.ngf + 3) & ~3, nfp = (nfl + 15) & ~15;
  return sizeof(double) * (2L * d.npf * d.ngf + kp * nfp) + sizeof(int) * static_cast<size_t>(d.npf) * d.nfe;
}

// After the solve: DinvF = -X_F, Ru = X_Ru (uequation.cpp:700-711).
template <typename TA, typename TD>
__global__ void k_post_schur(Dims d, int e0, int eb, const TA* aug, long augs, TD* DinvF, double* RuS) {
  const int n = d.n, m = d.m;
  const long per = static_cast<long>(n) * (m + 1);
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * eb) return;
  const long e = q / per;
  const long r = q % per;
  const TA v = aug[e * augs + static_cast<long>(n) * n + r];
  const long E = e0 + e;
  if (r < static_cast<long>(n) * m) DinvF[E * n * static_cast<long>(m) + r] = -v;
  else RuS[E * n + (r - static_cast<long>(n) * m)] = v;
}

// The same with a 2-D grid (x: entries of one element, y: element): no index division.
template <typename TA, typename TD>
__global__ void k_post_schur_2d(Dims d, int e0, const TA* aug, long augs, TD* DinvF, double* RuS) {
  const int n = d.n, m = d.m;
  const long r = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (r >= static_cast<long>(n) * (m + 1)) return;
  const long e = blockIdx.y;
  const TA v = aug[e * augs + static_cast<long>(n) * n + r];
  const long E = e0 + e;
  if (r < static_cast<long>(n) * m) DinvF[E * n * static_cast<long>(m) + r] = -v;
  else RuS[E * n + (r - static_cast<long>(n) * m)] = v;
}

// ---------------------------------------------------------------------------
// Global system
// ---------------------------------------------------------------------------
// PutElementFaceNodes twice (kokkosimpl.h:1032-1064) as one gather per face node.
__global__ void k_face_gather(long nnodes, int ncu, const long* src1, const long* src2,
                              const double* we, double* w) {
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= nnodes * ncu) return;
  const long g = q / ncu;
  const int c = static_cast<int>(q % ncu);
  double s = we[src1[g] + c];
  const long s2 = src2[g];
  if (s2 >= 0) s += we[s2 + c];
  w[q] = s;
}

// BlockJacobi (kokkosimpl.h:1067-1111) into an augmented [BE | I] for inversion.
template <typename TH>
__global__ void k_bj_assemble(Dims d, const int* f2e, const int* elemcon, const TH* H, double* aug) {
  const int ncf = d.ncf, ncu = d.ncu, npf = d.npf, nfe = d.nfe, m = d.m;
  const long per = static_cast<long>(ncf) * 2 * ncf;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * d.nf) return;
  const long f = q / per;
  const int rs = static_cast<int>(q % per);
  const int r = rs % ncf, s = rs / ncf;
  double v;
  if (s >= ncf) {
    v = (s - ncf == r) ? 1.0 : 0.0;
  } else {
    const int e1 = f2e[4 * f], l1 = f2e[4 * f + 1], e2 = f2e[4 * f + 2], l2 = f2e[4 * f + 3];
    v = H[static_cast<long>(e1) * m * m + (r + ncf * l1) + static_cast<long>(m) * (s + ncf * l1)];
    if (e2 >= 0) {
      const int ar = r % ncu, br = r / ncu, as = s % ncu, bs = s / ncu;
      const long base = npf * (l2 + static_cast<long>(nfe) * e2);
      const int dr = ar + ncu * static_cast<int>(elemcon[br + base] - npf * f);
      const int ds = as + ncu * static_cast<int>(elemcon[bs + base] - npf * f);
      v += H[static_cast<long>(e2) * m * m + (dr + ncf * l2) + static_cast<long>(m) * (ds + ncf * l2)];
    }
  }
  aug[f * per + rs] = v;
}

// Elemental additive Schwarz (hdgElementalAdditiveSchwarz, matvec.cpp:168-188;
// kokkosimpl.h:1113-1142): element e's m x m block of H with the diagonal block of
// each interior face replaced by the sum of both sides' blocks, written as the
// augmented [block | I] (m x 2m) for the batched LU inversion (FP64, or FP32 in the mixed form).
template <typename TH, typename TA>
__global__ void k_eas_assemble(Dims d, int e0, int eb, const TH* H, const int* f2e, const int* elemcon,
                               TA* aug) {
  const int m = d.m, ncf = d.ncf, ncu = d.ncu, npf = d.npf, ndf = d.ndf;
  const long per = static_cast<long>(m) * 2 * m;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * eb) return;
  const int le = static_cast<int>(q / per);
  const int rs = static_cast<int>(q % per);
  const int r = rs % m, c = rs / m;
  const long e = e0 + le;
  double v;
  if (c >= m) {
    v = (c - m == r) ? 1.0 : 0.0;
  } else {
    v = H[e * m * m + r + static_cast<long>(m) * c];
    const int l = r / ncf;
    if (c / ncf == l) {
      const int f = elemcon[npf * l + ndf * e] / npf;
      const int e1 = f2e[4 * f], l1 = f2e[4 * f + 1], e2 = f2e[4 * f + 2], l2 = f2e[4 * f + 3];
      if (e2 >= 0) {
        const int rr = r % ncf, cc = c % ncf;
        const int am = rr % ncu, bm = rr / ncu, an = cc % ncu, bn = cc / ncu;
        // the other side's position of face node bm: through e2's elemcon (an involution)
        const int om = am + ncu * (elemcon[bm + npf * l2 + ndf * e2] - npf * f);
        const int on = an + ncu * (elemcon[bn + npf * l2 + ndf * e2] - npf * f);
        const bool first = (e == e1 && l == l1);
        const long eo = first ? e2 : e1;
        const int lo = first ? l2 : l1;
        v += H[eo * m * m + (om + ncf * lo) + static_cast<long>(m) * (on + ncf * lo)];
      }
    }
  }
  aug[static_cast<long>(le) * per + rs] = static_cast<TA>(v);
}

template <typename T>
__global__ void k_eas_extract(Dims d, int eb, const T* aug, T* BEe) {
  const long per = static_cast<long>(d.m) * d.m;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * eb) return;
  const long e = q / per;
  BEe[q] = aug[e * 2 * per + per + q % per];
}

__global__ void k_bj_extract(Dims d, const double* aug, double* BE) {
  const long per = static_cast<long>(d.ncf) * d.ncf;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * d.nf) return;
  const long f = q / per;
  BE[q] = aug[f * 2 * per + per + q % per];
}

// x_f <- BE_f x_f (preconditioner.cpp:282-292); one thread per row.
template <typename T>
__global__ void k_bj_apply(Dims d, const T* BE, const double* x, double* y, const double* act) {
  if (act && act[0] == 0.0) return;
  const int ncf = d.ncf;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= static_cast<long>(ncf) * d.nf) return;
  const long f = q / ncf;
  const int r = static_cast<int>(q % ncf);
  const T* B = BE + f * ncf * static_cast<long>(ncf);
  const double* xf = x + f * ncf;
  double s = 0.0;
  for (int c = 0; c < ncf; ++c) s += static_cast<double>(B[r + static_cast<long>(ncf) * c]) * xf[c];
  y[q] = s;
}

// Block entries as doubles: FP64 / FP32 storage, or BF16 kept as the upper 16 bits of an FP32
// (unsigned short storage)
template <typename T>
__device__ __forceinline__ double blk_val(T x) { return static_cast<double>(x); }
template <>
__device__ __forceinline__ double blk_val<unsigned short>(unsigned short x) {
  return static_cast<double>(__uint_as_float(static_cast<unsigned int>(x) << 16));
}

// hdgMatVec element part (matvec.cpp:377-380): we_e = H_e * v_e with v_e gathered
// through elemcon.  One workgroup per element; each of the 4 waves owns a
// quarter of the columns, lanes own VW consecutive rows (16-byte vector loads of
// the column), accumulation is FP64, and the four partial sums are added in LDS
// in a fixed order.
// NRB > 0: exactly NRB row blocks of 64 VW rows (m > 64 VW (NRB - 1)), with CB columns' loads
// issued before their FMAs (more bytes in flight per wave; a row's FMAs still run over the columns
// in order); NRB = 0: any m <= 1024, one column at a time.
template <typename T, int VW, int NRB>
__global__ __launch_bounds__(256) void k_matvec_elem(Dims d, const T* H, const int* elemcon,
                                                     const double* v, double* we, const int* einst,
                                                     const double* emask) {
  typedef T vecT __attribute__((ext_vector_type(VW)));
  constexpr int MAXRB = NRB > 0 ? NRB : (1024 + 64 * VW - 1) / (64 * VW);  // m <= 1024
  constexpr int CB = NRB > 0 ? (sizeof(T) <= 4 ? 4 : 2) : 1;
  const int m = d.m, ncu = d.ncu, ndf = d.ndf;
  const long e = blockIdx.x;
  if (emask && einst && emask[einst[e]] == 0.0) {  // block-uniform: the element's instance is inactive
    for (int r = threadIdx.x; r < m; r += 256) we[e * m + r] = 0.0;
    return;
  }
  extern __shared__ double sm[];
  double* ve = sm;          // m
  double* red = sm + m;     // 4*m
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int t = tid; t < m; t += 256)
    ve[t] = v[(t % ncu) + ncu * static_cast<long>(elemcon[t / ncu + ndf * e])];
  __syncthreads();
  const int cper = (m + 3) / 4;
  const int c0 = wave * cper, c1 = min(m, c0 + cper);
  const int nrb = (m + 64 * VW - 1) / (64 * VW);
  double acc[MAXRB][VW];
#pragma unroll
  for (int rb = 0; rb < MAXRB; ++rb)
#pragma unroll
    for (int q = 0; q < VW; ++q) acc[rb][q] = 0.0;
  const T* He = H + e * static_cast<long>(m) * m;
  int cfirst = c0;
  if (CB > 1) {
    for (; cfirst + CB <= c1; cfirst += CB) {
      vecT h[CB][MAXRB];
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        const T* col = He + static_cast<long>(m) * (cfirst + j);
#pragma unroll
        for (int rb = 0; rb < MAXRB; ++rb) {
          const int r = VW * (lane + 64 * rb);
          if (r < m) h[j][rb] = *reinterpret_cast<const vecT*>(col + r);
        }
      }
#pragma unroll
      for (int j = 0; j < CB; ++j) {
        const double vc = ve[cfirst + j];
#pragma unroll
        for (int rb = 0; rb < MAXRB; ++rb) {
          const int r = VW * (lane + 64 * rb);
          if (r < m)
#pragma unroll
            for (int q = 0; q < VW; ++q) acc[rb][q] = fma(blk_val(h[j][rb][q]), vc, acc[rb][q]);
        }
      }
    }
  }
  for (int c = cfirst; c < c1; ++c) {
    const double vc = ve[c];
    const T* col = He + static_cast<long>(m) * c;
#pragma unroll
    for (int rb = 0; rb < MAXRB; ++rb) {
      const int r = VW * (lane + 64 * rb);
      if (rb < nrb && r < m) {
        const vecT h = *reinterpret_cast<const vecT*>(col + r);
#pragma unroll
        for (int q = 0; q < VW; ++q) acc[rb][q] = fma(blk_val(h[q]), vc, acc[rb][q]);
      }
    }
  }
#pragma unroll
  for (int rb = 0; rb < MAXRB; ++rb) {
    const int r = VW * (lane + 64 * rb);
    if (rb < nrb && r < m)
#pragma unroll
      for (int q = 0; q < VW; ++q) red[wave * m + r + q] = acc[rb][q];
  }
  __syncthreads();
  for (int r = tid; r < m; r += 256)
    we[e * m + r] = ((red[r] + red[m + r]) + red[2 * m + r]) + red[3 * m + r];
}

// The same product for small trace blocks (m <= 96, e.g. 2D elements): one thread per row
// of one element, the rows of consecutive elements in consecutive lanes, FP64 accumulation
// over the columns in order.  emask/einst (optional): elements of an instance whose mask is 0
// only write zeros (instances that have left a batched GMRES).
template <typename T>
__global__ __launch_bounds__(256) void k_matvec_small(Dims d, const T* H, const int* elemcon, const double* v,
                                                      double* we, const int* einst, const double* emask) {
  const int m = d.m, ncu = d.ncu, ndf = d.ndf;
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= static_cast<long>(m) * d.ne) return;
  const long e = q / m;
  const int r = static_cast<int>(q % m);
  if (emask && einst && emask[einst[e]] == 0.0) {
    we[q] = 0.0;
    return;
  }
  const T* He = H + e * static_cast<long>(m) * m + r;
  const int* ec = elemcon + ndf * e;
  double acc = 0.0;
  for (int c = 0; c < m; ++c)
    acc = fma(blk_val(He[static_cast<long>(m) * c]), v[(c % ncu) + ncu * static_cast<long>(ec[c / ncu])],
              acc);
  we[q] = acc;
}

// hdgGetDUDG (matvec.cpp:219-238): du_e = Ru_e + DinvF_e * x_e.  One workgroup per element: the
// element's trace values x_e are gathered into LDS once, then one thread per row reads DinvF_e
// along its columns (the sum runs over t in order).
template <typename T>
__global__ __launch_bounds__(1024) void k_dudg(Dims d, const double* RuS, const T* DinvF, const int* elemcon,
                                               const double* x, double* du) {
  extern __shared__ double dd_sm[];
  const int n = d.n, m = d.m, ncu = d.ncu, ndf = d.ndf;
  const long e = blockIdx.x;
  const int tid = threadIdx.x;
  for (int t = tid; t < m; t += blockDim.x)
    dd_sm[t] = x[(t % ncu) + ncu * static_cast<long>(elemcon[t / ncu + ndf * e])];
  __syncthreads();
  const T* F = DinvF + e * n * static_cast<long>(m);
  for (int r = tid; r < n; r += blockDim.x) {
    double s = RuS[e * n + r];
#pragma unroll 4
    for (int t = 0; t < m; ++t) s += static_cast<double>(F[r + static_cast<long>(n) * t]) * dd_sm[t];
    du[e * n + r] = s;
  }
}

// UpdateUDG (kokkosimpl.h:322-338) on the u components.
__global__ void k_update_udg(Dims d, const double* du, double alpha, double* udg) {
  const int npe = d.npe, ncu = d.ncu, nc = d.nc;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= static_cast<long>(npe) * ncu * d.ne) return;
  const long e = q / (npe * ncu);
  const int r = static_cast<int>(q % (npe * ncu));
  const int i = r % npe, c = r / npe;
  udg[i + npe * (c + nc * e)] += alpha * du[q];
}

// ---------------------------------------------------------------------------
// Superelement block-ILU(0): AssembleBlockILU0 (kokkosimpl.h:1264-1334),
// hdgBlockILU0 (matvec.cpp:119-166), ApplyBlockILU0 (preconditioner.cpp:184-270).
// BE[((t*nse) + n)*ncf*ncf + r + ncf*c]: CRS entry t of superelement n.
// ---------------------------------------------------------------------------
// grid: (superelement n, CRS entry t)
template <typename T, typename TH>
__global__ __launch_bounds__(256) void k_bilu_assemble(Dims d, IluDev I, const int* f2e, const int* elemcon,
                                                       const TH* H, T* BE) {
  // one workgroup per (superelement n, block slot t): the block's element(s), local faces and node
  // permutations are found once and kept in LDS as H row / column offsets; the ncf x ncf entries are
  // then gathered column by column (the diagonal block of a face shared by two elements adds both)
  __shared__ int rowoff[2][256], coloff[2][256];
  const int ncf = d.ncf, ncu = d.ncu, npf = d.npf, ndf = d.ndf, m = d.m;
  const int per = ncf * ncf;
  const int n = blockIdx.x, t = blockIdx.y, tid = threadIdx.x;
  const int i = I.row_of[t];
  const int fi = I.face[n + I.nse * i];
  const int e1 = f2e[4 * fi], l1 = f2e[4 * fi + 1], e2 = f2e[4 * fi + 2], l2 = f2e[4 * fi + 3];
  const int nfi = npf * fi;
  const bool diag = t == I.row_ptr[i];
  int ea, eb = -1;  // the element of each term
  if (diag) {
    ea = e1;
    eb = e2;
    for (int r = tid; r < ncf; r += 256) {
      const int am = r % ncu, bm = r / ncu;
      rowoff[0][r] = am + ncu * (elemcon[bm + npf * l1 + ndf * e1] - nfi) + ncf * l1;
      coloff[0][r] = rowoff[0][r];
      if (e2 >= 0) {
        rowoff[1][r] = am + ncu * (elemcon[bm + npf * l2 + ndf * e2] - nfi) + ncf * l2;
        coloff[1][r] = rowoff[1][r];
      }
    }
  } else {
    const int j = I.col_ind[t];
    const int fj = I.face[n + I.nse * j];
    const int je1 = f2e[4 * fj], je2 = f2e[4 * fj + 2];
    int k1, k2;
    if (je1 == e1) { ea = e1; k1 = l1; k2 = f2e[4 * fj + 1]; }
    else if (je1 == e2) { ea = e2; k1 = l2; k2 = f2e[4 * fj + 1]; }
    else if (je2 == e1) { ea = e1; k1 = l1; k2 = f2e[4 * fj + 3]; }
    else { ea = e2; k1 = l2; k2 = f2e[4 * fj + 3]; }
    for (int r = tid; r < ncf; r += 256) {
      const int am = r % ncu, bm = r / ncu;
      rowoff[0][r] = am + ncu * (elemcon[bm + npf * k1 + ndf * ea] - nfi) + ncf * k1;
      coloff[0][r] = am + ncu * (elemcon[bm + npf * k2 + ndf * ea] - npf * fj) + ncf * k2;
    }
  }
  __syncthreads();
  const TH* Ha = H + static_cast<long>(ea) * m * m;
  const TH* Hb = H + static_cast<long>(eb >= 0 ? eb : ea) * m * m;
  T* out = BE + (static_cast<long>(t) * I.nse + n) * per;
  for (int rc = tid; rc < per; rc += 256) {
    const int rr = rc % ncf, cc = rc / ncf;
    double v = Ha[rowoff[0][rr] + static_cast<long>(m) * coloff[0][cc]];
    if (eb >= 0) v += Hb[rowoff[1][rr] + static_cast<long>(m) * coloff[1][cc]];
    out[rc] = static_cast<T>(v);
  }
}

// Inverse of the diagonal block of face slot i of every superelement (one
// workgroup each): Gauss-Jordan with partial pivoting on [D | I] in LDS
// (row-major, 2ncf wide), the inverse written back in place.
template <typename T>
__global__ __launch_bounds__(256) void k_bilu_inv(Dims d, IluDev I, int i, T* BE, int* info) {
  const int ncf = d.ncf;
  const int n = blockIdx.x;
  const long per = static_cast<long>(ncf) * ncf;
  extern __shared__ double sm[];
  T* A = reinterpret_cast<T*>(sm);    // ncf x 2ncf, row-major: A[r*2ncf + c]
  T* fac = A + 2 * per;               // ncf
  __shared__ double wv[4];
  __shared__ int wi[4];
  __shared__ int pivs;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int w2 = 2 * ncf;
  T* D = BE + (static_cast<long>(I.diag[i]) * I.nse + n) * per;
  for (int idx = tid; idx < ncf * w2; idx += 256) {
    int r = idx / w2, c = idx % w2;
    A[idx] = (c < ncf) ? D[r + ncf * c] : ((c - ncf) == r ? T(1) : T(0));
  }
  __syncthreads();
  int singular = 0;
  for (int c = 0; c < ncf; ++c) {
    double best = -1.0;
    int bi = c;
    for (int r = c + tid; r < ncf; r += 256) {
      double vv = fabs(A[r * w2 + c]);
      if (vv > best) { best = vv; bi = r; }
    }
    argmax_wave(best, bi);
    if (lane == 0) { wv[wave] = best; wi[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      double b = wv[0];
      int p = wi[0];
      for (int w = 1; w < 4; ++w)
        if (wv[w] > b || (wv[w] == b && wi[w] < p)) { b = wv[w]; p = wi[w]; }
      pivs = p;
    }
    __syncthreads();
    const int p = pivs;
    if (p != c)
      for (int k = tid; k < w2; k += 256) {
        T tt = A[c * w2 + k];
        A[c * w2 + k] = A[p * w2 + k];
        A[p * w2 + k] = tt;
      }
    __syncthreads();
    const T pv = A[c * w2 + c];
    if (pv == 0.0) singular = 1;
    const T inv = T(1) / pv;
    for (int r = tid; r < ncf; r += 256) fac[r] = A[r * w2 + c];
    // every wave has read the pivot before row c (the pivot itself included) is scaled
    __syncthreads();
    for (int k = tid; k < w2; k += 256) A[c * w2 + k] *= inv;
    __syncthreads();
    // a wave per row at a time, a lane per column (no index division)
    for (int r = wave; r < ncf; r += 4) {
      if (r == c) continue;
      const T f = fac[r];
      for (int k = lane; k < w2; k += 64) A[r * w2 + k] = fma(-f, A[c * w2 + k], A[r * w2 + k]);
    }
    __syncthreads();
  }
  for (int idx = tid; idx < static_cast<int>(per); idx += 256) {
    int r = idx % ncf, c = idx / ncf;
    D[idx] = A[r * w2 + ncf + c];
  }
  if (tid == 0 && info) info[n] = (i == 0 ? 0 : info[n]) | singular;
}

// BE[elim_ji[k0 + k]][n] = Ltmp[k][n]
template <typename T>
__global__ void k_bilu_put_l(Dims d, IluDev I, int k0, int ne_i, const T* Ltmp, T* BE) {
  const long per = static_cast<long>(d.ncf) * d.ncf;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  const long tot = per * I.nse * ne_i;
  if (q >= tot) return;
  const long kn = q / per;
  const int k = static_cast<int>(kn / I.nse);
  const long n = kn % I.nse;
  BE[(static_cast<long>(I.elim_ji[k0 + k]) * I.nse + n) * per + q % per] = Ltmp[q];
}

// loads a thread keeps in flight per step of a row's sweep (raw storage words in registers)
template <typename T>
constexpr int bilu_unroll() { return sizeof(T) == 2 ? 12 : sizeof(T) == 4 ? 12 : 8; }

// One workgroup per superelement (ApplyBlockILU0, preconditioner.cpp:184-270):
// forward sweep y_i = b_i - sum L_ij y_j, backward sweep x_i = D_i^{-1} (y_i - sum
// U_il x_l).  The superelement's vector lives in LDS (y_i is overwritten by x_i).
// Thread (rr, part) accumulates rows VW*rr .. VW*rr+VW-1 over the part-th chunk of the
// columns of every block of the current row, with one VW-wide vector load per column, and
// all 256 threads issue loads in every step; the chunks are added in a fixed order.  The
// (block, column) pairs of a row are walked as one list, U at a time: the U loads are issued
// before their FMAs, so a thread keeps U loads in flight across block boundaries (the row
// sweep is bound by HBM bytes in flight).  The sweep structure is staged in LDS.  Xo is
// written as [slot = n + nse*i][ncf].
template <typename T, int VW>
__global__ __launch_bounds__(256) void k_bilu_apply(Dims d, IluDev I, const T* BE, const double* x,
                                                    double* Xo, const double* act) {
  if (act && act[0] == 0.0) return;  // a finished solve (grid-uniform)
  typedef T vecT __attribute__((ext_vector_type(VW)));
  constexpr int U = bilu_unroll<T>();
  const int ncf = d.ncf, nfse = I.nfse;
  const long n = blockIdx.x;
  const long per = static_cast<long>(ncf) * ncf;
  const int tid = threadIdx.x;
  const int nrow = ncf / VW;                 // row groups
  const int nparts = 256 / nrow;
  const int rr = tid % nrow, part = tid / nrow;
  const int r0 = VW * rr;
  const bool active = part < nparts;
  const int cw = (ncf + nparts - 1) / nparts;
  const int c0 = min(ncf, part * cw), c1 = min(ncf, c0 + cw);
  // the superelement's vector in LDS: FP64 with FP64 factors, FP32 with FP32 / BF16 factors (it
  // carries values the factors' own rounding already limits; the smaller LDS footprint lets 8
  // superelements share a CU instead of 6 at C2's 36 faces of 50 unknowns)
  using VT = typename std::conditional<sizeof(T) == 8, double, float>::type;
  extern __shared__ double sm[];
  double* red = sm;                        // nparts * ncf <= 256 * VW
  VT* v = reinterpret_cast<VT*>(red + 256 * VW);   // nfse * ncf
  VT* z = v + static_cast<long>(nfse) * ncf;       // ncf
  // the sweep structure (the same for every superelement) in LDS
  int* lptr = reinterpret_cast<int*>(z + ncf);
  int* uuptr = lptr + nfse + 1;
  int* diag = uuptr + nfse + 1;
  int* lidx = diag + nfse;
  int* lcol = lidx + I.nnz;
  int* uidx = lcol + I.nnz;
  int* ucol = uidx + I.nnz;
  for (int idx = tid; idx <= nfse; idx += 256) {
    lptr[idx] = I.lptr[idx];
    uuptr[idx] = I.uuptr[idx];
    if (idx < nfse) diag[idx] = I.diag[idx];
  }
  for (int idx = tid; idx < I.lptr[nfse]; idx += 256) {
    lidx[idx] = I.lidx[idx];
    lcol[idx] = I.lcol[idx];
  }
  for (int idx = tid; idx < I.uuptr[nfse]; idx += 256) {
    uidx[idx] = I.uidx[idx];
    ucol[idx] = I.ucol[idx];
  }
  auto blk = [&](int t) { return BE + (static_cast<long>(t) * I.nse + n) * per + r0; };
  for (int idx = tid; idx < nfse * ncf; idx += 256) {
    const int i = idx / ncf, q = idx % ncf;
    v[idx] = static_cast<VT>(x[static_cast<long>(I.face[n + static_cast<long>(I.nse) * i]) * ncf + q]);
  }
  __syncthreads();
  double acc[VW];
  // acc += sum over the blocks q in [qb, qe) of block idx[q] (this thread's rows, columns
  // c0 .. c1-1) times vec[col[q]], in the order (q, c); U loads are issued before their FMAs
  auto row_pass = [&](int qb, int qe, const int* idx, const int* col, const VT* vec) {
    const int cwn = c1 - c0;
    const int tot = (qe - qb) * cwn;
    int qq = qb, cc = c0;
    for (int t0 = 0; t0 < tot; t0 += U) {
      vecT raw[U];
      int yo[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (t0 + u < tot) {
          raw[u] = *reinterpret_cast<const vecT*>(blk(idx[qq]) + static_cast<long>(ncf) * cc);
          yo[u] = (col ? col[qq] * ncf : 0) + cc;
          if (++cc == c1) {
            cc = c0;
            ++qq;
          }
        } else {
          yo[u] = -1;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (yo[u] >= 0) {
          const double yc = static_cast<double>(vec[yo[u]]);
#pragma unroll
          for (int k = 0; k < VW; ++k) acc[k] = fma(blk_val<T>(raw[u][k]), yc, acc[k]);
        }
    }
  };
  for (int i = 0; i < nfse; ++i) {
#pragma unroll
    for (int k = 0; k < VW; ++k) acc[k] = 0.0;
    if (active) row_pass(lptr[i], lptr[i + 1], lidx, lcol, v);
    if (active)
#pragma unroll
      for (int k = 0; k < VW; ++k) red[part * ncf + r0 + k] = acc[k];
    __syncthreads();
    if (tid < ncf) {
      double sacc = v[i * ncf + tid];
      for (int p = 0; p < nparts; ++p) sacc -= red[p * ncf + tid];
      v[i * ncf + tid] = sacc;
    }
    __syncthreads();
  }
  for (int i = nfse - 1; i >= 0; --i) {
#pragma unroll
    for (int k = 0; k < VW; ++k) acc[k] = 0.0;
    if (active) row_pass(uuptr[i], uuptr[i + 1], uidx, ucol, v);
    if (active)
#pragma unroll
      for (int k = 0; k < VW; ++k) red[part * ncf + r0 + k] = acc[k];
    __syncthreads();
    if (tid < ncf) {
      double sacc = v[i * ncf + tid];
      for (int p = 0; p < nparts; ++p) sacc -= red[p * ncf + tid];
      z[tid] = sacc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < VW; ++k) acc[k] = 0.0;
    if (active) {
      row_pass(0, 1, diag + i, nullptr, z);
#pragma unroll
      for (int k = 0; k < VW; ++k) red[part * ncf + r0 + k] = acc[k];
    }
    __syncthreads();
    if (tid < ncf) {
      double sacc = 0.0;
      for (int p = 0; p < nparts; ++p) sacc += red[p * ncf + tid];
      v[i * ncf + tid] = sacc;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < nfse * ncf; idx += 256) {
    const int i = idx / ncf, q = idx % ncf;
    Xo[(n + static_cast<long>(I.nse) * i) * ncf + q] = v[idx];
  }
}

// The same sweeps with one wave per superelement (BF16 factors; ncf = nparts CW with
// nparts = 64 / (ncf / VW)): lane (rr, part) accumulates rows VW*rr .. VW*rr+VW-1 over columns
// part*CW .. part*CW+CW-1 of every block, the parts are added by lane shuffles in a fixed order,
// and no workgroup barrier is needed (the wave's LDS operations complete in order). The sweeps
// are one flat program of (block, vector offset) pairs in 3 nfse segments (forward rows
// 0..nfse-1, then for i = nfse-1..0 the U row and the diagonal block, whose vector z sits after
// v), read with uniform indices (scalar loads), so the LDS holds only the superelement's vector
// and 16 superelements share a CU. Blocks are walked one at a time (CW loads per lane), the next
// block's loads issued before the current block's FMAs, across segment ends: the loads never wait
// for a row's result.
template <typename T, int VW, int NCF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_bilu_apply_wave(Dims d, IluDev I, const T* BE, const double* x, double* Xo, const double* act) {
  if (act && act[0] == 0.0) return;  // a finished solve (grid-uniform)
  typedef T vecT __attribute__((ext_vector_type(VW)));
  constexpr int nrow = NCF / VW;
  constexpr int nparts = 64 / nrow;
  constexpr int CW = NCF / nparts;
  static_assert(NCF % VW == 0 && nparts * CW == NCF, "the wave's parts must tile the block's columns");
  constexpr int ncf = NCF;
  const int nfse = I.nfse;
  const long n = blockIdx.x;
  constexpr long per = static_cast<long>(NCF) * NCF;
  const long bstride = static_cast<long>(I.nse) * per;
  const int lane = threadIdx.x;
  const int rr = lane % nrow, part = lane / nrow;
  const bool active = part < nparts;
  const int r0 = VW * rr;
  const int c0 = active ? part * CW : 0;   // idle lanes (part >= nparts) read part 0's columns
  // the program is read-only for the kernel's lifetime: constant address space (scalar loads)
  typedef const int __attribute__((address_space(4)))* cptr;
  const cptr wseg = reinterpret_cast<cptr>(reinterpret_cast<uintptr_t>(I.wseg));
  const cptr wblk = reinterpret_cast<cptr>(reinterpret_cast<uintptr_t>(I.wblk));
  const cptr wvec = reinterpret_cast<cptr>(reinterpret_cast<uintptr_t>(I.wvec));
  extern __shared__ double sm[];
  float* v = reinterpret_cast<float*>(sm);  // nfse * ncf, then z (ncf)
  for (int idx = lane; idx < nfse * ncf; idx += 64) {
    const int i = idx / ncf, q = idx % ncf;
    v[idx] = static_cast<float>(x[static_cast<long>(I.face[n + static_cast<long>(I.nse) * i]) * ncf + q]);
  }
  __syncthreads();
  // this lane's column c0 of block 0 (the columns' offsets are compile-time constants)
  const T* base = BE + n * per + r0 + static_cast<long>(ncf) * c0;
  const int nseg = 3 * nfse;
  const int total = wseg[nseg];   // blocks in the program (>= nfse: every row has its diagonal)
  auto issue = [&](int q, vecT (&raw)[CW]) {
    const T* p = base + wblk[min(q, total - 1)] * bstride;
#pragma unroll
    for (int u = 0; u < CW; ++u) raw[u] = *reinterpret_cast<const vecT*>(p + ncf * u);
  };
  double acc[VW];
#pragma unroll
  for (int j = 0; j < VW; ++j) acc[j] = 0.0;
  auto consume = [&](int q, const vecT (&raw)[CW]) {
    const float* y = v + wvec[q] + c0;
#pragma unroll
    for (int u = 0; u < CW; ++u) {
      const double yc = static_cast<double>(y[u]);
#pragma unroll
      for (int j = 0; j < VW; ++j) acc[j] = fma(blk_val<T>(raw[u][j]), yc, acc[j]);
    }
  };
  // segment s is complete: add the parts and store its row's result
  auto finish = [&](int sg) {
    double tot[VW];
#pragma unroll
    for (int j = 0; j < VW; ++j) {
      const double a = active ? acc[j] : 0.0;
      acc[j] = a;
      tot[j] = 0.0;
    }
    for (int p = 0; p < nparts; ++p)
#pragma unroll
      for (int j = 0; j < VW; ++j) tot[j] += __shfl(acc[j], rr + p * nrow, 64);
    const int kind = sg < nfse ? 0 : 1 + ((sg - nfse) & 1);
    const int row = sg < nfse ? sg : nfse - 1 - (sg - nfse) / 2;
    if (lane < nrow) {
      float* vi = v + row * ncf + r0;
      float* zi = v + nfse * ncf + r0;
#pragma unroll
      for (int j = 0; j < VW; ++j) {
        if (kind == 0) vi[j] = static_cast<float>(static_cast<double>(vi[j]) - tot[j]);
        else if (kind == 1) zi[j] = static_cast<float>(static_cast<double>(vi[j]) - tot[j]);
        else vi[j] = static_cast<float>(tot[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < VW; ++j) acc[j] = 0.0;
  };
  int sg = 0;
  auto finish_through = [&](int q) {  // every segment that ends at or before block q
    while (sg < nseg && wseg[sg + 1] <= q) finish(sg++);
  };
  finish_through(0);  // segments with no blocks before the first one
  vecT ra[CW], rb[CW];
  issue(0, ra);
  // (the scheduling barriers keep the next block's loads ahead of the current block's FMAs)
  for (int q = 0; q < total; q += 2) {
    issue(q + 1, rb);   // past the last block: loads of the last block again, not used
    __builtin_amdgcn_sched_barrier(0);
    consume(q, ra);
    finish_through(q + 1);
    if (q + 1 >= total) break;
    issue(q + 2, ra);
    __builtin_amdgcn_sched_barrier(0);
    consume(q + 1, rb);
    finish_through(q + 2);
  }
  __syncthreads();
  for (int idx = lane; idx < nfse * ncf; idx += 64) {
    const int i = idx / ncf, qv = idx % ncf;
    Xo[(n + static_cast<long>(I.nse) * i) * ncf + qv] = v[idx];
  }
}

// x[f] = Xo[slot1[f]] + Xo[slot2[f]] (faces shared by two superelements get both)
__global__ void k_bilu_scatter(Dims d, IluDev I, const double* Xo, double* out, const double* act) {
  if (act && act[0] == 0.0) return;
  const int ncf = d.ncf;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= static_cast<long>(d.nf) * ncf) return;
  const long f = q / ncf;
  const int r = static_cast<int>(q % ncf);
  double s = Xo[static_cast<long>(I.slot1[f]) * ncf + r];
  if (I.slot2[f] >= 0) s += Xo[static_cast<long>(I.slot2[f]) * ncf + r];
  out[q] = s;
}

// u components of udg (DIRK stage bookkeeping, updatesource.cpp / updatesolution.cpp)
__global__ void k_u_axpby(Dims d, const double* udg, double a, double b, double* dst) {
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= static_cast<long>(d.npe) * d.ncu * d.ne) return;
  const long e = q / (d.npe * d.ncu);
  const int r = static_cast<int>(q % (d.npe * d.ncu));
  const int i = r % d.npe, c = r / d.npe;
  const double u = udg[i + d.npe * (c + static_cast<long>(d.nc) * e)];
  dst[q] = (b == 0.0) ? a * u : fma(a, u, b * dst[q]);
}

__global__ void k_u_insert(Dims d, const double* src, double* udg) {
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= static_cast<long>(d.npe) * d.ncu * d.ne) return;
  const long e = q / (d.npe * d.ncu);
  const int r = static_cast<int>(q % (d.npe * d.ncu));
  const int i = r % d.npe, c = r / d.npe;
  udg[i + d.npe * (c + static_cast<long>(d.nc) * e)] = src[q];
}

// ---------------------------------------------------------------------------
// Mixed precision: stored FP32 factors and the condensed RHS of a new residual
// ---------------------------------------------------------------------------
template <typename S>
__global__ void k_to_float(float* dst, const S* src, int rows, int cols, long ld, long sstride,
                           long dstride, int batch) {
  const long per = static_cast<long>(rows) * cols;
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= per * batch) return;
  const long b = q / per;
  const long rc = q % per;
  const int r = static_cast<int>(rc % rows);
  const long c = rc / rows;
  dst[b * dstride + rc] = static_cast<float>(src[b * sstride + r + ld * c]);
}

// Contiguous matrices (ld = rows) with a 2-D grid (x: entries of one matrix, y: matrix)
template <typename S>
__global__ void k_to_float_2d(float* dst, const S* src, long per, long sstride, long dstride) {
  const long rc = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (rc >= per) return;
  const long b = blockIdx.y;
  dst[b * dstride + rc] = static_cast<float>(src[b * sstride + rc]);
}

// One workgroup per element: RuS = D'^{-1} Ru from the stored FP32 LU factors of
// D' (replaying the blocked factorisation: each panel's row exchanges, then the
// unit-lower elimination), the upper-triangular solve, then Rh -= K' RuS.  The
// panel width nbp is the one lu_solve used for the factorisation.
__global__ __launch_bounds__(256) void k_refine_rhs(Dims d, int nbp, const float* LU, const int* piv,
                                                    const float* K, const double* Ru, double* RuS,
                                                    double* Rh) {
  const int n = d.n, m = d.m;
  const long e = blockIdx.x;
  extern __shared__ double sm[];
  double* y = sm;   // n
  const float* Le = LU + e * static_cast<long>(n) * n;
  const int* pe = piv + e * static_cast<long>(n);
  for (int t = threadIdx.x; t < n; t += 256) y[t] = Ru[e * n + t];
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += nbp) {
    const int nb = min(nbp, n - k0);
    if (threadIdx.x == 0)
      for (int kk = 0; kk < nb; ++kk) {
        const int p = k0 + pe[k0 + kk];
        if (p != k0 + kk) {
          const double t = y[k0 + kk];
          y[k0 + kk] = y[p];
          y[p] = t;
        }
      }
    __syncthreads();
    for (int k = k0; k < k0 + nb; ++k) {
      const double yk = y[k];
      for (int r = k + 1 + threadIdx.x; r < n; r += 256)
        y[r] = fma(-static_cast<double>(Le[r + static_cast<long>(n) * k]), yk, y[r]);
      __syncthreads();
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    if (threadIdx.x == 0) y[k] /= static_cast<double>(Le[k + static_cast<long>(n) * k]);
    __syncthreads();
    const double yk = y[k];
    for (int r = threadIdx.x; r < k; r += 256)
      y[r] = fma(-static_cast<double>(Le[r + static_cast<long>(n) * k]), yk, y[r]);
    __syncthreads();
  }
  for (int t = threadIdx.x; t < n; t += 256) RuS[e * n + t] = y[t];
  const float* Ke = K + e * static_cast<long>(m) * n;
  for (int r = threadIdx.x; r < m; r += 256) {
    double acc = 0.0;
    for (int c = 0; c < n; ++c) acc = fma(static_cast<double>(Ke[r + static_cast<long>(m) * c]), y[c], acc);
    Rh[e * m + r] -= acc;
  }
}

// panel width lu_solve uses for an n x n factorisation (see lu_solve).  64-wide panels (up to
// 150 KB of LDS at n = 300) were measured slower: one workgroup per CU leaves the latency-bound
// panel factorisation without company (C2 LU 0.36 s -> 0.55 s per assembly).
int lu_panel_width(int n) { return (static_cast<size_t>(n) * 32 * sizeof(double) <= 64000) ? 32 : 16; }
// the FP32 LU (the mixed form): 64-wide panels up to n = 250 (an n x 64 FP32 panel in 64 KB of LDS,
// 64 registers per row in the register panel), which halves the passes of the bandwidth-bound
// trailing update; 32 or 16 beyond
int lu_panel_width_f32(int n) {
  if (n <= 250) return 64;
  return (static_cast<size_t>(n) * 32 * sizeof(float) <= 64000) ? 32 : 16;
}

// ---------------------------------------------------------------------------
// Vector kernels (fixed-order reductions)
// ---------------------------------------------------------------------------
constexpr int DOT_BS = 256;
constexpr int DOT_MAXB = 1024;

__global__ __launch_bounds__(DOT_BS) void k_dot_partial(const double* x, const double* y, long n,
                                                        double* partial, const double* act) {
  if (act && act[0] == 0.0) return;  // a finished solve: nothing to do (grid-uniform)
  __shared__ double sh[DOT_BS];
  double s = 0.0;
  for (long i = blockIdx.x * static_cast<long>(DOT_BS) + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * DOT_BS)
    s = fma(x[i], y[i], s);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = DOT_BS / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// partial is [nvec][nb]; out[v] = sum_b partial[v][b] (sqrt when sqrt_out)
__global__ __launch_bounds__(DOT_BS) void k_reduce(const double* partial, int nb, double* out,
                                                   int sqrt_out, const double* act) {
  if (act && act[0] == 0.0) return;
  __shared__ double sh[DOT_BS];
  const double* pv = partial + static_cast<long>(blockIdx.x) * nb;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += DOT_BS) s += pv[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = DOT_BS / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt_out ? sqrt(sh[0]) : sh[0];
}

// partial[v][b] for v < k: dot(V_v, w) -- blockIdx.y = v
__global__ __launch_bounds__(DOT_BS) void k_multidot_partial(const double* V, long ldv,
                                                             const double* w, long n, double* partial,
                                                             const double* act) {
  if (act && act[0] == 0.0) return;
  __shared__ double sh[DOT_BS];
  const double* x = V + ldv * blockIdx.y;
  double s = 0.0;
  for (long i = blockIdx.x * static_cast<long>(DOT_BS) + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * DOT_BS)
    s = fma(x[i], w[i], s);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = DOT_BS / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[static_cast<long>(blockIdx.y) * gridDim.x + blockIdx.x] = sh[0];
}

__global__ void k_axpy_dev(double* y, const double* x, const double* coef, double scale, long n,
                           const double* act) {
  if (act && act[0] == 0.0) return;
  const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (i < n) y[i] = fma(scale * coef[0], x[i], y[i]);
}

__global__ void k_scale_inv_dev(double* x, const double* coef, long n, const double* act) {
  if (act && act[0] == 0.0) return;
  const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (i < n) x[i] = x[i] * (1.0 / coef[0]);
}

__global__ void k_axpby(double* y, const double* x, double a, double b, long n) {
  const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  // b == 0: y is not read (it may be fresh memory holding any bit pattern, NaN included)
  if (i < n) y[i] = (b == 0.0) ? a * x[i] : a * x[i] + b * y[i];
}

// lowest index i < n with info[i] != 0 into out[0] (INT_MAX when there is none)
__global__ void k_first_init(int* out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0x7fffffff;
}
__global__ void k_first_nonzero(const int* info, long n, int* out) {
  const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (i < n && info[i] != 0) atomicMin(out, static_cast<int>(i));
}

__global__ void k_copy(double* y, const double* x, long n) {
  const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// x += alpha * V[:, 0:k] y   (y on the device)
// SKIPZ: columns whose coefficient is zero are not read (a basis column the solve has not
// written in this cycle may hold anything)
template <bool SKIPZ>
__global__ void k_gemv_cols(double* x, const double* V, long ldv, int k, const double* y, double alpha,
                            long n, const double* act) {
  if (act && act[0] == 0.0) return;
  const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) {
    const double yj = y[j];
    if (SKIPZ && yj == 0.0) continue;
    s = fma(V[i + ldv * j], yj, s);
  }
  x[i] = fma(alpha, s, x[i]);
}

template <class M>
void launch_points(const Dims& d, const AsmArgs& a, bool jac, bool vol, hipStream_t s) {
  const int bs = 256;
  if (vol) {
    long P = static_cast<long>(a.eb) * d.nge;
    if (jac) k_vol_points<M, true><<<nblocks(P, bs), bs, 0, s>>>(d, a);
    else k_vol_points<M, false><<<nblocks(P, bs), bs, 0, s>>>(d, a);
  } else {
    long P = static_cast<long>(a.eb) * d.nfe * d.ngf;
    if (jac) k_face_points<M, true><<<nblocks(P, bs), bs, 0, s>>>(d, a);
    else k_face_points<M, false><<<nblocks(P, bs), bs, 0, s>>>(d, a);
    if (a.nbf > 0) {
      long Pb = static_cast<long>(a.nbf) * d.ngf;
      if (jac) k_face_bou_points<M, true><<<nblocks(Pb, 64), 64, 0, s>>>(d, a);
      else k_face_bou_points<M, false><<<nblocks(Pb, 64), 64, 0, s>>>(d, a);
    }
  }
  EXA_HIP(hipGetLastError());
}

template <template <int, int> class MT>
void dispatch_points(const Dims& d, const AsmArgs& a, bool jac, bool vol, hipStream_t s) {
  if (d.nd == 2 && d.ncu == 1) return launch_points<MT<2, 1>>(d, a, jac, vol, s);
  if (d.nd == 3 && d.ncu == 1) return launch_points<MT<3, 1>>(d, a, jac, vol, s);
  if (d.nd == 2 && d.ncu == 2) return launch_points<MT<2, 2>>(d, a, jac, vol, s);
  if (d.nd == 3 && d.ncu == 2) return launch_points<MT<3, 2>>(d, a, jac, vol, s);
  throw Error(5, "model instantiated for nd in {2,3}, ncu in {1,2} only");
}

// inputs of a model library at the volume points of the block (the interpolation of k_vol_points)
template <class M>
__global__ __launch_bounds__(256) void k_plug_vol_in(Dims d, AsmArgs a) {
  constexpr int ND = M::ND, NC = M::NC;
  const long P = static_cast<long>(a.eb) * d.nge;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const long E = a.e0 + p / d.nge;
  const int g = static_cast<int>(p % d.nge);
  const long PG = static_cast<long>(d.ne) * d.nge;
  const long pg = E * d.nge + g;
  double uq[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) uq[c] = 0.0;
  for (int i = 0; i < d.npe; ++i) {
    const double phi = a.shapegt[g + d.nge * i];
#pragma unroll
    for (int c = 0; c < NC; ++c) uq[c] += phi * a.udg[i + d.npe * (c + NC * E)];
  }
  double* in = a.pbuf;
#pragma unroll
  for (int b = 0; b < ND; ++b) in[b * P + p] = a.egeo[b * PG + pg];
#pragma unroll
  for (int c = 0; c < NC; ++c) in[(ND + c) * P + p] = uq[c];
  for (int c = 0; c < d.nco; ++c) in[(ND + NC + c) * P + p] = a.og[c * PG + pg];
}

// inputs of a model library at the face points of the block (face_state): x, uq, uq with u
// replaced by uhat (the interior flux is f(uhat, q)), uhat, n, v
template <class M>
__global__ __launch_bounds__(256) void k_plug_face_in(Dims d, AsmArgs a) {
  constexpr int ND = M::ND, NCU = M::NCU, NC = M::NC;
  const long P = static_cast<long>(a.eb) * d.nfe * d.ngf;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const int g = static_cast<int>(p % d.ngf);
  const long el = p / d.ngf;
  const int l = static_cast<int>(el % d.nfe);
  const long E = a.e0 + el / d.nfe;
  double uh[NCU], uq[NC], x[ND], nl[ND], v[8], jac;
  face_state<M>(d, a, E, l, g, uh, uq, x, nl, v, jac);
  double* in = a.pbuf;
#pragma unroll
  for (int b = 0; b < ND; ++b) in[b * P + p] = x[b];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    in[(ND + c) * P + p] = uq[c];
    in[(ND + NC + c) * P + p] = c < NCU ? uh[c] : uq[c];
  }
#pragma unroll
  for (int c = 0; c < NCU; ++c) in[(ND + 2 * NC + c) * P + p] = uh[c];
#pragma unroll
  for (int b = 0; b < ND; ++b) in[(ND + 2 * NC + NCU + b) * P + p] = nl[b];
  for (int c = 0; c < d.nco && c < 8; ++c) in[(2 * ND + 2 * NC + NCU + c) * P + p] = v[c];
}

// The model library's functions over the block's points (one instance: its physicsparam and tau),
// on the solver's stream, then the point kernels reading their values.  The boundary function runs
// once per condition over all face points of the block (the point kernel picks each boundary
// point's condition), and only for blocks with boundary faces.
template <class M>
void plugin_launch(const Dims& d, const AsmArgs& a, bool jac, bool vol, hipStream_t s) {
  constexpr int ND = M::ND, NCU = M::NCU, NC = M::NC;
  const PluginFns& pf = *a.plugin;
  const int nco = d.nco;
  if (pf.set_stream) pf.set_stream(s);
  if (vol) {
    const long P = static_cast<long>(a.eb) * d.nge;
    k_plug_vol_in<M><<<nblocks(P, 256), 256, 0, s>>>(d, a);
    EXA_HIP(hipGetLastError());
    const double* x = a.pbuf;
    const double* uq = x + ND * P;
    const double* v = uq + NC * P;
    double* f = a.pbuf + M::vol_in(nco) * P;
    double* f_uq = f + NCU * ND * P;
    double* sv = f_uq + NCU * ND * NC * P;
    double* s_uq = sv + NCU * P;
    const int n = static_cast<int>(P);
    pf.flux(f, f_uq, nullptr, x, uq, v, nullptr, pf.uinf, a.param, a.time, pf.modelnumber, n, NC, NCU, ND, ND, nco,
            0);
    pf.source(sv, s_uq, nullptr, x, uq, v, nullptr, pf.uinf, a.param, a.time, pf.modelnumber, n, NC, NCU, ND, ND,
              nco, 0);
  } else {
    const long P = static_cast<long>(a.eb) * d.nfe * d.ngf;
    k_plug_face_in<M><<<nblocks(P, 256), 256, 0, s>>>(d, a);
    EXA_HIP(hipGetLastError());
    const double* x = a.pbuf;
    const double* uq = x + ND * P;
    const double* uqh = uq + NC * P;
    const double* uh = uqh + NC * P;
    const double* nl = uh + NCU * P;
    const double* v = nl + ND * P;
    double* f = a.pbuf + M::face_in(nco) * P;
    const int n = static_cast<int>(P);
    pf.flux(f, f + NCU * ND * P, nullptr, x, uqh, v, nullptr, pf.uinf, a.param, a.time, pf.modelnumber, n, NC, NCU,
            ND, ND, nco, 0);
    if (a.nbf > 0)
      for (int ib = 1; ib <= pf.nbc; ++ib) {
        double* fb = f + (NCU * ND + NCU * ND * NC + (ib - 1) * M::face_bc()) * P;
        pf.fbou(fb, fb + NCU * P, nullptr, fb + (NCU + NCU * NC) * P, x, uq, v, nullptr, uh, nl, a.tau, pf.uinf,
                a.param, a.time, pf.modelnumber, ib, n, NC, NCU, ND, ND, nco, 0);
      }
  }
  EXA_HIP(hipGetLastError());
  launch_points<M>(d, a, jac, vol, s);
}

void plugin_points(const Dims& d, const AsmArgs& a, bool jac, bool vol, hipStream_t s) {
  if (!a.plugin || !a.pbuf) throw Error(1, "model library not loaded");
  switch (d.nd * 10 + d.ncu) {
    case 21: return plugin_launch<PluginModel<2, 1>>(d, a, jac, vol, s);
    case 22: return plugin_launch<PluginModel<2, 2>>(d, a, jac, vol, s);
    case 23: return plugin_launch<PluginModel<2, 3>>(d, a, jac, vol, s);
    case 24: return plugin_launch<PluginModel<2, 4>>(d, a, jac, vol, s);
    case 25: return plugin_launch<PluginModel<2, 5>>(d, a, jac, vol, s);
    case 31: return plugin_launch<PluginModel<3, 1>>(d, a, jac, vol, s);
    case 32: return plugin_launch<PluginModel<3, 2>>(d, a, jac, vol, s);
    case 33: return plugin_launch<PluginModel<3, 3>>(d, a, jac, vol, s);
    case 34: return plugin_launch<PluginModel<3, 4>>(d, a, jac, vol, s);
    case 35: return plugin_launch<PluginModel<3, 5>>(d, a, jac, vol, s);
    default: throw Error(5, "model libraries: nd in {2, 3} and ncu in 1..5 only");
  }
}

// GetW (wequation.cpp:226-240): the w equation at every element node with the nodal (u, q), the
// nodal w as the first guess and the nodal wsrc; wdg is updated in place (one thread per node).
// A pass from a converged w still moves it in the last bits, so a batch passes a.wmask to update
// only the instances whose (u, q) just changed.
template <class M>
__global__ __launch_bounds__(256) void k_get_w(Dims d, AsmArgs a, const double* xdg, const double* odg,
                                               double* wdg) {
  constexpr int NCW = model_ncw<M>::value, NC = M::NC, ND = M::ND;
  const long P = static_cast<long>(d.ne) * d.npe;
  const long p = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (p >= P) return;
  const long E = p / d.npe;
  if (a.wmask && a.wmask[a.einst[E]] == 0.0) return;  // an instance whose state did not change
  const int i = static_cast<int>(p % d.npe);
  double x[ND], uq[NC], v[8], w[NCW], ws[NCW], w_uq[NCW * NC];
#pragma unroll
  for (int b = 0; b < ND; ++b) x[b] = xdg[i + d.npe * (b + ND * E)];
#pragma unroll
  for (int c = 0; c < NC; ++c) uq[c] = a.udg[i + d.npe * (c + NC * E)];
  for (int c = 0; c < d.nco && c < 8; ++c) v[c] = odg[i + d.npe * (c + d.nco * E)];
#pragma unroll
  for (int c = 0; c < NCW; ++c) {
    w[c] = wdg[i + d.npe * (c + NCW * E)];
    ws[c] = a.wsrc[i + d.npe * (c + NCW * E)];
  }
  w_point<M>(a, x, uq, v, inst_param(a, E), inst_time(a, E), ws, w, w_uq);
#pragma unroll
  for (int c = 0; c < NCW; ++c) wdg[i + d.npe * (c + NCW * E)] = w[c];
}

template <class G>
void launch_get_w(const Dims& d, const AsmArgs& a, const double* xdg, const double* odg, double* wdg,
                  hipStream_t s) {
  if constexpr (model_ncw<G>::value > 0) {
    const long P = static_cast<long>(d.ne) * d.npe;
    k_get_w<G><<<nblocks(P, 256), 256, 0, s>>>(d, a, xdg, odg, wdg);
    EXA_HIP(hipGetLastError());
  } else {
    throw Error(1, "the model has no w equation");
  }
}

void points(const Dims& d, const AsmArgs& a, bool jac, bool vol, hipStream_t s) {
  if (d.nco > 8) throw Error(5, "at most 8 auxiliary components");
  switch (a.model) {
    case MODEL_PLUGIN: return plugin_points(d, a, jac, vol, s);
    case MODEL_POISSON: return dispatch_points<Poisson>(d, a, jac, vol, s);
    case MODEL_POISSON_POLY: return dispatch_points<PoissonPoly>(d, a, jac, vol, s);
    case MODEL_CONVDIFF: return dispatch_points<ConvDiff>(d, a, jac, vol, s);
#define EXA_GEN_CASE(ID, NAME, G) \
  case ID:                        \
    return launch_points<G>(d, a, jac, vol, s);
      EXA_GENERATED_MODELS(EXA_GEN_CASE)
#undef EXA_GEN_CASE
    default: throw Error(1, "unknown model");
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
void w_nodes(const Dims& d, const AsmArgs& a, const double* xdg, const double* odg, double* wdg, hipStream_t s) {
  switch (a.model) {
#define EXA_GEN_W(ID, NAME, G) \
  case ID:                     \
    return launch_get_w<G>(d, a, xdg, odg, wdg, s);
    EXA_GENERATED_MODELS(EXA_GEN_W)
#undef EXA_GEN_W
    default: throw Error(1, "the model has no w equation");
  }
}

int model_ncw_of(int model) {
  switch (model) {
#define EXA_GEN_NCW(ID, NAME, G) \
  case ID:                       \
    return model_ncw<G>::value;
    EXA_GENERATED_MODELS(EXA_GEN_NCW)
#undef EXA_GEN_NCW
    default: return 0;
  }
}

long plugin_buffer_doubles(const Dims& d, int nbc, int eb) {
  const long nd = d.nd, ncu = d.ncu, nc = d.nc, nco = d.nco;
  const long pv = static_cast<long>(eb) * d.nge, pf = static_cast<long>(eb) * d.nfe * d.ngf;
  const long vol = (nd + nc + nco + ncu * nd + ncu * nd * nc + ncu + ncu * nc) * pv;
  const long face = (2 * nd + 2 * nc + ncu + nco + ncu * nd + ncu * nd * nc + nbc * (ncu + ncu * nc + ncu * ncu)) * pf;
  return std::max(vol, face);
}

int generated_model(const std::string& name, int* nd, int* ncu, int* nv) {
#define EXA_GEN_FIND(ID, NAME, G) \
  if (name == NAME) {             \
    *nd = G::ND;                  \
    *ncu = G::NCU;                \
    *nv = G::NV;                  \
    return ID;                    \
  }
  EXA_GENERATED_MODELS(EXA_GEN_FIND)
#undef EXA_GEN_FIND
  return -1;
}

// operands of a GEMM inside the LU: FP64 storage and products, or FP32 storage and products
inline void lu_gemm_ptrs(Gemm& g, const double* A, const double* B, double* C) {
  g.A = A;
  g.B = B;
  g.C = C;
}
inline void lu_gemm_ptrs(Gemm& g, const float* A, const float* B, float* C) {
  g.Af = A;
  g.Bf = B;
  g.Cf = C;
  g.fp32 = 1;
}

template <typename T, int NB>
void blocked_lu(T* A, int n, int ncols, long stride, int batch, int* info, int* pivout, hipStream_t s,
                int* piv_ws, T* tri_ws) {
  // X[r0 : r0+nb, cols] = Tinv X[r0 : r0+nb, cols] in place (Tinv in tri_ws, ld NB)
  auto apply_tri = [&](int nb, int r0, long c0, int ncol) {
    Gemm g{};
    g.M = nb; g.N = ncol; g.K = nb; g.alpha = 1.0; g.beta = 0.0;
    T* X = A + r0 + static_cast<long>(n) * c0;
    lu_gemm_ptrs(g, tri_ws, X, X);
    g.lda = NB; g.sA0 = static_cast<long>(NB) * NB;
    g.ldb = n; g.sB0 = stride;
    g.ldc = n; g.sC0 = stride;
    g.nb0 = batch; g.nb1 = 1; g.nb2 = 1;
    gemm(g, s);
  };
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = std::min(NB, n - k0);
    if (piv_ws && n - k0 <= 256) {
      k_lu_panel_reg<T, NB><<<batch, 256, 0, s>>>(A, n, stride, k0, info, pivout, piv_ws);
      const int right = ncols - k0 - nb;
      if (right > 0) {
        k_lu_swap<T, NB><<<dim3(batch, (right + 63) / 64), 64, 0, s>>>(A, n, ncols, stride, k0, piv_ws);
        if (tri_ws) {
          k_lu_tri_inv<T, NB, true><<<batch, 64, 0, s>>>(A, n, stride, k0, tri_ws);
          apply_tri(nb, k0, k0 + nb, right);
        } else {
          k_lu_u12<T, NB><<<dim3(batch, (right + 255) / 256), 256, 0, s>>>(A, n, ncols, stride, k0, piv_ws);
        }
      }
    } else {
      size_t shm = static_cast<size_t>(n - k0) * nb * sizeof(T);
      k_lu_panel<T, NB><<<batch, 256, shm, s>>>(A, n, ncols, stride, k0, info, pivout);
    }
    EXA_HIP(hipGetLastError());
    if (k0 + nb < n) {  // A22 -= L21 * U12
      Gemm g{};
      g.M = n - k0 - nb; g.N = ncols - k0 - nb; g.K = nb; g.alpha = -1.0; g.beta = 1.0;
      lu_gemm_ptrs(g, A + (k0 + nb) + static_cast<long>(n) * k0, A + k0 + static_cast<long>(n) * (k0 + nb),
                   A + (k0 + nb) + static_cast<long>(n) * (k0 + nb));
      g.lda = n; g.sA0 = stride;
      g.ldb = n; g.sB0 = stride;
      g.ldc = n; g.sC0 = stride;
      g.nb0 = batch; g.nb1 = 1; g.nb2 = 1;
      gemm(g, s);
    }
  }
  const int last = ((n - 1) / NB) * NB;
  for (int k0 = last; k0 >= 0; k0 -= NB) {
    if (tri_ws) {
      k_lu_tri_inv<T, NB, false><<<batch, 64, 0, s>>>(A, n, stride, k0, tri_ws);
      EXA_HIP(hipGetLastError());
      apply_tri(std::min(NB, n - k0), k0, n, ncols - n);
    } else {
      dim3 grid(static_cast<unsigned>(batch), static_cast<unsigned>((ncols - n + 255) / 256));
      k_lu_back<T, NB><<<grid, 256, 0, s>>>(A, n, ncols, stride, k0);
      EXA_HIP(hipGetLastError());
    }
    if (k0 > 0) {  // X[0:k0] -= U[0:k0, k0:k0+nb] * X[k0:k0+nb]
      const int nb = std::min(NB, n - k0);
      Gemm g{};
      g.M = k0; g.N = ncols - n; g.K = nb; g.alpha = -1.0; g.beta = 1.0;
      lu_gemm_ptrs(g, A + static_cast<long>(n) * k0, A + k0 + static_cast<long>(n) * n, A + static_cast<long>(n) * n);
      g.lda = n; g.sA0 = stride;
      g.ldb = n; g.sB0 = stride;
      g.ldc = n; g.sC0 = stride;
      g.nb0 = batch; g.nb1 = 1; g.nb2 = 1;
      gemm(g, s);
    }
  }
}

template <typename T>
void lu_blocked_any(T* A, int n, int ncols, long stride, int batch, int* info, hipStream_t s, int* pivout,
                    int* piv_ws, T* tri_ws) {
  // 32-wide panels when an n x 32 panel fits the 64 KB of LDS, else 16-wide; with a pivot
  // workspace, panels of up to 256 rows are factored in registers (64-wide register panels
  // were measured: no gain at n = 250, slower at n = 100)
  const int w = sizeof(T) == 4 ? lu_panel_width_f32(n) : lu_panel_width(n);
  if constexpr (sizeof(T) == 4) {
    if (w == 64) {
      blocked_lu<T, 64>(A, n, ncols, stride, batch, info, pivout, s, piv_ws, tri_ws);
      return;
    }
  }
  if (w == 32)
    blocked_lu<T, 32>(A, n, ncols, stride, batch, info, pivout, s, piv_ws, tri_ws);
  else if (static_cast<size_t>(n) * 16 * sizeof(T) <= 64000)
    blocked_lu<T, 16>(A, n, ncols, stride, batch, info, pivout, s, piv_ws, tri_ws);
  else
    throw Error(5, "dense block larger than 500 rows");
}

void lu_solve(double* A, int n, int ncols, long stride, int batch, int* info, hipStream_t s, int* pivout,
              int* piv_ws, void* tri_ws) {
  if (batch <= 0) return;
  // the FP64 form keeps the substitution kernels (k_lu_u12, k_lu_back): explicit triangle inverses
  // amplify rounding by the triangle's condition number, which the FP64 form is not allowed to lose
  (void)tri_ws;
  if ((n >= BLU_MIN_N || pivout) && ncols > n) {
    lu_blocked_any(A, n, ncols, stride, batch, info, s, pivout, piv_ws, static_cast<double*>(nullptr));
    return;
  }
  if (pivout) throw Error(5, "pivot output needs right-hand-side columns");
  // dynamic LDS panel n x LU_NB doubles next to ~12 KB of static LDS, kept within 64 KB
  if (n > 400) throw Error(5, "dense block larger than 400 rows");
  size_t shm = static_cast<size_t>(n) * LU_NB * sizeof(double);
  k_lu_solve<<<batch, 256, shm, s>>>(A, n, ncols, stride, info);
  EXA_HIP(hipGetLastError());
}

void lu_solve(float* A, int n, int ncols, long stride, int batch, int* info, hipStream_t s, int* pivout,
              int* piv_ws, void* tri_ws) {
  if (batch <= 0) return;
  if (ncols <= n) throw Error(5, "FP32 LU needs right-hand-side columns");
  lu_blocked_any(A, n, ncols, stride, batch, info, s, pivout, piv_ws, static_cast<float*>(tri_ws));
}

void elem_geom(const Dims& d, const double* shapegt, const double* xdg, double* egeo, hipStream_t s) {
  long P = static_cast<long>(d.ne) * d.nge;
  k_elem_geom<<<nblocks(P, 256), 256, 0, s>>>(d, shapegt, xdg, egeo);
  EXA_HIP(hipGetLastError());
}

void face_geom(const Dims& d, const double* shapfgt, const int* perm, const double* xdg, double* fgeo,
               hipStream_t s) {
  long P = static_cast<long>(d.ne) * d.nfe * d.ngf;
  k_face_geom<<<nblocks(P, 256), 256, 0, s>>>(d, shapfgt, perm, xdg, fgeo);
  EXA_HIP(hipGetLastError());
}

void interp_elem(const Dims& d, int ncomp, const double* shapegt, const double* src, double* dst,
                 hipStream_t s) {
  if (ncomp <= 0) return;
  long P = static_cast<long>(d.ne) * d.nge;
  k_interp_elem<<<nblocks(P, 256), 256, 0, s>>>(d, ncomp, shapegt, src, dst);
  EXA_HIP(hipGetLastError());
}

void interp_face(const Dims& d, int ncomp, const double* shapfgt, const int* perm, const double* src,
                 double* dst, hipStream_t s) {
  if (ncomp <= 0) return;
  long P = static_cast<long>(d.ne) * d.nfe * d.ngf;
  k_interp_face<<<nblocks(P, 256), 256, 0, s>>>(d, ncomp, shapfgt, perm, src, dst);
  EXA_HIP(hipGetLastError());
}

void qsetup(const Dims& d, int e0, int eb, const double* shapegt, const double* shapegw,
            const double* shapfgt, const double* shapfgw, const int* perm, const double* egeo,
            const double* fgeo, double* aug, double* Mass, hipStream_t s, double* Aq) {
  if (Aq) {  // [Mass | C_1 .. C_nd] = Aq_e Phi^T on MFMA (Aq: eb * (nd + 1) * npe * nge doubles)
    const long per = static_cast<long>(d.npe) * d.nge;
    k_qsetup_a<<<nblocks(per * eb, 256), 256, 0, s>>>(d, e0, eb, shapegw, egeo, Aq);
    EXA_HIP(hipGetLastError());
    const long ncq = static_cast<long>(d.npe) + d.npe * d.nd + static_cast<long>(d.ndf) * d.nd;
    Gemm g{};
    g.M = d.npe; g.N = d.npe; g.K = d.nge; g.alpha = 1.0; g.beta = 0.0;
    g.A = Aq; g.lda = d.npe; g.sA0 = per; g.sA1 = per * (d.nd + 1);
    g.B = shapegt; g.ldb = d.nge; g.sB0 = 0; g.sB1 = 0;
    g.C = aug; g.ldc = d.npe; g.sC0 = static_cast<long>(d.npe) * d.npe; g.sC1 = d.npe * ncq;
    g.nb0 = d.nd + 1; g.nb1 = eb; g.nb2 = 1;
    gemm(g, s);
    const long pm = static_cast<long>(d.npe) * d.npe;
    k_qmass<<<nblocks(pm * eb, 256), 256, 0, s>>>(d, e0, eb, aug, Mass);
  } else {
    long n1 = static_cast<long>(d.npe) * (d.npe + d.npe * d.nd) * eb;
    k_qsetup_mc<<<nblocks(n1, 256), 256, 0, s>>>(d, e0, eb, shapegt, shapegw, egeo, aug, Mass);
  }
  long n2 = static_cast<long>(d.nfe) * d.npf * d.npf * d.nd * eb;
  k_qsetup_e<<<nblocks(n2, 256), 256, 0, s>>>(d, e0, eb, shapfgt, shapfgw, perm, fgeo, aug);
  EXA_HIP(hipGetLastError());
}

void qpack(const Dims& d, int e0, int eb, const double* aug, double* MCs, double* MEs, hipStream_t s) {
  long n1 = static_cast<long>(d.npe) * d.nd * (d.npe + d.ndf) * eb;
  k_qpack<<<nblocks(n1, 256), 256, 0, s>>>(d, e0, eb, aug, MCs, MEs);
  EXA_HIP(hipGetLastError());
}

void getq(const Dims& d, const double* MCs, const double* MEs, const int* elemcon, const double* uh,
          double* udg, hipStream_t s) {
  const long n1 = static_cast<long>(d.npe) * d.nd * d.ne;
  // one workgroup per element: the rows of M^-1 C rounded up to whole waves (at most 1024 threads)
  const int nt = static_cast<int>(std::min<long>(1024, (static_cast<long>(d.npe) * d.nd + 63) / 64 * 64));
  const size_t shm = static_cast<size_t>(d.npe + d.ndf) * d.ncu * sizeof(double);
  const unsigned ne = static_cast<unsigned>(d.ne);
  if (d.ne <= 0) return;
  switch (shm <= 48 * 1024 ? d.ncu : 0) {
    case 1: k_getq<1><<<ne, nt, shm, s>>>(d, MCs, MEs, elemcon, uh, udg); break;
    case 2: k_getq<2><<<ne, nt, shm, s>>>(d, MCs, MEs, elemcon, uh, udg); break;
    case 3: k_getq<3><<<ne, nt, shm, s>>>(d, MCs, MEs, elemcon, uh, udg); break;
    case 4: k_getq<4><<<ne, nt, shm, s>>>(d, MCs, MEs, elemcon, uh, udg); break;
    case 5: k_getq<5><<<ne, nt, shm, s>>>(d, MCs, MEs, elemcon, uh, udg); break;
    default:
      k_getq_any<<<nblocks(n1 * d.ncu, 256), 256, 0, s>>>(d, MCs, MEs, elemcon, uh, udg);
  }
  EXA_HIP(hipGetLastError());
}

void vol_points(const Dims& d, const AsmArgs& a, bool jac, hipStream_t s) { points(d, a, jac, true, s); }
void face_points(const Dims& d, const AsmArgs& a, bool jac, hipStream_t s) { points(d, a, jac, false, s); }

template <typename T>
void vol_res_t(const Dims& d, const AsmArgs& a, T* out, long ldo, hipStream_t s) {
  if (a.eb <= 0) return;
  const size_t shm = static_cast<size_t>(d.nd + 1) * d.ncu * d.nge * sizeof(double);
  if (shm > 48 * 1024) {
    const long n1 = static_cast<long>(d.npe) * d.ncu * a.eb;
    k_vol_res_direct<T><<<nblocks(n1, 256), 256, 0, s>>>(d, a, out, ldo);
  } else {
    const int nt = std::min(256, (d.npe * d.ncu + 63) / 64 * 64);
    k_vol_res<T><<<static_cast<unsigned>(a.eb), nt, shm, s>>>(d, a, out, ldo);
  }
  EXA_HIP(hipGetLastError());
}
void vol_res(const Dims& d, const AsmArgs& a, double* out, long ldo, hipStream_t s) { vol_res_t(d, a, out, ldo, s); }
void vol_res(const Dims& d, const AsmArgs& a, float* out, long ldo, hipStream_t s) { vol_res_t(d, a, out, ldo, s); }

template <typename T>
void vol_W_t(const Dims& d, const AsmArgs& a, T* W, hipStream_t s) {
  if (d.nd > 3) throw Error(5, "volume Jacobian for nd > 3");
  const int gp = std::max(1, 256 / d.npe);  // points per workgroup
  const long nwg = static_cast<long>(a.eb) * ((d.nge + gp - 1) / gp);
  const size_t shm = static_cast<size_t>(gp) * (d.nd + 1) * d.ncu * d.nc * sizeof(double);
  if (shm > 64 * 1024) throw Error(5, "volume Jacobian: too many components for the LDS");
  k_vol_W<T><<<static_cast<unsigned>(nwg), 256, shm, s>>>(d, a, W, gp);
  EXA_HIP(hipGetLastError());
}
void vol_W(const Dims& d, const AsmArgs& a, hipStream_t s) { vol_W_t(d, a, a.W, s); }
void vol_W(const Dims& d, const AsmArgs& a, float* W, hipStream_t s) { vol_W_t(d, a, W, s); }

template <typename T>
void face_res_t(const Dims& d, const AsmArgs& a, int l, bool jac, T* ru, long ldru, double* Rh, hipStream_t s) {
  int nf1 = jac ? (d.ncu + d.ncu * d.nc + d.ncu * d.ncu) : d.ncu;
  long n1 = static_cast<long>(d.npf) * d.ncu * a.eb;
  k_face_res<T><<<nblocks(n1, 256), 256, 0, s>>>(d, a, l, nf1, ru, ldru, Rh);
  EXA_HIP(hipGetLastError());
}
void face_res(const Dims& d, const AsmArgs& a, int l, bool jac, double* ru, long ldru, double* Rh, hipStream_t s) {
  face_res_t(d, a, l, jac, ru, ldru, Rh, s);
}
void face_res(const Dims& d, const AsmArgs& a, int l, bool jac, float* ru, long ldru, double* Rh, hipStream_t s) {
  face_res_t(d, a, l, jac, ru, ldru, Rh, s);
}

template <typename T>
void face_jac_t(const Dims& d, const AsmArgs& a, T* aug, T* Bfull, T* Kt, T* Gfull, T* H, hipStream_t s) {
  if (a.eb <= 0) return;
  const size_t lds = face_jac_elem_lds(d);
  if (lds <= 64000) {
    k_face_jac_elem<T><<<a.eb, FJ_THREADS, lds, s>>>(d, a, aug, Bfull, Kt, Gfull, H);
    EXA_HIP(hipGetLastError());
    return;
  }
  // face tables too large for LDS: one launch per local face (no two writes of a launch collide)
  long n1 = static_cast<long>(d.npf) * d.npf * d.ncu * (d.nc + d.ncu) * a.eb;
  for (int l = 0; l < d.nfe; ++l) {
    k_face_jac<T><<<nblocks(n1, 256), 256, 0, s>>>(d, a, l, aug, Bfull, Kt, Gfull, H);
    EXA_HIP(hipGetLastError());
  }
}
void face_jac(const Dims& d, const AsmArgs& a, double* aug, double* Bfull, double* Kt, double* Gfull, double* H,
              hipStream_t s) {
  face_jac_t(d, a, aug, Bfull, Kt, Gfull, H, s);
}
void face_jac(const Dims& d, const AsmArgs& a, float* aug, float* Bfull, float* Kt, float* Gfull, float* H,
              hipStream_t s) {
  face_jac_t(d, a, aug, Bfull, Kt, Gfull, H, s);
}

void post_schur(const Dims& d, int e0, int eb, const double* aug, long augs, double* DinvF, double* RuS,
                hipStream_t s) {
  const long per = static_cast<long>(d.n) * (d.m + 1);
  if (eb <= 0) return;
  if (eb <= 65535)
    k_post_schur_2d<double, double><<<dim3(nblocks(per, 256), eb), 256, 0, s>>>(d, e0, aug, augs, DinvF, RuS);
  else
    k_post_schur<double, double><<<nblocks(per * eb, 256), 256, 0, s>>>(d, e0, eb, aug, augs, DinvF, RuS);
  EXA_HIP(hipGetLastError());
}

void post_schur(const Dims& d, int e0, int eb, const float* aug, long augs, float* DinvF, double* RuS,
                hipStream_t s) {
  const long per = static_cast<long>(d.n) * (d.m + 1);
  if (eb <= 0) return;
  if (eb <= 65535)
    k_post_schur_2d<float, float><<<dim3(nblocks(per, 256), eb), 256, 0, s>>>(d, e0, aug, augs, DinvF, RuS);
  else
    k_post_schur<float, float><<<nblocks(per * eb, 256), 256, 0, s>>>(d, e0, eb, aug, augs, DinvF, RuS);
  EXA_HIP(hipGetLastError());
}

void u_axpby(const Dims& d, const double* udg, double a, double b, double* dst, hipStream_t s) {
  long n1 = static_cast<long>(d.npe) * d.ncu * d.ne;
  k_u_axpby<<<nblocks(n1, 256), 256, 0, s>>>(d, udg, a, b, dst);
  EXA_HIP(hipGetLastError());
}

void u_insert(const Dims& d, const double* src, double* udg, hipStream_t s) {
  long n1 = static_cast<long>(d.npe) * d.ncu * d.ne;
  k_u_insert<<<nblocks(n1, 256), 256, 0, s>>>(d, src, udg);
  EXA_HIP(hipGetLastError());
}

void to_float(float* dst, const double* src, int rows, int cols, long ld, long sstride, long dstride, int batch,
              hipStream_t s) {
  long n1 = static_cast<long>(rows) * cols * batch;
  if (n1 == 0) return;
  const long per = static_cast<long>(rows) * cols;
  if (ld == rows && batch <= 65535)
    k_to_float_2d<double><<<dim3(nblocks(per, 256), batch), 256, 0, s>>>(dst, src, per, sstride, dstride);
  else
    k_to_float<double><<<nblocks(n1, 256), 256, 0, s>>>(dst, src, rows, cols, ld, sstride, dstride, batch);
  EXA_HIP(hipGetLastError());
}

void to_float(float* dst, const float* src, int rows, int cols, long ld, long sstride, long dstride, int batch,
              hipStream_t s) {
  long n1 = static_cast<long>(rows) * cols * batch;
  if (n1 == 0) return;
  const long per = static_cast<long>(rows) * cols;
  if (ld == rows && batch <= 65535)
    k_to_float_2d<float><<<dim3(nblocks(per, 256), batch), 256, 0, s>>>(dst, src, per, sstride, dstride);
  else
    k_to_float<float><<<nblocks(n1, 256), 256, 0, s>>>(dst, src, rows, cols, ld, sstride, dstride, batch);
  EXA_HIP(hipGetLastError());
}

// BF16 (round to nearest even of the FP32 value), kept as unsigned short
__device__ __forceinline__ unsigned short bf16_round(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}
// four values per thread (the count's tail one by one)
template <typename S>
__global__ void k_to_bf16(unsigned short* dst, const S* src, long count) {
  const long q = (blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x) * 4;
  if (q + 4 <= count) {
    S v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = src[q + k];
    ushort4 o;
    o.x = bf16_round(static_cast<float>(v[0]));
    o.y = bf16_round(static_cast<float>(v[1]));
    o.z = bf16_round(static_cast<float>(v[2]));
    o.w = bf16_round(static_cast<float>(v[3]));
    if ((reinterpret_cast<uintptr_t>(dst + q) & 7) == 0) {
      *reinterpret_cast<ushort4*>(dst + q) = o;
    } else {
      dst[q] = o.x;
      dst[q + 1] = o.y;
      dst[q + 2] = o.z;
      dst[q + 3] = o.w;
    }
  } else {
    for (long k = q; k < count; ++k) dst[k] = bf16_round(static_cast<float>(src[k]));
  }
}

template <typename S>
void to_bf16_t(unsigned short* dst, const S* src, long count, hipStream_t s) {
  if (count <= 0) return;
  k_to_bf16<S><<<nblocks((count + 3) / 4, 256), 256, 0, s>>>(dst, src, count);
  EXA_HIP(hipGetLastError());
}
void to_bf16(unsigned short* dst, const double* src, long count, hipStream_t s) { to_bf16_t(dst, src, count, s); }
void to_bf16(unsigned short* dst, const float* src, long count, hipStream_t s) { to_bf16_t(dst, src, count, s); }

void refine_rhs(const Dims& d, const float* LU, const int* piv, const float* K, const double* Ru, double* RuS,
                double* Rh, hipStream_t s) {
  size_t shm = static_cast<size_t>(d.n) * sizeof(double);
  k_refine_rhs<<<d.ne, 256, shm, s>>>(d, lu_panel_width_f32(d.n), LU, piv, K, Ru, RuS, Rh);
  EXA_HIP(hipGetLastError());
}

void face_gather(long nnodes, int ncu, const long* src1, const long* src2, const double* we, double* w,
                 hipStream_t s) {
  k_face_gather<<<nblocks(nnodes * ncu, 256), 256, 0, s>>>(nnodes, ncu, src1, src2, we, w);
  EXA_HIP(hipGetLastError());
}

template <typename TH>
void bj_assemble_t(const Dims& d, const int* f2e, const int* elemcon, const TH* H, double* aug, hipStream_t s) {
  long n1 = static_cast<long>(d.ncf) * 2 * d.ncf * d.nf;
  k_bj_assemble<TH><<<nblocks(n1, 256), 256, 0, s>>>(d, f2e, elemcon, H, aug);
  EXA_HIP(hipGetLastError());
}
void bj_assemble(const Dims& d, const int* f2e, const int* elemcon, const double* H, double* aug,
                 hipStream_t s) {
  bj_assemble_t(d, f2e, elemcon, H, aug, s);
}
void bj_assemble(const Dims& d, const int* f2e, const int* elemcon, const float* H, double* aug, hipStream_t s) {
  bj_assemble_t(d, f2e, elemcon, H, aug, s);
}

template <typename TH, typename TA>
void eas_assemble_t(const Dims& d, int e0, int eb, const TH* H, const int* f2e, const int* elemcon, TA* aug,
                    hipStream_t s) {
  long n1 = static_cast<long>(d.m) * 2 * d.m * eb;
  k_eas_assemble<TH, TA><<<nblocks(n1, 256), 256, 0, s>>>(d, e0, eb, H, f2e, elemcon, aug);
  EXA_HIP(hipGetLastError());
}
void eas_assemble(const Dims& d, int e0, int eb, const double* H, const int* f2e, const int* elemcon,
                  double* aug, hipStream_t s) {
  eas_assemble_t(d, e0, eb, H, f2e, elemcon, aug, s);
}
void eas_assemble(const Dims& d, int e0, int eb, const float* H, const int* f2e, const int* elemcon,
                  double* aug, hipStream_t s) {
  eas_assemble_t(d, e0, eb, H, f2e, elemcon, aug, s);
}
void eas_assemble(const Dims& d, int e0, int eb, const float* H, const int* f2e, const int* elemcon,
                  float* aug, hipStream_t s) {
  eas_assemble_t(d, e0, eb, H, f2e, elemcon, aug, s);
}

template <typename T>
void eas_extract_t(const Dims& d, int eb, const T* aug, T* BEe, hipStream_t s) {
  long n1 = static_cast<long>(d.m) * d.m * eb;
  k_eas_extract<T><<<nblocks(n1, 256), 256, 0, s>>>(d, eb, aug, BEe);
  EXA_HIP(hipGetLastError());
}
void eas_extract(const Dims& d, int eb, const double* aug, double* BEe, hipStream_t s) {
  eas_extract_t(d, eb, aug, BEe, s);
}
void eas_extract(const Dims& d, int eb, const float* aug, float* BEe, hipStream_t s) {
  eas_extract_t(d, eb, aug, BEe, s);
}

void bj_extract(const Dims& d, const double* aug, double* BE, hipStream_t s) {
  long n1 = static_cast<long>(d.ncf) * d.ncf * d.nf;
  k_bj_extract<<<nblocks(n1, 256), 256, 0, s>>>(d, aug, BE);
  EXA_HIP(hipGetLastError());
}

template <typename T>
void bj_apply(const Dims& d, const T* BE, const double* x, double* y, hipStream_t s, const double* act) {
  long n1 = static_cast<long>(d.ncf) * d.nf;
  k_bj_apply<T><<<nblocks(n1, 256), 256, 0, s>>>(d, BE, x, y, act);
  EXA_HIP(hipGetLastError());
}

template <typename T>
void matvec_elem(const Dims& d, const T* H, const int* elemcon, const double* v, double* we, hipStream_t s,
                 const int* einst, const double* emask) {
  if (d.m > 1024) throw Error(5, "element trace block larger than 1024");
  if (d.m <= MATVEC_SMALL_M) {
    k_matvec_small<T><<<nblocks(static_cast<long>(d.m) * d.ne, 256), 256, 0, s>>>(d, H, elemcon, v, we, einst,
                                                                                   emask);
    EXA_HIP(hipGetLastError());
    return;
  }
  size_t shm = static_cast<size_t>(5) * d.m * sizeof(double);
  constexpr int VMAX = 16 / sizeof(T);  // 16-byte loads: 2 doubles, 4 floats or 8 BF16 per lane
  if (d.m % VMAX == 0) {
    const int nrb = (d.m + 64 * VMAX - 1) / (64 * VMAX);
    if (nrb == 2) k_matvec_elem<T, VMAX, 2><<<d.ne, 256, shm, s>>>(d, H, elemcon, v, we, einst, emask);
    else if (nrb == 3) k_matvec_elem<T, VMAX, 3><<<d.ne, 256, shm, s>>>(d, H, elemcon, v, we, einst, emask);
    else k_matvec_elem<T, VMAX, 0><<<d.ne, 256, shm, s>>>(d, H, elemcon, v, we, einst, emask);
  } else if (sizeof(T) == 2 && d.m % 4 == 0) {  // BF16 blocks: 8-byte loads
    k_matvec_elem<T, 4, 0><<<d.ne, 256, shm, s>>>(d, H, elemcon, v, we, einst, emask);
  } else if (d.m % 2 == 0) {
    k_matvec_elem<T, 2, 0><<<d.ne, 256, shm, s>>>(d, H, elemcon, v, we, einst, emask);
  } else {
    k_matvec_elem<T, 1, 0><<<d.ne, 256, shm, s>>>(d, H, elemcon, v, we, einst, emask);
  }
  EXA_HIP(hipGetLastError());
}

template <typename T>
void dudg(const Dims& d, const double* RuS, const T* DinvF, const int* elemcon, const double* x,
          double* du, hipStream_t s) {
  if (d.ne <= 0) return;
  if (d.m > 8192) throw Error(5, "element trace block larger than 8192 (the LDS of k_dudg)");
  const int nt = std::min(1024, (d.n + 63) / 64 * 64);  // one workgroup per element, a thread per row
  k_dudg<T><<<static_cast<unsigned>(d.ne), nt, static_cast<size_t>(d.m) * sizeof(double), s>>>(d, RuS, DinvF,
                                                                                             elemcon, x, du);
  EXA_HIP(hipGetLastError());
}

void update_udg(const Dims& d, const double* du, double alpha, double* udg, hipStream_t s) {
  long n1 = static_cast<long>(d.npe) * d.ncu * d.ne;
  k_update_udg<<<nblocks(n1, 256), 256, 0, s>>>(d, du, alpha, udg);
  EXA_HIP(hipGetLastError());
}

template <typename T, typename TH>
void bilu_assemble_t(const Dims& d, const IluDev& I, const int* f2e, const int* elemcon, const TH* H, T* BE,
                     hipStream_t s) {
  if (d.ncf > 256 || I.nnz > 65535) throw Error(5, "block-ILU(0) assembly grid too large");
  k_bilu_assemble<T, TH><<<dim3(I.nse, I.nnz), 256, 0, s>>>(d, I, f2e, elemcon, H, BE);
  EXA_HIP(hipGetLastError());
}

template <typename T>
void bilu_invert_diag_t(const Dims& d, const IluDev& I, int i, T* BE, int* info, hipStream_t s) {
  // the FP64 bound: FP32 blocks use the same limit, so both forms accept the same problems
  size_t shm = (2 * static_cast<size_t>(d.ncf) * d.ncf + d.ncf) * sizeof(double);
  if (shm > 60 * 1024) throw Error(5, "block-ILU(0) face blocks larger than 61 x 61");
  shm = (2 * static_cast<size_t>(d.ncf) * d.ncf + d.ncf) * sizeof(T);
  k_bilu_inv<T><<<I.nse, 256, shm, s>>>(d, I, i, BE, info);
  EXA_HIP(hipGetLastError());
}

template <typename T>
void bilu_put_l_t(const Dims& d, const IluDev& I, int k0, int ne_i, const T* Ltmp, T* BE, hipStream_t s) {
  long n1 = static_cast<long>(d.ncf) * d.ncf * I.nse * ne_i;
  if (n1 == 0) return;
  k_bilu_put_l<T><<<nblocks(n1, 256), 256, 0, s>>>(d, I, k0, ne_i, Ltmp, BE);
  EXA_HIP(hipGetLastError());
}

void bilu_assemble(const Dims& d, const IluDev& I, const int* f2e, const int* elemcon, const double* H,
                   double* BE, hipStream_t s) {
  bilu_assemble_t(d, I, f2e, elemcon, H, BE, s);
}
void bilu_assemble(const Dims& d, const IluDev& I, const int* f2e, const int* elemcon, const float* H,
                   float* BE, hipStream_t s) {
  bilu_assemble_t(d, I, f2e, elemcon, H, BE, s);
}
void bilu_invert_diag(const Dims& d, const IluDev& I, int i, double* BE, int* info, hipStream_t s) {
  bilu_invert_diag_t(d, I, i, BE, info, s);
}
void bilu_invert_diag(const Dims& d, const IluDev& I, int i, float* BE, int* info, hipStream_t s) {
  bilu_invert_diag_t(d, I, i, BE, info, s);
}
void bilu_put_l(const Dims& d, const IluDev& I, int k0, int ne_i, const double* Ltmp, double* BE,
                hipStream_t s) {
  bilu_put_l_t(d, I, k0, ne_i, Ltmp, BE, s);
}
void bilu_put_l(const Dims& d, const IluDev& I, int k0, int ne_i, const float* Ltmp, float* BE, hipStream_t s) {
  bilu_put_l_t(d, I, k0, ne_i, Ltmp, BE, s);
}

// the block-ILU(0) apply for BF16 factors: one wave per superelement (1) or the 256-thread kernel
// (0); -1 follows EXA_BILU_WAVE (default 1)
std::atomic<int> g_bilu_wave{-1};
void set_bilu_wave(int mode) { g_bilu_wave.store(mode); }
bool bilu_wave_enabled() {
  const int m = g_bilu_wave.load();
  if (m >= 0) return m != 0;
  static const bool env_on = [] {
    const char* e = std::getenv("EXA_BILU_WAVE");
    return !e || std::atoi(e) != 0;
  }();
  return env_on;
}

template <typename T>
void bilu_apply(const Dims& d, const IluDev& I, const T* BE, const double* x, double* Y, double* Xo,
                double* out, hipStream_t s, const double* act) {
  (void)Y;
  if (d.ncf > 256) throw Error(5, "block-ILU(0) face blocks larger than 256");
  // two rows per thread (8-byte FP32 / 16-byte FP64 loads) when the block rows pair up
  const int vw = (d.ncf % 2 == 0) ? 2 : 1;
  const size_t vt = sizeof(T) == 8 ? sizeof(double) : sizeof(float);  // the kernel's LDS vector type
  const size_t shm = 256 * vw * sizeof(double) + (static_cast<size_t>(I.nfse) * d.ncf + d.ncf) * vt +
                     (3 * static_cast<size_t>(I.nfse) + 2 + 4 * static_cast<size_t>(I.nnz)) * sizeof(int);
  if (shm > 160 * 1024) throw Error(5, "block-ILU(0) superelement vector larger than the LDS");
  auto launch = [&](auto kern) {
    if (shm > 64 * 1024)
      EXA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(shm)));
    kern<<<I.nse, 256, shm, s>>>(d, I, BE, x, Xo, act);
  };
  // one wave per superelement for FP32 / BF16 factors when a wave covers a block's rows
  const size_t shm_w = (static_cast<size_t>(I.nfse) * d.ncf + d.ncf) * sizeof(float);
  bool wave = false;
  if constexpr (sizeof(T) == 2) {
    // the one-wave kernel's shapes: ncf = 50 (p = 4, two components on hex faces): 25 row pairs,
    // 2 parts of 25 columns
    if (d.ncf == 50 && shm_w <= 64 * 1024 && I.wseg && bilu_wave_enabled()) {
      wave = true;
      k_bilu_apply_wave<T, 2, 50><<<I.nse, 64, shm_w, s>>>(d, I, BE, x, Xo, act);
    }
  }
  if (wave) {
  } else if (vw == 2) {
    launch(&k_bilu_apply<T, 2>);
  } else {
    launch(&k_bilu_apply<T, 1>);
  }
  k_bilu_scatter<<<nblocks(static_cast<long>(d.nf) * d.ncf, 256), 256, 0, s>>>(d, I, Xo, out, act);
  EXA_HIP(hipGetLastError());
}

int dot_partials_count(long n) {
  long b = (n + DOT_BS - 1) / DOT_BS;
  return static_cast<int>(std::max(1L, std::min<long>(DOT_MAXB, b)));
}

void dot(const double* x, const double* y, long n, double* partial, double* out, int sqrt_out,
         hipStream_t s, const double* act) {
  int nb = dot_partials_count(n);
  k_dot_partial<<<nb, DOT_BS, 0, s>>>(x, y, n, partial, act);
  k_reduce<<<1, DOT_BS, 0, s>>>(partial, nb, out, sqrt_out, act);
  EXA_HIP(hipGetLastError());
}

void multidot(const double* V, long ldv, int k, const double* w, long n, double* partial, double* out,
              hipStream_t s, const double* act) {
  int nb = dot_partials_count(n);
  dim3 grid(nb, k);
  k_multidot_partial<<<grid, DOT_BS, 0, s>>>(V, ldv, w, n, partial, act);
  k_reduce<<<k, DOT_BS, 0, s>>>(partial, nb, out, 0, act);
  EXA_HIP(hipGetLastError());
}

void axpy_dev(double* y, const double* x, const double* coef, double scale, long n, hipStream_t s,
              const double* act) {
  k_axpy_dev<<<nblocks(n, 256), 256, 0, s>>>(y, x, coef, scale, n, act);
  EXA_HIP(hipGetLastError());
}

void scale_inv_dev(double* x, const double* coef, long n, hipStream_t s, const double* act) {
  k_scale_inv_dev<<<nblocks(n, 256), 256, 0, s>>>(x, coef, n, act);
  EXA_HIP(hipGetLastError());
}

void axpby(double* y, const double* x, double a, double b, long n, hipStream_t s) {
  k_axpby<<<nblocks(n, 256), 256, 0, s>>>(y, x, a, b, n);
  EXA_HIP(hipGetLastError());
}

void first_nonzero(const int* info, long n, int* out, hipStream_t s) {
  k_first_init<<<1, 64, 0, s>>>(out);
  if (n > 0) k_first_nonzero<<<nblocks(n, 256), 256, 0, s>>>(info, n, out);
  EXA_HIP(hipGetLastError());
}

void copy(double* y, const double* x, long n, hipStream_t s) {
  k_copy<<<nblocks(n, 256), 256, 0, s>>>(y, x, n);
  EXA_HIP(hipGetLastError());
}

void gemv_cols(double* x, const double* V, long ldv, int k, const double* y, double alpha, long n,
               hipStream_t s, const double* act, bool skip_zero) {
  if (skip_zero) k_gemv_cols<true><<<nblocks(n, 256), 256, 0, s>>>(x, V, ldv, k, y, alpha, n, act);
  else k_gemv_cols<false><<<nblocks(n, 256), 256, 0, s>>>(x, V, ldv, k, y, alpha, n, act);
  EXA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Restarted GMRES with the Krylov scalars on the device (gmres.cpp:361-597): the Givens
// rotations, the convergence test and the back substitution of every instance run in one small
// kernel per Arnoldi step, so the host never waits on a step.  One block; thread t handles
// instances t, t + 256, ...; stat[0] receives the number of instances still iterating.
// ---------------------------------------------------------------------------
namespace {

constexpr int GM_BS = 1024;  // one block: 16 waves

// cpuApplyGivensRotation (backend/Common/cpuimpl.h:167-197) on the new Hessenberg column of one
// instance: the rotations k < i are applied in order with the running entry (the column's entry k
// after rotation k-1) kept in a register, so the chain waits on FMAs only, not on memory; the
// column comes from hcol (entry k at hc[k*ns]) and the rotated column goes to H.
__device__ void dev_givens(const double* __restrict__ hc, int ns, double* __restrict__ H, double* __restrict__ s,
                           double* __restrict__ cs, double* __restrict__ sn, int i) {
  double carry = hc[0];
  int k = 0;
  for (; k + 8 <= i; k += 8) {  // the chunk's loads are issued before its FMAs
    double h1[8], c[8], sg[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      h1[u] = hc[static_cast<long>(k + u + 1) * ns];
      c[u] = cs[k + u];
      sg[u] = sn[k + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      H[k + u] = c[u] * carry + sg[u] * h1[u];
      carry = -sg[u] * carry + c[u] * h1[u];
    }
  }
  for (; k < i; ++k) {
    const double h1 = hc[static_cast<long>(k + 1) * ns];
    const double c = cs[k], sg = sn[k];
    H[k] = c * carry + sg * h1;
    carry = -sg * carry + c * h1;
  }
  const double a = carry, b = hc[static_cast<long>(i + 1) * ns];
  double ci, si;
  if (b == 0.0) {
    ci = 1.0;
    si = 0.0;
  } else if (fabs(b) > fabs(a)) {
    const double t = a / b;
    si = 1.0 / sqrt(1.0 + t * t);
    ci = t * si;
  } else {
    const double t = b / a;
    ci = 1.0 / sqrt(1.0 + t * t);
    si = t * ci;
  }
  cs[i] = ci;
  sn[i] = si;
  const double t = ci * s[i];
  s[i + 1] = -si * s[i];
  s[i] = t;
  H[i] = ci * a + si * b;
  H[i + 1] = 0.0;
}

// cpuBackSolve (cpuimpl.h:199-208) by one wave: y (stride ys) = R^{-1} s over rows 0..i, R the
// leading (i+1) x (i+1) block of Hm (column-major, ld).  Column-oriented: lane l keeps the right-
// hand sides of rows l, l + 64, ...; for k = i .. 0 the owner of row k divides, the value goes to
// every lane, and each lane updates its rows above k; the columns' entries are loaded 8 columns
// ahead.  Rows beyond 64 * GM_ROWS fall back to one lane.
constexpr int GM_ROWS = 4;  // i + 1 <= 256 on the wave path
__device__ void dev_back_solve_wave(double* __restrict__ y, int ys, const double* __restrict__ Hm,
                                    const double* __restrict__ s, int i, int ld) {
  const int lane = threadIdx.x & 63;
  if (i + 1 > 64 * GM_ROWS) {
    if (lane == 0) {
      for (int j = i; j >= 0; --j) y[static_cast<long>(j) * ys] = s[j];
      for (int j = i; j >= 0; --j) {
        const double yj = y[static_cast<long>(j) * ys] / Hm[j + static_cast<long>(ld) * j];
        y[static_cast<long>(j) * ys] = yj;
        for (int k = j - 1; k >= 0; --k) y[static_cast<long>(k) * ys] -= Hm[k + static_cast<long>(ld) * j] * yj;
      }
    }
    return;
  }
  double r[GM_ROWS];
#pragma unroll
  for (int q = 0; q < GM_ROWS; ++q) {
    const int row = lane + 64 * q;
    r[q] = row <= i ? s[row] : 0.0;
  }
  for (int k0 = i; k0 >= 0; k0 -= 8) {
    double col[8][GM_ROWS], dg[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 - u;
      const int kk = k >= 0 ? k : 0;
      dg[u] = Hm[kk + static_cast<long>(ld) * kk];
#pragma unroll
      for (int q = 0; q < GM_ROWS; ++q) {
        const int row = lane + 64 * q;
        col[u][q] = row < kk ? Hm[row + static_cast<long>(ld) * kk] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 - u;
      if (k < 0) break;
      const int ql = k >> 6, owner = k & 63;
      double rk = 0.0;
#pragma unroll
      for (int q = 0; q < GM_ROWS; ++q)
        if (q == ql) rk = r[q];
      rk = __shfl(rk, owner);
      const double yk = rk / dg[u];
      if (lane == owner) y[static_cast<long>(k) * ys] = yk;
#pragma unroll
      for (int q = 0; q < GM_ROWS; ++q) r[q] -= col[u][q] * yk;
    }
  }
}

__device__ void block_count(int mine, int* stat) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  if (mine) atomicAdd(&cnt, mine);
  __syncthreads();
  if (threadIdx.x == 0) stat[0] = cnt;
}

// start of a solve: an instance iterates unless ||b|| is zero, non-finite, or (Newton's own
// solves) below the nonlinear tolerance (gmres.cpp:412-415)
__global__ __launch_bounds__(GM_BS) void k_gmres_begin(GmresDev g, const double* nb, double tol, double ntol,
                                                       int use_ntol) {
  if (threadIdx.x == 0) {
    g.tol[0] = tol;
    g.jcount[0] = 0;
    g.ticket[0] = 0;
  }
  int mine = 0;
  for (int sg = threadIdx.x; sg < g.ns; sg += GM_BS) {
    const double b = nb[sg];
    const bool a = isfinite(b) && !(b == 0.0 || (use_ntol && b < ntol));
    g.act[sg] = a ? 1.0 : 0.0;
    g.jc[sg] = 0.0;
    g.iters[sg] = 0;
    mine += a;
  }
  block_count(mine, g.stat);
}

// first cycle: nrmb = ||P r0|| (gmres.cpp:431-445), s = (nrmb, 0, ...), coef = 1/||r|| scale of v_0
__global__ __launch_bounds__(GM_BS) void k_gmres_start(GmresDev g, const double* nr, double* coef) {
  int mine = 0;
  const int n1 = g.n1;
  for (int sg = threadIdx.x; sg < g.ns; sg += GM_BS) {
    const double r = nr[sg];
    if (g.act[sg] != 0.0 && !isfinite(r)) g.act[sg] = 0.0;
    const bool a = g.act[sg] != 0.0;
    g.nrmb[sg] = r;
    double* sv = g.sv + static_cast<long>(sg) * (n1 + 1);
    sv[0] = a ? r : 0.0;
    for (int k = 1; k <= n1; ++k) sv[k] = 0.0;
    coef[sg] = a ? r : 0.0;
    mine += a;
  }
  block_count(mine, g.stat);
}

// Arnoldi step j (column i of the cycle): rotate the new column hcol[k*ns + sg] (k <= i+1), test
// |s_{i+1}| / nrmb < tol; a converged instance gets y = R^{-1} s in y[k*ns + sg] (k <= i), jc = 1,
// act = 0 and the reference's iteration count j; a non-finite rotation stops the instance.
// Arnoldi step j: one wave per instance, 16 instances per block.  The wave stages the new column
// (hcol[k*ns + sg], k <= i + 1) and the instance's rotations in LDS, lane 0 applies the rotations
// in order and forms the new one (cpuApplyGivensRotation, cpuimpl.h:167-197, the same operations
// as dev_givens), the lanes write the rotated column to H, and an instance with
// |s_{i+1}| / nrmb < tol gets y = R^{-1} s (zero above row i), jc = 1 (x += V y at the end of the
// cycle), act = 0 and the reference's iteration count j; a non-finite rotation stops the instance.
// The last block to finish adds the blocks' counts in block order, writes them to stat[j % ring]
// and advances jcount.
constexpr int GM_WAVES = GM_BS / 64;
constexpr int GM_LDS_N1 = 128;  // columns up to this length are staged in LDS
__global__ __launch_bounds__(GM_BS) void k_gmres_givens(GmresDev g, int i, const double* hcol) {
  __shared__ double lh[GM_WAVES][GM_LDS_N1 + 1], lc[GM_WAVES][GM_LDS_N1], ls[GM_WAVES][GM_LDS_N1];
  __shared__ int flag[GM_WAVES];
  __shared__ int cnt;
  __shared__ bool last;
  const int n1 = g.n1, ns = g.ns;
  const int wave = static_cast<int>(threadIdx.x >> 6), lane = static_cast<int>(threadIdx.x & 63);
  const int sg = static_cast<int>(blockIdx.x) * GM_WAVES + wave;
  const int j = g.jcount[0];
  const double tol = g.tol[0];
  const bool live = sg < ns && g.act[sg] != 0.0;
  const bool staged = n1 <= GM_LDS_N1;
  if (threadIdx.x == 0) cnt = 0;
  double* H = nullptr;
  double *sv = nullptr, *cs = nullptr, *sn = nullptr;
  if (live) {
    H = g.Hm + static_cast<long>(sg) * n1 * n1 + static_cast<long>(n1) * i;
    sv = g.sv + static_cast<long>(sg) * (n1 + 1);
    cs = g.cs + static_cast<long>(sg) * (n1 + 1);
    sn = g.sn + static_cast<long>(sg) * (n1 + 1);
    if (staged) {
      for (int k = lane; k <= i + 1; k += 64) lh[wave][k] = hcol[static_cast<long>(k) * ns + sg];
      for (int k = lane; k < i; k += 64) {
        lc[wave][k] = cs[k];
        ls[wave][k] = sn[k];
      }
    }
  }
  __syncthreads();
  if (live && lane == 0) {
    if (staged) {
      double* h = lh[wave];
      const double* c = lc[wave];
      const double* sg_ = ls[wave];
      double carry = h[0];
      for (int k = 0; k < i; ++k) {
        const double h1 = h[k + 1];
        h[k] = c[k] * carry + sg_[k] * h1;
        carry = -sg_[k] * carry + c[k] * h1;
      }
      const double a0 = carry, b0 = h[i + 1];
      double ci, si;
      if (b0 == 0.0) {
        ci = 1.0;
        si = 0.0;
      } else if (fabs(b0) > fabs(a0)) {
        const double t = a0 / b0;
        si = 1.0 / sqrt(1.0 + t * t);
        ci = t * si;
      } else {
        const double t = b0 / a0;
        ci = 1.0 / sqrt(1.0 + t * t);
        si = t * ci;
      }
      cs[i] = ci;
      sn[i] = si;
      const double t = ci * sv[i];
      sv[i + 1] = -si * sv[i];
      sv[i] = t;
      h[i] = ci * a0 + si * b0;
      h[i + 1] = 0.0;
    } else {
      dev_givens(hcol + sg, ns, H, sv, cs, sn, i);
    }
    g.iters[sg] = j + 1;
    const double s1 = sv[i + 1];
    int f = 1;  // 1 iterating, 2 converged, 0 stopped
    if (!isfinite(s1)) f = 0;
    else if (fabs(s1) / g.nrmb[sg] < tol) f = 2;
    if (f != 1) {
      g.act[sg] = 0.0;
      if (f == 2) {
        g.iters[sg] = j;
        g.jc[sg] = 1.0;
      }
    }
    flag[wave] = f;
  }
  __syncthreads();
  if (live) {
    if (staged)
      for (int k = lane; k <= i + 1; k += 64) H[k] = lh[wave][k];
    if (flag[wave] == 2) {  // y = R^{-1} s, zero above row i
      __threadfence_block();
      dev_back_solve_wave(g.y + sg, ns, g.Hm + static_cast<long>(sg) * n1 * n1, sv, i, n1);
      for (int k = i + 1 + lane; k < n1; k += 64) g.y[static_cast<long>(k) * ns + sg] = 0.0;
    }
    if (lane == 0 && flag[wave] == 1) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    g.bcount[blockIdx.x] = cnt;
    __threadfence();
    last = atomicAdd(g.ticket, 1) == static_cast<int>(gridDim.x) - 1;
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    int tot = 0;
    for (unsigned b = 0; b < gridDim.x; ++b) tot += atomicAdd(g.bcount + b, 0);
    g.stat[j % g.ring] = tot;
    g.jcount[0] = j + 1;
    g.ticket[0] = 0;
  }
}

// end of a cycle of i steps: y = R^{-1} s over rows 0..i-1 for the instances still iterating
// (for x += V y, gmres.cpp:552, UpdateSolution); xmask marks them and the instances that converged
// in this cycle (their y is set), one wave per instance
__global__ __launch_bounds__(GM_BS) void k_gmres_cycle_y(GmresDev g, int i) {
  const int n1 = g.n1, ns = g.ns;
  const int lane = threadIdx.x & 63;
  for (int sg = static_cast<int>(threadIdx.x >> 6); sg < ns; sg += GM_BS / 64) {
    if (g.act[sg] != 0.0) {
      dev_back_solve_wave(g.y + sg, ns, g.Hm + static_cast<long>(sg) * n1 * n1,
                          g.sv + static_cast<long>(sg) * (n1 + 1), i - 1, n1);
      for (int k = i + lane; k < n1; k += 64) g.y[static_cast<long>(k) * ns + sg] = 0.0;
    }
    if (lane == 0) g.xmask[sg] = (g.act[sg] != 0.0 || g.jc[sg] != 0.0) ? 1.0 : 0.0;
  }
}

// end of a solve: x += V y still due for the instances that converged in the last cycle
__global__ __launch_bounds__(GM_BS) void k_gmres_flush(GmresDev g) {
  for (int sg = threadIdx.x; sg < g.ns; sg += GM_BS) {
    g.xmask[sg] = g.jc[sg];
    g.jc[sg] = 0.0;
  }
}

// restart test (gmres.cpp:561-566): ||P(b - A x)|| / nrmb < tol stops the instance; the others
// restart with s = (||r||, 0, ...) and v_0 = r / ||r||
__global__ __launch_bounds__(GM_BS) void k_gmres_restart(GmresDev g, const double* nr, double* coef) {
  int mine = 0;
  const int n1 = g.n1;
  const double tol = g.tol[0];
  const int j = g.jcount[0] - 1;  // the step this cycle ended with
  for (int sg = threadIdx.x; sg < g.ns; sg += GM_BS) {
    g.jc[sg] = 0.0;  // their x update was made with the cycle's
    const double r = nr[sg];
    if (g.act[sg] != 0.0 && r / g.nrmb[sg] < tol) g.act[sg] = 0.0;
    const bool a = g.act[sg] != 0.0;
    if (a) {
      double* sv = g.sv + static_cast<long>(sg) * (n1 + 1);
      sv[0] = r;
      for (int k = 1; k <= n1; ++k) sv[k] = 0.0;
    }
    coef[sg] = a ? r : 0.0;
    mine += a;
  }
  block_count(mine, g.stat + j % g.ring);
}

__global__ void k_fill(double* y, double v, long n) {
  const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (i < n) y[i] = v;
}

}  // namespace

void gmres_dev_begin(const GmresDev& g, const double* nb, double tol, double newton_tol, int use_newton_tol,
                     hipStream_t s) {
  k_gmres_begin<<<1, GM_BS, 0, s>>>(g, nb, tol, newton_tol, use_newton_tol);
  EXA_HIP(hipGetLastError());
}
void gmres_dev_start(const GmresDev& g, const double* nr, double* coef, hipStream_t s) {
  k_gmres_start<<<1, GM_BS, 0, s>>>(g, nr, coef);
  EXA_HIP(hipGetLastError());
}
void gmres_dev_givens(const GmresDev& g, int i, const double* hcol, hipStream_t s) {
  k_gmres_givens<<<(g.ns + GM_WAVES - 1) / GM_WAVES, GM_BS, 0, s>>>(g, i, hcol);
  EXA_HIP(hipGetLastError());
}
void gmres_dev_cycle_y(const GmresDev& g, int i, hipStream_t s) {
  k_gmres_cycle_y<<<1, GM_BS, 0, s>>>(g, i);
  EXA_HIP(hipGetLastError());
}
void gmres_dev_flush(const GmresDev& g, hipStream_t s) {
  k_gmres_flush<<<1, GM_BS, 0, s>>>(g);
  EXA_HIP(hipGetLastError());
}
void gmres_dev_restart(const GmresDev& g, const double* nr, double* coef, hipStream_t s) {
  k_gmres_restart<<<1, GM_BS, 0, s>>>(g, nr, coef);
  EXA_HIP(hipGetLastError());
}
void fill(double* y, double v, long n, hipStream_t s) {
  if (n <= 0) return;
  k_fill<<<nblocks(n, 256), 256, 0, s>>>(y, v, n);
  EXA_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Vectors split into per-instance segments (batched solves).  Every block covers
// entries [bbeg, bend) of one segment, so a segment's sums and updates depend on
// its own entries and length only.
// ---------------------------------------------------------------------------
namespace {

constexpr int SEG_BS = 256;

// partial[v*nblk + b] = sum over block b of V_v[i] * w[i]   (blockIdx.y = v)
__global__ __launch_bounds__(SEG_BS) void k_seg_dot_partial(SegMap m, const double* V, long ldv,
                                                            const double* w, double* partial) {
  __shared__ double sh[SEG_BS];
  const int b = blockIdx.x;
  if (m.act && m.act[m.bseg[b]] == 0.0) {  // block-uniform
    if (threadIdx.x == 0) partial[static_cast<long>(blockIdx.y) * m.nblk + b] = 0.0;
    return;
  }
  const double* x = V + ldv * blockIdx.y;
  double acc = 0.0;
  for (long i = m.bbeg[b] + threadIdx.x; i < m.bend[b]; i += SEG_BS) acc = fma(x[i], w[i], acc);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int st = SEG_BS / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[static_cast<long>(blockIdx.y) * m.nblk + b] = sh[0];
}

// partial[v*nblk + b] = <V_v, w> over block b for every v < k: one block per chunk, the chunk of
// w staged in LDS once, wave q owns vectors q, q+4, ...; each lane sums its strided entries in
// order and the wave adds its lanes with a fixed butterfly.
__global__ __launch_bounds__(SEG_BS) void k_seg_multidot_partial(SegMap m, const double* V, long ldv, int k,
                                                                 const double* w, double* partial) {
  __shared__ double wl[SEG_CHUNK];
  const int b = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (m.act && m.act[m.bseg[b]] == 0.0) {  // block-uniform
    for (int v = threadIdx.x; v < k; v += SEG_BS) partial[static_cast<long>(v) * m.nblk + b] = 0.0;
    return;
  }
  const long i0 = m.bbeg[b];
  const int len = static_cast<int>(m.bend[b] - i0);
  for (int t = threadIdx.x; t < len; t += SEG_BS) wl[t] = w[i0 + t];
  __syncthreads();
  for (int v = wave; v < k; v += SEG_BS / 64) {
    const double* x = V + ldv * v + i0;
    double acc = 0.0;
    for (int t = lane; t < len; t += 64) acc = fma(x[t], wl[t], acc);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) partial[static_cast<long>(v) * m.nblk + b] = acc;
  }
}

// out[v*nseg + s] = sum of segment s's partials (fixed order), sqrt if asked  (blockIdx.y = v)
__global__ __launch_bounds__(SEG_BS) void k_seg_reduce(SegMap m, const double* partial, double* out,
                                                       int sqrt_out) {
  __shared__ double sh[SEG_BS];
  const int sg = blockIdx.x;
  const double* pv = partial + static_cast<long>(blockIdx.y) * m.nblk;
  double acc = 0.0;
  for (int i = m.sbptr[sg] + threadIdx.x; i < m.sbptr[sg + 1]; i += SEG_BS) acc += pv[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int st = SEG_BS / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[static_cast<long>(blockIdx.y) * m.nseg + sg] = sqrt_out ? sqrt(sh[0]) : sh[0];
}

// y += scale * coef[seg] * x
__global__ __launch_bounds__(SEG_BS) void k_seg_axpy(SegMap m, double* y, const double* x, const double* coef,
                                                     double scale) {
  const int b = blockIdx.x;
  if (m.act && m.act[m.bseg[b]] == 0.0) return;
  const double c = scale * coef[m.bseg[b]];
  if (c == 0.0) return;
  for (long i = m.bbeg[b] + threadIdx.x; i < m.bend[b]; i += SEG_BS) y[i] = fma(c, x[i], y[i]);
}

// mode 0: y = coef[seg] * x;  mode 1: y = x / coef[seg] (0 where coef[seg] == 0)
__global__ __launch_bounds__(SEG_BS) void k_seg_scale(SegMap m, double* y, const double* x, const double* coef,
                                                      int mode) {
  const int b = blockIdx.x;
  const double cv = coef[m.bseg[b]];
  const double c = mode == 0 ? cv : (cv != 0.0 ? 1.0 / cv : 0.0);
  for (long i = m.bbeg[b] + threadIdx.x; i < m.bend[b]; i += SEG_BS) y[i] = c * x[i];
}

// x += alpha * sum_j V_j * y[j*nseg + seg]
template <bool SKIPZ>
__global__ __launch_bounds__(SEG_BS) void k_seg_gemv_cols(SegMap m, double* x, const double* V, long ldv, int k,
                                                          const double* y, double alpha) {
  const int b = blockIdx.x;
  const int sg = m.bseg[b];
  if (m.act && m.act[sg] == 0.0) return;
  for (long i = m.bbeg[b] + threadIdx.x; i < m.bend[b]; i += SEG_BS) {
    double acc = 0.0;
    for (int j = 0; j < k; ++j) {
      const double yj = y[static_cast<long>(j) * m.nseg + sg];
      if (SKIPZ && yj == 0.0) continue;
      acc = fma(V[i + ldv * j], yj, acc);
    }
    x[i] = fma(alpha, acc, x[i]);
  }
}

}  // namespace

void seg_multidot(const SegMap& m, const double* V, long ldv, int k, const double* w, double* partial,
                  double* out, int sqrt_out, hipStream_t s) {
  if (m.nblk == 0 || k == 0) return;
  if (k >= 4) k_seg_multidot_partial<<<m.nblk, SEG_BS, 0, s>>>(m, V, ldv, k, w, partial);
  else k_seg_dot_partial<<<dim3(m.nblk, k), SEG_BS, 0, s>>>(m, V, ldv, w, partial);
  k_seg_reduce<<<dim3(m.nseg, k), SEG_BS, 0, s>>>(m, partial, out, sqrt_out);
  EXA_HIP(hipGetLastError());
}

void seg_axpy(const SegMap& m, double* y, const double* x, const double* coef, double scale, hipStream_t s) {
  if (m.nblk == 0) return;
  k_seg_axpy<<<m.nblk, SEG_BS, 0, s>>>(m, y, x, coef, scale);
  EXA_HIP(hipGetLastError());
}

void seg_scale(const SegMap& m, double* y, const double* x, const double* coef, int mode, hipStream_t s) {
  if (m.nblk == 0) return;
  k_seg_scale<<<m.nblk, SEG_BS, 0, s>>>(m, y, x, coef, mode);
  EXA_HIP(hipGetLastError());
}

void seg_gemv_cols(const SegMap& m, double* x, const double* V, long ldv, int k, const double* y, double alpha,
                   hipStream_t s, bool skip_zero) {
  if (m.nblk == 0 || k == 0) return;
  if (skip_zero) k_seg_gemv_cols<true><<<m.nblk, SEG_BS, 0, s>>>(m, x, V, ldv, k, y, alpha);
  else k_seg_gemv_cols<false><<<m.nblk, SEG_BS, 0, s>>>(m, x, V, ldv, k, y, alpha);
  EXA_HIP(hipGetLastError());
}

template void bj_apply<double>(const Dims&, const double*, const double*, double*, hipStream_t, const double*);
template void bj_apply<float>(const Dims&, const float*, const double*, double*, hipStream_t, const double*);
template void matvec_elem<double>(const Dims&, const double*, const int*, const double*, double*, hipStream_t,
                                  const int*, const double*);
template void matvec_elem<float>(const Dims&, const float*, const int*, const double*, double*, hipStream_t,
                                 const int*, const double*);
template void matvec_elem<unsigned short>(const Dims&, const unsigned short*, const int*, const double*, double*,
                                          hipStream_t, const int*, const double*);
template void dudg<double>(const Dims&, const double*, const double*, const int*, const double*, double*,
                           hipStream_t);
template void dudg<float>(const Dims&, const double*, const float*, const int*, const double*, double*,
                          hipStream_t);
template void bilu_apply<double>(const Dims&, const IluDev&, const double*, const double*, double*, double*,
                                 double*, hipStream_t, const double*);
template void bilu_apply<float>(const Dims&, const IluDev&, const float*, const double*, double*, double*,
                                double*, hipStream_t, const double*);
template void bilu_apply<unsigned short>(const Dims&, const IluDev&, const unsigned short*, const double*, double*,
                                         double*, double*, hipStream_t, const double*);

}  // namespace exa

