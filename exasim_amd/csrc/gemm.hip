// gemm.hip -- batched GEMMs of the assembly, the condensation LU and the set-up (strided batches
// of small dense products, uequation.cpp:559-775's schurMatrix* / the reference's batched BLAS
// calls) on gfx950, and the tile autotuner that picks each product's workgroup tile by measurement.
//
// Products of FP64 operands with M, N <= 32 run on a plain LDS-tiled kernel; all others on the
// matrix cores (v_mfma_f64_16x16x4f64, or v_mfma_f32_16x16x4f32 for FP32 products), with one of
// five workgroup tiles (BM x BN = 64x64, 32x64, 64x32, 128x64, 64x128; four waves, K staged through
// LDS 32 at a time).  Every tile runs the same MFMA instruction over the same K order for each
// output element, so the tiles give bitwise-identical results and differ only in speed: the tuner
// times them on the first call of a product shape -- (M, N, K, the batch count's power of two,
// FP32 or FP64 products, operand storage) -- on the caller's own operands (an accumulating product
// writes into a scratch copy of its C while it is timed), keeps the fastest for the process and
// launches it from then on.  No counterpart in the reference (its batched GEMMs
// are library calls).
#include "kernels.hpp"
#include "common.hpp"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

namespace exa {

namespace {

// ---------------------------------------------------------------------------
// Batched FP64 GEMM: LDS-tiled, 256 threads, each thread a TM x TN micro-tile (M, N <= 32).
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK, int TM, int TN>
__global__ __launch_bounds__(256) void k_gemm(Gemm g, int tilesM) {
  const int tile = blockIdx.x % tilesM;
  const long b = blockIdx.x / tilesM;
  const int b0 = static_cast<int>(b % g.nb0);
  const long t1 = b / g.nb0;
  const int b1 = static_cast<int>(t1 % g.nb1);
  const int b2 = static_cast<int>(t1 / g.nb1);
  const long a0 = g.iA ? g.iA[b0] : b0, bb0 = g.iB ? g.iB[b0] : b0, c0 = g.iC ? g.iC[b0] : b0;
  const double* A = g.A + g.sA0 * a0 + g.sA1 * b1 + g.sA2 * b2;
  const double* B = g.B + g.sB0 * bb0 + g.sB1 * b1 + g.sB2 * b2;
  double* C = g.C + g.sC0 * c0 + g.sC1 * b1 + g.sC2 * b2;
  const int row0 = tile * BM, col0 = blockIdx.y * BN;
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < g.K; k0 += BK) {
    for (int idx = threadIdx.x; idx < BM * BK; idx += 256) {
      int r = idx % BM, kk = idx / BM;
      int gr = row0 + r, gk = k0 + kk;
      As[kk][r] = (gr < g.M && gk < g.K) ? A[gr + g.lda * gk] : 0.0;
    }
    for (int idx = threadIdx.x; idx < BK * BN; idx += 256) {
      int kk = idx % BK, c = idx / BK;
      int gk = k0 + kk, gc = col0 + c;
      Bs[kk][c] = (gk < g.K && gc < g.N) ? B[gk + g.ldb * gc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
      double a[TM], bb[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < TN; ++j) bb[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = fma(a[i], bb[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      int r = row0 + tx + 16 * i, c = col0 + ty + 16 * j;
      if (r < g.M && c < g.N) {
        double v = g.alpha * acc[i][j];
        if (g.beta != 0.0) v = fma(g.beta, C[r + g.ldc * c], v);
        C[r + g.ldc * c] = v;
      }
    }
}

// ---------------------------------------------------------------------------
// Batched GEMM on the matrix cores.  Workgroup tile BM x BN, four waves in a 2 x 2 layout, each
// wave (BM/2) x (BN/2) as (BM/32) x (BN/32) MFMA tiles of 16 x 16; K staged through LDS MG_BK at a
// time with the next K-tile prefetched into registers.  Lane layout of the instruction: lane l
// supplies A[l%16][l/16] and B[l/16][l%16] of a 16x4 / 4x16 step.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
constexpr int MG_BK = 32;

// LDS is fragment-major: element (row r, k) of the A tile sits at
// (r/16)*16*MG_BK + k*16 + r%16, so the 64 lanes of one MFMA operand read 64
// consecutive values (no bank conflicts); B likewise with its columns.
__device__ inline int mg_off(int rc, int k) { return (rc >> 4) * (16 * MG_BK) + k * 16 + (rc & 15); }

// One 16x16x4 step.  FP64: register i of lane l holds D[l/16 + 4i][l%16];
// FP32 (inputs rounded to FP32, FP32 accumulation): D[4(l/16) + i][l%16].
__device__ inline dbl4 mfma_step(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline flt4 mfma_step(float a, float b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline int mfma_row(double, int lk, int i) { return lk + 4 * i; }
__device__ inline int mfma_row(float, int lk, int i) { return 4 * lk + i; }

// Workgroup w is dispatched to XCD w % 8 (each XCD has its own L2).  The logical
// tile id L = (w % 8) * (nwg / 8) + w / 8 gives every XCD a contiguous range of L,
// and L runs over the M/N tiles of one batch entry before the next, so the tiles
// that share an operand run together on one XCD and reuse it from its L2.
constexpr int NXCD = 8;

// Operand storage: the double pointer, or the float one when the operand is stored in FP32.
template <typename T> __device__ __forceinline__ const T* gsel(const double* d, const float* f);
template <> __device__ __forceinline__ const double* gsel<double>(const double* d, const float*) { return d; }
template <> __device__ __forceinline__ const float* gsel<float>(const double*, const float* f) { return f; }

// CT: MFMA type; TA, TB, TC: storage of A, B, C; BM x BN: the workgroup tile
template <typename CT, typename TA, typename TB, typename TC, int BM, int BN>
__global__ __launch_bounds__(256) void k_gemm_mfma(Gemm g, int tilesM, int tilesN, long ntiles) {
  typedef CT acc4 __attribute__((ext_vector_type(4)));
  constexpr int FM = BM / 32, FN = BN / 32;  // MFMA tiles per wave
  const long nwg = static_cast<long>(gridDim.x);   // a multiple of NXCD
  const long L = (blockIdx.x % NXCD) * (nwg / NXCD) + blockIdx.x / NXCD;
  if (L >= ntiles) return;
  const int tpb = tilesM * tilesN;
  const long b = L / tpb;
  const int tt = static_cast<int>(L % tpb);
  const int tile = tt % tilesM, tcol = tt / tilesM;
  const int b0 = static_cast<int>(b % g.nb0);
  const long t1 = b / g.nb0;
  const int b1 = static_cast<int>(t1 % g.nb1);
  const int b2 = static_cast<int>(t1 / g.nb1);
  const long a0 = g.iA ? g.iA[b0] : b0, bb0 = g.iB ? g.iB[b0] : b0, c0 = g.iC ? g.iC[b0] : b0;
  const TA* A = gsel<TA>(g.A, g.Af) + g.sA0 * a0 + g.sA1 * b1 + g.sA2 * b2;
  const TB* B = gsel<TB>(g.B, g.Bf) + g.sB0 * bb0 + g.sB1 * b1 + g.sB2 * b2;
  TC* C = const_cast<TC*>(gsel<TC>(g.C, g.Cf)) + g.sC0 * c0 + g.sC1 * b1 + g.sC2 * b2;
  const int row0 = tile * BM, col0 = tcol * BN;
  // the K tiles of A and B; the output tile is staged in the same LDS afterwards when it fits in
  // 48 KB (all FP32 tiles; the FP64 64x64, 32x64 and 64x32 tiles)
  constexpr int CLD = BM + 1;  // leading dimension of the staged output tile
  constexpr bool STAGE_C = static_cast<long>(BN) * CLD * sizeof(CT) <= 48 * 1024;
  constexpr int NAB = (BM + BN) * MG_BK, NCS = STAGE_C ? BN * CLD : 0;
  __shared__ CT smem[NAB > NCS ? NAB : NCS];
  CT* As = smem;
  CT* Bs = smem + BM * MG_BK;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave & 1, wc = wave >> 1;
  // global -> register staging of element e = tid + 256 j: A row e % BM, k = e / BM;
  // B k = e % MG_BK, column e / MG_BK
  constexpr int NLA = BM * MG_BK / 256, NLB = BN * MG_BK / 256;
  const int ar = tid % BM, ak = tid / BM;
  const int bk = tid % MG_BK, bc = tid / MG_BK;
  TA ra[NLA];
  TB rb[NLB];
  bool ma[NLA], mb[NLB];
  // out-of-range elements read a clamped (valid) address and are zeroed when stored to
  // LDS, so the loads need no branches and stay in flight through the MFMA loop
  const bool rowok = row0 + ar < g.M;
  const long arow = min(row0 + ar, g.M - 1);
  auto load = [&](int k0) {
#pragma unroll
    for (int j = 0; j < NLA; ++j) {
      const int gk = k0 + ak + (256 / BM) * j;
      ra[j] = A[arow + g.lda * min(gk, g.K - 1)];
      ma[j] = rowok & (gk < g.K);
    }
#pragma unroll
    for (int j = 0; j < NLB; ++j) {
      const int gk2 = k0 + bk, gc = col0 + bc + (256 / MG_BK) * j;
      rb[j] = B[min(gk2, g.K - 1) + g.ldb * static_cast<long>(min(gc, g.N - 1))];
      mb[j] = (gk2 < g.K) & (gc < g.N);
    }
  };
  acc4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc4{0, 0, 0, 0};
  load(0);
  const int li = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < g.K; k0 += MG_BK) {
#pragma unroll
    for (int j = 0; j < NLA; ++j) As[mg_off(ar, ak + (256 / BM) * j)] = ma[j] ? static_cast<CT>(ra[j]) : CT(0);
#pragma unroll
    for (int j = 0; j < NLB; ++j) Bs[mg_off(bc + (256 / MG_BK) * j, bk)] = mb[j] ? static_cast<CT>(rb[j]) : CT(0);
    __syncthreads();
    load(k0 + MG_BK);  // unconditional: past the end it reads clamped, unused values
#pragma unroll
    for (int kk = 0; kk < MG_BK; kk += 4) {
      const int kr = kk + lk;
      CT av[FM], bv[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) av[i] = As[mg_off((BM / 2) * wr + 16 * i + li, kr)];
#pragma unroll
      for (int j = 0; j < FN; ++j) bv[j] = Bs[mg_off((BN / 2) * wc + 16 * j + li, kr)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = mfma_step(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
  if constexpr (STAGE_C) {
    // the output tile through LDS (the K loop ended with a barrier): each column's rows are then
    // read and written by consecutive threads, where the MFMA layout puts 16 columns in a store
    CT* Cs = smem;
#pragma unroll
    for (int ti = 0; ti < FM; ++ti)
#pragma unroll
      for (int tj = 0; tj < FN; ++tj)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          Cs[((BN / 2) * wc + 16 * tj + li) * CLD + (BM / 2) * wr + 16 * ti + mfma_row(CT(0), lk, i)] = acc[ti][tj][i];
    __syncthreads();
    for (int idx = tid; idx < BM * BN; idx += 256) {
      const int rl = idx % BM, cl = idx / BM;
      const int r = row0 + rl, c = col0 + cl;
      if (r < g.M && c < g.N) {
        double v = g.alpha * static_cast<double>(Cs[cl * CLD + rl]);
        if (g.beta != 0.0) v = fma(g.beta, static_cast<double>(C[r + g.ldc * static_cast<long>(c)]), v);
        C[r + g.ldc * static_cast<long>(c)] = static_cast<TC>(v);
      }
    }
  } else {
#pragma unroll
    for (int ti = 0; ti < FM; ++ti)
#pragma unroll
      for (int tj = 0; tj < FN; ++tj)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = row0 + (BM / 2) * wr + 16 * ti + mfma_row(CT(0), lk, i);
          const int c = col0 + (BN / 2) * wc + 16 * tj + li;
          if (r < g.M && c < g.N) {
            double v = g.alpha * static_cast<double>(acc[ti][tj][i]);
            if (g.beta != 0.0) v = fma(g.beta, static_cast<double>(C[r + g.ldc * static_cast<long>(c)]), v);
            C[r + g.ldc * static_cast<long>(c)] = static_cast<TC>(v);
          }
        }
  }
}

// ---------------------------------------------------------------------------
// Launchers: one per (MFMA type, operand storage, tile)
// ---------------------------------------------------------------------------
struct TileShape {
  int bm, bn;
};
constexpr int NTILE = 5;
constexpr TileShape TILES[NTILE] = {{64, 64}, {32, 64}, {64, 32}, {128, 64}, {64, 128}};

typedef void (*MfmaLaunch)(const Gemm&, hipStream_t);

template <typename CT, typename TA, typename TB, typename TC, int BM, int BN>
void launch_mfma(const Gemm& g, hipStream_t s) {
  const long batch = static_cast<long>(g.nb0) * g.nb1 * g.nb2;
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const long ntiles = static_cast<long>(tm) * tn * batch;
  const long nwg = (ntiles + NXCD - 1) / NXCD * NXCD;
  if (nwg > 0x7fffffffL) throw Error(5, "GEMM batch too large for one launch");
  k_gemm_mfma<CT, TA, TB, TC, BM, BN><<<static_cast<unsigned>(nwg), 256, 0, s>>>(g, tm, tn, ntiles);
}

template <typename CT, typename TA, typename TB, typename TC>
constexpr std::array<MfmaLaunch, NTILE> tile_launchers() {
  return {launch_mfma<CT, TA, TB, TC, 64, 64>, launch_mfma<CT, TA, TB, TC, 32, 64>,
          launch_mfma<CT, TA, TB, TC, 64, 32>, launch_mfma<CT, TA, TB, TC, 128, 64>,
          launch_mfma<CT, TA, TB, TC, 64, 128>};
}

// [fp32][storage code: A float 1, B float 2, C float 4][tile]
template <typename CT>
std::array<std::array<MfmaLaunch, NTILE>, 8> storage_launchers() {
  return {tile_launchers<CT, double, double, double>(), tile_launchers<CT, float, double, double>(),
          tile_launchers<CT, double, float, double>(),  tile_launchers<CT, float, float, double>(),
          tile_launchers<CT, double, double, float>(),  tile_launchers<CT, float, double, float>(),
          tile_launchers<CT, double, float, float>(),   tile_launchers<CT, float, float, float>()};
}

MfmaLaunch mfma_launcher(int fp32, int store, int tile) {
  static const auto f64 = storage_launchers<double>();
  static const auto f32 = storage_launchers<float>();
  return fp32 ? f32[store][tile] : f64[store][tile];
}

// ---------------------------------------------------------------------------
// The autotuner
// ---------------------------------------------------------------------------
typedef std::tuple<int, int, int, int, int, int, int> TuneKey;  // device, M, N, K, log2 batch, fp32, store
struct TuneEntry {
  int tile = 0;
  long batch = 0;  // products in the timed call
  float us[NTILE] = {0};
};
std::mutex tune_mu;
std::map<TuneKey, TuneEntry>& tune_table() {
  static std::map<TuneKey, TuneEntry> t;
  return t;
}
int tune_mode() {  // 1: time the tiles on the first call of a shape; 0: the default tile
  static int m = [] {
    const char* e = std::getenv("EXA_GEMM_AUTOTUNE");
    return e ? std::atoi(e) : 1;
  }();
  return m;
}
int& tune_override() {  // set by set_gemm_autotune (-1: the environment's choice)
  static int v = -1;
  return v;
}
int& forced_tile() {  // set by set_gemm_tile: every MFMA product on this tile (-1: the tuner's)
  static int v = -1;
  return v;
}

int log2_bucket(long v) {
  int k = 0;
  while ((1L << (k + 1)) <= v) ++k;
  return k;
}

// A product may be run again on its own C for timing only when that leaves C as one run leaves
// it: beta = 0 and C apart from A and B (the tiles write identical values).
bool rerunnable(const Gemm& g) {
  if (g.beta != 0.0) return false;
  const void* c = g.Cf ? static_cast<const void*>(g.Cf) : static_cast<const void*>(g.C);
  const void* a = g.Af ? static_cast<const void*>(g.Af) : static_cast<const void*>(g.A);
  const void* b = g.Bf ? static_cast<const void*>(g.Bf) : static_cast<const void*>(g.B);
  return c != a && c != b;
}

// A product computed in place (C is B, as the LU's X = T^{-1} X) is correct only when one workgroup
// holds all rows of its columns (BM >= M: every row of B is read before any row of C is written);
// C is A likewise needs BN >= N.
bool tile_ok(const Gemm& g, int t) {
  const void* c = g.Cf ? static_cast<const void*>(g.Cf) : static_cast<const void*>(g.C);
  const void* a = g.Af ? static_cast<const void*>(g.Af) : static_cast<const void*>(g.A);
  const void* b = g.Bf ? static_cast<const void*>(g.Bf) : static_cast<const void*>(g.B);
  if (c == b && TILES[t].bm < g.M) return false;
  if (c == a && TILES[t].bn < g.N) return false;
  return true;
}

// Elements C spans (strided batches without an index map), or -1
long c_extent(const Gemm& g) {
  if (g.iC || g.sC0 < 0 || g.sC1 < 0 || g.sC2 < 0) return -1;
  return g.sC0 * (g.nb0 - 1) + g.sC1 * (g.nb1 - 1) + g.sC2 * (g.nb2 - 1) + g.ldc * static_cast<long>(g.N - 1) +
         g.M;
}

// the tile for this product: the table's, or timed now (every tile five times after one warm-up
// run, the fastest run kept, so that work of other streams inflates no tile's time), or the
// default when the shape cannot be timed.  A product that cannot be rerun on its own C is timed
// writing into a scratch copy of C's extent (its values are not used).
int pick_tile(const Gemm& g, int store, long batch, hipStream_t s) {
  if (forced_tile() >= 0) return tile_ok(g, forced_tile()) ? forced_tile() : 0;
  const int mode = tune_override() >= 0 ? tune_override() : tune_mode();
  if (mode == 0) return 0;
  const long work = static_cast<long>((g.M + 63) / 64) * ((g.N + 63) / 64) * batch;
  if (work < 64) return 0;  // a launch of a few workgroups: launch latency decides, not the tile
  int dev = 0;
  EXA_HIP(hipGetDevice(&dev));
  const TuneKey key{dev, g.M, g.N, g.K, log2_bucket(batch), g.fp32, store};
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tune_table().find(key);
    if (it != tune_table().end()) return tile_ok(g, it->second.tile) ? it->second.tile : 0;
  }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  EXA_HIP(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) return 0;  // no timing inside a graph capture
  Gemm gt = g;
  void* scratch = nullptr;
  if (!rerunnable(g)) {
    const long ext = c_extent(g);
    const size_t bytes = static_cast<size_t>(ext) * (g.Cf ? sizeof(float) : sizeof(double));
    if (ext <= 0 || bytes > (size_t(4) << 30)) return 0;
    EXA_HIP(hipStreamSynchronize(s));
    if (hipMalloc(&scratch, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return 0;
    }
    EXA_HIP(hipMemsetAsync(scratch, 0, bytes, s));
    if (g.Cf) gt.Cf = static_cast<float*>(scratch);
    else gt.C = static_cast<double*>(scratch);
  }
  TuneEntry ent;
  ent.batch = batch;
  hipEvent_t e0, e1;
  EXA_HIP(hipEventCreate(&e0));
  EXA_HIP(hipEventCreate(&e1));
  float best = 0.0f;
  for (int t = 0; t < NTILE; ++t) {
    if (!tile_ok(g, t)) {
      ent.us[t] = -1.0f;  // not a candidate for this product
      continue;
    }
    const MfmaLaunch f = mfma_launcher(g.fp32, store, t);
    f(gt, s);  // warm-up
    float fastest = 0.0f;
    for (int r = 0; r < 5; ++r) {
      float ms = 0.0f;
      EXA_HIP(hipEventRecord(e0, s));
      f(gt, s);
      EXA_HIP(hipEventRecord(e1, s));
      EXA_HIP(hipEventSynchronize(e1));
      EXA_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (r == 0 || ms < fastest) fastest = ms;
    }
    ent.us[t] = fastest * 1000.0f;
    if (t == 0 || ent.us[t] < best) {  // (tile 0 is always a candidate: in-place products have M, N <= 64)
      best = ent.us[t];
      ent.tile = t;
    }
  }
  EXA_HIP(hipEventDestroy(e0));
  EXA_HIP(hipEventDestroy(e1));
  if (scratch) {
    EXA_HIP(hipStreamSynchronize(s));
    EXA_HIP(hipFree(scratch));
  }
  EXA_HIP(hipGetLastError());
  std::lock_guard<std::mutex> lk(tune_mu);
  tune_table()[key] = ent;
  return ent.tile;
}

}  // namespace

void gemm(const Gemm& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  const long batch = static_cast<long>(g.nb0) * g.nb1 * g.nb2;
  if (batch <= 0) return;
  const int store = (g.Af ? 1 : 0) | (g.Bf ? 2 : 0) | (g.Cf ? 4 : 0);
  if (g.M <= 32 && g.N <= 32 && store == 0) {
    int tm = (g.M + 31) / 32, tn = (g.N + 31) / 32;
    dim3 grid(static_cast<unsigned>(tm * batch), tn);
    k_gemm<32, 32, 16, 2, 2><<<grid, 256, 0, s>>>(g, tm);
  } else {
    const int tile = pick_tile(g, store, batch, s);
    // (a product timed on its own C has left its result there; one more run writes the same values)
    mfma_launcher(g.fp32, store, tile)(g, s);
  }
  EXA_HIP(hipGetLastError());
}

void set_gemm_autotune(int mode) { tune_override() = mode; }

void set_gemm_tile(int tile) {
  if (tile < -1 || tile >= NTILE) throw Error(1, "gemm_tile: -1 (the tuner's choice) or 0..4");
  forced_tile() = tile;
}

std::string gemm_tuning_report() {
  std::lock_guard<std::mutex> lk(tune_mu);
  std::string out;
  char line[320];
  for (const auto& kv : tune_table()) {
    const TuneKey& k = kv.first;
    const TuneEntry& e = kv.second;
    std::snprintf(line, sizeof line, "%d %d %d %d %d %d %dx%d %ld", std::get<1>(k), std::get<2>(k), std::get<3>(k),
                  std::get<4>(k), std::get<5>(k), std::get<6>(k), TILES[e.tile].bm, TILES[e.tile].bn, e.batch);
    out += line;
    for (int t = 0; t < NTILE; ++t) {
      std::snprintf(line, sizeof line, " %.3f", e.us[t]);
      out += line;
    }
    out += "\n";
  }
  return out;
}

}  // namespace exa
