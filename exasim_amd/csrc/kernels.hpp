// kernels.hpp -- launch wrappers of the HIP kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace exa {

// Models generated from pdemodel.txt files (modelgen.py -> generated/registry.hpp):
// the model id for `name` and its nd, ncu and auxiliary-field width nv, or -1.
int generated_model(const std::string& name, int* nd, int* ncu, int* nv);

struct Dims {
  int nd, ncu, nc, nco, npe, npf, nfe, nge, ngf, ne, nf;
  int n, m, ncf, ndf;
  long N;  // trace unknowns ncu*npf*nf
};

// Column-major batched GEMM  C_b = alpha * A_b * B_b + beta * C_b  (no transposes).
// Batch index b = b0 + nb0*(b1 + nb1*b2); each operand offset is s0*i0 + s1*b1 + s2*b2
// with i0 = b0, or i0 = idx[b0] when that operand's gather index list is given.
struct Gemm {
  int M, N, K;
  double alpha, beta;
  const double* A; long lda, sA0, sA1, sA2;
  const double* B; long ldb, sB0, sB1, sB2;
  double* C;       long ldc, sC0, sC1, sC2;
  int nb0, nb1, nb2;
  const int* iA = nullptr;
  const int* iB = nullptr;
  const int* iC = nullptr;
  int fp32 = 0;  // 1: operands rounded to FP32, FP32 MFMA products and accumulation
  // operands stored in FP32 (set instead of A / B / C, in any combination)
  const float* Af = nullptr;
  const float* Bf = nullptr;
  float* Cf = nullptr;
};
void gemm(const Gemm& g, hipStream_t s);
// the tile autotuner of gemm (gemm.hip): mode 1 times the workgroup tiles on the first call of each
// product shape, 0 always takes the default tile, -1 follows EXA_GEMM_AUTOTUNE (default 1)
void set_gemm_autotune(int mode);
// the block-ILU(0) apply of FP32 / BF16 factors: 1 one wave per superelement, 0 the 256-thread
// kernel, -1 follows EXA_BILU_WAVE (default 1)
void set_bilu_wave(int mode);
// every MFMA product on tile 0..4 (64x64 32x64 64x32 128x64 64x128), or -1: the tuner's choice
void set_gemm_tile(int tile);
// one line per tuned shape: M N K log2(batch) fp32 storage chosen-tile and the microseconds of each tile
std::string gemm_tuning_report();

// In-place blocked LU with partial pivoting of the leading n x n block of each
// augmented matrix A_b (n x ncols, ld n, batch stride `stride`), then back
// substitution so that columns n..ncols-1 hold D^{-1} * RHS.  info[b] = 1 when a
// zero pivot is met.
// pivout (optional, [n, batch]): the panel-relative row exchanges of the blocked
// factorisation, for refine_rhs
void lu_solve(double* A, int n, int ncols, long stride, int batch, int* info, hipStream_t s,
              int* pivout = nullptr, int* piv_ws = nullptr, void* tri_ws = nullptr);
// The same in FP32 storage and arithmetic (blocked path only: ncols > n)
void lu_solve(float* A, int n, int ncols, long stride, int batch, int* info, hipStream_t s,
              int* pivout = nullptr, int* piv_ws = nullptr, void* tri_ws = nullptr);
// piv_ws (optional, device, >= LU_PIV_WS*batch ints, owned by the caller): with it, panels of up
// to 256 rows are factored with the panel in registers (k_lu_panel_reg + k_lu_u12)
// tri_ws (optional, device, >= LU_TRI_WS*batch doubles, owned by the caller): with it, the FP32
// LU runs the triangular solves of each panel (U12 = L11^{-1} A12, X = U11^{-1} X) as an explicit
// triangle inverse and an in-place GEMM (the FP64 LU always uses substitution)
constexpr int LU_PIV_WS = 64;
constexpr long LU_TRI_WS = 64L * 64L;

// ---- setup ----
void elem_geom(const Dims& d, const double* shapegt, const double* xdg, double* egeo, hipStream_t s);
void face_geom(const Dims& d, const double* shapfgt, const int* perm, const double* xdg, double* fgeo,
               hipStream_t s);
void interp_elem(const Dims& d, int ncomp, const double* shapegt, const double* src, double* dst,
                 hipStream_t s);
void interp_face(const Dims& d, int ncomp, const double* shapfgt, const int* perm, const double* src,
                 double* dst, hipStream_t s);
// Augmented [Mass | C_1..C_nd | E_1..E_nd] for elements [e0, e0+eb); the mass matrices are
// also written to Mass [npe, npe, ne] (elements e0..e0+eb-1).  With Aq (device workspace of
// eb*(nd+1)*npe*nge doubles) the volume blocks are one batched MFMA GEMM against the shape table;
// without it, one thread per entry.
void qsetup(const Dims& d, int e0, int eb, const double* shapegt, const double* shapegw,
            const double* shapfgt, const double* shapfgw, const int* perm, const double* egeo,
            const double* fgeo, double* aug, double* Mass, hipStream_t s, double* Aq = nullptr);
void qpack(const Dims& d, int e0, int eb, const double* aug, double* MCs, double* MEs, hipStream_t s);
void getq(const Dims& d, const double* MCs, const double* MEs, const int* elemcon, const double* uh,
          double* udg, hipStream_t s);

// ---- assembly (model-templated inside) ----
// Model functions loaded from a model library (model name "plugin:<path>"): text2code's GPU form
// with the signatures of backend/Model/libpdemodel.hpp:1-41 (HdgFlux, HdgSource, HdgFbou), each a
// host function that launches its own kernel over ng points, arrays [component * ng + point].
typedef void (*PlugFluxFn)(double* f, double* f_udg, double* f_wdg, const double* xdg, const double* udg,
                           const double* odg, const double* wdg, const double* uinf, const double* param,
                           double time, int modelnumber, int ng, int nc, int ncu, int nd, int ncx, int nco,
                           int ncw);
typedef void (*PlugFbouFn)(double* f, double* f_udg, double* f_wdg, double* f_uhg, const double* xdg,
                           const double* udg, const double* odg, const double* wdg, const double* uhg,
                           const double* nlg, const double* tau, const double* uinf, const double* param,
                           double time, int modelnumber, int ib, int ng, int nc, int ncu, int nd, int ncx,
                           int nco, int ncw);
typedef void (*PlugStreamFn)(hipStream_t);
struct PluginFns {
  PlugFluxFn flux = nullptr, source = nullptr;
  PlugFbouFn fbou = nullptr;
  PlugStreamFn set_stream = nullptr;  // optional
  int nbc = 0, modelnumber = 0;
  const double* uinf = nullptr;       // device
};
// doubles of the staging buffer a model library needs per block of eb elements
long plugin_buffer_doubles(const Dims& d, int nbc, int eb);

struct AsmArgs {
  int model;
  double time;
  const double* param;  // device, [nparam, ninst]
  const double* tau;    // device, [ntau, ninst]
  const int* einst;     // device, instance of each element (null: one instance)
  const double* tinst;  // device, [ninst]: each instance's model time (null: all at time)
  int nparam, ntau;
  int e0, eb;
  const double *shapegt, *shapegw, *shapfgt, *shapfgw;
  const int *perm, *elemcon, *bf;
  // boundary faces of the block: bface[bf0 .. bf0 + nbf), face ids E*nfe + l in ascending order
  const int* bface;
  int bf0, nbf;
  const double *udg, *uh, *egeo, *fgeo, *og, *ogf;
  double *rv, *rJ, *W, *fpt;
  long augs;  // per-element stride of the augmented system [D | F | Ru (| I)]
  // time-dependent problems: source += (sdg - dtfactor*u) T (sdg nodal [npe, ncu, ne]), T the
  // model's Tdfunc when tdfunc is 1 (and the model has one), else 1
  int tdep, tdfunc;
  double dtfactor;
  const double* sdg;
  // model library (MODEL_PLUGIN): its functions (host) and the staging buffer of the point
  // inputs and outputs (device, plugin_buffer_doubles)
  const PluginFns* plugin;
  double* pbuf;
  int plugin_nbc;
  // w equation (models with NCW > 0): nodal w (first guess) and its source, [npe, ncw, ne];
  // alpha dw/dt + beta w = Sourcew with dw/dt ~ dtfactor w - wsrc
  const double *wdg, *wsrc;
  double dae_alpha, dae_beta;
  const double* wmask;  // GetW in a batch: device [ninst], nonzero for the instances to update (null: all)
};
// GetW: the w equation at every element node with the nodal (u, q) (wdg updated in place)
void w_nodes(const Dims& d, const AsmArgs& a, const double* xdg, const double* odg, double* wdg, hipStream_t s);
// the NCW of a generated model (0 when it has no w equation or is not a generated model)
int model_ncw_of(int model);
void vol_points(const Dims& d, const AsmArgs& a, bool jac, hipStream_t s);
void face_points(const Dims& d, const AsmArgs& a, bool jac, hipStream_t s);
// Ru[e*ldo + i + npe*k] = sum phiw * rv   (Ru block of element e0+e at out + e*ldo)
void vol_res(const Dims& d, const AsmArgs& a, double* out, long ldo, hipStream_t s);
void vol_res(const Dims& d, const AsmArgs& a, float* out, long ldo, hipStream_t s);
// W into a.W, or into an FP32 buffer of the same layout
void vol_W(const Dims& d, const AsmArgs& a, hipStream_t s);
void vol_W(const Dims& d, const AsmArgs& a, float* W, hipStream_t s);
void face_res(const Dims& d, const AsmArgs& a, int l, bool jac, double* ru, long ldru, double* Rh,
              hipStream_t s);
void face_res(const Dims& d, const AsmArgs& a, int l, bool jac, float* ru, long ldru, double* Rh,
              hipStream_t s);
// Jacobian parts of every face of the block's elements (face by face, in local order)
void face_jac(const Dims& d, const AsmArgs& a, double* aug, double* Bfull, double* Kt, double* Gfull, double* H,
              hipStream_t s);
// the same into FP32 D/B/F, K, G and H blocks
void face_jac(const Dims& d, const AsmArgs& a, float* aug, float* Bfull, float* Kt, float* Gfull, float* H,
              hipStream_t s);
void post_schur(const Dims& d, int e0, int eb, const double* aug, long augs, double* DinvF, double* RuS,
                hipStream_t s);
void post_schur(const Dims& d, int e0, int eb, const float* aug, long augs, float* DinvF, double* RuS,
                hipStream_t s);
// BF16 copy (round to nearest even), stored as unsigned short (the upper half of an FP32)
void to_bf16(unsigned short* dst, const double* src, long count, hipStream_t s);
void to_bf16(unsigned short* dst, const float* src, long count, hipStream_t s);
// dst_b[r + rows*c] = (float) src_b[r + ld*c]
void to_float(float* dst, const double* src, int rows, int cols, long ld, long sstride, long dstride, int batch,
              hipStream_t s);
void to_float(float* dst, const float* src, int rows, int cols, long ld, long sstride, long dstride, int batch,
              hipStream_t s);
// Condensed right-hand side of a new residual with stored factors (mixed precision):
// RuS_e = D'^{-1} Ru_e from the FP32 LU factors and pivots of D', Rh_e -= K'_e RuS_e.
void refine_rhs(const Dims& d, const float* LU, const int* piv, const float* K, const double* Ru, double* RuS,
                double* Rh, hipStream_t s);

// ---- global system ----
void face_gather(long nnodes, int ncu, const long* src1, const long* src2, const double* we, double* w,
                 hipStream_t s);
void bj_assemble(const Dims& d, const int* f2e, const int* elemcon, const double* H, double* aug,
                 hipStream_t s);
void bj_assemble(const Dims& d, const int* f2e, const int* elemcon, const float* H, double* aug, hipStream_t s);
void bj_extract(const Dims& d, const double* aug, double* BE, hipStream_t s);
// act (optional, device): the apply, dot and update kernels below do nothing when act[0] == 0 (a
// solve that has finished while later steps were already queued)
template <typename T>
void bj_apply(const Dims& d, const T* BE, const double* x, double* y, hipStream_t s, const double* act = nullptr);
// elemental additive Schwarz: augmented [block | I] for elements [e0, e0+eb), and the
// inverse (columns m..2m-1 after lu_solve) into BEe (the block's first element)
void eas_assemble(const Dims& d, int e0, int eb, const double* H, const int* f2e, const int* elemcon,
                  double* aug, hipStream_t s);
void eas_assemble(const Dims& d, int e0, int eb, const float* H, const int* f2e, const int* elemcon,
                  double* aug, hipStream_t s);
void eas_assemble(const Dims& d, int e0, int eb, const float* H, const int* f2e, const int* elemcon,
                  float* aug, hipStream_t s);
void eas_extract(const Dims& d, int eb, const double* aug, double* BEe, hipStream_t s);
void eas_extract(const Dims& d, int eb, const float* aug, float* BEe, hipStream_t s);
template <typename T>
void matvec_elem(const Dims& d, const T* H, const int* elemcon, const double* v, double* we, hipStream_t s,
                 const int* einst = nullptr, const double* emask = nullptr);
template <typename T>
void dudg(const Dims& d, const double* RuS, const T* DinvF, const int* elemcon, const double* x,
          double* du, hipStream_t s);
void update_udg(const Dims& d, const double* du, double alpha, double* udg, hipStream_t s);
// u components of udg [npe, nc, ne] as [npe, ncu, ne]: dst = a*u + b*dst, and udg.u = src
void u_axpby(const Dims& d, const double* udg, double a, double b, double* dst, hipStream_t s);
void u_insert(const Dims& d, const double* src, double* udg, hipStream_t s);

// ---- superelement block-ILU(0) (preconditioner = 2) ----
struct IluDev {
  int nse, nfse, nnz;
  const int *face, *row_ptr, *col_ind, *row_of, *diag;
  const int *eptr, *elim_ji, *uptr, *upd_jl, *upd_il, *upd_ji;
  const int *lptr, *lidx, *lcol, *uuptr, *uidx, *ucol;
  const int *slot1, *slot2;
  // the apply's sweeps as one program of (block, vector offset) pairs in 3 nfse segments:
  // forward rows 0..nfse-1 (L blocks, offset col*ncf), then for i = nfse-1..0 the U row and
  // the diagonal block (offset nfse*ncf: the intermediate vector)
  const int *wseg, *wblk, *wvec;
};
void bilu_assemble(const Dims& d, const IluDev& I, const int* f2e, const int* elemcon, const double* H,
                   double* BE, hipStream_t s);
void bilu_assemble(const Dims& d, const IluDev& I, const int* f2e, const int* elemcon, const float* H,
                   float* BE, hipStream_t s);
// factorisation steps of face slot i, batched over the superelements: invert the
// diagonal block (info[n] |= singular), and copy the ne_i blocks L = A_ji D^{-1}
// computed into Ltmp[k][n] back into place; the products run on gemm()
void bilu_invert_diag(const Dims& d, const IluDev& I, int i, double* BE, int* info, hipStream_t s);
void bilu_invert_diag(const Dims& d, const IluDev& I, int i, float* BE, int* info, hipStream_t s);
void bilu_put_l(const Dims& d, const IluDev& I, int k0, int ne_i, const double* Ltmp, double* BE,
                hipStream_t s);
void bilu_put_l(const Dims& d, const IluDev& I, int k0, int ne_i, const float* Ltmp, float* BE, hipStream_t s);
template <typename T>
void bilu_apply(const Dims& d, const IluDev& I, const T* BE, const double* x, double* Y, double* Xo,
                double* out, hipStream_t s, const double* act = nullptr);

// trace blocks up to this size use the one-thread-per-row matvec
constexpr int MATVEC_SMALL_M = 96;

// ---- vector kernels (results stay on the device) ----
int dot_partials_count(long n);
void dot(const double* x, const double* y, long n, double* partial, double* out, int sqrt_out,
         hipStream_t s, const double* act = nullptr);
void axpy_dev(double* y, const double* x, const double* coef, double scale, long n, hipStream_t s,
              const double* act = nullptr);
void scale_inv_dev(double* x, const double* coef, long n, hipStream_t s, const double* act = nullptr);
// y = a x + b y; with b == 0 y is only written
void axpby(double* y, const double* x, double a, double b, long n, hipStream_t s);
void copy(double* y, const double* x, long n, hipStream_t s);
// out[0] = lowest i < n with info[i] != 0, or INT_MAX (device int)
void first_nonzero(const int* info, long n, int* out, hipStream_t s);
// skip_zero: columns with a zero coefficient are not read
void gemv_cols(double* x, const double* V, long ldv, int k, const double* y, double alpha, long n,
               hipStream_t s, const double* act = nullptr, bool skip_zero = false);
void multidot(const double* V, long ldv, int k, const double* w, long n, double* partial, double* out,
              hipStream_t s, const double* act = nullptr);

// y[0..n) = v
void fill(double* y, double v, long n, hipStream_t s);

// ---- restarted GMRES with its scalars on the device (one or several independent instances) ----
// Per instance sg: Hessenberg matrix Hm [ns][n1][n1] (column-major n1 x n1), rotations and
// right-hand side sv, cs, sn [ns][n1 + 1], nrmb [ns] (preconditioned initial residual norm), the
// activity mask act [ns] (1.0 iterating, 0.0 finished), jc [ns] (1.0: converged at the last
// step), y [n1][ns] (y[k*ns + sg]) and the iteration count iters [ns] (the reference's count).
// Each launcher writes the number of instances still iterating to stat[0].
// The step kernels read the solve's tolerance from tol[0] and the number of steps taken so far
// from jcount[0] (the Arnoldi step j), and write their count to stat[j % ring]: a step's launches
// do not depend on j, so one captured graph per column i serves every cycle.
// jc [ns]: 1.0 for an instance that converged in the current cycle (its x += V y is made with the
// cycle's, or by gmres_dev_flush at the end); xmask [ns]: the instances of the next x update;
// bcount [(ns + 15) / 16], ticket [1]: the step kernel's per-block counts and completion ticket.
struct GmresDev {
  int ns, n1, ring;
  double *Hm, *sv, *cs, *sn, *nrmb, *act, *jc, *y, *tol, *xmask;
  int *iters, *jcount, *stat, *bcount, *ticket;
};
// start of a solve: tol[0] = tol, jcount[0] = 0, the activity from ||b|| (nb); count to stat[0]
void gmres_dev_begin(const GmresDev& g, const double* nb, double tol, double newton_tol, int use_newton_tol,
                     hipStream_t s);
void gmres_dev_start(const GmresDev& g, const double* nr, double* coef, hipStream_t s);
void gmres_dev_givens(const GmresDev& g, int i, const double* hcol, hipStream_t s);
void gmres_dev_cycle_y(const GmresDev& g, int i, hipStream_t s);
// xmask = jc, jc = 0: the x update still due at the end of a solve
void gmres_dev_flush(const GmresDev& g, hipStream_t s);
void gmres_dev_restart(const GmresDev& g, const double* nr, double* coef, hipStream_t s);

// ---- vectors split into per-instance segments (batched solves) ----
// Segment s of a vector is entries [off[s], off[s+1]).  Blocks of at most SEG_CHUNK entries
// never straddle two segments (block b covers [bbeg[b], bend[b]) of segment bseg[b]; the blocks of
// segment s are sbptr[s] .. sbptr[s+1]), so a segment's reductions see only its own entries, in an
// order fixed by its length: an instance gets the same sums alone and in a batch.
constexpr long SEG_CHUNK = 512;
struct SegMap {
  int nseg = 0, nblk = 0;
  const int* bseg = nullptr;
  const long* bbeg = nullptr;
  const long* bend = nullptr;
  const int* sbptr = nullptr;
  // optional per-segment activity (device, 1/0): the dot, axpy and gemv kernels skip the blocks
  // of inactive segments (dots over them read as 0, their entries are left unchanged)
  const double* act = nullptr;
};
// out[v*nseg + s] = <V_v, w> over segment s (v < k; sqrt of it with sqrt_out); partial: k*nblk
void seg_multidot(const SegMap& m, const double* V, long ldv, int k, const double* w, double* partial,
                  double* out, int sqrt_out, hipStream_t s);
// y += scale * coef[s] * x on each segment s
void seg_axpy(const SegMap& m, double* y, const double* x, const double* coef, double scale, hipStream_t s);
// mode 0: y = coef[s] * x;  mode 1: y = x / coef[s], and 0 where coef[s] == 0
void seg_scale(const SegMap& m, double* y, const double* x, const double* coef, int mode, hipStream_t s);
// x += alpha * sum_{j<k} V_j * y[j*nseg + s]
void seg_gemv_cols(const SegMap& m, double* x, const double* V, long ldv, int k, const double* y, double alpha,
                   hipStream_t s, bool skip_zero = false);

}  // namespace exa
