// solver.hpp -- the solver handle behind the C ABI: device-resident HDG state of
// one problem and the host control of the Newton-GMRES loop.
#pragma once

#include "bilu0.hpp"
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <array>
#include <string>
#include <vector>

namespace exa {

enum TimerClass { T_MATVEC = 0, T_PRECOND = 1, T_ORTHO = 2, T_ASSEMBLY = 3, T_LU = 4, T_OTHER = 5, T_NCLASS = 6 };

// model id of a built-in or generated model name (Error for an unknown name)
int resolve_model(const std::string& name);

// slots of the per-step status ring of the device-resident GMRES
constexpr int GM_RING = 32;

struct Solver;
// no-allocation matvec (noalloc.hip): look up this device's scratch array (device current), and
// w = A v with the element products in that array, in chunks of at most max_chunk elements (0: as
// many as fit)
void noalloc_init(int device);
long noalloc_scratch_doubles();
void matvec_noalloc(Solver& S, const double* v, double* w, int max_chunk);

struct Solver {
  // problem description
  std::string datain, dataout, model_name;
  int model = 0;
  int device = 0;
  Dims d{};
  AppData app;
  MasterData master;
  MeshData mesh;
  Conn conn;
  hipStream_t s = nullptr;

  // options (app.bin defaults, overridable)
  int precond = 0, ortho = 0, restart = 50, gmres_maxit = 200, newton_maxit = 20, verbose = 0;
  double gmres_tol = 1e-8, newton_tol = 1e-6;
  // precision 0: FP64 throughout; 1: FP32 storage of the condensed operator, the local
  // solve factors and the preconditioner, with FP64 residuals and an iterative-refinement
  // loop to refine_rtol * (first residual norm), corrections solved to refine_tol
  int precision = 0, refine_maxit = 5;
  double refine_rtol = 1e-10, refine_tol = 1e-4, refine_first_tol = 1e-7;
  // option refine_cycles (mixed form): each GMRES solve of the Newton step and of its refinement
  // corrections stops after at most this many restart cycles (0: no cap), so that a correction
  // starts from the FP64 residual where plain restarted GMRES would restart
  int refine_cycles = 0;
  int last_refinements = 0;
  // time dependence (DIRK): steps app.dt, stage tables c, d (= A^{-1}, column-major), t
  int tdep = 0, torder = 1, tstages = 1, tscheme = 0, tdfunc = 0;
  std::vector<double> dts;
  double time0 = 0.0, tnow = 0.0, dtfactor = 0.0;
  double dirk_c[3] = {0, 0, 0}, dirk_d[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, dirk_t[3] = {0, 0, 0};
  DevBuf<double> sdg, udgprev, utmp, bdfhist;  // bdfhist: BDF history u^{n-k}, k < torder
  // w equation (ncw > 0): nodal w, its source and initial state, the DIRK step's start and stage
  // combination, the BDF history of w; alpha dw/dt + beta w = Sourcew (app.bin factor[1], [2])
  int ncw = 0;
  double dae_alpha = 1.0, dae_beta = 0.0;
  DevBuf<double> wdg, wsrc, wdg0, wprev, wtmp, wbdfhist;
  // mask (device, [ninst], may be null): update only the instances with a nonzero entry
  void get_w(const double* mask = nullptr);
  bool timing = false;
  double ws_budget_bytes = 16.0e9;

  // device data
  DevBuf<double> shapegt, shapegw, shapfgt, shapfgw, param, tau;
  // model library (model name "plugin:<path>"): its functions, uinf and the point staging buffer
  void* plugin_handle = nullptr;
  PluginFns plugin_fns;
  DevBuf<double> uinf_d, pbuf;
  void load_plugin();
  DevBuf<int> perm, elemcon, bf, f2e, info, ifirst;
  std::vector<int> bface_h;  // boundary faces E*nfe + l (bf > 0), ascending
  DevBuf<int> bface;
  DevBuf<long> src1, src2;
  DevBuf<double> xdg, egeo, fgeo, odg, og, ogf;
  DevBuf<double> udg, uh, udg0, uh0, MCs, MEs, Mass;
  DevBuf<double> H, DinvF, RuS, Rh, b, BE, bjaug, we, du, RuR, bres;
  // per-block workspace
  int neb = 0;
  DevBuf<double> aug, Bfull, Kt, Gfull, rv, rJ, W, fpt;
  // mixed precision: FP32 copies of H, DinvF, the LU factors of D', K' and the preconditioner blocks
  DevBuf<float> H32, DinvF32, LU32, K32, BE32, BEilu32, BEeas32, easaug32;
  DevBuf<float> aug32, Bfull32, Gfull32, W32;  // FP32 element-block workspace (mixed form)
  DevBuf<float> MCs32, MEs32, shapegt32;       // FP32 copies of M^-1 C, M^-1 E and the shape table
  bool q32_ready = false;
  bool dinvf_fp32 = false;  // the last assembly kept D^{-1}F in FP32 only (DinvF32)
  bool ilu_fp32 = false;    // the last assembly kept the block-ILU(0) factors in FP32 only (BEilu32)
  DevBuf<float> Lilu32;
  // option precond_bf16 (mixed form, block-ILU(0) or Schwarz blocks): the preconditioner's blocks
  // in BF16 (upper half of FP32)
  int precond_bf16 = 0;
  DevBuf<unsigned short> BEilu16, BEeas16;
  DevBuf<int> piv32;  // row exchanges of the blocked LU of D' ([n, ne], panel-relative)
  DevBuf<int> lupiv;     // this handle's pivots of the current LU panel (LU_PIV_WS per matrix of a batched LU)
  DevBuf<double> lutri;  // this handle's triangle inverses of the current LU panel (LU_TRI_WS per matrix)
  // elemental additive Schwarz: inverted element blocks, inversion scratch
  DevBuf<double> BEeas, easaug;
  int eas_chunk = 0;
  // superelement block-ILU(0)
  Bilu0Struct ilu;
  bool ilu_ready = false;
  DevBuf<int> iface, irow_ptr, icol_ind, irow_of, idiag, ieptr, ielim_ji, iuptr, iupd_jl, iupd_il, iupd_ji,
      ilptr, ilidx, ilcol, iuuptr, iuidx, iucol, islot1, islot2, iwseg, iwblk, iwvec;
  DevBuf<double> BEilu, Yilu, Xilu, Lilu;
  IluDev ilud{};
  // GMRES: Krylov basis V, vectors, dot partials; the Krylov scalars of every instance live on the
  // device (GmresDev: gHm, gsv, gcs, gsn, gnrmb, the activity mask segmask[0, ns), gjc, segy, giters)
  DevBuf<double> V, x, r, tmp, wv, partial, hcol, scal;
  DevBuf<double> gHm, gsv, gcs, gsn, gnrmb, gjc, gtol, gxmask;
  DevBuf<int> giters, gjcount, gbcount;
  // option gmres_graph (default 0): 1 replays the Arnoldi steps of the Newton loop's solves
  // (b -> x) from HIP graphs captured on first use, one per column i (the cycle's last column
  // with its restart work); gm_graphs[i] for the cycle length gm_graph_len.  Off by default: on
  // the batch workload the replays were measured slower than the direct launches (concurrent
  // batches 0.49 s against 0.43 s, one-instance solves 73 against 117 solves/s with the captures)
  int gm_graph = 0;
  std::vector<hipGraphExec_t> gm_graphs;
  int gm_graph_len = 0;
  void clear_graphs();
  // host memory the status kernels write the number of iterating instances into (one slot per
  // queued step), its device address, and an event per slot
  int* hstat = nullptr;
  int* dstat = nullptr;
  hipEvent_t gm_ev[GM_RING] = {};
  // option gmres_lag: steps queued ahead of the host's convergence check (1 .. GM_RING - 2)
  int gm_lag = 3;
  // option noalloc_chunk: largest element chunk of exa_matvec_noalloc (0: what the scratch holds)
  int noalloc_chunk = 0;
  // during a one-problem GMRES: its activity flag, read by the preconditioner apply and the
  // orthogonalisation kernels (steps queued after convergence do nothing); null otherwise
  const double* gm_act = nullptr;
  double* hpin = nullptr;  // pinned host scratch for small device->host copies
  Stager st;               // pinned staging of bulk host <-> device copies
  bool assembled = false;

  // instances: one problem, or a batch of independent problems with one element shape,
  // merged into one disconnected mesh (instance i owns elements eoff[i] .. eoff[i+1] and faces
  // foff[i] .. foff[i+1], each in its own standalone order) so that every element- and
  // face-local kernel serves the whole batch in one launch
  int ninst = 1;
  bool batch = false;
  int nparam = 0, ntau = 0;
  std::vector<double> inst_param, inst_tau;  // [nparam, ninst], [ntau, ninst]
  std::vector<long> eoff, foff;
  std::vector<std::string> datains;
  DevBuf<int> einst;
  struct SegDev {
    DevBuf<int> bseg, sbptr;
    DevBuf<long> bbeg, bend;
    SegMap map;
  };
  SegDev segT, segU;  // segments of trace vectors [N] and of element vectors [n, ne]
  DevBuf<double> segpart, segcoef, segy, segmask, dus;
  double* hseg = nullptr;  // pinned host scratch for per-instance scalars
  long hseg_len = 0;
  // during a batched GMRES: per-instance activity (1/0) read by the matvec and the Schwarz
  // preconditioner, whose elements of inactive instances only write zeros; null otherwise
  const double* elem_active = nullptr;
  std::vector<int> inst_newton, inst_gmres;
  std::vector<char> inst_failed;  // instances stopped by a Newton breakdown in this solve
  std::vector<double> inst_norm;
  // exa_batch_step state: instances still iterating (empty until the first step or refill), the
  // residual norms of the current state (valid unless a refill or a full solve came between) and
  // the Newton iterate
  std::vector<char> step_act;
  std::vector<double> step_nr;
  bool step_valid = false;
  DevBuf<double> step_u;
  // time-dependent steps: each instance's time and number of time steps done; while use_tinst the
  // assembly reads each instance's model time from tinst_d
  std::vector<double> inst_t;
  std::vector<int> inst_steps, inst_nsteps;  // time steps done, and each instance's number of steps
  DevBuf<double> tinst_d;
  bool use_tinst = false;

  // timing
  std::array<std::vector<std::pair<hipEvent_t, hipEvent_t>>, T_NCLASS> ev;
  std::array<double, T_NCLASS> ms{};
  std::array<long, T_NCLASS> cnt{};

  Solver(const std::string& datain, const std::string& dataout, const std::string& model, int device);
  // a batch: the datain directories of ninst instances with one element shape and model
  Solver(const std::vector<std::string>& datains, const std::string& model, int device);
  ~Solver();
  void set_dims();
  int* lu_ws(long batch);
  // FP32 LU of D' and of the Schwarz blocks: explicit triangle inverses applied by GEMMs (true) or
  // substitution (false, option lu_inverses = 0)
  bool lu_inverses = true;
  void* lu_tri(long batch);  // triangle-inverse workspace of the blocked LU (LU_TRI_WS doubles per matrix)
  void init(SolData& sol);
  void setup_q(long e0, long ne);  // q-equation matrices (MCs, MEs, Mass) of elements [e0, e0 + ne)
  // batch (segmented) Newton-GMRES
  void seg_setup();
  void build_seg(SegDev& sd, const std::vector<long>& off);
  void seg_alloc();
  double* host_region(int which);
  void seg_norms(const SegDev& sd, const double* v, double* out_dev);
  std::vector<double> residual_seg();
  int gmres_seg(const double* bvec, double* xvec, double tol_rel, std::vector<int>& iters, int maxit = -1);
  // only (may be null): the instances to iterate (the others keep their state)
  void newton_seg(int* newton_iters, int* gmres_iters, double* final_norm, const std::vector<char>* only = nullptr);
  void newton_seg_iter(std::vector<char>& act, std::vector<double>& nr, double* u, int* gmres_total);
  // iteration-level batching: one Newton iteration per call; a finished slot takes a new problem
  void step_init();
  int newton_seg_step(int* status);
  int dirk_seg_step(int* status);
  void dirk_step(double dt, double t, const std::vector<char>* only, int* nit, int* gtot, double* nrm, int istep);
  void refill(int inst, const std::string& datain);
  long get_inst_array(int inst, const std::string& name, double* dst, long len);
  void restore_instance(int inst, double* newton_iterate);

  void set_option(const std::string& name, double value);
  // stages
  void reset();
  void get_q();
  void assemble();
  void residual(double* nb, double* nru);
  void residual_fields();  // RuR, Rh and the assembled trace residual bres
  void matvec(const double* v, double* w);
  void apply_precond(double* x);
  void apply_precond_to(const double* in, double* out);
  int gmres(const double* bvec, double* xvec, double tol_rel = -1.0, int maxit = -1);
  // every instance's GMRES at once; iters[i] = instance i's count; returns the largest
  // maxit < 0: gmres_maxit
  int gmres_run(const double* bvec, double* xvec, double tol_rel, std::vector<int>& iters, int maxit = -1);
  // iteration cap of the mixed form's solves (refine_cycles restart cycles, or gmres_maxit)
  int refine_maxit_gmres() const {
    return refine_cycles > 0 ? std::min(gmres_maxit, refine_cycles * restart) : gmres_maxit;
  }
  long augcols() const { return static_cast<long>(d.n) + d.m + 1; }
  void alloc_mixed();
  void solve(int* newton_iters, int* gmres_iters, double* final_norm);
  void newton(int* newton_iters, int* gmres_iters, double* final_norm);
  void solve_bdf(int* newton_iters, int* gmres_iters, double* final_norm);
  void dirk_coefficients();
  void write_output();
  long get_array(const std::string& name, double* dst, long len);
  void set_array(const std::string& name, const double* src, long len);
  void device_ptr(const std::string& name, void** p, long* len);
  int timings(double* out_ms, long* counts, int len);
  void reset_timings();

 private:
  void setup();
  void setup_bilu0();
  void factor_bilu0();
  template <typename T>
  void factor_bilu0_t(T* BE, T* L);
  void setup_eas();
  void alloc_workspace();
  AsmArgs asm_args(int e0, int eb);
  double norm(const double* x, long n);
  // index of the first nonzero entry of the device array info[0..n), or -1 (synchronises)
  int first_nonzero(const int* info, long n);
  GmresDev gmres_state();
  void gmres_step(const GmresDev& g, int i, double* xv);
  void gmres_cycle_end(const GmresDev& g, int i, const double* bv, double* xv);
  void gmres_update_x(const GmresDev& g, int k, double* xv);
  void tic(int cls, hipEvent_t* a);
  void toc(int cls, hipEvent_t a);
  void collect_timings();
};

}  // namespace exa
