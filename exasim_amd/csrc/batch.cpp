// batch.cpp -- independent problems solved together: a batch of instances with one
// element shape (nd, p, quadrature, ncu, model) is merged into one disconnected mesh, so
// that assembly, condensation, the matvec and the block-Jacobi / Schwarz preconditioners
// run as single launches over every instance's elements and faces.  The Krylov and
// Newton scalars stay per instance: dot products, norms, Givens rotations, tolerances,
// damping and convergence are those of each instance alone (segmented reductions over the
// instance's own contiguous slice of each vector), so an instance's iterates and result do
// not depend on what else is in the batch.
//
// Per-instance algorithm: GMRES as backend/Solver/gmres.cpp:361-597 and NewtonSolver as
// backend/Solution/solution.cpp:125-242 (see solver.cpp for the single-instance forms).
#include "solver.hpp"
#include "models.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

namespace exa {

namespace {

// Time steps two instances of a batch can share: the same list, or lists of one step length
// (of any number of steps: exa_batch_step takes each instance through its own)
bool same_steps(const std::vector<double>& a, const std::vector<double>& b) {
  if (a == b) return true;
  if (a.empty() || b.empty() || a[0] != b[0]) return false;
  for (double v : a)
    if (v != a[0]) return false;
  for (double v : b)
    if (v != b[0]) return false;
  return true;
}

bool same_master(const MasterData& a, const MasterData& b) {
  return a.nd == b.nd && a.elemtype == b.elemtype && a.porder == b.porder && a.npe == b.npe &&
         a.npf == b.npf && a.nge == b.nge && a.ngf == b.ngf && a.shapegt == b.shapegt &&
         a.shapegw == b.shapegw && a.shapfgt == b.shapfgt && a.shapfgw == b.shapfgw;
}

template <typename T>
void append(std::vector<T>& dst, const std::vector<T>& src) {
  dst.insert(dst.end(), src.begin(), src.end());
}

}  // namespace

Solver::Solver(const std::vector<std::string>& dins, const std::string& mname, int dev)
    : datain(dins.empty() ? std::string() : dins[0]), model_name(mname), model(0), device(dev) {
  if (dins.empty()) throw Error(1, "a batch needs at least one datain directory");
  model = resolve_model(mname);
  batch = true;
  ninst = static_cast<int>(dins.size());
  datains = dins;
  SolData sol;
  std::vector<std::vector<double>> uhs;
  int npmax = 1, ntmax = 1;
  std::vector<AppData> apps;
  for (int i = 0; i < ninst; ++i) {
    const std::string& din = dins[i];
    AppData a = read_app(din + "/app.bin");
    MasterData ma = read_master(din + "/master.bin");
    MeshData me = read_mesh(din + "/mesh.bin");
    if (ma.nd != me.nd) throw Error(2, din + ": master.bin and mesh.bin disagree on nd");
    const int nc = static_cast<int>(a.ndims[5]), ncu = static_cast<int>(a.ndims[6]);
    if (i == 0) {
      app = a;
      master = ma;
      mesh = me;
      set_dims();
    } else {
      if (!same_master(master, ma))
        throw Error(1, din + ": a batch needs one master element (same nd, p and quadrature)");
      if (nc != d.nc || ncu != d.ncu || me.nfe != d.nfe || me.perm != mesh.perm)
        throw Error(1, din + ": a batch needs one (nc, ncu) and one face numbering");
      if (a.flag[0] != app.flag[0] || a.flag[6] != app.flag[6] || a.problem[2] != app.problem[2] ||
          a.problem[3] != app.problem[3] || a.problem[4] != app.problem[4] || !same_steps(a.dt, app.dt))
        throw Error(1, din + ": a batch needs one time scheme, time steps and preconditioner");
    }
    inst_nsteps.push_back(static_cast<int>(a.dt.size()));
    SolData so = read_sol(din + "/sol.bin", ma.npe, nc, ncu, me.ne);
    if (i > 0 && so.nco != sol.nco) throw Error(1, din + ": a batch needs one auxiliary-field width");
    Conn c = build_conn(ma, me);
    const long e0 = i == 0 ? 0 : eoff.back(), f0 = i == 0 ? 0 : foff.back();
    if (i == 0) {
      eoff = {0};
      foff = {0};
      sol.nco = so.nco;
      sol.udg_ncomp = d.nc;
      mesh.bf.clear();
      conn = Conn{};
    }
    eoff.push_back(e0 + me.ne);
    foff.push_back(f0 + c.nf);
    // element data: [npe, *, ne] arrays concatenate over elements; udg gets all nc components
    append(sol.xdg, so.xdg);
    std::vector<double> u(static_cast<size_t>(ma.npe) * d.nc * me.ne, 0.0);
    if (!so.udg.empty()) {
      const int k = so.udg_ncomp;
      for (long e = 0; e < me.ne; ++e)
        for (int cc = 0; cc < k; ++cc)
          for (int j = 0; j < ma.npe; ++j)
            u[j + ma.npe * (cc + static_cast<long>(d.nc) * e)] = so.udg[j + ma.npe * (cc + static_cast<long>(k) * e)];
    }
    append(sol.udg, u);
    append(sol.odg, so.odg);
    append(sol.wdg, so.wdg);
    if (i > 0 && (a.ndims.size() > 13 ? a.ndims[13] : 0.0) != (app.ndims.size() > 13 ? app.ndims[13] : 0.0))
      throw Error(1, din + ": a batch needs one w-equation width (ncw)");
    if (i > 0 && a.factor != app.factor)
      throw Error(1, din + ": a batch needs one initial time and one w-equation alpha, beta (app factor)");
    uhs.push_back(so.uh);
    append(mesh.bf, me.bf);
    // connectivity with element, face and face-node offsets
    for (int f = 0; f < c.nf; ++f) {
      conn.f2e.push_back(c.f2e[4 * f] + static_cast<int>(e0));
      conn.f2e.push_back(c.f2e[4 * f + 1]);
      conn.f2e.push_back(c.f2e[4 * f + 2] >= 0 ? c.f2e[4 * f + 2] + static_cast<int>(e0) : -1);
      conn.f2e.push_back(c.f2e[4 * f + 3]);
    }
    for (int v : c.elemcon) conn.elemcon.push_back(v + static_cast<int>(f0) * ma.npf);
    conn.nif += c.nif;
    npmax = std::max<int>(npmax, static_cast<int>(a.physicsparam.size()));
    ntmax = std::max<int>(ntmax, static_cast<int>(a.tau.size()));
    apps.push_back(std::move(a));
  }
  conn.nf = static_cast<int>(foff.back());
  mesh.ne = static_cast<int>(eoff.back());
  mesh.cartgridpart.clear();
  mesh.t.clear();
  d.ne = mesh.ne;
  // initial traces: all given, or all taken from the elements (init() does that)
  const bool any_uh = std::any_of(uhs.begin(), uhs.end(), [](const std::vector<double>& v) { return !v.empty(); });
  const bool all_uh = std::all_of(uhs.begin(), uhs.end(), [](const std::vector<double>& v) { return !v.empty(); });
  if (any_uh && !all_uh) throw Error(1, "a batch needs uhat in every sol.bin or in none");
  if (all_uh)
    for (auto& v : uhs) append(sol.uh, v);
  // per-instance physicsparam and tau, zero-padded to the longest
  nparam = npmax;
  ntau = ntmax;
  inst_param.assign(static_cast<size_t>(nparam) * ninst, 0.0);
  inst_tau.assign(static_cast<size_t>(ntau) * ninst, 0.0);
  for (int i = 0; i < ninst; ++i) {
    std::copy(apps[i].physicsparam.begin(), apps[i].physicsparam.end(), inst_param.begin() + static_cast<long>(i) * nparam);
    std::copy(apps[i].tau.begin(), apps[i].tau.end(), inst_tau.begin() + static_cast<long>(i) * ntau);
  }
  if (static_cast<int>(app.flag[6]) == 2)
    throw Error(5, "block-ILU(0) (preconditioner 2) needs one structured grid; a batch uses "
                   "preconditioner 0 or 1");
  init(sol);
}

// Block maps of the segmented vector kernels: blocks of at most SEG_CHUNK entries that start at
// each segment's first entry.
void Solver::build_seg(SegDev& sd, const std::vector<long>& off) {
  const int ns = static_cast<int>(off.size()) - 1;
  std::vector<int> bseg, sbptr(ns + 1, 0);
  std::vector<long> bbeg, bend;
  for (int sg = 0; sg < ns; ++sg) {
    sbptr[sg] = static_cast<int>(bseg.size());
    for (long i = off[sg]; i < off[sg + 1]; i += SEG_CHUNK) {
      bseg.push_back(sg);
      bbeg.push_back(i);
      bend.push_back(std::min(off[sg + 1], i + SEG_CHUNK));
    }
  }
  sbptr[ns] = static_cast<int>(bseg.size());
  if (bseg.empty()) {  // all segments empty: keep valid device pointers
    bseg.push_back(0);
    bbeg.push_back(0);
    bend.push_back(0);
  }
  sd.bseg.upload(bseg, st, s);
  sd.sbptr.upload(sbptr, st, s);
  sd.bbeg.upload(bbeg, st, s);
  sd.bend.upload(bend, st, s);
  sd.map.nseg = ns;
  sd.map.nblk = sbptr[ns] > 0 ? static_cast<int>(bseg.size()) : 0;
  sd.map.bseg = sd.bseg.p;
  sd.map.bbeg = sd.bbeg.p;
  sd.map.bend = sd.bend.p;
  sd.map.sbptr = sd.sbptr.p;
}

void Solver::seg_setup() {
  std::vector<long> ot(ninst + 1), ou(ninst + 1);
  for (int i = 0; i <= ninst; ++i) {
    ot[i] = foff[i] * d.ncf;
    ou[i] = eoff[i] * d.n;
  }
  build_seg(segT, ot);
  build_seg(segU, ou);
  if (batch) dus.alloc(static_cast<size_t>(d.n) * d.ne);
  seg_alloc();
}

// Buffers of the segmented GMRES, sized by the restart length (called again when it changes,
// with the stream idle).  Device: partial sums for up to restart+1 vectors, the Hessenberg
// column, y, norms and coefficients ([0, 2ns) norms and GMRES coefficients, [3ns, 4ns) Newton
// damping), masks ([0, ns) GMRES, [ns, 2ns) Newton).  Host (pinned): disjoint regions for the
// column pulled back each step and for each kind of upload, so an upload still in flight is
// never overwritten by the next one.
void Solver::seg_alloc() {
  const long ns = ninst, n1 = restart + 1;
  const long nb = std::max(segT.map.nblk, segU.map.nblk);
  EXA_HIP(hipStreamSynchronize(s));
  segpart.alloc(static_cast<size_t>(std::max(nb, 1L)) * (n1 + 1));
  hcol.alloc(static_cast<size_t>(n1 + 1) * ns);
  segy.alloc(static_cast<size_t>(n1) * ns);
  segcoef.alloc(static_cast<size_t>(4 * ns));
  segmask.alloc(static_cast<size_t>(2 * ns));
  // device-resident GMRES state (the Krylov scalars of every instance)
  gHm.alloc(static_cast<size_t>(ns) * n1 * n1);
  gsv.alloc(static_cast<size_t>(ns) * (n1 + 1));
  gcs.alloc(static_cast<size_t>(ns) * (n1 + 1));
  gsn.alloc(static_cast<size_t>(ns) * (n1 + 1));
  gnrmb.alloc(ns);
  gjc.alloc(ns);
  giters.alloc(ns);
  gtol.alloc(1);
  gjcount.alloc(2);  // step counter, completion ticket of the step kernel
  gxmask.alloc(ns);
  gbcount.alloc((ns + 15) / 16);
  clear_graphs();  // captured launches hold the old buffers
  if (!hstat) {  // per-step counts of iterating instances, written by the kernels into host memory
    EXA_HIP(hipHostMalloc(reinterpret_cast<void**>(&hstat), sizeof(int) * GM_RING,
                          hipHostMallocMapped | hipHostMallocCoherent));
    EXA_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&dstat), hstat, 0));
    for (int k = 0; k < GM_RING; ++k) {
      hstat[k] = -1;
      EXA_HIP(hipEventCreateWithFlags(&gm_ev[k], hipEventDisableTiming));
    }
  }
  const long need = (n1 + 1) * ns + 2 * ns + n1 * ns + 2 * ns;
  if (need > hseg_len) {
    if (hseg) EXA_HIP(hipHostFree(hseg));
    hseg = nullptr;
    EXA_HIP(hipHostMalloc(reinterpret_cast<void**>(&hseg), sizeof(double) * need, hipHostMallocDefault));
    hseg_len = need;
  }
}

double* Solver::host_region(int which) {
  const long ns = ninst, n1 = restart + 1;
  switch (which) {
    case 0: return hseg;                                   // pulls, (n1+1)*ns
    case 1: return hseg + (n1 + 1) * ns;                   // GMRES mask / coefficients, 2*ns
    case 2: return hseg + (n1 + 1) * ns + 2 * ns;          // GMRES y, n1*ns
    default: return hseg + (n1 + 1) * ns + 2 * ns + n1 * ns;  // Newton mask / coefficients, 2*ns
  }
}

void Solver::seg_norms(const SegDev& sd, const double* v, double* out_dev) {
  seg_multidot(sd.map, v, 0, 1, v, segpart.p, out_dev, 1, s);
}

// per instance ||b|| + ||Ru|| of the current state (hdgAssembleResidual, Newton's norm)
std::vector<double> Solver::residual_seg() {
  residual_fields();
  seg_norms(segT, bres.p, segcoef.p);
  seg_norms(segU, RuR.p, segcoef.p + ninst);
  double* h = host_region(0);
  EXA_HIP(hipMemcpyAsync(h, segcoef.p, sizeof(double) * 2 * ninst, hipMemcpyDeviceToHost, s));
  EXA_HIP(hipStreamSynchronize(s));
  std::vector<double> out(ninst);
  for (int i = 0; i < ninst; ++i) out[i] = h[i] + h[ninst + i];
  return out;
}

// Newton per instance over shared launches: an instance stops when its own residual norm is
// below newton_tol; its damping (alpha halved while its residual grows) and, in the mixed
// form, its refinement loop are its own.  Converged instances keep their state (their right-hand
// sides and updates are masked to zero).
void Solver::newton_seg(int* newton_iters, int* gmres_total, double* final_norm, const std::vector<char>* only) {
  const long N = d.N;
  const int ns = ninst;
  if (static_cast<int>(inst_newton.size()) != ns) {
    inst_newton.assign(ns, 0);
    inst_gmres.assign(ns, 0);
    inst_norm.assign(ns, 0.0);
  }
  if (static_cast<int>(inst_failed.size()) != ns) inst_failed.assign(ns, 0);
  get_q();
  std::vector<double> nr = residual_seg();
  if (verbose) {
    double mx = *std::max_element(nr.begin(), nr.end());
    std::printf("batch Newton Iteration: 0, largest residual norm %g over %d instances\n", mx, ns);
  }
  DevBuf<double> u;
  u.alloc(N);
  copy(u.p, uh.p, N, s);
  std::vector<char> act(ns, 1);
  for (int i = 0; i < ns; ++i) act[i] = (only ? (*only)[i] : 1) && !inst_failed[i];
  int it = 0, gtot = 0;
  for (it = 0; it < newton_maxit; ++it) {
    if (std::none_of(act.begin(), act.end(), [](char c) { return c != 0; })) break;
    newton_seg_iter(act, nr, u.p, &gtot);
    if (verbose) {
      int left = static_cast<int>(std::count(act.begin(), act.end(), 1));
      double mx = *std::max_element(nr.begin(), nr.end());
      std::printf("batch Newton Iteration: %d, largest residual norm %g, %d instances still iterating\n", it + 1,
                  mx, left);
      std::fflush(stdout);
    }
  }
  EXA_HIP(hipStreamSynchronize(s));
  if (newton_iters) *newton_iters = it;
  if (gmres_total) *gmres_total = gtot;
  if (final_norm) {
    double mx = 0.0;
    for (int i = 0; i < ns; ++i)
      if (!inst_failed[i]) mx = std::max(mx, nr[i]);
    *final_norm = mx;
  }
}

// One Newton iteration (solution.cpp:125-242, the body of NewtonSolver's loop) of every instance
// with act[i] set, from the state (udg, uh = u) whose residual norms are nr: assembly, GMRES,
// damping and, in the mixed form, the refinement with FP64 residuals.  On return nr holds the
// norms of the new state, act[i] is cleared for instances that converged (nr < newton_tol) or
// broke down, and the per-instance counts have grown.
void Solver::newton_seg_iter(std::vector<char>& act, std::vector<double>& nr, double* u, int* gtot) {
  const long N = d.N;
  const int ns = ninst;
  // Newton's own device mask (segmask + ns) and damping coefficients (segcoef + 3 ns), uploaded from
  // their own host region; each upload is followed by a synchronising pull before the next
  double* mask_d = segmask.p + ns;
  double* coef_d = segcoef.p + 3L * ns;
  auto push = [&](double* dst, const std::vector<double>& v) {
    double* hh = host_region(3) + (dst == mask_d ? 0 : ns);
    std::memcpy(hh, v.data(), sizeof(double) * v.size());
    EXA_HIP(hipMemcpyAsync(dst, hh, sizeof(double) * v.size(), hipMemcpyHostToDevice, s));
  };
  auto mask_of = [&](const std::vector<char>& a) {
    std::vector<double> v(ns);
    for (int i = 0; i < ns; ++i) v[i] = a[i] ? 1.0 : 0.0;
    return v;
  };
  // an instance whose Newton iteration breaks down (NaN residual, growth beyond 1e6) stops:
  // it is marked failed (final norm NaN), and its input state is put back so that the shared
  // launches keep seeing finite data; the other instances go on
  auto fail = [&](int i, const char* why) {
    inst_failed[i] = 1;
    act[i] = 0;
    inst_norm[i] = std::numeric_limits<double>::quiet_NaN();
    restore_instance(i, u);
    if (verbose) std::printf("batch: instance %d stopped: %s\n", i, why);
  };
  std::vector<int> its;
  assemble();
  push(mask_d, mask_of(act));
  seg_scale(segT.map, b.p, b.p, mask_d, 0, s);
  *gtot += gmres_seg(b.p, x.p, precision ? std::max(gmres_tol, refine_first_tol) : -1.0, its,
                     precision ? refine_maxit_gmres() : -1);
  for (int i = 0; i < ns; ++i)
    if (act[i]) {
      inst_gmres[i] += its[i];
      ++inst_newton[i];
    }
  axpby(u, x.p, 1.0, 1.0, N, s);
  copy(uh.p, u, N, s);
  if (precision) dudg(d, RuS.p, DinvF32.p, elemcon.p, x.p, du.p, s);
  else dudg(d, RuS.p, DinvF.p, elemcon.p, x.p, du.p, s);
  seg_scale(segU.map, du.p, du.p, mask_d, 0, s);
  update_udg(d, du.p, 1.0, udg.p, s);
  get_q();
  // GetW of the instances that moved: another pass over a converged w changes its last bits, so
  // an instance's w would depend on its neighbours' iterations
  get_w(mask_d);
  std::vector<double> nrm0 = nr;
  nr = residual_seg();
  std::vector<double> alpha(ns, 1.0);
  for (int i = 0; i < ns; ++i)
    if (act[i] && nr[i] > nrm0[i] && nr[i] > 1e6) fail(i, "residual norm increases more than 1e6");
  // damping, per instance: alpha halved while its residual grows (or is NaN)
  for (;;) {
    std::vector<double> coef(ns, 0.0);
    bool any = false;
    for (int i = 0; i < ns; ++i)
      if (act[i] && ((nr[i] > nrm0[i] && alpha[i] > 0.1) || (std::isnan(nr[i]) && alpha[i] > 1e-6))) {
        alpha[i] /= 2.0;
        coef[i] = -alpha[i];
        any = true;
      }
    if (!any) break;
    push(coef_d, coef);
    seg_axpy(segT.map, u, x.p, coef_d, 1.0, s);
    copy(uh.p, u, N, s);
    seg_scale(segU.map, dus.p, du.p, coef_d, 0, s);
    update_udg(d, dus.p, 1.0, udg.p, s);
    get_q();
    get_w(coef_d);  // the damped instances (nonzero coefficient)
    std::vector<double> nn = residual_seg();
    for (int i = 0; i < ns; ++i)
      if (coef[i] != 0.0) nr[i] = nn[i];
  }
  for (int i = 0; i < ns; ++i)
    if (act[i] && std::isnan(nr[i])) fail(i, "NaN residual");
  // mixed form: refinement with the FP64 residual, per instance
  last_refinements = 0;
  if (precision) {
    for (;;) {
      std::vector<char> ref(ns, 0);
      bool any = false;
      for (int i = 0; i < ns; ++i) {
        ref[i] = act[i] && nr[i] > refine_rtol * nrm0[i];
        any = any || ref[i];
      }
      if (!any || last_refinements >= refine_maxit) break;
      refine_rhs(d, LU32.p, piv32.p, K32.p, RuR.p, RuS.p, Rh.p, s);
      face_gather(static_cast<long>(d.npf) * d.nf, d.ncu, src1.p, src2.p, Rh.p, b.p, s);
      push(mask_d, mask_of(ref));
      seg_scale(segT.map, b.p, b.p, mask_d, 0, s);
      *gtot += gmres_seg(b.p, x.p, refine_tol, its, refine_maxit_gmres());
      for (int i = 0; i < ns; ++i)
        if (ref[i]) inst_gmres[i] += its[i];
      axpby(u, x.p, 1.0, 1.0, N, s);
      copy(uh.p, u, N, s);
      dudg(d, RuS.p, DinvF32.p, elemcon.p, x.p, du.p, s);
      seg_scale(segU.map, du.p, du.p, mask_d, 0, s);
      update_udg(d, du.p, 1.0, udg.p, s);
      get_q();
      get_w(mask_d);
      std::vector<double> nn = residual_seg();
      for (int i = 0; i < ns; ++i)
        if (ref[i]) nr[i] = nn[i];
      ++last_refinements;
      for (int i = 0; i < ns; ++i)
        if (ref[i] && std::isnan(nr[i])) fail(i, "NaN in refinement");
    }
  }
  for (int i = 0; i < ns; ++i)
    if (act[i]) {
      inst_norm[i] = nr[i];
      if (nr[i] < newton_tol) act[i] = 0;
    }
}

// Iteration-level batching: one Newton iteration of every instance that still iterates, from the
// state the previous step (or exa_batch_create / refill) left.  A sequence of steps gives every
// instance the iterates of exa_solve bit for bit: the step starts from the residual norms the
// previous step ended with (computed again after a refill), and each instance's damping,
// refinement and GMRES scalars are its own.  status[i]: 1 still iterating, 0 converged or at
// newton_maxit, -1 broken down.  Returns the number still iterating.
int Solver::newton_seg_step(int* status) {
  if (!batch) throw Error(1, "exa_batch_step needs a batch handle");
  if (tdep) return dirk_seg_step(status);
  const int ns = ninst;
  step_init();
  if (!step_valid) {
    get_q();
    step_nr = residual_seg();
    step_u.alloc(d.N);
    step_valid = true;
  }
  copy(step_u.p, uh.p, d.N, s);
  std::vector<char> act = step_act;
  for (int i = 0; i < ns; ++i) act[i] = act[i] && !inst_failed[i];
  if (std::any_of(act.begin(), act.end(), [](char c) { return c != 0; })) {
    int gtot = 0;
    newton_seg_iter(act, step_nr, step_u.p, &gtot);
  }
  int left = 0;
  for (int i = 0; i < ns; ++i) {
    if (act[i] && inst_newton[i] >= newton_maxit) act[i] = 0;  // NewtonSolver's iteration limit
    step_act[i] = act[i];
    left += act[i] ? 1 : 0;
    if (status) status[i] = act[i] ? 1 : (inst_failed[i] ? -1 : 0);
  }
  EXA_HIP(hipStreamSynchronize(s));
  return left;
}

// Time-dependent problems: one DIRK time step (all stages, each with its Newton solve) of every
// instance that still has time steps to take, each at its own time (an instance put in by refill
// starts at the initial time while the others go on).  The time-step length must be one for all
// steps; BDF keeps a history whose order depends on the steps taken, and model libraries receive
// one time for all points, so both run through exa_solve only.  status as for steady steps (0:
// all its time steps done).
int Solver::dirk_seg_step(int* status) {
  if (tscheme != 0) throw Error(5, "exa_batch_step runs DIRK time stepping; BDF batches use exa_solve");
  if (model == MODEL_PLUGIN)
    throw Error(5, "exa_batch_step: a model library takes one time for all points; use exa_solve");
  for (double v : dts)
    if (v != dts[0]) throw Error(5, "exa_batch_step needs one time-step length for every step");
  if (dts.empty()) throw Error(2, "app.bin has no time steps");
  const int ns = ninst;
  step_init();
  std::vector<char> act(ns);
  for (int i = 0; i < ns; ++i) act[i] = step_act[i] && !inst_failed[i];
  if (std::any_of(act.begin(), act.end(), [](char c) { return c != 0; })) {
    tinst_d.alloc(ns);
    use_tinst = true;
    int nit = 0, gtot = 0;
    double nrm = 0.0;
    try {
      dirk_step(dts[0], time0, &act, &nit, &gtot, &nrm, 0);
    } catch (...) {
      use_tinst = false;
      throw;
    }
    use_tinst = false;
  }
  int left = 0;
  for (int i = 0; i < ns; ++i) {
    if (act[i] && !inst_failed[i]) {
      inst_t[i] += dts[0];
      ++inst_steps[i];
    }
    step_act[i] = act[i] && !inst_failed[i] && inst_steps[i] < inst_nsteps[i];
    left += step_act[i] ? 1 : 0;
    if (status) status[i] = step_act[i] ? 1 : (inst_failed[i] ? -1 : 0);
  }
  EXA_HIP(hipStreamSynchronize(s));
  return left;
}

// Step state at the first exa_batch_step or refill after creation, exa_solve or reset: every
// instance is pending (it iterates at the next step) with its counts at zero.
void Solver::step_init() {
  if (static_cast<int>(step_act.size()) == ninst) return;
  step_act.assign(ninst, 1);
  inst_t.assign(ninst, time0);
  inst_steps.assign(ninst, 0);
  inst_newton.assign(ninst, 0);
  inst_gmres.assign(ninst, 0);
  inst_norm.assign(ninst, 0.0);
  inst_failed.assign(ninst, 0);
  step_valid = false;
}

// Load a new problem into instance slot i of a batch, whose previous problem has finished
// (converged, stopped or broken down): the new problem needs the slot's element shape and mesh
// (the same mesh.bin connectivity, boundary faces and sizes) and may differ in coordinates,
// initial state, auxiliary fields, physicsparam and tau.  The slot's geometry and q-equation
// matrices are computed again, its Newton state starts afresh, and the next exa_batch_step
// iterates it together with the instances already under way.
void Solver::refill(int i, const std::string& din) {
  if (!batch) throw Error(1, "refill needs a batch handle");
  if (i < 0 || i >= ninst) throw Error(1, "instance index out of range");
  if (tdep && tscheme != 0) throw Error(5, "refill runs steady and DIRK problems");
  step_init();
  if (step_act[i]) throw Error(1, "instance " + std::to_string(i) + " has not finished");
  AppData a = read_app(din + "/app.bin");
  MasterData ma = read_master(din + "/master.bin");
  MeshData me = read_mesh(din + "/mesh.bin");
  if (ma.nd != me.nd) throw Error(2, din + ": master.bin and mesh.bin disagree on nd");
  const int nc = static_cast<int>(a.ndims[5]), ncu = static_cast<int>(a.ndims[6]);
  if (!same_master(master, ma))
    throw Error(1, din + ": a batch needs one master element (same nd, p and quadrature)");
  if (nc != d.nc || ncu != d.ncu || me.nfe != d.nfe || me.perm != mesh.perm)
    throw Error(1, din + ": a batch needs one (nc, ncu) and one face numbering");
  if (a.flag[0] != app.flag[0] || a.flag[6] != app.flag[6] || a.problem[2] != app.problem[2] ||
      a.problem[3] != app.problem[3] || a.problem[4] != app.problem[4] || !same_steps(a.dt, app.dt))
    throw Error(1, din + ": a batch needs one time scheme, time steps and preconditioner");
  if ((a.ndims.size() > 13 ? a.ndims[13] : 0.0) != (app.ndims.size() > 13 ? app.ndims[13] : 0.0) ||
      a.factor != app.factor)
    throw Error(1, din + ": a batch needs one w-equation width and one app factor");
  const long e0 = eoff[i], ne_i = eoff[i + 1] - eoff[i];
  const long f0 = foff[i], nf_i = foff[i + 1] - foff[i];
  if (me.ne != ne_i) throw Error(1, din + ": a refill needs the mesh of the instance it replaces (ne differs)");
  Conn c = build_conn(ma, me);
  bool same = c.nf == nf_i;
  for (long f = 0; same && f < nf_i; ++f) {
    const int* g = &conn.f2e[4 * (f0 + f)];
    const int* h = &c.f2e[4 * f];
    same = g[0] == h[0] + e0 && g[1] == h[1] && g[2] == (h[2] >= 0 ? h[2] + e0 : -1) && g[3] == h[3];
  }
  const long nel = static_cast<long>(d.npf) * d.nfe;
  for (long k = 0; same && k < nel * ne_i; ++k)
    same = conn.elemcon[e0 * nel + k] == c.elemcon[k] + static_cast<int>(f0) * d.npf;
  for (long k = 0; same && k < static_cast<long>(d.nfe) * ne_i; ++k) same = mesh.bf[e0 * d.nfe + k] == me.bf[k];
  if (!same) throw Error(1, din + ": a refill needs the mesh of the instance it replaces (connectivity differs)");
  if (static_cast<int>(a.physicsparam.size()) > nparam || static_cast<int>(a.tau.size()) > ntau)
    throw Error(1, din + ": physicsparam or tau longer than the batch's");
  SolData so = read_sol(din + "/sol.bin", ma.npe, nc, ncu, me.ne);
  if (so.nco != d.nco) throw Error(1, din + ": a batch needs one auxiliary-field width");
  if (ncw > 0 && !so.wdg.empty() && static_cast<long>(so.wdg.size()) != static_cast<long>(d.npe) * ncw * ne_i)
    throw Error(2, din + ": wdg in sol.bin has the wrong size");
  if (!so.uh.empty() && static_cast<long>(so.uh.size()) != nf_i * d.ncf)
    throw Error(2, din + ": uhat in sol.bin has the wrong size");
  EXA_HIP(hipStreamSynchronize(s));
  // per-instance parameters
  std::fill(inst_param.begin() + static_cast<long>(i) * nparam, inst_param.begin() + static_cast<long>(i + 1) * nparam, 0.0);
  std::fill(inst_tau.begin() + static_cast<long>(i) * ntau, inst_tau.begin() + static_cast<long>(i + 1) * ntau, 0.0);
  std::copy(a.physicsparam.begin(), a.physicsparam.end(), inst_param.begin() + static_cast<long>(i) * nparam);
  std::copy(a.tau.begin(), a.tau.end(), inst_tau.begin() + static_cast<long>(i) * ntau);
  st.h2d(param.p, inst_param.data(), sizeof(double) * inst_param.size(), s);
  st.h2d(tau.p, inst_tau.data(), sizeof(double) * inst_tau.size(), s);
  // element fields of the slot: coordinates, udg (all nc components), auxiliary fields, w
  const long npe = d.npe;
  st.h2d(xdg.p + e0 * npe * d.nd, so.xdg.data(), sizeof(double) * so.xdg.size(), s);
  std::vector<double> hu(static_cast<size_t>(npe) * d.nc * ne_i, 0.0);
  if (!so.udg.empty()) {
    const int k = so.udg_ncomp;
    for (long e = 0; e < ne_i; ++e)
      for (int cc = 0; cc < k; ++cc)
        for (long j = 0; j < npe; ++j) hu[j + npe * (cc + d.nc * e)] = so.udg[j + npe * (cc + k * e)];
  }
  st.h2d(udg.p + e0 * npe * d.nc, hu.data(), sizeof(double) * hu.size(), s);
  st.h2d(udg0.p + e0 * npe * d.nc, hu.data(), sizeof(double) * hu.size(), s);
  std::vector<double> huh = so.uh;
  if (huh.empty()) {  // GetFaceNodes from the first element of each face, as at creation
    huh.assign(static_cast<size_t>(nf_i) * d.ncf, 0.0);
    for (long f = 0; f < nf_i; ++f) {
      const int e1 = c.f2e[4 * f], l1 = c.f2e[4 * f + 1];
      for (int j = 0; j < d.npf; ++j)
        for (int cc = 0; cc < d.ncu; ++cc)
          huh[cc + d.ncu * (j + d.npf * f)] = hu[mesh.perm[j + d.npf * l1] + npe * (cc + static_cast<long>(d.nc) * e1)];
    }
  }
  st.h2d(uh.p + f0 * d.ncf, huh.data(), sizeof(double) * huh.size(), s);
  st.h2d(uh0.p + f0 * d.ncf, huh.data(), sizeof(double) * huh.size(), s);
  if (d.nco > 0) st.h2d(odg.p + e0 * npe * d.nco, so.odg.data(), sizeof(double) * so.odg.size(), s);
  if (ncw > 0) {
    std::vector<double> hw = so.wdg;
    if (hw.empty()) hw.assign(static_cast<size_t>(npe) * ncw * ne_i, 0.0);
    st.h2d(wdg.p + e0 * npe * ncw, hw.data(), sizeof(double) * hw.size(), s);
    st.h2d(wdg0.p + e0 * npe * ncw, hw.data(), sizeof(double) * hw.size(), s);
  }
  // geometry and interpolated auxiliary fields (computed for every element: the other slots'
  // values come out unchanged), then the slot's q-equation matrices
  elem_geom(d, shapegt.p, xdg.p, egeo.p, s);
  face_geom(d, shapfgt.p, perm.p, xdg.p, fgeo.p, s);
  if (d.nco > 0) {
    interp_elem(d, d.nco, shapegt.p, odg.p, og.p, s);
    interp_face(d, d.nco, shapfgt.p, perm.p, odg.p, ogf.p, s);
  }
  setup_q(e0, ne_i);
  // the slot's Newton state starts afresh
  step_act[i] = 1;
  inst_t[i] = time0;
  inst_steps[i] = 0;
  inst_nsteps[i] = static_cast<int>(a.dt.size());
  inst_newton[i] = 0;
  inst_gmres[i] = 0;
  inst_norm[i] = 0.0;
  inst_failed[i] = 0;
  datains[i] = din;
  step_valid = false;  // the next step computes the residual norms again
  assembled = false;
  EXA_HIP(hipStreamSynchronize(s));
}

// Put instance i back to its input state: udg and uh (and the Newton iterate u, if given)
void Solver::restore_instance(int i, double* u) {
  const long ue = static_cast<long>(d.npe) * d.nc;
  EXA_HIP(hipMemcpyAsync(udg.p + eoff[i] * ue, udg0.p + eoff[i] * ue,
                         sizeof(double) * (eoff[i + 1] - eoff[i]) * ue, hipMemcpyDeviceToDevice, s));
  const long f0 = foff[i] * d.ncf, fn = (foff[i + 1] - foff[i]) * d.ncf;
  EXA_HIP(hipMemcpyAsync(uh.p + f0, uh0.p + f0, sizeof(double) * fn, hipMemcpyDeviceToDevice, s));
  if (u) EXA_HIP(hipMemcpyAsync(u + f0, uh0.p + f0, sizeof(double) * fn, hipMemcpyDeviceToDevice, s));
}

long Solver::get_inst_array(int inst, const std::string& name, double* dst, long len) {
  if (inst < 0 || inst >= ninst) throw Error(1, "instance index out of range");
  long off = 0, cnt = 0;
  const double* src = nullptr;
  if (name == "uh") {
    off = foff[inst] * d.ncf;
    cnt = (foff[inst + 1] - foff[inst]) * d.ncf;
    src = uh.p;
  } else if (name == "udg") {
    off = eoff[inst] * static_cast<long>(d.npe) * d.nc;
    cnt = (eoff[inst + 1] - eoff[inst]) * static_cast<long>(d.npe) * d.nc;
    src = udg.p;
  } else {
    throw Error(1, "per-instance arrays: uh, udg");
  }
  if (!dst) return cnt;
  if (len < cnt) throw Error(1, "destination too small for " + name);
  std::vector<double> tmpv(cnt);
  EXA_HIP(hipStreamSynchronize(s));
  st.d2h(tmpv.data(), src + off, cnt * sizeof(double), s);
  std::memcpy(dst, tmpv.data(), cnt * sizeof(double));
  return cnt;
}

}  // namespace exa
