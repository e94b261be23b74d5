// gmres.cpp -- restarted, left-preconditioned GMRES of the condensed trace system
// (backend/Solver/gmres.cpp:361-597, HDG overload; MGS getpoly.cpp:24-42 or CGS gmres.cpp:81-88)
// for one problem or for every instance of a batch at once.
//
// The host only queues work.  The Krylov scalars stay on the device: each Arnoldi step ends in
// one small kernel that applies the instance's Givens rotations to its new Hessenberg column,
// tests convergence, and for a converged instance back-solves y and marks it finished; the next
// launch adds V y to x for the instances that have just converged.  Finished instances drop out
// of every later launch through their activity mask (the element matvec, the Schwarz and block
// apply, the segmented dots and updates; for one problem, every kernel of the step).  A batch
// handle, a batch of one included, always takes the segmented reductions, so an instance's sums
// do not depend on the batch it is in.
//
// The status kernels write the number of instances still iterating into host memory, one slot
// per queued step; the host keeps gm_lag steps queued ahead of the step whose count it reads, so
// the device never waits for the host between steps, and steps queued after the last instance
// finished do no work.  The steps of the Newton loop's solves are replayed from HIP graphs.
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace exa {

GmresDev Solver::gmres_state() {
  GmresDev g{};
  g.ns = ninst;
  g.n1 = restart + 1;
  g.ring = GM_RING;
  g.Hm = gHm.p;
  g.sv = gsv.p;
  g.cs = gcs.p;
  g.sn = gsn.p;
  g.nrmb = gnrmb.p;
  g.act = segmask.p;
  g.jc = gjc.p;
  g.y = segy.p;
  g.tol = gtol.p;
  g.xmask = gxmask.p;
  g.iters = giters.p;
  g.jcount = gjcount.p;
  g.stat = dstat;
  g.bcount = gbcount.p;
  g.ticket = gjcount.p + 1;
  return g;
}

void Solver::clear_graphs() {
  if (!gm_graphs.empty() && s) (void)hipStreamSynchronize(s);
  for (auto& gx : gm_graphs)
    if (gx) (void)hipGraphExecDestroy(gx);
  gm_graphs.clear();
  gm_graph_len = 0;
}

// Arnoldi step (column i of the cycle): v_{i+1} = P A v_i orthogonalised against v_0..v_i,
// rotations and convergence test, x += V y for instances that converged at this step.
void Solver::gmres_step(const GmresDev& g, int i, double* xv) {
  const long N = d.N;
  const int ns = ninst;
  double* vi = V.p + N * i;
  double* vm = V.p + N * (i + 1);
  matvec(vi, tmp.p);
  apply_precond_to(tmp.p, vm);
  hipEvent_t t0{};
  tic(T_ORTHO, &t0);
  if (!batch) {  // one problem: the full-length kernels, skipped once the solve has finished
    const double* act = g.act;
    if (ortho == 0) {  // MGS (getpoly.cpp:24-42)
      for (int k = 0; k <= i; ++k) {
        dot(vm, V.p + N * k, N, partial.p, hcol.p + k, 0, s, act);
        axpy_dev(vm, V.p + N * k, hcol.p + k, -1.0, N, s, act);
      }
    } else {  // CGS (gmres.cpp:81-88)
      multidot(V.p, N, i + 1, vm, N, partial.p, hcol.p, s, act);
      gemv_cols(vm, V.p, N, i + 1, hcol.p, -1.0, N, s, act);
    }
    dot(vm, vm, N, partial.p, hcol.p + i + 1, 1, s, act);
    scale_inv_dev(vm, hcol.p + i + 1, N, s, act);
  } else {  // per instance: segmented kernels over each instance's own unknowns
    SegMap ma = segT.map;
    ma.act = g.act;
    if (ortho == 0) {
      for (int k = 0; k <= i; ++k) {
        seg_multidot(ma, V.p + N * k, N, 1, vm, segpart.p, hcol.p + static_cast<long>(k) * ns, 0, s);
        seg_axpy(ma, vm, V.p + N * k, hcol.p + static_cast<long>(k) * ns, -1.0, s);
      }
    } else {
      seg_multidot(ma, V.p, N, i + 1, vm, segpart.p, hcol.p, 0, s);
      seg_gemv_cols(ma, vm, V.p, N, i + 1, hcol.p, -1.0, s);
    }
    seg_multidot(ma, vm, N, 1, vm, segpart.p, hcol.p + static_cast<long>(i + 1) * ns, 1, s);
    seg_scale(segT.map, vm, vm, hcol.p + static_cast<long>(i + 1) * ns, 1, s);
  }
  toc(T_ORTHO, t0);
  gmres_dev_givens(g, i, hcol.p, s);
  (void)xv;  // instances that converge here update x at the end of the cycle
}

// x += V y over the first k basis vectors for the instances in xmask (zero coefficients skipped)
void Solver::gmres_update_x(const GmresDev& g, int k, double* xv) {
  if (k <= 0) return;
  const long N = d.N;
  if (!batch) {
    gemv_cols(xv, V.p, N, k, g.y, 1.0, N, s, g.xmask, true);
  } else {
    SegMap mx = segT.map;
    mx.act = g.xmask;
    seg_gemv_cols(mx, xv, V.p, N, k, g.y, 1.0, s, true);
  }
}

// End of a cycle of i steps (gmres.cpp:551-566): x += V y, r = P (b - A x), restart test, v_0.
void Solver::gmres_cycle_end(const GmresDev& g, int i, const double* bv, double* xv) {
  const long N = d.N;
  const int ns = ninst;
  gmres_dev_cycle_y(g, i, s);  // y of the instances still iterating; xmask: those and the converged
  gmres_update_x(g, i, xv);
  matvec(xv, tmp.p);
  axpby(tmp.p, bv, 1.0, -1.0, N, s);  // tmp = b - A x
  apply_precond_to(tmp.p, r.p);
  seg_scale(segT.map, r.p, r.p, g.act, 0, s);  // finished instances: r = 0
  seg_norms(segT, r.p, segcoef.p);
  gmres_dev_restart(g, segcoef.p, segcoef.p + ns, s);
  seg_scale(segT.map, V.p, r.p, segcoef.p + ns, 1, s);  // v_0 = r / ||r|| (0 when finished)
}

int Solver::gmres(const double* bv, double* xv, double tol_rel, int maxit) {
  std::vector<int> its;
  return gmres_run(bv, xv, tol_rel, its, maxit);
}

int Solver::gmres_run(const double* bv, double* xv, double tol_rel, std::vector<int>& iters, int maxit_in) {
  const long N = d.N;
  const int gmres_maxit = maxit_in < 0 ? this->gmres_maxit : maxit_in;
  const int ns = ninst;
  if (static_cast<long>(gHm.n) < static_cast<long>(ns) * (restart + 1) * (restart + 1)) seg_alloc();
  const GmresDev g = gmres_state();
  // the tolerance relative to the preconditioned initial residual, min(0.1, tol ||b|| / ||r0||)
  // with x0 = 0 (gmres.cpp:431-445)
  const double tol = std::min(0.1, tol_rel < 0.0 ? gmres_tol : tol_rel);
  // the matvec, the Schwarz apply (einst / elem_active) and, for one problem, the other step
  // kernels (gm_act) skip finished instances while this solve runs
  struct Scope {
    Solver& S;
    Scope(Solver& q, const double* act) : S(q) {
      S.elem_active = act;
      S.gm_act = S.batch ? nullptr : act;
    }
    ~Scope() {
      S.elem_active = nullptr;
      S.gm_act = nullptr;
    }
  };
  // the Newton loop's solves (b -> x) replay captured steps; other vectors run eagerly
  const bool graphs = gm_graph && !timing && !verbose && bv == b.p && xv == x.p;
  if (graphs && gm_graph_len != restart) {
    clear_graphs();
    gm_graphs.assign(restart, nullptr);
    gm_graph_len = restart;
  }
  fill(xv, 0.0, N, s);
  seg_norms(segT, bv, segcoef.p);
  gmres_dev_begin(g, segcoef.p, tol, newton_tol, tol_rel < 0.0 ? 1 : 0, s);
  int nact = 0;
  int icycle = 0;  // steps queued in the current cycle
  {
    Scope scope(*this, g.act);
    apply_precond_to(bv, r.p);
    seg_scale(segT.map, r.p, r.p, g.act, 0, s);
    seg_norms(segT, r.p, segcoef.p);
    gmres_dev_start(g, segcoef.p, segcoef.p + ns, s);
    seg_scale(segT.map, V.p, r.p, segcoef.p + ns, 1, s);  // v_0 = r / ||r||
    EXA_HIP(hipStreamSynchronize(s));
    nact = hstat[0];
    int j = 0, i = 0;
    while (nact > 0 && j < gmres_maxit) {
      const long unit = j;  // one queued unit per step; its count lands in hstat[j % GM_RING]
      const bool last = (i + 1 == restart);  // the cycle ends after this step
      if (graphs && (last || j + 1 < gmres_maxit)) {
        hipGraphExec_t& gx = gm_graphs[i];
        if (!gx) {  // capture this column's step (and, for the last column, the restart work)
          hipGraph_t gr = nullptr;
          EXA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
          gmres_step(g, i, xv);
          if (last) gmres_cycle_end(g, i + 1, bv, xv);
          EXA_HIP(hipStreamEndCapture(s, &gr));
          const hipError_t e = hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0);
          (void)hipGraphDestroy(gr);
          EXA_HIP(e);
        }
        EXA_HIP(hipGraphLaunch(gx, s));
      } else {
        gmres_step(g, i, xv);
        if (last || j + 1 == gmres_maxit) gmres_cycle_end(g, i + 1, bv, xv);
      }
      if (verbose > 1) {  // per-step trace of the first instance: |s_{i+1}| / nrmb
        double h[2];
        EXA_HIP(hipMemcpyAsync(hpin, gsv.p + i + 1, sizeof(double), hipMemcpyDeviceToHost, s));
        EXA_HIP(hipMemcpyAsync(hpin + 1, gnrmb.p, sizeof(double), hipMemcpyDeviceToHost, s));
        EXA_HIP(hipStreamSynchronize(s));
        h[0] = hpin[0];
        h[1] = hpin[1];
        std::printf("GMRES step %d: rel %g (tol %g), %d of %d instances iterating\n", j, std::fabs(h[0]) / h[1], tol,
                    hstat[unit % GM_RING], ns);
      }
      ++i;
      ++j;
      icycle = i;
      if (last || j == gmres_maxit) {
        i = 0;
        icycle = 0;
        if (verbose) {  // the restart line of the reference needs the count: wait for it
          EXA_HIP(hipStreamSynchronize(s));
          std::printf("GMRES restart after %d iterations: %d of %d instances still iterating\n", j,
                      hstat[unit % GM_RING], ns);
          std::fflush(stdout);
        }
      }
      EXA_HIP(hipEventRecord(gm_ev[unit % GM_RING], s));
      if (unit >= gm_lag) {  // read the count of the step gm_lag steps back
        const long u = unit - gm_lag;
        EXA_HIP(hipEventSynchronize(gm_ev[u % GM_RING]));
        nact = hstat[u % GM_RING];
      }
    }
    // x += V y for the instances that converged in the last, unfinished cycle
    gmres_dev_flush(g, s);
    gmres_update_x(g, icycle, xv);
    EXA_HIP(hipStreamSynchronize(s));
  }
  iters.assign(ns, 0);
  if (ns <= 8192) {
    EXA_HIP(hipMemcpyAsync(hpin, giters.p, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
    EXA_HIP(hipStreamSynchronize(s));
    std::copy(reinterpret_cast<const int*>(hpin), reinterpret_cast<const int*>(hpin) + ns, iters.begin());
  } else {
    st.d2h(iters.data(), giters.p, sizeof(int) * ns, s);
  }
  if (nact > 0 && verbose)
    std::printf("Warning: GMRES(%d) leaves %d instances above the tolerance %g after %d iterations\n", restart,
                nact, tol, gmres_maxit);
  return *std::max_element(iters.begin(), iters.end());
}

int Solver::gmres_seg(const double* bv, double* xv, double tol_rel, std::vector<int>& iters, int maxit) {
  return gmres_run(bv, xv, tol_rel, iters, maxit);
}

}  // namespace exa
