// models.hpp -- pointwise model functions evaluated inside the assembly kernels.
//
// These stand in for the text2code-generated plugin (backend/Model/libpdemodel.hpp:1-41):
// HdgFlux / HdgSource / HdgFbou for the pdemodel files the workloads use.  The
// layouts are those of the generated code (text2code CodeGenerator.cpp:903-910):
//   f[c + NCU*d], f_uq[(c + NCU*d) + NCU*ND*n], s[c], s_uq[c + NCU*n],
//   fb[c], fb_uq[c + NCU*n], fb_uh[c + NCU*n]
// where uq = (u[NCU], q[NCU*ND]) with q component (c, d) at NCU + c + NCU*d.
// Each model is a struct of static __device__ functions templated on the
// spatial dimension and component count, so the assembly kernels keep every
// pointwise array in registers.
#pragma once

#include <hip/hip_runtime.h>

namespace exa {

enum ModelId { MODEL_POISSON = 0, MODEL_POISSON_POLY = 1, MODEL_CONVDIFF = 2, MODEL_PLUGIN = 50 };

// examples/Poisson/poisson{2d,3d}/pdemodel.txt: Flux = kappa*q, Source =
// nd*pi^2*prod sin(pi x_d), FbouHdg = tau*(0 - uhat).
template <int ND_, int NCU_>
struct Poisson {
  static constexpr int ND = ND_, NCU = NCU_, NC = NCU_ * (ND_ + 1);
  __device__ static double boundary_value(const double* x) { return 0.0; }
  template <bool JAC>
  __device__ static void flux(const double* x, const double* uq, const double* v, const double* p,
                              double t, double* f, double* f_uq) {
    const double kappa = p[0];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int c = 0; c < NCU; ++c) {
        f[c + NCU * d] = kappa * uq[NCU + c + NCU * d];
        if (JAC) {
#pragma unroll
          for (int n = 0; n < NC; ++n) f_uq[(c + NCU * d) + NCU * ND * n] = 0.0;
          f_uq[(c + NCU * d) + NCU * ND * (NCU + c + NCU * d)] = kappa;
        }
      }
  }
  template <bool JAC>
  __device__ static void source(const double* x, const double* uq, const double* v, const double* p,
                                double t, double* s, double* s_uq) {
    const double pi = 3.141592653589793;
    double prod = 1.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) prod *= sin(pi * x[d]);
    s[0] = ND * pi * pi * prod;
#pragma unroll
    for (int c = 1; c < NCU; ++c) s[c] = 0.0;
    if (JAC)
#pragma unroll
      for (int i = 0; i < NCU * NC; ++i) s_uq[i] = 0.0;
  }
  template <bool JAC>
  __device__ static void fbou(int ib, const double* x, const double* uq, const double* v,
                              const double* uh, const double* nl, const double* tau,
                              const double* p, double t, double* fb, double* fb_uq, double* fb_uh) {
    const double g = boundary_value(x);
#pragma unroll
    for (int c = 0; c < NCU; ++c) fb[c] = tau[0] * (g - uh[c]);
    if (JAC) {
#pragma unroll
      for (int i = 0; i < NCU * NC; ++i) fb_uq[i] = 0.0;
#pragma unroll
      for (int i = 0; i < NCU * NCU; ++i) fb_uh[i] = 0.0;
#pragma unroll
      for (int c = 0; c < NCU; ++c) fb_uh[c + NCU * c] = -tau[0];
    }
  }
};

// Same flux with exact solution u = |x|^2: Source = -2*nd*kappa, FbouHdg = tau*(u(x) - uhat).
template <int ND_, int NCU_>
struct PoissonPoly : Poisson<ND_, NCU_> {
  static constexpr int ND = ND_, NCU = NCU_, NC = NCU_ * (ND_ + 1);
  __device__ static double boundary_value(const double* x) {
    double r = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) r += x[d] * x[d];
    return r;
  }
  template <bool JAC>
  __device__ static void source(const double* x, const double* uq, const double* v, const double* p,
                                double t, double* s, double* s_uq) {
    s[0] = -2.0 * ND * p[0];
#pragma unroll
    for (int c = 1; c < NCU; ++c) s[c] = 0.0;
    if (JAC)
#pragma unroll
      for (int i = 0; i < NCU * NC; ++i) s_uq[i] = 0.0;
  }
  template <bool JAC>
  __device__ static void fbou(int ib, const double* x, const double* uq, const double* v,
                              const double* uh, const double* nl, const double* tau,
                              const double* p, double t, double* fb, double* fb_uq, double* fb_uh) {
    const double g = boundary_value(x);
#pragma unroll
    for (int c = 0; c < NCU; ++c) fb[c] = tau[0] * (g - uh[c]);
    if (JAC) {
#pragma unroll
      for (int i = 0; i < NCU * NC; ++i) fb_uq[i] = 0.0;
#pragma unroll
      for (int i = 0; i < NCU * NCU; ++i) fb_uh[i] = 0.0;
#pragma unroll
      for (int c = 0; c < NCU; ++c) fb_uh[c + NCU * c] = -tau[0];
    }
  }
};

// Steady convection-diffusion, per component: Flux = kappa*q + vel*u, Source = v
// (the auxiliary field, vdgfile), FbouHdg: condition 1 Dirichlet tau*(0 - uhat),
// any other condition homogeneous Neumann (zero diffusive flux)
// kappa*q.n + tau*(u - uhat).  physicsparam = [kappa, vel_1 .. vel_ND].
template <int ND_, int NCU_>
struct ConvDiff {
  static constexpr int ND = ND_, NCU = NCU_, NC = NCU_ * (ND_ + 1);
  template <bool JAC>
  __device__ static void flux(const double* x, const double* uq, const double* v, const double* p,
                              double t, double* f, double* f_uq) {
    const double kappa = p[0];
    if (JAC)
#pragma unroll
      for (int i = 0; i < NCU * ND * NC; ++i) f_uq[i] = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int c = 0; c < NCU; ++c) {
        f[c + NCU * d] = kappa * uq[NCU + c + NCU * d] + p[1 + d] * uq[c];
        if (JAC) {
          f_uq[(c + NCU * d) + NCU * ND * (NCU + c + NCU * d)] = kappa;
          f_uq[(c + NCU * d) + NCU * ND * c] = p[1 + d];
        }
      }
  }
  template <bool JAC>
  __device__ static void source(const double* x, const double* uq, const double* v, const double* p,
                                double t, double* s, double* s_uq) {
#pragma unroll
    for (int c = 0; c < NCU; ++c) s[c] = v[c];
    if (JAC)
#pragma unroll
      for (int i = 0; i < NCU * NC; ++i) s_uq[i] = 0.0;
  }
  template <bool JAC>
  __device__ static void fbou(int ib, const double* x, const double* uq, const double* v,
                              const double* uh, const double* nl, const double* tau,
                              const double* p, double t, double* fb, double* fb_uq, double* fb_uh) {
    if (JAC) {
#pragma unroll
      for (int i = 0; i < NCU * NC; ++i) fb_uq[i] = 0.0;
#pragma unroll
      for (int i = 0; i < NCU * NCU; ++i) fb_uh[i] = 0.0;
    }
    if (ib == 1) {
#pragma unroll
      for (int c = 0; c < NCU; ++c) {
        fb[c] = tau[0] * (0.0 - uh[c]);
        if (JAC) fb_uh[c + NCU * c] = -tau[0];
      }
      return;
    }
    const double kappa = p[0];
#pragma unroll
    for (int c = 0; c < NCU; ++c) {
      double acc = tau[0] * (uq[c] - uh[c]);
      const double duh = -tau[0];
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        acc += kappa * uq[NCU + c + NCU * d] * nl[d];
        if (JAC) fb_uq[c + NCU * (NCU + c + NCU * d)] = kappa * nl[d];
      }
      fb[c] = acc;
      if (JAC) {
        fb_uq[c + NCU * c] = tau[0];
        fb_uh[c + NCU * c] = duh;
      }
    }
  }
};

}  // namespace exa
