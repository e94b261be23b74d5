// CudaHip.hpp -- compile support for model functions written in text2code's GPU form, so that
// they build with hipcc into a model library that libexahip.so loads (model name
// "plugin:<path to the .so>").
//
// text2code (text2code/text2code/CodeGenerator.cpp:995-1025, framework "hip") writes each model
// function as a host function whose body is
//     parallel_for("<name>", N, GPU_LAMBDA(const size_t i) { ... });
// over N points, every array component-major ([component * N + i]), and a header that includes
// "CudaHip.hpp".  The functions keep the signatures of backend/Model/libpdemodel.hpp:1-41:
//     HdgFlux(f, f_udg, f_wdg, xdg, udg, odg, wdg, uinf, param, time, modelnumber,
//             ng, nc, ncu, nd, ncx, nco, ncw)
//     HdgSource(... the same ...)
//     HdgFbou(f, f_udg, f_wdg, f_uhg, xdg, udg, odg, wdg, uhg, nlg, tau, uinf, param, time,
//             modelnumber, ib, ng, nc, ncu, nd, ncx, nco, ncw)
// with every pointer a device pointer.  Build a model library with
//     hipcc --offload-arch=gfx950 -O3 -fPIC -shared -I<this directory> HdgFlux.cpp HdgSource.cpp
//           HdgFbou.cpp -o libpdemodel.so
// (the functions need C linkage or the extern "C" wrappers libexahip.so looks up: HdgFlux,
// HdgSource, HdgFbou).
//
// parallel_for launches one thread per point (256-thread workgroups) on the HIP stream that
// libexahip.so hands over through exa_plugin_set_stream before each call, so the model's launches
// are ordered with the solver's own kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

typedef double dstype;

#define GPU_LAMBDA [=] __device__

namespace exa_plugin {
inline hipStream_t& current_stream() {
  static hipStream_t s = nullptr;
  return s;
}

template <class F>
__global__ __launch_bounds__(256) void k_points(size_t n, F f) {
  const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (i < n) f(i);
}
}  // namespace exa_plugin

template <class F>
inline void parallel_for(const char*, size_t n, F f) {
  if (n == 0) return;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(exa_plugin::k_points<F>, dim3(blocks), dim3(256), 0, exa_plugin::current_stream(), n, f);
}

// looked up by libexahip.so (optional: without it the model's kernels run on the null stream)
extern "C" __attribute__((weak, visibility("default"))) void exa_plugin_set_stream(hipStream_t s) {
  exa_plugin::current_stream() = s;
}
