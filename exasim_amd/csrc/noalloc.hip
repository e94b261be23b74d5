// noalloc.hip -- the condensed-system matvec (hdgMatVec, backend/Discretization/matvec.cpp:240-385)
// with its element-local products kept in a fixed device array instead of library-allocated
// memory, for callers that cannot allocate (code under HIP graph capture, serving loops that
// forbid hipMalloc).
//
// scratch is a file-scope device array of SCRATCH_DOUBLES doubles (32 MiB), one copy per GPU the
// library runs on, held until the process ends.  Its device address is looked up once per device
// (noalloc_init, called from handle set-up with the device current) and kept in scratch_dev[];
// host code passes that address to the kernels, never the array's name.  Elements are processed in
// chunks that fit the array: each chunk's products H_e v_e go to scratch and are added into w
// (zeroed first) at the global trace indices.  Every trace entry receives one or two terms, so the
// sum is the same, bit for bit, as the face gather of exa_matvec (a + b == b + a, 0 + a == a).
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two streams, two
// threads) race and must be serialised by the caller.
#include "solver.hpp"

#include <mutex>

namespace exa {

namespace {

constexpr long SCRATCH_DOUBLES = 4L << 20;  // 32 MiB
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

std::mutex scratch_mu;
double* scratch_dev[MAX_DEVICES] = {};

// w[c + ncu*elemcon[jl, e0+e]] += we[e*m + c + ncu*jl]  (the trace entries of elements e0..e0+eb-1)
__global__ void k_scatter_add(Dims d, long e0, int eb, const int* elemcon, const double* we, double* w) {
  const long q = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (q >= static_cast<long>(eb) * d.m) return;
  const long e = q / d.m;
  const int t = static_cast<int>(q - e * d.m);
  const int c = t % d.ncu, jl = t / d.ncu;
  const long g = elemcon[jl + static_cast<long>(d.ndf) * (e0 + e)];
  atomicAdd(w + c + static_cast<long>(d.ncu) * g, we[q]);
}

}  // namespace

void noalloc_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw Error(5, "device index beyond the no-allocation table");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[device]) return;
  int cur = -1;
  EXA_HIP(hipGetDevice(&cur));
  if (cur != device) throw Error(1, "noalloc_init: the device is not current");
  void* p = nullptr;
  EXA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw Error(3, "no device address for the scratch array");
  scratch_dev[device] = static_cast<double*>(p);
}

long noalloc_scratch_doubles() { return SCRATCH_DOUBLES; }

void matvec_noalloc(Solver& S, const double* v, double* w, int max_chunk) {
  const Dims& d = S.d;
  double* ws = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (S.device >= 0 && S.device < MAX_DEVICES) ws = scratch_dev[S.device];
  }
  if (!ws) throw Error(3, "the scratch array of this device was not looked up at set-up");
  if (d.m > SCRATCH_DOUBLES)
    throw Error(5, "element trace block of " + std::to_string(d.m) + " entries exceeds the scratch array (" +
                       std::to_string(SCRATCH_DOUBLES) + " doubles)");
  long chunk = SCRATCH_DOUBLES / d.m;
  if (max_chunk > 0) chunk = std::min<long>(chunk, max_chunk);
  fill(w, 0.0, d.N, S.s);
  for (long e0 = 0; e0 < d.ne; e0 += chunk) {
    const int eb = static_cast<int>(std::min<long>(chunk, d.ne - e0));
    Dims dc = d;
    dc.ne = eb;
    const int* ec = S.elemcon.p + static_cast<long>(d.ndf) * e0;
    const long off = e0 * static_cast<long>(d.m) * d.m;
    if (S.precision) matvec_elem(dc, S.H32.p + off, ec, v, ws, S.s);
    else matvec_elem(dc, S.H.p + off, ec, v, ws, S.s);
    k_scatter_add<<<static_cast<unsigned>((static_cast<long>(eb) * d.m + 255) / 256), 256, 0, S.s>>>(
        d, e0, eb, S.elemcon.p, ws, w);
    EXA_HIP(hipGetLastError());
  }
}

}  // namespace exa
