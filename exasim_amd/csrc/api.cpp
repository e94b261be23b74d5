// api.cpp -- the extern "C" entry points of libexahip.so (declared in
// include/exasim_hdg.h).  Every call catches the library's exceptions and turns
// them into an EXA_ERR_* code plus a message kept on the handle (or, when
// there is no handle yet, in a thread-local slot).
#include "exasim_hdg.h"
#include "solver.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

struct exa_handle {
  std::unique_ptr<exa::Solver> solver;
  std::string err;
};

namespace {
thread_local std::string g_create_error;

template <typename F>
int guard(exa_handle* h, F&& f) {
  if (!h || !h->solver) return EXA_ERR_ARG;
  try {
    f(*h->solver);
    h->err.clear();
    return EXA_OK;
  } catch (const exa::Error& e) {
    h->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    h->err = "host allocation failed";
    return EXA_ERR_HIP;
  } catch (const std::exception& e) {
    h->err = e.what();
    return EXA_ERR_ARG;
  }
}
}  // namespace

extern "C" {

int exa_create(const char* datain_dir, const char* dataout_prefix, const char* model, int device,
               exa_handle** out) {
  if (!out || !datain_dir || !model) return EXA_ERR_ARG;
  *out = nullptr;
  try {
    auto h = std::make_unique<exa_handle>();
    h->solver = std::make_unique<exa::Solver>(datain_dir, dataout_prefix ? dataout_prefix : "", model,
                                              device);
    *out = h.release();
    g_create_error.clear();
    return EXA_OK;
  } catch (const exa::Error& e) {
    g_create_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return EXA_ERR_ARG;
  }
}

int exa_batch_create(int ninst, const char* const* datain_dirs, const char* model, int device,
                     exa_handle** out) {
  if (!out || !datain_dirs || !model || ninst < 1) return EXA_ERR_ARG;
  *out = nullptr;
  try {
    std::vector<std::string> dirs;
    for (int i = 0; i < ninst; ++i) {
      if (!datain_dirs[i]) {
        g_create_error = "datain_dirs[" + std::to_string(i) + "] is null";
        return EXA_ERR_ARG;
      }
      dirs.emplace_back(datain_dirs[i]);
    }
    auto h = std::make_unique<exa_handle>();
    h->solver = std::make_unique<exa::Solver>(dirs, model, device);
    *out = h.release();
    g_create_error.clear();
    return EXA_OK;
  } catch (const exa::Error& e) {
    g_create_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return EXA_ERR_ARG;
  }
}

int exa_batch_size(const exa_handle* h) {
  if (!h || !h->solver) return 0;
  return h->solver->ninst;
}

int exa_batch_stats(exa_handle* h, int* newton_iters, int* gmres_iters, double* final_norm, int n) {
  return guard(h, [&](exa::Solver& s) {
    if (n < s.ninst) throw exa::Error(EXA_ERR_ARG, "exa_batch_stats: arrays shorter than the batch");
    for (int i = 0; i < s.ninst; ++i) {
      const bool have = static_cast<int>(s.inst_newton.size()) == s.ninst;
      if (newton_iters) newton_iters[i] = have ? s.inst_newton[i] : 0;
      if (gmres_iters) gmres_iters[i] = have ? s.inst_gmres[i] : 0;
      if (final_norm) final_norm[i] = have ? s.inst_norm[i] : 0.0;
    }
  });
}

long exa_batch_get(exa_handle* h, int inst, const char* name, double* dst, long len) {
  long out = -1;
  int rc = guard(h, [&](exa::Solver& s) {
    if (!name) throw exa::Error(EXA_ERR_ARG, "null name");
    out = s.get_inst_array(inst, name, dst, len);
  });
  return rc == EXA_OK ? out : -rc;
}

int exa_batch_step(exa_handle* h, int* status, int n, int* iterating) {
  return guard(h, [&](exa::Solver& s) {
    if (status && n < s.ninst) throw exa::Error(EXA_ERR_ARG, "exa_batch_step: status shorter than the batch");
    const int left = s.newton_seg_step(status);
    if (iterating) *iterating = left;
  });
}

int exa_batch_refill(exa_handle* h, int inst, const char* datain_dir) {
  return guard(h, [&](exa::Solver& s) {
    if (!datain_dir) throw exa::Error(EXA_ERR_ARG, "null datain directory");
    s.refill(inst, datain_dir);
  });
}

void exa_destroy(exa_handle* h) { delete h; }

const char* exa_last_error(const exa_handle* h) {
  if (!h) return g_create_error.c_str();
  return h->err.c_str();
}

int exa_sizes(const exa_handle* h, int* out, int len) {
  if (!h || !h->solver || !out) return 0;
  const exa::Solver& s = *h->solver;
  const exa::Dims& d = s.d;
  const bool ilu = s.ilu_ready;
  int v[22] = {d.nd, d.ncu, d.nc, d.npe, d.npf, d.nfe, d.ne, d.nf, d.n, d.m, static_cast<int>(d.N),
               d.nge, d.ngf, d.nco, s.restart, s.precond, s.precision, s.last_refinements, s.neb,
               ilu ? s.ilu.nse : 0, ilu ? s.ilu.nfse : 0, ilu ? s.ilu.nnz : 0};
  int k = len < 22 ? len : 22;
  for (int i = 0; i < k; ++i) out[i] = v[i];
  return k;
}

int exa_set_option(exa_handle* h, const char* name, double value) {
  if (!name) return EXA_ERR_ARG;
  return guard(h, [&](exa::Solver& s) { s.set_option(name, value); });
}

int exa_solve(exa_handle* h, int* newton_iters, int* gmres_iters, double* final_norm) {
  return guard(h, [&](exa::Solver& s) { s.solve(newton_iters, gmres_iters, final_norm); });
}

int exa_reset(exa_handle* h) { return guard(h, [&](exa::Solver& s) { s.reset(); }); }

int exa_get_q(exa_handle* h) { return guard(h, [&](exa::Solver& s) { s.get_q(); }); }

int exa_assemble(exa_handle* h) { return guard(h, [&](exa::Solver& s) { s.assemble(); }); }

int exa_residual(exa_handle* h, double* norm_b, double* norm_ru) {
  return guard(h, [&](exa::Solver& s) { s.residual(norm_b, norm_ru); });
}

int exa_matvec(exa_handle* h, const double* v_dev, double* w_dev) {
  if (!v_dev || !w_dev) return EXA_ERR_ARG;
  return guard(h, [&](exa::Solver& s) {
    if (!s.assembled) throw exa::Error(EXA_ERR_ARG, "exa_matvec before exa_assemble");
    s.matvec(v_dev, w_dev);
  });
}

int exa_matvec_noalloc(exa_handle* h, const double* v_dev, double* w_dev) {
  if (!v_dev || !w_dev) return EXA_ERR_ARG;
  return guard(h, [&](exa::Solver& s) {
    if (!s.assembled) throw exa::Error(EXA_ERR_ARG, "exa_matvec_noalloc before exa_assemble");
    exa::matvec_noalloc(s, v_dev, w_dev, s.noalloc_chunk);
  });
}

int exa_precond(exa_handle* h, double* x_dev) {
  if (!x_dev) return EXA_ERR_ARG;
  return guard(h, [&](exa::Solver& s) {
    if (!s.assembled) throw exa::Error(EXA_ERR_ARG, "exa_precond before exa_assemble");
    s.apply_precond(x_dev);
  });
}

int exa_gmres(exa_handle* h, const double* b_dev, double* x_dev, int* iters) {
  if (!b_dev || !x_dev) return EXA_ERR_ARG;
  return guard(h, [&](exa::Solver& s) {
    if (!s.assembled) throw exa::Error(EXA_ERR_ARG, "exa_gmres before exa_assemble");
    int it = s.gmres(b_dev, x_dev);
    if (iters) *iters = it;
  });
}

long exa_get_array(exa_handle* h, const char* name, double* dst, long len) {
  if (!name) return -EXA_ERR_ARG;
  long out = 0;
  int rc = guard(h, [&](exa::Solver& s) { out = s.get_array(name, dst, len); });
  return rc == EXA_OK ? out : -rc;
}

int exa_set_array(exa_handle* h, const char* name, const double* src, long len) {
  if (!name || !src) return EXA_ERR_ARG;
  return guard(h, [&](exa::Solver& s) { s.set_array(name, src, len); });
}

int exa_device_ptr(exa_handle* h, const char* name, void** ptr, long* len) {
  if (!name || !ptr) return EXA_ERR_ARG;
  return guard(h, [&](exa::Solver& s) { s.device_ptr(name, ptr, len); });
}

int exa_write_output(exa_handle* h) { return guard(h, [&](exa::Solver& s) { s.write_output(); }); }

int exa_timings(exa_handle* h, double* ms, long* counts, int len) {
  int n = 0;
  int rc = guard(h, [&](exa::Solver& s) { n = s.timings(ms, counts, len); });
  return rc == EXA_OK ? n : -rc;
}

int exa_reset_timings(exa_handle* h) { return guard(h, [&](exa::Solver& s) { s.reset_timings(); }); }

int exa_gemm_tuning(char* buf, int len) {
  if (!buf || len < 1) return EXA_ERR_ARG;
  const std::string r = exa::gemm_tuning_report();
  const size_t n = std::min(r.size(), static_cast<size_t>(len - 1));
  std::memcpy(buf, r.data(), n);
  buf[n] = 0;
  return static_cast<int>(r.size());
}

int exa_sync(exa_handle* h) {
  return guard(h, [&](exa::Solver& s) { EXA_HIP(hipStreamSynchronize(s.s)); });
}

// The reference library's entry point (backend/Main/libexasim.cpp: libmain -> premain,
// backend/Main/exasim.hpp:143-360 without text2code): argv = {program, model, nummodels, filein,
// fileout [, restart]}; filein is the datain prefix (a directory, "datain/"), fileout the output
// prefix ("dataout/" -> dataout/udg_np0.bin, dataout/uhat_np0.bin).  One model (nummodels = 1) and
// no restart file; the model is a built-in or generated model name, where the reference's
// built-in "Poisson3D" (backend/Main/BuiltIn/Poisson3D: flux mu q, source 3 pi^2 sin sin sin,
// Dirichlet zero) is the built-in "poisson".  Returns 0 or an EXA_ERR_* code.
int libmain(int argc, char** argv) {
  if (argc < 5 || !argv) {
    std::fprintf(stderr, "Usage: %s model nummodels InputFile OutputFile [restart]\n",
                 (argc > 0 && argv && argv[0]) ? argv[0] : "libmain");
    return EXA_ERR_ARG;
  }
  std::string model = argv[1] ? argv[1] : "";
  if (model == "Poisson3D" || model == "Poisson2D") model = "poisson";
  int nummodels = 0, restart = 0;
  try {
    nummodels = std::stoi(argv[2]);
  } catch (...) {
    std::fprintf(stderr, "Error: invalid nummodels argument\n");
    return EXA_ERR_ARG;
  }
  if (nummodels != 1) {
    std::fprintf(stderr, "Error: %d models: coupled multi-model problems are not supported\n", nummodels);
    return EXA_ERR_UNSUPPORTED;
  }
  // the optional restart step follows the in/out pair (premain's argv[2*nummodels + 2])
  try {
    if (argc > 5) restart = std::stoi(argv[5]);
  } catch (...) {
    std::fprintf(stderr, "Error: invalid restart argument\n");
    return EXA_ERR_ARG;
  }
  if (restart != 0) {
    std::fprintf(stderr, "Error: restarting from a saved time step is not supported\n");
    return EXA_ERR_UNSUPPORTED;
  }
  std::string filein = argv[3] ? argv[3] : "";
  while (filein.size() > 1 && filein.back() == '/') filein.pop_back();
  const std::string fileout = argv[4] ? argv[4] : "";
  std::printf("Running model %s: datain %s, output prefix %s\n", model.c_str(), filein.c_str(), fileout.c_str());
  exa_handle* h = nullptr;
  int rc = exa_create(filein.c_str(), fileout.c_str(), model.c_str(), -1, &h);
  if (rc != EXA_OK) {
    std::fprintf(stderr, "Error: %s\n", exa_last_error(h));
    exa_destroy(h);
    return rc;
  }
  exa_set_option(h, "verbose", 1);
  int newton = 0, gmres = 0;
  double norm = 0.0;
  rc = exa_solve(h, &newton, &gmres, &norm);
  if (rc == EXA_OK) rc = exa_write_output(h);
  if (rc != EXA_OK) std::fprintf(stderr, "Error: %s\n", exa_last_error(h));
  else std::printf("Newton iterations %d, GMRES iterations %d, residual norm %g\n", newton, gmres, norm);
  std::fflush(stdout);
  exa_destroy(h);
  return rc;
}

}  // extern "C"
