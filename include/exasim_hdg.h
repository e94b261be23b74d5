/*
 * exasim_hdg.h -- C ABI of libexahip.so, the MI355X-native HDG Newton-GMRES path.
 *
 * The library reads the reference's datain/ layout (app.bin, master.bin,
 * mesh.bin, sol.bin), runs the HDG Newton-GMRES solve on one GPU and writes
 * the reference's dataout/ layout (out{udg,uhat}_np0.bin).  All arguments are
 * plain pointers and sizes; host buffers are host memory, device buffers are
 * device memory of the handle's GPU.  Every entry point returns 0 on success
 * and a non-zero EXA_ERR_* code on failure; exa_last_error() then describes
 * the failure.  Calls on one handle must not overlap; different handles are
 * independent.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   exa_create     backend/Main/exasim.hpp:281-370 (t2c parse + CSolution ctor):
 *                  readInput (backend/Discretization/readbinaryfiles.cpp:625-699),
 *                  buildConn (backend/Discretization/connectivity.cpp:1212-1385),
 *                  compGeometry + qEquation (backend/Discretization/discretization.cpp:170-412)
 *   exa_solve      backend/Solution/solution.cpp:702-720 (SolveProblem -> SteadyProblem
 *                  -> NewtonSolver, solution.cpp:125-242)
 *   exa_write_output backend/Solution/solution.cpp:726-760 (SaveSolutions)
 *   exa_assemble   backend/Discretization/discretization.cpp:483-512 (hdgAssembleLinearSystem)
 *   exa_residual   backend/Discretization/discretization.cpp:576-591 (hdgAssembleResidual)
 *   exa_matvec     backend/Discretization/matvec.cpp:240-385 (hdgMatVec)
 *   exa_precond    backend/Preconditioning/preconditioner.cpp:273-325 (ApplyPreconditioner)
 *   exa_gmres      backend/Solver/gmres.cpp:361-597 (GMRES, HDG overload)
 *   exa_get_array  the debugmode=1 stage dumps (uequation.cpp:182-202, 542-556, 683-774;
 *                  qequation.cpp:152-165; ptcsolver.cpp:363-371)
 *   model names    the generated model plugin (backend/Model/libpdemodel.hpp:1-41) is
 *                  selected by name: built-in structs (exasim_amd/csrc/models.hpp) or
 *                  pdemodel.txt files compiled by exasim_amd/modelgen.py.
 */
#ifndef EXASIM_HDG_H
#define EXASIM_HDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct exa_handle exa_handle;

enum {
  EXA_OK = 0,
  EXA_ERR_ARG = 1,        /* bad argument (null pointer, wrong size, unknown name) */
  EXA_ERR_IO = 2,         /* datain/dataout file missing or malformed */
  EXA_ERR_HIP = 3,        /* a HIP runtime call failed */
  EXA_ERR_NUMERIC = 4,    /* singular element/face block, NaN residual, divergence */
  EXA_ERR_UNSUPPORTED = 5 /* option the library does not implement */
};

/* Build a solver from datain_dir/{app,master,mesh,sol}.bin for the named model
 * ("poisson", "poisson_poly", "convdiff", or a generated model such as "ns2d").  dataout_prefix is the prefix of the
 * output files (e.g. "dataout/out"); it may be NULL.  device < 0 selects the
 * current HIP device. */
int exa_create(const char* datain_dir, const char* dataout_prefix, const char* model,
               int device, exa_handle** out);
void exa_destroy(exa_handle* h);
const char* exa_last_error(const exa_handle* h);

/* Sizes and state: writes up to 22 ints: nd, ncu, nc, npe, npf, nfe, ne, nf, n, m,
 * N (= ndof uhat), nge, ngf, nco, restart, preconditioner, precision, refinements
 * (corrections made in the last Newton step of a mixed-precision solve), neb (elements per
 * assembly block, set by the workspace_bytes option), and the block-ILU(0) structure once it
 * is built: superelements, face slots per superelement, blocks per superelement (0 before).
 * Returns the count written. */
int exa_sizes(const exa_handle* h, int* out, int len);

/* Options (name, value): "gmres_tol", "gmres_restart" (1..2000), "gmres_maxit",
 * "newton_tol", "newton_maxit", "preconditioner" (0 block-Jacobi, 1 elemental
 * additive Schwarz, 2 superelement block-ILU(0), which needs cartgridpart in mesh.bin),
 * "ortho" (0 MGS, 1 CGS), "verbose" (1: Newton and GMRES restart lines on stdout),
 * "timing" (1: record per-kernel HIP events), "workspace_bytes" (element-block
 * workspace budget of the assembly, default 16e9), "precision" (0 FP64; 1 mixed: the element blocks
 * are assembled and condensed in FP32 (FP32 MFMA products, FP32 LU of D' with its pivots
 * kept), the condensed operator H, the local-solve factors and the preconditioner blocks
 * are stored in FP32, residuals and GMRES vectors stay in FP64, and each Newton step is
 * refined
 * with FP64 residuals until ||R|| <= "refine_rtol" (1e-10) times its first residual,
 * at most "refine_maxit" (5) corrections, each solved by GMRES to "refine_tol" (1e-4);
 * the first solve of each Newton step goes to max(gmres_tol, "refine_first_tol" (1e-7)));
 * "refine_cycles" (default 0 = no cap): in the mixed form, each of those GMRES solves stops after
 * at most this many restart cycles, so the next correction restarts from the FP64 residual;
 * "precond_bf16" (0/1): in the mixed form, keep the block-ILU(0) factors or the inverted Schwarz
 * blocks in BF16;
 * "gmres_lag" (0..30, default 3): GMRES steps queued ahead of the host's convergence check (the
 * Krylov scalars stay on the device; steps queued after convergence do no work);
 * "gmres_graph" (0/1, default 0): replay the Newton loop's GMRES steps from captured HIP graphs;
 * "noalloc_chunk" (elements, default 0 = as many as fit): element chunk of exa_matvec_noalloc;
 * "lu_inverses" (1 default, 0): the mixed form's FP32 LU of D' and of the Schwarz blocks applies
 * explicit triangle inverses by MFMA products (1) or solves by substitution (0);
 * "ilu_wave" (process-wide; 1 default, 0, or -1 for the EXA_BILU_WAVE environment value): the
 * block-ILU(0) apply of FP32 / BF16 factors with one wave per superelement (1) or with a
 * 256-thread workgroup per superelement (0). */
int exa_set_option(exa_handle* h, const char* name, double value);

/* Full steady solve (NewtonSolver).  newton_iters/gmres_iters/final_norm may be NULL. */
int exa_solve(exa_handle* h, int* newton_iters, int* gmres_iters, double* final_norm);

/* Restore udg and uh to the values read from sol.bin (device-to-device copies),
 * so that the next exa_solve starts from the input state. */
int exa_reset(exa_handle* h);

/* Stage entry points. */
int exa_get_q(exa_handle* h);                 /* hdgGetQ on the current (udg, uh) */
int exa_assemble(exa_handle* h);              /* H, DinvF, Ru, Rh, b and the preconditioner */
int exa_residual(exa_handle* h, double* norm_b, double* norm_ru);
/* w = A v (device pointers, length N). */
int exa_matvec(exa_handle* h, const double* v_dev, double* w_dev);
/* exa_matvec without library-allocated scratch: no counterpart in the reference; for callers that
 * cannot allocate (code under HIP graph capture, serving loops that forbid hipMalloc).  The element
 * products go to a fixed 32 MiB device array per GPU (elements are processed in chunks that fit it;
 * an element trace block larger than the array is rejected with EXA_ERR_UNSUPPORTED); w equals
 * exa_matvec's bit for bit.  The call allocates nothing.  Every call on one GPU shares that array:
 * calls that overlap on one GPU race and must be serialised by the caller. */
int exa_matvec_noalloc(exa_handle* h, const double* v_dev, double* w_dev);
/* x = P x in place (device pointer, length N). */
int exa_precond(exa_handle* h, double* x_dev);
/* Solve A x = b with the handle's assembled system (device pointers); x starts at 0. */
int exa_gmres(exa_handle* h, const double* b_dev, double* x_dev, int* iters);

/* Copy a named array to host memory (len = number of doubles available at dst).
 * Names: "udg" [npe,nc,ne], "uh" [ncu,npf,nf], "H" [m,m,ne], "DinvF" [n,m,ne],
 * "Ru" [n,ne], "Rh" [m,ne], "b" [N], "BE" (block-Jacobi inverse) [ncf,ncf,nf],
 * "BEilu" (factored block-ILU(0) blocks) [ncf,ncf,nse,nnz], "ilu_face" [nse,nfse],
 * "BEeas" (inverted elemental additive Schwarz blocks) [m,m,ne],
 * "ilu_row_ptr" [nfse+1], "ilu_col_ind" [nnz] (as doubles), "Mass" [npe,npe,ne],
 * "MinvC" [npe,nd,npe,ne], "MinvE" [npe,nd,ndf,ne], "elemcon" [npf*nfe*ne] (as doubles),
 * "f2e" [4*nf] (as doubles), "res_b" [N], "res_Ru" [n,ne], "x", "w" [N], "egeo",
 * "fgeo" (geometry at quadrature points).  After a mixed-precision assembly "H", "DinvF"
 * and "BEilu" are the FP32 arrays widened to FP64.  Returns the number of doubles of the array
 * (copies min(len, size)). */
long exa_get_array(exa_handle* h, const char* name, double* dst, long len);
/* Overwrite "udg", "uh", "x" or "w" from host memory (len must equal the array size). */
int exa_set_array(exa_handle* h, const char* name, const double* src, long len);
/* Device pointer of a named persistent device array (H, DinvF, uh, udg, b, x, w); the FP64
 * H and DinvF are those of the last FP64 assembly. */
int exa_device_ptr(exa_handle* h, const char* name, void** ptr, long* len);

/* dataout/out{udg,uhat}_np0.bin with the reference's 3-double headers. */
int exa_write_output(exa_handle* h);

/* Per-kernel-class timings recorded with HIP events on the handle's stream
 * while the "timing" option is on: for class index k, total milliseconds and
 * launch count.  Classes: 0 matvec, 1 precond apply, 2 orthogonalisation,
 * 3 assembly, 4 condensation (LU), 5 other.  Returns the number of classes. */
int exa_timings(exa_handle* h, double* ms, long* counts, int len);
int exa_reset_timings(exa_handle* h);

/* Synchronise the handle's stream. */
int exa_sync(exa_handle* h);

/* The batched GEMMs' tile autotuner (no counterpart in the reference, whose batched products are
 * library calls): the workgroup tile of every product shape timed so far in this process, one line
 * per shape, "M N K log2(batch) fp32 storage BMxBN" followed by the microseconds of each of the five
 * tiles (64x64 32x64 64x32 128x64 64x128; -1 for a tile the product cannot take: a product
 * computed in place needs a tile that holds all its rows).  Writes at most len - 1 characters and a
 * terminating zero; returns the length of the whole report.  exa_set_option(h, "gemm_autotune", 0)
 * turns the timing off for the process (1 on, the default; the environment variable
 * EXA_GEMM_AUTOTUNE sets the start value); "gemm_tile" 0..4 puts every product on one tile (-1: the
 * tuner's choice).  Every tile gives bitwise the same products. */
int exa_gemm_tuning(char* buf, int len);

/* ---- batches of independent problems (no counterpart in the reference, which runs one
 * problem per process; each instance follows exa_solve's algorithm: NewtonSolver,
 * backend/Solution/solution.cpp:125-242, and GMRES, backend/Solver/gmres.cpp:361-597).
 *
 * exa_batch_create reads ninst datain directories with one element shape (same nd, p,
 * quadrature, ncu and face numbering), one time scheme and one preconditioner (0 or 1),
 * and any mesh sizes, physicsparam and tau.  The instances are merged into one disconnected
 * mesh: assembly, condensation, the matvec and the preconditioner run as single launches over
 * all of them, while every instance keeps its own Krylov recurrence (segmented dot products
 * over its own unknowns), GMRES and Newton tolerances, damping and convergence.  An instance's
 * result is the same alone (a batch of one) and in a batch.  The handle takes every other
 * entry point; exa_solve solves all instances, exa_get_array returns the merged arrays. */
int exa_batch_create(int ninst, const char* const* datain_dirs, const char* model, int device,
                     exa_handle** out);
/* Number of instances of the handle (1 for a handle from exa_create, 0 for NULL). */
int exa_batch_size(const exa_handle* h);
/* Per instance, from the last exa_solve: Newton iterations, GMRES iterations and the final
 * residual norm ||b|| + ||Ru||; each array (may be NULL) holds at least n >= ninst entries. */
int exa_batch_stats(exa_handle* h, int* newton_iters, int* gmres_iters, double* final_norm, int n);
/* Instance inst's "uh" [ncu,npf,nf_i] or "udg" [npe,nc,ne_i]: copies it to dst (len >= size)
 * and returns its size; with dst NULL returns the size only.  Negative EXA_ERR_* on failure. */
long exa_batch_get(exa_handle* h, int inst, const char* name, double* dst, long len);
/* Iteration-level batching (no counterpart in the reference): for steady problems one Newton
 * iteration (the body of NewtonSolver's loop, backend/Solution/solution.cpp:125-242, with its
 * GMRES, damping and the mixed form's refinement), for DIRK problems one time step (its stages,
 * each with its Newton solve, solution.cpp:244-330), of every instance that still iterates, in
 * shared launches; each instance runs at its own model time.  status[i] (n >= ninst entries, may
 * be NULL): 1 still iterating, 0 converged, at the Newton iteration limit or through its time
 * steps, -1 broken down; *iterating (may be NULL) = the number still iterating.  Repeated steps
 * give each instance exa_solve's iterates bit for bit; exa_batch_stats reports the counts of the
 * problem now in each slot.  Read a finished slot's arrays (exa_batch_get) before the next step:
 * a time step's stage updates also touch finished slots.  exa_solve and exa_reset end the step
 * sequence (the next step starts every instance afresh).  Time-dependent batches need one
 * time-step length; BDF batches and model libraries ("plugin:") run through exa_solve only. */
int exa_batch_step(exa_handle* h, int* status, int n, int* iterating);
/* Put the problem of datain_dir into slot inst of a batch whose problem in that slot has
 * finished (status 0 or -1): the same element shape, model options, time scheme and steps and
 * mesh (mesh.bin connectivity and boundary faces) as the slot's problem, any coordinates, initial
 * state, auxiliary fields, physicsparam and tau (no longer than the batch's).  The slot's
 * geometry and q-equation matrices are set up again and its Newton iteration (or its first time
 * step, from the initial time) starts at the next exa_batch_step, beside the instances already
 * under way. */
int exa_batch_refill(exa_handle* h, int inst, const char* datain_dir);

/* ---- the reference library's entry point ----
 * libmain replaces backend/Main/libexasim.h:5 (int libmain(int argc, char** argv), defined in
 * backend/Main/libexasim.cpp, which hands argv without the model name to premain,
 * backend/Main/exasim.hpp:143-360).  argv = {program, model, nummodels, InputFile, OutputFile
 * [, restart]}, as in the reference's CI call "cpuEXASIMMAINLIB Poisson3D 1 datain/ dataout/"
 * (.github/workflows/exasim-ci.yml:49-54).  InputFile is the datain directory; OutputFile the prefix of
 * udg_np0.bin / uhat_np0.bin.  model: a model name of exa_create, or the reference's built-in
 * "Poisson3D" (= "poisson").  Only nummodels = 1 and restart = 0.  Solves the problem on the current
 * device, writes the output files and returns 0 or an EXA_ERR_* code (the reference exits on errors). */
int libmain(int argc, char** argv);

#ifdef __cplusplus
}
#endif
#endif
