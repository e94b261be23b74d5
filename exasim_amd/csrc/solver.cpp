// solver.cpp -- setup and host control of the HDG Newton-GMRES path.
//
// Setup follows the CDiscretization constructor (backend/Discretization/
// discretization.cpp:170-412): read datain, buildConn, geometry, qEquation,
// initial uhat.  assemble() is hdgAssembleLinearSystem (discretization.cpp:483-512)
// with uEquationHDG's element-block loop (uequation.cpp:778-787); gmres() is the
// HDG GMRES of backend/Solver/gmres.cpp:361-597; solve() is NewtonSolver
// (backend/Solution/solution.cpp:125-242).  All device work runs on one stream
// owned by the handle; the host only reads back the scalars GMRES and Newton
// branch on.
#include "solver.hpp"
#include "models.hpp"

#include <algorithm>
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstring>

namespace exa {

int resolve_model(const std::string& name) {
  if (name.rfind("plugin:", 0) == 0) return MODEL_PLUGIN;
  if (name == "poisson") return MODEL_POISSON;
  if (name == "poisson_poly") return MODEL_POISSON_POLY;
  if (name == "convdiff") return MODEL_CONVDIFF;
  int nd, ncu, nv;
  const int id = generated_model(name, &nd, &ncu, &nv);
  if (id >= 0) return id;
  throw Error(1, "unknown model '" + name + "'");
}

// A model library (model name "plugin:<path>"): text2code's GPU form of the model functions,
// built with include/CudaHip.hpp; HdgFlux, HdgSource and HdgFbou with the signatures of
// backend/Model/libpdemodel.hpp:1-41 (required), exa_plugin_set_stream (optional).  The library
// stays loaded until the handle is destroyed.  One instance per handle (the functions take one
// physicsparam array); the boundary conditions are 1 .. the largest one mesh.bin's bf uses.
void Solver::load_plugin() {
  if (batch && ninst > 1)
    throw Error(5, "a model library serves single problems (batches of one instance)");
  const std::string path = model_name.substr(7);
  plugin_handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!plugin_handle) {
    const char* e = dlerror();
    throw Error(1, "cannot load model library '" + path + "': " + (e ? e : "unknown error"));
  }
  auto sym = [&](const char* name, bool need) -> void* {
    void* f = dlsym(plugin_handle, name);
    if (!f && need) throw Error(1, "model library '" + path + "' has no " + name);
    return f;
  };
  plugin_fns.flux = reinterpret_cast<PlugFluxFn>(sym("HdgFlux", true));
  plugin_fns.source = reinterpret_cast<PlugFluxFn>(sym("HdgSource", true));
  plugin_fns.fbou = reinterpret_cast<PlugFbouFn>(sym("HdgFbou", true));
  plugin_fns.set_stream = reinterpret_cast<PlugStreamFn>(sym("exa_plugin_set_stream", false));
  int nbc = 0;
  for (int b : mesh.bf) nbc = std::max(nbc, b);
  plugin_fns.nbc = nbc;
  plugin_fns.modelnumber = app.flag.size() > 12 ? static_cast<int>(app.flag[12]) : 0;
}

Solver::Solver(const std::string& din, const std::string& dout, const std::string& mname, int dev)
    : datain(din), dataout(dout), model_name(mname), model(resolve_model(mname)), device(dev) {
  setup();
}

Solver::~Solver() {
  for (auto& v : ev)
    for (auto& p : v) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  if (s) (void)hipStreamSynchronize(s);
  if (hpin) (void)hipHostFree(hpin);
  if (hseg) (void)hipHostFree(hseg);
  for (auto& gx : gm_graphs)
    if (gx) (void)hipGraphExecDestroy(gx);
  if (hstat) {
    (void)hipHostFree(hstat);
    for (auto& e : gm_ev)
      if (e) (void)hipEventDestroy(e);
  }
  if (s) (void)hipStreamDestroy(s);
  if (plugin_handle) dlclose(plugin_handle);
}

void Solver::setup() {
  app = read_app(datain + "/app.bin");
  master = read_master(datain + "/master.bin");
  mesh = read_mesh(datain + "/mesh.bin");
  set_dims();
  SolData sol = read_sol(datain + "/sol.bin", d.npe, d.nc, d.ncu, d.ne);
  conn = build_conn(master, mesh);
  init(sol);
}

// pivot workspace of the batched LU for `batch` matrices (grown with the stream idle)
void* Solver::lu_tri(long batch) {
  if (!lu_inverses) return nullptr;
  const size_t need = static_cast<size_t>(std::max(1L, batch)) * LU_TRI_WS;
  if (lutri.n < need) {
    EXA_HIP(hipStreamSynchronize(s));
    lutri.alloc(need);
  }
  return lutri.p;
}

int* Solver::lu_ws(long batch) {
  const size_t need = static_cast<size_t>(std::max(1L, batch)) * LU_PIV_WS;
  if (lupiv.n < need) {
    EXA_HIP(hipStreamSynchronize(s));
    lupiv.alloc(need);
  }
  return lupiv.p;
}

void Solver::set_dims() {
  if (master.nd != mesh.nd) throw Error(2, "master.bin and mesh.bin disagree on nd");
  d.nd = master.nd;
  d.nc = static_cast<int>(app.ndims[5]);
  d.ncu = static_cast<int>(app.ndims[6]);
  d.npe = master.npe;
  d.npf = master.npf;
  d.nge = master.nge;
  d.ngf = master.ngf;
  d.ne = mesh.ne;
  d.nfe = mesh.nfe;
  if (d.nc != d.ncu * (d.nd + 1))
    throw Error(5, "only the HDG (u, q) formulation with nc = ncu*(nd+1) is supported");
}

void Solver::init(SolData& sol) {
  if (static_cast<long>(sol.xdg.size()) != static_cast<long>(d.npe) * d.nd * d.ne)
    throw Error(2, "xdg must have ncx = nd components");
  d.nco = sol.nco;
  // the point kernels hold the auxiliary fields of a quadrature point in registers
  if (d.nco > 8)
    throw Error(5, "sol.bin has " + std::to_string(d.nco) + " auxiliary fields (vdg); at most 8 are supported");
  if (model == MODEL_PLUGIN) load_plugin();
  if (model >= 100) {  // a generated model is compiled for one (nd, ncu)
    int nd, ncu, nv;
    generated_model(model_name, &nd, &ncu, &nv);
    if (nd != d.nd || ncu != d.ncu)
      throw Error(1, "model '" + model_name + "' is for nd = " + std::to_string(nd) + ", ncu = " +
                         std::to_string(ncu) + " but datain has nd = " + std::to_string(d.nd) +
                         ", ncu = " + std::to_string(d.ncu));
    if (nv > d.nco)
      throw Error(1, "model '" + model_name + "' reads " + std::to_string(nv) +
                         " auxiliary components (vdg) but sol.bin has " + std::to_string(d.nco));
  }
  d.nf = conn.nf;
  d.n = d.npe * d.ncu;
  d.ndf = d.npf * d.nfe;
  d.m = d.ndf * d.ncu;
  d.ncf = d.npf * d.ncu;
  d.N = static_cast<long>(d.ncf) * d.nf;

  // inputs are read and checked before the device is touched
  if (device >= 0) EXA_HIP(hipSetDevice(device));
  EXA_HIP(hipGetDevice(&device));
  EXA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  noalloc_init(device);
  EXA_HIP(hipHostMalloc(reinterpret_cast<void**>(&hpin), 4096 * sizeof(double), hipHostMallocDefault));
  st.init(size_t(32) << 20);

  std::vector<double>& fl = app.flag;
  std::vector<double>& pr = app.problem;
  // time dependence (setstructs.cpp:126-149, 194): flag[0] tdep, flag[1] wave, flag[10] tdfunc,
  // problem[2] temporal scheme (0 DIRK), problem[3] order, problem[4] stages
  tdep = static_cast<int>(fl[0]);
  tdfunc = static_cast<int>(fl[10]);  // 1: the model's Tdfunc multiplies the time derivative
  if (tdep) {
    if (static_cast<int>(fl[1]) != 0) throw Error(5, "wave problems are not supported");
    tscheme = static_cast<int>(pr[2]);
    if (tscheme != 0 && tscheme != 1) throw Error(5, "temporal scheme " + std::to_string(tscheme) +
                                                         ": 0 (DIRK) and 1 (BDF) are supported");
    torder = static_cast<int>(pr[3]);
    tstages = tscheme == 1 ? 1 : static_cast<int>(pr[4]);
    dts = app.dt;
    time0 = app.factor.empty() ? 0.0 : app.factor[0];
    if (tscheme == 0) dirk_coefficients();
    else if (torder < 1 || torder > 3) throw Error(5, "BDF" + std::to_string(torder) + " is not implemented");
  }
  tnow = time0;
  precond = static_cast<int>(fl[6]);
  if (precond < 0 || precond > 2)
    throw Error(5, "preconditioner " + std::to_string(precond) +
                       " is not implemented (0 block-Jacobi, 1 elemental additive Schwarz, "
                       "2 superelement block-ILU(0))");
  ortho = static_cast<int>(fl[5]);
  newton_maxit = static_cast<int>(pr[13]);
  gmres_maxit = static_cast<int>(pr[14]);
  restart = static_cast<int>(pr[15]);
  if (restart < 1 || restart > 2000) throw Error(2, "GMRES restart in app.bin must be in [1, 2000]");
  newton_tol = app.solversparam[0];
  gmres_tol = app.solversparam[1];
  if (static_cast<int>(pr[27]) > 1) throw Error(5, "polynomial preconditioner (ppdegree > 1) is not supported");
  // the w equation (wequation.cpp:31-240): ncw = ndims[13], dae_alpha = factor[1], dae_beta = factor[2]
  ncw = app.ndims.size() > 13 ? static_cast<int>(app.ndims[13]) : 0;
  dae_alpha = app.factor.size() > 1 ? app.factor[1] : 1.0;
  dae_beta = app.factor.size() > 2 ? app.factor[2] : 0.0;
  if (ncw != model_ncw_of(model))
    throw Error(1, "app.bin has ncw = " + std::to_string(ncw) + " but the model '" + model_name + "' has " +
                       std::to_string(model_ncw_of(model)) + " w components (a Sourcew function)");
  if (ncw > 0 && static_cast<long>(sol.wdg.size()) != 0 &&
      static_cast<long>(sol.wdg.size()) != static_cast<long>(d.npe) * ncw * d.ne)
    throw Error(2, "wdg in sol.bin has the wrong size");
  if (static_cast<int>(pr[16]) > 0) throw Error(5, "reduced-basis initial guess (RBdim > 0) is not supported");

  // master tables, mesh, fields
  shapegt.upload(master.shapegt, st, s);
  shapegw.upload(master.shapegw, st, s);
  shapfgt.upload(master.shapfgt, st, s);
  shapfgw.upload(master.shapfgw, st, s);
  if (!batch) {
    if (app.physicsparam.empty()) app.physicsparam.push_back(0.0);
    if (app.tau.empty()) app.tau.push_back(0.0);
    nparam = static_cast<int>(app.physicsparam.size());
    ntau = static_cast<int>(app.tau.size());
    inst_param = app.physicsparam;
    inst_tau = app.tau;
    eoff = {0, d.ne};
    foff = {0, d.nf};
  }
  {  // element -> instance map (all zero for one problem): the matvec and the Schwarz apply skip
     // the elements of instances that have left GMRES
    std::vector<int> ei(d.ne);
    for (int i = 0; i < ninst; ++i)
      for (long e = eoff[i]; e < eoff[i + 1]; ++e) ei[e] = i;
    einst.upload(ei, st, s);
  }
  param.upload(inst_param, st, s);
  tau.upload(inst_tau, st, s);
  if (model == MODEL_PLUGIN) {  // the model library's uinf argument (app.bin's externalparam)
    std::vector<double> ui = app.uinf;
    if (ui.empty()) ui.push_back(0.0);
    uinf_d.upload(ui, st, s);
    plugin_fns.uinf = uinf_d.p;
  }
  seg_setup();
  perm.upload(mesh.perm, st, s);
  elemcon.upload(conn.elemcon, st, s);
  bf.upload(mesh.bf, st, s);
  bface_h.clear();
  for (long f = 0; f < static_cast<long>(d.ne) * d.nfe; ++f)
    if (mesh.bf[f] > 0) bface_h.push_back(static_cast<int>(f));
  if (!bface_h.empty()) bface.upload(bface_h, st, s);
  f2e.upload(conn.f2e, st, s);
  xdg.upload(sol.xdg, st, s);

  // udg with all nc components (readsolstruct: missing q components are zero;
  // a missing udg is Initu, which is zero for the models here)
  std::vector<double> hudg(static_cast<size_t>(d.npe) * d.nc * d.ne, 0.0);
  if (!sol.udg.empty()) {
    int k = sol.udg_ncomp;
    for (long e = 0; e < d.ne; ++e)
      for (int c = 0; c < k; ++c)
        for (int i = 0; i < d.npe; ++i)
          hudg[i + d.npe * (c + d.nc * e)] = sol.udg[i + d.npe * (c + k * e)];
  }
  udg.upload(hudg, st, s);
  // initial uhat (GetFaceNodes from the first element of each face, kokkosimpl.h:901-915)
  std::vector<double> huh(static_cast<size_t>(d.N), 0.0);
  if (!sol.uh.empty()) {
    if (static_cast<long>(sol.uh.size()) != d.N) throw Error(2, "uhat in sol.bin has the wrong size");
    huh = sol.uh;
  } else {
    for (long f = 0; f < d.nf; ++f) {
      int e1 = conn.f2e[4 * f], l1 = conn.f2e[4 * f + 1];
      for (int j = 0; j < d.npf; ++j)
        for (int c = 0; c < d.ncu; ++c)
          huh[c + d.ncu * (j + d.npf * f)] =
              hudg[mesh.perm[j + d.npf * l1] + d.npe * (c + static_cast<long>(d.nc) * e1)];
    }
  }
  uh.upload(huh, st, s);
  udg0.upload(hudg, st, s);
  uh0.upload(huh, st, s);
  if (ncw > 0) {  // initial w from sol.bin (zero when absent); wsrc starts at zero
    std::vector<double> hw = sol.wdg;
    if (hw.empty()) hw.assign(static_cast<size_t>(d.npe) * ncw * d.ne, 0.0);
    wdg.upload(hw, st, s);
    wdg0.upload(hw, st, s);
    wsrc.alloc(hw.size());
    wsrc.zero(s);
  }

  // face-node gather sources for the atomic-free element -> face assembly
  long nn = static_cast<long>(d.npf) * d.nf;
  std::vector<long> h1(nn), h2(nn);
  for (long f = 0; f < d.nf; ++f) {
    int e1 = conn.f2e[4 * f], l1 = conn.f2e[4 * f + 1], e2 = conn.f2e[4 * f + 2], l2 = conn.f2e[4 * f + 3];
    for (int j = 0; j < d.npf; ++j) {
      long g = j + d.npf * f;
      h1[g] = static_cast<long>(e1) * d.m + d.ncu * (j + d.npf * l1);
      if (e2 >= 0) {
        long ii = conn.elemcon[j + d.npf * (l2 + static_cast<long>(d.nfe) * e2)] - d.npf * f;
        h2[g] = static_cast<long>(e2) * d.m + d.ncu * (ii + d.npf * l2);
      } else {
        h2[g] = -1;
      }
    }
  }
  src1.upload(h1, st, s);
  src2.upload(h2, st, s);

  // geometry
  egeo.alloc(static_cast<size_t>(d.ne) * d.nge * (d.nd + d.nd * d.nd + 1));
  fgeo.alloc(static_cast<size_t>(d.ne) * d.nfe * d.ngf * (2 * d.nd + 1));
  elem_geom(d, shapegt.p, xdg.p, egeo.p, s);
  face_geom(d, shapfgt.p, perm.p, xdg.p, fgeo.p, s);
  if (d.nco > 0) {
    odg.upload(sol.odg, st, s);
    og.alloc(static_cast<size_t>(d.ne) * d.nge * d.nco);
    ogf.alloc(static_cast<size_t>(d.ne) * d.nfe * d.ngf * d.nco);
    interp_elem(d, d.nco, shapegt.p, odg.p, og.p, s);
    interp_face(d, d.nco, shapfgt.p, perm.p, odg.p, ogf.p, s);
  }

  // qEquation: per element solve Mass X = [Mass | C_1..C_nd | E_1..E_nd]
  MCs.alloc(static_cast<size_t>(d.ne) * d.npe * d.nd * d.npe);
  MEs.alloc(static_cast<size_t>(d.ne) * d.npe * d.nd * d.ndf);
  Mass.alloc(static_cast<size_t>(d.ne) * d.npe * d.npe);
  info.alloc(std::max(d.ne, d.nf) + 1);
  setup_q(0, d.ne);

  // persistent system arrays
  H.alloc(static_cast<size_t>(d.ne) * d.m * d.m);
  DinvF.alloc(static_cast<size_t>(d.ne) * d.n * d.m);
  RuS.alloc(static_cast<size_t>(d.ne) * d.n);
  RuR.alloc(static_cast<size_t>(d.ne) * d.n);
  Rh.alloc(static_cast<size_t>(d.ne) * d.m);
  b.alloc(d.N);
  bres.alloc(d.N);
  we.alloc(static_cast<size_t>(d.ne) * d.m);
  du.alloc(static_cast<size_t>(d.ne) * d.n);
  BE.alloc(static_cast<size_t>(d.nf) * d.ncf * d.ncf);
  bjaug.alloc(static_cast<size_t>(d.nf) * d.ncf * 2 * d.ncf);
  scal.alloc(8);
  x.alloc(d.N);
  r.alloc(d.N);
  tmp.alloc(d.N);
  wv.alloc(d.N);
  alloc_workspace();
  if (precond == 2) setup_bilu0();
  if (precond == 1) setup_eas();
  if (tdep) {
    const size_t nu = static_cast<size_t>(d.npe) * d.ncu * d.ne;
    sdg.alloc(nu);
    udgprev.alloc(nu);
    utmp.alloc(nu);
    sdg.zero(s);
    if (tscheme == 1) bdfhist.alloc(nu * torder);
    if (ncw > 0) {
      const size_t nw = static_cast<size_t>(d.npe) * ncw * d.ne;
      wprev.alloc(nw);
      wtmp.alloc(nw);
      if (tscheme == 1) wbdfhist.alloc(nw * torder);
    }
  }
}

void Solver::setup_q(long e_begin, long ne) {
  const long ncq = static_cast<long>(d.npe) + d.npe * d.nd + static_cast<long>(d.ndf) * d.nd;
  const long per = d.npe * ncq;
  const int qb = static_cast<int>(std::max(1L, std::min<long>(ne, static_cast<long>(1.0e9 / 8 / per))));
  DevBuf<double> qaug, qA;
  qaug.alloc(static_cast<size_t>(qb) * per);
  qA.alloc(static_cast<size_t>(qb) * (d.nd + 1) * d.npe * d.nge);
  for (long e0 = e_begin; e0 < e_begin + ne; e0 += qb) {
    const int eb = static_cast<int>(std::min<long>(qb, e_begin + ne - e0));
    EXA_HIP(hipMemsetAsync(qaug.p, 0, static_cast<size_t>(eb) * per * sizeof(double), s));
    qsetup(d, static_cast<int>(e0), eb, shapegt.p, shapegw.p, shapfgt.p, shapfgw.p, perm.p, egeo.p, fgeo.p,
           qaug.p, Mass.p, s, qA.p);
    lu_solve(qaug.p, d.npe, static_cast<int>(ncq), per, eb, info.p, s, nullptr, lu_ws(eb), lu_tri(eb));
    const int bad = first_nonzero(info.p, eb);
    if (bad >= 0) throw Error(4, "singular mass matrix at element " + std::to_string(e0 + bad));
    qpack(d, static_cast<int>(e0), eb, qaug.p, MCs.p, MEs.p, s);
  }
  EXA_HIP(hipStreamSynchronize(s));
}

// DIRKcoeff (timestepcoeff.cpp:35-104) computed from the Butcher tableaux of the
// schemes it lists: d = A^{-1} (column-major d[i + S*j]), c = b^T A^{-1}, t = A 1.
void Solver::dirk_coefficients() {
  const int S = tstages;
  double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, b[3] = {0, 0, 0};
  if (S == 1 && torder == 1) {  // backward Euler
    A[0][0] = 1.0; b[0] = 1.0;
  } else if (S == 1 && torder == 2) {  // implicit midpoint
    A[0][0] = 0.5; b[0] = 1.0;
  } else if (S == 2 && torder == 2) {  // Alexander, L-stable
    const double g = 1.0 - 1.0 / std::sqrt(2.0);
    A[0][0] = g; A[1][0] = 1.0 - g; A[1][1] = g; b[0] = 1.0 - g; b[1] = g;
  } else if (S == 2 && torder == 3) {  // Crouzeix
    const double g = (3.0 + std::sqrt(3.0)) / 6.0;
    A[0][0] = g; A[1][0] = 1.0 - 2.0 * g; A[1][1] = g; b[0] = b[1] = 0.5;
  } else if (S == 3 && torder == 3) {  // Alexander, L-stable, stiffly accurate
    const double g = 0.4358665215;
    A[0][0] = g; A[1][0] = (1.0 - g) / 2.0; A[1][1] = g;
    A[2][0] = -(6.0 * g * g - 16.0 * g + 1.0) / 4.0; A[2][1] = (6.0 * g * g - 20.0 * g + 5.0) / 4.0; A[2][2] = g;
    b[0] = A[2][0]; b[1] = A[2][1]; b[2] = g;
  } else if (S == 3 && torder == 4) {  // Crouzeix
    const double g = std::cos(M_PI / 18.0) / std::sqrt(3.0) + 0.5;
    const double dl = 1.0 / (6.0 * (2.0 * g - 1.0) * (2.0 * g - 1.0));
    A[0][0] = g; A[1][0] = 0.5 - g; A[1][1] = g; A[2][0] = 2.0 * g; A[2][1] = 1.0 - 4.0 * g; A[2][2] = g;
    b[0] = dl; b[1] = 1.0 - 2.0 * dl; b[2] = dl;
  } else {
    throw Error(5, "DIRK(" + std::to_string(S) + "," + std::to_string(torder) + ") is not implemented");
  }
  // inverse of the lower-triangular A by forward substitution, column by column
  double Ai[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int j = 0; j < S; ++j)
    for (int i = j; i < S; ++i) {
      double v = (i == j) ? 1.0 : 0.0;
      for (int k = j; k < i; ++k) v -= A[i][k] * Ai[k][j];
      Ai[i][j] = v / A[i][i];
    }
  for (int i = 0; i < S; ++i) {
    dirk_t[i] = 0.0;
    for (int k = 0; k < S; ++k) dirk_t[i] += A[i][k];
    dirk_c[i] = 0.0;
    for (int k = 0; k < S; ++k) dirk_c[i] += b[k] * Ai[k][i];
    for (int j = 0; j < S; ++j) dirk_d[i + S * j] = Ai[i][j];
  }
}

// hdgBlockILU0 (matvec.cpp:133-165) batched over the superelements: for each face
// slot i in order, invert the diagonal block, form L_ji = A_ji D_i^{-1} for the
// blocks below it, and update A_jl -= L_ji A_il; the products are batched GEMMs
// whose operands are gathered through the elimination index lists.
namespace {
// operands of the block-ILU(0) factorisation products: FP64, or FP32 storage and products
void ilu_gemm_ptrs(Gemm& g, double* A, double* B, double* C) {
  g.A = A;
  g.B = B;
  g.C = C;
}
void ilu_gemm_ptrs(Gemm& g, float* A, float* B, float* C) {
  g.Af = A;
  g.Bf = B;
  g.Cf = C;
  g.fp32 = 1;
}
}  // namespace

template <typename T>
void Solver::factor_bilu0_t(T* BE, T* L) {
  const int ncf = d.ncf;
  const long per = static_cast<long>(ncf) * ncf, slab = per * ilu.nse;
  for (int i = 0; i < ilu.nfse; ++i) {
    bilu_invert_diag(d, ilud, i, BE, info.p, s);
    const int k0 = ilu.eptr[i], ne_i = ilu.eptr[i + 1] - k0;
    if (ne_i == 0) continue;
    Gemm g{};
    g.M = g.N = g.K = ncf;
    g.alpha = 1.0; g.beta = 0.0;
    ilu_gemm_ptrs(g, BE, BE + ilu.diag[i] * slab, L);
    g.iA = ilud.elim_ji + k0; g.lda = ncf; g.sA0 = slab; g.sA1 = per;
    g.ldb = ncf; g.sB0 = 0; g.sB1 = per;
    g.ldc = ncf; g.sC0 = slab; g.sC1 = per;
    g.nb0 = ne_i; g.nb1 = ilu.nse; g.nb2 = 1;
    gemm(g, s);
    bilu_put_l(d, ilud, k0, ne_i, L, BE, s);
    const int u0 = ilu.uptr[k0], nu = ilu.uptr[k0 + ne_i] - u0;
    if (nu == 0) continue;
    Gemm u{};
    u.M = u.N = u.K = ncf;
    u.alpha = -1.0; u.beta = 1.0;
    ilu_gemm_ptrs(u, BE, BE, BE);
    u.iA = ilud.upd_ji + u0; u.lda = ncf; u.sA0 = slab; u.sA1 = per;
    u.iB = ilud.upd_il + u0; u.ldb = ncf; u.sB0 = slab; u.sB1 = per;
    u.iC = ilud.upd_jl + u0; u.ldc = ncf; u.sC0 = slab; u.sC1 = per;
    u.nb0 = nu; u.nb1 = ilu.nse; u.nb2 = 1;
    gemm(u, s);
  }
}

void Solver::factor_bilu0() { factor_bilu0_t(BEilu.p, Lilu.p); }

void Solver::alloc_mixed() {
  const size_t ne = d.ne, n = d.n, m = d.m, ncf2 = static_cast<size_t>(d.ncf) * d.ncf;
  H32.alloc(m * m * ne);
  DinvF32.alloc(n * m * ne);
  LU32.alloc(n * n * ne);
  piv32.alloc(n * ne);
  K32.alloc(m * n * ne);
  // FP32 copies of the FP64 operands of the mixed-form assembly (its FP32 element-block
  // workspace is sized by alloc_workspace)
  if (!q32_ready) {
    MCs32.alloc(MCs.n);
    MEs32.alloc(MEs.n);
    shapegt32.alloc(shapegt.n);
    const int rq = d.npe * d.nd;  // rows of M^-1 C and M^-1 E
    to_float(MCs32.p, MCs.p, rq, d.npe * d.ne, rq, 0, 0, 1, s);
    to_float(MEs32.p, MEs.p, rq, d.ndf * d.ne, rq, 0, 0, 1, s);
    to_float(shapegt32.p, shapegt.p, static_cast<int>(shapegt.n), 1, static_cast<long>(shapegt.n), 0, 0, 1, s);
    q32_ready = true;
  }
  if (precond == 0) BE32.alloc(ncf2 * d.nf);
  if (precond == 1) {  // the mixed form inverts the Schwarz blocks in FP32 (applied in FP32 or BF16)
    BEeas32.alloc(m * m * ne);
    easaug32.alloc(static_cast<size_t>(std::max(eas_chunk, 1)) * 2 * m * m);
    if (precond_bf16) BEeas16.alloc(m * m * ne);
  }
  if (precond == 2) {  // FP32 factorisation (and its factors), BF16 copy of the factors
    BEilu32.alloc(ncf2 * ilu.nnz * ilu.nse);
    Lilu32.alloc(Lilu.n);
    if (precond_bf16) BEilu16.alloc(ncf2 * ilu.nnz * ilu.nse);
  }
}

void Solver::setup_eas() {
  if (d.m > 400) throw Error(5, "elemental additive Schwarz: element trace block larger than 400");
  const double per = 16.0 * d.m * d.m;  // bytes of one augmented [block | I]
  eas_chunk = static_cast<int>(std::max(1.0, std::min<double>(d.ne, ws_budget_bytes / per)));
  BEeas.alloc(static_cast<size_t>(d.m) * d.m * d.ne);
  easaug.alloc(static_cast<size_t>(eas_chunk) * 2 * d.m * d.m);
  if (info.n < static_cast<size_t>(eas_chunk)) info.alloc(eas_chunk + 1);
}

void Solver::setup_bilu0() {
  if (ilu_ready) return;
  ilu = build_bilu0(mesh.cartgridpart, conn, d.nfe, d.ne);
  std::vector<int> row_of(ilu.nnz);
  for (int i = 0; i < ilu.nfse; ++i)
    for (int t = ilu.row_ptr[i]; t < ilu.row_ptr[i + 1]; ++t) row_of[t] = i;
  auto up = [&](DevBuf<int>& b, const std::vector<int>& v) {
    static const std::vector<int> zero1{0};
    b.upload(v.empty() ? zero1 : v, st, s);
    return b.p;
  };
  ilud.nse = ilu.nse;
  ilud.nfse = ilu.nfse;
  ilud.nnz = ilu.nnz;
  ilud.face = up(iface, ilu.face);
  ilud.row_ptr = up(irow_ptr, ilu.row_ptr);
  ilud.col_ind = up(icol_ind, ilu.col_ind);
  ilud.row_of = up(irow_of, row_of);
  ilud.diag = up(idiag, ilu.diag);
  ilud.eptr = up(ieptr, ilu.eptr);
  ilud.elim_ji = up(ielim_ji, ilu.elim_ji);
  ilud.uptr = up(iuptr, ilu.uptr);
  ilud.upd_jl = up(iupd_jl, ilu.upd_jl);
  ilud.upd_il = up(iupd_il, ilu.upd_il);
  ilud.upd_ji = up(iupd_ji, ilu.upd_ji);
  ilud.lptr = up(ilptr, ilu.lptr);
  ilud.lidx = up(ilidx, ilu.lidx);
  ilud.lcol = up(ilcol, ilu.lcol);
  ilud.uuptr = up(iuuptr, ilu.uuptr);
  ilud.uidx = up(iuidx, ilu.uidx);
  ilud.ucol = up(iucol, ilu.ucol);
  ilud.slot1 = up(islot1, ilu.slot1);
  ilud.slot2 = up(islot2, ilu.slot2);
  {  // the apply's sweeps as one flat program (ApplyBlockILU0's loop order)
    std::vector<int> wseg(1, 0), wblk, wvec;
    for (int i = 0; i < ilu.nfse; ++i) {
      for (int q = ilu.lptr[i]; q < ilu.lptr[i + 1]; ++q) {
        wblk.push_back(ilu.lidx[q]);
        wvec.push_back(ilu.lcol[q] * d.ncf);
      }
      wseg.push_back(static_cast<int>(wblk.size()));
    }
    for (int i = ilu.nfse - 1; i >= 0; --i) {
      for (int q = ilu.uuptr[i]; q < ilu.uuptr[i + 1]; ++q) {
        wblk.push_back(ilu.uidx[q]);
        wvec.push_back(ilu.ucol[q] * d.ncf);
      }
      wseg.push_back(static_cast<int>(wblk.size()));
      wblk.push_back(ilu.diag[i]);
      wvec.push_back(ilu.nfse * d.ncf);
      wseg.push_back(static_cast<int>(wblk.size()));
    }
    ilud.wseg = up(iwseg, wseg);
    ilud.wblk = up(iwblk, wblk);
    ilud.wvec = up(iwvec, wvec);
  }
  BEilu.alloc(static_cast<size_t>(ilu.nnz) * ilu.nse * d.ncf * d.ncf);
  Yilu.alloc(static_cast<size_t>(ilu.nfse) * ilu.nse * d.ncf);
  Xilu.alloc(static_cast<size_t>(ilu.nfse) * ilu.nse * d.ncf);
  int max_elim = 1;
  for (int i = 0; i < ilu.nfse; ++i) max_elim = std::max(max_elim, ilu.eptr[i + 1] - ilu.eptr[i]);
  Lilu.alloc(static_cast<size_t>(max_elim) * ilu.nse * d.ncf * d.ncf);
  if (info.n < static_cast<size_t>(ilu.nse)) info.alloc(ilu.nse + 1);
  ilu_ready = true;
}

// Element-block workspace of the assembly, sized to ws_budget_bytes.  The point values (rv, rJ,
// fpt) are FP64 in both forms; the block matrices W, [D | F | Ru], B and G are FP64 in the FP64
// form and FP32 in the mixed form (which assembles K in place in its persistent K32), so the
// mixed form fits about twice the elements per block in the same budget.
void Solver::alloc_workspace() {
  const long n = d.n, m = d.m;
  const long pts = static_cast<long>(d.nd + 1) * d.ncu * d.nge + static_cast<long>(d.nd + 1) * d.ncu * d.nc * d.nge +
                   2L * (d.ncu + d.ncu * d.nc + d.ncu * d.ncu) * d.nfe * d.ngf;
  const long blocks = n * augcols() + n * n * d.nd + m * n * d.nd + static_cast<long>(d.npe) * d.nge * d.ncu * d.nc;
  const double per = precision ? 8.0 * pts + 4.0 * blocks : 8.0 * (pts + blocks + m * n);
  long cap = static_cast<long>(ws_budget_bytes / per);
  neb = static_cast<int>(std::max(1L, std::min<long>(d.ne, cap)));
  if (precision) {
    aug.release();
    Bfull.release();
    Kt.release();
    Gfull.release();
    W.release();
    aug32.alloc(static_cast<size_t>(neb) * n * augcols());
    Bfull32.alloc(static_cast<size_t>(neb) * n * n * d.nd);
    Gfull32.alloc(static_cast<size_t>(neb) * m * n * d.nd);
    W32.alloc(static_cast<size_t>(neb) * d.npe * d.nge * d.ncu * d.nc);
  } else {
    aug32.release();
    Bfull32.release();
    Gfull32.release();
    W32.release();
    aug.alloc(static_cast<size_t>(neb) * n * augcols());
    Bfull.alloc(static_cast<size_t>(neb) * n * n * d.nd);
    Kt.alloc(static_cast<size_t>(neb) * m * n);
    Gfull.alloc(static_cast<size_t>(neb) * m * n * d.nd);
    W.alloc(static_cast<size_t>(neb) * d.npe * d.nge * d.ncu * d.nc);
  }
  rv.alloc(static_cast<size_t>(neb) * (d.nd + 1) * d.ncu * d.nge);
  rJ.alloc(static_cast<size_t>(neb) * (d.nd + 1) * d.ncu * d.nc * d.nge);
  fpt.alloc(static_cast<size_t>(neb) * 2 * (d.ncu + d.ncu * d.nc + d.ncu * d.ncu) * d.nfe * d.ngf);
  if (model == MODEL_PLUGIN) pbuf.alloc(static_cast<size_t>(plugin_buffer_doubles(d, plugin_fns.nbc, neb)));
  V.alloc(static_cast<size_t>(restart + 1) * d.N);
  partial.alloc(static_cast<size_t>(dot_partials_count(d.N)) * (restart + 2));
  seg_alloc();
}

void Solver::set_option(const std::string& name, double v) {
  clear_graphs();  // an option can change what a step launches
  if (name == "gmres_tol") gmres_tol = v;
  else if (name == "gmres_restart") {
    if (v < 1 || v > 2000) throw Error(1, "gmres_restart out of range");
    restart = static_cast<int>(v);
    V.alloc(static_cast<size_t>(restart + 1) * d.N);
    partial.alloc(static_cast<size_t>(dot_partials_count(d.N)) * (restart + 2));
    seg_alloc();
  } else if (name == "gmres_maxit") gmres_maxit = static_cast<int>(v);
  else if (name == "newton_tol") newton_tol = v;
  else if (name == "newton_maxit") newton_maxit = static_cast<int>(v);
  else if (name == "preconditioner") {
    int pv = static_cast<int>(v);
    if (pv < 0 || pv > 2)
      throw Error(5, "preconditioners: 0 block-Jacobi, 1 elemental additive Schwarz, "
                     "2 superelement block-ILU(0)");
    if (pv == 2) setup_bilu0();
    if (pv == 1) setup_eas();
    precond = pv;
    assembled = false;
  } else if (name == "ortho") ortho = static_cast<int>(v);
  else if (name == "gmres_graph") gm_graph = v != 0.0;
  else if (name == "noalloc_chunk") {
    if (v < 0) throw Error(1, "noalloc_chunk must be >= 0");
    noalloc_chunk = static_cast<int>(v);
  } else if (name == "gmres_lag") {
    if (v < 0 || v > GM_RING - 2) throw Error(1, "gmres_lag must be in [0, " + std::to_string(GM_RING - 2) + "]");
    gm_lag = static_cast<int>(v);
  }
  else if (name == "verbose") verbose = static_cast<int>(v);
  else if (name == "gemm_autotune") {  // process-wide: 1 time the GEMM tiles per shape, 0 default tile
    if (v != 0.0 && v != 1.0 && v != -1.0) throw Error(1, "gemm_autotune: 1, 0 or -1 (the environment's choice)");
    set_gemm_autotune(static_cast<int>(v));
  } else if (name == "lu_inverses") {  // FP32 LU: 1 triangle inverses on MFMA, 0 substitution
    lu_inverses = v != 0.0;
  } else if (name == "ilu_wave") {  // process-wide: FP32/BF16 block-ILU(0) apply, 1 one wave per superelement
    if (v != 0.0 && v != 1.0 && v != -1.0) throw Error(1, "ilu_wave: 1, 0 or -1 (the environment's choice)");
    set_bilu_wave(static_cast<int>(v));
  } else if (name == "gemm_tile") {  // process-wide: every MFMA product on tile v (0..4), -1 the tuner's
    set_gemm_tile(static_cast<int>(v));
  } else if (name == "precision") {
    int pv = static_cast<int>(v);
    if (pv != 0 && pv != 1) throw Error(1, "precision: 0 (FP64) or 1 (mixed FP32/FP64)");
    if (pv != precision) {
      precision = pv;
      alloc_workspace();
      assembled = false;
    }
  } else if (name == "refine_rtol") refine_rtol = v;
  else if (name == "refine_tol") refine_tol = v;
  else if (name == "refine_first_tol") refine_first_tol = v;
  else if (name == "precond_bf16") {  // mixed form: block-ILU(0) factors or Schwarz blocks stored in BF16
    const int pv = v != 0.0;
    if (pv != precond_bf16) {
      precond_bf16 = pv;
      if (precision) alloc_mixed();
      assembled = false;
    }
  }
  else if (name == "refine_maxit") refine_maxit = static_cast<int>(v);
  else if (name == "refine_cycles") {
    if (v < 0) throw Error(1, "refine_cycles must be >= 0");
    refine_cycles = static_cast<int>(v);
  }
  else if (name == "timing") timing = v != 0.0;
  else if (name == "workspace_bytes") {
    ws_budget_bytes = v;
    alloc_workspace();
  } else throw Error(1, "unknown option '" + name + "'");
}

AsmArgs Solver::asm_args(int e0, int eb) {
  AsmArgs a{};
  a.augs = static_cast<long>(d.n) * augcols();
  a.tdep = tdep;
  a.tdfunc = tdfunc;
  a.dtfactor = dtfactor;
  a.sdg = sdg.p;
  a.model = model;
  a.time = tnow;
  a.tinst = use_tinst ? tinst_d.p : nullptr;
  a.param = param.p;
  a.tau = tau.p;
  a.einst = einst.p;
  a.nparam = nparam;
  a.ntau = ntau;
  a.e0 = e0;
  a.eb = eb;
  a.shapegt = shapegt.p;
  a.shapegw = shapegw.p;
  a.shapfgt = shapfgt.p;
  a.shapfgw = shapfgw.p;
  a.perm = perm.p;
  a.elemcon = elemcon.p;
  a.bf = bf.p;
  a.bface = bface.p;
  {
    const int lo = static_cast<int>(std::lower_bound(bface_h.begin(), bface_h.end(), e0 * d.nfe) -
                                    bface_h.begin());
    const int hi = static_cast<int>(std::lower_bound(bface_h.begin(), bface_h.end(), (e0 + eb) * d.nfe) -
                                    bface_h.begin());
    a.bf0 = lo;
    a.nbf = hi - lo;
  }
  a.wdg = wdg.p;
  a.wsrc = wsrc.p;
  a.dae_alpha = dae_alpha;
  a.dae_beta = dae_beta;
  a.plugin = model == MODEL_PLUGIN ? &plugin_fns : nullptr;
  a.pbuf = pbuf.p;
  a.plugin_nbc = plugin_fns.nbc;
  a.udg = udg.p;
  a.uh = uh.p;
  a.egeo = egeo.p;
  a.fgeo = fgeo.p;
  a.og = og.p;
  a.ogf = ogf.p;
  a.rv = rv.p;
  a.rJ = rJ.p;
  a.W = W.p;
  a.fpt = fpt.p;
  return a;
}

void Solver::tic(int cls, hipEvent_t* a) {
  if (!timing) return;
  EXA_HIP(hipEventCreate(a));
  EXA_HIP(hipEventRecord(*a, s));
  (void)cls;
}

void Solver::toc(int cls, hipEvent_t a) {
  if (!timing) return;
  hipEvent_t b;
  EXA_HIP(hipEventCreate(&b));
  EXA_HIP(hipEventRecord(b, s));
  ev[cls].push_back({a, b});
}

void Solver::collect_timings() {
  EXA_HIP(hipStreamSynchronize(s));
  for (int c = 0; c < T_NCLASS; ++c) {
    for (auto& p : ev[c]) {
      float t = 0.f;
      EXA_HIP(hipEventElapsedTime(&t, p.first, p.second));
      ms[c] += t;
      cnt[c] += 1;
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
    ev[c].clear();
  }
}

int Solver::timings(double* out_ms, long* counts, int len) {
  collect_timings();
  int k = std::min(len, static_cast<int>(T_NCLASS));
  for (int c = 0; c < k; ++c) {
    if (out_ms) out_ms[c] = ms[c];
    if (counts) counts[c] = cnt[c];
  }
  return T_NCLASS;
}

void Solver::reset_timings() {
  collect_timings();
  ms.fill(0.0);
  cnt.fill(0);
}

double Solver::norm(const double* v, long n) {
  dot(v, v, n, partial.p, scal.p, 1, s);
  EXA_HIP(hipMemcpyAsync(hpin, scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
  EXA_HIP(hipStreamSynchronize(s));
  return hpin[0];
}

int Solver::first_nonzero(const int* inf, long n) {
  if (!ifirst.p) ifirst.alloc(1);
  ::exa::first_nonzero(inf, n, ifirst.p, s);
  EXA_HIP(hipMemcpyAsync(hpin, ifirst.p, sizeof(int), hipMemcpyDeviceToHost, s));
  EXA_HIP(hipStreamSynchronize(s));
  const int k = *reinterpret_cast<const int*>(hpin);
  return k == 0x7fffffff ? -1 : k;
}

void Solver::reset() {
  copy(udg.p, udg0.p, static_cast<long>(udg.n), s);
  copy(uh.p, uh0.p, static_cast<long>(uh.n), s);
  if (ncw > 0) {
    copy(wdg.p, wdg0.p, static_cast<long>(wdg.n), s);
    wsrc.zero(s);
  }
  tnow = time0;
  dtfactor = 0.0;
  if (tdep) sdg.zero(s);
  assembled = false;
  step_act.clear();
  step_valid = false;
}

void Solver::get_q() { getq(d, MCs.p, MEs.p, elemcon.p, uh.p, udg.p, s); }

// GetW (solution.cpp:187-188): the nodal w from its equation with the nodal (u, q), after every
// change of udg in the Newton loop and at each time stage
void Solver::get_w(const double* mask) {
  if (ncw == 0) return;
  AsmArgs a = asm_args(0, d.ne);
  a.wmask = mask;
  w_nodes(d, a, xdg.p, odg.p, wdg.p, s);
}

void Solver::assemble() {
  if (precond < 0 || precond > 2)
    throw Error(5, "preconditioner " + std::to_string(precond) + " is not implemented");
  const long n = d.n, m = d.m, nd = d.nd, ncu = d.ncu, npe = d.npe, ndf = d.ndf;
  const long augs = n * augcols();
  const long blk = static_cast<long>(npe) * d.nge;
  if (precision) alloc_mixed();
  dinvf_fp32 = precision != 0;
  ilu_fp32 = precision != 0 && precond == 2;
  hipEvent_t t0{};
  // The mixed form assembles the element blocks (W, [D | F | Ru], B, K, G) in FP32 storage:
  // its operator, factors and u-recovery matrices are all FP32, and the Schur products and
  // the LU move half the bytes.  H is assembled in place in its FP32 copy H32 (the
  // preconditioner is built from H32); Rh stays FP64.
  const bool f32 = precision != 0;
  // operand of a GEMM: element offset `off` into the FP64 or the FP32 buffer
  auto opA = [&](Gemm& g, double* pd, float* pf, long off) { if (f32) g.Af = pf + off; else g.A = pd + off; };
  auto opB = [&](Gemm& g, double* pd, float* pf, long off) { if (f32) g.Bf = pf + off; else g.B = pd + off; };
  auto opC = [&](Gemm& g, double* pd, float* pf, long off) { if (f32) g.Cf = pf + off; else g.C = pd + off; };
  for (int e0 = 0; e0 < d.ne; e0 += neb) {
    const int eb = std::min(neb, d.ne - e0);
    AsmArgs a = asm_args(e0, eb);
    float* K32b = f32 ? K32.p + static_cast<long>(e0) * m * n : nullptr;  // K of this block, in place
    tic(T_ASSEMBLY, &t0);
    if (f32) {
      EXA_HIP(hipMemsetAsync(aug32.p, 0, sizeof(float) * augs * eb, s));
      EXA_HIP(hipMemsetAsync(K32b, 0, sizeof(float) * m * n * eb, s));
      EXA_HIP(hipMemsetAsync(Gfull32.p, 0, sizeof(float) * m * n * nd * eb, s));
    } else {
      EXA_HIP(hipMemsetAsync(aug.p, 0, sizeof(double) * augs * eb, s));
      EXA_HIP(hipMemsetAsync(Kt.p, 0, sizeof(double) * m * n * eb, s));
      EXA_HIP(hipMemsetAsync(Gfull.p, 0, sizeof(double) * m * n * nd * eb, s));
    }
    if (f32) EXA_HIP(hipMemsetAsync(H32.p + e0 * m * m, 0, sizeof(float) * m * m * eb, s));
    else EXA_HIP(hipMemsetAsync(H.p + e0 * m * m, 0, sizeof(double) * m * m * eb, s));
    // volume terms
    vol_points(d, a, true, s);
    if (f32) {
      vol_res(d, a, aug32.p + n * (n + m), augs, s);
      vol_W(d, a, W32.p, s);
    } else {
      vol_res(d, a, aug.p + n * (n + m), augs, s);
      vol_W(d, a, s);
    }
    {  // D block: A_kc = -W_kc * Phi for the u components
      Gemm g{};
      g.M = npe; g.N = npe; g.K = d.nge; g.alpha = -1.0; g.beta = 0.0;
      opA(g, W.p, W32.p, 0);
      g.lda = npe; g.sA0 = blk; g.sA1 = ncu * blk; g.sA2 = static_cast<long>(d.nc) * ncu * blk;
      opB(g, shapegt.p, shapegt32.p, 0);
      g.ldb = d.nge; g.sB0 = g.sB1 = g.sB2 = 0;
      opC(g, aug.p, aug32.p, 0);
      g.ldc = n; g.sC0 = npe; g.sC1 = n * npe; g.sC2 = augs;
      g.nb0 = static_cast<int>(ncu); g.nb1 = static_cast<int>(ncu); g.nb2 = eb;
      g.fp32 = precision;
      gemm(g, s);
    }
    for (int dd = 0; dd < nd; ++dd) {  // B_d blocks for the q components
      Gemm g{};
      g.M = npe; g.N = npe; g.K = d.nge; g.alpha = -1.0; g.beta = 0.0;
      opA(g, W.p, W32.p, (ncu + ncu * dd) * ncu * blk);
      g.lda = npe; g.sA0 = blk; g.sA1 = ncu * blk; g.sA2 = static_cast<long>(d.nc) * ncu * blk;
      opB(g, shapegt.p, shapegt32.p, 0);
      g.ldb = d.nge;
      opC(g, Bfull.p, Bfull32.p, n * npe * dd);
      g.ldc = n; g.sC0 = npe; g.sC1 = n * npe * nd; g.sC2 = n * n * nd;
      g.nb0 = static_cast<int>(ncu); g.nb1 = static_cast<int>(ncu); g.nb2 = eb;
      g.fp32 = precision;
      gemm(g, s);
    }
    // face terms; residual parts one local face at a time (no two writes of a launch collide)
    face_points(d, a, true, s);
    for (int l = 0; l < d.nfe; ++l) {
      if (f32) face_res(d, a, l, true, aug32.p + n * (n + m), augs, Rh.p, s);
      else face_res(d, a, l, true, aug.p + n * (n + m), augs, Rh.p, s);
    }
    if (f32) face_jac(d, a, aug32.p, Bfull32.p, K32b, Gfull32.p, H32.p, s);
    else face_jac(d, a, aug.p, Bfull.p, Kt.p, Gfull.p, H.p, s);
    // eliminate q: D += B*MinvC, F -= B*MinvE, K += G*MinvC, H -= G*MinvE (uequation.cpp:614-681)
    {
      Gemm g{};
      g.M = n; g.N = npe; g.K = npe * nd; g.alpha = 1.0; g.beta = 1.0;
      opA(g, Bfull.p, Bfull32.p, 0);
      g.lda = n; g.sA0 = n * npe * nd; g.sA1 = n * n * nd;
      opB(g, MCs.p, MCs32.p, e0 * npe * nd * npe);
      g.ldb = npe * nd; g.sB0 = 0; g.sB1 = npe * nd * npe;
      opC(g, aug.p, aug32.p, 0);
      g.ldc = n; g.sC0 = n * npe; g.sC1 = augs;
      g.nb0 = static_cast<int>(ncu); g.nb1 = eb; g.nb2 = 1;
      g.fp32 = precision;
      gemm(g, s);
      g.N = static_cast<int>(ndf); g.alpha = -1.0;
      opB(g, MEs.p, MEs32.p, e0 * npe * nd * ndf);
      g.sB1 = npe * nd * ndf;
      opC(g, aug.p, aug32.p, n * n);
      g.ldc = n * ncu; g.sC0 = n; g.sC1 = augs;
      gemm(g, s);
      Gemm k{};
      k.M = m; k.N = npe; k.K = npe * nd; k.alpha = 1.0; k.beta = 1.0;
      opA(k, Gfull.p, Gfull32.p, 0);
      k.lda = m; k.sA0 = m * npe * nd; k.sA1 = m * n * nd;
      opB(k, MCs.p, MCs32.p, e0 * npe * nd * npe);
      k.ldb = npe * nd; k.sB0 = 0; k.sB1 = npe * nd * npe;
      opC(k, Kt.p, K32b, 0);
      k.ldc = m; k.sC0 = m * npe; k.sC1 = m * n;
      k.nb0 = static_cast<int>(ncu); k.nb1 = eb; k.nb2 = 1;
      k.fp32 = precision;
      gemm(k, s);
      k.N = static_cast<int>(ndf); k.alpha = -1.0;
      opB(k, MEs.p, MEs32.p, e0 * npe * nd * ndf);
      k.sB1 = npe * nd * ndf;
      opC(k, H.p, H32.p, e0 * m * m);
      k.ldc = m * ncu; k.sC0 = m; k.sC1 = m * m;
      gemm(k, s);
    }
    toc(T_ASSEMBLY, t0);
    // condense u: [D | F | Ru] -> [LU | D^{-1}F | D^{-1}Ru] (the mixed form in FP32, with the
    // same blocked algorithm; its pivots are kept for the refinement)
    tic(T_LU, &t0);
    if (f32)
      lu_solve(aug32.p, static_cast<int>(n), static_cast<int>(augcols()), augs, eb, info.p, s,
               piv32.p + static_cast<long>(e0) * n, lu_ws(eb), lu_tri(eb));
    else
      lu_solve(aug.p, static_cast<int>(n), static_cast<int>(augcols()), augs, eb, info.p, s, nullptr, lu_ws(eb),
               lu_tri(eb));
    toc(T_LU, t0);
    tic(T_ASSEMBLY, &t0);
    {  // H += K * DinvF (= -K X_F), Rh -= K * (D^{-1} Ru)
      Gemm g{};
      g.M = m; g.N = m; g.K = n; g.alpha = -1.0; g.beta = 1.0;
      opA(g, Kt.p, K32b, 0);
      g.lda = m; g.sA0 = m * n;
      opB(g, aug.p, aug32.p, n * n);
      g.ldb = n; g.sB0 = augs;
      opC(g, H.p, H32.p, e0 * m * m);
      g.ldc = m; g.sC0 = m * m;
      g.nb0 = eb; g.nb1 = 1; g.nb2 = 1;
      g.fp32 = precision;
      gemm(g, s);
      g.N = 1;
      g.fp32 = 0;
      g.Cf = nullptr;
      opB(g, aug.p, aug32.p, n * (n + m));
      g.C = Rh.p + e0 * m; g.sC0 = m;
      gemm(g, s);
    }
    if (f32) {  // DinvF straight into FP32; FP32 copy of this block's LU factors of D'
      const long E0 = e0;
      post_schur(d, e0, eb, aug32.p, augs, DinvF32.p, RuS.p, s);
      to_float(LU32.p + E0 * n * n, aug32.p, static_cast<int>(n), static_cast<int>(n), n, augs, n * n, eb, s);
    } else {
      post_schur(d, e0, eb, aug.p, augs, DinvF.p, RuS.p, s);
    }
    toc(T_ASSEMBLY, t0);
    const int bad = first_nonzero(info.p, eb);
    if (bad >= 0) throw Error(4, "singular element matrix D at element " + std::to_string(e0 + bad));
  }
  // hdgAssembleRHS (matvec.cpp:103-117)
  face_gather(static_cast<long>(d.npf) * d.nf, d.ncu, src1.p, src2.p, Rh.p, b.p, s);
  tic(T_LU, &t0);
  if (precond == 1) {
    // hdgElementalAdditiveSchwarz (matvec.cpp:168-188), inverted in element chunks; the mixed
    // form assembles, inverts and stores the blocks in FP32 (its apply reads FP32 blocks)
    const long mm = static_cast<long>(d.m) * d.m;
    for (int e0 = 0; e0 < d.ne; e0 += eas_chunk) {
      const int eb = std::min(eas_chunk, d.ne - e0);
      if (precision) {
        eas_assemble(d, e0, eb, H32.p, f2e.p, elemcon.p, easaug32.p, s);
        lu_solve(easaug32.p, d.m, 2 * d.m, 2 * mm, eb, info.p, s, nullptr, lu_ws(eb), lu_tri(eb));
      } else {
        eas_assemble(d, e0, eb, H.p, f2e.p, elemcon.p, easaug.p, s);
        lu_solve(easaug.p, d.m, 2 * d.m, 2 * mm, eb, info.p, s, nullptr, lu_ws(eb), lu_tri(eb));
      }
      const int bad = first_nonzero(info.p, eb);
      if (bad >= 0) throw Error(4, "singular Schwarz block at element " + std::to_string(e0 + bad));
      if (precision) eas_extract(d, eb, easaug32.p, BEeas32.p + e0 * mm, s);
      else eas_extract(d, eb, easaug.p, BEeas.p + e0 * mm, s);
    }
    if (precision && precond_bf16) to_bf16(BEeas16.p, BEeas32.p, mm * d.ne, s);
  } else if (precond == 0) {
    // hdgBlockJacobi (matvec.cpp:190-217)
    if (precision) bj_assemble(d, f2e.p, elemcon.p, H32.p, bjaug.p, s);
    else bj_assemble(d, f2e.p, elemcon.p, H.p, bjaug.p, s);
    lu_solve(bjaug.p, d.ncf, 2 * d.ncf, 2L * d.ncf * d.ncf, d.nf, info.p, s, nullptr, lu_ws(d.nf));
    bj_extract(d, bjaug.p, BE.p, s);
  } else {
    // hdgBlockILU0 (matvec.cpp:119-166)
    // the mixed form assembles and factors in FP32 (its factors are applied in FP32 or BF16)
    if (precision) {
      bilu_assemble(d, ilud, f2e.p, elemcon.p, H32.p, BEilu32.p, s);
      factor_bilu0_t(BEilu32.p, Lilu32.p);
    } else {
      bilu_assemble(d, ilud, f2e.p, elemcon.p, H.p, BEilu.p, s);
      factor_bilu0();
    }
  }
  toc(T_LU, t0);
  if (precision) {
    const long c2 = static_cast<long>(d.ncf) * d.ncf;
    if (precond == 0) to_float(BE32.p, BE.p, d.ncf, d.ncf, d.ncf, c2, c2, d.nf, s);
    if (precond == 2) {
      if (precond_bf16) to_bf16(BEilu16.p, BEilu32.p, c2 * ilu.nnz * ilu.nse, s);
    }
  }
  if (precond != 1) {
    const int bad = first_nonzero(info.p, precond == 2 ? ilu.nse : d.nf);
    if (bad >= 0)
      throw Error(4, std::string("singular preconditioner block at ") +
                         (precond == 2 ? "superelement " : "face ") + std::to_string(bad));
  }
  assembled = true;
}

void Solver::residual(double* nb, double* nru) {
  residual_fields();
  double a1 = norm(bres.p, d.N);
  double a2 = norm(RuR.p, static_cast<long>(d.ne) * d.n);
  if (nb) *nb = a1;
  if (nru) *nru = a2;
}

void Solver::residual_fields() {
  const long n = d.n;
  for (int e0 = 0; e0 < d.ne; e0 += neb) {
    const int eb = std::min(neb, d.ne - e0);
    AsmArgs a = asm_args(e0, eb);
    vol_points(d, a, false, s);
    vol_res(d, a, RuR.p + e0 * n, n, s);
    face_points(d, a, false, s);
    for (int l = 0; l < d.nfe; ++l) face_res(d, a, l, false, RuR.p + e0 * n, n, Rh.p, s);
  }
  face_gather(static_cast<long>(d.npf) * d.nf, d.ncu, src1.p, src2.p, Rh.p, bres.p, s);
}

void Solver::matvec(const double* v, double* w) {
  hipEvent_t t0{};
  tic(T_MATVEC, &t0);
  if (precision) matvec_elem(d, H32.p, elemcon.p, v, we.p, s, einst.p, elem_active);
  else matvec_elem(d, H.p, elemcon.p, v, we.p, s, einst.p, elem_active);
  toc(T_MATVEC, t0);
  tic(T_OTHER, &t0);
  face_gather(static_cast<long>(d.npf) * d.nf, d.ncu, src1.p, src2.p, we.p, w, s);
  toc(T_OTHER, t0);
}

void Solver::apply_precond(double* xv) {
  copy(tmp.p, xv, d.N, s);
  apply_precond_to(tmp.p, xv);
}

// out = P in (in and out distinct)
void Solver::apply_precond_to(const double* in, double* out) {
  hipEvent_t t0{};
  tic(T_PRECOND, &t0);
  if (precond == 2) {
    if (precision && precond_bf16) bilu_apply(d, ilud, BEilu16.p, in, Yilu.p, Xilu.p, out, s, gm_act);
    else if (precision) bilu_apply(d, ilud, BEilu32.p, in, Yilu.p, Xilu.p, out, s, gm_act);
    else bilu_apply(d, ilud, BEilu.p, in, Yilu.p, Xilu.p, out, s, gm_act);
  } else if (precond == 1) {  // x = sum_e scatter(BE_e^{-1} gather_e(x)), as hdgMatVec with BE
    if (precision && precond_bf16) matvec_elem(d, BEeas16.p, elemcon.p, in, we.p, s, einst.p, elem_active);
    else if (precision) matvec_elem(d, BEeas32.p, elemcon.p, in, we.p, s, einst.p, elem_active);
    else matvec_elem(d, BEeas.p, elemcon.p, in, we.p, s, einst.p, elem_active);
    face_gather(static_cast<long>(d.npf) * d.nf, d.ncu, src1.p, src2.p, we.p, out, s);
  } else {
    if (precision) bj_apply(d, BE32.p, in, out, s, gm_act);
    else bj_apply(d, BE.p, in, out, s, gm_act);
  }
  toc(T_PRECOND, t0);
}

// Steady problems: one Newton solve.  Time-dependent problems: the t2c time loop
// (exasim.hpp:412-470) with DIRK stages -- PreviousSolutions, UpdateSourceDIRK,
// a Newton solve per stage, UpdateSolutionDIRK (previoussolutions.cpp:23-58,
// updatesource.cpp:32-72, updatesolution.cpp:48-70).
void Solver::solve(int* newton_iters, int* gmres_total, double* final_norm) {
  if (batch) {
    inst_newton.assign(ninst, 0);
    inst_gmres.assign(ninst, 0);
    inst_norm.assign(ninst, 0.0);
    inst_failed.assign(ninst, 0);
    step_act.clear();  // a full solve ends an exa_batch_step sequence
    step_valid = false;
  }
  if (!tdep) {
    if (batch) newton_seg(newton_iters, gmres_total, final_norm);
    else newton(newton_iters, gmres_total, final_norm);
    return;
  }
  if (batch)
    for (int n : inst_nsteps)
      if (n != static_cast<int>(dts.size()))
        throw Error(1, "the instances of this batch take different numbers of time steps: exa_batch_step "
                       "runs them, exa_solve needs one list of steps");
  if (tscheme == 1) {
    solve_bdf(newton_iters, gmres_total, final_norm);
    return;
  }
  int nit = 0, gtot = 0;
  double nrm = 0.0;
  double t = tnow;
  for (size_t istep = 0; istep < dts.size(); ++istep) {
    const double dt = dts[istep];
    dirk_step(dt, t, nullptr, &nit, &gtot, &nrm, static_cast<int>(istep));
    t += dt;
  }
  tnow = t;
  EXA_HIP(hipStreamSynchronize(s));
  if (newton_iters) *newton_iters = nit;
  if (gmres_total) *gmres_total = gtot;
  if (final_norm) *final_norm = nrm;
}

// One DIRK time step of length dt from time t (the stages of the reference's time loop,
// solution.cpp:244-330, with PreviousSolutions, UpdateSource and UpdateSolution): per stage j the
// model time t + dt t_j, dtfactor = d_jj / dt, the stage source in sdg and a Newton solve.  With
// only (a batch) the Newton solves iterate the instances it marks; with use_tinst each instance
// runs at its own time inst_t[i] + dt t_j.  Adds the Newton and GMRES iterations to *nit, *gtot.
void Solver::dirk_step(double dt, double t, const std::vector<char>* only, int* nit, int* gtot, double* nrm,
                       int istep) {
  const int S = tstages;
  double csum = 0.0;
  for (int j = 0; j < S; ++j) csum += dirk_c[j];
  u_axpby(d, udg.p, 1.0, 0.0, udgprev.p, s);          // udgprev = u
  axpby(utmp.p, udgprev.p, 1.0 - csum, 0.0, static_cast<long>(udgprev.n), s);
  const long nw = static_cast<long>(wdg.n);
  if (ncw > 0) {  // PreviousSolutions (previoussolutions.cpp:55-60): wprev = w, wtmp = (1 - sum c) w
    copy(wprev.p, wdg.p, nw, s);
    axpby(wtmp.p, wdg.p, 1.0 - csum, 0.0, nw, s);
  }
  for (int j = 0; j < S; ++j) {
    tnow = t + dt * dirk_t[j];
    if (use_tinst) {
      std::vector<double> ti(ninst);
      for (int i = 0; i < ninst; ++i) ti[i] = inst_t[i] + dt * dirk_t[j];
      st.h2d(tinst_d.p, ti.data(), sizeof(double) * ninst, s);
    }
    dtfactor = dirk_d[j * S + j] / dt;
    if (j == 0) {
      u_axpby(d, udg.p, dirk_d[0] / dt, 0.0, sdg.p, s);
    } else if (j == 1) {
      u_axpby(d, udg.p, -dirk_d[1] / dt, (dirk_d[1] + dirk_d[S + 1]) / dirk_d[0], sdg.p, s);
    } else {
      const double sc = dirk_d[2] / dirk_d[1];
      const double pc = dirk_d[2] + dirk_d[S + 2] + dirk_d[2 * S + 2] - sc * (dirk_d[1] + dirk_d[S + 1]);
      u_axpby(d, udg.p, -dirk_d[S + 2] / dt, sc, sdg.p, s);
      axpby(sdg.p, udgprev.p, pc / dt, 1.0, static_cast<long>(sdg.n), s);
    }
    if (ncw > 0) {  // the w equation's stage source (updatesource.cpp:78-100), as for u
      if (j == 0) {
        axpby(wsrc.p, wdg.p, dirk_d[0] / dt, 0.0, nw, s);
      } else if (j == 1) {
        axpby(wsrc.p, wdg.p, -dirk_d[1] / dt, (dirk_d[1] + dirk_d[S + 1]) / dirk_d[0], nw, s);
      } else {
        const double sc = dirk_d[2] / dirk_d[1];
        const double pc = dirk_d[2] + dirk_d[S + 2] + dirk_d[2 * S + 2] - sc * (dirk_d[1] + dirk_d[S + 1]);
        axpby(wsrc.p, wdg.p, -dirk_d[S + 2] / dt, sc, nw, s);
        axpby(wsrc.p, wprev.p, pc / dt, 1.0, nw, s);
      }
    }
    int ni = 0, gi = 0;
    if (batch) newton_seg(&ni, &gi, nrm, only);
    else newton(&ni, &gi, nrm);
    *nit += ni;
    *gtot += gi;
    u_axpby(d, udg.p, dirk_c[j], 1.0, utmp.p, s);       // utmp += c_j u
    if (j == S - 1) u_insert(d, utmp.p, udg.p, s);
    if (ncw > 0) {  // UpdateSolution (updatesolution.cpp:142-172): GetW, wtmp += c_j w
      get_w();
      axpby(wtmp.p, wdg.p, dirk_c[j], 1.0, nw, s);
      if (j == S - 1) copy(wdg.p, wtmp.p, nw, s);
    }
    if (verbose) {
      std::printf("Timestep %d, stage %d, time %g: Newton %d, GMRES %d, residual %g\n", istep + 1, j + 1, tnow, ni,
                  gi, *nrm);
      std::fflush(stdout);
    }
  }
}

// BDF1-3 (PreviousSolutions, previoussolutions.cpp:62-100; UpdateSourceBDF, updatesource.cpp:104-140;
// BDFcoeff, timestepcoeff.cpp:106-130): per step the history takes u^n, dtfactor = c0/dt and
// sdg = -sum_k (c_k/dt) u^{n+1-k}, and one Newton solve at t + dt gives u^{n+1}.  The first steps
// use the highest order the stored history allows (the reference reads history it has not
// written in its first torder - 1 steps).
void Solver::solve_bdf(int* newton_iters, int* gmres_total, double* final_norm) {
  static const double C[3][4] = {{1.0, -1.0, 0, 0}, {1.5, -2.0, 0.5, 0}, {11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0}};
  const long nu = static_cast<long>(d.npe) * d.ncu * d.ne;
  int nit = 0, gtot = 0, nhist = 0;
  double nrm = 0.0, t = tnow;
  for (size_t istep = 0; istep < dts.size(); ++istep) {
    const double dt = dts[istep];
    // history: slot k holds u^{n-k}; shift by one and store u^n in slot 0
    for (int k = std::min(nhist, torder - 1); k >= 1; --k) copy(bdfhist.p + k * nu, bdfhist.p + (k - 1) * nu, nu, s);
    u_axpby(d, udg.p, 1.0, 0.0, bdfhist.p, s);
    nhist = std::min(nhist + 1, torder);
    const int q = std::min(torder, nhist);
    const double* c = C[q - 1];
    dtfactor = c[0] / dt;
    axpby(sdg.p, bdfhist.p, -c[1] / dt, 0.0, nu, s);
    for (int k = 1; k < q; ++k) axpby(sdg.p, bdfhist.p + k * nu, -c[k + 1] / dt, 1.0, nu, s);
    if (ncw > 0) {  // the w equation's history source, as for u (the reference's UpdateSourceBDF
                    // leaves wsrc unchanged for non-wave problems)
      const long nw = static_cast<long>(wdg.n);
      for (int k = std::min(nhist - 1, torder - 1); k >= 1; --k)
        copy(wbdfhist.p + k * nw, wbdfhist.p + (k - 1) * nw, nw, s);
      copy(wbdfhist.p, wdg.p, nw, s);
      axpby(wsrc.p, wbdfhist.p, -c[1] / dt, 0.0, nw, s);
      for (int k = 1; k < q; ++k) axpby(wsrc.p, wbdfhist.p + k * nw, -c[k + 1] / dt, 1.0, nw, s);
    }
    tnow = t + dt;
    int ni = 0, gi = 0;
    if (batch) newton_seg(&ni, &gi, &nrm);
    else newton(&ni, &gi, &nrm);
    nit += ni;
    gtot += gi;
    if (verbose) {
      std::printf("Timestep %zu (BDF%d), time %g: Newton %d, GMRES %d, residual %g\n", istep + 1, q, tnow, ni, gi,
                  nrm);
      std::fflush(stdout);
    }
    t += dt;
  }
  tnow = t;
  EXA_HIP(hipStreamSynchronize(s));
  if (newton_iters) *newton_iters = nit;
  if (gmres_total) *gmres_total = gtot;
  if (final_norm) *final_norm = nrm;
}

void Solver::newton(int* newton_iters, int* gmres_total, double* final_norm) {
  const long N = d.N, nU = static_cast<long>(d.n) * d.ne;
  get_q();
  double nb = 0, nru = 0;
  residual(&nb, &nru);
  double nrmr = nb + nru;
  if (verbose) std::printf("Newton Iteration: 0,  Residual Norm: %g\n", nrmr);
  int it = 0, gtot = 0;
  DevBuf<double> u;
  u.alloc(N);
  copy(u.p, uh.p, N, s);
  for (it = 0; it < newton_maxit; ++it) {
    assemble();
    // mixed precision: the FP32 operator limits the first solve to about 1e-7, so it is
    // solved to refine_first_tol and the refinement loop below supplies the rest
    int gi = precision ? gmres(b.p, x.p, std::max(gmres_tol, refine_first_tol), refine_maxit_gmres())
                       : gmres(b.p, x.p);
    gtot += gi;
    double alpha = 1.0;
    axpby(u.p, x.p, 1.0, 1.0, N, s);  // u += x
    copy(uh.p, u.p, N, s);
    if (precision) dudg(d, RuS.p, DinvF32.p, elemcon.p, x.p, du.p, s);
    else dudg(d, RuS.p, DinvF.p, elemcon.p, x.p, du.p, s);
    update_udg(d, du.p, alpha, udg.p, s);
    get_q();
    get_w();
    double nrm0 = nrmr;
    residual(&nb, &nru);
    nrmr = nb + nru;
    if (nrmr > nrm0 && nrmr > 1e6) throw Error(4, "Residual norm increases more than 1e6");
    // the reference's rule (solution.cpp:213-227), with a floor on alpha for a residual that
    // stays NaN, which would otherwise halve alpha forever
    while ((nrmr > nrm0 && alpha > 0.1) || (std::isnan(nrmr) && alpha > 1e-6)) {
      if (verbose) std::printf("Newton Iteration: %d, Alpha: %g, Original Norm: %g,  Updated Norm: %g\n",
                               it + 1, alpha, nrm0, nrmr);
      alpha /= 2.0;
      axpby(u.p, x.p, -alpha, 1.0, N, s);
      copy(uh.p, u.p, N, s);
      update_udg(d, du.p, -alpha, udg.p, s);
      get_q();
      get_w();
      residual(&nb, &nru);
      nrmr = nb + nru;
    }
    if (std::isnan(nrmr)) throw Error(4, "NaN residual");
    if (verbose)
      std::printf("Newton Iteration: %d, Alpha: %g, Original Norm: %g,  Updated Norm: %g  (GMRES %d)\n",
                  it + 1, alpha, nrm0, nrmr, gi);
    // mixed precision: iterative refinement with the FP64 residual and the stored FP32
    // factors, until the residual is refine_rtol times this iteration's first residual
    last_refinements = 0;
    while (precision && nrmr > refine_rtol * nrm0 && last_refinements < refine_maxit) {
      refine_rhs(d, LU32.p, piv32.p, K32.p, RuR.p, RuS.p, Rh.p, s);
      face_gather(static_cast<long>(d.npf) * d.nf, d.ncu, src1.p, src2.p, Rh.p, b.p, s);
      gi = gmres(b.p, x.p, refine_tol, refine_maxit_gmres());
      gtot += gi;
      axpby(u.p, x.p, 1.0, 1.0, N, s);
      copy(uh.p, u.p, N, s);
      dudg(d, RuS.p, DinvF32.p, elemcon.p, x.p, du.p, s);
      update_udg(d, du.p, 1.0, udg.p, s);
      get_q();
      get_w();
      residual(&nb, &nru);
      nrmr = nb + nru;
      ++last_refinements;
      if (std::isnan(nrmr)) throw Error(4, "NaN residual in refinement");
      if (verbose) {
        std::printf("  refinement %d: residual norm %g (GMRES %d)\n", last_refinements, nrmr, gi);
        std::fflush(stdout);
      }
    }
    if (nrmr < newton_tol) {
      ++it;
      break;
    }
  }
  (void)nU;
  EXA_HIP(hipStreamSynchronize(s));
  if (newton_iters) *newton_iters = it;
  if (gmres_total) *gmres_total = gtot;
  if (final_norm) *final_norm = nrmr;
}

void Solver::write_output() {
  if (dataout.empty()) throw Error(1, "no dataout prefix was given");
  std::vector<double> hu(static_cast<size_t>(d.npe) * d.nc * d.ne), hh(d.N);
  st.d2h(hu.data(), udg.p, hu.size() * sizeof(double), s);
  st.d2h(hh.data(), uh.p, hh.size() * sizeof(double), s);
  ::exa::write_output(dataout + "udg_np0.bin", d.npe, d.nc, d.ne, hu);
  ::exa::write_output(dataout + "uhat_np0.bin", d.ncu, d.npf, d.nf, hh);
}

long Solver::get_array(const std::string& name, double* dst, long len) {
  const double* src = nullptr;
  long size = 0;
  std::vector<double> host;
  auto from_int = [&](const std::vector<int>& v) {
    host.assign(v.begin(), v.end());
    size = static_cast<long>(host.size());
  };
  if (name == "udg") { src = udg.p; size = static_cast<long>(udg.n); }
  else if (name == "uh") { src = uh.p; size = static_cast<long>(uh.n); }
  else if (name == "H" && dinvf_fp32) {  // the mixed form assembles H in FP32 only
    std::vector<float> f(H32.n);
    EXA_HIP(hipMemcpyAsync(f.data(), H32.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost, s));
    EXA_HIP(hipStreamSynchronize(s));
    host.assign(f.begin(), f.end());
    size = static_cast<long>(host.size());
  } else if (name == "H") { src = H.p; size = static_cast<long>(H.n); }
  else if (name == "DinvF" && dinvf_fp32) {  // the mixed form keeps D^{-1}F in FP32 only
    std::vector<float> f(DinvF32.n);
    EXA_HIP(hipMemcpyAsync(f.data(), DinvF32.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost, s));
    EXA_HIP(hipStreamSynchronize(s));
    host.assign(f.begin(), f.end());
    size = static_cast<long>(host.size());
  } else if (name == "DinvF") { src = DinvF.p; size = static_cast<long>(DinvF.n); }
  else if (name == "Ru") { src = RuS.p; size = static_cast<long>(RuS.n); }
  else if (name == "Rh") { src = Rh.p; size = static_cast<long>(Rh.n); }
  else if (name == "b") { src = b.p; size = static_cast<long>(b.n); }
  else if (name == "BE") { src = BE.p; size = static_cast<long>(BE.n); }
  else if (name == "Mass") { src = Mass.p; size = static_cast<long>(Mass.n); }
  else if (name == "MinvC") { src = MCs.p; size = static_cast<long>(MCs.n); }
  else if (name == "MinvE") { src = MEs.p; size = static_cast<long>(MEs.n); }
  else if (name == "res_b") { src = bres.p; size = static_cast<long>(bres.n); }
  else if (name == "BEilu" && ilu_fp32) {  // the mixed form keeps its block-ILU(0) factors in FP32
    std::vector<float> f(BEilu32.n);
    EXA_HIP(hipMemcpyAsync(f.data(), BEilu32.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost, s));
    EXA_HIP(hipStreamSynchronize(s));
    host.assign(f.begin(), f.end());
    size = static_cast<long>(host.size());
  } else if (name == "BEilu") { src = BEilu.p; size = static_cast<long>(BEilu.n); }
  else if (name == "BEeas" && dinvf_fp32) {  // the mixed form inverts its Schwarz blocks in FP32
    std::vector<float> f(BEeas32.n);
    EXA_HIP(hipMemcpyAsync(f.data(), BEeas32.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost, s));
    EXA_HIP(hipStreamSynchronize(s));
    host.assign(f.begin(), f.end());
    size = static_cast<long>(host.size());
  } else if (name == "BEeas") { src = BEeas.p; size = static_cast<long>(BEeas.n); }
  else if (name == "ilu_face") from_int(ilu.face);
  else if (name == "ilu_row_ptr") from_int(ilu.row_ptr);
  else if (name == "ilu_col_ind") from_int(ilu.col_ind);
  else if (name == "res_Ru") { src = RuR.p; size = static_cast<long>(RuR.n); }
  else if (name == "x") { src = x.p; size = static_cast<long>(x.n); }
  else if (name == "w") { src = wv.p; size = static_cast<long>(wv.n); }
  else if (name == "egeo") { src = egeo.p; size = static_cast<long>(egeo.n); }
  else if (name == "wdg") { src = wdg.p; size = static_cast<long>(wdg.n); }
  else if (name == "fgeo") { src = fgeo.p; size = static_cast<long>(fgeo.n); }
  else if (name == "elemcon") from_int(conn.elemcon);
  else if (name == "f2e") from_int(conn.f2e);
  else throw Error(1, "unknown array '" + name + "'");
  if (!dst) return size;
  long k = std::min(len, size);
  if (src) {
    st.d2h(dst, src, sizeof(double) * k, s);
  } else {
    std::memcpy(dst, host.data(), sizeof(double) * k);
  }
  return size;
}

void Solver::set_array(const std::string& name, const double* src, long len) {
  DevBuf<double>* t = nullptr;
  if (name == "udg") t = &udg;
  else if (name == "uh") t = &uh;
  else if (name == "x") t = &x;
  else if (name == "w") t = &wv;
  else throw Error(1, "array '" + name + "' cannot be set");
  if (len != static_cast<long>(t->n)) throw Error(1, "wrong length for '" + name + "'");
  st.h2d(t->p, src, sizeof(double) * len, s);
}

void Solver::device_ptr(const std::string& name, void** p, long* len) {
  DevBuf<double>* t = nullptr;
  if (name == "H") t = &H;
  else if (name == "DinvF") t = &DinvF;
  else if (name == "uh") t = &uh;
  else if (name == "udg") t = &udg;
  else if (name == "b") t = &b;
  else if (name == "x") t = &x;
  else if (name == "w") t = &wv;
  else throw Error(1, "no device pointer for '" + name + "'");
  *p = t->p;
  if (len) *len = static_cast<long>(t->n);
}

}  // namespace exa
