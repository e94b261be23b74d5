"""numpy restatement of the reference's HDG Newton-GMRES path (test infrastructure only).

Only tests, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
use this module, and only as the checker / the timed CPU baseline.

Conventions.  The reference stores every array column-major with the first
index fastest.  Here arrays are numpy arrays in *natural* order (element index
first); element matrices are ``A[e, row, col]`` with the reference's row and
column numbering:

* u-DOF of an element:      r = i + npe*c           (node i, component c)
* trace DOF of an element:  t = c + ncu*(j + npf*l) (component c, face node j, local face l)

These are the layouts of backend/Discretization/uequation.cpp:559-775 after the
``schurMatrix*`` permutations (backend/Common/kokkosimpl.h:1797-2020).

Parity status of each function is stated in DESIGN.md; every function here is
pinned by a test in tests/test_oracle*.py (closed forms, invariants, dense
brute force, reference fixtures).
"""
import numpy as np

from .models import make_model


# --------------------------------------------------------------------------
# datain readers -- backend/Discretization/readbinaryfiles.cpp
# --------------------------------------------------------------------------

def _segments(a, offset, sizes):
    out = []
    for s in sizes:
        out.append(a[offset:offset + s])
        offset += s
    return out, offset


def read_app(path):
    """readbinaryfiles.cpp:65-173 (``readappstruct``)."""
    a = np.fromfile(path, dtype=np.float64)
    lsize = int(a[0])
    nsize = a[1:1 + lsize].astype(int)
    names = ["ndims", "flag", "problem", "uinf", "dt", "factor", "physicsparam",
             "solversparam", "tau", "stgdata", "stgparam", "stgib", "vindx", "dae_dt",
             "interfacefluxmap", "avparam"]
    segs, _ = _segments(a, 1 + lsize, nsize[:len(names)])
    return dict(zip(names, segs))


def read_master(path):
    """readbinaryfiles.cpp:208-277 (``readmasterstruct``) -> natural-order arrays."""
    a = np.fromfile(path, dtype=np.float64)
    lsize = int(a[0])
    nsize = a[1:1 + lsize].astype(int)
    names = ["ndims", "shapegt", "shapegw", "shapfgt", "shapfgw", "shapent", "shapen",
             "shapfnt", "shapfn", "xpe", "gpe", "gwe", "xpf", "gpf", "gwf", "shap1dgt",
             "shap1dgw", "shap1dnt", "shap1dnl", "xp1d", "gp1d", "gw1d"]
    segs, _ = _segments(a, 1 + lsize, nsize[:len(names)])
    raw = dict(zip(names, segs))
    nd_ = raw["ndims"].astype(int)
    nd, elemtype, porder, npe, npf, nge, ngf = nd_[0], nd_[1], nd_[3], nd_[5], nd_[6], nd_[7], nd_[8]
    m = dict(nd=nd, elemtype=elemtype, porder=porder, npe=npe, npf=npf, nge=nge, ngf=ngf)
    # shapegt[g + nge*j + nge*npe*d] -> phi_e[d, g, j]
    m["phi_e"] = raw["shapegt"].reshape(nd + 1, npe, nge).transpose(0, 2, 1).copy()
    m["phiw_e"] = raw["shapegw"].reshape(nd + 1, nge, npe).copy()       # [d, g, j]
    m["phi_f"] = raw["shapfgt"].reshape(nd, npf, ngf).transpose(0, 2, 1).copy()
    m["phiw_f"] = raw["shapfgw"].reshape(nd, ngf, npf).copy()
    m["xpe"] = raw["xpe"].reshape(nd, npe).T.copy()
    m["xpf"] = raw["xpf"].reshape(max(nd - 1, 1), npf).T.copy()
    m["gwe"] = raw["gwe"].copy()
    m["gwf"] = raw["gwf"].copy()
    return m


def read_mesh(path):
    """readbinaryfiles.cpp:321-411 (``readmeshstruct``): the raw arrays from
    which buildConn derives the connectivity."""
    a = np.fromfile(path, dtype=np.float64)
    lsize = int(a[0])
    nsize = a[1:1 + lsize].astype(int)
    segs, _ = _segments(a, 1 + lsize, nsize[:29])
    ndims = segs[0].astype(int)
    ne, nfe = ndims[1], ndims[4]
    ti = segs[26].astype(np.int64)
    nve = ti.size // ne
    return dict(nd=ndims[0], ne=ne, nfe=nfe, nve=nve,
                elempart=segs[9].astype(np.int64),
                perm=segs[23].astype(np.int64),
                bf=segs[24].astype(np.int64).reshape(ne, nfe),
                cartgridpart=segs[25].astype(np.int64),
                t=ti.reshape(ne, nve),
                boundaryConditions=segs[27].astype(np.int64))


def read_sol(path, npe, nc, ncu, ne):
    """readbinaryfiles.cpp:499-622 (``readsolstruct``)."""
    a = np.fromfile(path, dtype=np.float64)
    lsize = int(a[0])
    nsize = a[1:1 + lsize].astype(int)
    segs, _ = _segments(a, 1 + lsize, nsize[:7])
    out = {}
    x = segs[1]
    ncx = x.size // (npe * ne)
    out["xdg"] = x.reshape(ne, ncx, npe).transpose(0, 2, 1).copy()
    u = segs[2]
    if u.size:
        k = u.size // (npe * ne)
        u = u.reshape(ne, k, npe).transpose(0, 2, 1)
        udg = np.zeros((ne, npe, nc))
        udg[:, :, :k] = u
        out["udg"] = udg
    else:
        out["udg"] = None
    o = segs[3]
    out["odg"] = o.reshape(ne, o.size // (npe * ne), npe).transpose(0, 2, 1).copy() if o.size else None
    w = segs[4]
    out["wdg"] = w.reshape(ne, w.size // (npe * ne), npe).transpose(0, 2, 1).copy() if w.size else None
    out["uh"] = segs[5].copy() if segs[5].size else None
    return out


# --------------------------------------------------------------------------
# connectivity -- backend/Discretization/connectivity.cpp
# --------------------------------------------------------------------------

LOCAL_FACES = {
    (2, 1): [[0, 1], [1, 2], [2, 3], [3, 0]],
    (3, 1): [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [2, 3, 7, 6], [1, 2, 6, 5], [3, 0, 4, 7]],
    (2, 0): [[1, 2], [2, 0], [0, 1]],
    (3, 0): [[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]],
}


def permindex(xpf, nd, elemtype):
    """connectivity.cpp:1128-1205: face-node permutations for the matching
    orientations of a shared face (each is an involution)."""
    npf = xpf.shape[0]
    if nd == 2:
        return np.arange(npf)[::-1].reshape(npf, 1)
    s, t = xpf[:, 0], xpf[:, 1]
    if elemtype == 1:
        maps = [np.stack([t, s], 1), np.stack([s, 1 - t], 1),
                np.stack([1 - t, 1 - s], 1), np.stack([1 - s, t], 1)]
    else:
        maps = [np.stack([t, s], 1), np.stack([1 - s - t, t], 1), np.stack([s, 1 - s - t], 1)]
    out = np.zeros((npf, len(maps)), dtype=np.int64)
    for k, B in enumerate(maps):
        for i in range(npf):
            j = np.where(np.abs(B - xpf[i][None, :]).max(axis=1) < 1e-8)[0]
            out[i, k] = j[0]
    return out


def build_conn(master, mesh):
    """buildConn (connectivity.cpp:1212-1385) for one process.

    Returns f2e[nf, 4] (e1, l1, e2, l2), elemcon[ne, nfe, npf] (global face-node
    index f*npf + j), facecon, the number of interior faces, and the element
    blocks.  Interior faces are numbered in the order ``mkf2e_hash`` discovers
    them (second visit while scanning elements then local faces); boundary faces
    follow, stably sorted by boundary-condition number (connectivity.cpp:426-497).
    Within one condition this numbering takes boundary faces in order of first
    visit; the reference takes them in std::unordered_map iteration order.
    """
    nd, elemtype = master["nd"], master["elemtype"]
    npe, npf = master["npe"], master["npf"]
    t = mesh["t"]
    ne, nfe = mesh["ne"], mesh["nfe"]
    faces = LOCAL_FACES[(nd, elemtype)]
    perm = mesh["perm"].reshape(nfe, npf)
    seen = {}
    interior = []
    for e in range(ne):
        for l in range(nfe):
            key = tuple(sorted(t[e, faces[l]]))
            if key in seen:
                e0, l0 = seen.pop(key)
                interior.append((e0, l0, e, l))
            else:
                seen[key] = (e, l)
    boundary = [(e0, l0, -1, -1) for (e0, l0) in seen.values()]
    bfa = [mesh["bf"][e0, l0] for (e0, l0, _, _) in boundary]
    order = sorted(range(len(boundary)), key=lambda i: bfa[i])   # stable
    f2e = np.array(interior + [boundary[i] for i in order], dtype=np.int64).reshape(-1, 4)
    nf = f2e.shape[0]
    # fix_f2t_consistency (connectivity.cpp:301-335)
    ep = mesh["elempart"]
    for f in range(len(interior)):
        if ep[f2e[f, 2]] < ep[f2e[f, 0]]:
            f2e[f] = f2e[f, [2, 3, 0, 1]]
    permind = permindex(master["xpf"], nd, elemtype)
    elemcon = -np.ones((ne, nfe, npf), dtype=np.int64)
    facecon = -np.ones((nf, npf, 2), dtype=np.int64)
    for f in range(nf):
        e1, l1, e2, l2 = f2e[f]
        facecon[f, :, 0] = e1 * npe + perm[l1]
        elemcon[e1, l1, :] = f * npf + np.arange(npf)
        if e2 < 0:
            facecon[f, :, 1] = facecon[f, :, 0]
            continue
        if nd <= 2:
            pk = permind[:, 0]
        else:
            f1 = t[e1, faces[l1]]
            f2 = t[e2, faces[l2]]
            if elemtype == 0:
                conds = [(0, 0, 1, 2, 2, 1), (0, 1, 1, 0, 2, 2), (0, 2, 1, 1, 2, 0)]
                k = next(i for i, c in enumerate(conds)
                         if f1[c[0]] == f2[c[1]] and f1[c[2]] == f2[c[3]] and f1[c[4]] == f2[c[5]])
            else:
                conds = [(0, 3, 2, 1), (3, 2, 1, 0), (2, 1, 0, 3), (1, 0, 3, 2)]
                k = next(i for i, c in enumerate(conds)
                         if all(f1[a] == f2[c[a]] for a in range(4)))
            pk = permind[:, k]
        facecon[f, :, 1] = e2 * npe + perm[l2][pk]
        elemcon[e2, l2, :] = f * npf + pk
    nebmax = 4096 if nd <= 2 else 1024
    nblk = (ne + nebmax - 1) // nebmax
    base, rem = divmod(ne, nblk)
    eblks, s = [], 0
    for b in range(nblk):
        n = base + (1 if b < rem else 0)
        eblks.append((s, s + n))
        s += n
    return dict(f2e=f2e, elemcon=elemcon, facecon=facecon, nf=nf, nif=len(interior),
                perm=perm, eblks=eblks, bf=mesh["bf"])


# --------------------------------------------------------------------------
# geometry -- backend/Discretization/geometry.cpp, kokkosimpl.h:2321-2416
# --------------------------------------------------------------------------

def _adjugate(J):
    """Xx[..., j, m] = jac * d xi_m / d x_j for Jmat[..., a, b] = d x_b / d xi_a."""
    nd = J.shape[-1]
    if nd == 2:
        jac = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
        Xx = np.empty(J.shape)
        Xx[..., 0, 0] = J[..., 1, 1]
        Xx[..., 1, 0] = -J[..., 1, 0]
        Xx[..., 0, 1] = -J[..., 0, 1]
        Xx[..., 1, 1] = J[..., 0, 0]
        return jac, Xx
    jac = np.linalg.det(J)
    inv = np.linalg.inv(J)             # inv[b, a] = d xi_a / d x_b
    return jac, jac[..., None, None] * inv


def elem_geom(master, xdg):
    """ElemGeomBlock (geometry.cpp:44-85): xg[e,g,d], Xx[e,g,j,m], jac[e,g]."""
    phi = master["phi_e"]
    nd = master["nd"]
    xg = np.einsum("gi,eid->egd", phi[0], xdg)
    J = np.stack([np.einsum("gi,eid->egd", phi[1 + a], xdg) for a in range(nd)], axis=2)
    jac, Xx = _adjugate(J)
    return xg, Xx, jac


def elemface_geom(master, xdg, perm):
    """ElemFaceGeomBlock (geometry.cpp:168-205): per element face, at face
    quadrature points: xg[e,l,g,d], nl[e,l,g,d], jac[e,l,g]."""
    phi = master["phi_f"]
    nd = master["nd"]
    xf = xdg[:, perm, :]                                   # [e, l, j, d]
    xg = np.einsum("gj,eljd->elgd", phi[0], xf)
    if nd == 2:
        T = np.einsum("gj,eljd->elgd", phi[1], xf)
        jac = np.sqrt(T[..., 0] ** 2 + T[..., 1] ** 2)
        nl = np.stack([T[..., 1] / jac, -T[..., 0] / jac], axis=-1)
    else:
        T0 = np.einsum("gj,eljd->elgd", phi[1], xf)
        T1 = np.einsum("gj,eljd->elgd", phi[2], xf)
        c = np.cross(T0, T1)
        jac = np.sqrt((c * c).sum(-1))
        nl = c / jac[..., None]
    return xg, nl, jac


# --------------------------------------------------------------------------
# q-equation -- backend/Discretization/qequation.cpp
# --------------------------------------------------------------------------

def q_equation(master, Xx, jacf_nl, jac, perm):
    """qEquationElem (qequation.cpp:66-165) and qEquationElemFaceBlock (:258-340).

    Returns Mass[e,i,j], Minv[e,i,j], MinvC[d,e,k,j] (npe x npe) and
    MinvE[d,e,k,t] (npe x ndf, t = j + npf*l)."""
    phi, phiw = master["phi_e"], master["phiw_e"]
    nd, npe, npf = master["nd"], master["npe"], master["npf"]
    nfe = perm.shape[0]
    ne = jac.shape[0]
    Mass = np.einsum("gj,gi,eg->eji", phiw[0], phi[0], jac)
    Minv = np.linalg.inv(Mass)
    MinvC = np.empty((nd, ne, npe, npe))
    for d in range(nd):
        # C_d[j', i] = sum_g sum_m w dphi_j'/dxi_m phi_i Xx[g, d, m]
        C = np.einsum("mgj,gi,egm->eji", phiw[1:], phi[0], Xx[:, :, d, :])
        MinvC[d] = Minv @ C
    phifw, phif = master["phiw_f"], master["phi_f"]
    nl, jacf = jacf_nl
    MinvE = np.empty((nd, ne, npe, npf * nfe))
    for d in range(nd):
        Et = np.einsum("gk,gi,elg->elki", phifw[0], phif[0], nl[..., d] * jacf)
        E = np.zeros((ne, npe, npf * nfe))
        for l in range(nfe):
            # assembleMatrixE (kokkosimpl.h:1634-1648): E[perm[k,l], i + npf*l] += Et
            np.add.at(E, (slice(None), perm[l][:, None], (np.arange(npf) + npf * l)[None, :]),
                      Et[:, l])
        MinvE[d] = Minv @ E
    return Mass, Minv, MinvC, MinvE


# --------------------------------------------------------------------------
# the HDG problem object
# --------------------------------------------------------------------------

class HDGProblem:
    """State of one HDG solve, built from a datain directory.

    Follows CDiscretization's constructor (discretization.cpp:170-412) for the
    parts on the HDG path: readInput, buildConn, geometry, qEquation, initial uh.
    """

    def __init__(self, datain, model_name):
        app = read_app(datain + "/app.bin")
        master = read_master(datain + "/master.bin")
        mesh = read_mesh(datain + "/mesh.bin")
        nd_app = app["ndims"].astype(int)
        self.nc, self.ncu, self.nco = nd_app[5], nd_app[6], nd_app[9]
        self.app, self.master, self.mesh = app, master, mesh
        self.nd, self.npe, self.npf = master["nd"], master["npe"], master["npf"]
        self.ne, self.nfe = mesh["ne"], mesh["nfe"]
        sol = read_sol(datain + "/sol.bin", self.npe, self.nc, self.ncu, self.ne)
        self.model = make_model(model_name, self.nd, self.ncu)
        self.conn = build_conn(master, mesh)
        self.nf = self.conn["nf"]
        self.perm = self.conn["perm"]
        self.elemcon = self.conn["elemcon"]
        flag = app["flag"].astype(int)
        prob = app["problem"].astype(int)
        self.preconditioner = flag[6]
        self.ortho = flag[5]
        self.newton_maxit = prob[13]
        self.gmres_maxit = prob[14]
        self.restart = prob[15]
        self.newton_tol = app["solversparam"][0]
        self.gmres_tol = app["solversparam"][1]
        self.tau = app["tau"]
        self.param = app["physicsparam"]
        # time dependence (setstructs.cpp:126-149): flag[0] tdep, flag[10] tdfunc,
        # problem[3] torder, problem[4] DIRK stages; the time steps are app.dt and the
        # initial time app.factor[0] (setstructs.cpp:194)
        self.tdep = int(flag[0])
        self.tdfunc = int(flag[10])
        self.torder, self.tstages = int(prob[3]), int(prob[4])
        self.tscheme = int(prob[2])                 # 0 DIRK, 1 BDF
        self.dt = np.asarray(app["dt"], dtype=float)
        self.time = float(app["factor"][0]) if len(app["factor"]) else 0.0
        self.dtfactor = 0.0
        # w-equation (wequation.cpp:31-240): ncw = ndims[13], alpha dw/dt + beta w = Sourcew with
        # dae_alpha = factor[1], dae_beta = factor[2] (setstructs.cpp:196-197)
        self.ncw = int(nd_app[13]) if len(nd_app) > 13 else 0
        fac = app["factor"]
        self.dae_alpha = float(fac[1]) if len(fac) > 1 else 1.0
        self.dae_beta = float(fac[2]) if len(fac) > 2 else 0.0
        if int(flag[1]) != 0:
            raise NotImplementedError("wave problems")
        self.xdg = sol["xdg"]
        self.udg = sol["udg"] if sol["udg"] is not None else np.zeros((self.ne, self.npe, self.nc))
        self.sdg = np.zeros((self.ne, self.npe, self.ncu))       # DIRK stage source (ncs = ncu)
        self.odg = sol["odg"] if sol["odg"] is not None else np.zeros((self.ne, self.npe, 0))
        if self.ncw:
            self.wdg = sol["wdg"] if sol["wdg"] is not None else np.zeros((self.ne, self.npe, self.ncw))
            self.wsrc = np.zeros((self.ne, self.npe, self.ncw))
        self.xg, self.Xx, self.jac = elem_geom(master, self.xdg)
        self.xgf, self.nlf, self.jacf = elemface_geom(master, self.xdg, self.perm)
        self.Mass, self.Minv, self.MinvC, self.MinvE = q_equation(
            master, self.Xx, (self.nlf, self.jacf), self.jac, self.perm)
        self.og = np.einsum("gi,eic->egc", master["phi_e"][0], self.odg)
        self.ogf = np.einsum("gj,eljc->elgc", master["phi_f"][0], self.odg[:, self.perm, :])
        f2e = self.conn["f2e"]
        # initial uh from the first element of each face (GetFaceNodes, kokkosimpl.h:901-915)
        if sol["uh"] is not None:
            self.uh = sol["uh"].reshape(self.nf, self.npf, self.ncu).copy()
        else:
            self.uh = self.udg[f2e[:, 0][:, None], self.perm[f2e[:, 1]], :self.ncu].copy()

    # ------------------------------------------------------------------
    @property
    def n(self):
        return self.npe * self.ncu

    @property
    def m(self):
        return self.npf * self.nfe * self.ncu

    def gather_uh(self, uh):
        """GetElementFaceNodes with elemcon (kokkosimpl.h:999-1030): [e, l, j, c]."""
        return uh.reshape(-1, self.ncu)[self.elemcon]

    def get_q(self, udg, uh):
        """hdgGetQ (qequation.cpp:372-487): q_d = MinvC_d u - MinvE_d uh."""
        ncu, nd = self.ncu, self.nd
        u = udg[:, :, :ncu]
        uhe = self.gather_uh(uh).reshape(self.ne, self.npf * self.nfe, ncu)
        for d in range(nd):
            q = self.MinvC[d] @ u - self.MinvE[d] @ uhe
            udg[:, :, ncu * (d + 1):ncu * (d + 2)] = q

    # ------------------------------------------------------------------
    def w_solve(self, x, uq, v, w0, wsrc):
        """wEquation (wequation.cpp:108-224) at N points: Newton on
        R(w) = Sourcew(u, q, w) + alpha wsrc - (alpha dtfactor + beta) w = 0 from w0, steps
        dw = J^{-1} R with J = (alpha dtfactor + beta) I - dSourcew/dw, until max |dw| < 1e-8 (at
        most 20 steps); then w_uq = J^{-1} dSourcew/duq.  Returns w [N, ncw], w_uq [N, ncw, nc]."""
        scal = self.dae_alpha * self.dtfactor + self.dae_beta
        w = np.array(w0, dtype=float, copy=True)
        eye = np.eye(self.ncw)[None]
        for _ in range(20):
            sw, sw_uq, sw_w = self.model.sourcew(x, uq, v, w, self.param, self.time)
            R = sw + self.dae_alpha * wsrc - scal * w
            J = scal * eye - sw_w
            dw = np.linalg.solve(J, R[..., None])[..., 0]
            w = w + dw
            if np.abs(dw).max() < 1e-8:
                break
        else:
            raise FloatingPointError("Newton in the w equation does not converge to 1e-8")
        sw, sw_uq, sw_w = self.model.sourcew(x, uq, v, w, self.param, self.time)
        w_uq = np.linalg.solve(scal * eye - sw_w, sw_uq)
        return w, w_uq

    def get_w(self):
        """GetW (wequation.cpp:226-240): the w equation at the element nodes with the nodal
        (u, q), the nodal w as the first guess and the nodal wsrc; updates wdg."""
        if not self.ncw:
            return
        N = self.ne * self.npe
        w, _ = self.w_solve(self.xdg.reshape(N, self.nd), self.udg.reshape(N, self.nc),
                            self.odg.reshape(N, -1), self.wdg.reshape(N, self.ncw),
                            self.wsrc.reshape(N, self.ncw))
        self.wdg = w.reshape(self.ne, self.npe, self.ncw)

    def _volume(self, udg, jacobian):
        """uEquationElemBlock (uequation.cpp:45-203) / RuEquationElemBlock (:790-880)."""
        m_ = self.master
        phi, phiw = m_["phi_e"], m_["phiw_e"]
        ne, npe, ncu, nc, nd = self.ne, self.npe, self.ncu, self.nc, self.nd
        uqg = np.einsum("gi,eic->egc", phi[0], udg)
        N = ne * phi.shape[1]
        x = self.xg.reshape(N, nd)
        uq = uqg.reshape(N, nc)
        v = self.og.reshape(N, -1)
        if self.ncw:  # w at the points, and the chain rule through w_uq (uequation.cpp:105-165)
            w0 = np.einsum("gi,eic->egc", phi[0], self.wdg).reshape(N, self.ncw)
            wsg = np.einsum("gi,eic->egc", phi[0], self.wsrc).reshape(N, self.ncw)
            w, w_uq = self.w_solve(x, uq, v, w0, wsg)
            s, s_uq, s_w = self.model.source(x, uq, v, self.param, self.time, w=w)
            s_uq = s_uq + np.einsum("nkw,nwc->nkc", s_w, w_uq)
        else:
            s, s_uq = self.model.source(x, uq, v, self.param, self.time)
        if self.tdep:
            # uequation.cpp:126-151: s += (sdg - dtfactor*u) T, s_u -= dtfactor T, with T the
            # model's Tdfunc values when flag tdfunc is 1 and ones otherwise (the derivative of T
            # does not enter the Jacobian)
            sdgg = np.einsum("gi,eic->egc", phi[0], self.sdg).reshape(N, ncu)
            T = np.ones((N, ncu))
            if self.tdfunc == 1 and hasattr(self.model, "tdfunc"):
                T = self.model.tdfunc(x, uq, v, self.param, self.time)
            s = s + (sdgg - self.dtfactor * uq[:, :ncu]) * T
            s_uq = s_uq.copy()
            for c in range(ncu):
                s_uq[:, c, c] -= self.dtfactor * T[:, c]
        if self.ncw:
            f, f_uq, f_w = self.model.flux(x, uq, v, self.param, self.time, w=w)
            f_uq = f_uq + np.einsum("nkdw,nwc->nkdc", f_w, w_uq)
        else:
            f, f_uq = self.model.flux(x, uq, v, self.param, self.time)
        ng = phi.shape[1]
        s = s.reshape(ne, ng, ncu)
        f = f.reshape(ne, ng, ncu, nd)
        # ApplyXxJac (kokkosimpl.h:2105-2128): r0 = s*jac, r_{m+1} = sum_j f_j Xx[j, m]
        r = np.empty((nd + 1, ne, ng, ncu))
        r[0] = s * self.jac[..., None]
        r[1:] = np.einsum("egkj,egjm->megk", f, self.Xx)
        Ru = np.einsum("dgi,degk->eik", phiw, r)           # [e, i, k]
        if not jacobian:
            return Ru
        s_uq = s_uq.reshape(ne, ng, ncu, nc)
        f_uq = f_uq.reshape(ne, ng, ncu, nd, nc)
        rJ = np.empty((nd + 1, ne, ng, ncu, nc))
        rJ[0] = s_uq * self.jac[..., None, None]
        rJ[1:] = np.einsum("egkjn,egjm->megkn", f_uq, self.Xx)
        # D/B[e, i, k, j, n] = -sum_g sum_d' phiw[d',g,i] phi[g,j] rJ[d',g,k,n]
        W = np.einsum("dgi,degkn->egikn", phiw, rJ)
        A = -np.einsum("egikn,gj->eikjn", W, phi[0])
        return Ru, A

    def _face_fluxes(self, udg, uh, jacobian):
        """FhatDriver (Model/KokkosDrivers.cpp:874-980) at every element face."""
        m_ = self.master
        phif = m_["phi_f"]
        ne, nfe, ncu, nc, nd = self.ne, self.nfe, self.ncu, self.nc, self.nd
        ngf = phif.shape[1]
        uhg = np.einsum("gj,eljc->elgc", phif[0], self.gather_uh(uh))
        udgf = np.einsum("gj,eljc->elgc", phif[0], udg[:, self.perm, :])
        N = ne * nfe * ngf
        x = self.xgf.reshape(N, nd)
        nl = self.nlf.reshape(N, nd)
        v = self.ogf.reshape(N, -1)
        uq = udgf.reshape(N, nc).copy()
        uhv = uhg.reshape(N, ncu)
        u_save = uq[:, :ncu].copy()
        uq[:, :ncu] = uhv
        jacf = self.jacf.reshape(N)
        out = dict(x=x, nl=nl, v=v, uq=udgf.reshape(N, nc), uh=uhv, jac=jacf)
        if self.ncw:
            # w at the face points with (uhat, q) in place of (u, q) (uequation.cpp:262-280): its
            # derivative in the u slot is a derivative in uhat
            w0 = np.einsum("gj,eljc->elgc", phif[0], self.wdg[:, self.perm, :]).reshape(N, self.ncw)
            wsf = np.einsum("gj,eljc->elgc", phif[0], self.wsrc[:, self.perm, :]).reshape(N, self.ncw)
            w, w_uq = self.w_solve(x, uq, v, w0, wsf)
            out["w"], out["w_uq"] = w, w_uq
            f, f_uq, f_w = self.model.flux(x, uq, v, self.param, self.time, w=w)
        else:
            f, f_uq = self.model.flux(x, uq, v, self.param, self.time)
        fh = np.einsum("nkd,nd->nk", f, nl) + self.tau[0] * (u_save - uhv)
        out["fh"] = fh * jacf[:, None]
        if jacobian:
            fh_uq = np.einsum("nkdc,nd->nkc", f_uq, nl)
            fh_uh = fh_uq[:, :, :ncu].copy()
            fh_uq[:, :, :ncu] = 0.0
            if self.ncw:  # fh_uh += fh_w w_u, fh_q += fh_w w_q
                fh_w = np.einsum("nkdw,nd->nkw", f_w, nl)
                fh_uh += np.einsum("nkw,nwc->nkc", fh_w, w_uq[:, :, :ncu])
                fh_uq[:, :, ncu:] += np.einsum("nkw,nwc->nkc", fh_w, w_uq[:, :, ncu:])
            for k in range(ncu):
                fh_uq[:, k, k] += self.tau[0]
                fh_uh[:, k, k] -= self.tau[0]
            out["fh_uq"] = fh_uq * jacf[:, None, None]
            out["fh_uh"] = fh_uh * jacf[:, None, None]
        return out

    def _boundary_fluxes(self, ff, jacobian):
        """Boundary faces (uequation.cpp:324-395): FbouDriver per condition number."""
        ne, nfe, ncu = self.ne, self.nfe, self.ncu
        ngf = self.master["phi_f"].shape[1]
        bf = self.conn["bf"]
        res = {}
        for ib in np.unique(bf[bf > 0]):
            ee, ll = np.nonzero(bf == ib)
            idx = ((ee * nfe + ll)[:, None] * ngf + np.arange(ngf)[None, :]).ravel()
            if self.ncw:  # the face points' w (from (uhat, q)) and the same chain rule
                fb, fb_uq, fb_uh, fb_w = self.model.fbou(int(ib), ff["x"][idx], ff["uq"][idx], ff["v"][idx],
                                                         ff["uh"][idx], ff["nl"][idx], self.tau,
                                                         self.param, self.time, w=ff["w"][idx])
                wu = ff["w_uq"][idx]
                fb_uh = fb_uh + np.einsum("nkw,nwc->nkc", fb_w, wu[:, :, :ncu])
                fb_uq = fb_uq.copy()
                fb_uq[:, :, ncu:] += np.einsum("nkw,nwc->nkc", fb_w, wu[:, :, ncu:])
            else:
                fb, fb_uq, fb_uh = self.model.fbou(int(ib), ff["x"][idx], ff["uq"][idx], ff["v"][idx],
                                                   ff["uh"][idx], ff["nl"][idx], self.tau,
                                                   self.param, self.time)
            jb = ff["jac"][idx]
            res[int(ib)] = (ee, ll, fb * jb[:, None], fb_uq * jb[:, None, None],
                            fb_uh * jb[:, None, None])
        return res

    def _face_project(self, ff, bres, jacobian):
        """Face quadrature -> face-node arrays (Gauss2Node with shapfgw and
        shapfgwdotshapfg, uequation.cpp:306-322), with the interior formula on
        every face ("t") and with boundary faces replaced by FbouHdg ("b",
        uequation.cpp:372-394).  Layouts: R[e,l,a,k], Q[e,l,a,k,b,c],
        F[e,l,a,k,b,c] with test face node a, equation k, trial node b, unknown c."""
        phif, phifw = self.master["phi_f"], self.master["phiw_f"]
        ne, nfe, ncu, nc = self.ne, self.nfe, self.ncu, self.nc
        ngf = phif.shape[1]
        Rt = np.einsum("gj,elgk->eljk", phifw[0], ff["fh"].reshape(ne, nfe, ngf, ncu))
        out = {"Rt": Rt, "Rb": Rt.copy()}
        pp = np.einsum("ga,gb->gab", phifw[0], phif[0])
        if jacobian:
            out["Qt"] = np.einsum("gab,elgkc->elakbc", pp,
                                  ff["fh_uq"].reshape(ne, nfe, ngf, ncu, nc))
            out["Ft"] = np.einsum("gab,elgkc->elakbc", pp,
                                  ff["fh_uh"].reshape(ne, nfe, ngf, ncu, ncu))
            out["Qb"], out["Fb"] = out["Qt"].copy(), out["Ft"].copy()
        for ib, (ee, ll, fb, fb_uq, fb_uh) in bres.items():
            nb = ee.size
            out["Rb"][ee, ll] = np.einsum("gj,zgk->zjk", phifw[0], fb.reshape(nb, ngf, ncu))
            if jacobian:
                out["Qb"][ee, ll] = np.einsum("gab,zgkc->zakbc", pp, fb_uq.reshape(nb, ngf, ncu, nc))
                out["Fb"][ee, ll] = np.einsum("gab,zgkc->zakbc", pp, fb_uh.reshape(nb, ngf, ncu, ncu))
        return out

    # ------------------------------------------------------------------
    def element_matrices(self, udg=None, uh=None):
        """uEquationElemBlock + uEquationElemFaceBlock: per-element D, B_d, F,
        K, G_d, H (face-block diagonal), Ru, Rh in Schur layouts, before the
        q-elimination of uEquationSchurBlock.  The element equation (Ru, D, B, F)
        uses the interior flux on every face; the trace equation (Rh, K, G, H)
        uses FbouHdg on boundary faces (uequation.cpp:300-420)."""
        udg = self.udg if udg is None else udg
        uh = self.uh if uh is None else uh
        ne, npe, npf, nfe, ncu, nc, nd = (self.ne, self.npe, self.npf, self.nfe,
                                          self.ncu, self.nc, self.nd)
        n, m = self.n, self.m
        Ru, A = self._volume(udg, True)                     # Ru[e,i,k]; A[e,i,k,j,c]
        ff = self._face_fluxes(udg, uh, True)
        bres = self._boundary_fluxes(ff, True)
        P = self._face_project(ff, bres, True)
        perm = self.perm
        Ff = np.zeros((ne, ncu, npe, nfe, npf, ncu))        # [e, k, node, l, b, c]
        Kf = np.zeros((ne, nfe, npf, ncu, ncu, npe))        # [e, l, a, k, c, node]
        Gf = np.zeros((nd, ne, nfe, npf, ncu, ncu, npe))
        H = np.zeros((ne, m, m))
        for l in range(nfe):
            pa = perm[l]
            # assembleRu / assembleMatrixBD / assembleMatrixF (kokkosimpl.h:1650-1712)
            np.subtract.at(Ru, (slice(None), pa, slice(None)), P["Rt"][:, l])
            np.add.at(A, (slice(None), pa[:, None], slice(None), pa[None, :], slice(None)),
                      P["Qt"][:, l].transpose(1, 3, 0, 2, 4))
            Ff[:, :, pa, l, :, :] += P["Ft"][:, l].transpose(0, 2, 1, 3, 4)
            # assembleMatrixGK / assembleMatrixH (kokkosimpl.h:1680-1730)
            Qb = P["Qb"][:, l]                               # [e, a, k, b, c]
            Kf[:, l][..., pa] += Qb[:, :, :, :, :ncu].transpose(0, 1, 2, 4, 3)
            for d in range(nd):
                Gf[d][:, l][..., pa] += Qb[:, :, :, :, ncu + ncu * d:ncu + ncu * (d + 1)].transpose(0, 1, 2, 4, 3)
            s = slice(l * npf * ncu, (l + 1) * npf * ncu)
            H[:, s, s] = P["Fb"][:, l].reshape(ne, npf * ncu, npf * ncu)
        D = A[..., :ncu].transpose(0, 2, 1, 4, 3).reshape(ne, n, n)
        B = np.stack([A[..., ncu + ncu * d:ncu + ncu * (d + 1)].transpose(0, 2, 1, 4, 3).reshape(ne, n, n)
                      for d in range(nd)])
        F = Ff.reshape(ne, n, m)
        K = Kf.reshape(ne, m, n)
        G = Gf.reshape(nd, ne, m, n)
        Rh = -P["Rb"].reshape(ne, m)
        Ru = Ru.transpose(0, 2, 1).reshape(ne, n)            # (i + npe*k)
        return dict(D=D, B=B, F=F, K=K, G=G, H=H, Ru=Ru, Rh=Rh)

    def schur(self, E):
        """uEquationSchurBlock (uequation.cpp:559-775): eliminate q and u."""
        ncu, nd, npe, npf, nfe = self.ncu, self.nd, self.npe, self.npf, self.nfe
        n, m, ne = self.n, self.m, self.ne
        D, F, K, H = E["D"].copy(), E["F"].copy(), E["K"].copy(), E["H"].copy()
        eye = np.eye(ncu)
        for d in range(nd):
            MC = np.einsum("ab,ekj->eakbj", eye, self.MinvC[d]).reshape(ne, n, n)
            # trace index of MinvE columns: (c + ncu*t) for component c
            ME = np.einsum("ab,ekt->eaktb", eye, self.MinvE[d]).reshape(ne, n, m)
            D += E["B"][d] @ MC
            F -= E["B"][d] @ ME
            K += E["G"][d] @ MC
            H -= E["G"][d] @ ME
        Dinv = np.linalg.inv(D)
        DinvF = -Dinv @ F
        Ru = np.einsum("eab,eb->ea", Dinv, E["Ru"])
        H = H + K @ DinvF
        Rh = E["Rh"] - np.einsum("eab,eb->ea", K, Ru)
        return dict(D=D, F=F, K=K, DinvF=DinvF, Ru=Ru, H=H, Rh=Rh)

    def assemble(self):
        """hdgAssembleLinearSystem (discretization.cpp:483-512)."""
        E = self.element_matrices()
        S = self.schur(E)
        b = self.assemble_rhs(S["Rh"])
        return E, S, b

    def residual(self, udg=None, uh=None, elem=False):
        """hdgAssembleResidual (discretization.cpp:576-591): (b, Ru[e, i + npe*k]);
        with elem=True also the element trace residuals Rh[e, t]."""
        udg = self.udg if udg is None else udg
        uh = self.uh if uh is None else uh
        ne, nfe, npf, ncu = self.ne, self.nfe, self.npf, self.ncu
        Ru = self._volume(udg, False)
        ff = self._face_fluxes(udg, uh, False)
        bres = self._boundary_fluxes(ff, False)
        P = self._face_project(ff, bres, False)
        for l in range(nfe):
            np.subtract.at(Ru, (slice(None), self.perm[l], slice(None)), P["Rt"][:, l])
        Rh = -P["Rb"].reshape(ne, self.m)
        if elem:
            return self.assemble_rhs(Rh), Ru.transpose(0, 2, 1).reshape(ne, self.n), Rh
        return self.assemble_rhs(Rh), Ru.transpose(0, 2, 1).reshape(ne, self.n)

    # ------------------------------------------------------------------
    def assemble_rhs(self, Rh):
        """hdgAssembleRHS (matvec.cpp:103-117): sum of the two sides' Rh per face node."""
        b = np.zeros((self.nf * self.npf, self.ncu))
        np.add.at(b, self.elemcon.reshape(-1), Rh.reshape(self.ne * self.nfe * self.npf, self.ncu))
        return b.reshape(-1)

    def matvec(self, H, v):
        """hdgMatVec (matvec.cpp:240-252, 377-385): gather, per-element GEMV, scatter."""
        ve = v.reshape(-1, self.ncu)[self.elemcon].reshape(self.ne, self.m)
        we = np.einsum("eab,eb->ea", H, ve)
        w = np.zeros((self.nf * self.npf, self.ncu))
        np.add.at(w, self.elemcon.reshape(-1), we.reshape(-1, self.ncu))
        return w.reshape(-1)

    def block_jacobi(self, H):
        """hdgBlockJacobi (matvec.cpp:190-217): inverse of the diagonal ncf x ncf
        face blocks of the assembled condensed matrix."""
        ncu, npf, nfe, nf, ne = self.ncu, self.npf, self.nfe, self.nf, self.ne
        ncf = ncu * npf
        BE = np.zeros((nf, ncf, ncf))
        for l in range(nfe):
            g = self.elemcon[:, l, :]                      # [e, j] global face nodes
            f = g[:, 0] // npf
            loc = g - f[:, None] * npf                      # [e, j] position in face
            rows = (np.arange(ncu)[None, None, :] + ncu * loc[:, :, None]).reshape(ne, ncf)
            src = (np.arange(ncu)[None, :] + ncu * (np.arange(npf)[:, None] + npf * l)).ravel()
            blk = H[:, src[:, None], src[None, :]]
            np.add.at(BE, (f[:, None, None], rows[:, :, None], rows[:, None, :]), blk)
        return np.linalg.inv(BE)

    def face_diagonal_blocks(self, H):
        """Assembled diagonal ncf x ncf face blocks of the condensed matrix, in the
        face's own node order: [nf, ncf, ncf]."""
        ncu, npf, nfe, nf, ne = self.ncu, self.npf, self.nfe, self.nf, self.ne
        ncf = ncu * npf
        A = np.zeros((nf, ncf, ncf))
        for l in range(nfe):
            g = self.elemcon[:, l, :]
            f = g[:, 0] // npf
            loc = g - f[:, None] * npf
            rows = (np.arange(ncu)[None, None, :] + ncu * loc[:, :, None]).reshape(ne, ncf)
            src = (np.arange(ncu)[None, :] + ncu * (np.arange(npf)[:, None] + npf * l)).ravel()
            np.add.at(A, (f[:, None, None], rows[:, :, None], rows[:, None, :]),
                      H[:, src[:, None], src[None, :]])
        return A

    def elemental_schwarz(self, H):
        """hdgElementalAdditiveSchwarz (matvec.cpp:168-188, kokkosimpl.h:1113-1142):
        each element's m x m block with the diagonal block of every face replaced by
        the assembled (both-sides) face block, inverted."""
        ncu, npf, nfe, ne = self.ncu, self.npf, self.nfe, self.ne
        ncf = ncu * npf
        A = self.face_diagonal_blocks(H)
        BE = H.copy()
        for l in range(nfe):
            g = self.elemcon[:, l, :]
            f = g[:, 0] // npf
            loc = g - f[:, None] * npf
            rows = (np.arange(ncu)[None, None, :] + ncu * loc[:, :, None]).reshape(ne, ncf)
            src = (np.arange(ncu)[None, :] + ncu * (np.arange(npf)[:, None] + npf * l)).ravel()
            blk = A[f[:, None, None], rows[:, :, None], rows[:, None, :]]
            BE[:, src[:, None], src[None, :]] = blk
        return np.linalg.inv(BE)

    def make_preconditioner(self, H):
        """CPreconditioner for the HDG scheme: 0 block-Jacobi, 1 elemental additive
        Schwarz, 2 superelement block-ILU(0) (discretization.cpp:296-320,
        matvec.cpp:119-217, preconditioner.cpp:282-318)."""
        if self.preconditioner == 0:
            BE = self.block_jacobi(H)
            return lambda r: self.apply_block_jacobi(BE, r)
        if self.preconditioner == 1:
            BEinv = self.elemental_schwarz(H)
            return lambda r: self.matvec(BEinv, r)
        if self.preconditioner == 2:
            from . import bilu0
            if not hasattr(self, "_ilu"):
                cart = self.mesh["cartgridpart"]
                if cart.size == 0:
                    raise ValueError("preconditioner 2 needs cartgridpart in mesh.bin")
                elem = bilu0.gridpartition(cart)
                row_ptr, col_ind, face = bilu0.crs_faceordering(elem, self.conn["f2e"], self.nfe)
                ind_ii, elim, L, U = bilu0.crs_indexingilu0(row_ptr, col_ind, face.shape[1])
                self._ilu = (row_ptr, col_ind, face, ind_ii, elim, L, U)
            row_ptr, col_ind, face, ind_ii, elim, L, U = self._ilu
            BE = bilu0.assemble(H, self.conn["f2e"], self.elemcon, face, row_ptr, col_ind,
                                self.npf, self.nfe, self.ncu)
            bilu0.factor(BE, ind_ii, elim)
            ncf = self.ncu * self.npf
            return lambda r: bilu0.apply(BE, ind_ii, L, U, face, r, ncf, self.nf)
        raise NotImplementedError("oracle preconditioner %d" % self.preconditioner)

    def apply_block_jacobi(self, BEinv, x):
        """ApplyPreconditioner, preconditioner == 0 (preconditioner.cpp:282-292)."""
        ncf = self.ncu * self.npf
        return np.einsum("fab,fb->fa", BEinv, x.reshape(self.nf, ncf)).reshape(-1)

    def get_dudg(self, Ru, DinvF, x):
        """hdgGetDUDG (matvec.cpp:219-238): du_e = Ru_e + DinvF_e * x_e."""
        xe = x.reshape(-1, self.ncu)[self.elemcon].reshape(self.ne, self.m)
        return Ru + np.einsum("eab,eb->ea", DinvF, xe)

    def update_udg(self, udg, du, alpha):
        """UpdateUDG (kokkosimpl.h:322-338) on the u components."""
        udg[:, :, :self.ncu] += alpha * du.reshape(self.ne, self.ncu, self.npe).transpose(0, 2, 1)


# --------------------------------------------------------------------------
# GMRES -- backend/Solver/gmres.cpp:361-597, getpoly.cpp:24-64, cpuimpl.h:167-208
# --------------------------------------------------------------------------

def givens(Hc, s, cs, sn, i):
    """cpuApplyGivensRotation (cpuimpl.h:167-197)."""
    for k in range(i):
        t = cs[k] * Hc[k] + sn[k] * Hc[k + 1]
        Hc[k + 1] = -sn[k] * Hc[k] + cs[k] * Hc[k + 1]
        Hc[k] = t
    if Hc[i + 1] == 0.0:
        cs[i], sn[i] = 1.0, 0.0
    elif abs(Hc[i + 1]) > abs(Hc[i]):
        t = Hc[i] / Hc[i + 1]
        sn[i] = 1.0 / np.sqrt(1.0 + t * t)
        cs[i] = t * sn[i]
    else:
        t = Hc[i + 1] / Hc[i]
        cs[i] = 1.0 / np.sqrt(1.0 + t * t)
        sn[i] = t * cs[i]
    t = cs[i] * s[i]
    s[i + 1] = -sn[i] * s[i]
    s[i] = t
    Hc[i] = cs[i] * Hc[i] + sn[i] * Hc[i + 1]
    Hc[i + 1] = 0.0


def back_solve(Hm, s, i):
    """cpuBackSolve (cpuimpl.h:199-208): Hm[k, j] is column j of the Hessenberg matrix."""
    y = s[:i + 1].copy()
    for j in range(i, -1, -1):
        y[j] /= Hm[j, j]
        y[:j] -= Hm[:j, j] * y[j]
    return y


def gmres(matvec, precond, b, restart, maxit, tol, nonlinear_tol, ortho=0, x0=None):
    """Left-preconditioned restarted GMRES as gmres.cpp:361-597 (HDG overload,
    spatialScheme = 1).  Returns (x, iterations)."""
    N = b.size
    nrmb = np.linalg.norm(b)
    if nrmb < nonlinear_tol:
        return np.zeros(N), 0
    x = np.zeros(N) if x0 is None else x0.copy()
    if np.linalg.norm(x) > 1e-12:
        r = b - matvec(x)
        nrmr = np.linalg.norm(r)
    else:
        r = b.copy()
        nrmr = nrmb
    tol = min(0.1, tol * nrmb / nrmr)
    r = precond(r)
    nrmb = np.linalg.norm(r)
    nrmr = nrmb
    n1 = restart + 1
    j = 0
    V = np.zeros((n1, N))
    Hm = np.zeros((n1, n1))        # Hm[k, i] = H[k + n1*i]
    while j < maxit:
        V[0] = r / nrmr
        s = np.zeros(n1)
        s[0] = nrmr
        cs = np.zeros(n1)
        sn = np.zeros(n1)
        i = 0
        while i < restart and j < maxit:
            mm = i + 1
            w = precond(matvec(V[i]))
            Hc = np.zeros(n1)
            if ortho == 0:
                for k in range(mm):                     # MGS (getpoly.cpp:24-42)
                    Hc[k] = np.dot(w, V[k])
                    w = w - Hc[k] * V[k]
            else:
                Hc[:mm] = V[:mm] @ w                    # CGS (gmres.cpp:81-88)
                w = w - V[:mm].T @ Hc[:mm]
            Hc[mm] = np.sqrt(np.dot(w, w))
            V[mm] = w * (1.0 / Hc[mm])
            givens(Hc, s, cs, sn, i)
            Hm[:, i] = Hc
            rel = abs(s[i + 1]) / nrmb
            if rel < tol:
                y = back_solve(Hm, s, i)
                x += V[:i + 1].T @ y
                return x, j                             # the reference's count (gmres.cpp:546)
            i += 1
            j += 1
        y = back_solve(Hm, s, i - 1)
        x += V[:i].T @ y
        r = precond(b - matvec(x))
        nrmr = np.linalg.norm(r)
        if nrmr / nrmb < tol:
            return x, j
    return x, j


# --------------------------------------------------------------------------
# Newton -- backend/Solution/solution.cpp:125-242, ptcsolver.cpp:347-396
# --------------------------------------------------------------------------

def newton_solve(prob, log=None):
    """CSolution::NewtonSolver for the HDG scheme.  Mutates prob.udg / prob.uh.
    Returns a dict with the Newton/GMRES history."""
    ncu, ncq = prob.ncu, prob.nc - prob.ncu
    u = prob.uh.reshape(-1).copy()
    if ncq > 0:
        prob.get_q(prob.udg, prob.uh)
    b, Ru = prob.residual()
    nrmr = np.linalg.norm(b) + np.linalg.norm(Ru)
    hist = dict(norms=[nrmr], gmres_iters=[], alphas=[])
    for it in range(prob.newton_maxit):
        E, S, b = prob.assemble()
        precond = prob.make_preconditioner(S["H"])
        mv = lambda v: prob.matvec(S["H"], v)
        x, its = gmres(mv, precond, b, prob.restart, prob.gmres_maxit, prob.gmres_tol,
                       prob.newton_tol, prob.ortho)
        hist["gmres_iters"].append(its)
        alpha = 1.0
        u += x
        prob.uh = u.reshape(prob.nf, prob.npf, ncu).copy()
        du = prob.get_dudg(S["Ru"], S["DinvF"], x)
        prob.update_udg(prob.udg, du, alpha)
        if ncq > 0:
            prob.get_q(prob.udg, prob.uh)
        prob.get_w()
        nrm0 = nrmr
        b, Ru = prob.residual()
        nrmr = np.linalg.norm(b) + np.linalg.norm(Ru)
        if nrmr > nrm0 and nrmr > 1e6:
            raise FloatingPointError("Residual norm increases more than 1e6")
        while (nrmr > nrm0 and alpha > 0.1) or np.isnan(nrmr):
            alpha /= 2.0
            u -= alpha * x
            prob.uh = u.reshape(prob.nf, prob.npf, ncu).copy()
            prob.update_udg(prob.udg, du, -alpha)
            if ncq > 0:
                prob.get_q(prob.udg, prob.uh)
            prob.get_w()
            b, Ru = prob.residual()
            nrmr = np.linalg.norm(b) + np.linalg.norm(Ru)
        hist["norms"].append(nrmr)
        hist["alphas"].append(alpha)
        if log:
            log("Newton %d alpha %g norm %g gmres %d" % (it + 1, alpha, nrmr, its))
        if nrmr < prob.newton_tol:
            break
    hist["newton_iters"] = len(hist["alphas"])
    return hist


# --------------------------------------------------------------------------
# DIRK time stepping -- backend/Solution/timestepcoeff.cpp:35-104 (DIRKcoeff),
# updatesource.cpp:32-102, previoussolutions.cpp:23-58, updatesolution.cpp:48-70,
# backend/Main/exasim.hpp:412-470 (the time loop)
# --------------------------------------------------------------------------

def dirk_tableau(nstage, torder):
    """Butcher matrix A and weights b of the DIRK schemes the reference offers:
    backward Euler (1,1), implicit midpoint (1,2), Alexander's L-stable (2,2) and
    (3,3), Crouzeix's (2,3) and (3,4)."""
    if (nstage, torder) == (1, 1):
        return np.array([[1.0]]), np.array([1.0])
    if (nstage, torder) == (1, 2):
        return np.array([[0.5]]), np.array([1.0])
    if (nstage, torder) == (2, 2):
        g = 1.0 - 1.0 / np.sqrt(2.0)
        return np.array([[g, 0.0], [1.0 - g, g]]), np.array([1.0 - g, g])
    if (nstage, torder) == (2, 3):
        g = (3.0 + np.sqrt(3.0)) / 6.0
        return np.array([[g, 0.0], [1.0 - 2.0 * g, g]]), np.array([0.5, 0.5])
    if (nstage, torder) == (3, 3):
        g = 0.4358665215
        b1 = -(6.0 * g * g - 16.0 * g + 1.0) / 4.0
        b2 = (6.0 * g * g - 20.0 * g + 5.0) / 4.0
        A = np.array([[g, 0.0, 0.0], [(1.0 - g) / 2.0, g, 0.0], [b1, b2, g]])
        return A, A[2].copy()
    if (nstage, torder) == (3, 4):
        g = np.cos(np.pi / 18.0) / np.sqrt(3.0) + 0.5
        dl = 1.0 / (6.0 * (2.0 * g - 1.0) ** 2)
        A = np.array([[g, 0.0, 0.0], [0.5 - g, g, 0.0], [2.0 * g, 1.0 - 4.0 * g, g]])
        return A, np.array([dl, 1.0 - 2.0 * dl, dl])
    raise NotImplementedError("DIRK(%d,%d)" % (nstage, torder))


def dirk_coeffs(nstage, torder):
    """The reference's form of the tableau: d = A^{-1} (column-major, d[i + S*j]),
    c = b^T A^{-1}, t = A 1 (stage abscissae)."""
    A, b = dirk_tableau(nstage, torder)
    Ai = np.linalg.inv(A)
    return b @ Ai, Ai.T.reshape(-1), A.sum(1)


def _dirk_source(src, cur, prev, stage, dt, d, S):
    """The stage source of UpdateSourceDIRK for one field: src the previous stage's source, cur
    the current solution, prev the step's starting solution."""
    if stage == 0:
        return (d[0] / dt) * cur
    if stage == 1:
        return ((d[1] + d[S + 1]) / d[0]) * src - (d[1] / dt) * cur
    if stage == 2:
        sc = d[2] / d[1]
        pc = d[2] + d[S + 2] + d[2 * S + 2] - sc * (d[1] + d[S + 1])
        return sc * src - (d[S + 2] / dt) * cur + (pc / dt) * prev
    raise NotImplementedError("DIRK with more than 3 stages")


def update_source_dirk(P, stage, dt, d, udgprev, wprev=None):
    """UpdateSourceDIRK (updatesource.cpp:32-100) for ncs = ncu, and for the w equation's
    source wsrc from wdg (updatesource.cpp:78-100)."""
    S = P.tstages
    P.dtfactor = d[stage * S + stage] / dt
    P.sdg = _dirk_source(P.sdg, P.udg[:, :, :P.ncu], udgprev, stage, dt, d, S)
    if P.ncw:
        P.wsrc = _dirk_source(P.wsrc, P.wdg, wprev, stage, dt, d, S)


def bdf_coeffs(order):
    """BDFcoeff (timestepcoeff.cpp:106-130): u_t ~ (c0 u^{n+1} + c1 u^n + ... ) / dt."""
    return {1: [1.0, -1.0], 2: [1.5, -2.0, 0.5], 3: [11.0 / 6.0, -3.0, 1.5, -1.0 / 3.0]}[order]


def bdf_solve(P, nsteps=None):
    """Time loop with BDF1-3 (PreviousSolutions, previoussolutions.cpp:62-100; UpdateSourceBDF,
    updatesource.cpp:104-140): dtfactor = c0/dt, sdg = -sum_k (c_k/dt) u^{n+1-k}, one Newton solve
    per step at time t + dt.  The first steps use the highest order the stored history allows
    (BDF1, then BDF2, ...): the reference reads history it has not written in its first torder - 1
    steps.  The w equation gets wsrc = -sum_k (c_k/dt) w^{n+1-k} from its own history (the
    reference's UpdateSourceBDF leaves wsrc unchanged for non-wave problems, which drops w's
    time derivative history)."""
    nsteps = len(P.dt) if nsteps is None else nsteps
    time = P.time
    hist, whist = [], []                            # hist[k] = u^{n-k}, whist[k] = w^{n-k}
    gmres_total = 0
    for istep in range(nsteps):
        dt = P.dt[istep]
        hist.insert(0, P.udg[:, :, :P.ncu].copy())
        del hist[P.torder:]
        order = min(P.torder, len(hist))
        c = bdf_coeffs(order)
        P.dtfactor = c[0] / dt
        P.sdg = sum(-(c[k + 1] / dt) * hist[k] for k in range(order))
        if P.ncw:  # the w equation's history source, as for u (see the module docstring of bdf_solve)
            whist.insert(0, P.wdg.copy())
            del whist[P.torder:]
            P.wsrc = sum(-(c[k + 1] / dt) * whist[k] for k in range(order))
        P.time = time + dt
        h = newton_solve(P)
        gmres_total += sum(h["gmres_iters"])
        time += dt
    P.time = time
    return gmres_total


def time_solve(P, nsteps=None):
    """The time loop of the problem's scheme (DIRK or BDF)."""
    return bdf_solve(P, nsteps) if P.tscheme == 1 else dirk_solve(P, nsteps)


def dirk_solve(P, nsteps=None):
    """Time loop of the t2c driver (exasim.hpp:412-470) with DIRK stages: each stage
    is a steady Newton solve with the stage source; the new solution is the
    c-weighted combination of the stages (UpdateSolutionDIRK)."""
    c, d, t = dirk_coeffs(P.tstages, P.torder)
    nsteps = len(P.dt) if nsteps is None else nsteps
    time = P.time
    gmres_total = 0
    for istep in range(nsteps):
        dt = P.dt[istep]
        udgprev = P.udg[:, :, :P.ncu].copy()
        utmp = (1.0 - c.sum()) * udgprev
        if P.ncw:  # PreviousSolutions (previoussolutions.cpp:55-60)
            wprev = P.wdg.copy()
            wtmp = (1.0 - c.sum()) * wprev
        for j in range(P.tstages):
            P.time = time + dt * t[j]
            update_source_dirk(P, j, dt, d, udgprev, wprev if P.ncw else None)
            h = newton_solve(P)
            gmres_total += sum(h["gmres_iters"])
            utmp = utmp + c[j] * P.udg[:, :, :P.ncu]
            if j == P.tstages - 1:
                P.udg[:, :, :P.ncu] = utmp
            if P.ncw:  # UpdateSolution (updatesolution.cpp:142-172): GetW, then the stage combination
                P.get_w()
                wtmp = wtmp + c[j] * P.wdg
                if j == P.tstages - 1:
                    P.wdg = wtmp
        time += dt
    P.time = time
    return gmres_total
