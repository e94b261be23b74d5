"""Builders of ``datain/`` directories for the workloads this package is measured
and tested on (structured quad/hex meshes, mpiprocs = 1).

Each builder writes app.bin, master.bin, mesh.bin and sol.bin in the layout
text2code writes (see ``datain.py``) and returns a small dict describing the case
(model name and sizes).  The pdeapp/pdemodel settings they encode:

* ``poisson2d`` -- examples/Poisson/poisson2d (pdeapp.txt, pdemodel.txt) on an
  n x n quad mesh of the unit square, Dirichlet on all sides.
* ``convdiff_unsteady`` -- time-dependent convection-diffusion (a heat equation
  when the velocity is zero) on the unit square/cube, Dirichlet u = 0 on every
  side, initial state prod_d sin(pi x_d), zero source, DIRK time stepping.
* ``convdiff3d`` -- steady convection-diffusion on an n^3 hex mesh of the unit
  cube, ncu components, velocity c, diffusivity kappa, per-node uniform[-1, 1]
  source passed as the auxiliary field v (``vdgfile``); Dirichlet u = 0 on x = 0
  (condition 1), homogeneous Neumann (zero diffusive flux) elsewhere (condition 2).
* ``navierstokes2d`` -- compressible Navier-Stokes (model ``ns2d``, a pdemodel file
  under models/) in a [0, 10] x [0, 1] channel-like domain with a wall-normal
  stretched quad mesh, an isothermal no-slip wall at y = 0 and a characteristic far
  field elsewhere, free stream plus seeded Gaussian density bumps, DIRK time stepping.
"""
import os
import numpy as np

from . import datain as dw

PRECOND = {"bjacobi": 0, "eas": 1, "bilu0": 2}


def _app_flags(precond, ortho):
    return {5: ortho, 6: PRECOND[precond] if isinstance(precond, str) else precond}


def poisson2d(path, n=16, porder=3, pgauss=14, model="poisson", kappa=1.0, tau=1.0,
              precond="bjacobi", restart=50, gmres_tol=1e-12, gmres_maxit=1000,
              newton_iter=1, newton_tol=1e-6, ortho=0, nd=2, cartgridpart=None, elemtype=1):
    """Poisson on the unit square (nd=2) or cube (nd=3), all boundaries Dirichlet; quadrilaterals /
    hexahedra (elemtype 1) or triangles / tetrahedra (elemtype 0, datain.simplex_grid)."""
    os.makedirs(path, exist_ok=True)
    m = dw.make_master(nd, porder, pgauss, elemtype=elemtype)
    p, t = dw.structured_grid([n] * nd) if elemtype == 1 else dw.simplex_grid([n] * nd)
    if nd == 2:
        exprs = [lambda v: np.all(np.abs(v[:, 1]) < 1e-8), lambda v: np.all(np.abs(v[:, 0] - 1) < 1e-8),
                 lambda v: np.all(np.abs(v[:, 1] - 1) < 1e-8), lambda v: np.all(np.abs(v[:, 0]) < 1e-8)]
    else:
        exprs = [lambda v, d=d, s=s: np.all(np.abs(v[:, d] - s) < 1e-8)
                 for (d, s) in [(1, 0), (0, 1), (1, 1), (0, 0), (2, 0), (2, 1)]]
    bnd = dw.boundary_face_tags(p, t, nd, exprs)
    bcs = [1] * len(exprs)
    bf = np.where(bnd >= 0, np.array(bcs)[np.maximum(bnd, 0)], 0)
    dw.write_app(os.path.join(path, "app.bin"), nd=nd, ncu=1,
                 flag=_app_flags(precond, ortho),
                 problem={13: newton_iter, 14: gmres_maxit, 15: restart},
                 physicsparam=[kappa], tau=[tau], solversparam=[newton_tol, gmres_tol, 1e-3, 0.0])
    dw.write_master(os.path.join(path, "master.bin"), m)
    dw.write_mesh(os.path.join(path, "mesh.bin"), m, t, bf, bcs,
                  cartgridpart=[] if cartgridpart is None else list(cartgridpart))
    xdg = dw.compute_dgnodes(m, p, t)
    dw.write_sol(os.path.join(path, "sol.bin"), m, xdg, ncu=1, nc=nd + 1)
    return dict(model=model, nd=nd, ne=t.shape[0], npe=m["npe"], npf=m["npf"], ncu=1)


def convdiff3d(path, n=32, porder=4, pgauss=None, ncu=2, kappa=1e-2, vel=(1.0, 0.5, 0.25),
               tau=1.0, seed=20261202, precond="bjacobi", restart=100, gmres_tol=1e-10,
               gmres_maxit=2000, newton_iter=1, newton_tol=1e-8, ortho=0, cartgridpart=None,
               inflow_dirichlet=False):
    """Steady convection-diffusion on an n^3 (or n[0] x n[1] x n[2]) hex mesh of the unit cube:
    u = 0 on x = 0 and zero diffusive flux elsewhere, or, with ``inflow_dirichlet``, u = 0 on
    every inflow face (c . n < 0) and zero diffusive flux on the outflow faces.  A zero-flux
    inflow face leaves the convective term without control of u there (its boundary term
    c . n u^2 / 2 has the wrong sign), which makes GMRES stagnate at small kappa for some
    velocity directions."""
    os.makedirs(path, exist_ok=True)
    nn = [n] * 3 if np.isscalar(n) else list(n)
    pgauss = 2 * porder if pgauss is None else pgauss
    m = dw.make_master(3, porder, pgauss)
    p, t = dw.structured_grid(nn)
    planes = [(0, 0.0)]
    if inflow_dirichlet:
        planes = [(i, 0.0 if vel[i] > 0 else 1.0) for i in range(3) if abs(vel[i]) > 1e-12] or planes
    exprs = [lambda v: any(np.all(np.abs(v[:, i] - x) < 1e-8) for i, x in planes), lambda v: True]
    bnd = dw.boundary_face_tags(p, t, 3, exprs)
    bcs = [1, 2]
    bf = np.where(bnd >= 0, np.array(bcs)[np.maximum(bnd, 0)], 0)
    cg = [] if cartgridpart is None else list(cartgridpart)
    dw.write_app(os.path.join(path, "app.bin"), nd=3, ncu=ncu, nco=ncu,
                 flag=_app_flags(precond, ortho),
                 problem={13: newton_iter, 14: gmres_maxit, 15: restart},
                 physicsparam=[kappa] + list(vel), tau=[tau],
                 solversparam=[newton_tol, gmres_tol, 1e-3, 0.0])
    dw.write_master(os.path.join(path, "master.bin"), m)
    dw.write_mesh(os.path.join(path, "mesh.bin"), m, t, bf, bcs, cartgridpart=cg)
    xdg = dw.compute_dgnodes(m, p, t)
    rng = np.random.default_rng(seed)
    odg = rng.uniform(-1.0, 1.0, size=(t.shape[0], m["npe"], ncu))
    dw.write_sol(os.path.join(path, "sol.bin"), m, xdg, ncu=ncu, nc=4 * ncu, odg=odg)
    return dict(model="convdiff", nd=3, ne=t.shape[0], npe=m["npe"], npf=m["npf"], ncu=ncu)


def wmodel2d(path, model="wnl2d", n=4, porder=3, pgauss=None, mu=(1.0, 1.0, 1.0), alpha=0.0, beta=0.0,
             tau=1.0, neumann=True, dt=None, torder=1, nstage=1, scheme=0, w0=1.0, u0=0.0,
             precond="bjacobi", restart=50, gmres_tol=1e-12, gmres_maxit=1000, newton_iter=20,
             newton_tol=1e-10):
    """A model with a w equation (ncw = 1; models wnl2d, wsub2d, or wsub2d_ref without w) on the
    unit square, n x n quads: condition 1 (u = 0) on y = 0 and x = 0, condition 2 (zero total
    diffusive flux) on the other sides when ``neumann``, else condition 1 everywhere; physicsparam
    ``mu``; app.bin factor[1:3] = (dae_alpha, dae_beta) = (alpha, beta); the initial state u = u0
    and w = w0 at the nodes.  ``dt`` (a list of steps) makes the problem time dependent with DIRK
    (scheme 0; torder, nstage) or BDF (scheme 1; torder)."""
    os.makedirs(path, exist_ok=True)
    pgauss = 2 * porder if pgauss is None else pgauss
    m = dw.make_master(2, porder, pgauss)
    p, t = dw.structured_grid([n, n])
    if neumann:
        exprs = [lambda v: np.all(np.abs(v[:, 1]) < 1e-8) or np.all(np.abs(v[:, 0]) < 1e-8), lambda v: True]
    else:
        exprs = [lambda v: True]
    bnd = dw.boundary_face_tags(p, t, 2, exprs)
    bcs = [1, 2] if neumann else [1]
    bf = np.where(bnd >= 0, np.array(bcs)[np.maximum(bnd, 0)], 0)
    ncw = 0 if model.endswith("_ref") else 1
    tdep = dt is not None
    dw.write_app(os.path.join(path, "app.bin"), nd=2, ncu=1, ncw=ncw,
                 flag={0: 1 if tdep else 0, 6: PRECOND[precond] if isinstance(precond, str) else precond},
                 problem={2: scheme, 3: torder, 4: 1 if scheme == 1 else nstage, 13: newton_iter,
                          14: gmres_maxit, 15: restart},
                 physicsparam=list(mu), tau=[tau], dt=list(dt) if tdep else [0.0],
                 factor=(0.0, alpha, beta, 0.0, 0.0), solversparam=[newton_tol, gmres_tol, 1e-3, 0.0])
    dw.write_master(os.path.join(path, "master.bin"), m)
    dw.write_mesh(os.path.join(path, "mesh.bin"), m, t, bf, bcs)
    xdg = dw.compute_dgnodes(m, p, t)
    udg = np.zeros(xdg.shape[:2] + (3,))
    udg[:, :, 0] = u0
    wdg = np.full(xdg.shape[:2] + (1,), float(w0)) if ncw else None
    dw.write_sol(os.path.join(path, "sol.bin"), m, xdg, ncu=1, nc=3, udg=udg, wdg=wdg)
    return dict(model=model, nd=2, ne=t.shape[0], npe=m["npe"], npf=m["npf"], ncu=1, ncw=ncw)


def convdiff_unsteady(path, nd=2, n=4, porder=3, pgauss=None, ncu=1, kappa=1.0, vel=None, tau=1.0,
                      dt=0.01, nsteps=5, torder=3, nstage=3, precond="bjacobi", restart=50,
                      gmres_tol=1e-12, gmres_maxit=1000, newton_iter=2, newton_tol=1e-10, scheme=0,
                      capacity=None):
    """u_t + div(c u - kappa grad u) = 0 with u = 0 on the boundary and
    u(x, 0) = prod_d sin(pi x_d) (exact for c = 0: exp(-nd pi^2 kappa t) u(x, 0)).
    scheme: 0 DIRK (torder, nstage), 1 BDF (order torder).  capacity = (c0, c1): the model
    heat2d_tdfunc, (c0 + c1 x) u_t = kappa lap u, with Tdfunc switched on (flag 10); the exact
    decay for c1 = 0 is exp(-2 pi^2 kappa t / c0)."""
    os.makedirs(path, exist_ok=True)
    vel = [0.0] * nd if vel is None else list(vel)
    pgauss = 2 * porder if pgauss is None else pgauss
    m = dw.make_master(nd, porder, pgauss)
    p, t = dw.structured_grid([n] * nd)
    bnd = dw.boundary_face_tags(p, t, nd, [lambda v: True])
    bf = np.where(bnd >= 0, 1, 0)
    flags = {0: 1, 6: PRECOND[precond] if isinstance(precond, str) else precond}
    param = [kappa] + vel
    if capacity is not None:
        flags[10] = 1
        param = [kappa] + list(capacity)
    dw.write_app(os.path.join(path, "app.bin"), nd=nd, ncu=ncu, nco=0 if capacity is not None else ncu,
                 flag=flags,
                 problem={2: scheme, 3: torder, 4: 1 if scheme == 1 else nstage, 13: newton_iter,
                          14: gmres_maxit, 15: restart},
                 physicsparam=param, tau=[tau], dt=[dt] * nsteps,
                 solversparam=[newton_tol, gmres_tol, 1e-3, 0.0])
    dw.write_master(os.path.join(path, "master.bin"), m)
    dw.write_mesh(os.path.join(path, "mesh.bin"), m, t, bf, [1])
    xdg = dw.compute_dgnodes(m, p, t)
    u0 = np.prod(np.sin(np.pi * xdg), axis=2)
    udg = np.repeat(u0[:, :, None], ncu, axis=2)
    odg = None if capacity is not None else np.zeros((t.shape[0], m["npe"], ncu))
    dw.write_sol(os.path.join(path, "sol.bin"), m, xdg, ncu=ncu, nc=ncu * (nd + 1), udg=udg, odg=odg)
    return dict(model="heat2d_tdfunc" if capacity is not None else "convdiff", nd=nd, ne=t.shape[0],
                npe=m["npe"], npf=m["npf"], ncu=ncu)


def ns_freestream(gam=1.4, mach=0.5):
    """(rho, rho u, rho v, rho E) of the nondimensional free stream: rho = 1, velocity (1, 0),
    temperature 1, so p = 1/(gam M^2)."""
    p = 1.0 / (gam * mach * mach)
    return np.array([1.0, 1.0, 0.0, p / (gam - 1.0) + 0.5])


def navierstokes2d(path, nx=256, ny=32, porder=4, pgauss=None, length=(10.0, 1.0), stretch=1.05,
                   gam=1.4, mach=0.5, re=1000.0, pr=0.72, twall=1.0, nbumps=8, amp=1e-3, width=0.2,
                   seed=20261202, dt=1e-2, nsteps=20, torder=3, nstage=3, tau=1.0, precond="bilu0",
                   restart=100, gmres_tol=1e-8, gmres_maxit=2000, newton_iter=10, newton_tol=1e-8,
                   ortho=0, cartgridpart=True, se=2):
    """Compressible Navier-Stokes (model ``ns2d``) on [0, Lx] x [0, Ly], nx x ny quads with the
    cell heights growing geometrically by ``stretch`` away from the wall y = 0.

    Condition 1 (y = 0): no-slip isothermal wall at temperature ``twall``; condition 2
    (every other side): characteristic far field.  The initial state is the free stream
    (rho, u, v, T) = (1, 1, 0, 1) with ``nbumps`` Gaussian density bumps
    amp * exp(-|x - x_k|^2 / width^2) at seeded uniform centres x_k, at constant velocity and
    pressure; q = -grad u of that state.  DIRK(torder, nstage) with nsteps steps of dt.
    ``cartgridpart`` = True writes the element grid (nx, ny) with se x se superelements for the
    block-ILU(0) preconditioner.
    """
    os.makedirs(path, exist_ok=True)
    pgauss = 2 * porder if pgauss is None else pgauss
    m = dw.make_master(2, porder, pgauss)
    p, t = dw.structured_grid([nx, ny], lo=[0.0, 0.0], hi=list(length), stretch=[None, stretch])
    exprs = [lambda v: np.all(np.abs(v[:, 1]) < 1e-10), lambda v: True]
    bnd = dw.boundary_face_tags(p, t, 2, exprs)
    bcs = [1, 2]
    bf = np.where(bnd >= 0, np.array(bcs)[np.maximum(bnd, 0)], 0)
    uinf = ns_freestream(gam, mach)
    param = [gam, re, pr, mach] + list(uinf) + [twall]
    dw.write_app(os.path.join(path, "app.bin"), nd=2, ncu=4,
                 flag={0: 1, 5: ortho, 6: PRECOND[precond] if isinstance(precond, str) else precond},
                 problem={2: 0, 3: torder, 4: nstage, 13: newton_iter, 14: gmres_maxit, 15: restart},
                 physicsparam=param, tau=[tau], dt=[dt] * nsteps,
                 solversparam=[newton_tol, gmres_tol, 1e-3, 0.0])
    dw.write_master(os.path.join(path, "master.bin"), m)
    if cartgridpart is True:
        if nx % se or ny % se:
            raise ValueError("superelement size %d must divide the grid %d x %d" % (se, nx, ny))
        cg = [2, nx, ny, se, se, 0]
    else:
        cg = [] if not cartgridpart else list(cartgridpart)
    dw.write_mesh(os.path.join(path, "mesh.bin"), m, t, bf, bcs, cartgridpart=cg)
    xdg = dw.compute_dgnodes(m, p, t)                     # [ne, npe, 2]
    rng = np.random.default_rng(seed)
    centres = rng.uniform([0.0, 0.0], list(length), size=(nbumps, 2))
    rho = np.ones(xdg.shape[:2])
    drho = np.zeros(xdg.shape)                            # grad rho
    for c in centres:
        dx = xdg - c
        gk = amp * np.exp(-np.sum(dx * dx, axis=2) / width ** 2)
        rho += gk
        drho += gk[:, :, None] * (-2.0 * dx / width ** 2)
    pres = 1.0 / (gam * mach * mach)
    u = np.stack([rho, rho, 0.0 * rho, pres / (gam - 1.0) + 0.5 * rho], axis=2)
    du = np.stack([drho, drho, 0.0 * drho, 0.5 * drho], axis=2)     # [ne, npe, 4, 2]
    q = -du
    udg = np.concatenate([u, q[..., 0], q[..., 1]], axis=2)          # (u, q_x, q_y)
    dw.write_sol(os.path.join(path, "sol.bin"), m, xdg, ncu=4, nc=12, udg=udg)
    return dict(model="ns2d", nd=2, ne=t.shape[0], npe=m["npe"], npf=m["npf"], ncu=4)


def main(argv=None):
    """python -m exasim_amd.cases BUILDER DIR [key=value ...] [--solve] [--precision 0|1]
    [--precond N] [--no-verbose]: write a datain directory with one of the builders above
    and optionally solve it on the GPU, printing the Newton log, the wall time and the
    per-class kernel timings.  Values are parsed as Python literals (e.g. nx=256 dt=0.01
    length=(10.0,1.0) precond='bilu0')."""
    import argparse
    import ast
    import json
    import sys
    import time
    ap = argparse.ArgumentParser(description=main.__doc__.split("\n")[0])
    ap.add_argument("builder")
    ap.add_argument("dir")
    ap.add_argument("params", nargs="*")
    ap.add_argument("--solve", action="store_true")
    ap.add_argument("--model", default=None)
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--precond", type=int, default=None)
    ap.add_argument("--no-verbose", action="store_true")
    ap.add_argument("--save", default=None, help="write the final uh and udg to this .npz file")
    args = ap.parse_args(argv)
    fn = globals().get(args.builder)
    if fn is None or args.builder.startswith("_") or args.builder in ("main", "ns_freestream"):
        ap.error("unknown builder " + args.builder)
    kw = {}
    for p in args.params:
        k, v = p.split("=", 1)
        try:
            kw[k] = ast.literal_eval(v)
        except (ValueError, SyntaxError):
            kw[k] = v
    t0 = time.perf_counter()
    info = fn(args.dir, **kw)
    print("datain %s written in %.2f s: %s" % (args.dir, time.perf_counter() - t0, info), flush=True)
    if not args.solve:
        return 0
    from .hdg import HDGSolver
    with HDGSolver(args.dir, args.model or info["model"]) as S:
        S.set_option("verbose", 0 if args.no_verbose else 1)
        S.set_option("precision", args.precision)
        if args.precond is not None:
            S.set_option("preconditioner", args.precond)
        S.set_option("timing", 1)
        t0 = time.perf_counter()
        newton, gmres, norm = S.solve()
        wall = time.perf_counter() - t0
        print(json.dumps({"newton": newton, "gmres": gmres, "residual": norm, "wall_s": wall,
                          "sizes": S.sizes, "timings_ms": S.timings()}), flush=True)
        if args.save:
            np.savez(args.save, uh=S.get("uh"), udg=S.get("udg"))
    sys.stdout.flush()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())


def convdiff2d_param(path, n=16, porder=3, pgauss=None, kappa=0.1, angle=0.0, k=1, tau=1.0,
                     precond="eas", restart=50, gmres_tol=1e-10, gmres_maxit=2000, newton_iter=1,
                     newton_tol=1e-8, ortho=0):
    """One parametrized 2D convection-diffusion instance (model ``cd2d``): n x n quads of the
    unit square, velocity (cos angle, sin angle), diffusivity kappa, source
    sin(k pi x) cos(k pi y), u = 0 on the boundary."""
    os.makedirs(path, exist_ok=True)
    pgauss = 2 * porder if pgauss is None else pgauss
    m = dw.make_master(2, porder, pgauss)
    p, t = dw.structured_grid([n, n])
    bnd = dw.boundary_face_tags(p, t, 2, [lambda v: True])
    bf = np.where(bnd >= 0, 1, 0)
    dw.write_app(os.path.join(path, "app.bin"), nd=2, ncu=1, flag=_app_flags(precond, ortho),
                 problem={13: newton_iter, 14: gmres_maxit, 15: restart},
                 physicsparam=[kappa, np.cos(angle), np.sin(angle), float(k)], tau=[tau],
                 solversparam=[newton_tol, gmres_tol, 1e-3, 0.0])
    dw.write_master(os.path.join(path, "master.bin"), m)
    dw.write_mesh(os.path.join(path, "mesh.bin"), m, t, bf, [1])
    xdg = dw.compute_dgnodes(m, p, t)
    dw.write_sol(os.path.join(path, "sol.bin"), m, xdg, ncu=1, nc=3)
    return dict(model="cd2d", nd=2, ne=t.shape[0], npe=m["npe"], npf=m["npf"], ncu=1, n=n, porder=porder,
                kappa=kappa, angle=angle, k=k)


def convdiff2d_batch(root, count=256, sizes=(8, 16, 24, 32, 48, 64), orders=(2, 3, 4, 5), seed=20261202,
                     select=None, **kw):
    """``count`` parametrized instances of ``convdiff2d_param`` under root/inst####: n uniform
    over ``sizes``, p uniform over ``orders``, kappa log-uniform in [1e-3, 1], velocity angle
    uniform in [0, 2 pi) (magnitude 1), k uniform in {1, 2, 3, 4}.  Returns the instance
    descriptions (with their datain paths), in instance order.  ``select(index, n, porder)``, if
    given, picks the instances to write (the others are drawn but neither written nor returned)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        porder = int(rng.choice(orders))
        kappa = float(10.0 ** rng.uniform(-3.0, 0.0))
        angle = float(rng.uniform(0.0, 2.0 * np.pi))
        k = int(rng.integers(1, 5))
        if select is not None and not select(i, n, porder):
            continue
        path = os.path.join(root, "inst%04d" % i)
        info = convdiff2d_param(path, n=n, porder=porder, kappa=kappa, angle=angle, k=k, **kw)
        info["path"] = path
        out.append(info)
    return out


def convdiff3d_batch(root, count=48, sizes=(4, 6, 8, 12), orders=(2, 3, 4), seed=20261202, ncu=2,
                     precond="eas", restart=100, gmres_tol=1e-10, newton_tol=1e-8, ortho=1):
    """``count`` parametrized 3D convection-diffusion instances (``convdiff3d``: n^3 hexes of the
    unit cube, ncu components) under root/inst####: n uniform over ``sizes``, p uniform over
    ``orders``, kappa log-uniform in [1e-2, 1e-1], velocity a uniform random direction of magnitude
    1, a per-instance seeded uniform[-1, 1] nodal source, Dirichlet u = 0 on the inflow faces and
    zero diffusive flux on the outflow faces; elemental additive Schwarz.  Returns the instance
    descriptions (with their datain paths)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        porder = int(rng.choice(orders))
        kappa = float(10.0 ** rng.uniform(-2.0, -1.0))
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        path = os.path.join(root, "inst%04d" % i)
        info = convdiff3d(path, n=n, porder=porder, ncu=ncu, kappa=kappa, vel=tuple(float(x) for x in v),
                          seed=seed + 1 + i, precond=precond, restart=restart, gmres_tol=gmres_tol,
                          newton_tol=newton_tol, ortho=ortho, inflow_dirichlet=True)
        info.update(path=path, n=n, porder=porder, kappa=kappa)
        out.append(info)
    return out


def serve_stream(root, seconds=60.0, rate=50.0, frac_ns=0.7, sizes=(32, 64, 128), orders=(2, 3, 4, 5),
                 seed=20261202, ns_steps=2, ns_dt=5e-3):
    """A timed stream of parametrized solve requests: Poisson arrivals at ``rate`` per second
    for ``seconds``; each request is, with probability ``frac_ns``, a compressible
    Navier-Stokes solve (model ``ns2d``: n x n quads of the unit square, wall at y = 0 and far
    field elsewhere, free-stream start, Mach uniform in [0.2, 0.6], Reynolds number
    log-uniform in [100, 5000], ``ns_steps`` DIRK(2,2) steps of ``ns_dt``, stabilization tau =
    1 + 1/M, the free-stream wave speed |u| + c; GMRES and Newton tolerances 1e-10, which keep
    the mixed form within 1e-8 of the FP64 form), and otherwise a
    steady convection-diffusion solve (model ``cd2d``, kappa log-uniform in [1e-3, 1], velocity
    angle uniform, source sin(k pi x) cos(k pi y), k uniform in 1..4); n uniform over ``sizes``,
    p uniform over ``orders``.  Mesh, master element and node coordinates are written once per
    (model, n, p) under root/tmpl_* and linked into each request's directory.  Returns
    [(request id, datain path, model, arrival seconds)]."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    tmpl = {}
    gam, pr = 1.4, 0.72
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        rid = len(out)
        ns = rng.random() < frac_ns
        n = int(rng.choice(sizes))
        porder = int(rng.choice(orders))
        path = os.path.join(root, "req%05d" % rid)
        os.makedirs(path, exist_ok=True)
        key = ("ns2d" if ns else "cd2d", n, porder)
        if key not in tmpl:
            tdir = os.path.join(root, "tmpl_%s_%d_%d" % key)
            if ns:
                navierstokes2d(tdir, nx=n, ny=n, porder=porder, length=(1.0, 1.0), stretch=1.0, nbumps=0,
                               nsteps=ns_steps, torder=2, nstage=2, dt=ns_dt, precond="eas", restart=50,
                               gmres_tol=1e-10, newton_tol=1e-10, ortho=1, cartgridpart=False)
            else:
                convdiff2d_param(tdir, n=n, porder=porder, precond="eas", restart=100, gmres_tol=1e-10,
                                 newton_tol=1e-8, ortho=1)
            m = dw.make_master(2, porder, 2 * porder)
            p, tt = dw.structured_grid([n, n])
            tmpl[key] = (tdir, m, dw.compute_dgnodes(m, p, tt))
        tdir, m, xdg = tmpl[key]
        for f in ("master.bin", "mesh.bin"):
            os.symlink(os.path.join(tdir, f), os.path.join(path, f))
        if ns:
            mach = float(rng.uniform(0.2, 0.6))
            re = float(10.0 ** rng.uniform(2.0, np.log10(5000.0)))
            uinf = ns_freestream(gam, mach)
            dw.write_app(os.path.join(path, "app.bin"), nd=2, ncu=4,
                         flag={0: 1, 5: 1, 6: PRECOND["eas"]},
                         problem={2: 0, 3: 2, 4: 2, 13: 8, 14: 600, 15: 50},
                         physicsparam=[gam, re, pr, mach] + list(uinf) + [1.0], tau=[1.0 + 1.0 / mach],
                         dt=[ns_dt] * ns_steps, solversparam=[1e-10, 1e-10, 1e-3, 0.0])
            udg = np.broadcast_to(uinf, xdg.shape[:2] + (4,))
            dw.write_sol(os.path.join(path, "sol.bin"), m, xdg, ncu=4, nc=12, udg=np.ascontiguousarray(udg))
        else:
            kappa = float(10.0 ** rng.uniform(-3.0, 0.0))
            angle = float(rng.uniform(0.0, 2.0 * np.pi))
            k = int(rng.integers(1, 5))
            dw.write_app(os.path.join(path, "app.bin"), nd=2, ncu=1, flag=_app_flags("eas", 1),
                         problem={13: 1, 14: 2000, 15: 100},
                         physicsparam=[kappa, np.cos(angle), np.sin(angle), float(k)], tau=[1.0],
                         solversparam=[1e-8, 1e-10, 1e-3, 0.0])
            os.symlink(os.path.join(tdir, "sol.bin"), os.path.join(path, "sol.bin"))
        out.append((rid, path, key[0], t))
    return out
