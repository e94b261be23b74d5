"""The model bridge on the CPU: pdemodel files under exasim_amd/models/, the oracle's
interpreter of them (oracle/dslmodel.py) and the HIP code generator's front end.

Pins of the interpreter:
* model files restating the hand-written oracle models (poisson2d.txt, convdiff3d.txt)
  give the same values and Jacobians to round-off;
* for ns2d.txt, which has no hand-written twin: finite differences of every Jacobian,
  closed forms of the inviscid flux, of the shear stress and of the heat flux, the
  boundary states that make each condition vanish, and the far-field sign matrix
  against a numerical eigendecomposition of the Euler normal-flux Jacobian.
"""
import os

import numpy as np
import pytest

from oracle import models as om
from oracle.dslmodel import DSLModel
from exasim_amd import modelgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = os.path.join(ROOT, "exasim_amd", "models")

GAM, RE, PR, MINF = 1.4, 1000.0, 0.72, 0.5


def ns_param(twall=1.0):
    from exasim_amd.cases import ns_freestream
    return np.array([GAM, RE, PR, MINF] + list(ns_freestream(GAM, MINF)) + [twall])


def _pts(rng, P, nd, ncu, nc, base=None):
    x = rng.random((P, nd))
    uq = 0.5 * rng.standard_normal((P, nc))
    if base is not None:
        uq[:, :ncu] = base + 0.1 * rng.standard_normal((P, ncu))
    v = rng.standard_normal((P, ncu))
    uh = (base if base is not None else 0.0) + 0.1 * rng.standard_normal((P, ncu))
    th = rng.random(P) * 2 * np.pi
    nl = np.stack([np.cos(th), np.sin(th)] + ([np.zeros(P)] if nd == 3 else []), 1)
    nl /= np.linalg.norm(nl, axis=1, keepdims=True)
    return x, uq, v, uh, nl


@pytest.mark.parametrize("fname,ref,param,ibs", [
    ("poisson2d.txt", om.Poisson(2, 1), np.array([0.7]), [1]),
    ("convdiff3d.txt", om.ConvDiff(3, 2), np.array([0.3, 1.0, 0.5, 0.25]), [1, 2]),
])
def test_dsl_matches_handwritten_model(fname, ref, param, ibs):
    M = DSLModel(os.path.join(MODELS, fname))
    assert (M.nd, M.ncu, M.nc) == (ref.nd, ref.ncu, ref.nc)
    rng = np.random.default_rng(3)
    x, uq, v, uh, nl = _pts(rng, 40, ref.nd, ref.ncu, ref.nc)
    tau = np.array([1.7])
    for a, b in zip(M.flux(x, uq, v, param, 0.0), ref.flux(x, uq, v, param, 0.0)):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-13)
    for a, b in zip(M.source(x, uq, v, param, 0.0), ref.source(x, uq, v, param, 0.0)):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)
    for ib in ibs:
        for a, b in zip(M.fbou(ib, x, uq, v, uh, nl, tau, param, 0.0),
                        ref.fbou(ib, x, uq, v, uh, nl, tau, param, 0.0)):
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-13)


@pytest.fixture(scope="module")
def ns():
    return DSLModel(os.path.join(MODELS, "ns2d.txt"))


def _ns_points(seed=1, P=24):
    from exasim_amd.cases import ns_freestream
    rng = np.random.default_rng(seed)
    return _pts(rng, P, 2, 4, 12, base=ns_freestream(GAM, MINF))


def test_ns_jacobians_match_finite_differences(ns):
    x, uq, v, uh, nl = _ns_points()
    mu, tau, eps = ns_param(), np.array([1.0]), 1e-6
    f, f_uq = ns.flux(x, uq, None, mu, 0.0)
    for k in range(12):
        d = np.zeros(12)
        d[k] = eps
        fd = (ns.flux(x, uq + d, None, mu, 0.0)[0] - ns.flux(x, uq - d, None, mu, 0.0)[0]) / (2 * eps)
        np.testing.assert_allclose(f_uq[..., k], fd, rtol=0, atol=1e-7 * max(1.0, np.abs(fd).max()))
    for ib in (1, 2):
        fb, fb_uq, fb_uh = ns.fbou(ib, x, uq, None, uh, nl, tau, mu, 0.0)
        for k in range(12):
            d = np.zeros(12)
            d[k] = eps
            fd = (ns.fbou(ib, x, uq + d, None, uh, nl, tau, mu, 0.0)[0] -
                  ns.fbou(ib, x, uq - d, None, uh, nl, tau, mu, 0.0)[0]) / (2 * eps)
            np.testing.assert_allclose(fb_uq[..., k], fd, rtol=0, atol=1e-7)
        for k in range(4):
            d = np.zeros(4)
            d[k] = eps
            fd = (ns.fbou(ib, x, uq, None, uh + d, nl, tau, mu, 0.0)[0] -
                  ns.fbou(ib, x, uq, None, uh - d, nl, tau, mu, 0.0)[0]) / (2 * eps)
            np.testing.assert_allclose(fb_uh[..., k], fd, rtol=0, atol=1e-7)
    s, s_uq = ns.source(x, uq, None, mu, 0.0)
    assert not s.any() and not s_uq.any()


def test_ns_flux_closed_forms(ns):
    """Inviscid flux of a uniform state; shear stress of u = (a y, 0); heat flux of a
    fluid at rest with T = 1 + b x (q = -grad u throughout)."""
    mu = ns_param()
    P = 3
    x = np.zeros((P, 2))
    r, vx, vy, p = 1.3, 0.7, -0.4, 2.1
    rE = p / (GAM - 1) + 0.5 * r * (vx * vx + vy * vy)
    uq = np.zeros((P, 12))
    uq[:, :4] = [r, r * vx, r * vy, rE]
    f, _ = ns.flux(x, uq, None, mu, 0.0)                       # f[P, ncu, nd]
    Fx = [r * vx, r * vx * vx + p, r * vx * vy, (rE + p) * vx]
    Fy = [r * vy, r * vx * vy, r * vy * vy + p, (rE + p) * vy]
    np.testing.assert_allclose(f[0, :, 0], Fx, rtol=1e-14)
    np.testing.assert_allclose(f[0, :, 1], Fy, rtol=1e-14)
    # shear: rho = 1, u = a*y (at y = 0), v = 0: F_x(rho v) = -tau_xy = -a/Re, F_y(rho u) = -a/Re
    a = 0.8
    uq = np.zeros((P, 12))
    uq[:, :4] = [1.0, 0.0, 0.0, 1.0 / (GAM * (GAM - 1) * MINF ** 2)]
    uq[:, 8 + 1] = -a                                          # q_y(rho u) = -d(rho u)/dy
    f, _ = ns.flux(x, uq, None, mu, 0.0)
    p0 = 1.0 / (GAM * MINF ** 2)
    np.testing.assert_allclose(f[0, :, 0], [0.0, p0, -a / RE, 0.0], atol=1e-15)
    np.testing.assert_allclose(f[0, :, 1], [0.0, -a / RE, p0, 0.0], atol=1e-15)
    # heat flux: at rest, rho = 1, T = 1 + b x -> F_x(rho E) = -k dT/dx, k = 1/((gam-1) M^2 Re Pr)
    b = 0.3
    uq = np.zeros((P, 12))
    uq[:, :4] = [1.0, 0.0, 0.0, 1.0 / (GAM * (GAM - 1) * MINF ** 2)]
    uq[:, 4 + 3] = -b / (GAM * (GAM - 1) * MINF ** 2)          # q_x(rho E)
    f, _ = ns.flux(x, uq, None, mu, 0.0)
    k = 1.0 / ((GAM - 1) * MINF ** 2 * RE * PR)
    np.testing.assert_allclose(f[0, 3, 0], -k * b, rtol=1e-13)


def test_ns_boundary_states(ns):
    """Wall: uhat = (rho, 0, 0, rho T_w/(gam (gam-1) M^2)) zeroes condition 1; far field:
    u = uhat = u_inf zeroes condition 2; and the far-field matrix 2 d fb/d u - I is
    V tanh(100 Lambda) V^-1 of the Euler normal-flux Jacobian at uhat."""
    x, uq, v, uh, nl = _ns_points(seed=5)
    mu, tau = ns_param(twall=1.3), np.array([1.0])
    r = uq[:, 0]
    uhw = np.stack([r, 0 * r, 0 * r, r * 1.3 / (GAM * (GAM - 1) * MINF ** 2)], 1)
    assert np.abs(ns.fbou(1, x, uq, None, uhw, nl, tau, mu, 0.0)[0]).max() < 1e-14
    uq0 = uq.copy()
    uq0[:, :4] = mu[4:8]
    fb = ns.fbou(2, x, uq0, None, np.tile(mu[4:8], (len(r), 1)), nl, tau, mu, 0.0)[0]
    assert np.abs(fb).max() < 1e-14
    fb, fb_uq, _ = ns.fbou(2, x, uq, None, uh, nl, tau, mu, 0.0)

    def euler_normal_flux(u, n):
        rho, ru, rv, rE = u
        p = (GAM - 1) * (rE - 0.5 * (ru * ru + rv * rv) / rho)
        un = (ru * n[0] + rv * n[1]) / rho
        return np.array([rho * un, ru * un + p * n[0], rv * un + p * n[1], (rE + p) * un])

    for i in range(4):
        J = np.zeros((4, 4))
        for k in range(4):
            d = np.zeros(4)
            d[k] = 1e-7
            J[:, k] = (euler_normal_flux(uh[i] + d, nl[i]) - euler_normal_flux(uh[i] - d, nl[i])) / 2e-7
        lam, V = np.linalg.eig(J)
        sign = (V @ np.diag(np.tanh(100 * lam.real)) @ np.linalg.inv(V)).real
        np.testing.assert_allclose(2 * fb_uq[i, :, :4] - np.eye(4), sign, atol=1e-6)


def test_generated_headers_are_current():
    """csrc/generated/ holds exactly what modelgen writes for the files under models/."""
    gen = os.path.join(ROOT, "exasim_amd", "csrc", "generated")
    names = sorted(f[:-4] for f in os.listdir(MODELS) if f.endswith(".txt"))
    for name in names:
        src = os.path.join(MODELS, name + ".txt")
        code, info = modelgen.emit_cpp(open(src).read(), name, "exasim_amd/models/%s.txt" % name)
        assert open(os.path.join(gen, info["ident"] + ".hpp")).read() == code, name
    reg = open(os.path.join(gen, "registry.hpp")).read()
    for k, name in enumerate(names):
        assert 'X(%d, "%s", Gen_%s)' % (100 + k, name, name) in reg


@pytest.mark.parametrize("text,msg", [
    ("vectors x(2), uq(3), uhat(1), w(1)\nfunction Flux(x, uq)\n output_size(f) = 2;\n f[0] = uq[1];\n"
     " f[1] = uq[2];\nend\nfunction Source(x, uq)\n output_size(s) = 1;\n zeros(s);\nend\n"
     "function FbouHdg(x, uq, uhat)\n output_size(fb) = 1;\n fb[0] = uhat[0];\nend\n", "w equation"),
    ("vectors x(2), uq(3), uhat(1)\nfunction Flux(x, uq)\n output_size(f) = 2;\n f[0] = uq[1];\n"
     " f[1] = uq[2];\nend\nfunction Source(x, uq)\n output_size(s) = 1;\n zeros(s);\nend\n", "FbouHdg"),
    ("vectors x(2), uq(3), uhat(1)\nfunction Flux(x, uq)\n output_size(f) = 2;\n f[0] = foo;\n"
     " f[1] = uq[2];\nend\nfunction Source(x, uq)\n output_size(s) = 1;\n zeros(s);\nend\n"
     "function FbouHdg(x, uq, uhat)\n output_size(fb) = 1;\n fb[0] = uhat[0];\nend\n", "unknown name"),
    ("vectors x(2), uq(3), uhat(1)\nfunction Flux(x, uq)\n output_size(f) = 2;\n f[0] = uq[1];\n"
     " f[1] = uq[2];\nend\nfunction Source(x, uq)\n output_size(s) = 1;\n zeros(s);\nend\n"
     "function Tdfunc(x, uq)\n output_size(m) = 2;\n m[0] = uq[0];\n m[1] = 1.0;\nend\n"
     "function FbouHdg(x, uq, uhat)\n output_size(fb) = 1;\n fb[0] = uhat[0];\nend\n",
     "Tdfunc must have ncu outputs"),
])
def test_modelgen_rejects_unsupported(text, msg):
    with pytest.raises(modelgen.ModelSyntaxError, match=msg):
        modelgen.emit_cpp(text, "bad")


def test_generated_model_checks_sizes(case_dir):
    """exa_create rejects a generated model whose (nd, ncu) differ from datain's, before
    any device work (EXA_ERR_ARG)."""
    from exasim_amd import hdg
    if not os.path.exists(hdg.LIB_PATH):
        pytest.skip("libexahip.so not built")
    path, _ = case_dir("poisson2d_small", "poisson2d", n=2, porder=1, pgauss=2)
    with pytest.raises(hdg.ExaError) as e:
        hdg.HDGSolver(path, "ns2d", device=-2)
    assert e.value.code == 1 and "is for nd = 2, ncu = 4" in str(e.value)
