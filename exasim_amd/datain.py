"""Writers and readers for the Exasim ``datain/`` and ``dataout/`` binary layout.

This mirrors what the reference's text2code writes for a ``pdeapp.txt`` with
``mpiprocs = 1`` on tensor-product elements (quadrilaterals and hexahedra) and on simplices
(triangles and tetrahedra; their quadrature rules are this module's own, see make_master):

* ``app.bin``    -- text2code/text2code/readpdeapp.cpp:830-915 (``writepde``)
* ``master.bin`` -- text2code/text2code/makemaster.cpp:1019-1225 (``initializeMaster``,
  ``writemaster``)
* ``mesh.bin``   -- text2code/text2code/writebinaryfiles.cpp:155-263 (``writemesh``; the
  connectivity arrays are left empty so that the backend builds them, exactly as
  text2code does with ``callbuildConn = false``)
* ``sol.bin``    -- text2code/text2code/writebinaryfiles.cpp:51-132 (``writesol``)

Every file is a stream of float64; integers are stored as doubles.  Arrays are
column-major with the first index fastest.  This module is plain numpy: it
prepares inputs on the host, it is not on the device path.
"""
import os
import numpy as np

# 1D node sets of the tensor-product master element, porder 1..8, as tabulated in
# text2code/text2code/masternodes.bin (block 0, elemtype 1, dim 1).  Kept
# value-identical so that master.bin matches what text2code writes.
NODES_1D = {
    1: [0.0, 1.0],
    2: [0.0, 0.49999999999999994, 1.0],
    3: [0.0, 0.2928932188134524, 0.7071067811865475, 1.0],
    4: [0.0, 0.19098300562505255, 0.49999999999999994, 0.8090169943749475, 1.0],
    5: [0.0, 0.13397459621556135, 0.36602540378443865, 0.6339745962155613,
        0.8660254037844386, 1.0],
    6: [0.0, 0.09903113209758085, 0.27747906604368555, 0.49999999999999994,
        0.7225209339563143, 0.9009688679024189, 1.0],
    7: [0.0, 0.07612046748871326, 0.21677275132473922, 0.400543816310171,
        0.599456183689829, 0.7832272486752606, 0.9238795325112867, 1.0],
    8: [0.0, 0.06030737921409157, 0.1736481776669303, 0.3263518223330696,
        0.49999999999999994, 0.6736481776669302, 0.8263518223330697,
        0.9396926207859083, 1.0],
}

# 1D Gauss-Legendre rules on [0, 1] with n = 1..8 points, as tabulated in
# text2code/text2code/gaussnodes.bin (elemtype 1, dim 1).  pgauss -> n = pgauss//2 + 1.
GAUSS_1D = {
    1: ([0.5], [1.0]),
    2: ([0.21132486540518708, 0.7886751345948129],
        [0.4999999999999999, 0.4999999999999999]),
    3: ([0.11270166537925824, 0.5, 0.8872983346207417],
        [0.2777777777777778, 0.44444444444444425, 0.2777777777777778]),
    4: ([0.06943184420297371, 0.3300094782075718, 0.6699905217924281, 0.9305681557970263],
        [0.173927422568727, 0.326072577431273, 0.326072577431273, 0.17392742256872695]),
    5: ([0.046910077030668074, 0.23076534494715845, 0.5, 0.7692346550528415,
         0.9530899229693319],
        [0.11846344252809453, 0.23931433524968337, 0.28444444444444433,
         0.23931433524968343, 0.11846344252809445]),
    6: ([0.033765242898423975, 0.16939530676686776, 0.3806904069584015,
         0.6193095930415985, 0.8306046932331322, 0.9662347571015761],
        [0.08566224618958529, 0.18038078652406928, 0.2339569672863453,
         0.23395696728634602, 0.18038078652406928, 0.08566224618958522]),
    7: ([0.025446043828620812, 0.12923440720030277, 0.29707742431130146,
         0.49999999999999994, 0.7029225756886985, 0.8707655927996973,
         0.9745539561713792],
        [0.06474248308443485, 0.13985269574463838, 0.19091502525255957,
         0.2089795918367343, 0.19091502525255938, 0.1398526957446382,
         0.06474248308443485]),
    8: ([0.019855071751231967, 0.10166676129318664, 0.2372337950418354,
         0.4082826787521751, 0.5917173212478251, 0.7627662049581644,
         0.8983332387068135, 0.9801449282487682],
        [0.05061426814518811, 0.11119051722668728, 0.15685332293894375,
         0.1813418916891813, 0.18134189168918077, 0.1568533229389435,
         0.11119051722668724, 0.05061426814518823]),
}

# Reference-element vertices of the quadrilateral and hexahedron, in the order the
# bilinear/trilinear vertex basis of makemaster.cpp:848-938 (``localbasis``) uses.
QUAD_VERTS = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float64)
HEX_VERTS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                      [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)
# Vertices of each local face (0-based), backend/Discretization/connectivity.cpp:1010-1106
# (``getelemface``); face node orientation follows from this ordering.
QUAD_FACES = [[0, 1], [1, 2], [2, 3], [3, 0]]
HEX_FACES = [[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [2, 3, 7, 6], [1, 2, 6, 5], [3, 0, 4, 7]]
# Triangle and tetrahedron (elemtype 0): vertices of the reference simplex and the local faces
# of connectivity.cpp:1010-1106 (face l is opposite vertex l)
TRI_VERTS = np.array([[0, 0], [1, 0], [0, 1]], dtype=np.float64)
TET_VERTS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
TRI_FACES = [[1, 2], [2, 0], [0, 1]]
TET_FACES = [[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]]


def _lagrange_1d(nodes, x):
    """Values and first derivatives of the 1D Lagrange basis on ``nodes`` at ``x``.

    Returns (val[nx, nn], der[nx, nn]).  Mathematically identical to the
    Legendre-Vandermonde construction of makemaster.cpp:774-846 (``mkshape``).
    """
    nodes = np.asarray(nodes, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    nn = nodes.size
    val = np.ones((x.size, nn))
    der = np.zeros((x.size, nn))
    for k in range(nn):
        others = [m for m in range(nn) if m != k]
        den = np.prod(nodes[k] - nodes[others])
        for m in others:
            val[:, k] *= (x - nodes[m])
        val[:, k] /= den
        for q in others:
            term = np.ones(x.size)
            for m in others:
                if m != q:
                    term *= (x - nodes[m])
            der[:, k] += term
        der[:, k] /= den
    return val, der


def _tensor_points(x1, nd):
    """Tensor points with the first coordinate fastest."""
    x1 = np.asarray(x1, dtype=np.float64)
    n = x1.size
    if nd == 1:
        return x1[:, None].copy()
    if nd == 2:
        return np.stack([np.tile(x1, n), np.repeat(x1, n)], axis=1)
    return np.stack([np.tile(x1, n * n), np.tile(np.repeat(x1, n), n),
                     np.repeat(x1, n * n)], axis=1)


def _tensor_weights(w1, nd):
    w1 = np.asarray(w1, dtype=np.float64)
    n = w1.size
    if nd == 1:
        return w1.copy()
    w2 = np.tile(w1, n) * np.repeat(w1, n)
    if nd == 2:
        return w2
    return np.tile(w2, n) * np.repeat(w1, n * n)


def _tensor_shape(nodes1, pts, nd):
    """shap[d, g, j]: d=0 value, d>=1 derivative along coordinate d-1 of basis j at pts[g]."""
    ng = pts.shape[0]
    nn = len(nodes1)
    vals, ders = [], []
    for d in range(nd):
        v, dv = _lagrange_1d(nodes1, pts[:, d])
        vals.append(v)
        ders.append(dv)
    idx = _tensor_points(np.arange(nn, dtype=np.float64), nd).astype(int)  # [npe, nd]
    npe = idx.shape[0]
    shap = np.ones((nd + 1, ng, npe))
    for j in range(npe):
        for d in range(nd):
            shap[0, :, j] *= vals[d][:, idx[j, d]]
        for a in range(nd):
            col = np.ones(ng)
            for d in range(nd):
                col *= (ders[d] if d == a else vals[d])[:, idx[j, d]]
            shap[1 + a, :, j] = col
    return shap


def _simplex_points(nd, p):
    """Equispaced nodes of the reference simplex of degree p, the first coordinate fastest
    (the order of text2code's simplex node tables: masternodes.bin, elemtype 0)."""
    if nd == 1:
        return (np.arange(p + 1, dtype=np.float64) / p)[:, None]
    pts = []
    if nd == 2:
        for j in range(p + 1):
            for i in range(p + 1 - j):
                pts.append((i / p, j / p))
    else:
        for k in range(p + 1):
            for j in range(p + 1 - k):
                for i in range(p + 1 - k - j):
                    pts.append((i / p, j / p, k / p))
    return np.array(pts, dtype=np.float64)


def _simplex_quadrature(nd, degree):
    """Points [n, nd] and weights of a rule exact for total degree ``degree`` on the reference
    simplex: conical (collapsed-coordinate) products of Gauss-Legendre and Gauss-Jacobi rules
    (Stroud), the Jacobi weights (1 - b), (1 - c)^2 absorbing the collapse.  Weights sum to the
    simplex's measure (1/2, 1/6)."""
    from scipy.special import roots_jacobi, roots_legendre
    n = max(1, (degree + 2) // 2)
    a, wa = roots_legendre(n)
    b, wb = roots_jacobi(n, 1.0, 0.0)
    pts, w = [], []
    if nd == 2:
        for j in range(n):
            for i in range(n):
                pts.append(((1 + a[i]) * (1 - b[j]) / 4.0, (1 + b[j]) / 2.0))
                w.append(wa[i] * wb[j] / 8.0)
    else:
        c, wc = roots_jacobi(n, 2.0, 0.0)
        for k in range(n):
            for j in range(n):
                for i in range(n):
                    pts.append(((1 + a[i]) * (1 - b[j]) * (1 - c[k]) / 8.0, (1 + b[j]) * (1 - c[k]) / 4.0,
                                (1 + c[k]) / 2.0))
                    w.append(wa[i] * wb[j] * wc[k] / 64.0)
    return np.array(pts, dtype=np.float64), np.array(w, dtype=np.float64)


def _simplex_shape(nodes, pts, nd, p):
    """shap[d, g, j] of the Lagrange basis on the simplex ``nodes`` (degree p) at ``pts``: d = 0
    values, d >= 1 derivatives.  The basis is the total-degree space spanned by products of
    Legendre polynomials L_i(2x-1) L_j(2y-1) (L_k(2z-1)), i + j (+ k) <= p, made nodal by
    inverting its Vandermonde matrix at the nodes (the construction of makemaster.cpp:774-846,
    with another modal basis of the same space, so the nodal basis is the same)."""
    from numpy.polynomial import legendre as L
    if nd == 2:
        exps = [(i, j) for j in range(p + 1) for i in range(p + 1 - j)]
    else:
        exps = [(i, j, k) for k in range(p + 1) for j in range(p + 1 - k) for i in range(p + 1 - k - j)]

    def modal(x):
        x = np.asarray(x, dtype=np.float64)
        val = np.ones((nd + 1, x.shape[0], len(exps)))
        for m, e in enumerate(exps):
            for d in range(nd):
                c = np.zeros(e[d] + 1)
                c[-1] = 1.0
                v = L.legval(2.0 * x[:, d] - 1.0, c)
                dv = 2.0 * L.legval(2.0 * x[:, d] - 1.0, L.legder(c)) if e[d] > 0 else np.zeros(x.shape[0])
                val[0, :, m] *= v
                for a in range(nd):
                    val[1 + a, :, m] *= dv if a == d else v
        return val

    V = modal(nodes)[0]
    if V.shape[0] != V.shape[1]:
        raise ValueError("simplex nodes do not match the degree")
    Vinv = np.linalg.inv(V)
    return np.einsum("dgm,mj->dgj", modal(pts), Vinv)


def _face_perm_simplex(xpe, xpf, nd):
    """perm[l, j] for the triangle / tetrahedron: face node j of local face l at
    V0 + s (V1 - V0) (+ t (V2 - V0)) for the face's vertices (V0, V1 (, V2))."""
    verts, faces = (TRI_VERTS, TRI_FACES) if nd == 2 else (TET_VERTS, TET_FACES)
    perm = np.zeros((len(faces), xpf.shape[0]), dtype=np.int64)
    for l, fv in enumerate(faces):
        V = verts[fv]
        for j in range(xpf.shape[0]):
            pos = V[0] + xpf[j, 0] * (V[1] - V[0])
            if nd == 3:
                pos = pos + xpf[j, 1] * (V[2] - V[0])
            d = np.abs(xpe - pos[None, :]).max(axis=1)
            k = int(np.argmin(d))
            if d[k] > 1e-8:
                raise RuntimeError("face node not found on the master element")
            perm[l, j] = k
    return perm


def _face_perm(xpe, xpf, nd):
    """perm[l, j]: element node that sits at face node j of local face l."""
    if nd == 2:
        verts, faces = QUAD_VERTS, QUAD_FACES
    else:
        verts, faces = HEX_VERTS, HEX_FACES
    nfe = len(faces)
    npf = xpf.shape[0]
    perm = np.zeros((nfe, npf), dtype=np.int64)
    for l, fv in enumerate(faces):
        V = verts[fv]
        for j in range(npf):
            if nd == 2:
                s = xpf[j, 0]
                pos = (1 - s) * V[0] + s * V[1]
            else:
                s, t = xpf[j, 0], xpf[j, 1]
                pos = ((1 - s) * (1 - t) * V[0] + s * (1 - t) * V[1]
                       + s * t * V[2] + (1 - s) * t * V[3])
            d = np.abs(xpe - pos[None, :]).max(axis=1)
            k = int(np.argmin(d))
            if d[k] > 1e-8:
                raise RuntimeError("face node not found on the master element")
            perm[l, j] = k
    return perm


def make_master(nd, porder, pgauss, elemtype=1):
    """Master element for quadrilaterals / hexahedra (elemtype 1) or triangles / tetrahedra
    (elemtype 0), nd = 2, 3.

    Follows text2code/text2code/makemaster.cpp:1019-1146 (``initializeMaster``).
    Returns a dict of numpy arrays in natural index order plus ``ndims``.
    """
    if nd not in (2, 3):
        raise ValueError("master elements are built for nd = 2, 3")
    if elemtype == 0:
        return _make_master_simplex(nd, porder, pgauss)
    if porder not in NODES_1D:
        raise ValueError("porder must be in 1..8")
    ng1 = pgauss // 2 + 1
    if ng1 not in GAUSS_1D:
        raise ValueError("pgauss must be in 1..15")
    x1 = np.array(NODES_1D[porder])
    g1, w1 = (np.array(a) for a in GAUSS_1D[ng1])
    xpe = _tensor_points(x1, nd)
    xpf = _tensor_points(x1, nd - 1)
    gpe, gwe = _tensor_points(g1, nd), _tensor_weights(w1, nd)
    gpf, gwf = _tensor_points(g1, nd - 1), _tensor_weights(w1, nd - 1)
    shapeg = _tensor_shape(x1, gpe, nd)          # [nd+1, nge, npe]
    shapen = _tensor_shape(x1, xpe, nd)          # [nd+1, npe, npe]
    shapfg = _tensor_shape(x1, gpf, nd - 1)      # [nd, ngf, npf]
    shapfn = _tensor_shape(x1, xpf, nd - 1)
    shap1g = _tensor_shape(x1, g1[:, None], 1)   # [2, ng1, np1]
    shap1n = _tensor_shape(x1, x1[:, None], 1)
    perm = _face_perm(xpe, xpf, nd)
    return dict(nd=nd, elemtype=1, nodetype=1, porder=porder, pgauss=pgauss,
                npe=xpe.shape[0], npf=xpf.shape[0], nge=gpe.shape[0], ngf=gpf.shape[0],
                np1d=x1.size, ng1d=g1.size, nfe=2 * nd, nve=2 ** nd,
                xpe=xpe, xpf=xpf, gpe=gpe, gwe=gwe, gpf=gpf, gwf=gwf,
                shapeg=shapeg, shapen=shapen, shapfg=shapfg, shapfn=shapfn,
                shap1g=shap1g, shap1n=shap1n, xp1d=x1, gp1d=g1, gw1d=w1, perm=perm)


def _make_master_simplex(nd, porder, pgauss):
    """Triangle / tetrahedron master element: equispaced nodes, conical-product quadrature of
    degree pgauss on the element and its faces (1D faces: the Gauss-Legendre rule of the tensor
    elements), the nodal basis of _simplex_shape."""
    if porder < 1 or porder > 8:
        raise ValueError("porder must be in 1..8")
    ng1 = pgauss // 2 + 1
    if ng1 not in GAUSS_1D:
        raise ValueError("pgauss must be in 1..15")
    xpe = _simplex_points(nd, porder)
    xpf = _simplex_points(nd - 1, porder)
    x1 = _simplex_points(1, porder)[:, 0]
    g1, w1 = (np.array(a) for a in GAUSS_1D[ng1])
    gpe, gwe = _simplex_quadrature(nd, pgauss)
    if nd == 2:
        gpf, gwf = g1[:, None].copy(), w1.copy()
        shapfg = np.stack(_lagrange_1d(x1, g1), 0)
        shapfn = np.stack(_lagrange_1d(x1, x1), 0)
    else:
        gpf, gwf = _simplex_quadrature(2, pgauss)
        shapfg = _simplex_shape(xpf, gpf, 2, porder)
        shapfn = _simplex_shape(xpf, xpf, 2, porder)
    shapeg = _simplex_shape(xpe, gpe, nd, porder)
    shapen = _simplex_shape(xpe, xpe, nd, porder)
    shap1g = np.stack(_lagrange_1d(x1, g1), 0)
    shap1n = np.stack(_lagrange_1d(x1, x1), 0)
    perm = _face_perm_simplex(xpe, xpf, nd)
    return dict(nd=nd, elemtype=0, nodetype=1, porder=porder, pgauss=pgauss,
                npe=xpe.shape[0], npf=xpf.shape[0], nge=gpe.shape[0], ngf=gpf.shape[0],
                np1d=x1.size, ng1d=g1.size, nfe=nd + 1, nve=nd + 1,
                xpe=xpe, xpf=xpf, gpe=gpe, gwe=gwe, gpf=gpf, gwf=gwf,
                shapeg=shapeg, shapen=shapen, shapfg=shapfg, shapfn=shapfn,
                shap1g=shap1g, shap1n=shap1n, xp1d=x1, gp1d=g1, gw1d=w1, perm=perm)


def _cm(a):
    """Flatten a numpy array (natural order, last index fastest) column-major
    with the reversed index order, i.e. the reference's first index fastest."""
    return np.ascontiguousarray(a).ravel()


def write_master(path, m):
    """makemaster.cpp:1148-1225 (``writemaster``)."""
    nd = m["nd"]
    # shapegt[g + nge*j + nge*npe*d] <- shapeg[d, g, j] ; shapegw[j + npe*g + ...] = * gwe[g]
    shapegt = np.transpose(m["shapeg"], (0, 2, 1))                 # [d, j, g]
    shapegw = m["shapeg"] * m["gwe"][None, :, None]                 # [d, g, j]
    shapfgt = np.transpose(m["shapfg"], (0, 2, 1))
    shapfgw = m["shapfg"] * m["gwf"][None, :, None]
    shapent = np.transpose(m["shapen"], (0, 2, 1))
    shapen = m["shapen"]
    shapfnt = np.transpose(m["shapfn"], (0, 2, 1))
    shapfn = m["shapfn"]
    shap1dgt = np.transpose(m["shap1g"], (0, 2, 1))
    shap1dgw = m["shap1g"] * m["gw1d"][None, :, None]
    shap1dnt = np.transpose(m["shap1n"], (0, 2, 1))
    shap1dn = m["shap1n"]
    xpe = m["xpe"].T
    gpe = m["gpe"].T
    xpf = m["xpf"].T
    gpf = m["gpf"].T
    arrays = [shapegt, shapegw, shapfgt, shapfgw, shapent, shapen, shapfnt, shapfn,
              xpe, gpe, m["gwe"], xpf, gpf, m["gwf"], shap1dgt, shap1dgw, shap1dnt,
              shap1dn, m["xp1d"], m["gp1d"], m["gw1d"]]
    ndims = np.zeros(20)
    ndims[:11] = [nd, m["elemtype"], m["nodetype"], m["porder"], m["pgauss"], m["npe"],
                  m["npf"], m["nge"], m["ngf"], m["np1d"], m["ng1d"]]
    nsize = np.zeros(22)
    nsize[0] = ndims.size
    for i, a in enumerate(arrays):
        nsize[i + 1] = np.asarray(a).size
    with open(path, "wb") as f:
        np.array([nsize.size], dtype=np.float64).tofile(f)
        nsize.astype(np.float64).tofile(f)
        ndims.astype(np.float64).tofile(f)
        for a in arrays:
            _cm(np.asarray(a, dtype=np.float64)).tofile(f)


def structured_grid(n, lo=None, hi=None, stretch=None):
    """Vertices p[np, nd] and elements t[ne, nve] (0-based) of an n[0] x n[1] (x n[2]) grid.

    Vertex and element numbering are lexicographic with x fastest, as in the
    reference's squaregrid/cubegrid front-end generators.  ``stretch`` is an
    optional per-direction geometric growth ratio of the cell sizes.
    """
    n = list(n)
    nd = len(n)
    lo = [0.0] * nd if lo is None else list(lo)
    hi = [1.0] * nd if hi is None else list(hi)
    axes = []
    for d in range(nd):
        r = None if stretch is None else stretch[d]
        if r is None or r == 1.0:
            s = np.linspace(0.0, 1.0, n[d] + 1)
        else:
            h = r ** np.arange(n[d])
            s = np.concatenate([[0.0], np.cumsum(h)]) / h.sum()
        axes.append(lo[d] + (hi[d] - lo[d]) * s)
    if nd == 2:
        X, Y = np.meshgrid(axes[0], axes[1], indexing="xy")
        p = np.stack([X.ravel(), Y.ravel()], axis=1)
        nx = n[0] + 1
        t = []
        for j in range(n[1]):
            for i in range(n[0]):
                a = i + nx * j
                t.append([a, a + 1, a + 1 + nx, a + nx])
        return p, np.array(t, dtype=np.int64)
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    nx, ny = n[0] + 1, n[1] + 1
    t = []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                a = i + nx * j + nx * ny * k
                b = a + nx * ny
                t.append([a, a + 1, a + 1 + nx, a + nx, b, b + 1, b + 1 + nx, b + nx])
    return p, np.array(t, dtype=np.int64)


def simplex_grid(n, lo=None, hi=None, stretch=None):
    """Vertices and triangles (nd = 2) or tetrahedra (nd = 3) of the structured grid of
    ``structured_grid``: each quadrilateral split into 2 triangles along its (i, j)-(i+1, j+1)
    diagonal, each hexahedron into the 6 tetrahedra around its (0,0,0)-(1,1,1) diagonal (Kuhn),
    which match across shared faces; every element positively oriented.  Elements of a cell
    are numbered together, cells in the order of ``structured_grid``."""
    p, q = structured_grid(n, lo, hi, stretch)
    nd = p.shape[1]
    t = []
    if nd == 2:
        for a, b, c, d in q:
            t += [[a, b, c], [a, c, d]]
    else:
        for v in q:
            for path in ((1, 2), (1, 5), (4, 5), (4, 7), (3, 7), (3, 2)):
                tet = [v[0], v[path[0]], v[path[1]], v[6]]
                x = p[tet]
                if np.linalg.det(np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]])) < 0:
                    tet[2], tet[3] = tet[3], tet[2]
                t.append(tet)
    return p, np.array(t, dtype=np.int64)


def compute_dgnodes(m, p, t):
    """xdg[ne, npe, nd] = vertex basis at the master nodes times vertex coordinates.

    text2code/text2code/makemesh.cpp:809-827 (``compute_dgnodes``) with the
    bilinear/trilinear ``phielem`` of makemaster.cpp:848-938 (barycentric for simplices).
    """
    xpe = m["xpe"]
    nd = m["nd"]
    if m.get("elemtype", 1) == 0:
        phi = np.concatenate([1.0 - xpe.sum(axis=1, keepdims=True), xpe], axis=1)
        return np.einsum("iv,evd->eid", phi, p[t])
    if nd == 2:
        xi, eta = xpe[:, 0], xpe[:, 1]
        phi = np.stack([(1 - xi) * (1 - eta), xi * (1 - eta), xi * eta, (1 - xi) * eta], 1)
    else:
        xi, eta, ze = xpe[:, 0], xpe[:, 1], xpe[:, 2]
        phi = np.stack([(1 - xi) * (1 - eta) * (1 - ze), xi * (1 - eta) * (1 - ze),
                        xi * eta * (1 - ze), (1 - xi) * eta * (1 - ze),
                        (1 - xi) * (1 - eta) * ze, xi * (1 - eta) * ze,
                        xi * eta * ze, (1 - xi) * eta * ze], 1)
    return np.einsum("iv,evd->eid", phi, p[t])


def boundary_face_tags(p, t, nd, exprs):
    """bnd[ne, nfe]: index of the first boundary expression a boundary face
    satisfies, -1 for interior faces (makemesh.cpp:555-590 ``setboundaryfaces``).

    ``exprs`` are callables taking the face-vertex coordinates [nvf, nd] and
    returning True when the face lies on that boundary.
    """
    if t.shape[1] == nd + 1:
        faces = TRI_FACES if nd == 2 else TET_FACES
    else:
        faces = QUAD_FACES if nd == 2 else HEX_FACES
    ne = t.shape[0]
    nfe = len(faces)
    keys = {}
    for e in range(ne):
        for l, fv in enumerate(faces):
            k = tuple(sorted(t[e, fv]))
            keys[k] = keys.get(k, 0) + 1
    bnd = -np.ones((ne, nfe), dtype=np.int64)
    for e in range(ne):
        for l, fv in enumerate(faces):
            k = tuple(sorted(t[e, fv]))
            if keys[k] == 1:
                pv = p[t[e, fv]]
                for ib, fn in enumerate(exprs):
                    if fn(pv):
                        bnd[e, l] = ib
                        break
                else:
                    raise RuntimeError("boundary face matches no boundary expression")
    return bnd


def write_app(path, *, nd, ncu, ncw=0, nco=0, flag=None, problem=None, physicsparam=(),
              tau=(1.0,), solversparam=(1e-6, 1e-8, 1e-3, 0.0), dt=(), factor=(0, 1, 0, 0, 0),
              externalparam=(0.0, 0.0)):
    """readpdeapp.cpp:830-915 (``writepde``); ``flag``/``problem`` are dicts of
    overrides of the defaults below (index meanings: setstructs.cpp:123-181; the defaults are the
    reference's PDE defaults, readpdeapp.cpp:360-440, where the path does not set them: nch = ncu,
    AVsmoothingIter 2, frozenAVflag 1, dae_alpha 1)."""
    nc = ncu * (nd + 1)
    ndims = np.zeros(40)
    ndims[0] = 1                    # mpiprocs
    ndims[1] = nd
    ndims[5] = nc
    ndims[6] = ncu
    ndims[7] = nc - ncu
    ndims[9] = nco
    ndims[10] = ncu                 # nch
    ndims[11] = nd                  # ncx
    ndims[13] = ncw
    fl = np.zeros(17)
    fl[4] = 1          # matvecorder
    fl[10] = 1         # tdfunc
    fl[11] = 1         # source
    for k, v in (flag or {}).items():
        fl[k] = v
    pr = np.zeros(32)
    pr[0] = 1          # hybrid (HDG)
    pr[3] = 1          # torder
    pr[4] = 1          # nstage
    pr[13] = 20        # NewtonIter
    pr[14] = 200       # GMRESiter
    pr[15] = 50        # GMRESrestart
    pr[17] = 1
    pr[18] = 1         # saveSolOpt
    pr[25] = 2         # AVsmoothingIter
    pr[26] = 1         # frozenAVflag
    pr[27] = 1         # ppdegree
    for k, v in (problem or {}).items():
        pr[k] = v
    arrays = [ndims, fl, pr, np.asarray(externalparam, float), np.asarray(dt, float),
              np.asarray(factor, float), np.asarray(physicsparam, float),
              np.asarray(solversparam, float), np.asarray(tau, float)]
    nsize = np.zeros(30)
    for i, a in enumerate(arrays):
        nsize[i] = a.size
    with open(path, "wb") as f:
        np.array([nsize.size], dtype=np.float64).tofile(f)
        nsize.tofile(f)
        for a in arrays:
            a.astype(np.float64).tofile(f)


def write_mesh(path, m, t, bf, boundary_conditions, cartgridpart=()):
    """writebinaryfiles.cpp:155-263 (``writemesh``) for one process with the
    connectivity left to the backend's ``buildConn`` (readbinaryfiles.cpp:389-393)."""
    ne, nve = t.shape
    nd = m["nd"]
    ndims = np.zeros(20)
    ndims[0] = nd
    ndims[1] = ne
    ndims[3] = int(t.max()) + 1
    ndims[4] = m["nfe"]
    nsize = np.zeros(50)
    nsize[0] = 20
    nsize[9] = ne
    nsize[10] = 1
    nsize[23] = m["perm"].size
    nsize[24] = bf.size
    nsize[25] = len(cartgridpart)
    nsize[26] = ne * nve
    nsize[27] = len(boundary_conditions)
    nsize[28] = 4
    with open(path, "wb") as f:
        np.array([nsize.size], dtype=np.float64).tofile(f)
        nsize.tofile(f)
        ndims.tofile(f)
        np.arange(ne, dtype=np.float64).tofile(f)                    # elempart
        np.array([ne], dtype=np.float64).tofile(f)                   # elempartpts
        _cm(m["perm"].astype(np.float64)).tofile(f)                  # perm[npf, nfe]
        _cm(np.asarray(bf, dtype=np.float64)).tofile(f)              # bf[nfe, ne]
        np.asarray(cartgridpart, dtype=np.float64).tofile(f)
        _cm(t.astype(np.float64)).tofile(f)                          # ti[nve, ne], 0-based
        np.asarray(boundary_conditions, dtype=np.float64).tofile(f)
        np.array([ne, 0, 0, 0], dtype=np.float64).tofile(f)          # intepartpts


def write_sol(path, m, xdg, ncu, nc, udg=None, odg=None, uhat=None, wdg=None):
    """writebinaryfiles.cpp:51-132 (``writesol``).  xdg[ne, npe, nd],
    udg[ne, npe, ncu|nc], odg[ne, npe, nco], wdg[ne, npe, ncw] (natural order)."""
    ne, npe, nd = xdg.shape
    ndims = np.zeros(12)
    ndims[:12] = [ne, 0, m["nfe"], npe, m["npf"], nc, ncu, nc - ncu, 0,
                  0 if odg is None else odg.shape[2], 0, nd]
    nsize = np.zeros(20)
    nsize[0] = ndims.size
    nsize[1] = xdg.size
    nsize[2] = 0 if udg is None else udg.size
    nsize[3] = 0 if odg is None else odg.size
    nsize[4] = 0 if wdg is None else wdg.size
    nsize[5] = 0 if uhat is None else np.asarray(uhat).size
    with open(path, "wb") as f:
        np.array([nsize.size], dtype=np.float64).tofile(f)
        nsize.tofile(f)
        ndims.tofile(f)
        # xdg[npe, nd, ne] column-major == natural [ne, nd, npe]
        np.ascontiguousarray(np.transpose(xdg, (0, 2, 1))).tofile(f)
        if udg is not None:
            np.ascontiguousarray(np.transpose(udg, (0, 2, 1))).tofile(f)
        if odg is not None:
            np.ascontiguousarray(np.transpose(odg, (0, 2, 1))).tofile(f)
        if wdg is not None:
            np.ascontiguousarray(np.transpose(wdg, (0, 2, 1))).tofile(f)
        if uhat is not None:
            np.asarray(uhat, dtype=np.float64).tofile(f)


def read_output(path):
    """dataout ``out{udg,uhat}_np0.bin``: 3-double header then the array
    (solution.h:45-59)."""
    a = np.fromfile(path, dtype=np.float64)
    hdr = a[:3].astype(int)
    return hdr, a[3:]


def read_sizes(path):
    """Sizes from the headers of a datain directory's app.bin, master.bin and mesh.bin (the
    layouts written above): nd, ncu, nc, element type, p, pgauss, npe, npf, nge, ngf, ne,
    nfe and the GMRES restart length."""
    a = np.fromfile(os.path.join(path, "app.bin"), dtype=np.float64)
    na = int(a[0])
    nsa = a[1:1 + na].astype(int)
    o = 1 + na
    ndims = a[o:o + nsa[0]]
    problem = a[o + nsa[0] + nsa[1]:o + nsa[0] + nsa[1] + nsa[2]]
    m = np.fromfile(os.path.join(path, "master.bin"), dtype=np.float64, count=64)
    nm = int(m[0])
    md = m[1 + nm:1 + nm + 11].astype(int)
    h = np.fromfile(os.path.join(path, "mesh.bin"), dtype=np.float64, count=80)
    nh = int(h[0])
    hd = h[1 + nh:1 + nh + 20].astype(int)
    return dict(nd=int(ndims[1]), nc=int(ndims[5]), ncu=int(ndims[6]), elemtype=int(md[1]), porder=int(md[3]),
                pgauss=int(md[4]), npe=int(md[5]), npf=int(md[6]), nge=int(md[7]), ngf=int(md[8]),
                ne=int(hd[1]), nfe=int(hd[4]), restart=int(problem[15]))
