"""Superelement block-ILU(0) preconditioner (preconditioner = 2) -- oracle restatement.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* gridpartition3d / 2d     backend/Common/cpuimpl.h:1332-1496
* crs_faceordering         cpuimpl.h:1266-1330 with mke2f (:440-458), mkfelem (:529-560),
                           mkf2eelem (:562-600), facereorder (:1118-1183),
                           sortrecurrence (:1024-1094), sortinteriorfaces (:946-1022),
                           crs_array (:1198-1252), mkface (:1254-1264)
* crs_indexingilu0         cpuimpl.h:1498-1600
* AssembleBlockILU0        backend/Common/kokkosimpl.h:1264-1334
* hdgBlockILU0 (factor)    backend/Discretization/matvec.cpp:119-166
* ApplyBlockILU0 (apply)   backend/Preconditioning/preconditioner.cpp:184-270, :296-318

Vectors in the ILU are laid out [ncf, nse, nfse] (face slot slowest); blocks
BE[nnz, nse] are ncf x ncf matrices A[row, col].
"""
import numpy as np


def gridpartition(cart):
    """elem[n, i]: i-th element of superelement n (cart = [nd, N..., M..., opt])."""
    nd = int(cart[0])
    N = [int(v) for v in cart[1:1 + nd]]
    M = [int(v) for v in cart[1 + nd:1 + 2 * nd]]
    opt = int(cart[1 + 2 * nd]) if len(cart) > 1 + 2 * nd else 0
    if any(N[d] % M[d] for d in range(nd)):
        raise ValueError("grid dimensions must be divisible by the block sizes")
    nb = [N[d] // M[d] for d in range(nd)]
    if nd == 2:
        Nx, Ny = N
        grid = np.zeros((Ny, Nx), dtype=np.int64)          # grid[j, i]
        for j in range(Ny):
            for i in range(Nx):
                grid[j, i] = i + Nx * j if opt == 0 else j + Ny * i
        blocks = []
        order = [(jj, ii) for jj in range(nb[1]) for ii in range(nb[0])] if opt == 0 else \
                [(jj, ii) for ii in range(nb[0]) for jj in range(nb[1])]
        for (jj, ii) in order:
            blocks.append([grid[jj * M[1] + j, ii * M[0] + i] for j in range(M[1]) for i in range(M[0])])
        return np.array(blocks, dtype=np.int64)
    Nx, Ny, Nz = N
    blocks = []
    if opt == 0:
        order = [(kk, jj, ii) for kk in range(nb[2]) for jj in range(nb[1]) for ii in range(nb[0])]
        gid = lambda i, j, k: i + Nx * j + Nx * Ny * k
    else:
        order = [(kk, jj, ii) for ii in range(nb[0]) for jj in range(nb[1]) for kk in range(nb[2])]
        gid = lambda i, j, k: k + Nz * j + Nz * Ny * i
    for (kk, jj, ii) in order:
        blocks.append([gid(ii * M[0] + i, jj * M[1] + j, kk * M[2] + k)
                       for k in range(M[2]) for j in range(M[1]) for i in range(M[0])])
    return np.array(blocks, dtype=np.int64)


def _stable_sort_perm(vals):
    """Indices of a stable ascending sort (simple_bubble_sort, cpuimpl.h:912-934)."""
    return sorted(range(len(vals)), key=lambda i: vals[i])


def _e2f(f2e, nfe, ne):
    e2f = -np.ones((ne, nfe), dtype=np.int64)
    for i, (e1, l1, e2, l2) in enumerate(f2e):
        e2f[e1, l1] = i
        if e2 >= 0:
            e2f[e2, l2] = i
    return e2f


def _f2f(f2e, e2f, nfe):
    nbf = 2 * (nfe - 1)
    f2f = -np.ones((len(f2e), nbf), dtype=np.int64)
    for i, (e1, l1, e2, l2) in enumerate(f2e):
        k = 0
        for l in range(nfe):
            if l != l1:
                f2f[i, k] = e2f[e1, l]
                k += 1
        if e2 >= 0:
            for l in range(nfe):
                if l != l2:
                    f2f[i, k] = e2f[e2, l]
                    k += 1
    return f2f


def crs_faceordering(elem, f2e, nfe):
    """Face ordering and CRS pattern of one superelement, applied to all of them.
    Returns row_ptr, col_ind, face[nse, nfse] (global face ids)."""
    ne_tot = int(max(f2e[:, 0].max(), f2e[:, 2].max())) + 1
    e2f = _e2f(f2e, nfe, ne_tot)
    el0 = list(elem[0])
    pos = {e: i for i, e in enumerate(el0)}
    felem = []
    seen = set()
    for e in el0:
        for j in range(nfe):
            f = int(e2f[e, j])
            if f not in seen:
                seen.add(f)
                felem.append(f)
    loc = []
    for f in felem:
        e1, l1, e2, l2 = (int(v) for v in f2e[f])
        in1, in2 = e1 in pos, e2 in pos
        if in1:
            a = [pos[e1], l1, pos[e2] if in2 else -1, l2 if in2 else -1]
        else:
            a = [pos[e2], l2, -1, -1]
        loc.append(a)
    loc = np.array(loc, dtype=np.int64)
    nf = len(loc)
    # facereorder: superelement-boundary faces sorted by recurrence of their element
    ind0 = [i for i in range(nf) if loc[i, 3] == -1]
    first = [int(loc[i, 0]) for i in ind0]
    cnt = [first.count(v) for v in first]
    order0 = []
    for k in range(1, max(cnt) + 1 if cnt else 1):
        grp = [i for i in range(len(first)) if cnt[i] == k]
        grp = [grp[p] for p in _stable_sort_perm([first[i] for i in grp])]
        order0 += [ind0[i] for i in grp]
    ind1 = [i for i in range(nf) if i not in set(ind0)]
    loc = loc[order0 + ind1]
    # sortinteriorfaces
    nbf = 2 * (nfe - 1)
    ne = int(max(loc[:, 0].max(), loc[:, 2].max())) + 1
    nfs = len(ind0)
    while nfs < nf:
        e2 = _e2f(loc, nfe, ne)
        f2f = _f2f(loc, e2, nfe)
        unsorted = list(range(nfs, nf))
        counts = [int(np.sum(f2f[i] >= nfs)) for i in unsorted]
        p = _stable_sort_perm(counts)
        loc[nfs:] = loc[[unsorted[q] for q in p]]
        nfs += sum(1 for c in counts if c < nbf)
    # crs_array
    e2 = _e2f(loc, nfe, ne)
    f2f = _f2f(loc, e2, nfe)
    row_ptr = [0]
    col_ind = []
    for i in range(nf):
        nb = sorted(int(v) for v in f2f[i] if v >= 0)
        col_ind += [i] + nb
        row_ptr.append(len(col_ind))
    # mkface
    nse = elem.shape[0]
    face = np.zeros((nse, nf), dtype=np.int64)
    for n in range(nse):
        for i in range(nf):
            face[n, i] = e2f[elem[n, loc[i, 0]], loc[i, 1]]
    return np.array(row_ptr), np.array(col_ind), face


def crs_indexingilu0(row_ptr, col_ind, N):
    """Elimination index lists of the block ILU(0) (crs_indexingilu0)."""
    ind_ii = np.zeros(N, dtype=np.int64)
    elim = []                                        # per row i: list of (idx_ji, [(idx_jl, idx_il)])
    for i in range(N):
        r1, r2 = row_ptr[i], row_ptr[i + 1]
        cols = list(col_ind[r1:r2])
        ind_ii[i] = r1 + cols.index(i)
        lst = []
        for p in range(r1, r2):
            j = col_ind[p]
            if j <= i:
                continue
            rj = list(col_ind[row_ptr[j]:row_ptr[j + 1]])
            if i not in rj:
                continue
            idx_ji = row_ptr[j] + rj.index(i)
            upd = []
            for pp in range(r1, r2):
                ell = col_ind[pp]
                if ell <= i or ell not in rj:
                    continue
                upd.append((row_ptr[j] + rj.index(ell), pp))
            lst.append((idx_ji, upd))
        elim.append(lst)
    L = [[(p, col_ind[p]) for p in range(row_ptr[i], row_ptr[i + 1]) if col_ind[p] < i] for i in range(N)]
    U = [[(p, col_ind[p]) for p in range(row_ptr[i], row_ptr[i + 1]) if col_ind[p] > i] for i in range(N)]
    return ind_ii, elim, L, U


def assemble(H, f2e, elemcon, face, row_ptr, col_ind, npf, nfe, ncu):
    """AssembleBlockILU0: BE[t, n] = block of the condensed matrix for CRS entry t of
    superelement n (rows/cols c + ncu*j of the two faces)."""
    ncf = ncu * npf
    nse, nfse = face.shape
    nnz = row_ptr[-1]
    BE = np.zeros((nnz, nse, ncf, ncf))

    def loc_rows(e, l, f):
        g = elemcon[e, l] - npf * f                        # global face node for local j -> its position
        return (np.arange(ncu)[None, :] + ncu * g[:, None]).reshape(-1), \
               (np.arange(ncu)[None, :] + ncu * (np.arange(npf)[:, None] + npf * l)).reshape(-1)

    for n in range(nse):
        for i in range(nfse):
            fi = face[n, i]
            e1, l1, e2, l2 = f2e[fi]
            blk = np.zeros((ncf, ncf))
            for (e, l) in ((e1, l1), (e2, l2)):
                if e < 0:
                    continue
                pos, src = loc_rows(e, l, fi)
                blk[np.ix_(pos, pos)] += H[e][np.ix_(src, src)]
            BE[row_ptr[i], n] = blk
            for t in range(row_ptr[i] + 1, row_ptr[i + 1]):
                j = col_ind[t]
                fj = face[n, j]
                je1, jl1, je2, jl2 = f2e[fj]
                if je1 == e1:
                    e, k1, k2 = e1, l1, jl1
                elif je1 == e2:
                    e, k1, k2 = e2, l2, jl1
                elif je2 == e1:
                    e, k1, k2 = e1, l1, jl2
                else:
                    e, k1, k2 = e2, l2, jl2
                pr, sr = loc_rows(e, k1, fi)
                pc, sc = loc_rows(e, k2, fj)
                b = np.zeros((ncf, ncf))
                b[np.ix_(pr, pc)] = H[e][np.ix_(sr, sc)]
                BE[t, n] = b
    return BE


def factor(BE, ind_ii, elim):
    """hdgBlockILU0 (matvec.cpp:133-165): invert diagonal blocks in place,
    L_ji = A_ji * inv(A_ii), A_jl -= L_ji * A_il, batched over superelements."""
    for i in range(len(ind_ii)):
        d = ind_ii[i]
        BE[d] = np.linalg.inv(BE[d])
        for idx_ji, upd in elim[i]:
            BE[idx_ji] = BE[idx_ji] @ BE[d]
            for idx_jl, idx_il in upd:
                BE[idx_jl] -= BE[idx_ji] @ BE[idx_il]
    return BE


def apply(BE, ind_ii, L, U, face, x, ncf, nf_total):
    """ApplyPreconditioner, preconditioner == 2: gather per superelement, ILU
    forward/backward sweeps, scatter-add (preconditioner.cpp:296-318)."""
    nse, nfse = face.shape
    xv = x.reshape(-1, ncf)
    b = xv[face.T]                                   # [nfse, nse, ncf]
    y = b.copy()
    for i in range(nfse):
        for (p, j) in L[i]:
            y[i] -= np.einsum("nab,nb->na", BE[p], y[j])
    out = np.zeros_like(y)
    for i in range(nfse - 1, -1, -1):
        yi = y[i].copy()
        for (p, j) in U[i]:
            yi -= np.einsum("nab,nb->na", BE[p], out[j])
        out[i] = np.einsum("nab,nb->na", BE[ind_ii[i]], yi)
    res = np.zeros((nf_total, ncf))
    np.add.at(res, face.T.reshape(-1), out.reshape(-1, ncf))
    return res.reshape(-1)
