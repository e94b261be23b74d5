"""Extract small fixtures from the reference checkout (run in the build container only).

Writes tests/golden/text2code_tables.npz with the tensor-element blocks of the
reference's data tables text2code/text2code/masternodes.bin and gaussnodes.bin
(node coordinates, face-node permutations, Gauss points and weights), and
tests/golden/poisson2d_xdg.npz with examples/Poisson/poisson2d/xdg.bin (the
p = 3 nodes of the 16 x 16 quad mesh of configuration C1), and
tests/golden/poisson3d.npz with the reference's committed poisson3d run
(examples/Poisson/poisson3d: datain/{app,master,mesh}.bin and grid.bin verbatim,
plus every 8th interior face of dataout/outuhat_np0.bin).

Usage:  python tests/golden/make_fixtures.py [/root/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_table(fn):
    t = np.fromfile(fn, dtype=np.float64)
    nd = int(t[0])
    na = t[1:1 + nd].astype(int)
    off = 1 + nd
    tb = int(np.prod(na))
    s1 = t[off:off + tb].astype(int)
    s2 = t[off + tb:off + 2 * tb].astype(int)
    sz = s1 * s2
    lz = np.concatenate([[0], np.cumsum(sz)])
    ds = off + 2 * tb

    def blk(i, e, p, d):
        idx = i + na[0] * (e + na[1] * (p + na[2] * d))
        return t[ds + lz[idx]:ds + lz[idx + 1]].reshape(s2[idx], s1[idx]).T
    return blk


def main(ref):
    t2c = os.path.join(ref, "text2code", "text2code")
    nodes = load_table(os.path.join(t2c, "masternodes.bin"))
    gauss = load_table(os.path.join(t2c, "gaussnodes.bin"))
    out = {}
    for dim in (1, 2, 3):
        for p in range(1, 9):
            out["xpe_d%d_p%d" % (dim, p)] = nodes(0, 1, p - 1, dim - 1)
            if dim > 1:
                out["xpf_d%d_p%d" % (dim, p)] = nodes(2, 1, p - 1, dim - 1)
                out["perm_d%d_p%d" % (dim, p)] = nodes(4, 1, p - 1, dim - 1).astype(np.int64) - 1
        for pg in range(1, 16):
            out["gp_d%d_g%d" % (dim, pg)] = gauss(0, 1, pg - 1, dim - 1)
            out["gw_d%d_g%d" % (dim, pg)] = gauss(1, 1, pg - 1, dim - 1).ravel()
    np.savez_compressed(os.path.join(HERE, "text2code_tables.npz"), **out)
    xdg = np.fromfile(os.path.join(ref, "examples", "Poisson", "poisson2d", "xdg.bin"), dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "poisson2d_xdg.npz"), header=xdg[:3], xdg=xdg[3:])
    P3 = os.path.join(ref, "examples", "Poisson", "poisson3d")
    rd = lambda *a: np.fromfile(os.path.join(P3, *a), dtype=np.float64)
    uhat = rd("dataout", "outuhat_np0.bin")
    nf = int(uhat[2])
    faces = np.arange(0, nf, 8)
    np.savez_compressed(os.path.join(HERE, "poisson3d.npz"), app=rd("datain", "app.bin"),
                        master=rd("datain", "master.bin"), mesh=rd("datain", "mesh.bin"),
                        grid=rd("grid.bin"), uhat_header=uhat[:3], faces=faces,
                        uhat_faces=uhat[3:].reshape(nf, -1)[faces])
    print("wrote", len(out), "table blocks")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
