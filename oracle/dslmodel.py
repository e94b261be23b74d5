"""Model functions from a pdemodel.txt, evaluated on the CPU -- oracle side of the
model bridge.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The file is parsed with exasim_amd.modelgen.parse (the same front end as the
HIP code generator) and interpreted here on numpy arrays of points with
forward-mode dual numbers for the Jacobians, independently of the generated
C++.  The interpreter is pinned against the hand-written oracle models
(tests/test_oracle_dsl.py): for a model file restating one of them, values and
Jacobians agree to round-off; finite differences pin the Jacobians of models
that have no hand-written twin.
"""
import ast

import numpy as np

from exasim_amd.modelgen import parse, model_sizes


class D:
    """Dual number over points: value v [P], derivatives g [nv, P]."""
    __slots__ = ("v", "g")

    def __init__(self, v, g):
        self.v, self.g = v, g

    @staticmethod
    def const(x, nv, P):
        return D(np.broadcast_to(np.asarray(x, dtype=float), (P,)).copy(), np.zeros((nv, P)))

    def __add__(self, o):
        return D(self.v + o.v, self.g + o.g)

    def __sub__(self, o):
        return D(self.v - o.v, self.g - o.g)

    def __mul__(self, o):
        return D(self.v * o.v, self.g * o.v + self.v * o.g)

    def __truediv__(self, o):
        q = self.v / o.v
        return D(q, (self.g - q * o.g) / o.v)

    def __neg__(self):
        return D(-self.v, -self.g)

    def chain(self, fv, dfv):
        return D(fv, dfv * self.g)


def _pow(a, b):
    pv = a.v ** b.v
    da = np.where(b.v == 0.0, 0.0, b.v * a.v ** (b.v - 1.0))
    g = da * a.g
    if np.any(b.g != 0.0):
        g = g + pv * np.log(a.v) * b.g
    return D(pv, g)


MATH = {
    "sin": lambda a: a.chain(np.sin(a.v), np.cos(a.v)),
    "cos": lambda a: a.chain(np.cos(a.v), -np.sin(a.v)),
    "tan": lambda a: a.chain(np.tan(a.v), 1.0 + np.tan(a.v) ** 2),
    "exp": lambda a: a.chain(np.exp(a.v), np.exp(a.v)),
    "log": lambda a: a.chain(np.log(a.v), 1.0 / a.v),
    "sqrt": lambda a: a.chain(np.sqrt(a.v), 0.5 / np.sqrt(a.v)),
    "tanh": lambda a: a.chain(np.tanh(a.v), 1.0 - np.tanh(a.v) ** 2),
    "sinh": lambda a: a.chain(np.sinh(a.v), np.cosh(a.v)),
    "cosh": lambda a: a.chain(np.cosh(a.v), np.sinh(a.v)),
    "atan": lambda a: a.chain(np.arctan(a.v), 1.0 / (1.0 + a.v ** 2)),
    "abs": lambda a: a.chain(np.abs(a.v), np.where(a.v < 0, -1.0, 1.0)),
    "fabs": lambda a: a.chain(np.abs(a.v), np.where(a.v < 0, -1.0, 1.0)),
}


def _inv(M, nv, P):
    """Gauss-Jordan with partial pivoting per point (pivot chosen on the values)."""
    n = len(M)
    W = [[D(M[i][j].v.copy(), M[i][j].g.copy()) for j in range(n)] +
         [D.const(1.0 if i == j else 0.0, nv, P) for j in range(n)] for i in range(n)]
    for c in range(n):
        # per-point pivot: permute rows point by point
        vals = np.stack([np.abs(W[r][c].v) for r in range(c, n)])
        piv = c + np.argmax(vals, axis=0)
        for r in range(c + 1, n):
            sel = piv == r
            if np.any(sel):
                for j in range(2 * n):
                    a, b = W[c][j], W[r][j]
                    av, bv, ag, bg = a.v.copy(), b.v.copy(), a.g.copy(), b.g.copy()
                    a.v[sel], b.v[sel] = bv[sel], av[sel]
                    a.g[:, sel], b.g[:, sel] = bg[:, sel], ag[:, sel]
        p = W[c][c]
        W[c] = [W[c][j] / p for j in range(2 * n)]
        for r in range(n):
            if r != c:
                f = W[r][c]
                W[r] = [W[r][j] - f * W[c][j] for j in range(2 * n)]
    return [[W[i][n + j] for j in range(n)] for i in range(n)]


class _Interp:
    def __init__(self, funcs, header, nv, P):
        self.funcs, self.header, self.nv, self.P = funcs, header, nv, P

    def run(self, fname, args):
        f = self.funcs[fname]
        env = dict(zip(f.args, args))
        oname, osize = f.output
        env[oname] = [D.const(0.0, self.nv, self.P) for _ in range(osize)]
        self.block(f.body, env, {})
        return env[oname]

    def c(self, x):
        return D.const(x, self.nv, self.P)

    def ival(self, node, loops):
        if isinstance(node, ast.Constant):
            return int(node.value)
        if isinstance(node, ast.Name):
            return loops[node.id]
        if isinstance(node, ast.BinOp):
            a, b = self.ival(node.left, loops), self.ival(node.right, loops)
            return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b}[type(node.op)]
        raise ValueError("bad index")

    def ev(self, node, env, loops):
        if isinstance(node, ast.Constant):
            return self.c(float(node.value))
        if isinstance(node, ast.Name):
            if node.id == "pi":
                return self.c(np.pi)
            if node.id in loops:
                return self.c(float(loops[node.id]))
            return env[node.id]
        if isinstance(node, ast.UnaryOp):
            x = self.ev(node.operand, env, loops)
            return -x if isinstance(node.op, ast.USub) else x
        if isinstance(node, ast.BinOp):
            a, b = self.ev(node.left, env, loops), self.ev(node.right, env, loops)
            if isinstance(node.op, ast.Pow):
                return _pow(a, b)
            if isinstance(a, list) and isinstance(b, list) and isinstance(node.op, ast.Mult):
                n, k, m = len(a), len(b), len(b[0])
                return [[self._sum([a[i][t] * b[t][j] for t in range(k)]) for j in range(m)] for i in range(n)]
            return {ast.Add: a.__add__, ast.Sub: a.__sub__, ast.Mult: a.__mul__,
                    ast.Div: a.__truediv__}[type(node.op)](b)
        if isinstance(node, ast.Subscript):
            ids = []
            base = node
            while isinstance(base, ast.Subscript):
                ids.insert(0, self.ival(base.slice, loops))
                base = base.value
            val = env[base.id]
            for i in ids:
                val = val[i]
            return val
        if isinstance(node, ast.Call):
            fn = node.func.id
            if fn in MATH:
                return MATH[fn](self.ev(node.args[0], env, loops))
            if fn == "pow":
                return _pow(self.ev(node.args[0], env, loops), self.ev(node.args[1], env, loops))
            if fn == "mul":
                return self.ev(node.args[0], env, loops) * self.ev(node.args[1], env, loops)
            if fn in ("Expression", "integer", "RealDouble"):
                return self.ev(node.args[0], env, loops)
            if fn == "inv":
                return _inv(self.ev(node.args[0], env, loops), self.nv, self.P)
            if fn in self.funcs:
                return self.run(fn, [env[a.id] for a in node.args])
        raise ValueError("unsupported expression %s" % ast.dump(node))

    def _sum(self, terms):
        acc = terms[0]
        for t in terms[1:]:
            acc = acc + t
        return acc

    def block(self, body, env, loops):
        for st in body:
            k = st[0]
            if k == "size":
                continue
            if k == "fill":
                v = env[st[1]]
                if isinstance(v[0], list):
                    env[st[1]] = [[self.c(st[2]) for _ in row] for row in v]
                else:
                    env[st[1]] = [self.c(st[2]) for _ in v]
            elif k == "matrix":
                env[st[1]] = [[self.c(0.0) for _ in range(st[3])] for _ in range(st[2])]
            elif k == "vector":
                env[st[1]] = [self.c(0.0) for _ in range(st[2])]
            elif k == "for":
                for i in range(st[2], st[3] + 1):
                    self.block(st[4], env, dict(loops, **{st[1]: i}))
            elif k == "assign":
                lhs, val = st[1], self.ev(st[2], env, loops)
                if lhs[0] == "name":
                    if isinstance(val, list) and lhs[1] in env and isinstance(env[lhs[1]], list):
                        # matrix assignment keeps the declared shape
                        env[lhs[1]] = val
                    else:
                        env[lhs[1]] = val
                else:
                    ids = [self.ival(i, loops) for i in lhs[2]]
                    tgt = env[lhs[1]]
                    for i in ids[:-1]:
                        tgt = tgt[i]
                    tgt[ids[-1]] = val


class DSLModel:
    """Oracle model (the interface of oracle/models.py) defined by a pdemodel.txt."""

    def __init__(self, path_or_text, nd=None, ncu=None):
        text = open(path_or_text).read() if "\n" not in path_or_text else path_or_text
        self.header, self.funcs = parse(text)
        self.nd, self.ncu, self.nc, self.nco = model_sizes(self.header)
        if nd is not None and nd != self.nd or ncu is not None and ncu != self.ncu:
            raise ValueError("model file sizes differ from the problem's")
        self.nbc = self.funcs["FbouHdg"].output[1] // self.ncu
        self.ncw = int(self.header["vectors"].get("w", 0))
        if self.ncw and "Sourcew" not in self.funcs:
            raise ValueError("a model with w(ncw > 0) needs a Sourcew function")
        self.name = "dsl"

    def _args(self, fname, P, nv, x, uq, v, param, time, uh=None, nl=None, tau=None, w=None, woff=0):
        """Dual-number arguments: uq seeded in directions 0..nc-1, uhat in nc..nc+ncu-1, w (when
        given) in woff..woff+ncw-1."""
        vec = self.header["vectors"]
        out = []
        for a in self.funcs[fname].args:
            if a == "x":
                out.append([D(x[:, i].copy(), np.zeros((nv, P))) for i in range(x.shape[1])])
            elif a == "uq":
                out.append([D(uq[:, k].copy(), self._seed(nv, P, k)) for k in range(self.nc)])
            elif a == "uhat":
                out.append([D(uh[:, k].copy(), self._seed(nv, P, self.nc + k)) for k in range(self.ncu)])
            elif a == "v":
                nvv = max(vec.get("v", 0), 1)
                out.append([D.const(v[:, i] if v is not None and i < v.shape[1] else 0.0, nv, P) for i in range(nvv)])
            elif a == "n":
                out.append([D.const(nl[:, i], nv, P) for i in range(nl.shape[1])])
            elif a == "tau":
                out.append([D.const(tau[i], nv, P) for i in range(max(vec.get("tau", 1), 1))])
            elif a == "mu":
                out.append([D.const(param[i] if i < len(param) else 0.0, nv, P)
                            for i in range(max(vec.get("mu", 0), 1))])
            elif a == "w" and w is not None:
                out.append([D(w[:, i].copy(), self._seed(nv, P, woff + i)) for i in range(self.ncw)])
            elif a in ("w", "eta"):
                out.append([D.const(0.0, nv, P) for _ in range(max(vec.get(a, 0), 1))])
            elif a == "t":
                out.append(D.const(time, nv, P))
            else:
                raise ValueError("unsupported argument %s" % a)
        return out

    def _seed(self, nv, P, k):
        g = np.zeros((nv, P))
        if k < nv:
            g[k] = 1.0
        return g

    def flux(self, x, uq, v, param, time, w=None):
        """f [P, ncu, nd], f_uq [P, ncu, nd, nc] (and f_w [P, ncu, nd, ncw] when w is given)."""
        P, nd, ncu, nc = x.shape[0], self.nd, self.ncu, self.nc
        nv = nc + (self.ncw if w is not None else 0)
        F = _Interp(self.funcs, self.header, nv, P).run(
            "Flux", self._args("Flux", P, nv, x, uq, v, param, time, w=w, woff=nc))
        f = np.stack([F[c + ncu * d].v for d in range(nd) for c in range(ncu)], 1).reshape(P, nd, ncu)
        g = np.stack([F[c + ncu * d].g for d in range(nd) for c in range(ncu)], 0).reshape(nd, ncu, nv, P)
        g = g.transpose(3, 1, 0, 2)
        if w is None:
            return f.transpose(0, 2, 1), g
        return f.transpose(0, 2, 1), g[..., :nc], g[..., nc:]

    def source(self, x, uq, v, param, time, w=None):
        """s [P, ncu], s_uq [P, ncu, nc] (and s_w [P, ncu, ncw] when w is given)."""
        P, ncu, nc = x.shape[0], self.ncu, self.nc
        nv = nc + (self.ncw if w is not None else 0)
        S = _Interp(self.funcs, self.header, nv, P).run(
            "Source", self._args("Source", P, nv, x, uq, v, param, time, w=w, woff=nc))
        s = np.stack([np.broadcast_to(S[c].v, (P,)) for c in range(ncu)], 1)
        g = np.stack([S[c].g for c in range(ncu)], 0).transpose(2, 0, 1)          # [P, ncu, nv]
        if w is None:
            return s, g
        return s, g[..., :nc], g[..., nc:]

    def sourcew(self, x, uq, v, w, param, time):
        """Sourcew (the w-equation's right-hand side): sw [P, ncw], sw_uq [P, ncw, nc],
        sw_w [P, ncw, ncw]."""
        P, nc, ncw = x.shape[0], self.nc, self.ncw
        nv = nc + ncw
        S = _Interp(self.funcs, self.header, nv, P).run(
            "Sourcew", self._args("Sourcew", P, nv, x, uq, v, param, time, w=w, woff=nc))
        sw = np.stack([np.broadcast_to(S[c].v, (P,)) for c in range(ncw)], 1)
        g = np.stack([S[c].g for c in range(ncw)], 0).transpose(2, 0, 1)        # [P, ncw, nv]
        return sw, g[..., :nc], g[..., nc:]

    def tdfunc(self, x, uq, v, param, time):
        """Tdfunc values [N, ncu] (the time-derivative coefficient; ones when the file has no
        Tdfunc).  Only the values enter the equations (uequation.cpp:129-145)."""
        P, ncu, nc = x.shape[0], self.ncu, self.nc
        if "Tdfunc" not in self.funcs:
            return np.ones((P, ncu))
        T = _Interp(self.funcs, self.header, nc, P).run("Tdfunc", self._args("Tdfunc", P, nc, x, uq, v, param, time))
        return np.stack([np.broadcast_to(T[c].v, (P,)) for c in range(ncu)], 1)

    def fbou(self, ib, x, uq, v, uh, nl, tau, param, time, w=None):
        """fb [P, ncu], fb_uq, fb_uh (and fb_w [P, ncu, ncw] when w is given)."""
        P, ncu, nc = x.shape[0], self.ncu, self.nc
        nv = nc + ncu + (self.ncw if w is not None else 0)
        Fb = _Interp(self.funcs, self.header, nv, P).run(
            "FbouHdg", self._args("FbouHdg", P, nv, x, uq, v, param, time, uh=uh, nl=nl, tau=tau, w=w,
                                  woff=nc + ncu))
        o = (ib - 1 if 1 <= ib <= self.nbc else 0) * ncu
        fb = np.stack([np.broadcast_to(Fb[o + c].v, (P,)) for c in range(ncu)], 1)
        g = np.stack([Fb[o + c].g for c in range(ncu)], 0)    # [ncu, nv, P]
        out = fb, g[:, :nc].transpose(2, 0, 1), g[:, nc:nc + ncu].transpose(2, 0, 1)
        if w is None:
            return out
        return out + (g[:, nc + ncu:].transpose(2, 0, 1),)
