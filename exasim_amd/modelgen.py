"""Model bridge: text2code's pdemodel.txt language -> HIP model structs.

The reference generates its pointwise model functions (HdgFlux, HdgSource,
HdgFbou, ...) from a pdemodel.txt with text2code (text2code/CodeGenerator.cpp,
libpdemodel.hpp:1-41).  This module reads the same language and writes a
C++ struct with the interface of the hand-written models in
``csrc/models.hpp`` (``flux<JAC>``, ``source<JAC>``, ``fbou<JAC>``).  The
Jacobians come from forward-mode dual numbers (``csrc/dual.hpp``) instead of
symbolic differentiation, so matrix inverses and loops in a model cost no
expression swell.

    python -m exasim_amd.modelgen path/to/pdemodel.txt --name ns2d

writes ``csrc/generated/ns2d.hpp`` and refreshes ``csrc/generated/registry.hpp``
(the sizes nd, ncu and nv come from the ``vectors`` line);
``python -m exasim_amd.build`` then compiles it into libexahip.so, and
``exa_create(..., model="ns2d", ...)`` selects it.  The build regenerates every
model under ``exasim_amd/models/`` (``--all``), named by its file name.

Language subset (the one the reference's examples use): header lines
``scalars``, ``vectors``, ``jacobian``, ``hessian``, ``batch``, ``outputs``;
``function Name(args) ... end`` blocks with ``output_size(a) = k;``,
``zeros(a);``, ``ones(a);``, ``matrix M(r,c);``, ``vector v(k);``,
assignments to names, ``a[i]`` and ``M[i][j]``, ``for i in a:b ... endfor``,
calls of other model functions (``f = Flux(...)``), ``inv(M)``, matrix
products ``A * B``, and the functions sin, cos, tan, exp, log, sqrt, tanh,
sinh, cosh, atan, abs, pow, mul, plus ``pi`` and ``SymEngine::integer(k)``.
"""
import argparse
import ast
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
GEN_DIR = os.path.join(HERE, "csrc", "generated")

MATH1 = {"sin", "cos", "tan", "exp", "log", "sqrt", "tanh", "sinh", "cosh", "atan", "abs", "fabs"}


class ModelSyntaxError(ValueError):
    pass


# ---------------------------------------------------------------------------
# Parser: pdemodel.txt -> header dict + {name: Function}
# ---------------------------------------------------------------------------

class Function:
    def __init__(self, name, args, body):
        self.name, self.args, self.body = name, args, body
        self.output = None
        for st in body:
            if st[0] == "size":
                self.output = (st[1], st[2])
        if self.output is None:
            raise ModelSyntaxError("function %s has no output_size" % name)


def _expr(src, lineno):
    s = src.strip()
    s = re.sub(r"SymEngine::integer\(([^()]*)\)", r"(\1)", s)
    s = s.replace("SymEngine::", "")
    s = s.replace("^", "**")
    try:
        return ast.parse(s, mode="eval").body
    except SyntaxError as e:
        raise ModelSyntaxError("line %d: cannot parse expression %r (%s)" % (lineno, src, e))


def _lhs(src, lineno):
    node = _expr(src, lineno)
    if isinstance(node, ast.Name):
        return ("name", node.id)
    if isinstance(node, ast.Subscript):
        idx = []
        while isinstance(node, ast.Subscript):
            idx.insert(0, node.slice)
            node = node.value
        if not isinstance(node, ast.Name) or len(idx) > 2:
            raise ModelSyntaxError("line %d: bad assignment target %r" % (lineno, src))
        return ("index", node.id, idx)
    raise ModelSyntaxError("line %d: bad assignment target %r" % (lineno, src))


def _statements(lines, i, end_words):
    """Parse statements until a line whose first word is in end_words."""
    body = []
    while i < len(lines):
        lineno, text = lines[i]
        word = text.split()[0] if text.split() else ""
        if word in end_words:
            return body, i
        m = re.match(r"output_size\((\w+)\)\s*=\s*(\d+)\s*;?$", text)
        if m:
            body.append(("size", m.group(1), int(m.group(2))))
            i += 1
            continue
        m = re.match(r"(zeros|ones)\((\w+)\)\s*;?$", text)
        if m:
            body.append(("fill", m.group(2), 0.0 if m.group(1) == "zeros" else 1.0))
            i += 1
            continue
        m = re.match(r"matrix\s+(\w+)\((\d+)\s*,\s*(\d+)\)\s*;?$", text)
        if m:
            body.append(("matrix", m.group(1), int(m.group(2)), int(m.group(3))))
            i += 1
            continue
        m = re.match(r"vector\s+(\w+)\((\d+)\)\s*;?$", text)
        if m:
            body.append(("vector", m.group(1), int(m.group(2))))
            i += 1
            continue
        m = re.match(r"for\s+(\w+)\s+in\s+(-?\d+)\s*:\s*(-?\d+)\s*$", text)
        if m:
            inner, j = _statements(lines, i + 1, {"endfor"})
            if j >= len(lines):
                raise ModelSyntaxError("line %d: for without endfor" % lineno)
            body.append(("for", m.group(1), int(m.group(2)), int(m.group(3)), inner))
            i = j + 1
            continue
        m = re.match(r"(.+?)\s*=\s*(.+?)\s*;?$", text)
        if m and "==" not in text:
            body.append(("assign", _lhs(m.group(1), lineno), _expr(m.group(2), lineno), lineno))
            i += 1
            continue
        raise ModelSyntaxError("line %d: cannot parse %r" % (lineno, text))
    return body, i


def parse(text):
    """-> (header, functions).  header: {'vectors': {name: size}, 'scalars': [...], ...}."""
    lines = []
    for k, raw in enumerate(text.splitlines(), 1):
        t = raw.split("//")[0].split("#")[0].strip()
        if t:
            lines.append((k, t))
    header = {"vectors": {}, "scalars": [], "jacobian": [], "outputs": []}
    funcs = {}
    i = 0
    while i < len(lines):
        lineno, t = lines[i]
        word = t.split()[0]
        if word == "function":
            m = re.match(r"function\s+(\w+)\s*\(([^)]*)\)\s*$", t)
            if not m:
                raise ModelSyntaxError("line %d: bad function header" % lineno)
            body, j = _statements(lines, i + 1, {"end"})
            if j >= len(lines):
                raise ModelSyntaxError("line %d: function without end" % lineno)
            args = [a.strip() for a in m.group(2).split(",") if a.strip()]
            funcs[m.group(1)] = Function(m.group(1), args, body)
            i = j + 1
            continue
        rest = t[len(word):].strip()
        if word == "vectors":
            for name, size in re.findall(r"(\w+)\((\d+)\)", rest):
                header["vectors"][name] = int(size)
        elif word == "scalars":
            header["scalars"] = [s.strip() for s in rest.split(",") if s.strip()]
        elif word in ("jacobian", "batch", "outputs", "hessian"):
            header[word] = [s.strip() for s in rest.split(",") if s.strip()]
        else:
            raise ModelSyntaxError("line %d: unknown statement %r" % (lineno, t))
        i += 1
    for need in ("Flux", "Source", "FbouHdg"):
        if need not in funcs:
            raise ModelSyntaxError("the model has no %s function" % need)
    return header, funcs


def model_sizes(header):
    """(nd, ncu, nc, nco) from the vectors line: x(nd), uq(nc), uhat(ncu), v(nco)."""
    v = header["vectors"]
    nd, nc, ncu = v.get("x", 0), v.get("uq", 0), v.get("uhat", 0)
    if nd < 1 or ncu < 1 or nc < ncu:
        raise ModelSyntaxError("vectors must declare x(nd), uq(nc) and uhat(ncu)")
    return nd, ncu, nc, v.get("v", 0)


# ---------------------------------------------------------------------------
# C++ emitter
# ---------------------------------------------------------------------------

_BINOP = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}


class _Emitter:
    """Translates one Function to a templated C++ function on scalar type S.
    Every local is an S (a dual number for the Jacobian instantiations).  Model
    names get a trailing underscore in C++ (``kappa`` -> ``kappa_``), so no name in a
    model file can collide with a C++ keyword or with the emitter's own names."""

    def __init__(self, funcs, fname):
        self.funcs = funcs
        self.f = funcs[fname]
        self.locals = {}      # name -> ('scalar',) | ('array', n) | ('matrix', r, c)
        self.decls = []       # declarations, hoisted to the top of the function
        self.out = []
        self.loopvars = set()

    def expr(self, node):
        if isinstance(node, ast.Constant):
            if isinstance(node.value, (int, float)):
                return "S(%r)" % float(node.value)
            raise ModelSyntaxError("unsupported constant %r" % node.value)
        if isinstance(node, ast.Name):
            if node.id == "pi":
                return "S(M_PI)"
            if node.id in self.loopvars:
                return "S(static_cast<double>(%s_))" % node.id
            kind = self.locals.get(node.id)
            if kind is None:
                raise ModelSyntaxError("%s: unknown name %s" % (self.f.name, node.id))
            return node.id + "_"
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
            return "(-%s)" % self.expr(node.operand)
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.UAdd):
            return self.expr(node.operand)
        if isinstance(node, ast.BinOp):
            if isinstance(node.op, ast.Pow):
                return "dpow(%s, %s)" % (self.expr(node.left), self.expr(node.right))
            op = _BINOP.get(type(node.op))
            if op is None:
                raise ModelSyntaxError("unsupported operator in %s" % self.f.name)
            return "(%s %s %s)" % (self.expr(node.left), op, self.expr(node.right))
        if isinstance(node, ast.Subscript):
            return self.index(node)
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
            fn = node.func.id
            args = node.args
            if fn in MATH1:
                return "d%s(%s)" % ("abs" if fn == "fabs" else fn, self.expr(args[0]))
            if fn == "pow":
                return "dpow(%s, %s)" % (self.expr(args[0]), self.expr(args[1]))
            if fn == "mul":
                return "(%s * %s)" % (self.expr(args[0]), self.expr(args[1]))
            if fn in ("Expression", "integer", "RealDouble"):
                return self.expr(args[0])
            raise ModelSyntaxError("%s: unsupported function %s" % (self.f.name, fn))
        raise ModelSyntaxError("%s: unsupported expression %s" % (self.f.name, ast.dump(node)))

    def idx(self, node):
        if isinstance(node, ast.Constant):
            return str(int(node.value))
        if isinstance(node, ast.Name) and node.id in self.loopvars:
            return node.id + "_"
        if isinstance(node, ast.BinOp) and type(node.op) in (ast.Add, ast.Sub, ast.Mult):
            return "(%s %s %s)" % (self.idx(node.left), _BINOP[type(node.op)], self.idx(node.right))
        raise ModelSyntaxError("%s: index must be an integer expression" % self.f.name)

    def index(self, node):
        ids = []
        base = node
        while isinstance(base, ast.Subscript):
            ids.insert(0, base.slice)
            base = base.value
        name = base.id
        kind = self.locals.get(name)
        if kind is None:
            raise ModelSyntaxError("%s: unknown array %s" % (self.f.name, name))
        if kind[0] == "matrix":
            if len(ids) != 2:
                raise ModelSyntaxError("%s: matrix %s needs two indices" % (self.f.name, name))
            return "%s_[%s][%s]" % (name, self.idx(ids[0]), self.idx(ids[1]))
        return "%s_[%s]" % (name, self.idx(ids[0]))

    def call_model(self, target, call, ind):
        g = self.funcs.get(call.func.id)
        if g is None:
            raise ModelSyntaxError("%s: unknown function %s" % (self.f.name, call.func.id))
        names = []
        for a in call.args:
            if not isinstance(a, ast.Name) or a.id not in self.locals:
                raise ModelSyntaxError("%s: arguments of %s must be names" % (self.f.name, g.name))
            names.append(a.id + "_")
        size = g.output[1]
        self.declare(target, ("array", size), ind)
        self.out.append(ind + "fn_%s<S>(%s, %s_);" % (g.name, ", ".join(names), target))

    def declare(self, name, kind, ind):
        old = self.locals.get(name)
        if old is not None:
            if old != kind:
                raise ModelSyntaxError("%s: %s changes shape" % (self.f.name, name))
            return
        self.locals[name] = kind
        if kind[0] == "scalar":
            self.decls.append("  S %s_;" % name)
        elif kind[0] == "array":
            self.decls.append("  S %s_[%d];" % (name, kind[1]))
        else:
            self.decls.append("  S %s_[%d][%d];" % (name, kind[1], kind[2]))

    def matrix_expr(self, node):
        """inv(M) or A * B with matrix operands -> ('inv', M) / ('mul', A, B) or None."""
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id == "inv":
            a = node.args[0]
            if isinstance(a, ast.Name) and self.locals.get(a.id, ("",))[0] == "matrix":
                return ("inv", a.id)
        if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Mult) and \
                isinstance(node.left, ast.Name) and isinstance(node.right, ast.Name):
            kl, kr = self.locals.get(node.left.id, ("",)), self.locals.get(node.right.id, ("",))
            if kl[0] == "matrix" and kr[0] == "matrix":
                return ("mul", node.left.id, node.right.id)
        return None

    def stmts(self, body, ind):
        for st in body:
            kind = st[0]
            if kind == "size":
                continue
            if kind == "fill":
                k = self.locals[st[1]]
                n = k[1] if k[0] == "array" else k[1] * k[2]
                tgt = st[1] + "_" if k[0] == "array" else "(&%s_[0][0])" % st[1]
                self.out.append(ind + "for (int i_ = 0; i_ < %d; ++i_) %s[i_] = S(%r);" % (n, tgt, st[2]))
            elif kind == "matrix":
                self.declare(st[1], ("matrix", st[2], st[3]), ind)
            elif kind == "vector":
                self.declare(st[1], ("array", st[2]), ind)
            elif kind == "for":
                var, a, b, inner = st[1], st[2], st[3], st[4]
                self.out.append(ind + "for (int %s_ = %d; %s_ <= %d; ++%s_) {" % (var, a, var, b, var))
                self.loopvars.add(var)
                self.stmts(inner, ind + "  ")
                self.loopvars.discard(var)
                self.out.append(ind + "}")
            elif kind == "assign":
                lhs, rhs = st[1], st[2]
                if isinstance(rhs, ast.Call) and isinstance(rhs.func, ast.Name) and rhs.func.id in self.funcs:
                    if lhs[0] != "name":
                        raise ModelSyntaxError("%s: a function result must go to a name" % self.f.name)
                    self.call_model(lhs[1], rhs, ind)
                    continue
                mx = self.matrix_expr(rhs)
                if mx is not None:
                    if lhs[0] != "name" or self.locals.get(lhs[1], ("",))[0] != "matrix":
                        raise ModelSyntaxError("%s: matrix result needs a declared matrix" % self.f.name)
                    tk = self.locals[lhs[1]]
                    if mx[0] == "inv":
                        self.out.append(ind + "mat_inv<S, %d>(%s_, %s_);" % (tk[1], mx[1], lhs[1]))
                    else:
                        ka = self.locals[mx[1]]
                        self.out.append(ind + "mat_mul<S, %d, %d, %d>(%s_, %s_, %s_);"
                                        % (ka[1], ka[2], tk[2], mx[1], mx[2], lhs[1]))
                    continue
                val = self.expr(rhs)
                if lhs[0] == "name":
                    if self.locals.get(lhs[1], ("scalar",))[0] != "scalar":
                        raise ModelSyntaxError("%s: cannot assign a scalar to array %s" % (self.f.name, lhs[1]))
                    self.declare(lhs[1], ("scalar",), ind)
                    self.out.append(ind + "%s_ = %s;" % (lhs[1], val))
                else:
                    node = ast.Subscript(value=ast.Name(id=lhs[1]), slice=lhs[2][0])
                    if len(lhs[2]) == 2:
                        node = ast.Subscript(value=node, slice=lhs[2][1])
                    self.out.append(ind + "%s = %s;" % (self.index(node), val))

    def emit(self, header):
        f = self.f
        params = []
        for a in f.args:
            if a in header["scalars"]:
                self.locals[a] = ("scalar",)
                params.append("S %s_" % a)
            else:
                size = header["vectors"].get(a)
                if size is None:
                    raise ModelSyntaxError("%s: unknown argument %s" % (f.name, a))
                self.locals[a] = ("array", max(size, 1))
                params.append("const S* %s_" % a)
        oname, osize = f.output
        self.locals[oname] = ("array", osize)
        params.append("S* %s_" % oname)
        self.stmts(f.body, "  ")
        head = ["template <typename S>", "__device__ __forceinline__ void fn_%s(%s) {" % (f.name, ", ".join(params))]
        return "\n".join(head + self.decls + self.out + ["}"])


def _call_order(funcs, roots):
    """Functions reachable from roots, callees first."""
    order, seen = [], set()

    def visit(name, stack):
        if name in seen:
            return
        if name in stack:
            raise ModelSyntaxError("recursive model function %s" % name)
        stack = stack | {name}
        for st in _walk(funcs[name].body):
            if st[0] == "assign" and isinstance(st[2], ast.Call) and isinstance(st[2].func, ast.Name) \
                    and st[2].func.id in funcs:
                visit(st[2].func.id, stack)
        seen.add(name)
        order.append(name)
    for r in roots:
        visit(r, set())
    return order


def _walk(body):
    for st in body:
        yield st
        if st[0] == "for":
            yield from _walk(st[4])


def emit_cpp(text, name, source="a pdemodel.txt"):
    header, funcs = parse(text)
    nd, ncu, nc, nco = model_sizes(header)
    ncw = int(header["vectors"].get("w", 0))
    if ncw and "Sourcew" not in funcs:
        raise ModelSyntaxError("a model with w(ncw > 0) needs a Sourcew function (the w equation)")
    if ncw > 5:
        raise ModelSyntaxError("the w equation supports at most 5 components")
    if ncw and funcs["Sourcew"].output[1] != ncw:
        raise ModelSyntaxError("Sourcew must have ncw outputs")
    has_td = "Tdfunc" in funcs
    if has_td and funcs["Tdfunc"].output[1] != ncu:
        raise ModelSyntaxError("Tdfunc must have ncu outputs")
    nbc = funcs["FbouHdg"].output[1] // ncu
    if funcs["FbouHdg"].output[1] != nbc * ncu:
        raise ModelSyntaxError("FbouHdg output size must be a multiple of ncu")
    if funcs["Flux"].output[1] != ncu * nd or funcs["Source"].output[1] != ncu:
        raise ModelSyntaxError("Flux must have ncu*nd and Source ncu outputs")
    ident = re.sub(r"\W", "_", name)
    parts = ["// Generated by exasim_amd/modelgen.py from %s -- do not edit." % source,
             "#pragma once", "#include \"dual.hpp\"", "", "namespace exa {", "namespace gen_%s {" % ident, ""]
    for fname in _call_order(funcs, ["Flux", "Source", "FbouHdg"] + (["Tdfunc"] if has_td else []) +
                             (["Sourcew"] if ncw else [])):
        parts.append(_Emitter(funcs, fname).emit(header))
        parts.append("")
    vec = header["vectors"]

    def argv(fn, kind, wname="W0"):
        """call arguments of a generated function from the kernel-side names."""
        m = {"x": "X", "uq": "U", "v": "V", "w": wname, "uhat": "UH", "n": "NL", "tau": "TAU", "mu": "MU",
             "eta": "ETA", "t": "T"}
        out = []
        for a in funcs[fn].args:
            if a not in m:
                raise ModelSyntaxError("%s: argument %s is not supported" % (fn, a))
            out.append(m[a])
        return ", ".join(out)

    sz = lambda a: max(vec.get(a, 0), 1)
    parts.append("}  // namespace gen_%s" % ident)
    parts.append("")
    nv = vec.get("v", 0)
    parts.append("// %s: nd = %d, ncu = %d, nc = %d, nv = %d, %d boundary conditions" % (name, nd, ncu, nc, nv, nbc))
    parts.append("struct Gen_%s {" % ident)
    parts.append("  static constexpr int ND = %d, NCU = %d, NC = %d, NV = %d, NBC = %d, NCW = %d;"
                 % (nd, ncu, nc, nv, nbc, ncw))
    prologue = """    S X[{nx}], V[{nv}], W0[{nw}], MU[{nmu}], ETA[{neta}], TAU[{ntau}], NL[{nn}];
    for (int i = 0; i < {nx}; ++i) X[i] = S(x[i]);
    for (int i = 0; i < {nv}; ++i) V[i] = S(i < {nvr} ? v[i] : 0.0);
    for (int i = 0; i < {nw}; ++i) W0[i] = S(0.0);
    for (int i = 0; i < {nmu}; ++i) MU[i] = S(p[i]);
    for (int i = 0; i < {neta}; ++i) ETA[i] = S(0.0);
    for (int i = 0; i < {ntau}; ++i) TAU[i] = S(0.0);
    for (int i = 0; i < {nn}; ++i) NL[i] = S(0.0);
    const S T = S(t);""".format(nx=sz("x"), nv=sz("v"), nvr=vec.get("v", 0), nw=sz("w"), nmu=sz("mu"), neta=sz("eta"),
                                ntau=sz("tau"), nn=sz("n"))
    # flux
    parts.append("  template <bool JAC>")
    parts.append("  __device__ static void flux(const double* x, const double* uq, const double* v, const double* p,")
    parts.append("                              double t, double* f, double* f_uq) {")
    parts.append("    using S = typename std::conditional<JAC, Dual<NC>, Dual<0>>::type;")
    parts.append(prologue)
    parts.append("    S U[NC];")
    parts.append("    for (int k = 0; k < NC; ++k) { U[k] = S(uq[k]); if (JAC) U[k].seed(k); }")
    parts.append("    S F[NCU * ND];")
    parts.append("    gen_%s::fn_Flux<S>(%s, F);" % (ident, argv("Flux", "flux")))
    parts.append("    for (int i = 0; i < NCU * ND; ++i) {")
    parts.append("      f[i] = F[i].v;")
    parts.append("      if (JAC) for (int k = 0; k < NC; ++k) f_uq[i + NCU * ND * k] = F[i].deriv(k);")
    parts.append("    }")
    parts.append("  }")
    # source
    parts.append("  template <bool JAC>")
    parts.append("  __device__ static void source(const double* x, const double* uq, const double* v, const double* p,")
    parts.append("                                double t, double* s, double* s_uq) {")
    parts.append("    using S = typename std::conditional<JAC, Dual<NC>, Dual<0>>::type;")
    parts.append(prologue)
    parts.append("    S U[NC];")
    parts.append("    for (int k = 0; k < NC; ++k) { U[k] = S(uq[k]); if (JAC) U[k].seed(k); }")
    parts.append("    S F[NCU];")
    parts.append("    gen_%s::fn_Source<S>(%s, F);" % (ident, argv("Source", "source")))
    parts.append("    for (int i = 0; i < NCU; ++i) {")
    parts.append("      s[i] = F[i].v;")
    parts.append("      if (JAC) for (int k = 0; k < NC; ++k) s_uq[i + NCU * k] = F[i].deriv(k);")
    parts.append("    }")
    parts.append("  }")
    if has_td:  # the time-derivative coefficient (values only: its derivative is not used)
        parts.append("  __device__ static void tdfunc(const double* x, const double* uq, const double* v, const double* p,")
        parts.append("                                double t, double* tcoef) {")
        parts.append("    using S = Dual<0>;")
        parts.append(prologue)
        parts.append("    S U[NC];")
        parts.append("    for (int k = 0; k < NC; ++k) U[k] = S(uq[k]);")
        parts.append("    S F[NCU];")
        parts.append("    gen_%s::fn_Tdfunc<S>(%s, F);" % (ident, argv("Tdfunc", "tdfunc")))
        parts.append("    for (int i = 0; i < NCU; ++i) tcoef[i] = F[i].v;")
        parts.append("  }")
    # fbou: the Jacobian in passes of KB derivative directions (of NC + NCU), so the
    # dual numbers of a large boundary function stay small
    parts.append("  static constexpr int NT = NC + NCU, KB = NT < 4 ? NT : 4;")
    parts.append("  template <int K0, int W>")
    parts.append("  __device__ static void fbou_pass(int ib, const double* x, const double* uq, const double* v,")
    parts.append("                                   const double* uh, const double* nl, const double* tau,")
    parts.append("                                   const double* p, double t, double* fb, double* fb_uq,")
    parts.append("                                   double* fb_uh) {")
    parts.append("    using S = Dual<W>;")
    parts.append(prologue)
    parts.append("    for (int i = 0; i < %d; ++i) TAU[i] = S(tau[i]);" % sz("tau"))
    parts.append("    for (int i = 0; i < %d; ++i) NL[i] = S(i < ND ? nl[i] : 0.0);" % sz("n"))
    parts.append("    S U[NC], UH[NCU];")
    parts.append("#pragma unroll")
    parts.append("    for (int k = 0; k < NC; ++k) U[k] = S(uq[k]);")
    parts.append("#pragma unroll")
    parts.append("    for (int k = 0; k < NCU; ++k) UH[k] = S(uh[k]);")
    parts.append("#pragma unroll")
    parts.append("    for (int i = 0; i < W; ++i) {  // seed directions K0 .. K0 + W - 1")
    parts.append("      const int k = K0 + i;")
    parts.append("      if (k < NC) U[k < NC ? k : 0].d[i] = 1.0;")
    parts.append("      else if (k < NT) UH[k < NT ? k - NC : 0].d[i] = 1.0;")
    parts.append("    }")
    parts.append("    S F[NCU * NBC];")
    parts.append("    gen_%s::fn_FbouHdg<S>(%s, F);" % (ident, argv("FbouHdg", "fbou")))
    parts.append("    // condition ib (1-based; out of range -> 1); compile-time indices keep F in registers")
    parts.append("    const int sel = (ib >= 1 && ib <= NBC) ? ib - 1 : 0;")
    parts.append("#pragma unroll")
    parts.append("    for (int b = 0; b < NBC; ++b) {")
    parts.append("      if (b != sel) continue;")
    parts.append("#pragma unroll")
    parts.append("      for (int r = 0; r < NCU; ++r) {")
    parts.append("        fb[r] = F[b * NCU + r].v;")
    parts.append("#pragma unroll")
    parts.append("        for (int i = 0; i < W; ++i) {")
    parts.append("          const int k = K0 + i;")
    parts.append("          if (k < NC) fb_uq[r + NCU * (k < NC ? k : 0)] = F[b * NCU + r].d[i];")
    parts.append("          else if (k < NT) fb_uh[r + NCU * (k < NT ? k - NC : 0)] = F[b * NCU + r].d[i];")
    parts.append("        }")
    parts.append("      }")
    parts.append("    }")
    parts.append("    if constexpr (W > 0 && K0 + W < NT)")
    parts.append("      fbou_pass<K0 + W, W>(ib, x, uq, v, uh, nl, tau, p, t, fb, fb_uq, fb_uh);")
    parts.append("  }")
    parts.append("  template <bool JAC>")
    parts.append("  __device__ static void fbou(int ib, const double* x, const double* uq, const double* v,")
    parts.append("                              const double* uh, const double* nl, const double* tau,")
    parts.append("                              const double* p, double t, double* fb, double* fb_uq, double* fb_uh) {")
    parts.append("    if constexpr (JAC) fbou_pass<0, KB>(ib, x, uq, v, uh, nl, tau, p, t, fb, fb_uq, fb_uh);")
    parts.append("    else fbou_pass<0, 0>(ib, x, uq, v, uh, nl, tau, p, t, fb, fb_uq, fb_uh);")
    parts.append("  }")
    if ncw:
        parts += _emit_w_methods(ident, prologue, argv, sz)
    parts.append("};")
    parts.append("")
    parts.append("}  // namespace exa")
    return "\n".join(parts) + "\n", dict(name=name, ident=ident, nd=nd, ncu=ncu, nc=nc, nv=nv, nbc=nbc)


def _emit_w_methods(ident, prologue, argv, sz):
    """The w-equation forms (wequation.cpp:31-240) of a model with ncw > 0: flux_w, source_w,
    fbou_w take the point's w and give the derivatives in w as well (dual directions: uq, then w;
    for fbou_w: uq, uhat, w), and sourcew evaluates Sourcew with its derivatives in uq and w."""
    out = []
    pw = prologue + """
    S W[NCW];"""

    def head(name, args):
        out.append("  template <bool JAC>")
        out.append("  __device__ static void %s(%s) {" % (name, args))

    # flux_w
    head("flux_w", "const double* x, const double* uq, const double* v, const double* w, const double* p, "
                   "double t, double* f, double* f_uq, double* f_w")
    out.append("    using S = typename std::conditional<JAC, Dual<NC + NCW>, Dual<0>>::type;")
    out.append(pw)
    out.append("    for (int k = 0; k < NCW; ++k) { W[k] = S(w[k]); if (JAC) W[k].seed(NC + k); }")
    out.append("    S U[NC];")
    out.append("    for (int k = 0; k < NC; ++k) { U[k] = S(uq[k]); if (JAC) U[k].seed(k); }")
    out.append("    S F[NCU * ND];")
    out.append("    gen_%s::fn_Flux<S>(%s, F);" % (ident, argv("Flux", "flux", "W")))
    out.append("    for (int i = 0; i < NCU * ND; ++i) {")
    out.append("      f[i] = F[i].v;")
    out.append("      if (JAC) {")
    out.append("        for (int k = 0; k < NC; ++k) f_uq[i + NCU * ND * k] = F[i].deriv(k);")
    out.append("        for (int k = 0; k < NCW; ++k) f_w[i + NCU * ND * k] = F[i].deriv(NC + k);")
    out.append("      }")
    out.append("    }")
    out.append("  }")
    # source_w
    head("source_w", "const double* x, const double* uq, const double* v, const double* w, const double* p, "
                     "double t, double* s, double* s_uq, double* s_w")
    out.append("    using S = typename std::conditional<JAC, Dual<NC + NCW>, Dual<0>>::type;")
    out.append(pw)
    out.append("    for (int k = 0; k < NCW; ++k) { W[k] = S(w[k]); if (JAC) W[k].seed(NC + k); }")
    out.append("    S U[NC];")
    out.append("    for (int k = 0; k < NC; ++k) { U[k] = S(uq[k]); if (JAC) U[k].seed(k); }")
    out.append("    S F[NCU];")
    out.append("    gen_%s::fn_Source<S>(%s, F);" % (ident, argv("Source", "source", "W")))
    out.append("    for (int i = 0; i < NCU; ++i) {")
    out.append("      s[i] = F[i].v;")
    out.append("      if (JAC) {")
    out.append("        for (int k = 0; k < NC; ++k) s_uq[i + NCU * k] = F[i].deriv(k);")
    out.append("        for (int k = 0; k < NCW; ++k) s_w[i + NCU * k] = F[i].deriv(NC + k);")
    out.append("      }")
    out.append("    }")
    out.append("  }")
    # sourcew (always with its derivatives: the point Newton needs them)
    out.append("  __device__ static void sourcew(const double* x, const double* uq, const double* v, const double* w,")
    out.append("                                 const double* p, double t, double* sw, double* sw_uq, double* sw_w) {")
    out.append("    using S = Dual<NC + NCW>;")
    out.append(pw)
    out.append("    for (int k = 0; k < NCW; ++k) { W[k] = S(w[k]); W[k].seed(NC + k); }")
    out.append("    S U[NC];")
    out.append("    for (int k = 0; k < NC; ++k) { U[k] = S(uq[k]); U[k].seed(k); }")
    out.append("    S F[NCW];")
    out.append("    gen_%s::fn_Sourcew<S>(%s, F);" % (ident, argv("Sourcew", "sourcew", "W")))
    out.append("    for (int i = 0; i < NCW; ++i) {")
    out.append("      sw[i] = F[i].v;")
    out.append("      for (int k = 0; k < NC; ++k) sw_uq[i + NCW * k] = F[i].deriv(k);")
    out.append("      for (int k = 0; k < NCW; ++k) sw_w[i + NCW * k] = F[i].deriv(NC + k);")
    out.append("    }")
    out.append("  }")
    # fbou_w: one pass over all NC + NCU + NCW directions (the dual numbers stay within the limits
    # of models small enough to have a w equation here)
    head("fbou_w", "int ib, const double* x, const double* uq, const double* v, const double* w, "
                   "const double* uh, const double* nl, const double* tau, const double* p, double t, double* fb, "
                   "double* fb_uq, double* fb_uh, double* fb_w")
    out.append("    constexpr int NTW = NC + NCU + NCW;")
    out.append("    using S = typename std::conditional<JAC, Dual<NTW>, Dual<0>>::type;")
    out.append(pw)
    out.append("    for (int i = 0; i < %d; ++i) TAU[i] = S(tau[i]);" % sz("tau"))
    out.append("    for (int i = 0; i < %d; ++i) NL[i] = S(i < ND ? nl[i] : 0.0);" % sz("n"))
    out.append("    S U[NC], UH[NCU];")
    out.append("    for (int k = 0; k < NC; ++k) { U[k] = S(uq[k]); if (JAC) U[k].seed(k); }")
    out.append("    for (int k = 0; k < NCU; ++k) { UH[k] = S(uh[k]); if (JAC) UH[k].seed(NC + k); }")
    out.append("    for (int k = 0; k < NCW; ++k) { W[k] = S(w[k]); if (JAC) W[k].seed(NC + NCU + k); }")
    out.append("    S F[NCU * NBC];")
    out.append("    gen_%s::fn_FbouHdg<S>(%s, F);" % (ident, argv("FbouHdg", "fbou", "W")))
    out.append("    const int sel = (ib >= 1 && ib <= NBC) ? ib - 1 : 0;")
    out.append("#pragma unroll")
    out.append("    for (int b = 0; b < NBC; ++b) {")
    out.append("      if (b != sel) continue;")
    out.append("#pragma unroll")
    out.append("      for (int r = 0; r < NCU; ++r) {")
    out.append("        fb[r] = F[b * NCU + r].v;")
    out.append("        if (JAC) {")
    out.append("          for (int k = 0; k < NC; ++k) fb_uq[r + NCU * k] = F[b * NCU + r].deriv(k);")
    out.append("          for (int k = 0; k < NCU; ++k) fb_uh[r + NCU * k] = F[b * NCU + r].deriv(NC + k);")
    out.append("          for (int k = 0; k < NCW; ++k) fb_w[r + NCU * k] = F[b * NCU + r].deriv(NC + NCU + k);")
    out.append("        }")
    out.append("      }")
    out.append("    }")
    out.append("  }")
    return out


def write_registry():
    """csrc/generated/registry.hpp: the generated models compiled into the library."""
    os.makedirs(GEN_DIR, exist_ok=True)
    entries = []
    for fn in sorted(os.listdir(GEN_DIR)):
        if not fn.endswith(".hpp") or fn == "registry.hpp":
            continue
        text = open(os.path.join(GEN_DIR, fn)).read()
        m = re.search(r"// (\S+): nd = (\d+), ncu = (\d+), nc = (\d+), nv = (\d+), (\d+) boundary", text)
        s = re.search(r"struct Gen_(\w+) \{", text)
        if m and s:
            entries.append((m.group(1), s.group(1), int(m.group(2)), int(m.group(3)), fn))
    lines = ["// Generated by exasim_amd/modelgen.py -- the generated models compiled into libexahip.so.",
             "#pragma once"]
    for e in entries:
        lines.append('#include "%s"' % e[4])
    lines.append("")
    lines.append("// X(id, name, struct): ids follow the built-in models (100 + position)")
    lines.append("#define EXA_GENERATED_MODELS(X) \\")
    for k, e in enumerate(entries):
        lines.append("  X(%d, \"%s\", Gen_%s) \\" % (100 + k, e[0], e[1]))
    lines.append("  /* end of list */")
    open(os.path.join(GEN_DIR, "registry.hpp"), "w").write("\n".join(lines) + "\n")
    return entries


def generate(path, name):
    text = open(path).read()
    code, info = emit_cpp(text, name, os.path.relpath(os.path.abspath(path), os.path.dirname(HERE)))
    os.makedirs(GEN_DIR, exist_ok=True)
    out = os.path.join(GEN_DIR, info["ident"] + ".hpp")
    open(out, "w").write(code)
    write_registry()
    return out, info


def generate_all(models_dir=None):
    """Regenerate csrc/generated/ from every <name>.txt under models_dir (the package's models/
    by default); each model is registered under its file name without the extension."""
    models_dir = models_dir or os.path.join(HERE, "models")
    os.makedirs(GEN_DIR, exist_ok=True)
    keep = set()
    out = []
    for fn in sorted(os.listdir(models_dir)):
        if fn.endswith(".txt"):
            path, info = generate(os.path.join(models_dir, fn), fn[:-4])
            keep.add(os.path.basename(path))
            out.append(info)
    for fn in os.listdir(GEN_DIR):
        if fn.endswith(".hpp") and fn != "registry.hpp" and fn not in keep:
            os.remove(os.path.join(GEN_DIR, fn))
    write_registry()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("pdemodel", nargs="?")
    ap.add_argument("--name", help="model name used by exa_create")
    ap.add_argument("--all", action="store_true", help="regenerate every model under exasim_amd/models/")
    args = ap.parse_args(argv)
    if args.all:
        infos = generate_all()
    elif args.pdemodel and args.name:
        infos = [generate(args.pdemodel, args.name)[1]]
    else:
        ap.error("give a pdemodel.txt and --name, or --all")
    for info in infos:
        print("%s: nd %d, ncu %d, nc %d, nv %d, %d boundary conditions" % (
            info["name"], info["nd"], info["ncu"], info["nc"], info["nv"], info["nbc"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
