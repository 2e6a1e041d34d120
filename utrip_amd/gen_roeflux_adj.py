"""Generates csrc/cans_roeflux_adj.hpp, the reverse mode (vector-Jacobian
product) of the 2D Roe face flux roeflux12 (csrc/cans_roeflux.hpp;
cans2d/hdroe/roeflux_m.f:48-242).

The generated function first runs the flux's own statements, copied from
cans_roeflux.hpp with T = double (so every forward value is exactly the
flux's), then walks them backwards: for a statement v = f(x1, .., xk) it adds
a_v * df/dxi to the adjoint of each xi.  The local partial derivatives come
from sympy.  The flux's branch switches are differentiated as the reference
evaluates them: SIGN(...) factors (the degenerate-field switches sp, sp2 and
the sign of bx) are piecewise constant and carry no derivative; MAX and ABS
take the derivative of the branch taken (at a tie, the first argument of
MAX).  Run from the repository root:  python3 -m utrip_amd.gen_roeflux_adj
"""
from __future__ import annotations

import re
import sys
from pathlib import Path

import sympy as sp
from sympy.printing.c import C99CodePrinter

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc" / "cans_roeflux.hpp"
OUT = HERE / "csrc" / "cans_roeflux_adj.hpp"

INPUTS = ["rhol", "prl", "vxl", "vyl", "bxl", "byl", "rhor", "prr", "vxr", "vyr", "bxr", "byr"]
CONSTS = ["gm", "pi4", "pi4i", "pi8i", "one", "half", "zero", "sqh"]


def body_of(src: str, name: str) -> str:
    i = src.index(f"CANS_DEV void {name}(")
    i = src.index("{", i) + 1
    depth, j = 1, i
    while depth:
        if src[j] == "{":
            depth += 1
        elif src[j] == "}":
            depth -= 1
        j += 1
    return src[i:j - 1]


def split_top(s: str, sep: str):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == sep and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    out.append(cur)
    return out


def statements(body: str):
    """(name, rhs) in order; outputs are named out0..out4."""
    text = re.sub(r"//[^\n]*", "", body)
    res = []
    for st in text.split(";"):
        st = " ".join(st.split())
        if not st or st.startswith("(void)"):
            continue
        m = re.match(r"^const T (.*)$", st)
        if m:
            for part in split_top(m.group(1), ","):
                name, rhs = part.split("=", 1)
                res.append((name.strip(), rhs.strip()))
            continue
        m = re.match(r"^out\[(\d)\] = (.*)$", st)
        if m:
            res.append((f"out{m.group(1)}", m.group(2).strip()))
            continue
        raise ValueError(f"unparsed statement: {st}")
    return res


class Sgn(sp.Function):
    """signF(a, b) = |a| sign(b): piecewise constant, no derivative."""

    def fdiff(self, argindex=1):
        return sp.Integer(0)


class Gt(sp.Function):
    """1 if a > b else 0 (a derivative factor; itself constant)."""

    def fdiff(self, argindex=1):
        return sp.Integer(0)


class SgnD(sp.Function):
    """sign(x) as the derivative of |x| (0 at 0)."""

    def fdiff(self, argindex=1):
        return sp.Integer(0)


class FMax(sp.Function):
    """fmaxF(a, b) = a > b ? a : b: the derivative of the argument taken."""

    def fdiff(self, argindex=1):
        a, b = self.args
        return Gt(a, b) if argindex == 1 else 1 - Gt(a, b)


class FAbs(sp.Function):
    def fdiff(self, argindex=1):
        return SgnD(self.args[0])


def to_sympy(rhs: str, syms: dict):
    e = rhs
    e = re.sub(r"T\(RoePi<T>::(\w+)\)", lambda m: {"pi4d": "pi4", "pi4id": "pi4i", "pi8id": "pi8i",
                                                    "sqh": "sqh"}[m.group(1)], e)
    e = re.sub(r"\bT\(([0-9.eE+-]+)\)", r"\1", e)
    e = e.replace("tsqrt(", "sqrt(").replace("tabs(", "FAbs(").replace("fmaxF(", "FMax(").replace("signF(", "Sgn(")
    loc = dict(syms)
    loc.update({"sqrt": sp.sqrt, "FAbs": FAbs, "FMax": FMax, "Sgn": Sgn})
    return sp.sympify(e, locals=loc, rational=False)


class Printer(C99CodePrinter):
    def _print_Pow(self, expr):
        b, e = expr.base, expr.exp
        if e.is_Integer:
            n = int(e)
            if 1 <= n <= 4:
                return "(" + "*".join([self.parenthesize(b, 100)] * n) + ")"
            if -4 <= n <= -1:
                return "(1.0/(" + "*".join([self.parenthesize(b, 100)] * (-n)) + "))"
        raise ValueError(f"power {expr} (square roots are separate statements)")

    def _print_Gt(self, expr):
        return f"adj_gt({self._print(expr.args[0])}, {self._print(expr.args[1])})"

    def _print_SgnD(self, expr):
        return f"adj_sgn({self._print(expr.args[0])})"

    def _print_Sgn(self, expr):
        return f"signF({self._print(expr.args[0])}, {self._print(expr.args[1])})"

    def _print_FMax(self, expr):
        return f"fmaxF({self._print(expr.args[0])}, {self._print(expr.args[1])})"

    def _print_FAbs(self, expr):
        return f"fabs({self._print(expr.args[0])})"

    def _print_Float(self, expr):
        return repr(float(expr))

    def _print_Integer(self, expr):
        return f"{int(expr)}.0"

    def _print_Rational(self, expr):
        return f"({expr.p}.0/{expr.q}.0)"


def flatten_sqrt(exprs):
    """Each square root becomes a statement of its own: tq = arg, sq =
    sqrt(tq); the statement that used it reads sq.  Returns the new list of
    (name, expr, kind) with kind 'orig' (a statement of the flux), 'tmp'
    (an argument temporary) or 'sqrt'."""
    out = []
    k = 0
    for name, ex in exprs:
        while True:
            roots = [p for p in sp.preorder_traversal(ex) if isinstance(p, sp.Pow) and p.exp == sp.Rational(1, 2)]
            roots += [p for p in sp.preorder_traversal(ex) if isinstance(p, sp.Pow) and p.exp == sp.Rational(-1, 2)]
            if not roots:
                break
            # innermost first: a root whose argument holds no root
            r = next(p for p in roots if not any(isinstance(q, sp.Pow) and q.exp in (sp.Rational(1, 2),
                                                                                  sp.Rational(-1, 2))
                                                 for q in sp.preorder_traversal(p.base) if q is not p.base
                                                 or isinstance(q, sp.Pow)))
            arg = r.base
            if isinstance(arg, sp.Symbol):
                tq = arg
            else:
                tq = sp.Symbol(f"tq{k}", real=True)
                out.append((tq.name, arg, "tmp"))
            sq = sp.Symbol(f"sq{k}", real=True)
            out.append((sq.name, sp.sqrt(tq), "sqrt"))
            k += 1
            ex = ex.subs({sp.sqrt(arg): sq, 1 / sp.sqrt(arg): 1 / sq})
        out.append((name, ex, "orig"))
    return out


def generate() -> str:
    src = SRC.read_text()
    body = body_of(src, "roeflux12")
    sts = statements(body)
    syms = {n: sp.Symbol(n, real=True) for n in INPUTS + CONSTS}
    exprs = []
    for name, rhs in sts:
        if name in CONSTS:   # the flux's own constants (pi4, one, half, ...)
            continue
        if name in syms:
            raise ValueError(f"reassigned {name}")
        ex = to_sympy(rhs, syms)
        syms[name] = sp.Symbol(name, real=True)
        exprs.append((name, ex))
    # constants: statements whose value does not depend on a differentiable
    # quantity once the SIGN switches are taken as constants
    const_names = set(CONSTS)
    for name, ex in exprs:
        ex0 = ex.xreplace({g: sp.Integer(1) for g in ex.atoms(Sgn)})
        if not ({x.name for x in ex0.free_symbols} - const_names):
            const_names.add(name)
    flat = flatten_sqrt(exprs)
    pr = Printer()
    lines = []
    fwd = body.rstrip()
    fwd = re.sub(r"\bconst T\b", "const double", fwd)
    fwd = re.sub(r"RoePi<T>", "RoePi<double>", fwd)
    fwd = re.sub(r"\bT\(", "double(", fwd)
    fwd = re.sub(r"out\[(\d)\] = ", r"const double out\1 = ", fwd)
    fwd = fwd.replace("(void)gq;", "")
    lines.append(fwd)
    lines.append("    // square roots and their arguments as separate values")
    for name, ex, kind in flat:
        if kind in ("tmp", "sqrt"):
            lines.append(f"    const double {name} = {pr.doprint(ex) if kind == 'tmp' else 'sqrt(' + pr.doprint(ex.args[0]) + ')'};")
    # reverse sweep: only statements that can carry an adjoint
    has = {f"out{k}" for k in range(5)}
    body_rev = []
    for name, ex, kind in reversed(flat):
        if name in const_names or name not in has:
            continue
        args = sorted([x for x in ex.free_symbols if x.name not in const_names], key=lambda x: x.name)
        if not args:
            continue
        body_rev.append(f"    // {name}")
        for x in args:
            if kind == "sqrt":
                d = "adj_hsq(" + name + ")"
            else:
                dd = sp.diff(ex, x)
                if dd == 0:
                    continue
                d = pr.doprint(dd)
            body_rev.append(f"    a_{x.name} += a_{name} * ({d});")
            has.add(x.name)
    live = [n for n, _, _ in flat if n in has] + INPUTS
    lines.append("    // adjoints")
    lines.append("    double " + ", ".join(f"a_{n} = 0.0" for n in live) + ";")
    for k in range(5):
        lines.append(f"    a_out{k} = lo[{k}];")
    lines += body_rev
    lines.append("    " + " ".join(f"li[{k}] = a_{n};" for k, n in enumerate(INPUTS)))
    lines.append("    " + " ".join(f"fo[{k}] = out{k};" for k in range(5)))
    head = f'''// cans_roeflux_adj.hpp -- GENERATED by utrip_amd/gen_roeflux_adj.py from
// roeflux12 in cans_roeflux.hpp; do not edit (run the generator instead).
//
// roeflux12_vjp: the reverse mode of the 2D Roe face flux.  Given lo = dL/d
// (fro, fee, frx, fry, fby) of one face, it returns li = dL/d (rhol, prl,
// vxl, vyl, bxl, byl, rhor, prr, vxr, vyr, bxr, byr), the left and right
// states in the sweep frame, and fo = the flux itself.  The forward statements are the flux's own (T =
// double); SIGN switches carry no derivative, MAX and ABS the derivative of
// the branch taken, and every square root is a value of its own whose
// derivative is 0.5/s (0 at s = 0).
#pragma once
#include "cans_roeflux.hpp"

namespace cans {{

CANS_DEV double adj_gt(double a, double b) {{ return a > b ? 1.0 : 0.0; }}
CANS_DEV double adj_sgn(double x) {{ return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }}
// d sqrt(u) / du from s = sqrt(u); 0 where s = 0 (a genuine singularity of
// the flux, reached only on measure-zero states such as bx = 0 exactly)
CANS_DEV double adj_hsq(double s) {{ return s > 0.0 ? 0.5 / s : 0.0; }}

CANS_DEV void roeflux12_vjp(double gm, double rhol, double prl, double vxl, double vyl, double bxl, double byl,
                            double rhor, double prr, double vxr, double vyr, double bxr, double byr,
                            const double *lo, double *li, double *fo) {{
'''
    return head + "\n".join(lines) + "\n}\n\n}  // namespace cans\n"


def main():
    text = generate()
    if len(sys.argv) > 1 and sys.argv[1] == "--check":
        if OUT.read_text() != text:
            raise SystemExit(f"{OUT} is stale: rerun python3 -m utrip_amd.gen_roeflux_adj")
        return
    OUT.write_text(text)
    print(f"wrote {OUT} ({len(text.splitlines())} lines)")


if __name__ == "__main__":
    main()
