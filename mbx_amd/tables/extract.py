"""Extract the MB-pol / MB-nrg polynomial tables from the reference's C++ *text*.

This is a one-off generator that runs only where the reference checkout exists
(``/root/reference``); its output (``*.tbl``) is committed and is what the
kernel generator and the oracle read.  Nothing here imports or executes
reference code: the generated energy-only polynomial routines are parsed as
straight-line arithmetic and expanded symbolically.

What is extracted, per polynomial:
  * ``coef``  -- the fitted linear coefficients ``thefit[]`` of the holder
    (e.g. src/potential/2b/x2b-v9x.cpp:164-1318, src/potential/3b/x3b-v2x.cpp:80-1244);
  * ``monomials`` -- the expansion  E(x) = sum_k a_k * sum_{m in S_k} prod_{i in m} x_i
    of the energy-only routine (src/potential/2b/poly-2b-v6x-nogrd.cpp:41,
    src/potential/3b/poly-3b-v2x-nogrd.cpp:39).  Every monomial of those
    routines belongs to exactly one coefficient with weight 1; the script
    checks this and aborts otherwise.

Usage:  python mbx_amd/tables/extract.py [/root/reference]
"""
import ast
import os
import re
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))

# (table name, polynomial text, holder text with thefit[], nvars, ncoef, routine regex)
POLYS = [
    ("mbpol_2b", "src/potential/2b/poly-2b-v6x-nogrd.cpp", "src/potential/2b/x2b-v9x.cpp", 31, 1153),
    ("mbpol_3b", "src/potential/3b/poly-3b-v2x-nogrd.cpp", "src/potential/3b/x3b-v2x.cpp", 36, 1163),
]


class Poly:
    """Sparse polynomial, linear in the coefficient vector ``a``.

    Terms are keyed by (k, mono): k is the coefficient index (-1 for an
    a-free term) and mono a sorted tuple of variable indices.
    """

    __slots__ = ("d",)

    def __init__(self, d):
        self.d = d

    def __add__(self, o):
        if isinstance(o, (int, float)):
            o = Poly({(-1, ()): float(o)})
        r = dict(self.d)
        for key, v in o.d.items():
            r[key] = r.get(key, 0.0) + v
        return Poly(r)

    __radd__ = __add__

    def __mul__(self, o):
        if isinstance(o, (int, float)):
            return Poly({key: v * o for key, v in self.d.items()})
        r = defaultdict(float)
        for (ka, ma), va in self.d.items():
            for (kb, mb), vb in o.d.items():
                if ka >= 0 and kb >= 0:
                    raise ValueError("product of two coefficients: not linear in a")
                r[(max(ka, kb), tuple(sorted(ma + mb)))] += va * vb
        return Poly(dict(r))

    __rmul__ = __mul__

    def __neg__(self):
        return Poly({key: -v for key, v in self.d.items()})

    def __sub__(self, o):
        return self + (-o)


class _Coef:
    def __getitem__(self, k):
        return Poly({(int(k), ()): 1.0})


class _Var:
    def __getitem__(self, j):
        return Poly({(-1, (int(j),)): 1.0})


def _routine_body(src):
    """Body of the scalar energy-only eval routine (the first ``eval(`` with a braces block)."""
    m = re.search(r"eval\s*\(([^)]*)\)\s*\{", src)
    if not m:
        raise RuntimeError("eval routine not found")
    i = m.end()
    depth = 1
    j = i
    while depth:
        c = src[j]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
        j += 1
    return src[i:j - 1]


def expand(text):
    body = re.sub(r"//[^\n]*", "", _routine_body(text))
    # the 2B routine loops over dimers: x[nv + nd * 30] -> x[30]
    body = re.sub(r"x\[\s*nv\s*\+\s*nd\s*\*\s*(\d+)\s*\]", r"x[\1]", body)
    body = re.sub(r"for\s*\([^)]*\)\s*\{", ";", body)
    body = body.replace("}", ";")
    stmts = [s.strip() for s in " ".join(body.split()).split(";") if s.strip()]
    env = {"a": _Coef(), "x": _Var()}
    result = None
    for s in stmts:
        if s.startswith("const double"):
            name, expr = s[len("const double"):].split("=", 1)
            env[name.strip()] = eval(compile(ast.parse(expr.strip(), mode="eval"), "<poly>", "eval"), env)
        elif s.startswith("energy[nv]"):
            result = eval(compile(ast.parse(s.split("=", 1)[1].strip(), mode="eval"), "<poly>", "eval"), env)
        elif s.startswith("return") and result is None:
            expr = s[len("return"):].strip()
            if expr and expr != "energy":
                result = eval(compile(ast.parse(expr, mode="eval"), "<poly>", "eval"), env)
    if result is None:
        raise RuntimeError("no energy expression found")
    return result


def parse_thefit(text, ncoef):
    m = re.search(r"thefit\[\]\s*=\s*\{(.*?)\};", text, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    vals = [float(v) for v in body.replace(",", " ").split()]
    if len(vals) != ncoef:
        raise RuntimeError("thefit has %d entries, expected %d" % (len(vals), ncoef))
    return vals


def write_table(path, name, nvars, coef, poly):
    terms = []
    for (k, mono), w in poly.d.items():
        if w == 0.0:
            continue
        if k < 0:
            raise RuntimeError("a-free term in %s" % name)
        if w != 1.0:
            raise RuntimeError("monomial weight %r != 1 in %s" % (w, name))
        terms.append((k, mono))
    monos = [m for _, m in terms]
    if len(set(monos)) != len(monos):
        raise RuntimeError("a monomial belongs to two coefficients in %s" % name)
    if len({k for k, _ in terms}) != len(coef):
        raise RuntimeError("not every coefficient is used in %s" % name)
    terms.sort(key=lambda t: (len(t[1]), t[1]))
    with open(path, "w") as f:
        f.write("# %s polynomial table: E(x) = sum_k a_k sum_{m in S_k} prod x_i\n" % name)
        f.write("# generated by mbx_amd/tables/extract.py from the reference C++ text\n")
        f.write("nvars %d\nncoef %d\nnmono %d\n" % (nvars, len(coef), len(terms)))
        f.write("coef\n")
        for c in coef:
            f.write("%.17g\n" % c)
        f.write("monomials  # k degree i1 .. id\n")
        for k, mono in terms:
            f.write("%d %d %s\n" % (k, len(mono), " ".join(str(i) for i in mono)))
    return len(terms)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for name, poly_path, holder_path, nvars, ncoef in POLYS:
        poly = expand(open(os.path.join(ref, poly_path)).read())
        coef = parse_thefit(open(os.path.join(ref, holder_path)).read(), ncoef)
        n = write_table(os.path.join(HERE, name + ".tbl"), name, nvars, coef, poly)
        print("%s: %d monomials, %d coefficients" % (name, n, len(coef)))


if __name__ == "__main__":
    main()
