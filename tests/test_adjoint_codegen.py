"""CPU checks of the generated differentiation code (mbx_amd/gen/polygen.py: emit_dual, emit_basis).

The generated C is interpreted line by line in Python and compared with an independent
evaluation of the polynomial from its monomial table (mbx_amd/tables/<name>.tbl), by complex-step
differentiation: with z = x + i h xd (h = 1e-30), Re P(z) = P(x) and Im P(z) / h = grad P . xd
exactly up to rounding, and the same for grad P (whose tangent is the Hessian-vector product
H xd) and for each coefficient's basis function B_k = sum of its monomials.  The dual-number
parts must add up to P, grad P . xd, grad P and H xd (the tangent-only parts, which the backward
pass runs, to H xd alone); the basis slices must give B_k and grad B_k . xd for every coefficient
of the slice.
"""
import re
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, ROOT + "/mbx_amd/gen")
import polygen  # noqa: E402

H = 1e-30


def _fma(a, b, c):
    return a * b + c


def _body(src, fname):
    lines = src.split("\n")
    i = next(i for i, l in enumerate(lines) if l.startswith("__device__") and (" %s(" % fname) in l)
    out = []
    for l in lines[i + 1:]:
        if l == "}":
            return out
        out.append(l.strip())
    raise AssertionError("unterminated function " + fname)


def _translate(lines):
    py = []
    for t in lines:
        if not t or t in ("{", "}") or t.startswith(("MBXG_NODE_FENCE", "ad_opaque", "__builtin", "ad_basis_flush")):
            continue
        t = t.rstrip(";")
        m = re.match(r"^red\[\d+ \* 65 \+ lane\] = ", t)
        if m:
            py.append("out.append((TA, TS))")
            continue
        t = re.sub(r"^(const )?double ", "", t)
        t = t.replace("TA = 0.0, TS = 0.0", "TA = 0.0; TS = 0.0")
        py.append(t.replace("__fma_rn", "fma"))
    return py


def dual_fn(src, fname):
    code = "def f(C, x, xd, g, gd, E, dE):\n" + "\n".join("    " + l for l in _translate(_body(src, fname))) + \
        "\n    return E, dE\n"
    ns = {"fma": _fma}
    exec(compile(code, fname, "exec"), ns)
    return ns["f"]


def tan_fn(src, fname):
    code = "def f(C, x, xd, gd):\n" + "\n".join("    " + l for l in _translate(_body(src, fname))) + "\n"
    ns = {"fma": _fma}
    exec(compile(code, fname, "exec"), ns)
    return ns["f"]


def basis_fn(src, fname):
    code = "def f(x, xd):\n    out = []\n" + "\n".join("    " + l for l in _translate(_body(src, fname))) + \
        "\n    return out\n"
    ns = {"fma": _fma}
    exec(compile(code, fname, "exec"), ns)
    return ns["f"]


class Reference:
    """Monomials of a table, evaluated at a complex point."""

    def __init__(self, name):
        self.nv, self.coef, monos = polygen.load_table(name)
        self.k = np.array([k for k, m in monos])
        self.e = np.zeros((len(monos), self.nv), dtype=np.int64)
        for r, (k, m) in enumerate(monos):
            for v in m:
                self.e[r, v] += 1

    def at(self, x, xd):
        z = x + 1j * H * xd
        mono = np.prod(z[None, :] ** self.e, axis=1)
        dmono = self.e * (mono[:, None] / z[None, :])  # d mono / d z_i (x_i > 0)
        return mono, dmono

    def dual(self, x, xd, a):
        mono, dmono = self.at(x, xd)
        w = a[self.k]
        p = (w * mono).sum()
        g = (w[:, None] * dmono).sum(axis=0)
        scale = np.abs(w * mono.real).sum()
        return p.real, p.imag / H, g.real, g.imag / H, scale

    def basis(self, x, xd):
        mono, _ = self.at(x, xd)
        nc = len(self.coef)
        return (np.bincount(self.k, weights=mono.real, minlength=nc),
                np.bincount(self.k, weights=mono.imag, minlength=nc) / H)


def test_dual_rules():
    """Tangent statements of every statement form the schedule emits."""
    assert polygen.dual_stmt("const double P2_0 = x[1] * x[4];") == [
        "const double P2_0 = x[1] * x[4];", "const double dP2_0 = __fma_rn(xd[1], x[4], x[1] * xd[4]);"]
    assert polygen.dual_stmt("double R3_0 = {c} * x[6];")[1] == "double dR3_0 = {c} * xd[6];"
    assert polygen.dual_stmt("double R2_1 = {c};")[1] == "double dR2_1 = 0.0;"
    assert polygen.dual_stmt("g[3] = __fma_rn({c}, P3_2, g[3]);")[1] == "gd[3] = __fma_rn({c}, dP3_2, gd[3]);"
    assert polygen.dual_stmt("g[3] = __fma_rn(P2_1, R3_2, g[3]);")[1] == \
        "gd[3] = __fma_rn(dP2_1, R3_2, __fma_rn(P2_1, dR3_2, gd[3]));"
    assert polygen.dual_stmt("g[2] += {c};") == ["g[2] += {c};"]
    assert polygen.dual_stmt("g[2] += R1_0;")[1] == "gd[2] += dR1_0;"
    assert polygen.dual_stmt("E = __fma_rn(x[0], R1_0, E);")[1] == "dE = __fma_rn(xd[0], R1_0, __fma_rn(x[0], dR1_0, dE));"


@pytest.mark.parametrize("name", ["mbpol_2b", "mbpol_3b"])
def test_dual_parts(name):
    ref = Reference(name)
    monos = _monos(ref)
    s = polygen.Schedule(ref.nv, list(zip(ref.k.tolist(), monos)))
    assign = s.partition(polygen.AD_PARTS)
    rng = np.random.default_rng(3)
    x, xd = rng.uniform(0.1, 0.9, ref.nv), rng.normal(size=ref.nv)
    a = rng.normal(size=len(ref.coef))
    E, dE, g, gd = 0.0, 0.0, np.zeros(ref.nv), np.zeros(ref.nv)
    for p in range(polygen.AD_PARTS):
        f = dual_fn(polygen.emit_dual(s, "f", p, assign), "f")
        gp, gdp = [0.0] * ref.nv, [0.0] * ref.nv
        e, de = f(list(a), list(x), list(xd), gp, gdp, 0.0, 0.0)
        E, dE, g, gd = E + e, dE + de, g + np.array(gp), gd + np.array(gdp)
    p_ref, dp_ref, g_ref, gd_ref, scale = ref.dual(x, xd, a)
    assert E == pytest.approx(p_ref, rel=1e-12, abs=1e-14 * scale)
    assert dE == pytest.approx(dp_ref, rel=1e-11, abs=1e-13 * scale)
    np.testing.assert_allclose(g, g_ref, rtol=1e-11, atol=1e-13 * scale)
    np.testing.assert_allclose(gd, gd_ref, rtol=1e-10, atol=1e-12 * scale)


@pytest.mark.parametrize("name", ["mbpol_2b", "mbpol_3b"])
def test_tangent_only_parts(name):
    """The kernel the backward pass runs: only H xd, with E, dE and g dropped (the generated
    function must not reference them at all)."""
    ref = Reference(name)
    s = polygen.Schedule(ref.nv, list(zip(ref.k.tolist(), _monos(ref))))
    assign = s.partition(polygen.AD_PARTS)
    rng = np.random.default_rng(4)
    x, xd = rng.uniform(0.1, 0.9, ref.nv), rng.normal(size=ref.nv)
    a = rng.normal(size=len(ref.coef))
    gd = np.zeros(ref.nv)
    for p in range(polygen.AD_PARTS):
        src = polygen.emit_dual(s, "f", p, assign, tangent_only=True)
        body = "\n".join(_body(src, "f"))
        assert not re.search(r"\b(E|dE)\b|\bg\[", body)
        gdp = [0.0] * ref.nv
        tan_fn(src, "f")(list(a), list(x), list(xd), gdp)
        gd += np.array(gdp)
    _, _, _, gd_ref, scale = ref.dual(x, xd, a)
    np.testing.assert_allclose(gd, gd_ref, rtol=1e-10, atol=1e-12 * scale)


def _monos(ref):
    return [tuple(int(v) for v in np.repeat(np.arange(ref.nv), row)) for row in ref.e]


@pytest.mark.parametrize("name,slices", [("mbpol_2b", None), ("mbpol_3b", [0, 15])])
def test_basis_slices(name, slices):
    ref = Reference(name)
    monos = list(zip(ref.k.tolist(), _monos(ref)))
    bounds = polygen.basis_slices(monos, len(ref.coef), polygen.AD_POLYS[name])
    assert bounds[0][0] == 0 and bounds[-1][1] == len(ref.coef)
    assert all(b0[1] == b1[0] for b0, b1 in zip(bounds, bounds[1:]))
    rng = np.random.default_rng(4)
    x, xd = rng.uniform(0.1, 0.9, ref.nv), rng.normal(size=ref.nv)
    b_ref, bd_ref = ref.basis(x, xd)
    for i in (slices if slices is not None else range(len(bounds))):
        c0, c1 = bounds[i]
        out = basis_fn(polygen.emit_basis("f", monos, c0, c1), "f")(list(x), list(xd))
        assert len(out) == c1 - c0
        ta, ts = np.array(out).T
        np.testing.assert_allclose(ta, b_ref[c0:c1], rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(ts, bd_ref[c0:c1], rtol=1e-12, atol=1e-14)
