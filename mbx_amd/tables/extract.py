"""Extract the MB-pol / MB-nrg polynomial tables from the reference's C++ *text*.

This is a one-off generator that runs only where the reference checkout exists
(``/root/reference``); its output (``*.tbl``) is committed and is what the
kernel generator and the oracle read.  Nothing here imports or executes
reference code: the generated energy-only polynomial routines are parsed as
straight-line arithmetic and expanded symbolically.

What is extracted, per polynomial (MB-pol 2B/3B and the MB-nrg CH4 families):
  * ``coef``  -- the fitted linear coefficients ``thefit[]`` of the holder
    (e.g. src/potential/2b/x2b-v9x.cpp:164-1318, src/potential/3b/x3b-v2x.cpp:80-1244);
  * ``monomials`` -- the expansion  E(x) = sum_k a_k * sum_{m in S_k} prod_{i in m} x_i
    of the energy-only routine (src/potential/2b/poly-2b-v6x-nogrd.cpp:41,
    src/potential/3b/poly-3b-v2x-nogrd.cpp:39).  Every monomial of those
    routines belongs to exactly one coefficient with weight 1; the script
    checks this and aborts otherwise.
and per MB-nrg holder (``<name>.holder.json``, 2B/3B and the CH4 1B): its variables as data (form, k, d, sites),
its lone-pair sites and switching radii; and the data of the MB-pol 1B water potential
(``ps_1b.json``: the Partridge-Schwenke constants and coefficient tables of
src/potential/1b/ps.cpp).

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
    # MB-nrg CH4 families (energy-only routines; coefficients of the holders' single parameter set)
    ("ch4h2o_2b", "src/potential/2b/poly_2b_A1B2Z2_C1D4_deg3_exp0_v1.cpp",
     "src/potential/2b/x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp", 38, 603),
    ("ch4ch4_2b", "src/potential/2b/poly_2b_A1B4_A1B4_deg4_exp0_v1.cpp",
     "src/potential/2b/x2b_A1B4_A1B4_deg4_exp0_v1x.cpp", 45, 1345),
    ("ch4h2o2_3b", "src/potential/3b/poly_3b_A1B4_C1D2_C1D2_deg3_nograd_v1.cpp",
     "src/potential/3b/mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.cpp", 55, 1254),
    # MB-nrg CH4 1B (degree 5)
    ("ch4_1b", "src/potential/1b/poly_1b_A1B4_deg5_exp0_v1.cpp",
     "src/potential/1b/x1b_A1B4_deg5_exp0_v1x.cpp", 10, 207),
    # MB-nrg CO2 1B (degree 4; the value-and-gradient routine, its energy expression)
    ("co2_1b", "src/potential/1b/poly_1b_A1B2_deg4_v1x.cpp", "src/potential/1b/x1b_A1B2_deg4_v1x.cpp", 3, 21),
]

# 1B holders: the monomer, its atom count and the holder text (x1b_A1B4_deg5_exp0_v1x.cpp:309-343)
HOLDERS_1B = [
    ("ch4_1b", "src/potential/1b/x1b_A1B4_deg5_exp0_v1x.cpp", "ch4", 5),
    ("co2_1b", "src/potential/1b/x1b_A1B2_deg4_v1x.cpp", "co2", 3),
]

# Holders of the MB-nrg families: argument order of the monomers, and how the variables are
# built (x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp:821-949, x2b_A1B4_A1B4_deg4_exp0_v1x.cpp:1470-1600,
# mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.cpp:1336-1480).
HOLDERS = [
    ("ch4h2o_2b", "src/potential/2b/x2b_A1B2Z2_C1D4_deg3_exp0_v1x.cpp", ["h2o", "ch4"], [3, 5]),
    ("ch4ch4_2b", "src/potential/2b/x2b_A1B4_A1B4_deg4_exp0_v1x.cpp", ["ch4", "ch4"], [5, 5]),
    ("ch4h2o2_3b", "src/potential/3b/mbnrg_3b_A1B4_C1D2_C1D2_deg3_v1.cpp", ["ch4", "h2o", "h2o"], [5, 3, 3]),
]


def parse_holder(text, mons, nats):
    """Variables of a generated MB-nrg holder, as data: sites (real atoms by position in the
    concatenated coordinates, plus the water's two lone-pair sites when the holder builds them),
    and for each variable its form (exp0: exp(k (d - r)); coul0: exp(k (d - r)) / r; exp:
    exp(-k r)), k, d and its two sites; the switching radii."""
    src = re.sub(r"//[^\n]*", "", text)
    ev = src[src.index("::eval("):]
    ev = ev[:ev.index("::eval(", 10)] if "::eval(" in ev[10:] else ev
    sites = {}
    for name, off in re.findall(r"const double\*\s*(?:coords_)?([A-Z]_\d+_[a-z])\s*=\s*(?:xcrd|xyz\.data\(\))\s*\+\s*(\d+);", ev):
        sites[name] = int(off) // 3
    nreal = sum(nats)
    vsite = None
    m = re.search(r"m\.setup\(([A-Z]_\d+_[a-z]),\s*(\w+),\s*(\w+),\s*([A-Z]_\d+_[a-z]),\s*([A-Z]_\d+_[a-z])\)", ev)
    if m:
        w = {k: float(v) for k, v in re.findall(r"double\s+(w12|w13|wcross)\s*=\s*([-+0-9.eE]+);", ev)}
        sites[m.group(4)] = nreal
        sites[m.group(5)] = nreal + 1
        vsite = {"oxygen_site": sites[m.group(1)], "in_plane": w[m.group(2)], "out_of_plane": w[m.group(3)]}
    # parameters of the holder's (single) parameter set
    ctor = src[:src.index("::f_switch")]
    par = {k: float(v) for k, v in re.findall(r"\b(m_\w+)\s*=\s*([-+0-9.eE]+)\s*;", ctor)}
    vars_ = {}
    for idx, form, args in re.findall(r"(?:v|xs)\[(\d+)\]\s*=\s*(?:vr|vs)\[\d+\]\.(v_exp0|v_coul0|v_exp)\(([^)]*)\);", ev):
        a = [t.strip() for t in args.split(",")]
        if form == "v_exp":
            d, k, s1, s2 = 0.0, par[a[0]], a[1], a[2]
        else:
            d, k, s1, s2 = par[a[0]], par[a[1]], a[2], a[3]
        s1, s2 = s1.replace("coords_", ""), s2.replace("coords_", "")
        vars_[int(idx)] = [form[2:], k, d, sites[s1], sites[s2]]
    nv = max(vars_) + 1
    assert sorted(vars_) == list(range(nv))
    sw = {"inner": par.get("m_r2i", par.get("m_ri")), "outer": par.get("m_r2f", par.get("m_ro"))}
    return {"monomers": mons, "atoms": nats, "nreal": nreal, "nsites": nreal + (2 if vsite else 0),
            "vsite": vsite, "switch": sw, "vars": [vars_[i] for i in range(nv)]}


# Further MB-nrg families, evaluated on the GPU as data-described holders (eval_xfam.hip):
# (holder name, holder text, monomer names in the holder's argument order, atoms per monomer,
#  polynomial structure name).  A holder text may hold several parameter sets (one per ion, or
#  per CO2 model); the set is the constructor's block for these monomer names
# (e.g. mbnrg_2b_A1_B1C2X2_deg6_v1.cpp:12 "f-", :929 "cl-", :1846 "br-", :2763 "i-").
XHOLDERS = [
    ("f_h2o_2b", "src/potential/2b/mbnrg_2b_A1_B1C2X2_deg6_v1.cpp", ["f-", "h2o"], [1, 3], "halide_h2o_2b"),
    ("cl_h2o_2b", "src/potential/2b/mbnrg_2b_A1_B1C2X2_deg6_v1.cpp", ["cl-", "h2o"], [1, 3], "halide_h2o_2b"),
    ("br_h2o_2b", "src/potential/2b/mbnrg_2b_A1_B1C2X2_deg6_v1.cpp", ["br-", "h2o"], [1, 3], "halide_h2o_2b"),
    ("i_h2o_2b", "src/potential/2b/mbnrg_2b_A1_B1C2X2_deg6_v1.cpp", ["i-", "h2o"], [1, 3], "halide_h2o_2b"),
    ("f_h2o2_3b", "src/potential/3b/mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.cpp", ["f-", "h2o", "h2o"], [1, 3, 3],
     "halide_h2o2_3b"),
    ("cl_h2o2_3b", "src/potential/3b/mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.cpp", ["cl-", "h2o", "h2o"], [1, 3, 3],
     "halide_h2o2_3b"),
    ("br_h2o2_3b", "src/potential/3b/mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.cpp", ["br-", "h2o", "h2o"], [1, 3, 3],
     "halide_h2o2_3b"),
    ("i_h2o2_3b", "src/potential/3b/mbnrg_3b_A1_B1C2X2_B1C2X2_deg4_v1.cpp", ["i-", "h2o", "h2o"], [1, 3, 3],
     "halide_h2o2_3b"),
    ("co2co2_2b", "src/potential/2b/x2b_A1B2_A1B2_v1x.cpp", ["co2", "co2"], [3, 3], "co2co2_2b"),
    ("co2x3_3b", "src/potential/3b/mbnrg_3b_A1B2_A1B2_A1B2_deg4_v1.cpp", ["co2", "co2", "co2"], [3, 3, 3],
     "co2x3_3b"),
    ("h2ox4_4b", "src/potential/4b/mbnrg_4b_A1B2_A1B2_A1B2_A1B2_deg4_v1.cpp", ["h2o"] * 4, [3] * 4, "h2ox4_4b"),
]

# Alkali-metal-ion-water holders, in the older hand-written style (named non-linear parameters,
# one parameter set per ion): x2b_h2o_ion_v2x (x2b-A1B2Z2_C-v1x.cpp:176-2416, eval with
# gradients :2546-2770) and x3b_h2o_ion_v1x_deg4_filtered (x3b-h2o-ion-v1x-deg4-filtered.cpp:
# 206-482, parameter sets :484-).  Both take water(s) first and the ion last
# (energy2b.cpp:258-282, energy3b.cpp:223-228).
ALKALI = ("li+", "na+", "k+", "rb+", "cs+")
ALKALI_HOLDERS = [("%s_h2o_2b" % ion[:-1], "src/potential/2b/x2b-A1B2Z2_C-v1x.cpp", ["h2o", ion], [3, 1],
                   "alkali_h2o_2b") for ion in ALKALI] + \
                 [("%s_h2o2_3b" % ion[:-1], "src/potential/3b/x3b-h2o-ion-v1x-deg4-filtered.cpp",
                   ["h2o", "h2o", ion], [3, 3, 1], "alkali_h2o2_3b") for ion in ALKALI]


def _named_params(blk):
    return {k: float(v) for k, v in re.findall(r"\b([A-Za-z_]\w*)\s*=\s*([-+0-9.eE]+)\s*;", blk)}


def parse_alkali_2b(text, mons, nats):
    """x2b_h2o_ion_v2x for one ion as holder data: sites O H H X lp1 lp2, the eight variables of
    eval() (v_exp: exp(k (r0 - r)), v_coul: the same over r; x2b-A1B2Z2_C-v1x.cpp:58-96), the
    water lone pairs (monomer::setup, :123-141), the cosine switch on the O-X separation with
    the dimer skipped beyond r2f, and the ion's twobodyfit coefficients."""
    ion = mons[1]
    src = re.sub(r"//[^\n]*", "", text)
    m = re.search(r'if\s*\(\s*mon1\s*==\s*"h2o"\s*&&\s*mon2\s*==\s*"%s"\s*\)' % re.escape(ion), src)
    blk = _block(src, m.end())
    c = re.search(r"twobodyfit\s*=\s*std::vector<double>\s*\{(.*?)\};", blk, re.S)
    coef = [float(v) for v in c.group(1).replace(",", " ").split()]
    par = _named_params(blk[c.end():])
    i = src.index("x2b_h2o_ion_v2x::eval(const double* w1, const double* x, double* g1")
    ev = src[i:]
    site = {n: int(o) // 3 for n, o in re.findall(r"const int (\w+) = (\d+);", ev[:ev.index("poly_2b_h2o_ion")])}
    assert site == {"O": 0, "H1": 1, "H2": 2, "X": 3, "lp1": 4, "lp2": 5}, site
    assert re.search(r"setup\(xcrd \+ O, in_plane_gamma, out_of_plane_gamma, xcrd \+ lp1, xcrd \+ lp2\)", ev)
    vars_ = {}
    for idx, form, d, k, a, b in re.findall(
            r"v\[(\d+)\]\s*=\s*ctxt\[[^\]]*\]\.(v_exp|v_coul)\((\w+),\s*(\w+),\s*xcrd,\s*(\w+),\s*(\w+)\)", ev):
        vars_[int(idx)] = [{"v_exp": "exp0", "v_coul": "coul0"}[form], par[k], par[d], site[a], site[b]]
    assert sorted(vars_) == list(range(8)), vars_
    assert "if (rab[i] > r2f) continue;" in ev and "const double sw = f_switch(rab[sh], gsw);" in ev
    return {"monomers": mons, "atoms": nats, "nreal": 4, "nsites": 6,
            "vsites": [{"oxygen_site": 0, "in_plane": par["in_plane_gamma"], "out_of_plane": par["out_of_plane_gamma"]}],
            "switch": {"inner": par["r2i"], "outer": par["r2f"], "pairs": [[0, 1]], "terms": [[1.0, [0]]]},
            "skip": "pairs_beyond_outer", "skip_min": 1, "vars": [vars_[i] for i in range(8)], "coef": coef}


def parse_alkali_3b(text, mons, nats):
    """x3b_h2o_ion_v1x_deg4_filtered for one ion as holder data: sites Oa Ha1 Ha2 Ob Hb1 Hb2 X,
    the 21 variables of operator() (v_intra / v_main: exp(-k (d - r0)), which is exp(k (r0 - d))
    bit for bit; x3b-h2o-ion-v1x-deg4-filtered.cpp:43-93), the switch combination
    s_ab s_ac + s_ab s_bc + s_ac s_bc over the O-O and O-X separations, the trimer skipped when two
    of them exceed r3f, and the ion's coefficients."""
    ion = mons[2]
    src = re.sub(r"//[^\n]*", "", text)
    m = re.search(r'if\s*\(\s*m3\s*==\s*"%s"\s*\)' % re.escape(ion), src)
    blk = _block(src, m.end())
    c = re.search(r"the_poly\[\]\s*=\s*\{(.*?)\};", blk, re.S)
    coef = [float(v) for v in c.group(1).replace(",", " ").split()]
    par = _named_params(blk[c.end():])
    i = src.index("double* grad1, double* grad2, double* grad3")
    ev = src[i:src.index("x3b_h2o_ion_v1x_deg4_filtered::x3b_h2o_ion_v1x_deg4_filtered")]
    base = {"w1": 0, "w2": 3, "x": 6}
    site = {}
    for n, b, o in re.findall(r"const double\*\s*(\w+)\s*=\s*(w1|w2|x)\s*(?:\+\s*(\d+))?;", ev):
        if n not in ("w1", "w2", "x"):
            site[n] = base[b] + int(o or 0) // 3
    assert site == {"Oa": 0, "Ha1": 1, "Ha2": 2, "Ob": 3, "Hb1": 4, "Hb2": 5, "X": 6}, site
    vars_ = {}
    for idx, form, k, d, a, b in re.findall(r"v\[(\d+)\]\s*=\s*(v_intra|v_main)\((m_\w+),\s*(m_\w+),\s*(\w+),\s*(\w+)\)", ev):
        vars_[int(idx)] = ["exp0", par[k], par[d], site[a], site[b]]
    assert sorted(vars_) == list(range(21)), vars_
    assert "const double s = sab * sac + sab * sbc + sac * sbc;" in ev
    assert re.search(r"if \(\(drab > m_r3f and drac > m_r3f\) or \(drab > m_r3f and drbc > m_r3f\) or "
                     r"\(drac > m_r3f and drbc > m_r3f\)\)\s*continue;", ev)
    return {"monomers": mons, "atoms": nats, "nreal": 7, "nsites": 7, "vsites": [],
            "switch": {"inner": par["m_r3i"], "outer": par["m_r3f"], "pairs": [[0, 1], [0, 2], [1, 2]],
                       "terms": [[1.0, [0, 1]], [1.0, [0, 2]], [1.0, [1, 2]]]},
            "skip": "pairs_beyond_outer", "skip_min": 2, "vars": [vars_[i] for i in range(21)], "coef": coef}


def xholder_data(path, text, mons, nats):
    if path.endswith("x2b-A1B2Z2_C-v1x.cpp"):
        return parse_alkali_2b(text, mons, nats)
    if path.endswith("x3b-h2o-ion-v1x-deg4-filtered.cpp"):
        return parse_alkali_3b(text, mons, nats)
    return parse_xholder(text, mons, nats)


# Their polynomial structures: (name, energy-only routine, variables, coefficients)
XPOLYS = [
    ("halide_h2o_2b", "src/potential/2b/poly_2b_A1_B1C2X2_deg6_nograd_v1.cpp", 8, 906),
    ("halide_h2o2_3b", "src/potential/3b/poly_3b_A1_B1C2X2_B1C2X2_deg4_nograd_v1.cpp", 41, 3173),
    ("co2co2_2b", "src/potential/2b/poly_2b_A1B2_A1B2_v1.cpp", 15, 2269),
    ("co2x3_3b", "src/potential/3b/poly_3b_A1B2_A1B2_A1B2_deg4_nograd_v1.cpp", 27, 1067),
    ("h2ox4_4b", "src/potential/4b/poly_4b_A1B2_A1B2_A1B2_A1B2_deg4_nograd_v1.cpp", 66, 398),
    ("alkali_h2o_2b", "src/potential/2b/poly-2b-A1B2Z2_C-v1.cpp", 8, 429),
    # value-and-gradient routine; its energy expression (double e)
    ("alkali_h2o2_3b", "src/potential/3b/poly-3b-h2o-ion-v1x_deg4_filtered.cpp", 21, 1016),
]

# pairs of an n-mer's monomers in the holders' switch naming (sw12, sw13, sw14, sw23, sw24, sw34)
PAIRS = {2: [(0, 1)], 3: [(0, 1), (0, 2), (1, 2)], 4: [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]}


class _Multi:
    """Multilinear polynomial in the pair switches: {frozenset(pair indices): coefficient}."""

    def __init__(self, d):
        self.d = d

    def __add__(self, o):
        o = o if isinstance(o, _Multi) else _Multi({frozenset(): float(o)})
        r = dict(self.d)
        for k, v in o.d.items():
            r[k] = r.get(k, 0.0) + v
        return _Multi(r)

    __radd__ = __add__

    def __neg__(self):
        return _Multi({k: -v for k, v in self.d.items()})

    def __sub__(self, o):
        return self + (-o if isinstance(o, _Multi) else -float(o))

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if not isinstance(o, _Multi):
            return _Multi({k: v * float(o) for k, v in self.d.items()})
        r = {}
        for ka, va in self.d.items():
            for kb, vb in o.d.items():
                if ka & kb:
                    raise ValueError("a switch appears twice in one term")
                r[ka | kb] = r.get(ka | kb, 0.0) + va * vb
        return _Multi(r)

    __rmul__ = __mul__


def _block(src, start):
    """Text of the brace block whose opening brace is the first one at or after `start`."""
    i = src.index("{", start) + 1
    depth, j = 1, i
    while depth:
        depth += src[j] == "{"
        depth -= src[j] == "}"
        j += 1
    return src[i:j - 1]


def parameter_set(src, mons):
    """The constructor block of the parameter set for these monomer names: its coefficients and
    its m_* parameters."""
    cond = r"\s*(?:and|&&)\s*".join(r'mon%d\s*==\s*"%s"' % (i + 1, re.escape(m)) for i, m in enumerate(mons))
    m = re.search(r"if\s*\(\s*" + cond + r"\s*\)", src)
    if not m:
        raise RuntimeError("no parameter set for %s" % "-".join(mons))
    blk = _block(src, m.end())
    c = re.search(r"coefficients\s*=\s*std::vector<double>\s*\{(.*?)\};", blk, re.S)
    coef = [float(v) for v in c.group(1).replace(",", " ").split()]
    par = {k: float(v) for k, v in re.findall(r"\b(m_\w+)\s*=\s*([-+0-9.eE]+)\s*;", blk)}
    return coef, par


def parse_xholder(text, mons, nats):
    """A generated MB-nrg holder as data: sites (real atoms of the monomers in argument order,
    then the lone-pair sites of every water it builds, two each), variables (form, k, d, site
    pair), the switching radii, the pair-switch combination and the skip rule, plus the
    parameter set of these monomers (coefficients)."""
    src = re.sub(r"//[^\n]*", "", text)
    coef, par = parameter_set(src, mons)
    ev = src[src.index("::eval("):]
    ev = ev[:ev.index("::eval(", 10)]
    sites = {}
    for name, off in re.findall(
            r"const double\*\s*(?:coords_)?([A-Z]_\d+_[a-z])\s*=\s*(?:xcrd|xyz\.data\(\))\s*\+\s*(\d+);", ev):
        sites[name] = int(off) // 3
    nreal = sum(nats)
    assert sorted(sites.values()) == list(range(nreal)), sites
    w = {k: float(v) for k, v in re.findall(r"double\s+(w12|w13|wcross)\s*=\s*([-+0-9.eE]+);", ev)}
    vsites = []
    for o, g1, g2, x1, x2 in re.findall(
            r"\w+\.setup\((?:coords_)?([A-Z]_\d+_[a-z]),\s*(\w+),\s*(\w+),\s*(?:coords_)?([A-Z]_\d+_[a-z]),"
            r"\s*(?:coords_)?([A-Z]_\d+_[a-z])\)", ev):
        sites[x1] = nreal + 2 * len(vsites)
        sites[x2] = nreal + 2 * len(vsites) + 1
        vsites.append({"oxygen_site": sites[o], "in_plane": w[g1], "out_of_plane": w[g2]})
    vars_ = {}
    for idx, form, args in re.findall(r"(?:v|xs)\[(\d+)\]\s*=\s*(?:vr|vs)\[\d+\]\.(v_exp0|v_coul0|v_exp|v_coul)\(([^)]*)\);",
                                      ev):
        a = [t.strip() for t in args.split(",")]
        if len(a) == 4:  # (r0, k, p1, p2): exp(k (r0 - r)) [/ r]; x2b_A1B2_A1B2_v1x.cpp:62's own v_exp too
            f = {"v_exp0": "exp0", "v_coul0": "coul0", "v_exp": "exp0", "v_coul": "coul0"}[form]
            d, k, s1, s2 = par[a[0]], par[a[1]], a[2], a[3]
        else:  # (k, p1, p2): exp(-k r) [/ r]
            f = {"v_exp": "exp", "v_coul": "coul"}[form]
            d, k, s1, s2 = 0.0, par[a[0]], a[1], a[2]
        vars_[int(idx)] = [f, k, d, sites[s1.replace("coords_", "")], sites[s2.replace("coords_", "")]]
    nv = max(vars_) + 1
    assert sorted(vars_) == list(range(nv))
    k = len(mons)
    pairs = PAIRS[k]
    names = {"sw%d%d" % (a + 1, b + 1): i for i, (a, b) in enumerate(pairs)}
    expr = [" ".join(e.split()) for e in re.findall(r"(?<![\w.])sw\s*=\s*([^;]+);", ev) if e.strip() != "0.0"][0]
    if expr.startswith("f_switch("):
        assert k == 2
        terms = [[1.0, [0]]]
    else:
        env = {n: _Multi({frozenset([i]): 1.0}) for n, i in names.items()}
        val = eval(compile(ast.parse(expr, mode="eval"), "<sw>", "eval"), env)
        terms = [[v, sorted(t)] for t, v in sorted(val.d.items(), key=lambda kv: sorted(kv[0])) if v != 0.0]
    # skip rule: the n-mer is skipped when every first-atom separation exceeds the outer radius
    outer = par.get("m_ro", par.get("m_r2f"))
    inner = par.get("m_ri", par.get("m_r2i"))
    skip = re.search(r"if\s*\(\s*(true\s*&&[^)]*|r12\s*>\s*m_r2f)\)", ev).group(1)
    want = {"d%d%dr" % (a + 1, b + 1) for a, b in pairs}
    got = set(re.findall(r"(d\d\dr)\s*>\s*m_ro", skip)) or ({"d12r"} if "r12" in skip else set())
    assert got == want, (skip, want)
    return {"monomers": mons, "atoms": nats, "nreal": nreal, "nsites": nreal + 2 * len(vsites), "vsites": vsites,
            "switch": {"inner": inner, "outer": outer, "pairs": pairs, "terms": terms},
            "skip": "all_pairs_beyond_outer", "vars": [vars_[i] for i in range(nv)], "coef": coef}


def parse_holder_1b(text, mon, nat):
    """Variables of a generated MB-nrg 1B holder over the monomer's atoms (sites named A_1, B_1, ...
    or A0, B0, ... at xcrd + 3 * site), with the numeric parameters of the holder's constructor:
    v[i] = vr[i].v_exp0(d, k, P, Q) (x1b_A1B4_deg5_exp0_v1x.cpp) or x[i] = v_intra(k, d, P, Q)
    (x1b_A1B2_deg4_v1x.cpp:41-47, exp(-k (r - d))).  Both are form exp0: exp(k (d - r))."""
    src = re.sub(r"//[^\n]*", "", text)
    ev = src[src.index("::eval("):]
    ev = ev[:ev.index("::eval(", 10)]
    sites = {name: int(off) // 3 for name, off in re.findall(r"const double\*\s*([A-Z]_?\d+)\s*=\s*xcrd\s*\+\s*(\d+);", ev)}
    ctor = src[:src.index("::eval(")]
    par = {k: float(v) for k, v in re.findall(r"\b(\w+)\s*=\s*([-+0-9.eE]+)\s*;", ctor)}
    vars_ = {}
    for idx, args in re.findall(r"v\[(\d+)\]\s*=\s*vr\[\d+\]\.v_exp0\(([^)]*)\);", ev):
        a = [t.strip() for t in args.split(",")]
        vars_[int(idx)] = ["exp0", par[a[1]], par[a[0]], sites[a[2]], sites[a[3]]]
    for idx, args in re.findall(r"x\[(\d+)\]\s*=\s*v_intra\(([^)]*)\);", ev):
        a = [t.strip() for t in args.split(",")]
        vars_[int(idx)] = ["exp0", par[a[0]], par[a[1]], sites[a[2]], sites[a[3]]]
    nv = max(vars_) + 1
    assert sorted(vars_) == list(range(nv)) and len(sites) == nat
    return {"monomers": [mon], "atoms": [nat], "nreal": nat, "nsites": nat, "vsite": None,
            "switch": {"inner": 0.0, "outer": 0.0}, "vars": [vars_[i] for i in range(nv)]}


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
    body = re.sub(r"^\s*#[^\n]*", "", body, flags=re.M)  # #ifdef DEBUG ... #endif lines
    # the 2B routine loops over dimers: x[nv + nd * 30] -> x[30]
    body = re.sub(r"x\[\s*nv\s*\+\s*nd\s*\*\s*(\d+)\s*\]", r"x[\1]", body)
    body = re.sub(r"for\s*\([^)]*\)\s*\{", ";", body)
    body = body.replace("}", ";")
    stmts = [s.strip() for s in " ".join(body.split()).split(";") if s.strip()]
    env = {"a": _Coef(), "x": _Var()}
    result = None
    for s in stmts:
        if s.startswith("const double") or s.startswith("double "):
            name, expr = s[len("const double") if s.startswith("const") else len("double"):].split("=", 1)
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


def expand_tarray(text):
    """Routines split into helpers fmgN(x, a, t) that fill a temporary array t[]
    (src/potential/2b/poly_2b_A1B4_A1B4_deg4_exp0_v1.cpp:44982): eval() calls them in order
    and returns an expression of t[]."""
    src = re.sub(r"//[^\n]*", "", text)
    m = re.search(r"double\s+\w+::eval\s*\(\s*const double a\[\d+\]\s*,\s*const double x\[\d+\]\s*\)\s*\{", src)
    body = src[m.end():src.index("}", m.end())]
    calls = re.findall(r"(fmg\d+)\s*\(", body)
    ret = re.search(r"return\s*(.*?);", body, re.S).group(1)
    T = {}

    class _T:
        def __getitem__(self, i):
            return T[int(i)]

    env = {"a": _Coef(), "x": _Var(), "t": _T()}
    for fn in calls:
        fm = re.search(r"void\s+%s\s*\([^)]*\)\s*\{" % fn, src)
        i = j = fm.end()
        depth = 1
        while depth:
            depth += src[j] == "{"
            depth -= src[j] == "}"
            j += 1
        for stmt in " ".join(src[i:j - 1].split()).split(";"):
            stmt = stmt.strip()
            if stmt:
                lhs, expr = stmt.split("=", 1)
                idx = int(re.match(r"t\[(\d+)\]", lhs.strip()).group(1))
                T[idx] = eval(compile(ast.parse(expr.strip(), mode="eval"), "<poly>", "eval"), env)
    return eval(compile(ast.parse(" ".join(ret.split()), mode="eval"), "<poly>", "eval"), env)


def parse_thefit(text, ncoef):
    """Fitted linear coefficients: ``thefit[] = {...}`` (MB-pol holders) or the first
    ``coefficients = std::vector<double>{...}`` (MB-nrg holders, one parameter set each)."""
    m = re.search(r"thefit\[\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        m = re.search(r"coefficients\s*=\s*std::vector<double>\s*\{(.*?)\};", text, re.S)
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


def parse_ps(text, consts_text):
    """The Partridge-Schwenke water monomer potential's data (src/potential/1b/ps.cpp:46-89 and
    the tables c5zA, cbasis, ccore, crest, idx1-3 at :61-295; the cm-1 -> kcal/mol constants of
    src/tools/constants.h:43-61): scalars and arrays, as given in the text."""
    src = re.sub(r"//[^\n]*", "", text)
    out = {}
    for name in ("reoh", "thetae", "b1", "roh", "alphaoh", "deohA", "phh1A", "phh2", "f5z", "fbasis", "fcore",
                 "frest"):
        m = re.search(r"const double %s\s*=\s*([^;]+);" % name, src)
        out[name] = float(m.group(1))
    for name, typ in (("idx1", int), ("idx2", int), ("idx3", int), ("c5zA", float), ("cbasis", float),
                      ("ccore", float), ("crest", float)):
        m = re.search(r"const (?:size_t|double) %s\[245\]\s*=\s*\{([^}]*)\}" % name, src)
        vals = [typ(v.strip()) for v in m.group(1).split(",") if v.strip()]
        assert len(vals) == 245, (name, len(vals))
        out[name] = vals
    m = re.search(r"const double costhe\s*=\s*([^;]+);", src)
    out["costhe"] = float(m.group(1))
    m = re.search(r"e1\[nv\]\s*\+=\s*([0-9.eE+-]+);", src)
    out["correction"] = float(m.group(1))
    csrc = re.sub(r"//[^\n]*", "", consts_text)
    for name in ("Eh_J", "Na", "kcal_J", "c0", "h_Js"):
        m = re.search(r"const double %s\s*=\s*([^;]+);" % name, csrc)
        out[name] = float(m.group(1))
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    import json
    ps = parse_ps(open(os.path.join(ref, "src/potential/1b/ps.cpp")).read(),
                  open(os.path.join(ref, "src/tools/constants.h")).read())
    with open(os.path.join(HERE, "ps_1b.json"), "w") as f:
        json.dump(ps, f, indent=0)
    print("ps_1b: %d coefficients" % len(ps["c5zA"]))
    for name, holder_path, mons, nats in HOLDERS:
        h = parse_holder(open(os.path.join(ref, holder_path)).read(), mons, nats)
        h["name"] = name
        with open(os.path.join(HERE, name + ".holder.json"), "w") as f:
            json.dump(h, f, indent=1)
        print("%s holder: %d variables, %d sites" % (name, len(h["vars"]), h["nsites"]))
    for name, holder_path, mon, nat in HOLDERS_1B:
        h = parse_holder_1b(open(os.path.join(ref, holder_path)).read(), mon, nat)
        h["name"] = name
        with open(os.path.join(HERE, name + ".holder.json"), "w") as f:
            json.dump(h, f, indent=1)
        print("%s holder: %d variables, %d sites" % (name, len(h["vars"]), h["nsites"]))
    for name, poly_path, nvars, ncoef in XPOLYS:
        text = open(os.path.join(ref, poly_path)).read()
        poly = expand_tarray(text) if "fmg1(" in text else expand(text)
        first = [h for h in XHOLDERS + ALKALI_HOLDERS if h[4] == name][0]
        coef = xholder_data(first[1], open(os.path.join(ref, first[1])).read(), first[2], first[3])["coef"]
        if len(coef) != ncoef:
            raise RuntimeError("%s: %d coefficients, expected %d" % (name, len(coef), ncoef))
        n = write_table(os.path.join(HERE, name + ".tbl"), name, nvars, coef, poly)
        print("%s: %d monomials, %d coefficients" % (name, n, len(coef)))
    for name, holder_path, mons, nats, poly in XHOLDERS + ALKALI_HOLDERS:
        h = xholder_data(holder_path, open(os.path.join(ref, holder_path)).read(), mons, nats)
        h["name"], h["poly"] = name, poly
        with open(os.path.join(HERE, name + ".holder.json"), "w") as f:
            json.dump(h, f, indent=0)
        print("%s holder: %d variables, %d sites, %d switch terms, %d coefficients" % (
            name, len(h["vars"]), h["nsites"], len(h["switch"]["terms"]), len(h["coef"])))
    for name, poly_path, holder_path, nvars, ncoef in POLYS:
        text = open(os.path.join(ref, poly_path)).read()
        poly = expand_tarray(text) if "fmg1(" in text else expand(text)
        coef = parse_thefit(open(os.path.join(ref, holder_path)).read(), ncoef)
        n = write_table(os.path.join(HERE, name + ".tbl"), name, nvars, coef, poly)
        print("%s: %d monomials, %d coefficients" % (name, n, len(coef)))


if __name__ == "__main__":
    main()
