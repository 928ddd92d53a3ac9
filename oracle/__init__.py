"""ORACLE -- test infrastructure, not product code.

ctypes bindings over ``oracle/liboracle.so`` (the C restatement of the
reference's MB-pol 2B/3B path, see oracle.h) and, where it was built, over
``oracle/_ref/libmbxref.so`` (the reference's own leaves compiled in place,
see ref_shim.cpp).  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
CPU-baseline leg import this package, and only as the checker.

Parity status per oracle function (what pins it) is recorded in oracle.h.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TABLES = os.path.join(ROOT, "mbx_amd", "tables")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_szp = ctypes.POINTER(ctypes.c_size_t)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


class _OrcPoly(ctypes.Structure):
    _fields_ = [("nvars", ctypes.c_int), ("ncoef", ctypes.c_int), ("nmono", ctypes.c_int),
                ("coef", _dp), ("mono_k", _ip), ("mono_deg", _ip), ("mono_idx", _ip), ("stride", ctypes.c_int)]


class Table:
    """A polynomial monomial table (mbx_amd/tables/<name>.tbl)."""

    def __init__(self, name):
        self.name = name
        path = os.path.join(TABLES, name + ".tbl")
        lines = [l.split("#")[0].strip() for l in open(path)]
        lines = [l for l in lines if l]
        it = iter(lines)
        hdr = {}
        for l in it:
            if l == "coef":
                break
            k, v = l.split()
            hdr[k] = int(v)
        self.nvars, self.ncoef, self.nmono = hdr["nvars"], hdr["ncoef"], hdr["nmono"]
        self.coef = np.array([float(next(it)) for _ in range(self.ncoef)])
        assert next(it).startswith("monomials")
        rows = [[int(t) for t in next(it).split()] for _ in range(self.nmono)]
        self.stride = max(4, max(v[1] for v in rows))
        k = np.zeros(self.nmono, np.int32)
        deg = np.zeros(self.nmono, np.int32)
        idx = -np.ones((self.nmono, self.stride), np.int32)
        for m, v in enumerate(rows):
            k[m], deg[m] = v[0], v[1]
            idx[m, :v[1]] = v[2:]
        self.mono_k, self.mono_deg, self.mono_idx = k, deg, idx
        self.c = _OrcPoly(self.nvars, self.ncoef, self.nmono, _d(self.coef), _i(self.mono_k),
                          _i(self.mono_deg), _i(self.mono_idx), self.stride)


class _OrcHolder(ctypes.Structure):
    _fields_ = [("nmon", ctypes.c_int), ("nat", ctypes.c_int * 3), ("nreal", ctypes.c_int),
                ("nsites", ctypes.c_int), ("vsite", ctypes.c_int), ("osite", ctypes.c_int),
                ("in_plane", ctypes.c_double), ("out_of_plane", ctypes.c_double), ("r_in", ctypes.c_double),
                ("r_out", ctypes.c_double), ("nvars", ctypes.c_int), ("form", _ip), ("k", _dp), ("d", _dp),
                ("s1", _ip), ("s2", _ip)]


FORMS = {"exp0": 0, "coul0": 1, "exp": 2}
NRG = ("ch4h2o_2b", "ch4ch4_2b", "ch4h2o2_3b")


class Holder:
    """An MB-nrg holder description (mbx_amd/tables/<name>.holder.json)."""

    def __init__(self, name):
        import json

        h = json.load(open(os.path.join(TABLES, name + ".holder.json")))
        self.name, self.desc = name, h
        v = h["vars"]
        self.form = np.array([FORMS[t[0]] for t in v], np.int32)
        self.k = np.array([t[1] for t in v], np.float64)
        self.d = np.array([t[2] for t in v], np.float64)
        self.s1 = np.array([t[3] for t in v], np.int32)
        self.s2 = np.array([t[4] for t in v], np.int32)
        nat = (ctypes.c_int * 3)(*(h["atoms"] + [0] * (3 - len(h["atoms"]))))
        vs = h["vsite"] or {"oxygen_site": 0, "in_plane": 0.0, "out_of_plane": 0.0}
        self.c = _OrcHolder(len(h["monomers"]), nat, h["nreal"], h["nsites"], int(h["vsite"] is not None),
                            vs["oxygen_site"], vs["in_plane"], vs["out_of_plane"], h["switch"]["inner"],
                            h["switch"]["outer"], len(v), _i(self.form), _d(self.k), _d(self.d), _i(self.s1),
                            _i(self.s2))


_TABLES = {}
_HOLDERS = {}


def holder(name):
    if name not in _HOLDERS:
        _HOLDERS[name] = Holder(name)
    return _HOLDERS[name]


def table(name):
    if name not in _TABLES:
        _TABLES[name] = Table(name)
    return _TABLES[name]


def build():
    """Compile liboracle.so (and the reference library when the checkout is present)."""
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-j8", "-C", HERE, "ref"], stdout=subprocess.DEVNULL)


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(_OrcPoly)
        L.orc_poly_eval.restype = ctypes.c_double
        L.orc_poly_eval.argtypes = [P, _dp, _dp, _dp]
        L.orc_poly_basis.argtypes = [P, _dp, _dp]
        L.orc_mbpol2b.restype = ctypes.c_double
        L.orc_mbpol2b.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        L.orc_mbpol3b.restype = ctypes.c_double
        L.orc_mbpol3b.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        L.orc_mbpol2b_t.restype = ctypes.c_double
        L.orc_mbpol2b_t.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        L.orc_mbpol3b_t.restype = ctypes.c_double
        L.orc_mbpol3b_t.argtypes = [P, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        L.orc_mbpol_nmers.argtypes = [ctypes.c_int, P, _dp, _dp, ctypes.c_long, _dp, _dp, _dp]
        L.orc_invert_unit_cell.argtypes = [_dp, _dp]
        L.orc_fix_monomer_coordinates.argtypes = [_dp, _dp, _dp, ctypes.c_int, _ip, _ip]
        L.orc_close_image.argtypes = [_dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_clusters.restype = ctypes.c_long
        L.orc_clusters.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip,
                                   _ip, ctypes.c_int, _ip, ctypes.c_long]
        L.orc_water_system.argtypes = [P, P, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                       _dp, _dp, _lp, _ip, ctypes.c_long, _ip, ctypes.c_long, _dp, _dp]
        H = ctypes.POINTER(_OrcHolder)
        L.orc_nrg_holder.restype = ctypes.c_double
        L.orc_nrg_holder.argtypes = [H, P, _dp, ctypes.POINTER(_dp), ctypes.POINTER(_dp), _dp, _ip]
        L.orc_system.argtypes = [P, P, ctypes.POINTER(H), ctypes.POINTER(P), ctypes.c_int, ctypes.c_int, _ip, _ip, _ip,
                                 _dp, _dp,
                                 ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp, _lp]
        XH = ctypes.POINTER(_OrcXHolder)
        L.orc_xholder_eval.restype = ctypes.c_double
        L.orc_xholder_eval.argtypes = [XH, P, _dp, ctypes.POINTER(_dp), ctypes.POINTER(_dp), _dp, _ip]
        _LIB = L
    return _LIB


def _arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def poly_eval(tab, x, a=None, grad=True):
    x = _arr(x)
    g = np.zeros(tab.nvars) if grad else None
    a = _arr(a) if a is not None else None
    e = lib().orc_poly_eval(ctypes.byref(tab.c), _d(a), _d(x), _d(g))
    return (e, g) if grad else e


def poly_basis(tab, x):
    x = _arr(x)
    b = np.zeros(tab.ncoef)
    lib().orc_poly_basis(ctypes.byref(tab.c), _d(x), _d(b))
    return b


def mbpol2b(w1, w2, a=None, theta=None):
    """One dimer: returns (E, g1, g2, virial, evaluated).  theta: the ten nonlinear parameters
    (k_HH_intra k_OH_intra k_HH_coul k_OH_coul k_OO_coul k_XH_main k_XO_main k_XX_main
    in_plane_gamma out_of_plane_gamma), None = the reference's."""
    w1, w2 = _arr(w1), _arr(w2)
    g1, g2, v = np.zeros(9), np.zeros(9), np.zeros(9)
    ev = ctypes.c_int(0)
    a = _arr(a) if a is not None else None
    th = _arr(theta) if theta is not None else None
    e = lib().orc_mbpol2b_t(ctypes.byref(table("mbpol_2b").c), _d(a), _d(th), _d(w1), _d(w2), _d(g1), _d(g2),
                            _d(v), ctypes.byref(ev))
    return e, g1, g2, v, ev.value


def mbpol3b(w1, w2, w3, a=None, theta=None):
    """One trimer: returns (E, g1, g2, g3, virial, evaluated).  theta: kHH_intra kOH_intra kHH kOH
    kOO dHH_intra dOH_intra dHH dOH dOO, None = the reference's."""
    w1, w2, w3 = _arr(w1), _arr(w2), _arr(w3)
    g1, g2, g3, v = np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(9)
    ev = ctypes.c_int(0)
    a = _arr(a) if a is not None else None
    th = _arr(theta) if theta is not None else None
    e = lib().orc_mbpol3b_t(ctypes.byref(table("mbpol_3b").c), _d(a), _d(th), _d(w1), _d(w2), _d(w3), _d(g1),
                            _d(g2), _d(g3), _d(v), ctypes.byref(ev))
    return e, g1, g2, g3, v, ev.value


MBPOL_THETA = {  # the reference's nonlinear parameters (x2b-v9x.cpp:1330-1345, x3b-v2x.cpp:1256-1271)
    "mbpol_2b": np.array([-6.480884773303821e-01, 1.674518993682975e+00, 1.148231864355956e+00,
                          1.205989761123099e+00, 1.395357065790959e+00, 7.347036852042255e-01,
                          7.998249864422826e-01, 7.960663960630585e-01, -9.721486914088159e-02,
                          9.859272078406150e-02]),
    "mbpol_3b": np.array([-2.254303257872797e+00, -2.564771901404151e-01, 4.247920544074333e-01,
                          8.128985941165371e-01, 3.749457984616480e-02, 1.690594510379166e+00,
                          -2.999999868517452e+00, 3.499031358429095e+00, 4.854042963514281e+00,
                          4.816312044947604e-08]),
}


def mbpol_nmers(family, xyz, a=None, theta=None, grad=True):
    """Independent dimers ([n][18]) or trimers ([n][27]) as given: (E[n], G[n][18|27] or None)."""
    order = 2 if family == "mbpol_2b" else 3
    x = _arr(xyz).reshape(-1, 9 * order)
    n = len(x)
    e = np.zeros(n)
    g = np.zeros((n, 9 * order)) if grad else None
    a = _arr(a) if a is not None else None
    th = _arr(theta) if theta is not None else None
    lib().orc_mbpol_nmers(order, ctypes.byref(table(family).c), _d(a), _d(th), n, _d(x), _d(e), _d(g))
    return e, g


def invert_unit_cell(box):
    box = _arr(box)
    bi = np.zeros(9)
    if lib().orc_invert_unit_cell(_d(box), _d(bi)):
        raise ValueError("singular box")
    return bi


def fix_monomer_coordinates(xyz, box, nat):
    xyz = _arr(xyz).copy()
    box = _arr(box)
    nat = np.ascontiguousarray(nat, np.int32)
    fi = np.concatenate([[0], np.cumsum(nat)[:-1]]).astype(np.int32)
    lib().orc_fix_monomer_coordinates(_d(xyz), _d(box), _d(invert_unit_cell(box)), len(nat), _i(nat), _i(fi))
    return xyz


def close_image(box, xyz1, xyz2):
    box = _arr(box)
    xyz2 = _arr(xyz2).copy()
    lib().orc_close_image(_d(box), _d(invert_unit_cell(box)), _d(_arr(xyz1)), len(xyz2) // 3, _d(xyz2))
    return xyz2


def clusters(nmax, cutoff, first_xyz, box=None, islocal=None, tags=None, use_ghost=False):
    """AddClusters restatement.  Returns an (n, nmax) int array in the oracle's emission order."""
    first_xyz = _arr(first_xyz).reshape(-1)
    nmon = len(first_xyz) // 3
    loc = np.ones(nmon, np.int32) if islocal is None else np.ascontiguousarray(islocal, np.int32)
    tag = np.arange(nmon, dtype=np.int32) if tags is None else np.ascontiguousarray(tags, np.int32)
    b = _arr(box) if box is not None and len(box) else None
    bi = invert_unit_cell(b) if b is not None else None
    L = lib()
    n = L.orc_clusters(nmax, cutoff, nmon, int(b is not None), _d(b), _d(bi), _d(first_xyz), _i(loc), _i(tag),
                       int(use_ghost), None, 0)
    out = np.zeros((max(n, 1), nmax), np.int32)
    L.orc_clusters(nmax, cutoff, nmon, int(b is not None), _d(b), _d(bi), _d(first_xyz), _i(loc), _i(tag),
                   int(use_ghost), _i(out), n)
    return out[:n]


def water_system(xyz, box, cutoff2b, cutoff3b, lists=False, per_nmer=False):
    """2B+3B energy/gradient of an all-water system (9 doubles per water, O H H)."""
    xyz = _arr(xyz).reshape(-1).copy()
    nmon = len(xyz) // 9
    b = _arr(box) if box is not None and len(box) else None
    e2, e3 = ctypes.c_double(), ctypes.c_double()
    grad = np.zeros(9 * nmon)
    vir = np.zeros(9)
    counts = np.zeros(4, np.int64)
    res = {}
    d = t = ed = et = None
    cap2 = cap3 = 0
    if lists or per_nmer:
        # first pass for the counts
        lib().orc_water_system(ctypes.byref(table("mbpol_2b").c), ctypes.byref(table("mbpol_3b").c), nmon,
                               _d(xyz.copy()), _d(b), cutoff2b, 0.0, ctypes.byref(e2), ctypes.byref(e3),
                               _d(grad), _d(vir), counts.ctypes.data_as(_lp), None, 0, None, 0, None, None)
        cap2 = int(counts[0])
        n3 = len(clusters(3, cutoff3b, fix_first(xyz, b), b))
        cap3 = n3
        d = np.zeros((max(cap2, 1), 2), np.int32)
        t = np.zeros((max(cap3, 1), 3), np.int32)
        if per_nmer:
            ed = np.zeros(max(cap2, 1))
            et = np.zeros(max(cap3, 1))
    lib().orc_water_system(ctypes.byref(table("mbpol_2b").c), ctypes.byref(table("mbpol_3b").c), nmon, _d(xyz),
                           _d(b), cutoff2b, cutoff3b, ctypes.byref(e2), ctypes.byref(e3), _d(grad), _d(vir),
                           counts.ctypes.data_as(_lp), _i(d), cap2, _i(t), cap3, _d(ed), _d(et))
    res.update(e2b=e2.value, e3b=e3.value, grad=grad, virial=vir, xyz=xyz, n2=int(counts[0]),
               n2_eval=int(counts[1]), n3=int(counts[2]), n3_eval=int(counts[3]))
    if lists or per_nmer:
        res["dimers"] = d[:cap2]
        res["trimers"] = t[:cap3]
    if per_nmer:
        res["e_dimer"] = ed[:cap2]
        res["e_trimer"] = et[:cap3]
    return res


def nrg(name, mons, a=None):
    """One n-mer through an MB-nrg holder: mons in the holder's monomer order (already imaged).
    Returns (E, [grad per monomer], virial(9), evaluated)."""
    h = holder(name)
    tab = table(name)
    ms = [_arr(m).reshape(-1).copy() for m in mons]
    gs = [np.zeros_like(m) for m in ms]
    vir = np.zeros(9)
    ev = ctypes.c_int()
    mp = (_dp * 3)(*[_d(m) for m in ms], *([None] * (3 - len(ms))))
    gp = (_dp * 3)(*[_d(g) for g in gs], *([None] * (3 - len(gs))))
    aa = _arr(a) if a is not None else None
    e = lib().orc_nrg_holder(ctypes.byref(h.c), ctypes.byref(tab.c), _d(aa), mp, gp, _d(vir), ctypes.byref(ev))
    return e, gs, vir, ev.value


KIND_NAT = {0: 3, 1: 5}


FAMILIES = ("mbpol_2b", "ch4h2o_2b", "ch4ch4_2b", "mbpol_3b", "ch4h2o2_3b")


def system(xyz, kinds, box, cutoff2b, cutoff3b, tags=None, families=FAMILIES, conditioning=False):
    """2B+3B energy/gradient of a mixed H2O (kind 0) / CH4 (kind 1) system: 3 doubles per atom,
    monomers in the given order; tags default to each monomer's first-atom index.  Families not
    in `families` are listed but not evaluated (the System's ignore lists).  conditioning=True
    evaluates with |coefficients| (see abs_table): the scale of the cancelling terms."""
    tab = abs_table if conditioning else table
    kinds = np.ascontiguousarray(kinds, np.int32)
    nat = np.array([KIND_NAT[int(k)] for k in kinds], np.int32)
    first = np.concatenate([[0], np.cumsum(nat)[:-1]]).astype(np.int32)
    tag = first.copy() if tags is None else np.ascontiguousarray(tags, np.int32)
    xyz = _arr(xyz).reshape(-1).copy()
    b = _arr(box) if box is not None and len(box) else None
    e2, e3 = ctypes.c_double(), ctypes.c_double()
    grad = np.zeros(len(xyz))
    vir = np.zeros(9)
    counts = np.zeros(4, np.int64)
    H = ctypes.POINTER(_OrcHolder)
    P = ctypes.POINTER(_OrcPoly)
    hs = (H * 3)(*[ctypes.pointer(holder(n).c) for n in NRG])
    ps = (P * 3)(*[ctypes.pointer(tab(n).c) for n in NRG])
    mask = sum(1 << i for i, f in enumerate(FAMILIES) if f in families)
    rc = lib().orc_system(ctypes.byref(tab("mbpol_2b").c), ctypes.byref(tab("mbpol_3b").c), hs, ps, mask, len(kinds),
                          _i(kinds), _i(nat), _i(tag), _d(xyz), _d(b), cutoff2b, cutoff3b, ctypes.byref(e2),
                          ctypes.byref(e3), _d(grad), _d(vir), counts.ctypes.data_as(_lp))
    if rc:
        raise ValueError("orc_system failed (%d)" % rc)
    return dict(e2b=e2.value, e3b=e3.value, grad=grad, virial=vir, xyz=xyz, n2=int(counts[0]),
                n2_eval=int(counts[1]), n3=int(counts[2]), n3_eval=int(counts[3]))


def abs_table(name):
    """The table `name` with every coefficient replaced by its magnitude: evaluated through the
    holders it gives sum_n sw_n sum_k |a_k| M_k(x_n), the scale of the cancelling terms behind
    an energy (its floating-point conditioning)."""
    key = name + "|abs"
    if key not in _TABLES:
        t = table(name)
        a = Table.__new__(Table)
        a.__dict__.update(t.__dict__)
        a.coef = np.abs(t.coef)
        a.c = _OrcPoly(t.nvars, t.ncoef, t.nmono, _d(a.coef), _i(t.mono_k), _i(t.mono_deg), _i(t.mono_idx),
                       t.stride)
        _TABLES[key] = a
    return _TABLES[key]


def water_conditioning(xyz, box, cutoff2b, cutoff3b):
    """(E2_abs, E3_abs): the 2B / 3B energies with |coefficients| (see abs_table)."""
    xyz = _arr(xyz).reshape(-1).copy()
    nmon = len(xyz) // 9
    b = _arr(box) if box is not None and len(box) else None
    e2, e3 = ctypes.c_double(), ctypes.c_double()
    grad, vir, counts = np.zeros(9 * nmon), np.zeros(9), np.zeros(4, np.int64)
    lib().orc_water_system(ctypes.byref(abs_table("mbpol_2b").c), ctypes.byref(abs_table("mbpol_3b").c), nmon,
                           _d(xyz), _d(b), cutoff2b, cutoff3b, ctypes.byref(e2), ctypes.byref(e3), _d(grad),
                           _d(vir), counts.ctypes.data_as(_lp), None, 0, None, 0, None, None)
    return e2.value, e3.value


def fix_first(xyz, box):
    """First-atom coordinates of each water after FixMonomerCoordinates (if periodic)."""
    xyz = _arr(xyz).reshape(-1)
    nmon = len(xyz) // 9
    if box is not None:
        xyz = fix_monomer_coordinates(xyz, box, [3] * nmon)
    return xyz.reshape(nmon, 9)[:, :3].copy()


# --------------------------------------------------------------------------------------------
# The compiled reference leaves (oracle/_ref/libmbxref.so), when built.
# --------------------------------------------------------------------------------------------

_REF = None
REF_PATH = os.path.join(HERE, "_ref", "libmbxref.so")


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _REF
    if _REF is None:
        L = ctypes.CDLL(REF_PATH)
        L.ref_poly2b.argtypes = [ctypes.c_size_t, _dp, _dp, _dp, _dp]
        L.ref_poly2b_nogrd.argtypes = [ctypes.c_size_t, _dp, _dp, _dp]
        L.ref_poly3b.argtypes = [_dp, _dp, _dp, _dp]
        L.ref_poly3b_nogrd.restype = ctypes.c_double
        L.ref_poly3b_nogrd.argtypes = [_dp, _dp]
        L.ref_x2b.restype = ctypes.c_double
        L.ref_x2b.argtypes = [_dp, _dp, ctypes.c_size_t, _dp, _dp, _dp]
        L.ref_x3b.restype = ctypes.c_double
        L.ref_x3b.argtypes = [_dp, _dp, _dp, ctypes.c_size_t, _dp, _dp, _dp, _dp]
        L.ref_invert_unit_cell.argtypes = [_dp, _dp]
        L.ref_fix_monomer_coordinates.argtypes = [_dp, ctypes.c_size_t, _dp, _dp, ctypes.c_size_t, _szp, _szp]
        L.ref_close_dimer_image.argtypes = [_dp, _dp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _dp, _dp]
        L.ref_add_clusters.restype = ctypes.c_size_t
        L.ref_add_clusters.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_int, _dp, _dp, _dp,
                                       ctypes.c_size_t, _szp, _szp, _ip, ctypes.c_int, _szp, ctypes.c_size_t]
        L.ref_water_2b3b.argtypes = [ctypes.c_size_t, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_size_t,
                                     ctypes.c_int, _dp, _dp, _dp, _dp, _szp]
        _REF = L
    return _REF


def ref_add_clusters(nmax, cutoff, first_xyz, box=None, islocal=None, use_ghost=False):
    first_xyz = _arr(first_xyz).reshape(-1)
    nmon = len(first_xyz) // 3
    fi = np.arange(nmon, dtype=np.uintp)
    loc = np.ones(nmon, np.uintp) if islocal is None else np.ascontiguousarray(islocal, np.uintp)
    tag = np.arange(nmon, dtype=np.int32)
    b = _arr(box) if box is not None else None
    bi = invert_unit_cell(b) if b is not None else None
    sz = lambda a: a.ctypes.data_as(_szp)
    L = ref()
    n = L.ref_add_clusters(nmax, cutoff, nmon, int(b is not None), _d(b), _d(bi), _d(first_xyz), nmon, sz(fi),
                           sz(loc), _i(tag), int(use_ghost), None, 0)
    out = np.zeros((max(n, 1), nmax), np.uintp)
    L.ref_add_clusters(nmax, cutoff, nmon, int(b is not None), _d(b), _d(bi), _d(first_xyz), nmon, sz(fi), sz(loc),
                       _i(tag), int(use_ghost), sz(out), n)
    return out[:n].astype(np.int64)


def ref_invert_unit_cell(box):
    """The compiled InvertUnitCell (tools/math_tools.cpp:37-64)."""
    bi = np.zeros(9)
    ref().ref_invert_unit_cell(_d(_arr(box)), _d(bi))
    return bi


def ref_fix_monomer_coordinates(xyz, box, nat):
    """The compiled systools::FixMonomerCoordinates (sys_tools.cpp:257-320); monomers of nat[i]
    atoms, back to back."""
    xyz = _arr(xyz).copy()
    box = _arr(box)
    nat = np.ascontiguousarray(nat, np.uintp)
    fi = np.ascontiguousarray(np.concatenate([[0], np.cumsum(nat)[:-1]]), np.uintp)
    sz = lambda a: a.ctypes.data_as(_szp)
    ref().ref_fix_monomer_coordinates(_d(xyz), len(xyz) // 3, _d(box), _d(ref_invert_unit_cell(box)), len(nat),
                                      sz(nat), sz(fi))
    return xyz


def ref_close_image(box, xyz1, xyz2):
    """The compiled systools::GetCloseDimerImage (sys_tools.cpp:322-365) for one dimer: xyz2 moved
    next to xyz1's first atom."""
    box = _arr(box)
    x1 = _arr(xyz1).copy()
    x2 = _arr(xyz2).copy()
    ref().ref_close_dimer_image(_d(box), _d(ref_invert_unit_cell(box)), len(x1) // 3, len(x2) // 3, 1, _d(x1), _d(x2))
    return x2


def ref_water_2b3b(xyz, box, cutoff2b, cutoff3b, batch=8, nthreads=1):
    xyz = _arr(xyz).reshape(-1).copy()
    nmon = len(xyz) // 9
    b = _arr(box) if box is not None else None
    e2, e3 = ctypes.c_double(), ctypes.c_double()
    grad, vir = np.zeros(9 * nmon), np.zeros(9)
    counts = np.zeros(4, np.uintp)
    ref().ref_water_2b3b(nmon, _d(xyz), _d(b), cutoff2b, cutoff3b, batch, nthreads, ctypes.byref(e2),
                         ctypes.byref(e3), _d(grad), _d(vir), counts.ctypes.data_as(_szp))
    return dict(e2b=e2.value, e3b=e3.value, grad=grad, virial=vir, n2=int(counts[0]), n3=int(counts[1]), xyz=xyz)


# ---------------------------------------------------------------------------------------------
# MB-pol 1B: the Partridge-Schwenke water monomer (potential/1b/ps.cpp:346-633)

class _Ps(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("reoh", "b1", "roh", "alphaoh", "deoh", "phh1", "phh2", "costhe",
                                               "correction", "cm1")] + \
               [("idx1", ctypes.c_int * 245), ("idx2", ctypes.c_int * 245), ("idx3", ctypes.c_int * 245),
                ("c5z", ctypes.c_double * 245)]


_PS = None


def ps_table():
    """The 1B data of mbx_amd/tables/ps_1b.json in the reference's derived form: c5z
    (ps.cpp:445-447), deoh and phh1 (:450-451) and the cm-1 -> kcal/mol factor (constants.h:47-61),
    each evaluated in the reference's operation order."""
    global _PS
    if _PS is None:
        import json
        import math
        d = json.load(open(os.path.join(ROOT, "mbx_amd", "tables", "ps_1b.json")))
        p = _Ps()
        for n in ("reoh", "b1", "roh", "alphaoh", "phh2", "costhe", "correction"):
            setattr(p, n, d[n])
        p.deoh = d["f5z"] * d["deohA"]
        p.phh1 = d["f5z"] * d["phh1A"] * math.exp(d["phh2"])
        eh_kcalmol = d["Eh_J"] * d["Na"] / d["kcal_J"]
        eh_cm1 = 1.0e-2 * d["Eh_J"] / (d["c0"] * d["h_Js"])
        p.cm1 = eh_kcalmol / eh_cm1
        for j in range(245):
            p.idx1[j], p.idx2[j], p.idx3[j] = d["idx1"][j], d["idx2"][j], d["idx3"][j]
            p.c5z[j] = d["f5z"] * d["c5zA"][j] + d["fbasis"] * d["cbasis"][j] + d["fcore"] * d["ccore"][j] + \
                d["frest"] * d["crest"][j]
        L = lib()
        L.orc_ps_1b.restype = ctypes.c_double
        L.orc_ps_1b.argtypes = [ctypes.POINTER(_Ps), _dp, _dp, _dp]
        _PS = p
    return _PS


def ps_1b(w, grad=True):
    """MB-pol 1B energies of waters w [n][9] (O H1 H2): (E[n], G[n][9] or None, virial[6])."""
    p = ps_table()
    w = _arr(w).reshape(-1, 9)
    e = np.zeros(len(w))
    g = np.zeros_like(w) if grad else None
    vir = np.zeros(6)
    for i in range(len(w)):
        wi = np.ascontiguousarray(w[i])
        gi = np.zeros(9) if grad else None
        e[i] = lib().orc_ps_1b(ctypes.byref(p), _d(wi), _d(gi) if grad else None, _d(vir) if grad else None)
        if grad:
            g[i] = gi
    return e, g, vir


def nrg_1b(name, w, a=None, grad=True):
    """MB-nrg 1B energies (orc_1b_nrg; name "ch4_1b" or "co2_1b") of monomers w [n][3*atoms]:
    (E[n], G[n][3*atoms] or None, virial[6])."""
    L = lib()
    L.orc_1b_nrg.restype = ctypes.c_double
    L.orc_1b_nrg.argtypes = [ctypes.POINTER(_OrcHolder), ctypes.POINTER(_OrcPoly), _dp, _dp, _dp, _dp]
    h, t = holder(name), table(name)
    na = 3 * h.desc["nreal"]
    w = _arr(w).reshape(-1, na)
    e = np.zeros(len(w))
    g = np.zeros_like(w) if grad else None
    vir = np.zeros(6)
    aa = _arr(a) if a is not None else None
    for i in range(len(w)):
        wi = np.ascontiguousarray(w[i])
        gi = np.zeros(na) if grad else None
        e[i] = L.orc_1b_nrg(ctypes.byref(h.c), ctypes.byref(t.c), _d(aa), _d(wi), _d(gi) if grad else None,
                            _d(vir) if grad else None)
        if grad:
            g[i] = gi
    return e, g, vir


# Atoms per monomer kind (the engine's kind numbers: h2o, ch4, -, co2, f-, cl-, br-, i-).
NAT_OF_KIND = {0: 3, 1: 5, 3: 3, 4: 1, 5: 1, 6: 1, 7: 1, 8: 1, 9: 1, 10: 1, 11: 1, 12: 1}


def one_body(xyz, kinds, box=None):
    """The 1B term of a system of H2O (kind 0, 3 atoms), CH4 (kind 1, 5 atoms), CO2 (kind 3, 3
    atoms) and ion (halides kinds 4-7, alkali metals 8-12, 1 atom) monomers in the given order:
    e1b::get_1b_energy per
    monomer (potential/1b/energy1b.cpp:118-177; ions have no 1B term) on the coordinates
    System::SetPBC wraps first (FixMonomerCoordinates) when box is given.
    Returns (E, grad[3 * atoms], virial[9]); E sums in monomer order."""
    kinds = np.asarray(kinds)
    nat = np.array([NAT_OF_KIND[int(k)] for k in kinds], np.int32)
    x = _arr(xyz).reshape(-1)
    if box is not None and len(box):
        x = fix_monomer_coordinates(x, box, nat)
    first = np.concatenate([[0], np.cumsum(nat)[:-1]])
    e = 0.0
    g = np.zeros_like(x)
    vir = np.zeros(6)
    for m, (k, f) in enumerate(zip(kinds, first)):
        if k >= 4:
            continue
        w = x[3 * f:3 * f + 3 * nat[m]]
        em, gm, vm = ps_1b(w) if k == 0 else nrg_1b("ch4_1b" if k == 1 else "co2_1b", w)
        e += em[0]
        g[3 * f:3 * f + 3 * nat[m]] = gm[0]
        vir += vm
    xx, xy, xz, yy, yz, zz = vir
    return e, g, np.array([xx, xy, xz, xy, yy, yz, xz, yz, zz])


def ref_ch4_1b(w):
    """The compiled reference's x1b_A1B4_v1x("ch4")::eval on CH4 monomers w [n][15]:
    (E[n], G[n][15], virial[9])."""
    L = ref()
    L.ref_ch4_1b.restype = None
    L.ref_ch4_1b.argtypes = [_dp, _dp, ctypes.c_size_t, _dp, _dp]
    w = np.ascontiguousarray(_arr(w).reshape(-1))
    n = len(w) // 15
    g = np.zeros(15 * n)
    e = np.zeros(n)
    v = np.zeros(9)
    L.ref_ch4_1b(_d(w), _d(g), n, _d(e), _d(v))
    return e, g.reshape(n, 15), v


def ref_co2_1b(w):
    """The compiled reference's x1b_A1B2_v1x("co2")::eval on CO2 monomers w [n][9]:
    (E[n], G[n][9], virial[9])."""
    L = refx()
    L.ref_co2_1b.restype = None
    L.ref_co2_1b.argtypes = [_dp, _dp, ctypes.c_size_t, _dp, _dp]
    w = np.ascontiguousarray(_arr(w).reshape(-1))
    n = len(w) // 9
    g = np.zeros(9 * n)
    e = np.zeros(n)
    v = np.zeros(9)
    L.ref_co2_1b(_d(w), _d(g), n, _d(e), _d(v))
    return e, g.reshape(n, 9), v


def ref_ps_1b(w, virial=None):
    """The compiled reference's ps::pot_nasa on waters w [n][9]: (E[n], G[n][9], virial[9])."""
    L = ref()
    L.ref_ps_1b.restype = None
    L.ref_ps_1b.argtypes = [_dp, _dp, ctypes.c_size_t, _dp, _dp]
    w = np.ascontiguousarray(_arr(w).reshape(-1))
    nw = len(w) // 9
    g = np.zeros(9 * nw)
    e = np.zeros(nw)
    v = np.zeros(9) if virial is None else np.array(virial, np.float64)
    L.ref_ps_1b(_d(w), _d(g), nw, _d(e), _d(v))
    return e, g.reshape(nw, 9), v


# ------------------------------------------------------------------------------------------
# Further MB-nrg families (halide-water 2B / 3B, CO2 2B / 3B, water 4B, alkali-ion-water 2B / 3B) and the system-level
# orchestration over monomers of any kind.

class _OrcXHolder(ctypes.Structure):
    _fields_ = [("nmon", ctypes.c_int), ("nat", ctypes.c_int * 4), ("nreal", ctypes.c_int), ("nsites", ctypes.c_int),
                ("nvsite", ctypes.c_int), ("vo", ctypes.c_int * 2), ("in_plane", ctypes.c_double),
                ("out_of_plane", ctypes.c_double), ("r_in", ctypes.c_double), ("r_out", ctypes.c_double),
                ("nvars", ctypes.c_int), ("form", _ip), ("k", _dp), ("d", _dp), ("s1", _ip), ("s2", _ip),
                ("npair", ctypes.c_int), ("pa", ctypes.c_int * 6), ("pb", ctypes.c_int * 6), ("nterm", ctypes.c_int),
                ("tcoef", _dp), ("tmask", _ip), ("skip_min", ctypes.c_int)]


XFORMS = {"exp0": 0, "coul0": 1, "exp": 2, "coul": 3}


class XHolder:
    """A further MB-nrg holder (mbx_amd/tables/<name>.holder.json) with its parameter set."""

    def __init__(self, name):
        import json

        h = json.load(open(os.path.join(TABLES, name + ".holder.json")))
        self.name, self.desc = name, h
        self.poly = h["poly"]
        self.coef = np.array(h["coef"], np.float64)
        v = h["vars"]
        self.form = np.array([XFORMS[t[0]] for t in v], np.int32)
        self.k = np.array([t[1] for t in v], np.float64)
        self.d = np.array([t[2] for t in v], np.float64)
        self.s1 = np.array([t[3] for t in v], np.int32)
        self.s2 = np.array([t[4] for t in v], np.int32)
        sw = h["switch"]
        self.tcoef = np.array([t[0] for t in sw["terms"]], np.float64)
        self.tmask = np.array([sum(1 << q for q in t[1]) for t in sw["terms"]], np.int32)
        vs = h["vsites"]
        pairs = sw["pairs"]
        self.c = _OrcXHolder(len(h["monomers"]), (ctypes.c_int * 4)(*(h["atoms"] + [0] * (4 - len(h["atoms"])))),
                             h["nreal"], h["nsites"], len(vs),
                             (ctypes.c_int * 2)(*([w["oxygen_site"] for w in vs] + [0] * (2 - len(vs)))),
                             vs[0]["in_plane"] if vs else 0.0, vs[0]["out_of_plane"] if vs else 0.0,
                             sw["inner"], sw["outer"], len(v), _i(self.form), _d(self.k), _d(self.d), _i(self.s1),
                             _i(self.s2), len(pairs), (ctypes.c_int * 6)(*([a for a, b in pairs] + [0] * (6 - len(pairs)))),
                             (ctypes.c_int * 6)(*([b for a, b in pairs] + [0] * (6 - len(pairs)))),
                             len(self.tcoef), _d(self.tcoef), _i(self.tmask), h.get("skip_min", len(pairs)))


_XHOLDERS = {}


def xholder(name):
    if name not in _XHOLDERS:
        _XHOLDERS[name] = XHolder(name)
    return _XHOLDERS[name]


def xnmer(name, mons, a=None):
    """One n-mer through a further MB-nrg holder: mons in the holder's monomer order (already
    imaged).  a: coefficients (None = the holder's parameter set).  Returns (E, [grad per
    monomer], virial(9), evaluated)."""
    h = xholder(name)
    tab = table(h.poly)
    ms = [_arr(m).reshape(-1).copy() for m in mons]
    gs = [np.zeros_like(m) for m in ms]
    vir = np.zeros(9)
    ev = ctypes.c_int()
    mp = (_dp * 4)(*[_d(m) for m in ms], *([None] * (4 - len(ms))))
    gp = (_dp * 4)(*[_d(g) for g in gs], *([None] * (4 - len(gs))))
    aa = _arr(a) if a is not None else h.coef
    e = lib().orc_xholder_eval(ctypes.byref(h.c), ctypes.byref(tab.c), _d(aa), mp, gp, _d(vir), ctypes.byref(ev))
    return e, gs, vir, ev.value


# Atoms per monomer of the ids the n-mer families here know (systools::SetUpMonomers,
# sys_tools.cpp:133-255).
NAT_OF_NAME = {"h2o": 3, "ch4": 5, "co2": 3, "f-": 1, "cl-": 1, "br-": 1, "i-": 1, "li+": 1, "na+": 1, "k+": 1,
               "rb+": 1, "cs+": 1}

# The holder of each sorted name tuple (energy2b.cpp:253-374, energy3b.cpp:219-260 with the
# alkali ions at energy2b.cpp:261-264, 279-282 and energy3b.cpp:223-228,
# energy4b.cpp:214-221); the holder's argument order is taken from its monomer list.
NMER_HOLDERS = {
    ("h2o", "h2o"): "mbpol_2b", ("ch4", "h2o"): "ch4h2o_2b", ("ch4", "ch4"): "ch4ch4_2b",
    ("f-", "h2o"): "f_h2o_2b", ("cl-", "h2o"): "cl_h2o_2b", ("br-", "h2o"): "br_h2o_2b", ("h2o", "i-"): "i_h2o_2b",
    ("co2", "co2"): "co2co2_2b",
    ("h2o", "h2o", "h2o"): "mbpol_3b", ("ch4", "h2o", "h2o"): "ch4h2o2_3b",
    ("f-", "h2o", "h2o"): "f_h2o2_3b", ("cl-", "h2o", "h2o"): "cl_h2o2_3b", ("br-", "h2o", "h2o"): "br_h2o2_3b",
    ("h2o", "h2o", "i-"): "i_h2o2_3b", ("co2", "co2", "co2"): "co2x3_3b",
    ("h2o", "h2o", "h2o", "h2o"): "h2ox4_4b",
    ("h2o", "li+"): "li_h2o_2b", ("h2o", "na+"): "na_h2o_2b", ("h2o", "k+"): "k_h2o_2b", ("h2o", "rb+"): "rb_h2o_2b",
    ("cs+", "h2o"): "cs_h2o_2b",
    ("h2o", "h2o", "li+"): "li_h2o2_3b", ("h2o", "h2o", "na+"): "na_h2o2_3b", ("h2o", "h2o", "k+"): "k_h2o2_3b",
    ("h2o", "h2o", "rb+"): "rb_h2o2_3b", ("cs+", "h2o", "h2o"): "cs_h2o2_3b",
}


def sort_names(names):
    """The name-sorting networks of e2b::get_2b_energy (energy2b.cpp:243-254), e3b::get_3b_energy
    (energy3b.cpp:185-216) and e4b::get_4b_energy (energy4b.cpp:178-214): returns the positions
    of the n-mer's monomers in sorted order (strict comparisons: equal names keep their order)."""
    n = list(names)
    ix = list(range(len(n)))

    def swp(i, j):
        n[i], n[j] = n[j], n[i]
        ix[i], ix[j] = ix[j], ix[i]

    if len(n) == 2:
        if n[1] < n[0]:
            swp(0, 1)
    elif len(n) == 3:
        if n[2] < n[0] and n[2] < n[1]:
            swp(0, 2)
        if n[0] > n[1] and n[0] > n[2]:
            swp(0, 2)
        elif n[1] > n[0] and n[1] > n[2]:
            swp(1, 2)
        if n[0] > n[1]:
            swp(0, 1)
    else:
        for i, j in ((0, 1), (1, 2), (2, 3), (0, 1), (1, 2), (0, 1)):
            if n[i] > n[j]:
                swp(i, j)
    return ix


def holder_order(names):
    """(holder name, positions of the n-mer's monomers in the holder's argument order) or
    (None, None) when the reference has no term for these names."""
    s = sort_names(names)
    key = tuple(names[i] for i in s)
    hn = NMER_HOLDERS.get(key)
    if hn is None:
        return None, None
    if hn.startswith("mbpol"):
        return hn, s
    import json

    mons = (holder(hn).desc if hn in NRG else json.load(open(os.path.join(TABLES, hn + ".holder.json"))))["monomers"]
    used, perm = set(), []
    for m in mons:  # each holder slot takes the first unused sorted monomer of its name
        j = next(i for i in s if i not in used and names[i] == m)
        used.add(j)
        perm.append(j)
    return hn, perm


def xsystem(xyz, names, box, cutoff2b, cutoff3b, cutoff4b=0.0, tags=None, ignore=(), nmers=False, cond=False):
    """2B + 3B (+ 4B) energy, gradient and virial of a system of any of the monomers above
    (names in input order, 3 doubles per atom), as System::Energy's n-body part does it:
    FixMonomerCoordinates under PBC, AddClusters lists, each n-mer imaged next to its first
    listed monomer (GetCloseNmerImage, sys_tools.cpp:453-468) and passed to the holder of its
    names in the holder's argument order.  ignore: sorted name tuples whose polynomials are
    skipped (the System's ignore lists).  Returns dict(e2b, e3b, e4b, grad, virial, counts
    {order: (listed, evaluated)}, [per-n-mer (order, ids, E)], [cond]).  cond: per order, the sum
    over n-mers of the magnitude of the polynomial's cancelling terms (sum_k |a_k| M_k sw, the
    further families' holders evaluated with |a|; |E| for the MB-pol and CH4 holders), the scale
    of the rounding error any evaluation order leaves; cond_mon: per monomer, the same sum over
    the n-mers that hold it."""
    names = list(names)
    nat = np.array([NAT_OF_NAME[n] for n in names], np.int32)
    first = np.concatenate([[0], np.cumsum(nat)[:-1]]).astype(np.int64)
    x = _arr(xyz).reshape(-1).copy()
    b = _arr(box) if box is not None and len(box) else None
    if b is not None:
        x = fix_monomer_coordinates(x, b, nat)
    tg = first.astype(np.int32) if tags is None else np.ascontiguousarray(tags, np.int32)
    fx = np.array([x[3 * f:3 * f + 3] for f in first]).reshape(-1)
    grad = np.zeros(len(x))
    vir = np.zeros(9)
    energies = {2: 0.0, 3: 0.0, 4: 0.0}
    counts, per = {}, []
    conds = {2: 0.0, 3: 0.0, 4: 0.0}
    cond_mon = np.zeros(len(names))
    ign = {tuple(sorted(t)) for t in ignore}
    for order, cut in ((2, cutoff2b), (3, cutoff3b), (4, cutoff4b)):
        if cut <= 0.0:
            counts[order] = (0, 0)
            continue
        lst = clusters(order, cut, fx, b, tags=tg)
        nev = 0
        for ids in lst:
            ids = [int(i) for i in ids]
            nm = [names[i] for i in ids]
            if tuple(sorted(nm)) in ign:
                continue
            hn, perm = holder_order(nm)
            if hn is None:
                continue
            w = [x[3 * first[i]:3 * (first[i] + nat[i])].copy() for i in ids]
            if b is not None:
                for j in range(1, order):
                    w[j] = close_image(b, w[0], w[j])
            ws = [w[j] for j in perm]
            if hn == "mbpol_2b":
                e, g1, g2, v, ev = mbpol2b(ws[0], ws[1])
                gs = [g1, g2]
            elif hn == "mbpol_3b":
                e, g1, g2, g3, v, ev = mbpol3b(ws[0], ws[1], ws[2])
                gs = [g1, g2, g3]
            elif hn in NRG:
                e, gs, v, ev = nrg(hn, ws)
            else:
                e, gs, v, ev = xnmer(hn, ws)
            if cond:
                if hn in NRG or hn.startswith("mbpol"):
                    cn = abs(e)
                else:
                    cn = abs(xnmer(hn, ws, a=np.abs(xholder(hn).coef))[0])
                conds[order] += cn
                cond_mon[ids] += cn
            nev += ev
            energies[order] += e
            vir += v
            for j, p in enumerate(perm):
                i = ids[p]
                grad[3 * first[i]:3 * (first[i] + nat[i])] += gs[j]
            if nmers:
                per.append((order, ids, e))
        counts[order] = (len(lst), nev)
    out = dict(e2b=energies[2], e3b=energies[3], e4b=energies[4], grad=grad, virial=vir, counts=counts, xyz=x)
    if nmers:
        out["nmers"] = per
    if cond:
        out["cond"] = conds
        out["cond_mon"] = cond_mon
    return out


# The compiled reference's further holders (oracle/_ref/libmbxref_x.so), when built.
_REFX = None
REFX_PATH = os.path.join(HERE, "_ref", "libmbxref_x.so")


def refx_available():
    return os.path.exists(REFX_PATH)


def refx():
    global _REFX
    if _REFX is None:
        L = ctypes.CDLL(REFX_PATH)
        sz = ctypes.c_size_t
        for name in ("ref_x2b_halide_h2o", "ref_x3b_halide_h2o2", "ref_x2b_co2", "ref_x3b_co2", "ref_x4b_h2o",
                     "ref_x2b_alkali_h2o", "ref_x3b_alkali_h2o2"):
            getattr(L, name).restype = ctypes.c_double
        L.ref_x2b_halide_h2o.argtypes = [ctypes.c_char_p, _dp, _dp, sz, _dp, _dp, _dp]
        L.ref_x3b_halide_h2o2.argtypes = [ctypes.c_char_p, _dp, _dp, _dp, sz, _dp, _dp, _dp, _dp]
        L.ref_x2b_co2.argtypes = [_dp, _dp, sz, _dp, _dp, _dp]
        L.ref_x3b_co2.argtypes = [_dp, _dp, _dp, sz, _dp, _dp, _dp, _dp]
        L.ref_x4b_h2o.argtypes = [_dp, _dp, _dp, _dp, sz, _dp, _dp, _dp, _dp, _dp]
        L.ref_x2b_alkali_h2o.argtypes = [ctypes.c_char_p, _dp, _dp, sz, _dp, _dp, _dp]
        L.ref_x3b_alkali_h2o2.argtypes = [ctypes.c_char_p, _dp, _dp, _dp, sz, _dp, _dp, _dp, _dp]
        _REFX = L
    return _REFX


def ref_xnmer(name, mons):
    """One n-mer through the compiled reference holder of a further family (holder argument
    order, as xnmer): (E, [grad per monomer], virial(9))."""
    L = refx()
    ms = [_arr(m).reshape(-1).copy() for m in mons]
    gs = [np.zeros_like(m) for m in ms]
    vir = np.zeros(9)
    h = xholder(name).desc
    if h["poly"] == "alkali_h2o_2b":
        e = L.ref_x2b_alkali_h2o(h["monomers"][1].encode(), _d(ms[0]), _d(ms[1]), 1, _d(gs[0]), _d(gs[1]), _d(vir))
    elif h["poly"] == "alkali_h2o2_3b":
        e = L.ref_x3b_alkali_h2o2(h["monomers"][2].encode(), _d(ms[0]), _d(ms[1]), _d(ms[2]), 1, _d(gs[0]),
                                  _d(gs[1]), _d(gs[2]), _d(vir))
    elif name.endswith("_h2o_2b"):
        e = L.ref_x2b_halide_h2o(h["monomers"][0].encode(), _d(ms[0]), _d(ms[1]), 1, _d(gs[0]), _d(gs[1]), _d(vir))
    elif name.endswith("_h2o2_3b"):
        e = L.ref_x3b_halide_h2o2(h["monomers"][0].encode(), _d(ms[0]), _d(ms[1]), _d(ms[2]), 1, _d(gs[0]),
                                  _d(gs[1]), _d(gs[2]), _d(vir))
    elif name == "co2co2_2b":
        e = L.ref_x2b_co2(_d(ms[0]), _d(ms[1]), 1, _d(gs[0]), _d(gs[1]), _d(vir))
    elif name == "co2x3_3b":
        e = L.ref_x3b_co2(_d(ms[0]), _d(ms[1]), _d(ms[2]), 1, _d(gs[0]), _d(gs[1]), _d(gs[2]), _d(vir))
    elif name == "h2ox4_4b":
        e = L.ref_x4b_h2o(_d(ms[0]), _d(ms[1]), _d(ms[2]), _d(ms[3]), 1, _d(gs[0]), _d(gs[1]), _d(gs[2]), _d(gs[3]),
                          _d(vir))
    else:
        raise KeyError(name)
    return e, gs, vir
