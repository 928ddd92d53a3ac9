"""ctypes binding of the mbxg C ABI (include/mbxg.h) -- the GPU engine behind bblock::System.

The shared library ``mbx_amd/libmbxg.so`` is built in-tree (``__graft_entry__.build()`` or
``make -C mbx_amd/csrc``).  There is no CPU fallback: if the library or a GPU is missing,
constructing an :class:`Engine` raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmbxg.so")

H2O, CH4, OTHER, CO2, F, CL, BR, I = 0, 1, 2, 3, 4, 5, 6, 7
LI, NA, K, RB, CS = 8, 9, 10, 11, 12
TERM_2B, TERM_3B, LISTS_ONLY, TERM_1B, TERM_4B = 1, 2, 4, 8, 16
KIND_OF = {"h2o": H2O, "ch4": CH4, "co2": CO2, "f-": F, "cl-": CL, "br-": BR, "i-": I, "li+": LI, "na+": NA,
           "k+": K, "rb+": RB, "cs+": CS}
NAT_OF = {H2O: 3, CH4: 5, CO2: 3, F: 1, CL: 1, BR: 1, I: 1, LI: 1, NA: 1, K: 1, RB: 1, CS: 1}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

_LIB = None

# name: (restype, argtypes) for every function include/mbxg.h declares
PROTOTYPES = {
    "mbxg_last_error": (ctypes.c_char_p, []),
    "mbxg_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    "mbxg_destroy": (ctypes.c_int, [_vp]),
    "mbxg_set_topology": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_ubyte)]),
    "mbxg_set_box": (ctypes.c_int, [_vp, _dp]),
    "mbxg_set_cutoffs": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
    "mbxg_set_coefficients": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp, ctypes.c_int]),
    "mbxg_set_coords": (ctypes.c_int, [_vp, _dp]),
    "mbxg_set_coords_device": (ctypes.c_int, [_vp, _vp]),
    "mbxg_compute": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "mbxg_get_results": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "mbxg_get_counts": (ctypes.c_int, [_vp, _i64p]),
    "mbxg_get_lists": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _i32p, ctypes.c_int64]),
    "mbxg_get_nmer_energies": (ctypes.c_int, [_vp, ctypes.c_int, _i32p, _dp, ctypes.c_int64]),
    "mbxg_get_coords": (ctypes.c_int, [_vp, _dp]),
    "mbxg_device_outputs": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "mbxg_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "mbxg_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mbxg_get_timings": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "mbxg_synchronize": (ctypes.c_int, [_vp]),
    "mbxg_set_batch": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "mbxg_set_families": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mbxg_set_use_ghost": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mbxg_eval_1b": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _dp]),
    "mbxg_get_energies": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "mbxg_fix_coordinates": (ctypes.c_int, [_vp, _dp, _dp]),
    "mbxg_set_tuning": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int]),
    "mbxg_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mbxg_get_tuning": (ctypes.c_int, [_vp, ctypes.c_char_p, _ip]),
    "mbxg_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "mbxg_get_batch_results": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "mbxg_set_cutoff4b": (ctypes.c_int, [_vp, ctypes.c_double]),
    "mbxg_get_term_energies": (ctypes.c_int, [_vp, _dp]),
    "mbxg_get_xcounts": (ctypes.c_int, [_vp, _i64p]),
    "mbxg_get_tetramers": (ctypes.c_int, [_vp, _i32p, ctypes.c_int64, _i64p]),
    "mbxg_ad_shape": (ctypes.c_int, [ctypes.c_int, _ip, _ip, _ip]),
    "mbxg_ad_defaults": (ctypes.c_int, [ctypes.c_int, _dp, _dp]),
    "mbxg_ad_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int, ctypes.c_int]),
    "mbxg_ad_forward": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "mbxg_ad_forward_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mbxg_ad_backward": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                        ctypes.c_size_t, ctypes.c_int, _vp]),
}

FAMILIES = ("mbpol_2b", "ch4h2o_2b", "ch4ch4_2b", "mbpol_3b", "ch4h2o2_3b",
            # further MB-nrg families (bits 5.. of mbxg_set_families)
            "f_h2o_2b", "cl_h2o_2b", "br_h2o_2b", "i_h2o_2b", "co2co2_2b", "f_h2o2_3b", "cl_h2o2_3b", "br_h2o2_3b",
            "i_h2o2_3b", "co2x3_3b", "h2ox4_4b",
            "li_h2o_2b", "na_h2o_2b", "k_h2o_2b", "rb_h2o_2b", "cs_h2o_2b",
            "li_h2o2_3b", "na_h2o2_3b", "k_h2o2_3b", "rb_h2o2_3b", "cs_h2o2_3b")
XFAMILIES = FAMILIES[5:]
TIMING_STAGES = ("fix", "lists_count", "lists_emit", "vars_2b", "poly_2b", "chain_2b", "vars_3b", "poly_3b",
                 "chain_3b", "total", "ch4h2o2_3b")


class MbxgError(RuntimeError):
    pass


def lib():
    """Load libmbxg.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise MbxgError("mbx_amd/libmbxg.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise MbxgError(lib().mbxg_last_error().decode() or ("mbxg status %d" % rc))


def _d(a):
    return a.ctypes.data_as(_dp)


class Engine:
    """One mbxg context (one HIP stream, its own device memory)."""

    def __init__(self, device=0):
        self._c = _vp()
        _check(lib().mbxg_create(ctypes.byref(self._c), device))
        self.natoms = 0
        self.nmon = 0

    def close(self):
        if self._c:
            lib().mbxg_destroy(self._c)
            self._c = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_topology(self, kinds, nats=None, tags=None, islocal=None):
        kinds = np.ascontiguousarray(kinds, np.int32)
        n = len(kinds)
        nats = np.ascontiguousarray([NAT_OF[k] for k in kinds] if nats is None else nats, np.int32)
        tags = np.ascontiguousarray(np.arange(n) if tags is None else tags, np.int32)
        loc = None if islocal is None else np.ascontiguousarray(islocal, np.uint8)
        _check(lib().mbxg_set_topology(self._c, n, kinds.ctypes.data_as(_ip), nats.ctypes.data_as(_ip),
                                       tags.ctypes.data_as(_ip),
                                       loc.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)) if loc is not None else None))
        self.nmon = n
        self.natoms = int(nats.sum())
        self.nconf = 1

    def set_families(self, families=FAMILIES):
        """Evaluate only these n-mer families (names as FAMILIES); the others stay listed."""
        _check(lib().mbxg_set_families(self._c, sum(1 << i for i, f in enumerate(FAMILIES) if f in families)))

    def set_cutoff4b(self, c4):
        """The 4B tetramer cutoff (0 disables the 4B term)."""
        _check(lib().mbxg_set_cutoff4b(self._c, float(c4)))

    def term_energies(self):
        """(E1B, E2B, E3B, E4B) of the last compute."""
        e = np.zeros(4)
        _check(lib().mbxg_get_term_energies(self._c, e.ctypes.data_as(_dp)))
        return e

    def xcounts(self):
        """{further family: evaluated n-mers}, and the listed tetramers under "tetramers"."""
        c = np.zeros(len(XFAMILIES) + 1, np.int64)
        _check(lib().mbxg_get_xcounts(self._c, c.ctypes.data_as(_i64p)))
        out = {f: int(c[i]) for i, f in enumerate(XFAMILIES)}
        out["tetramers"] = int(c[-1])
        return out

    def tetramers(self):
        """Listed tetramers of the last compute, (n, 4) internal monomer indices."""
        n = ctypes.c_int64()
        _check(lib().mbxg_get_tetramers(self._c, None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 4), np.int32)
        _check(lib().mbxg_get_tetramers(self._c, out.ctypes.data_as(_i32p), n.value, ctypes.byref(n)))
        return out[:n.value]

    def set_use_ghost(self, on=True):
        """Neighbour search over local and ghost monomers, n-mers found from local ones only
        (see mbxg_set_use_ghost)."""
        _check(lib().mbxg_set_use_ghost(self._c, int(bool(on))))

    def energies(self):
        """(E1B, E2B, E3B) of the last compute (see mbxg_get_energies)."""
        e = [ctypes.c_double() for _ in range(3)]
        _check(lib().mbxg_get_energies(self._c, *[ctypes.byref(v) for v in e]))
        return tuple(v.value for v in e)

    def fix_coordinates(self, xyz):
        """FixMonomerCoordinates of host coordinates under the current box (mbxg_fix_coordinates)."""
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        out = np.zeros_like(x)
        _check(lib().mbxg_fix_coordinates(self._c, _d(x), _d(out)))
        return out

    def eval_1b(self, do_grads=True):
        """1B energy of the local monomers (MB-pol water, MB-nrg CH4) at the current coordinates:
        (E1B, gradient per atom or None, virial[9]) (see mbxg_eval_1b)."""
        e = ctypes.c_double()
        g = np.zeros(3 * self.natoms) if do_grads else None
        v = np.zeros(9)
        _check(lib().mbxg_eval_1b(self._c, int(bool(do_grads)), ctypes.byref(e), _d(g) if do_grads else None,
                                  _d(v)))
        return e.value, g, v

    def set_batch(self, mon_offsets):
        """Split the monomers into configurations: configuration c = monomers
        [mon_offsets[c], mon_offsets[c+1])."""
        off = np.ascontiguousarray(mon_offsets, np.int32)
        _check(lib().mbxg_set_batch(self._c, len(off) - 1, off.ctypes.data_as(_ip)))
        self.nconf = len(off) - 1

    def batch_results(self):
        n = getattr(self, "nconf", 1)
        e2, e3, v = np.zeros(n), np.zeros(n), np.zeros(9 * n)
        _check(lib().mbxg_get_batch_results(self._c, _d(e2), _d(e3), _d(v)))
        return e2, e3, v.reshape(n, 9)

    def set_graph(self, on=True):
        """Replay the steady-state pipeline as one captured HIP graph (see mbxg_set_graph)."""
        _check(lib().mbxg_set_graph(self._c, int(bool(on))))

    def set_tuning(self, key, value):
        _check(lib().mbxg_set_tuning(self._c, key.encode(), int(value)))

    def tuning(self, key):
        v = ctypes.c_int()
        _check(lib().mbxg_get_tuning(self._c, key.encode(), ctypes.byref(v)))
        return v.value

    def autotune(self, terms=TERM_2B | TERM_3B, do_grads=True, reps=3):
        """Pick the fastest polynomial schedule and fusions for the current inputs (see mbxg_autotune)."""
        _check(lib().mbxg_autotune(self._c, int(terms), int(bool(do_grads)), int(reps)))
        return {k: self.tuning(k) for k in ("split_2b", "split_3b", "split_ch4h2o2_3b", "fuse_chain", "fuse_vars")}

    def set_box(self, box):
        if box is None or len(box) == 0:
            _check(lib().mbxg_set_box(self._c, None))
        else:
            b = np.ascontiguousarray(box, np.float64)
            if b.size != 9:
                raise ValueError("box needs 9 numbers")
            _check(lib().mbxg_set_box(self._c, _d(b)))

    def set_cutoffs(self, c2, c3):
        _check(lib().mbxg_set_cutoffs(self._c, float(c2), float(c3)))

    def set_coefficients(self, model, a=None):
        if a is None:
            _check(lib().mbxg_set_coefficients(self._c, model.encode(), None, 0))
        else:
            a = np.ascontiguousarray(a, np.float64)
            _check(lib().mbxg_set_coefficients(self._c, model.encode(), _d(a), len(a)))

    def set_coords(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        if xyz.size != 3 * self.natoms:
            raise ValueError("expected %d coordinates, got %d" % (3 * self.natoms, xyz.size))
        _check(lib().mbxg_set_coords(self._c, _d(xyz)))

    def set_coords_device(self, ptr):
        _check(lib().mbxg_set_coords_device(self._c, ctypes.c_void_p(ptr)))

    def compute(self, terms=TERM_2B | TERM_3B, do_grads=True):
        _check(lib().mbxg_compute(self._c, int(terms), int(bool(do_grads))))

    def results(self, grad=True):
        e2, e3 = ctypes.c_double(), ctypes.c_double()
        g = np.zeros(3 * self.natoms) if grad else None
        v = np.zeros(9)
        _check(lib().mbxg_get_results(self._c, ctypes.byref(e2), ctypes.byref(e3), _d(g) if grad else None, _d(v)))
        return e2.value, e3.value, g, v

    def counts(self):
        c = np.zeros(4, np.int64)
        _check(lib().mbxg_get_counts(self._c, c.ctypes.data_as(_i64p)))
        return c

    def lists(self):
        c = self.counts()
        d = np.zeros((max(int(c[0]), 1), 2), np.int32)
        t = np.zeros((max(int(c[2]), 1), 3), np.int32)
        _check(lib().mbxg_get_lists(self._c, d.ctypes.data_as(_i32p), int(c[0]), t.ctypes.data_as(_i32p), int(c[2])))
        return d[:c[0]], t[:c[2]]

    def nmer_energies(self, order):
        c = self.counts()
        n = int(c[1] if order == 2 else c[3])
        idx = np.zeros((max(n, 1), order), np.int32)
        e = np.zeros(max(n, 1))
        _check(lib().mbxg_get_nmer_energies(self._c, order, idx.ctypes.data_as(_i32p), _d(e), n))
        return idx[:n], e[:n]

    def coords(self):
        x = np.zeros(3 * self.natoms)
        _check(lib().mbxg_get_coords(self._c, _d(x)))
        return x

    def device_outputs(self):
        g, t = _vp(), _vp()
        _check(lib().mbxg_device_outputs(self._c, ctypes.byref(g), ctypes.byref(t)))
        return g.value, t.value

    def stream(self):
        s = _vp()
        _check(lib().mbxg_stream(self._c, ctypes.byref(s)))
        return s.value

    def set_timing(self, on=True):
        _check(lib().mbxg_set_timing(self._c, int(bool(on))))

    def timings(self):
        """Stage durations (ms) of the last compute, measured with HIP events on the engine stream."""
        t = np.zeros(len(TIMING_STAGES))
        _check(lib().mbxg_get_timings(self._c, _d(t), len(t)))
        return dict(zip(TIMING_STAGES, t))

    def synchronize(self):
        _check(lib().mbxg_synchronize(self._c))


def water_engine(nmon, box, cutoff2b, cutoff3b, device=0):
    """An engine set up for nmon waters (tags = input order)."""
    e = Engine(device)
    e.set_topology([H2O] * nmon)
    e.set_box(box)
    e.set_cutoffs(cutoff2b, cutoff3b)
    return e
