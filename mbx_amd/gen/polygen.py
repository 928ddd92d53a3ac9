"""Generate straight-line FP64 energy+gradient code for a permutationally-invariant polynomial.

The polynomial is given by a monomial table (mbx_amd/tables/<name>.tbl):

    E(x) = sum_k a_k * sum_{m in S_k} prod_{i in m} x_i ,   deg(m) <= 4.

The generated device function evaluates E and dE/dx for ONE n-mer per lane,
with every variable and gradient accumulator in registers and every
coefficient a uniform (scalar) load.  The schedule is a reverse-mode walk of a
monomial trie built so that the adjoint of every trie node is a product of its
ancestors' edge variables, which the forward walk already has:

    level-3 node (p,q,r):  R = c_pqr + sum_l C_l x_l          (leaves l)
                           g_l += C_l * pi_pqr                 (pi_pqr = x_p x_q x_r)
                           g_r += pi_pq * R ;  R_pq += x_r * R
    level-2 node (p,q):    g_q += x_p * R_pq ; R_p += x_q * R_pq
    level-1 node p:        g_p += R_p ;        E += x_p * R_p

so a degree-4 monomial costs two FMAs (one for E, one for the gradient) and
every interior node three.  Degree-4 monomials are attached to level-3 nodes
by a greedy set cover (each monomial may drop any one of its four variables),
which keeps the number of level-3 nodes low.

Output: a C++ include with
    __device__ __forceinline__ double <fn>(mbxg_cptr C, const double* x, double* g)
where C is the derived coefficient table: C[q] = a[ck[q]] * cs[q] with the
(ck, cs) map emitted alongside (the scale cs is 1 except for symmetry-reduced
tables).  A second function <fn>_e evaluates E only.

Regions of the op stream are separated by MBXG_NODE_FENCE(C, v) (poly_support.hpp): a
scheduling barrier plus an opaque redefinition of the coefficient pointer that consumes the
previous region's last result v.  It keeps each node's coefficient
loads and temporaries local to the node, so the register allocator never sees tens of
thousands of simultaneously live values, and keeps compile time linear in the code size.
"""
import collections
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
TABLES = os.path.join(ROOT, "mbx_amd", "tables")


def load_table(name):
    """Returns (nvars, coef list, [(k, mono tuple)])."""
    lines = [l.split("#")[0].strip() for l in open(os.path.join(TABLES, name + ".tbl"))]
    lines = [l for l in lines if l]
    it = iter(lines)
    hdr = {}
    for l in it:
        if l == "coef":
            break
        key, v = l.split()
        hdr[key] = int(v)
    coef = [float(next(it)) for _ in range(hdr["ncoef"])]
    assert next(it).startswith("monomials")
    monos = []
    for _ in range(hdr["nmono"]):
        v = [int(t) for t in next(it).split()]
        monos.append((v[0], tuple(v[2:])))
    return hdr["nvars"], coef, monos


class Schedule:
    """Trie schedule over a list of (coefficient slot, monomial) terms."""

    def __init__(self, nvars, terms):
        self.nvars = nvars
        # coefficient slots are opaque ints indexing the derived table C
        by_deg = collections.defaultdict(list)
        for q, m in terms:
            by_deg[len(m)].append((q, tuple(sorted(m))))
        d4 = by_deg[4]
        # greedy cover of degree-4 monomials by level-3 nodes
        cand = collections.defaultdict(set)
        for idx, (_, m) in enumerate(d4):
            for r in range(4):
                cand[m[:r] + m[r + 1:]].add(idx)
        leaves3 = {}  # node -> list of (q, leaf var)
        left = set(range(len(d4)))
        heap_order = sorted(cand, key=lambda n: (-len(cand[n]), n))
        # lazy greedy: re-evaluate the best candidate until its size is current
        import heapq
        heap = [(-len(cand[n]), n) for n in heap_order]
        heapq.heapify(heap)
        while left:
            negsz, n = heapq.heappop(heap)
            cur = cand[n] & left
            if len(cur) < -negsz:
                if cur:
                    heapq.heappush(heap, (-len(cur), n))
                continue
            lst = []
            for idx in sorted(cur):
                q, m = d4[idx]
                rest = list(m)
                for v in n:
                    rest.remove(v)
                lst.append((q, rest[0]))
            leaves3[n] = lst
            left -= cur
        self.const3 = {}
        leaves2 = collections.defaultdict(list)
        for q, m in by_deg[3]:
            if m in leaves3:
                self.const3[m] = q
            else:
                leaves2[m[:2]].append((q, m[2]))
        nodes2 = set(n[:2] for n in leaves3) | set(leaves2)
        self.const2 = {}
        leaves1 = collections.defaultdict(list)
        for q, m in by_deg[2]:
            if m in nodes2:
                self.const2[m] = q
            else:
                leaves1[m[:1]].append((q, m[1]))
        nodes1 = set(n[:1] for n in nodes2) | set(leaves1)
        self.const1 = {}
        self.leaves0 = []
        for q, m in by_deg[1]:
            if m in nodes1:
                self.const1[m] = q
            else:
                self.leaves0.append((q, m[0]))
        self.const0 = [q for q, m in by_deg[0]]
        self.leaves3 = {n: sorted(l, key=lambda t: t[1]) for n, l in leaves3.items()}
        self.leaves2 = {n: sorted(l, key=lambda t: t[1]) for n, l in leaves2.items()}
        self.leaves1 = {n: sorted(l, key=lambda t: t[1]) for n, l in leaves1.items()}
        self.nodes1 = sorted(nodes1)
        self.children1 = collections.defaultdict(list)
        for n in sorted(nodes2):
            self.children1[n[:1]].append(n)
        self.children2 = collections.defaultdict(list)
        for n in sorted(leaves3):
            self.children2[n[:2]].append(n)

    def counts(self):
        n4 = sum(len(v) for v in self.leaves3.values())
        n3l = sum(len(v) for v in self.leaves2.values())
        n2l = sum(len(v) for v in self.leaves1.values())
        ops = (2 * n4 + 3 * len(self.leaves3) + len(self.const3) + 2 * n3l + 3 * sum(len(v) for v in self.children1.values())
               + len(self.const2) + 2 * n2l + 2 * len(self.nodes1) + len(self.const1) + 2 * len(self.leaves0))
        return dict(level3_nodes=len(self.leaves3), deg4=n4, deg3_leaves=n3l, deg3_const=len(self.const3),
                    level2_nodes=sum(len(v) for v in self.children1.values()), deg2_leaves=n2l,
                    deg2_const=len(self.const2), level1_nodes=len(self.nodes1), fp64_ops_eg=ops)

    # ---------------------------------------------------------------------------------------
    def partition(self, nparts):
        """Split the schedule into nparts pieces of near-equal op count.  Items are level-2
        subtrees, each level-1 node's own terms ('x1'), and the degree-0/1 terms ('z');
        everything is linear in the pieces, so value and gradient are sums over parts.
        Longest-processing-time assignment; returns {item: part}."""
        items = []
        for n1 in self.nodes1:
            for n2 in self.children1.get(n1, []):
                c = 3 + 2 * len(self.leaves2.get(n2, [])) + (1 if n2 in self.const2 else 0)
                for n3 in self.children2.get(n2, []):
                    c += 3 + 2 * len(self.leaves3[n3]) + (1 if n3 in self.const3 else 0)
                items.append((c, ("n2", n2)))
            items.append((2 + 2 * len(self.leaves1.get(n1, [])) + (1 if n1 in self.const1 else 0), ("x1", n1)))
        items.append((1 + 2 * len(self.leaves0), ("z",)))
        load = [0] * nparts
        assign = {}
        for c, it in sorted(items, key=lambda t: (-t[0], t[1])):
            k = min(range(nparts), key=lambda j: (load[j], j))
            assign[it] = k
            load[k] += c
        self.part_load = load
        return assign

    def ops(self, grad=True, part=None, assign=None):
        """Linear op list.  Each op is (text template, coefficient slot or None); the template
        contains '{c}' where the coefficient value goes.  Temporaries get unique names, so the
        whole body lives at function scope.  With `part`, only the terms assigned to that part
        (see partition) are emitted and E starts from 0."""
        o = []
        w = lambda t, q=None: o.append((t, q))
        mine = lambda it: part is None or assign[it] == part
        if part is not None:
            w("double E = E0;")
        if mine(("z",)):
            if part is None:
                w("double E = %s;" % ("{c}" if self.const0 else "0.0"), self.const0[0] if self.const0 else None)
            elif self.const0:
                w("E += {c};", self.const0[0])
            for q, l in self.leaves0:
                w("E = __fma_rn({c}, x[%d], E);" % l, q)
                if grad:
                    w("g[%d] += {c};" % l, q)
        i1 = i2 = i3 = 0
        for n1 in self.nodes1:
            p = n1[0]
            kids = [n2 for n2 in self.children1.get(n1, []) if mine(("n2", n2))]
            own = mine(("x1", n1))
            if not kids and not own:
                continue
            R1 = "R1_%d" % i1
            i1 += 1
            if own and n1 in self.const1:
                w("double %s = {c};" % R1, self.const1[n1])
            else:
                w("double %s = 0.0;" % R1)
            for n2 in kids:
                qv = n2[1]
                P2, R2 = "P2_%d" % i2, "R2_%d" % i2
                i2 += 1
                w("const double %s = x[%d] * x[%d];" % (P2, p, qv))
                if n2 in self.const2:
                    w("double %s = {c};" % R2, self.const2[n2])
                else:
                    w("double %s = 0.0;" % R2)
                for n3 in self.children2.get(n2, []):
                    r = n3[2]
                    P3, R3 = "P3_%d" % i3, "R3_%d" % i3
                    i3 += 1
                    if grad:
                        w("const double %s = %s * x[%d];" % (P3, P2, r))
                    first = True
                    for q, l in self.leaves3[n3]:
                        if first and n3 not in self.const3:
                            w("double %s = {c} * x[%d];" % (R3, l), q)
                        else:
                            if first:
                                w("double %s = {c};" % R3, self.const3[n3])
                            w("%s = __fma_rn({c}, x[%d], %s);" % (R3, l, R3), q)
                        first = False
                        if grad:
                            w("g[%d] = __fma_rn({c}, %s, g[%d]);" % (l, P3, l), q)
                    if grad:
                        w("g[%d] = __fma_rn(%s, %s, g[%d]);" % (r, P2, R3, r))
                    w("%s = __fma_rn(x[%d], %s, %s);" % (R2, r, R3, R2))
                for q, l in self.leaves2.get(n2, []):
                    w("%s = __fma_rn({c}, x[%d], %s);" % (R2, l, R2), q)
                    if grad:
                        w("g[%d] = __fma_rn({c}, %s, g[%d]);" % (l, P2, l), q)
                if grad:
                    w("g[%d] = __fma_rn(x[%d], %s, g[%d]);" % (qv, p, R2, qv))
                w("%s = __fma_rn(x[%d], %s, %s);" % (R1, qv, R2, R1))
            for q, l in (self.leaves1.get(n1, []) if own else []):
                w("%s = __fma_rn({c}, x[%d], %s);" % (R1, l, R1), q)
                if grad:
                    w("g[%d] = __fma_rn({c}, x[%d], g[%d]);" % (l, p, l), q)
            if grad:
                w("g[%d] += %s;" % (p, R1))
            w("E = __fma_rn(x[%d], %s, E);" % (p, R1))
        return o

    def emit(self, fn, grad=True, region_coefs=12, region_ops=64, part=None, assign=None):
        """C++ source of the evaluation function.

        The op list is cut into regions of at most `region_coefs` distinct coefficients and
        about `region_ops` arithmetic ops.  Region r's coefficients are loaded (as named SGPR
        values) during region r-1, and a fence separates regions, so scalar-load latency is
        covered by the previous region's FMAs while only two regions' coefficients are live.
        """
        ops = self.ops(grad, part, assign)
        regions = [[]]
        coefs = [[]]
        nar = 0
        for t, q in ops:
            cur = coefs[-1]
            arith = True
            if q is not None and q not in cur and (len(cur) >= region_coefs or nar >= region_ops):
                regions.append([])
                coefs.append([])
                nar = 0
                cur = coefs[-1]
            if q is not None and q not in cur:
                cur.append(q)
            regions[-1].append((t, q))
            nar += arith
        out = []
        w = out.append
        e0 = ", double E0" if part is not None else ""
        if grad:
            w("__device__ __forceinline__ double %s(mbxg_cptr C, const double* x, double* g%s) {" % (fn, e0))
            if part is None:
                w("    for (int i = 0; i < %d; ++i) g[i] = 0.0;" % self.nvars)
        else:
            w("__device__ __forceinline__ double %s(mbxg_cptr C, const double* x%s) {" % (fn, e0))
        name = lambda r, q: "k%d_%d" % (r, coefs[r].index(q))
        addr = lambda r, q: q
        for j, q in enumerate(coefs[0]):
            w("    const double %s = C[%d];" % (name(0, q), addr(0, q)))
        last = None
        for r, reg in enumerate(regions):
            # the fence consumes the previous region's last result, so no load of this or the
            # next region can be hoisted above that region's arithmetic
            w("    MBXG_NODE_FENCE(C, %s);" % (last or ("E0" if part is not None else "x[0]")))
            if r + 1 < len(regions):
                for q in coefs[r + 1]:
                    w("    const double %s = C[%d];" % (name(r + 1, q), addr(r + 1, q)))
            for t, q in reg:
                w("    " + (t.replace("{c}", name(r, q)) if q is not None else t))
                if q is not None:
                    # a result that needs this region's coefficients: the next fence cannot be
                    # hoisted above it
                    last = _lhs(t)
        w("    return E;")
        w("}")
        return "\n".join(out) + "\n"

def _lhs(stmt):
    """Name assigned by a generated statement ('double R3_1 = ...' -> 'R3_1', 'g[4] = ...' -> 'g[4]')."""
    lhs = stmt.split("=")[0].strip()
    for kw in ("const double ", "double "):
        if lhs.startswith(kw):
            lhs = lhs[len(kw):]
    return lhs.replace("+", "").strip()


def full_terms(monos):
    """Terms of the full polynomial: coefficient slot q == coefficient index k, scale 1."""
    return [(k, m) for k, m in monos], None


# The polynomial is cut into NPARTS pieces so that 1, 2, 4 or 8 wavefronts can share one
# block of 64 n-mers (each wavefront evaluating NPARTS / P pieces).  A piece takes the running
# value E0 and returns it with its own terms added; its first fence consumes E0, so no
# arithmetic of the previous piece can sink into it.
NPARTS = 8


def write_if_changed(path, text):
    """Write only when the content differs, so make does not rebuild unchanged kernels."""
    try:
        if open(path).read() == text:
            return
    except OSError:
        pass
    with open(path, "w") as f:
        f.write(text)


def coef_source(name, coef):
    out = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.tbl -- do not edit." % name,
           "#pragma once", "static const int kNCoef_%s = %d;" % (name, len(coef)),
           "static const double kFit_%s[%d] = {" % (name, len(coef))]
    for i in range(0, len(coef), 4):
        out.append("    " + ", ".join("%.17g" % v for v in coef[i:i + 4]) + ",")
    out.append("};")
    return "\n".join(out) + "\n"


# MB-nrg polynomials: each piece is its own translation unit (compiled in parallel) with its
# own kernel writing a partial value and gradient; the chain-rule stage sums the pieces.  The
# CH4-H2O-H2O 3B polynomial instead takes the blocked form of the MB-pol ones (generate(): one
# translation unit, NPARTS parts shared by the wavefronts of a block, one output slice).
NRG_PARTS = {"ch4h2o_2b": 1, "ch4ch4_2b": 8}
NRG_BLOCKED = ("ch4h2o2_3b",)


def generate_nrg(name, out_dir):
    nvars, coef, monos = load_table(name)
    s = Schedule(nvars, [(k, m) for k, m in monos])
    q = NRG_PARTS[name]
    assign = s.partition(q)
    c = s.counts()
    for k in range(q):
        hdr = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.tbl -- do not edit." % name,
               "// Schedule: %s" % ", ".join("%s=%d" % kv for kv in c.items()),
               "// Piece %d of %d (op-count balance: %s)." % (k, q, ", ".join(str(v) for v in s.part_load)),
               "#pragma once", "#include \"poly_support.hpp\"", ""]
        src = "\n".join(hdr) + s.emit("poly_%s_p%d" % (name, k), grad=True, part=k, assign=assign) + "\n" + \
            s.emit("poly_%s_e_p%d" % (name, k), grad=False, part=k, assign=assign)
        write_if_changed(os.path.join(out_dir, "poly_%s_p%d.inc" % (name, k)), src)
        tu = ["// GENERATED by mbx_amd/gen/polygen.py -- do not edit.",
              "// Piece %d of %d of the %s polynomial (%d variables): its own kernel, so the pieces "
              "compile in parallel." % (k, q, name, nvars),
              "#include \"gen/poly_%s_p%d.inc\"" % (name, k), "#include \"nrg_kernel.hpp\"", "",
              "namespace mbxg {", "",
              "hipError_t launch_poly_%s_p%d(const double* xs, const int* n_dev, int cap, int ld, const double* coef, "
              "int do_grads, double* gs, double* ps, hipStream_t st) {" % (name, k),
              "    return launch_poly_piece<%d, poly_%s_p%d, poly_%s_e_p%d>(xs, n_dev, cap, ld, coef, do_grads, gs, ps, "
              "st);" % (nvars, name, k, name, k),
              "}", "", "}  // namespace mbxg", ""]
        write_if_changed(os.path.join(out_dir, "poly_%s_p%d.hip" % (name, k)), "\n".join(tu))
    write_if_changed(os.path.join(out_dir, "coef_%s.inc" % name), coef_source(name, coef))
    return c


def generate_holder(name, out_dir):
    """gen/holder_<name>.inc: the holder description (mbx_amd/tables/<name>.holder.json) as
    constexpr tables for the MB-nrg variable and chain-rule kernels."""
    import json

    h = json.load(open(os.path.join(TABLES, name + ".holder.json")))
    forms = {"exp0": 0, "coul0": 1, "exp": 2}
    vs = h["vsite"] or {"oxygen_site": 0, "in_plane": 0.0, "out_of_plane": 0.0}
    nat = h["atoms"] + [0] * (3 - len(h["atoms"]))
    kinds = ["MON_%s" % m.upper() for m in h["monomers"]] + ["-1"] * (3 - len(h["monomers"]))
    out = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.holder.json -- do not edit." % name,
           "#pragma once", "#include \"mbxg_device.hpp\"", "#include \"nrg_kernel.hpp\"", "",
           "namespace mbxg {", "",
           "struct Holder_%s {" % name,
           "    static constexpr int kMon = %d;" % len(h["monomers"]),
           "    static constexpr int kNat[3] = {%s};" % ", ".join(str(v) for v in nat),
           "    static constexpr int kKind[3] = {%s};" % ", ".join(kinds),
           "    static constexpr int kNReal = %d;" % h["nreal"],
           "    static constexpr int kNSites = %d;" % h["nsites"],
           "    static constexpr bool kVsite = %s;" % ("true" if h["vsite"] else "false"),
           "    static constexpr int kOSite = %d;" % vs["oxygen_site"],
           "    static constexpr double kInPlane = %.17g;" % vs["in_plane"],
           "    static constexpr double kOutOfPlane = %.17g;" % vs["out_of_plane"],
           "    static constexpr double kRIn = %.17g;" % h["switch"]["inner"],
           "    static constexpr double kROut = %.17g;" % h["switch"]["outer"],
           "    static constexpr int kNVars = %d;" % len(h["vars"]),
           "    static constexpr int kPieces = %d;" % NRG_PARTS.get(name, 1),
           "    static constexpr NrgVar kVars[%d] = {" % len(h["vars"])]
    for f, k, d, s1, s2 in h["vars"]:
        out.append("        {%d, %.17g, %.17g, %d, %d}," % (forms[f], k, d, s1, s2))
    out += ["    };", "};", "", "}  // namespace mbxg", ""]
    write_if_changed(os.path.join(out_dir, "holder_%s.inc" % name), "\n".join(out))


def generate(name, fn, out_dir):
    nvars, coef, monos = load_table(name)
    terms, _ = full_terms(monos)
    s = Schedule(nvars, terms)
    c = s.counts()
    hdr = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.tbl -- do not edit." % name,
           "// Schedule: %s" % ", ".join("%s=%d" % kv for kv in c.items()),
           "// Coefficient slots index the fitted coefficient vector directly (%d slots)." % len(coef),
           "#pragma once",
           "#include \"poly_support.hpp\"", ""]
    assign = s.partition(NPARTS)
    hdr.insert(2, "// %d parts (op-count balance: %s); <fn>_parts(w, P, ...) evaluates parts w, w+P, ... "
                  "into g and returns their share of the value." % (NPARTS, ", ".join(str(v) for v in s.part_load)))
    body = []
    for grad, suffix in ((True, ""), (False, "_e")):
        for k in range(NPARTS):
            body.append(s.emit("%s%s_p%d" % (fn, suffix, k), grad=grad, part=k, assign=assign))
        args, call = (("mbxg_cptr C, const double* x, double* g", "C, x, g, E") if grad
                      else ("mbxg_cptr C, const double* x", "C, x, E"))
        # wavefront w of P evaluates pieces w, w + P, ...: a straight sequence of uniform
        # branches (a switch over the pieces inside a loop costs the register allocator dearly)
        d = ["__device__ __forceinline__ double %s%s_parts(int w, int P, %s) {" % (fn, suffix, args),
             "    double E = 0.0;"]
        for k in range(NPARTS):
            d.append("    if (%d %% P == w) E = %s%s_p%d(%s);" % (k, fn, suffix, k, call))
        d += ["    return E;", "}", ""]
        body.append("\n".join(d))
    src = "\n".join(hdr) + "\n".join(body)
    path = os.path.join(out_dir, "poly_%s.inc" % name)
    write_if_changed(path, src)
    # fitted coefficients, for the host runtime's default table
    write_if_changed(os.path.join(out_dir, "coef_%s.inc" % name), coef_source(name, coef))
    return path, c


# ------------------------------------------------------------------------------------------
# Differentiation: the polynomial in dual numbers, and the coefficient basis.
#
# The adjoint of an energy+force loss needs, per n-mer and along the tangent direction xd of
# the variables, (i) P, dP = grad P . xd, grad P and d(grad P) = H xd (forward-over-reverse:
# the reverse-mode trie schedule above, every value carried with its tangent), and (ii) the
# coefficient gradient sum_n [alpha_n M_k(x_n) + sw_n dM_k(x_n; xd_n)], where M_k is the
# basis function of coefficient k (the sum of its monomials) and dM_k its tangent.

_DECL = re.compile(r"^(const )?double (\w+) = (.*);$")
_FMA = re.compile(r"^(\S+) = __fma_rn\((.*)\);$")
_ADD = re.compile(r"^(\S+) \+= (.*);$")


def _tan(a):
    """Tangent of an operand of a generated statement (None for a constant)."""
    a = a.strip()
    if a == "{c}" or re.match(r"^[0-9.]+$", a):
        return None
    m = re.match(r"^(x|g)\[(\d+)\]$", a)
    if m:
        return "%sd[%s]" % m.groups()
    if a == "E":
        return "dE"
    assert re.match(r"^\w+$", a), a
    return "d" + a


def _tan_product(u, v):
    """Tangent of u * v."""
    du, dv = _tan(u), _tan(v)
    if du and dv:
        return "__fma_rn(%s, %s, %s * %s)" % (du, v, u, dv)
    if du:
        return "%s * %s" % (du, v)
    if dv:
        return "%s * %s" % (u, dv)
    return None


def dual_stmt(t):
    """The statement followed by the statement(s) that carry its tangent."""
    m = _DECL.match(t)
    if m:
        const, name, expr = m.groups()
        if " * " in expr:
            u, v = expr.split(" * ")
            d = _tan_product(u, v)
        else:
            d = _tan(expr)
        return [t, "%sdouble d%s = %s;" % (const or "", name, d or "0.0")]
    m = _FMA.match(t)
    if m:
        target, args = m.groups()
        u, v, acc = [s.strip() for s in args.split(",")]
        assert acc == target, t
        d = _tan(target)
        for p, q in ((u, _tan(v)), (_tan(u), v)):
            if p is not None and q is not None:
                d = "__fma_rn(%s, %s, %s)" % (p, q, d)
        return [t, "%s = %s;" % (_tan(target), d)] if d != _tan(target) else [t]
    m = _ADD.match(t)
    if m:
        target, a = m.groups()
        da = _tan(a)
        return [t, "%s += %s;" % (_tan(target), da)] if da else [t]
    raise ValueError("no tangent rule for: " + t)


def emit_dual(s, fn, part, assign, region_coefs=12, region_ops=96, tangent_only=False):
    """Dual-number version of Schedule.emit for one part: value, tangent, gradient and gradient
    tangent accumulate into E, dE, g, gd (passed in, not cleared).

    With tangent_only, only the gradient tangent gd (the Hessian-vector product H xd) is kept:
    the statements accumulating E, dE and g are dropped (the plain E+G kernel provides E and g,
    and dE = g . xd), which frees a third of the accumulator registers; the trie values and their
    tangents that gd needs stay."""
    drop = lambda t: tangent_only and _lhs(t) in ("E", "dE") or (tangent_only and _lhs(t).startswith("g["))
    ops = []
    for t, q in s.ops(True, part, assign):
        if t == "double E = E0;":
            continue
        ops += [(d, q if "{c}" in d else None) for d in dual_stmt(t) if not drop(d)]
    regions, coefs, nar = [[]], [[]], 0
    for t, q in ops:
        cur = coefs[-1]
        if q is not None and q not in cur and (len(cur) >= region_coefs or nar >= region_ops):
            regions.append([])
            coefs.append([])
            nar = 0
            cur = coefs[-1]
        if q is not None and q not in cur:
            cur.append(q)
        regions[-1].append((t, q))
        nar += 1
    if tangent_only:
        out = ["__device__ __forceinline__ void %s(mbxg_cptr C, const double* x, const double* xd, double* gd) {" % fn]
    else:
        out = ["__device__ __forceinline__ void %s(mbxg_cptr C, const double* x, const double* xd, double* g, "
               "double* gd, double& E, double& dE) {" % fn]
    w = out.append
    name = lambda r, q: "k%d_%d" % (r, coefs[r].index(q))
    for q in coefs[0]:
        w("    const double %s = C[%d];" % (name(0, q), q))
    last = "x[0]" if tangent_only else "E"
    for r, reg in enumerate(regions):
        w("    MBXG_NODE_FENCE(C, %s);" % last)
        if r + 1 < len(regions):
            for q in coefs[r + 1]:
                w("    const double %s = C[%d];" % (name(r + 1, q), q))
        for t, q in reg:
            w("    " + (t.replace("{c}", name(r, q)) if q is not None else t))
            if q is not None:
                last = _lhs(t)
    w("}")
    return "\n".join(out) + "\n"


def basis_slices(monos, ncoef, nslices):
    """Contiguous coefficient ranges of near-equal cost (one per slice)."""
    cost = [0] * ncoef
    for k, m in monos:
        cost[k] += 2 * len(m) + 1
    total = sum(cost)
    bounds, acc, k0 = [], 0, 0
    for k in range(ncoef):
        acc += cost[k]
        if acc >= total * (len(bounds) + 1) / nslices and len(bounds) < nslices - 1:
            bounds.append((k0, k + 1))
            k0 = k + 1
    bounds.append((k0, ncoef))
    return bounds


def emit_basis(fn, monos, c0, c1):
    """Slice function: for coefficients c0..c1-1, each lane's alpha * M_k + sw * dM_k goes to
    the block's LDS tile red[64][65] (row = coefficient within a chunk of 64, column = lane);
    after every chunk the column sums (over the wave's 64 n-mers) go to out[k]."""
    by = collections.defaultdict(list)
    for k, m in monos:
        if c0 <= k < c1:
            by[k].append(tuple(sorted(m)))
    out = ["__device__ __forceinline__ void %s(double* x, double* xd, double alpha, double sw, "
           "double* red, double* out, int lane) {" % fn]
    w = out.append
    tmp = [0]
    nv = 1 + max(max(m) for k, m in monos if m)

    def fresh(p):
        tmp[0] += 1
        return "%s%d" % (p, tmp[0])

    for k in range(c0, c1):
        j = (k - c0) % 64
        w("    {")
        w("        double TA = 0.0, TS = 0.0;")
        prefix = {}  # sorted prefix tuple -> (value name, tangent name), shared inside a coefficient

        def prod(head):
            if len(head) == 1:
                return "x[%d]" % head[0], "xd[%d]" % head[0]
            if head not in prefix:
                pv, pt = prod(head[:-1])
                i, nv, nt = head[-1], fresh("p"), fresh("q")
                w("        const double %s = %s * x[%d];" % (nv, pv, i))
                w("        const double %s = __fma_rn(%s, x[%d], %s * xd[%d]);" % (nt, pt, i, pv, i))
                prefix[head] = (nv, nt)
            return prefix[head]

        for m in sorted(by.get(k, [])):
            if len(m) == 0:
                w("        TA += 1.0;")
                continue
            if len(m) == 1:
                w("        TA += x[%d];" % m[0])
                w("        TS += xd[%d];" % m[0])
                continue
            pv, pt = prod(m[:-1])  # value and tangent of all but the last variable
            l = m[-1]
            w("        TA = __fma_rn(%s, x[%d], TA);" % (pv, l))
            w("        TS = __fma_rn(%s, xd[%d], __fma_rn(%s, x[%d], TS));" % (pv, l, pt, l))
        w("        red[%d * 65 + lane] = __fma_rn(alpha, TA, sw * TS);" % j)
        w("    }")
        # the variables are redefined opaquely between coefficients: products repeated in
        # other coefficients are recomputed there, not kept live from here
        w("    ad_opaque<%d>(x, xd);" % nv)
        if j == 63 or k == c1 - 1:
            base = k - j
            w("    ad_basis_flush(red, out + %d, %d, lane);" % (base, j + 1))
    w("}")
    return "\n".join(out) + "\n"


# Differentiated polynomials: gradient-tangent parts (the schedule partitioned into AD_PARTS
# pieces, one wavefront each; kAdParts in ad_kernel.hpp) and coefficient-basis slices.
AD_PARTS = 16
AD_POLYS = {"mbpol_2b": 16, "mbpol_3b": 16}  # name -> basis slices


def generate_ad(name, out_dir):
    """gen/ad_<name>_tan.{inc,hip} and gen/ad_<name>_basis.{inc,hip}: two translation units
    (compiled in parallel), each with its kernel's launcher."""
    nvars, coef, monos = load_table(name)
    s = Schedule(nvars, [(k, m) for k, m in monos])
    assign = s.partition(AD_PARTS)
    fn = "ad_%s" % name
    sl = basis_slices(monos, len(coef), AD_POLYS[name])
    head = lambda what: ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.tbl -- do not edit." % name,
                         "// %s" % what, "#pragma once", "#include \"poly_support.hpp\"", "#include \"ad_kernel.hpp\"",
                         "", "namespace mbxg {", ""]
    # gradient tangent (Hessian-vector product), AD_PARTS parts
    body = [emit_dual(s, "%s_tan_p%d" % (fn, k), k, assign, tangent_only=True) for k in range(AD_PARTS)]
    d = ["__device__ __forceinline__ void %s_tan(int part, mbxg_cptr C, const double* x, const double* xd, "
         "double* gd) {" % fn]
    d += ["    if (part == %d) %s_tan_p%d(C, x, xd, gd);" % (k, fn, k) for k in range(AD_PARTS)]
    d += ["}", "", "}  // namespace mbxg", ""]
    write_if_changed(os.path.join(out_dir, "ad_%s_tan.inc" % name),
                     "\n".join(head("Gradient tangent of the polynomial (H xd) in %d parts." % AD_PARTS))
                     + "\n".join(body) + "\n".join(d))
    # coefficient basis, in slices of near-equal cost
    body = [emit_basis("%s_basis_s%d" % (fn, i), monos, c0, c1) for i, (c0, c1) in enumerate(sl)]
    d = ["__device__ __forceinline__ void %s_basis(int slice, double* x, double* xd, double alpha, "
         "double sw, double* red, double* out, int lane) {" % fn]
    d += ["    if (slice == %d) %s_basis_s%d(x, xd, alpha, sw, red, out, lane);" % (i, fn, i) for i in range(len(sl))]
    d += ["}", "", "}  // namespace mbxg", ""]
    write_if_changed(os.path.join(out_dir, "ad_%s_basis.inc" % name),
                     "\n".join(head("Coefficient basis and its tangent in %d slices (coefficient ranges %s)."
                                    % (len(sl), ", ".join("%d-%d" % (c0, c1 - 1) for c0, c1 in sl))))
                     + "\n".join(body) + "\n".join(d))
    tu = lambda what, inc, launcher: "\n".join([
        "// GENERATED by mbx_amd/gen/polygen.py -- do not edit.",
        "// %s of the %s polynomial (%d variables, %d coefficients)." % (what, name, nvars, len(coef)),
        "#include \"gen/%s\"" % inc, "", "namespace mbxg {", "", launcher, "", "}  // namespace mbxg", ""])
    write_if_changed(os.path.join(out_dir, "ad_%s_tan.hip" % name), tu(
        "Gradient tangent", "ad_%s_tan.inc" % name,
        "hipError_t launch_ad_tan_%s(const double* xs, const double* xds, int n, int ld, const double* coef, "
        "double* gds, hipStream_t st) {\n"
        "    static_assert(kAdParts == %d, \"part count of the generated code\");\n"
        "    if (n <= 0) return hipSuccess;\n"
        "    hipLaunchKernelGGL((k_ad_tan<%d, %d, %s_tan>), dim3((n + 63) / 64, %d), dim3(64), 0, st, xs, xds, n, ld,\n"
        "                       (uint64_t)(uintptr_t)coef, gds);\n"
        "    return hipGetLastError();\n}" % (name, AD_PARTS, nvars, AD_PARTS, fn, AD_PARTS)))
    write_if_changed(os.path.join(out_dir, "ad_%s_basis.hip" % name), tu(
        "Coefficient basis", "ad_%s_basis.inc" % name,
        "hipError_t launch_ad_basis_%s(const double* xs, const double* xds, const double* sw, const double* swd, "
        "const double* g_energy, int n, int ld, double* partial, hipStream_t st) {\n"
        "    if (n <= 0) return hipSuccess;\n"
        "    hipLaunchKernelGGL((k_ad_basis<%d, %d, %s_basis>), dim3((n + 63) / 64, %d), dim3(64), 0, st, xs, xds, sw, "
        "swd,\n                       g_energy, n, ld, partial);\n"
        "    return hipGetLastError();\n}" % (name, nvars, len(coef), fn, len(sl))))
    return dict(parts=AD_PARTS, slices=len(sl))


def generate_ps(out_dir):
    """gen/ps_1b.inc: the MB-pol 1B (Partridge-Schwenke) data of mbx_amd/tables/ps_1b.json as
    constants, the derived ones evaluated in the reference's operation order (c5z: ps.cpp:445-447;
    deoh, phh1: :450-451; cm-1 -> kcal/mol: constants.h:47-61), and the four coefficient sums of
    ps.cpp:505-526 as straight-line code in the reference's order (no contraction)."""
    import json
    import math

    d = json.load(open(os.path.join(TABLES, "ps_1b.json")))
    c5z = [d["f5z"] * d["c5zA"][j] + d["fbasis"] * d["cbasis"][j] + d["fcore"] * d["ccore"][j] + d["frest"] * d["crest"][j]
           for j in range(245)]
    consts = {
        "kReoh": d["reoh"], "kB1": d["b1"], "kRoh": d["roh"], "kAlphaoh": d["alphaoh"],
        "kDeoh": d["f5z"] * d["deohA"], "kPhh1": d["f5z"] * d["phh1A"] * math.exp(d["phh2"]), "kPhh2": d["phh2"],
        "kCosthe": d["costhe"], "kCorrection": d["correction"],
        "kCm1": (d["Eh_J"] * d["Na"] / d["kcal_J"]) / (1.0e-2 * d["Eh_J"] / (d["c0"] * d["h_Js"])),
        "kC5z0": c5z[0],
    }
    out = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/ps_1b.json -- do not edit.",
           "#pragma once", "", "namespace ps1b {", ""]
    for k, v in consts.items():
        out.append("constexpr double %s = %s;" % (k, repr(float(v))))
    out += ["", "// s0..s3 of ps.cpp:505-526 from the power tables f0, f1, f2 (fmat[0..2][0..15])",
            "__device__ __forceinline__ void sums(const double* f0, const double* f1, const double* f2, double& s0,",
            "                                     double& s1, double& s2, double& s3) {"]
    for j in range(1, 245):
        I, J, K = d["idx1"][j], d["idx2"][j], d["idx3"][j]
        c = repr(float(c5z[j]))
        out.append("    s0 += %s * (f0[%d] * f1[%d] + f0[%d] * f1[%d]) * f2[%d];" % (c, I, J, J, I, K))
        out.append("    s1 += %s * (%r * f0[%d] * f1[%d] + %r * f0[%d] * f1[%d]) * f2[%d];" % (
            c, float(I - 1), I - 1, J, float(J - 1), J - 1, I, K))
        out.append("    s2 += %s * (%r * f0[%d] * f1[%d] + %r * f0[%d] * f1[%d]) * f2[%d];" % (
            c, float(J - 1), I, J - 1, float(I - 1), J, I - 1, K))
        out.append("    s3 += %s * (f0[%d] * f1[%d] + f0[%d] * f1[%d]) * %r * f2[%d];" % (c, I, J, J, I, float(K - 1), K - 1))
    out += ["}", "", "}  // namespace ps1b", ""]
    path = os.path.join(out_dir, "ps_1b.inc")
    write_if_changed(path, "\n".join(out))
    return path


# ------------------------------------------------------------------------------------------
# 1B polynomials (any degree): reverse mode over the full monomial trie.  Node n (a sorted
# prefix of variable indices, reached from its parent over the edge variable v) has the product
# P(n) = P(parent) * x_v; E = sum_n c(n) P(n), with c(n) the coefficient of the monomial equal
# to n (zero if none).  A depth-first walk returns the adjoint R(n) = c(n) + sum_children
# x_v' R(child) and adds P(parent) R(n) to g_v on the way back, so the value and gradient come
# out of one pass with the path's products and adjoints as the only live temporaries.
ONE_BODY = ("ch4_1b", "co2_1b")


def emit_trie(fn, nvars, monos):
    coef_of = {m: k for k, m in monos}
    children = collections.defaultdict(list)
    for _, m in monos:
        for d in range(1, len(m) + 1):
            parent, v = m[:d - 1], m[d - 1]
            if v not in children[parent]:
                children[parent].append(v)
    for kids in children.values():
        kids.sort()
    out = ["__device__ __forceinline__ double %s(const double* __restrict__ C, const double* x, double* g) {" % fn,
           "    for (int i = 0; i < %d; ++i) g[i] = 0.0;" % nvars, "    double E = 0.0;"]
    w = out.append
    counter = [0]

    def walk(node, pname):
        """Statements for the subtree of `node`; returns the name holding R(node)."""
        i = counter[0]
        counter[0] += 1
        r = "R%d" % i
        k = coef_of.get(node)
        w("    double %s = %s;" % (r, "C[%d]" % k if k is not None else "0.0"))
        for v in children.get(node, []):
            child = node + (v,)
            if children.get(child):
                pc = "P%d" % counter[0]
                w("    const double %s = %s;" % (pc, ("%s * x[%d]" % (pname, v)) if pname else "x[%d]" % v))
            else:
                pc = None
            rc = walk(child, pc)
            w("    g[%d] %s;" % (v, ("= __fma_rn(%s, %s, g[%d])" % (pname, rc, v)) if pname else "+= %s" % rc))
            w("    %s = __fma_rn(x[%d], %s, %s);" % (r, v, rc, r))
        return r

    for v in children[()]:
        pc = "P%d" % counter[0] if children.get((v,)) else None
        if pc:
            w("    const double %s = x[%d];" % (pc, v))
        rc = walk((v,), pc)
        w("    g[%d] += %s;" % (v, rc))
        w("    E = __fma_rn(x[%d], %s, E);" % (v, rc))
    w("    return E;")
    w("}")
    return "\n".join(out) + "\n"


def generate_1b(name, out_dir):
    nvars, coef, monos = load_table(name)
    hdr = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.tbl -- do not edit." % name,
           "// %d monomials, %d coefficients: value and gradient by reverse mode over the monomial trie."
           % (len(monos), len(coef)), "#pragma once", ""]
    src = "\n".join(hdr) + emit_trie("poly_%s" % name, nvars, monos)
    write_if_changed(os.path.join(out_dir, "poly_%s.inc" % name), src)
    write_if_changed(os.path.join(out_dir, "coef_%s.inc" % name), coef_source(name, coef))
    generate_holder(name, out_dir)
    return len(monos)


# ------------------------------------------------------------------------------------------
# Further MB-nrg families (halide-water 2B / 3B, CO2 2B / 3B, water 4B, alkali-ion-water 2B / 3B):
# holders as data (gen/xholder_<holder>.inc), each parameter set's coefficients
# (gen/xcoef_<holder>.inc) and the polynomial structures in pieces (gen/xpoly_<structure>_p<k>.{inc,hip},
# one kernel each).
ALKALI = ("li", "na", "k", "rb", "cs")
XHOLDERS = ("f_h2o_2b", "cl_h2o_2b", "br_h2o_2b", "i_h2o_2b", "f_h2o2_3b", "cl_h2o2_3b", "br_h2o2_3b", "i_h2o2_3b",
            "co2co2_2b", "co2x3_3b", "h2ox4_4b") + tuple("%s_h2o_2b" % a for a in ALKALI) + \
    tuple("%s_h2o2_3b" % a for a in ALKALI)
XPOLY_PIECES = {"halide_h2o_2b": 1, "co2co2_2b": 2, "halide_h2o2_3b": 4, "co2x3_3b": 4, "h2ox4_4b": 8,
                "alkali_h2o_2b": 1, "alkali_h2o2_3b": 2}
XKIND = {"h2o": "MON_H2O", "ch4": "MON_CH4", "co2": "MON_CO2", "f-": "MON_F", "cl-": "MON_CL", "br-": "MON_BR",
         "i-": "MON_I", "li+": "MON_LI", "na+": "MON_NA", "k+": "MON_K", "rb+": "MON_RB", "cs+": "MON_CS"}


def emit_trie_piece(fn, nvars, monos, fence_every=48):
    """emit_trie for monomials of any degree as a piece function (mbxg_cptr coefficients, value
    returned on top of E0), with a region fence every `fence_every` statements so coefficient
    loads stay near their FMAs."""
    body = emit_trie(fn + "_trie", nvars, monos).replace("const double* __restrict__ C", "mbxg_cptr C")
    lines = body.split("\n")
    out, n, last = [], 0, "x[0]"
    for ln in lines:
        out.append(ln)
        m = re.match(r"\s+(?:const )?double (\w+) = ", ln) or re.match(r"\s+(\w+) = __fma_rn", ln)
        if m:
            last = m.group(1)
            n += 1
            if n % fence_every == 0:
                out.append("    MBXG_NODE_FENCE(C, %s);" % last)
    src = "\n".join(out)
    src += ("__device__ __forceinline__ double %s(mbxg_cptr C, const double* x, double* g, double E0) {\n"
            "    return E0 + %s_trie(C, x, g);\n}\n" % (fn, fn))
    return src


def generate_xpoly(name, out_dir):
    nvars, coef, monos = load_table(name)
    q = XPOLY_PIECES[name]
    deg = max(len(m) for _, m in monos)
    hdr = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.tbl -- do not edit." % name,
           "#pragma once", "#include \"poly_support.hpp\"", ""]
    if deg > 4:
        assert q == 1 or name == "co2co2_2b"
        # degree 5-6: reverse mode over the full monomial trie, split by the first variable
        firsts = sorted({m[0] for _, m in monos})
        load = [0] * q
        groups = [[] for _ in range(q)]
        for v in sorted(firsts, key=lambda v: -sum(1 for _, m in monos if m[0] == v)):
            k = min(range(q), key=lambda j: load[j])
            groups[k].append(v)
            load[k] += sum(1 for _, m in monos if m[0] == v)
        parts = [[(k, m) for k, m in monos if m[0] in set(g)] for g in groups]
        for k in range(q):
            src = "\n".join(hdr) + "// Piece %d of %d: the monomials whose first variable is one of %s (%d of %d).\n" % (
                k, q, groups[k], len(parts[k]), len(monos)) + emit_trie_piece("xpoly_%s_p%d" % (name, k), nvars,
                                                                            parts[k])
            write_if_changed(os.path.join(out_dir, "xpoly_%s_p%d.inc" % (name, k)), src)
    else:
        s = Schedule(nvars, [(k, m) for k, m in monos])
        assign = s.partition(q)
        c = s.counts()
        for k in range(q):
            src = "\n".join(hdr) + "// Schedule: %s\n// Piece %d of %d (op-count balance: %s).\n" % (
                ", ".join("%s=%d" % kv for kv in c.items()), k, q, ", ".join(str(v) for v in s.part_load)) + \
                s.emit("xpoly_%s_p%d" % (name, k), grad=True, part=k, assign=assign)
            write_if_changed(os.path.join(out_dir, "xpoly_%s_p%d.inc" % (name, k)), src)
    for k in range(q):
        tu = ["// GENERATED by mbx_amd/gen/polygen.py -- do not edit.",
              "// Piece %d of %d of the %s polynomial (%d variables, degree %d): its own kernel." % (
                  k, q, name, nvars, deg),
              "#include \"gen/xpoly_%s_p%d.inc\"" % (name, k), "#include \"nrg_kernel.hpp\"", "",
              "namespace mbxg {", "",
              "hipError_t launch_xpoly_%s_p%d(const double* xs, const int* n_dev, int cap, int ld, const double* coef, "
              "double* gs, double* ps, hipStream_t st) {" % (name, k),
              "    return launch_xpoly_piece<%d, xpoly_%s_p%d>(xs, n_dev, cap, ld, coef, gs, ps, st);" % (nvars, name, k),
              "}", "", "}  // namespace mbxg", ""]
        write_if_changed(os.path.join(out_dir, "xpoly_%s_p%d.hip" % (name, k)), "\n".join(tu))
    return dict(pieces=q, degree=deg, monomials=len(monos))


def generate_xholder(name, out_dir):
    """gen/xholder_<name>.inc and gen/xcoef_<name>.inc from mbx_amd/tables/<name>.holder.json."""
    import json

    h = json.load(open(os.path.join(TABLES, name + ".holder.json")))
    forms = {"exp0": 0, "coul0": 1, "exp": 2, "coul": 3}
    nm = len(h["monomers"])
    vs = h["vsites"]
    sw = h["switch"]
    pairs = sw["pairs"]
    pad = lambda v, n, z: list(v) + [z] * (n - len(v))
    out = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.holder.json -- do not edit." % name,
           "#pragma once", "#include \"mbxg_device.hpp\"", "#include \"nrg_kernel.hpp\"", "",
           "namespace mbxg {", "",
           "struct XHolder_%s {" % name,
           "    static constexpr int kMon = %d;" % nm,
           "    static constexpr int kNat[4] = {%s};" % ", ".join(str(v) for v in pad(h["atoms"], 4, 0)),
           "    static constexpr int kKind[4] = {%s};" % ", ".join(pad([XKIND[m] for m in h["monomers"]], 4, "-1")),
           "    static constexpr int kNReal = %d;" % h["nreal"],
           "    static constexpr int kNSites = %d;" % h["nsites"],
           "    static constexpr int kNVsite = %d;" % len(vs),
           "    static constexpr int kVO[2] = {%s};" % ", ".join(str(v) for v in pad([w["oxygen_site"] for w in vs], 2, 0)),
           "    static constexpr double kInPlane = %.17g;" % (vs[0]["in_plane"] if vs else 0.0),
           "    static constexpr double kOutOfPlane = %.17g;" % (vs[0]["out_of_plane"] if vs else 0.0),
           "    static constexpr double kRIn = %.17g;" % sw["inner"],
           "    static constexpr double kROut = %.17g;" % sw["outer"],
           "    static constexpr int kNPair = %d;" % len(pairs),
           "    // skipped when at least kSkipMin first-atom separations exceed kROut",
           "    static constexpr int kSkipMin = %d;" % h.get("skip_min", len(pairs)),
           "    static constexpr int kPA[6] = {%s};" % ", ".join(str(v) for v in pad([a for a, b in pairs], 6, 0)),
           "    static constexpr int kPB[6] = {%s};" % ", ".join(str(v) for v in pad([b for a, b in pairs], 6, 0)),
           "    static constexpr int kNTerm = %d;" % len(sw["terms"]),
           "    static constexpr double kTCoef[%d] = {%s};" % (len(sw["terms"]), ", ".join("%.17g" % t[0] for t in sw["terms"])),
           "    static constexpr int kTMask[%d] = {%s};" % (len(sw["terms"]), ", ".join(
               str(sum(1 << q for q in t[1])) for t in sw["terms"])),
           "    static constexpr int kNVars = %d;" % len(h["vars"]),
           "    static constexpr int kPieces = %d;" % XPOLY_PIECES[h["poly"]],
           "    static constexpr NrgVar kVars[%d] = {" % len(h["vars"])]
    for f, k, d, s1, s2 in h["vars"]:
        out.append("        {%d, %.17g, %.17g, %d, %d}," % (forms[f], k, d, s1, s2))
    out += ["    };", "};", "", "}  // namespace mbxg", ""]
    write_if_changed(os.path.join(out_dir, "xholder_%s.inc" % name), "\n".join(out))
    cs = ["// GENERATED by mbx_amd/gen/polygen.py from mbx_amd/tables/%s.holder.json -- do not edit." % name,
          "#pragma once", "static const int kNCoefX_%s = %d;" % (name, len(h["coef"])),
          "static const double kFitX_%s[%d] = {" % (name, len(h["coef"]))]
    for i in range(0, len(h["coef"]), 4):
        cs.append("    " + ", ".join("%.17g" % v for v in h["coef"][i:i + 4]) + ",")
    cs.append("};")
    write_if_changed(os.path.join(out_dir, "xcoef_%s.inc" % name), "\n".join(cs) + "\n")
    return h["poly"]


def main():
    out_dir = os.path.join(ROOT, "mbx_amd", "csrc", "gen")
    os.makedirs(out_dir, exist_ok=True)
    for name, fn in (("mbpol_2b", "poly_mbpol_2b"), ("mbpol_3b", "poly_mbpol_3b")):
        path, c = generate(name, fn, out_dir)
        print(path, c)
    for name in AD_POLYS:
        generate_ad(name, out_dir)
    for name in NRG_PARTS:
        print(name, generate_nrg(name, out_dir))
        generate_holder(name, out_dir)
    for name in NRG_BLOCKED:
        print(generate(name, "poly_" + name, out_dir))
        generate_holder(name, out_dir)
    print(generate_ps(out_dir))
    for name in ONE_BODY:
        print(name, generate_1b(name, out_dir))
    for name in XHOLDERS:
        generate_xholder(name, out_dir)
    for name in XPOLY_PIECES:
        print(name, generate_xpoly(name, out_dir))


if __name__ == "__main__":
    main()
