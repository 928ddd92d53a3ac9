"""Count the reference's algorithmic floating-point operations per evaluated n-mer.

Reads the reference's generated energy+gradient polynomial routines as TEXT (one pass of the
8-lane blocks f0..fN, i.e. one n-mer) and counts multiplications and additions/subtractions in
their arithmetic statements; array-index arithmetic (inside [...]) is excluded.  An FMA in our
kernels is two of these operations.  The totals define the "algorithmic FLOPs" that the
roofline figure divides by kernel time: they describe the work the reference does, independent
of how this project schedules it.

    poly_2b_v6x  E+G   potential/2b/poly-2b-v6x.cpp:1-12934     (f0..f99)
    poly_3b_v2x  E+G   potential/3b/poly-3b-v2x.cpp:1-26980     (f0..f130)
    poly_2b_v6x  E     potential/2b/poly-2b-v6x-nogrd.cpp:41
    poly_3b_v2x  E     potential/3b/poly-3b-v2x-nogrd.cpp:39
    x3b_v2x chain rule potential/3b/x3b-v2x.cpp:62-76, 1620-1726 (counted by hand, see CHAIN_3B)

Writes mbx_amd/tables/flops.json.  Usage: python mbx_amd/tables/flopcount.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


# The 3B holder's work after the polynomial, per evaluated trimer (x3b-v2x.cpp:1620-1726), in the
# same units (multiplications, additions/subtractions, and divisions; sqrt, exp and negations not
# counted): the 36 gradients scaled by the switch product (36 mul); 36 calls of g_var (:62-76):
# the separation (3 sub), its square (3 mul + 2 add), gg = -k g exp(-k (d - r0)) / d (3 mul,
# 1 sub, 1 div) and the two atom updates (6 mul + 6 add/sub), 25 each; the three switch
# derivatives (3 x (1 add + 2 mul + 1 div)); e += s E (1 mul + 1 add); the switch-derivative
# gradient of the three oxygens (9 x (2 mul + 2 add)); the virial (6 x (9 mul + 9 add)).
CHAIN_3B = dict(mul=36 + 36 * 12 + 6 + 1 + 18 + 54, add=36 * 12 + 3 + 1 + 18 + 54, div=36 + 3,
                source="src/potential/3b/x3b-v2x.cpp:62-76, 1620-1726")
CHAIN_3B["flops"] = CHAIN_3B["mul"] + CHAIN_3B["add"] + CHAIN_3B["div"]


def count_ops(text):
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"\[[^\]]*\]", "[]", text)  # drop index arithmetic
    mul = add = 0
    for stmt in text.split(";"):
        if "=" not in stmt or "for" in stmt or "#pragma" in stmt:
            continue
        rhs = stmt.split("=", 1)[1]
        if "==" in stmt or "<" in rhs:
            continue
        mul += rhs.count("*")
        # binary + / -: an operator preceded by an operand character
        add += len(re.findall(r"[\w\)\]]\s*[+-]\s*[\w\(]", rhs))
    return mul, add


def block_bodies(text):
    """Bodies of the f0..fN lane-block functions."""
    out = []
    for m in re.finditer(r"void f\d+\([^)]*\)\s*\{", text):
        i = m.end()
        depth = 1
        j = i
        while depth:
            depth += {"{": 1, "}": -1}.get(text[j], 0)
            j += 1
        out.append(text[i:j - 1])
    return out


def nogrd_body(text):
    m = re.search(r"eval\s*\([^)]*\)\s*\{", text)
    i = m.end()
    depth = 1
    j = i
    while depth:
        depth += {"{": 1, "}": -1}.get(text[j], 0)
        j += 1
    return text[i:j - 1]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    res = {}
    for key, path in (("mbpol_2b", "src/potential/2b/poly-2b-v6x.cpp"), ("mbpol_3b", "src/potential/3b/poly-3b-v2x.cpp")):
        text = open(os.path.join(ref, path)).read()
        mul = add = 0
        for b in block_bodies(text):
            m, a = count_ops(b)
            mul += m
            add += a
        res[key + "_eg"] = dict(mul=mul, add=add, flops=mul + add, source=path)
    for key, path in (("mbpol_2b", "src/potential/2b/poly-2b-v6x-nogrd.cpp"),
                      ("mbpol_3b", "src/potential/3b/poly-3b-v2x-nogrd.cpp")):
        text = open(os.path.join(ref, path)).read()
        m, a = count_ops(nogrd_body(text))
        res[key + "_e"] = dict(mul=m, add=a, flops=m + a, source=path)
    res["mbpol_3b_chain"] = CHAIN_3B
    with open(os.path.join(HERE, "flops.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
