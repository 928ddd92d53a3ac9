"""Helpers for the NRG reader tests: write fixture systems as NRG files and run the single_point
driver (mbx_amd/single_point, built with the library)."""
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mbx_amd", "single_point")


def driver():
    assert os.path.exists(DRIVER), "mbx_amd/single_point is not built (run __graft_entry__.build())"
    return DRIVER


def write_nrg(path, systems, per_molecule=1, keyword_case=str.upper):
    """systems: list of (mon ids, atoms per monomer, atom names, xyz).  `per_molecule` monomers
    share each MOLECULE block.  Coordinates are written with 17 significant digits (exact)."""
    K = keyword_case
    lines = []
    for mons, nat, names, xyz in systems:
        xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
        first = np.concatenate([[0], np.cumsum(nat)[:-1]])
        lines.append(K("system"))
        for m0 in range(0, len(mons), per_molecule):
            lines.append(K("molecule"))
            for m in range(m0, min(m0 + per_molecule, len(mons))):
                lines.append("%s %s" % (K("monomer"), mons[m]))
                for a in range(first[m], first[m] + nat[m]):
                    lines.append("%s %.17g %.17g %.17g" % (names[a], *xyz[a]))
                lines.append(K("endmon"))
            lines.append(K("endmol"))
        lines.append(K("endsys"))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_json(path, box, c2, c3):
    mbx = {"twobody_cutoff": float(c2), "threebody_cutoff": float(c3)}
    if box is not None and len(box):
        mbx["box"] = [float(v) for v in np.asarray(box).reshape(-1)]
    with open(path, "w") as f:
        json.dump({"MBX": mbx}, f)


def run(*args, timeout=120):
    return subprocess.run([driver(), *args], capture_output=True, text=True, timeout=timeout)


def dump(path):
    """single_point --dump: [{"monomers": [(id, [names], xyz array)], "molecules": [[...]]}]."""
    r = run("--dump", path)
    assert r.returncode == 0, r.stderr
    it = iter(r.stdout.splitlines())
    n = int(next(it).split()[1])
    out = []
    for _ in range(n):
        w = next(it).split()
        nmon, nmol = int(w[3]), int(w[5])
        mols = [[int(v) for v in next(it).split()[1:]] for _ in range(nmol)]
        mons = []
        for _ in range(nmon):
            _, mid, nat = next(it).split()
            rows = [next(it).split() for _ in range(int(nat))]
            mons.append((mid, [r[0] for r in rows], np.array([[float(v) for v in r[1:]] for r in rows])))
        out.append({"monomers": mons, "molecules": mols})
    return out


def energies(stdout):
    """Energy / 1B / 2B / 3B values and gradient rows printed by single_point."""
    e, grads = {}, []
    for line in stdout.splitlines():
        w = line.split()
        if w[0] == "Energy=":
            e["total"] = float(w[1])
        elif len(w) >= 4 and w[1] == "Energy":
            e[w[0]] = float(w[-1])
        elif len(w) == 3:
            grads.append([float(v) for v in w])
    return e, np.array(grads)


def read_first_system(path):
    """Monomers of the first SYSTEM block of an NRG file, in file order: (id, atom names, xyz)
    each (a plain tokenisation of MONOMER / atom / ENDMON lines)."""
    out, cur = [], None
    for line in open(path):
        w = line.split()
        if not w:
            continue
        k = w[0].lower()
        if k == "endsys":
            break
        if k == "monomer":
            cur = (w[1].lower(), [], [])
        elif k == "endmon":
            out.append(cur)
            cur = None
        elif cur is not None:
            cur[1].append(w[0])
            cur[2].extend(float(v) for v in w[1:4])
    return out
