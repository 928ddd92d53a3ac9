"""Seeded synthetic water systems for the benchmark and parity workloads.

Periodic boxes (the plain path): oxygen atoms on a lattice of nx*ny*nz sites filling a cubic
box of edge L (spacings L/nx, L/ny, L/nz), the first `nmon` sites in raster order (x fastest),
each displaced by U(-0.3, 0.3) A per axis.  Each water gets O-H bonds of 0.9572 + U(-0.02,
0.02) A, an H-O-H angle of 104.52 + U(-3, 3) degrees and a uniformly random orientation
(Haar-random rotation from a normalised Gaussian quaternion).

    water_box(256,  19.7, (7, 7, 6))     256 waters, 1 g/cm^3
    water_box(2048, 39.4, (14, 14, 12))  2048 waters, 1 g/cm^3 -- the same recipe with the
                                         lattice doubled along each axis

Coordinates are returned as a flat array, 9 doubles per water (O, H, H), in Angstrom.
"""
import numpy as np

R_OH, ANGLE = 0.9572, 104.52


def random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], axis=-1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], axis=-1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], axis=-1),
    ], axis=1)


def water_geometries(rng, n):
    """Body-frame O, H1, H2 of n perturbed waters: (n, 3, 3)."""
    r1 = R_OH + rng.uniform(-0.02, 0.02, n)
    r2 = R_OH + rng.uniform(-0.02, 0.02, n)
    th = np.radians(ANGLE + rng.uniform(-3.0, 3.0, n))
    g = np.zeros((n, 3, 3))
    g[:, 1, 0] = r1 * np.cos(th / 2)
    g[:, 1, 1] = r1 * np.sin(th / 2)
    g[:, 2, 0] = r2 * np.cos(th / 2)
    g[:, 2, 1] = -r2 * np.sin(th / 2)
    return g


def place_waters(rng, oxygens):
    n = len(oxygens)
    geo = water_geometries(rng, n)
    rot = random_rotations(rng, n)
    atoms = np.einsum("nij,naj->nai", rot, geo) + oxygens[:, None, :]
    return atoms.reshape(-1)


def water_box(nmon, edge, lattice, seed=2048, jitter=0.3):
    """(xyz, box9) of nmon waters on a jittered lattice in a cubic periodic box."""
    nx, ny, nz = lattice
    if nmon > nx * ny * nz:
        raise ValueError("lattice too small")
    rng = np.random.default_rng(seed)
    idx = np.arange(nmon)
    ix, iy, iz = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    o = np.stack([(ix + 0.5) * edge / nx, (iy + 0.5) * edge / ny, (iz + 0.5) * edge / nz], axis=1)
    o += rng.uniform(-jitter, jitter, size=o.shape)
    xyz = place_waters(rng, o)
    box = np.array([edge, 0, 0, 0, edge, 0, 0, 0, edge], dtype=np.float64)
    return xyz, box


# The two periodic workloads (2B cutoff 9 A, 3B cutoff 7 A).
C1 = dict(name="water_256_box19.7", nmon=256, edge=19.7, lattice=(7, 7, 6), seed=256, cutoff2b=9.0, cutoff3b=7.0)
C2 = dict(name="water_2048_box39.4", nmon=2048, edge=39.4, lattice=(14, 14, 12), seed=2048, cutoff2b=9.0,
          cutoff3b=7.0)


def build(cfg):
    xyz, box = water_box(cfg["nmon"], cfg["edge"], cfg["lattice"], seed=cfg["seed"])
    return xyz, box, cfg["cutoff2b"], cfg["cutoff3b"]


# Gas-phase clusters for batches: oxygen atoms by random sequential addition (O-O >= 2.6 A)
# inside a sphere sized for 1 g/cm^3 (0.03343 waters per A^3), random orientations.
WATER_DENSITY = 0.03343


def water_cluster(rng, n, dmin=2.6):
    """(9n,) coordinates of an n-water gas-phase cluster."""
    radius = (3.0 * n / (4.0 * np.pi * WATER_DENSITY)) ** (1.0 / 3.0)
    pts = np.zeros((n, 3))
    k = 0
    d2 = dmin * dmin
    while k < n:
        cand = rng.uniform(-radius, radius, size=(256, 3))
        cand = cand[np.einsum("ij,ij->i", cand, cand) <= radius * radius]
        # clear of every point placed before this batch (vectorised), then of those accepted in it
        if k > 0:
            dd = ((cand[:, None, :] - pts[None, :k, :]) ** 2).sum(-1)
            cand = cand[dd.min(axis=1) >= d2]
        k0 = k
        for c in cand:
            if k == n:
                break
            if k == k0 or np.min(((pts[k0:k] - c) ** 2).sum(-1)) >= d2:
                pts[k] = c
                k += 1
    return place_waters(rng, pts)


def cluster_batch(nconf, nmin=16, nmax=512, seed=512):
    """nconf clusters with log-uniform sizes in [nmin, nmax]: (list of coordinate arrays, sizes)."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax + 1), nconf))).astype(int)
    sizes = np.clip(sizes, nmin, nmax)
    return [water_cluster(rng, int(n)) for n in sizes], sizes


# The batched workload: 512 gas-phase clusters of 16..512 waters, cutoffs 9 / 7 A.
C3 = dict(name="water_clusters_512x16-512_gas", nconf=512, nmin=16, nmax=512, seed=512, cutoff2b=9.0,
          cutoff3b=7.0)


# Path-integral beads of a mixed CH4 / H2O box: molecule centres on an 8 x 8 x 9 lattice over
# the 21.6 x 21.6 x 30 A cell (U(+-0.2) A jitter), 57 of them methanes (C-H 1.09 A, ideal
# tetrahedron, random orientation), the rest waters; each bead displaces every atom of the
# centroid by N(0, 0.05 A).
def methane(rng, center):
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) / np.sqrt(3.0) * 1.09
    rot = random_rotations(rng, 1)[0]
    return np.concatenate([center, (tet @ rot.T + center).reshape(-1)])


def mixed_box(n_ch4=57, n_h2o=513, cell=(21.6, 21.6, 30.0), lattice=(8, 8, 9), seed=57, jitter=0.2):
    """(xyz, kinds, box9) of the centroid system; kinds 0 = h2o, 1 = ch4, methanes first."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = lattice
    n = n_ch4 + n_h2o
    sites = rng.permutation(nx * ny * nz)[:n]
    ix, iy, iz = sites % nx, (sites // nx) % ny, sites // (nx * ny)
    c = np.stack([(ix + 0.5) * cell[0] / nx, (iy + 0.5) * cell[1] / ny, (iz + 0.5) * cell[2] / nz], axis=1)
    c += rng.uniform(-jitter, jitter, size=c.shape)
    parts = [methane(rng, c[i]) for i in range(n_ch4)]
    parts.append(place_waters(rng, c[n_ch4:]))
    kinds = np.array([1] * n_ch4 + [0] * n_h2o)
    box = np.array([cell[0], 0, 0, 0, cell[1], 0, 0, 0, cell[2]], dtype=np.float64)
    return np.concatenate(parts), kinds, box


def beads(xyz, nbeads, sigma=0.05, seed=32):
    rng = np.random.default_rng(seed)
    return [xyz + rng.normal(0.0, sigma, xyz.shape) for _ in range(nbeads)]


# The batched path-integral workload: 32 beads of 57 CH4 + 513 H2O, cutoffs 9 / 7 A.
C5 = dict(name="pimd_32beads_57ch4_513h2o_box21.6x21.6x30", nbeads=32, n_ch4=57, n_h2o=513, cutoff2b=9.0,
          cutoff3b=7.0, seed=57)


# Independent water dimers and trimers for the differentiable path: dimer O-O separations
# U(2.4, 9) A; trimer O-O separations each U(2.5, 7) A, the three drawn together and redrawn
# until they form a triangle; random orientations (molecules and the whole n-mer).
C4 = dict(name="water_4096_dimers_4096_trimers", ndimers=4096, ntrimers=4096, seed=4096, dimer_range=(2.4, 9.0),
          trimer_range=(2.5, 7.0))


def water_dimers(n, seed, r_range=(2.4, 9.0)):
    """(n, 18): O H H O H H per dimer."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(r_range[0], r_range[1], n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = np.zeros((n, 2, 3))
    o[:, 1] = r[:, None] * u
    return place_waters(rng, o.reshape(-1, 3)).reshape(n, 18)


def water_trimers(n, seed, r_range=(2.5, 7.0)):
    """(n, 27): O H H per water, three waters per trimer."""
    rng = np.random.default_rng(seed)
    d = np.empty((0, 3))
    while len(d) < n:
        c = rng.uniform(r_range[0], r_range[1], size=(2 * n, 3))  # r_ab, r_ac, r_bc
        ok = (c[:, 0] < c[:, 1] + c[:, 2]) & (c[:, 1] < c[:, 0] + c[:, 2]) & (c[:, 2] < c[:, 0] + c[:, 1])
        d = np.concatenate([d, c[ok]])
    d = d[:n]
    ab, ac, bc = d.T
    cosa = (ab * ab + ac * ac - bc * bc) / (2 * ab * ac)
    o = np.zeros((n, 3, 3))
    o[:, 1, 0] = ab
    o[:, 2, 0] = ac * cosa
    o[:, 2, 1] = ac * np.sqrt(np.clip(1 - cosa * cosa, 0.0, None))
    o = np.einsum("nij,naj->nai", random_rotations(rng, n), o)
    return place_waters(rng, o.reshape(-1, 3)).reshape(n, 27)


def build_c4(cfg=C4):
    return (water_dimers(cfg["ndimers"], cfg["seed"], cfg["dimer_range"]),
            water_trimers(cfg["ntrimers"], cfg["seed"] + 1, cfg["trimer_range"]))
