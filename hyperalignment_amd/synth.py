"""Deterministic synthetic inputs for searchlight Procrustes.

Everything here is seeded and reproducible bit-for-bit on any machine that
runs the same numpy: an icosphere surface, great-circle searchlights in the
flat-array layout that ``ha_searchlight_procrustes_dense`` consumes
(``hyperalignment_c.py:161-169``: int32 indices concatenated, int32 offsets =
cumsum of sizes, float64 distances), and data matrices X, Y.

Random numbers come from splitmix64 with Box-Muller normals, the generator
family the reference's C tests use (``csrc/ha_test.c:33-59``), vectorised.

Workloads (n_samples x n_vertices, radius in mm on a 100 mm sphere):

* ``c0``: order-4 icosphere (2,562 vertices), 15 mm searchlights, n=200, float64.
* ``c1``: order-5 icosphere (10,242 vertices), 20 mm searchlights, n=300, float32.
* ``c2``: order-6 icosphere (40,962 vertices), 20 mm searchlights, n=500, float32.
* ``c4``: 16 template subjects on the order-5 mesh, n=400, float32.

X ~ N(0,1) z-scored per column (ddof=0); Y = X @ M + noise * E with M a
row-normalised random positive mixing over the vertices within 5 mm.
"""

from __future__ import annotations

import dataclasses
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
SPHERE_RADIUS_MM = 100.0


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Outputs start..start+n-1 of the splitmix64 stream seeded with ``seed``."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + _GAMMA * idx
    return _mix(state)


def stream_seed(seed: int, tag: int) -> int:
    """Independent sub-stream seed for (seed, tag)."""
    v = _mix(np.array([(seed * 0x100000001B3 + tag * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF],
                      dtype=np.uint64))
    return int(v[0])


def uniform(seed: int, n: int) -> np.ndarray:
    """Uniform [0,1) doubles: (next >> 11) * 2^-53."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def randn(seed: int, n: int) -> np.ndarray:
    """Standard normals, one per pair of uniforms (Box-Muller, cosine branch)."""
    u = uniform(seed, 2 * n)
    u1 = np.maximum(u[0::2], 1e-15)
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


# ---------------------------------------------------------------------------
# Surface
# ---------------------------------------------------------------------------

def icosphere(order: int, radius: float = SPHERE_RADIUS_MM) -> np.ndarray:
    """Vertices (10*4^order + 2, 3) of a subdivided icosahedron on a sphere.

    Vertex order: the 12 icosahedron vertices, then each level's edge
    midpoints in lexicographic order of their (lo, hi) endpoint ids.
    """
    return icosphere_mesh(order, radius)[0]


def icosphere_mesh(order: int, radius: float = SPHERE_RADIUS_MM):
    """(vertices (10*4^order + 2, 3), faces (20*4^order, 3) int32) of the subdivided
    icosahedron of icosphere()."""
    p = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0],
                  [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
                  [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(order):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        ukey, inv = np.unique(key, return_inverse=True)
        lo, hi = ukey // nv, ukey % nv
        mid = v[lo] + v[hi]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        nf = f.shape[0]
        ab = nv + inv[0:nf]
        bc = nv + inv[nf:2 * nf]
        ca = nv + inv[2 * nf:3 * nf]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)], axis=0)
        v = np.concatenate([v, mid], axis=0)
    return v * radius, f.astype(np.int32)


def _neighbourhoods(coords: np.ndarray, radius_mm: float, per_centre: np.ndarray | None = None):
    """Great-circle neighbourhoods: for each vertex, sorted ids with d < radius
    (d < per_centre[c] for centre c when given, every value <= radius_mm)."""
    from scipy.spatial import cKDTree

    R = float(np.linalg.norm(coords[0]))
    chord = 2.0 * R * np.sin(min(radius_mm / (2.0 * R), np.pi / 2)) * (1.0 + 1e-9) + 1e-9
    tree = cKDTree(coords)
    lists = tree.query_ball_point(coords, chord, return_sorted=True)
    sizes = np.fromiter((len(l) for l in lists), dtype=np.int64, count=len(lists))
    flat = np.fromiter((i for l in lists for i in l), dtype=np.int64, count=int(sizes.sum()))
    centre = np.repeat(np.arange(len(lists), dtype=np.int64), sizes)
    cosang = np.einsum("ij,ij->i", coords[flat], coords[centre]) / (R * R)
    d = R * np.arccos(np.clip(cosang, -1.0, 1.0))
    keep = d < (radius_mm if per_centre is None else per_centre[centre])
    flat, centre, d = flat[keep], centre[keep], d[keep]
    sizes = np.bincount(centre, minlength=len(lists))
    return flat, sizes, d


@dataclasses.dataclass
class Searchlights:
    """Flat searchlight arrays, the layout of ``hyperalignment_c.py:161-169``."""
    indices: np.ndarray   # int32, concatenated sorted vertex ids
    offsets: np.ndarray   # int32, count + 1, offsets[0] = 0
    dists: np.ndarray     # float64, great-circle distance to the centre (mm)
    radius: float
    nv: int

    @property
    def count(self) -> int:
        return int(self.offsets.shape[0] - 1)

    @property
    def sizes(self) -> np.ndarray:
        return np.diff(self.offsets)

    def lists(self):
        o = self.offsets
        return [self.indices[o[s]:o[s + 1]] for s in range(self.count)]

    def dist_lists(self):
        o = self.offsets
        return [self.dists[o[s]:o[s + 1]] for s in range(self.count)]


def tail_radii(nv: int, sphere_radius: float, seed: int, kmin: float = 50.0, kmax: float = 500.0):
    """Per-centre searchlight radii for sizes with a long tail: target size
    k = kmin (kmax / kmin)^(u^2), u uniform per centre (median about 90 for
    50-500), and the radius of a disc of k vertices' area, 4 pi R^2 / nv each."""
    u = uniform(seed, nv)
    k = kmin * (kmax / kmin) ** (u * u)
    area = 4.0 * np.pi * sphere_radius ** 2 / nv
    return np.sqrt(k * area / np.pi)


def searchlights(order: int, radius_mm: float, tail_seed: int | None = None) -> tuple[np.ndarray, Searchlights]:
    """Searchlights of every vertex: great-circle radius radius_mm, or with
    tail_seed the per-centre radii of tail_radii (capped at radius_mm, which
    stays the radius of the distance weights)."""
    coords = icosphere(order)
    per = None
    if tail_seed is not None:
        R = float(np.linalg.norm(coords[0]))
        per = np.minimum(tail_radii(coords.shape[0], R, tail_seed), radius_mm)
    flat, sizes, d = _neighbourhoods(coords, radius_mm, per)
    offsets = np.zeros(len(sizes) + 1, dtype=np.int64)
    np.cumsum(sizes, out=offsets[1:])
    if offsets[-1] >= 2 ** 31:
        raise ValueError("searchlight arrays exceed int32")
    sls = Searchlights(indices=flat.astype(np.int32), offsets=offsets.astype(np.int32),
                       dists=d.astype(np.float64), radius=float(radius_mm), nv=len(sizes))
    return coords, sls


def mixing_matrix(coords: np.ndarray, seed: int, radius_mm: float = 5.0):
    """Row-normalised random positive mixing over neighbours within radius_mm."""
    import scipy.sparse as sp

    flat, sizes, _ = _neighbourhoods(coords, radius_mm)
    rows = np.repeat(np.arange(len(sizes), dtype=np.int64), sizes)
    w = uniform(seed, flat.shape[0]) + 1e-3
    rs = np.bincount(rows, weights=w, minlength=len(sizes))
    w = w / rs[rows]
    nv = len(sizes)
    return sp.csr_matrix((w, (rows, flat)), shape=(nv, nv))


def zscore_columns(X: np.ndarray) -> np.ndarray:
    mu = X.mean(axis=0, keepdims=True)
    sd = X.std(axis=0, keepdims=True)
    sd[sd < 1e-15] = 1.0
    return (X - mu) / sd


def normal_matrix(seed: int, n: int, nv: int) -> np.ndarray:
    """n x nv normals, drawn vertex by vertex (column v = draws v*n .. v*n+n-1)."""
    return randn(seed, n * nv).reshape(nv, n).T.copy()


@dataclasses.dataclass
class Workload:
    name: str
    X: np.ndarray          # n x nv float64 (values exactly representable in the stated dtype)
    Y: np.ndarray
    sls: Searchlights
    dtype: str
    seed: int
    noise: float


_SL_CACHE: dict = {}
_MIX_CACHE: dict = {}


def _cached_searchlights(order, radius, tail_seed=None):
    key = (order, float(radius), tail_seed)
    if key not in _SL_CACHE:
        _SL_CACHE.clear()
        _SL_CACHE[key] = searchlights(order, radius, tail_seed)
    return _SL_CACHE[key]


def _cached_mixing(order, radius, seed):
    key = (order, seed)
    if key not in _MIX_CACHE:
        if len(_MIX_CACHE) > 4:
            _MIX_CACHE.clear()
        coords, _ = _cached_searchlights(order, radius)
        _MIX_CACHE[key] = mixing_matrix(coords, seed)
    return _MIX_CACHE[key]


CONFIGS = {
    "c0": dict(order=4, radius=15.0, n=200, dtype="float64", noise=0.3),
    "c1": dict(order=5, radius=20.0, n=300, dtype="float32", noise=0.3),
    "c2": dict(order=6, radius=20.0, n=500, dtype="float32", noise=0.3),
    # c2 with uneven searchlights: per-centre radii for sizes of about 50-500
    # with a long tail (median about 90), every kp bucket populated
    "c2tail": dict(order=6, radius=23.0, n=500, dtype="float32", noise=0.3, tail_seed=2024),
}


def make_workload(name: str, seed: int = 0, n: int | None = None, noise: float | None = None,
                  order: int | None = None, radius: float | None = None) -> Workload:
    """Build one of the named workloads (optionally overriding n / noise / mesh)."""
    cfg = dict(CONFIGS[name])
    if n is not None:
        cfg["n"] = n
    if noise is not None:
        cfg["noise"] = noise
    if order is not None:
        cfg["order"] = order
    if radius is not None:
        cfg["radius"] = radius
    coords, sls = _cached_searchlights(cfg["order"], cfg["radius"], cfg.get("tail_seed"))
    nv = sls.nv
    X = zscore_columns(normal_matrix(stream_seed(seed, 1), cfg["n"], nv))
    M = _cached_mixing(cfg["order"], cfg["radius"], stream_seed(seed, 2))
    E = normal_matrix(stream_seed(seed, 3), cfg["n"], nv)
    Y = np.asarray((M.T @ X.T).T) + cfg["noise"] * E
    if cfg["dtype"] == "float32":
        X = X.astype(np.float32).astype(np.float64)
        Y = Y.astype(np.float32).astype(np.float64)
    return Workload(name=name, X=X, Y=Y, sls=sls, dtype=cfg["dtype"], seed=seed,
                    noise=cfg["noise"])


def template_subjects(n_subj: int = 16, n: int = 400, order: int = 5, seed: int = 0,
                      noise: float = 0.5, radius: float = 20.0):
    """Template workload: subject i = S @ M_i + noise * E_i, float32 values."""
    coords, sls = searchlights(order, radius)
    nv = sls.nv
    S = normal_matrix(stream_seed(seed, 10), n, nv)
    dms = []
    for i in range(n_subj):
        M = mixing_matrix(coords, stream_seed(seed, 100 + i))
        E = normal_matrix(stream_seed(seed, 200 + i), n, nv)
        D = np.asarray((M.T @ S.T).T) + noise * E
        dms.append(D.astype(np.float32).astype(np.float64))
    return dms, sls


def toy_searchlights(lists, nv=None, dists=None, radius=0.0) -> Searchlights:
    """Flat arrays from Python lists (e.g. the reference's hand-written toy cases)."""
    sizes = np.array([len(l) for l in lists], dtype=np.int64)
    offsets = np.zeros(len(lists) + 1, dtype=np.int64)
    np.cumsum(sizes, out=offsets[1:])
    idx = np.concatenate([np.asarray(l, dtype=np.int64) for l in lists]).astype(np.int32)
    if dists is None:
        d = np.zeros(idx.shape[0], dtype=np.float64)
    else:
        d = np.concatenate([np.asarray(x, dtype=np.float64) for x in dists])
    if nv is None:
        nv = int(idx.max()) + 1
    return Searchlights(indices=idx, offsets=offsets.astype(np.int32), dists=d,
                        radius=float(radius), nv=int(nv))
