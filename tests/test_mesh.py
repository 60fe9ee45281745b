"""Searchlights on a triangle mesh (host code: no GPU).

The reference takes real-mesh searchlights from neuroboros
(benchmark_neuroboros.py:153), which is not in this image. Two distances:
  * along the edges (ha_mesh_searchlights): checked against a full Dijkstra
    in the oracle, a closed form on a triangulated grid, symmetry and the
    triangle inequality, to 1e-12 (sums of edge lengths in path order);
  * geodesic across the faces (ha_mesh_searchlights_geodesic, fast marching):
    on a flat triangulated grid every distance is the Euclidean one to 1e-12,
    and on the icosphere every distance lies within 0.1% below the
    great-circle distance (the polyhedron is inscribed in the sphere), so the
    searchlights equal the great-circle ones except at vertices within 0.1% of
    the radius.
Both keep the weight invariant of the reference's own searchlight tests
(tests/test_searchlight.py: weights summed per vertex are 1).
"""

import base64
import zlib

import numpy as np
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import synth
from oracle import ha_oracle as O


def grid_mesh(n):
    """n x n unit grid, each square split by its (i, j)-(i+1, j+1) diagonal."""
    ij = np.array([(i, j) for i in range(n) for j in range(n)], dtype=np.float64)
    coords = np.concatenate([ij, np.zeros((n * n, 1))], axis=1)
    faces = []
    for i in range(n - 1):
        for j in range(n - 1):
            a, b, c, d = i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1
            faces += [(a, b, c), (a, c, d)]
    return coords, np.array(faces, dtype=np.int32)


def grid_dist(i, j):
    """Edge distance on that grid from (0, 0) to (i, j): diagonal steps only along
    (1, 1) or (-1, -1)."""
    i, j = np.asarray(i), np.asarray(j)
    same = np.sign(i) * np.sign(j) >= 0
    m = np.minimum(np.abs(i), np.abs(j))
    return np.where(same, np.sqrt(2.0) * m + np.abs(np.abs(i) - np.abs(j)), np.abs(i) + np.abs(j))


def test_matches_oracle_dijkstra_icosphere():
    v, f = synth.icosphere_mesh(2)
    sls, ds = ha.mesh_searchlights(v, f, 40.0, return_dists=True, geodesic=False)
    osl, ods = O.mesh_searchlights(v, f, 40.0)
    assert len(sls) == len(osl) == v.shape[0]
    for a, b, da, db in zip(sls, osl, ds, ods):
        assert np.array_equal(a, b)
        assert np.abs(da - db).max() < 1e-12
    assert all(np.all(np.diff(sl) > 0) for sl in sls)


def test_grid_closed_form():
    coords, faces = grid_mesh(9)
    sls, ds = ha.mesh_searchlights(coords, faces, 100.0, centers=[0], return_dists=True, geodesic=False)
    sl, d = sls[0], ds[0]
    assert len(sl) == 81
    assert np.abs(d - grid_dist(sl // 9, sl % 9)).max() < 1e-12
    # strictly below the radius: the centre, 4 axis neighbours (1) and 2 diagonal
    # ones (sqrt 2); every other point is at 2 or more
    sls, ds = ha.mesh_searchlights(coords, faces, 2.0, centers=[40], return_dists=True, geodesic=False)
    sl, d = sls[0], ds[0]
    assert len(sl) == 7
    assert np.abs(d - grid_dist(sl // 9 - 4, sl % 9 - 4)).max() < 1e-12


def test_symmetric_and_triangle_inequality():
    v, f = synth.icosphere_mesh(3)
    idx, off, d = ha.mesh_searchlights_flat(v, f, 25.0, geodesic=False)
    pairs = {}
    for c in range(len(off) - 1):
        for j in range(off[c], off[c + 1]):
            pairs[(c, int(idx[j]))] = d[j]
    for (c, u), dd in pairs.items():
        assert abs(pairs[(u, c)] - dd) < 1e-12
        assert dd >= np.linalg.norm(v[c] - v[u]) - 1e-12  # a path is never shorter than the chord
        if c == u:
            assert dd == 0.0


@pytest.mark.parametrize("geodesic", [True, False])
def test_sparse_centers_match_full(geodesic):
    """Centres taken from a subset (the reference's center_space case,
    tests/test_searchlight.py:33-43) give the same lists as from every vertex."""
    v, f = synth.icosphere_mesh(3)
    full = ha.mesh_searchlights(v, f, 20.0, geodesic=geodesic)
    centers = np.arange(0, v.shape[0], 7)
    sub = ha.mesh_searchlights(v, f, 20.0, centers=centers, geodesic=geodesic)
    for c, sl in zip(centers, sub):
        assert np.array_equal(sl, full[c])


@pytest.mark.parametrize("center", [0, 40, 44])
def test_geodesic_flat_grid_is_euclidean(center):
    """On a flat triangulated grid the geodesic distances are the Euclidean
    ones: every shortest path is a straight ray across the faces, which the
    unfolding update follows exactly (the edge distances are up to 8% longer)."""
    coords, faces = grid_mesh(9)
    sls, ds = ha.mesh_searchlights(coords, faces, 100.0, centers=[center], return_dists=True)
    sl, d = sls[0], ds[0]
    assert len(sl) == 81
    assert np.abs(d - np.linalg.norm(coords[sl] - coords[center], axis=1)).max() < 1e-12
    # a radius that cuts through the grid keeps exactly the vertices strictly inside it
    sls = ha.mesh_searchlights(coords, faces, 2.5, centers=[40])
    inside = np.nonzero(np.linalg.norm(coords - coords[40], axis=1) < 2.5)[0]
    assert np.array_equal(sls[0], inside)


def test_geodesic_icosphere_matches_great_circle():
    """On the 10,242-vertex icosphere (radius 100 mm) the geodesic distances lie
    within 0.1% below the great-circle ones, and the 20 mm searchlights equal
    the great-circle searchlights of the synthetic workloads except at
    vertices within 0.1% of the radius (edge distances: 5-24% longer)."""
    w = synth.make_workload("c1")
    s = w.sls
    v, f = synth.icosphere_mesh(5)
    R = np.linalg.norm(v, axis=1).mean()
    sls, ds = ha.mesh_searchlights(v, f, s.radius, return_dists=True)
    for c, (sl, d) in enumerate(zip(sls, ds)):
        # (arccos near 1 resolves angles to about 1e-8 rad: 1e-5 mm of slack)
        gc = R * np.arccos(np.clip(v[sl] @ v[c] / R ** 2, -1.0, 1.0))
        assert np.all(d <= gc + 1e-5)
        assert np.all(d >= gc * (1 - 1e-3) - 1e-5)
    glists = s.lists()
    for c in range(0, len(sls), 7):
        for u in set(sls[c].tolist()) ^ set(glists[c].tolist()):
            gc = R * np.arccos(np.clip(v[u] @ v[c] / R ** 2, -1.0, 1.0))
            assert abs(gc - s.radius) < 1e-3 * s.radius


@pytest.mark.parametrize("uniform", [True, False])
def test_weights_sum_to_one_on_mesh(uniform):
    """tests/test_searchlight.py:8-31 on mesh searchlights: the weights of each
    vertex over the searchlights holding it sum to 1."""
    v, f = synth.icosphere_mesh(3)
    idx, off, d = ha.mesh_searchlights_flat(v, f, 20.0)
    w = O.searchlight_weights_flat(idx, off, None if uniform else d, 20.0, v.shape[0])
    tot = np.zeros(v.shape[0])
    np.add.at(tot, idx, w)
    assert np.abs(tot - 1.0).max() < 1e-12


def test_invalid_mesh_is_an_error():
    v, f = synth.icosphere_mesh(1)
    bad = f.copy()
    bad[0, 0] = v.shape[0]
    with pytest.raises(ha.HaError):
        ha.mesh_searchlights(v, bad, 10.0)
    with pytest.raises(ha.HaError):
        ha.mesh_searchlights(v, f, -1.0)


def test_freesurfer_surface_roundtrip(tmp_path):
    v, f = synth.icosphere_mesh(2)
    p = tmp_path / "lh.test"
    with open(p, "wb") as fh:
        fh.write(b"\xff\xff\xfe" + b"created by test\n\n")
        fh.write(np.array([v.shape[0], f.shape[0]], dtype=">i4").tobytes())
        fh.write(v.astype(">f4").tobytes())
        fh.write(f.astype(">i4").tobytes())
    rv, rf = ha.read_surface(str(p))
    assert np.array_equal(rv, v.astype(np.float32).astype(np.float64)) and np.array_equal(rf, f)


@pytest.mark.parametrize("encoding", ["ASCII", "Base64Binary", "GZipBase64Binary"])
def test_gifti_surface_roundtrip(tmp_path, encoding):
    v, f = synth.icosphere_mesh(1)
    v32 = v.astype(np.float32)

    def data(a):
        if encoding == "ASCII":
            return " ".join(str(x) for x in a.ravel())
        raw = a.astype(a.dtype.newbyteorder("<")).tobytes()
        if encoding == "GZipBase64Binary":
            raw = zlib.compress(raw)  # the GIFTI standard's zlib stream
        return base64.b64encode(raw).decode()

    def da(intent, a, dtype):
        return (f'<DataArray Intent="{intent}" DataType="{dtype}" ArrayIndexingOrder="RowMajorOrder" '
                f'Dimensionality="2" Dim0="{a.shape[0]}" Dim1="{a.shape[1]}" Encoding="{encoding}" '
                f'Endian="LittleEndian" ExternalFileName="" ExternalFileOffset=""><Data>{data(a)}</Data></DataArray>')

    xml = ('<?xml version="1.0" encoding="UTF-8"?><GIFTI Version="1.0" NumberOfDataArrays="2">'
           + da("NIFTI_INTENT_POINTSET", v32, "NIFTI_TYPE_FLOAT32")
           + da("NIFTI_INTENT_TRIANGLE", f, "NIFTI_TYPE_INT32") + "</GIFTI>")
    p = tmp_path / "lh.test.surf.gii"
    p.write_text(xml)
    rv, rf = ha.read_surface(str(p))
    assert np.allclose(rv, v32, atol=1e-5 if encoding == "ASCII" else 0) and np.array_equal(rf, f)
