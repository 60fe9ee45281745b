"""Searchlight Procrustes on searchlights built from a triangle mesh
(ha_mesh_searchlights_geodesic: geodesic distances across the faces, the lists
neuroboros' nb.sls provides to benchmark_neuroboros.py:153), against the CPU
FP64 oracle. Tolerance: max-abs
1e-4 on T, the bar of the other polar tests."""

import numpy as np
import pytest

import hyperalignment_amd as ha
from hyperalignment_amd import synth
from oracle import ha_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def hip():
    ha.init()
    yield
    ha.set_options(method="auto")


@pytest.mark.parametrize("method", ["polar", "jacobi64"])
def test_procrustes_on_mesh_searchlights(method):
    w = synth.make_workload("c0")
    v, f = synth.icosphere_mesh(4)
    idx, off, d = ha.mesh_searchlights_flat(v, f, 15.0)
    print(f"[mesh] {len(off) - 1} searchlights, k {np.diff(off).min()}-{np.diff(off).max()}", flush=True)
    ha.set_options(method=method)
    T = ha.searchlight_procrustes_flat(w.X, w.Y, idx, off, d, 15.0)
    ref = O.searchlight_procrustes_dense(np.asarray(w.X, np.float64), np.asarray(w.Y, np.float64), idx, off, d, 15.0)
    assert np.abs(T - ref).max() < 1e-4
