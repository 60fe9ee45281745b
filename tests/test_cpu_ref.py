"""The C/OpenMP restatement of the reference's CPU path (oracle/cpu_ref.c, the
bench's CPU baseline) against the numpy oracle: the same LAPACK drivers in
the same order (dgesdd, dgesvd on failure) give the same local transforms to
rounding, with and without reflections, on one thread and on several."""

import numpy as np
import pytest

from hyperalignment_amd import synth
from oracle import cpu_ref, ha_oracle as O


@pytest.mark.parametrize("reflection", [True, False])
@pytest.mark.parametrize("threads", [1, 4])
def test_cpu_ref_matches_oracle(reflection, threads):
    w = synth.make_workload("c0")
    s = w.sls
    which = np.arange(0, s.count, 23)
    got = cpu_ref.procrustes_searchlights(w.X, w.Y, s.indices, s.offsets, which, threads, reflection=reflection)
    lists = s.lists()
    for T, j in zip(got, which):
        ref = O.procrustes(w.X[:, lists[j]], w.Y[:, lists[j]], reflection=reflection)
        assert T.shape == ref.shape
        assert np.abs(T - ref).max() < 1e-12
        if not reflection:
            assert np.linalg.det(T) > 0
