"""A caller's CSC pattern and searchlight lists are checked on the host before
anything reaches the GPU: a malformed indptr or row list, an empty searchlight
or a vertex index out of range makes the entry points return kHaErrorArg (-3)
instead of letting a kernel read or write out of bounds. No GPU is needed: the
checks run before any device call."""

import types

import numpy as np
import pytest

import hyperalignment_amd as ha


def _mat(indptr, indices, nrows, ncols):
    indptr = np.asarray(indptr, dtype=np.int32)
    indices = np.asarray(indices, dtype=np.int32)
    return types.SimpleNamespace(indptr=indptr, indices=indices, data=np.zeros(len(indices)),
                                 shape=(nrows, ncols))


# three vertices, one searchlight over all of them; the good pattern is full
GOOD = ([0, 3, 6, 9], [0, 1, 2] * 3)
BAD = {
    "indptr_not_from_zero": ([1, 3, 6, 9], [0, 1, 2] * 3),
    "indptr_decreasing": ([0, 3, 2, 9], [0, 1, 2] * 3),
    "indptr_past_nnz": ([0, 3, 6, 12], [0, 1, 2] * 3),
    "row_out_of_range": ([0, 3, 6, 9], [0, 1, 2, 0, 1, 3, 0, 1, 2]),
    "negative_row": ([0, 3, 6, 9], [0, 1, 2, -1, 1, 2, 0, 1, 2]),
    "rows_unsorted": ([0, 3, 6, 9], [0, 2, 1, 0, 1, 2, 0, 1, 2]),
    "rows_repeated": ([0, 3, 6, 9], [0, 1, 1, 0, 1, 2, 0, 1, 2]),
}


@pytest.mark.parametrize("case", sorted(BAD))
@pytest.mark.parametrize("entry", ["procrustes", "ridge"])
def test_malformed_pattern_is_rejected(case, entry):
    rng = np.random.default_rng(0)
    X = rng.standard_normal((8, 3))
    Y = rng.standard_normal((8, 3))
    m = _mat(*BAD[case], 3, 3)
    fn = ha.searchlight_procrustes_csc if entry == "procrustes" else ha.searchlight_ridge_csc
    with pytest.raises(ha.HaError) as ei:
        fn(X, Y, [[0, 1, 2]], m)
    assert ei.value.rc == -3, case


def test_good_pattern_passes_the_check():
    """The good pattern is not rejected as an argument error (without a GPU the
    call then fails at device set-up instead)."""
    rng = np.random.default_rng(0)
    X = rng.standard_normal((8, 3))
    Y = rng.standard_normal((8, 3))
    m = _mat(*GOOD, 3, 3)
    try:
        ha.searchlight_procrustes_csc(X, Y, [[0, 1, 2]], m)
    except ha.HaError as e:
        assert e.rc != -3


EMPTY = [np.array([0, 1], dtype=np.int32), np.array([], dtype=np.int32), np.array([2], dtype=np.int32)]


def test_empty_searchlight_is_rejected_dense():
    """An empty searchlight (size 0) is an argument error of the dense entry
    point (ha_searchlight_procrustes_dense), reported before any device call."""
    X = np.random.default_rng(1).standard_normal((8, 3))
    with pytest.raises(ha.HaError) as ei:
        ha.searchlight_procrustes(X, X.copy(), EMPTY, None, 0.0)
    assert ei.value.rc == -3


@pytest.mark.parametrize("entry", ["procrustes", "ridge"])
def test_empty_searchlight_is_rejected_csc(entry):
    X = np.random.default_rng(1).standard_normal((8, 3))
    fn = ha.searchlight_procrustes_csc if entry == "procrustes" else ha.searchlight_ridge_csc
    with pytest.raises(ha.HaError) as ei:
        fn(X, X.copy(), [[0, 1, 2], []], _mat(*GOOD, 3, 3))
    assert ei.value.rc == -3


def test_vertex_index_out_of_range_is_rejected():
    X = np.random.default_rng(1).standard_normal((8, 3))
    with pytest.raises(ha.HaError) as ei:
        ha.searchlight_procrustes(X, X.copy(), [np.array([0, 3], dtype=np.int32)], None, 0.0)
    assert ei.value.rc == -3


def test_unknown_backend_is_a_value_error():
    X = np.zeros((4, 2))
    with pytest.raises(ValueError):
        ha.searchlight_procrustes(X, X, [np.array([0, 1])], None, 0.0, backend="cuda9")


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("entry", ["dense", "csc", "ridge", "template", "ha_procrustes", "ha_ridge"])
def test_non_finite_data_is_rejected(entry, bad):
    """NaN or Inf in X, Y or a subject is an argument error of the searchlight
    entry points and of the single-problem ha_procrustes and ha_ridge, found by
    a host scan before anything reaches the GPU."""
    rng = np.random.default_rng(2)
    X = rng.standard_normal((8, 3))
    Y = rng.standard_normal((8, 3))
    Y[5, 1] = bad
    lists = [np.array([0, 1, 2], dtype=np.int32)]
    with pytest.raises(ha.HaError) as ei:
        if entry == "dense":
            ha.searchlight_procrustes(X, Y, lists, None, 0.0)
        elif entry == "csc":
            ha.searchlight_procrustes_csc(X, Y, [[0, 1, 2]], _mat(*GOOD, 3, 3))
        elif entry == "ridge":
            ha.searchlight_ridge_csc(X, Y, [[0, 1, 2]], _mat(*GOOD, 3, 3))
        elif entry == "template":
            ha.searchlight_template_flat([X, Y], lists[0], np.array([0, 3], dtype=np.int32), None, 0.0)
        elif entry == "ha_procrustes":
            ha.procrustes(X, Y)
        else:
            ha.ridge(X, Y, 1.0)
    assert ei.value.rc == -3
