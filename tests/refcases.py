"""Inputs of the reference's C known-answer tests (csrc/ha_test.c), rebuilt
with the same generator: splitmix64 seeded by test_rng_seed(seed), Box-Muller
normals filled row-major (csrc/ha_test.c:33-65)."""

import numpy as np

from hyperalignment_amd import synth


class RefRng:
    """Sequential splitmix64 / Box-Muller stream of csrc/ha_test.c:33-59."""

    def __init__(self, seed):
        self.seed = seed
        self.pos = 0

    def randn(self, n):
        u = (synth.splitmix64(self.seed, 2 * n, start=self.pos) >> np.uint64(11)).astype(np.float64)
        u *= 1.0 / 9007199254740992.0
        self.pos += 2 * n
        u1 = np.maximum(u[0::2], 1e-15)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[1::2])

    def mat(self, rows, cols):
        return self.randn(rows * cols).reshape(rows, cols)


def identity_case():
    """test_procrustes_identity: seed 77, 50 x 20, T(X, X) = I (csrc/ha_test.c:234-256)."""
    return RefRng(77).mat(50, 20)


def rotation_case():
    """test_procrustes_known_rotation: seed 55, 50 x 10, Givens 0.7 rad in the (0, 1)
    plane, Y = X R (csrc/ha_test.c:258-299)."""
    X = RefRng(55).mat(50, 10)
    R = np.eye(10)
    a = 0.7
    R[0, 0], R[0, 1], R[1, 0], R[1, 1] = np.cos(a), -np.sin(a), np.sin(a), np.cos(a)
    return X, X @ R, R


def no_reflection_case():
    """test_procrustes_no_reflection: seed 33, X then Y 50 x 10 (csrc/ha_test.c:301-332)."""
    r = RefRng(33)
    X = r.mat(50, 10)
    Y = r.mat(50, 10)
    return X, Y


def zscore_case():
    """test_zscore: seed 200, 100 x 10 (csrc/ha_test.c:474-499)."""
    return RefRng(200).mat(100, 10)


def template_case():
    """test_template_identical_subjects: seed 300, nt 30, nv 10, 5 identical subjects
    (csrc/ha_test.c:503-535)."""
    base = RefRng(300).mat(30, 10)
    return base, [base.copy() for _ in range(5)]


# csrc/ha_test.c:369-402 (two searchlights on five vertices, nnz == 21)
SPARSE_TOY = [[0, 1, 2], [1, 2, 3, 4]]
# csrc/ha_test.c:439-470 (three searchlights on six vertices)
WEIGHTS_TOY = [[0, 1, 2], [1, 2, 3, 4], [3, 4, 5]]


def ridge_case():
    """test_ridge_small_alpha: seed 101, X 50 x 10 then true betas 10 x 5, Y = X B;
    ridge with alpha 1e-6 recovers B to 1e-6 (csrc/ha_test.c:336-362)."""
    r = RefRng(101)
    X = r.mat(50, 10)
    B = r.mat(10, 5)
    return X, X @ B, B
