"""Pins of the oracle's L-mode product (oracle/matrix_free.py)."""
import json
import os

import numpy as np
import pytest
from scipy import sparse as sp

from igaops_amd.common import kron_nonzero_indices
from oracle import matrix_free as mf

GOLD = os.path.join(os.path.dirname(__file__), "golden", "kron_known_answers.json")


def test_kron_known_answer_matrixfree():
    g = json.load(open(GOLD))["matrixfree_kron"]
    A, B, Cm = (np.array(g[k], dtype=float) for k in "ABC")
    v = np.random.default_rng(0).random(12)
    ref = np.kron(Cm, np.kron(B, A)) @ v
    np.testing.assert_allclose(mf.apply([A, B, Cm], v), ref, rtol=1e-14)
    # the same with the dimension-1 factor applied first by hand
    arr = v.reshape(6, 2)
    t = np.einsum("ij,kj->ki", A, arr)
    np.testing.assert_allclose(np.ravel(np.kron(Cm, B) @ t), ref, rtol=1e-14)


def test_kron_nonzero_indices_known_answer():
    g = json.load(open(GOLD))["kron_nonzero_indices"]
    v1, v2, v3 = (np.array(v) for v in g["vectors"])
    ref = np.nonzero(np.kron(np.kron(v1, v2), v3))[0]
    got = kron_nonzero_indices(g["indices_list"], g["nnz_list"])
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("ndim", [1, 2, 3, 4])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("sparse", [False, True])
def test_apply_matches_brute_force_kron(ndim, transpose, sparse):
    rng = np.random.default_rng(ndim)
    mats = [rng.standard_normal((rng.integers(2, 5), rng.integers(2, 5))) for _ in range(ndim)]
    K = mats[0]
    for m in mats[1:]:
        K = np.kron(m, K)
    if transpose:
        K = K.T
    ops = [sp.csr_array(m) for m in mats] if sparse else mats
    v = rng.standard_normal(K.shape[1])
    np.testing.assert_allclose(mf.apply(ops, v, is_transpose=transpose), K @ v, rtol=1e-12, atol=1e-12)
    V = rng.standard_normal((K.shape[1], 3))
    np.testing.assert_allclose(mf.apply(ops, V, is_transpose=transpose), K @ V, rtol=1e-12, atol=1e-12)


def test_mode_product_transposition_is_caught():
    """A transposed factor must change the answer (guards a swapped operand)."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 3))
    B = rng.standard_normal((3, 3))
    v = rng.standard_normal(9)
    assert not np.allclose(mf.apply([A, B], v), mf.apply([A.T, B], v))
    assert not np.allclose(mf.apply([A, B], v), mf.apply([B, A], v))
