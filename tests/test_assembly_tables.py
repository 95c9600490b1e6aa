"""Host tables of the sum-factorised assembly (igaops_amd/assembly.py), CPU only:
the 1-D (i, j, q) products and their union pattern reproduce W[i, q] B[q, j]."""
import numpy as np

from igaops_amd.assembly import AsmTerm, _one_dim_products, _tables, asm_terms_gradu_gradv
from igaops_amd.quadrature import StandardGauss, WeightedQuadrature, make_open_knotvector


def test_one_dim_products_match_dense():
    q = WeightedQuadrature(3, make_open_knotvector(3, 5), "1")
    for z, b in [(0, 0), (1, 1), (3, 1), (2, 0)]:
        W, B = q.weights[z].toarray(), q.basis[b].toarray()
        I, J, Q, V = _one_dim_products(q.weights[z], q.basis[b])
        dense = np.zeros((W.shape[0], B.shape[1], W.shape[1]))
        dense[I, J, Q] = V
        np.testing.assert_array_equal(dense, W[:, None, :] * B.T[None, :, :])
        keys = (I * B.shape[1] + J) * W.shape[1] + Q
        assert np.all(np.diff(keys) > 0)  # sorted by (i, j, q), no duplicates


def test_tables_cover_every_term():
    p, nel = 2, 3
    qs = [StandardGauss(p, make_open_knotvector(p, nel), "legendre")] * 2 + \
        [WeightedQuadrature(p, make_open_knotvector(p, nel + 1), "2")]
    nq = int(np.prod([q.nbquadpts for q in qs]))
    C = np.ones((3, 3, nq))
    terms = asm_terms_gradu_gradv(qs, C)
    nr, nqs, nc, pat_ptr, pat_col, lst_ptr, lst_q, lst_v = _tables(qs, terms)
    for d, q in enumerate(qs):
        assert pat_ptr[d][-1] == pat_col[d].size
        for t, term in enumerate(terms):
            W, B = q.weights[term.weights[d]].toarray(), q.basis[term.basis[d]].toarray()
            # rebuild W[i, q] B[q, j] from the pattern and the term's lists
            dense = np.zeros((nr[d], nc[d], nqs[d]))
            for i in range(nr[d]):
                for p_ in range(pat_ptr[d][i], pat_ptr[d][i + 1]):
                    j = pat_col[d][p_]
                    lo, hi = lst_ptr[t][d][p_], lst_ptr[t][d][p_ + 1]
                    dense[i, j, lst_q[d][lo:hi]] = lst_v[d][lo:hi]
            np.testing.assert_array_equal(dense, W[:, None, :] * B.T[None, :, :])


def test_planted_table_error_is_detected():
    """Swapping the trial and test indices of one term changes the tables."""
    q = WeightedQuadrature(2, make_open_knotvector(2, 4), "1")
    a = _tables([q], [AsmTerm([1], [1], np.ones(q.nbquadpts))])
    b = _tables([q], [AsmTerm([2], [0], np.ones(q.nbquadpts))])
    assert not (np.array_equal(a[4], b[4]) and np.array_equal(a[7][0], b[7][0]))


def test_oracle_sumfact_terms_reproduces_mass_and_stiffness():
    from oracle import assembly as asm

    q = WeightedQuadrature(2, make_open_knotvector(2, 3), "1")
    qs = [q] * 3
    nq = q.nbquadpts ** 3
    rng = np.random.default_rng(2)
    C = rng.uniform(0.5, 1.5, (3, 3, nq))
    m = rng.uniform(0.5, 1.5, nq)
    terms = []
    for j in range(3):
        beta = np.eye(3, dtype=int)[j]
        for i in range(3):
            terms.append((beta + 2 * np.eye(3, dtype=int)[i], beta, C[i, j], 1.0))
    np.testing.assert_allclose(asm.sumfact_terms(qs, terms).toarray(), asm.bspline_stiffness(qs, C).toarray(),
                               rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(asm.sumfact_terms(qs, [([0] * 3, [0] * 3, m, 1.0)]).toarray(),
                               asm.bspline_mass(qs, m).toarray(), rtol=1e-13, atol=1e-15)


def test_spacetime_gradu_gradv_term_list_decomposes_into_pinned_operators():
    """The four-direction term list of the space-time ``assemble_scalar_gradu_gradv``
    (``bsplines_ops.py:161-191``, time_ders (0, 0)) splits, for rules whose test
    weights ignore the trial derivative (WQ type "2": weights[1] == weights[0],
    weights[3] == weights[2] in every direction), into operators the oracle states
    independently: the spatial stiffness with C[:3, :3], ``gradu_v`` (test gradient)
    with C[i][3], ``u_gradv`` (trial gradient) with C[3][j] and the mass with C[3][3].
    A swapped C index or a dropped time term changes the product."""
    from scipy import sparse as sp

    from oracle import assembly as asm
    from oracle import operators as oo

    qs = [WeightedQuadrature(2, make_open_knotvector(2, 2), "2")] * 3 + \
        [WeightedQuadrature(1, make_open_knotvector(1, 3), "2")]
    for q in qs:
        for a, b in ((1, 0), (3, 2)):
            assert abs(sp.csr_array(q.weights[a]) - sp.csr_array(q.weights[b])).sum() == 0
    nq = int(np.prod([q.nbquadpts for q in qs]))
    rng = np.random.default_rng(21)
    C = rng.uniform(0.5, 1.5, (4, 4, nq))
    td = (0, 0)
    terms = [(t.weights, t.basis, np.asarray(t.coef), t.scale) for t in asm_terms_gradu_gradv(qs, C, td)]
    K = asm.sumfact_terms(qs, terms)
    v = rng.standard_normal(K.shape[1])
    ref = (oo.bspline_gradu_gradv(qs, C[:3, :3], v, time_ders=td)
           + oo.bspline_gradu_v(qs, C[:3, 3], v, time_ders=td)
           + oo.bspline_u_gradv(qs, C[3, :3], v, time_ders=td)
           + oo.bspline_u_v(qs, C[3, 3], v, time_ders=td))
    np.testing.assert_allclose(K @ v, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    # the decomposition is not vacuous: the C[3][j] and C[i][3] blocks are not symmetric
    Ct = C.copy()
    Ct[3, :3], Ct[:3, 3] = C[:3, 3], C[3, :3]
    terms_t = [(t.weights, t.basis, np.asarray(t.coef), t.scale) for t in asm_terms_gradu_gradv(qs, Ct, td)]
    assert np.abs(asm.sumfact_terms(qs, terms_t) @ v - ref).max() > 1e-6 * np.abs(ref).max()
