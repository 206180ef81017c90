"""Pins of the FP64 oracle's pair indexing, normalisation, F(Q) and gradient.

None of these retypes the oracle's formula: they use closed forms (dimer,
square, small-Q limit), an independent special function (spherical Bessel
j0), a second code path (N x N sums), invariances and finite differences.
"""
import numpy as np
import pytest
from scipy.special import spherical_jn

from oracle import scatter as S


def rand_atoms(n, seed, box=8.0):
    rs = np.random.RandomState(seed)
    return rs.random_sample((n, 3)) * box


def rand_table(n, nq, seed, n_elem=3):
    rs = np.random.RandomState(seed)
    tab = 1.0 + rs.random_sample((n_elem, nq)) * 5.0
    idx = rs.randint(0, n_elem, n)
    return tab[idx]


def test_k_to_ij_bruteforce():
    n = 60
    ks = []
    for i in range(n):
        for j in range(i):
            ks.append((S.ij_to_k(i, j), i, j))
    ks.sort()
    assert [k for k, _, _ in ks] == list(range(n * (n - 1) // 2))
    i, j = S.k_to_ij(np.array([k for k, _, _ in ks]))
    assert np.array_equal(i, [a for _, a, _ in ks]) and np.array_equal(j, [b for _, _, b in ks])
    # large indices (beyond where float32 decoding breaks, N ~ 4.6k)
    big = np.array([10619135, 22434950, 199989999])
    i, j = S.k_to_ij(big)
    assert np.all(S.ij_to_k(i, j) == big) and np.all(j < i)


def test_pair_geometry_flat_vs_nxn():
    q = rand_atoms(17, 1)
    d, r = S.pair_geometry(q)
    dn, rn = S.pair_geometry_nxn(q)
    i, j = S.k_to_ij(np.arange(len(r)))
    np.testing.assert_array_equal(d, -dn[i, j])     # flat d = q_i - q_j, nxn d = q_j - q_i
    np.testing.assert_array_equal(r, rn[i, j])


def test_normalization_closed_form_vs_pairs():
    f = rand_table(23, 40, 2)
    np.testing.assert_allclose(S.normalization(f), S.normalization_bruteforce(f), rtol=1e-12)


def test_fq_dimer_closed_form():
    q = np.array([[0.0, 0.0, 0.0], [2.5, 0.3, -0.7]])
    f = rand_table(2, 250, 3, n_elem=2)
    Q = S.q_values(0.1, 250)
    r = np.linalg.norm(q[1] - q[0])
    np.testing.assert_allclose(S.fq(q, f, 0.1, cast_f32=False), np.sin(Q * r) / r,
                               rtol=1e-12, atol=1e-14)


def test_fq_square_closed_form():
    # Au4 square of side 3 A (pyiid/tests/__init__.py:131-141): 4 edges, 2 diagonals
    q = np.array([[0, 0, 0], [3, 0, 0], [0, 3, 0], [3, 3, 0]], float)
    f = np.full((4, 250), 79.0)
    Q = S.q_values(0.1, 250)
    d = 3 * np.sqrt(2)
    want = 2.0 / 4 * (4 * np.sin(3 * Q) / 3 + 2 * np.sin(d * Q) / d)
    np.testing.assert_allclose(S.fq(q, f, 0.1), want, rtol=1e-12, atol=1e-14)


def test_fq_vs_spherical_bessel():
    q = rand_atoms(15, 4)
    f = rand_table(15, 120, 5)
    Q = S.q_values(0.2, 120)
    _, r = S.pair_geometry(q)
    i, j = S.k_to_ij(np.arange(len(r)))
    ff = np.asarray(f, np.float32).astype(np.float64)
    s = np.sum(ff[i] * ff[j] * Q[None, :] * spherical_jn(0, np.outer(r, Q)), axis=0)
    want = 2 * s / S.normalization(f)
    np.testing.assert_allclose(S.fq(q, f, 0.2), want, rtol=1e-10, atol=1e-12)


def test_fq_flat_vs_nxn_and_chunking():
    q = rand_atoms(19, 6)
    f = rand_table(19, 80, 7)
    a = S.fq(q, f, 0.15)
    np.testing.assert_allclose(a, S.fq_nxn(q, f, 0.15), rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(a, S.fq(q, f, 0.15, chunk_elems=37), rtol=1e-12, atol=1e-14)
    qi = np.array([3, 17, 79])
    np.testing.assert_allclose(a[qi], S.fq(q, f, 0.15, q_index=qi), rtol=1e-12)


def test_fq_invariances_and_limits():
    q = rand_atoms(12, 8)
    f = np.full((12, 100), 10.0)
    a = S.fq(q, f, 0.05, cast_f32=False)
    rs = np.random.RandomState(9)
    rot, _ = np.linalg.qr(rs.randn(3, 3))
    b = S.fq(q @ rot.T + np.array([1.5, -2.0, 0.3]), f, 0.05, cast_f32=False)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-11)
    assert a[0] == 0.0
    # small Q: sin(Qr)/r -> Q, so F -> 2/N * K * Q = (N - 1) Q
    small = S.fq(q, np.full((12, 3), 10.0), 1e-3, cast_f32=False)
    Q = S.q_values(1e-3, 3)
    np.testing.assert_allclose(small[1:], (12 - 1) * Q[1:], rtol=1e-4)


def test_fq_nan_to_num_zero_table():
    q = rand_atoms(5, 10)
    f = np.zeros((5, 30))
    assert not np.any(S.fq(q, f, 0.1))


def test_grad_is_minus_half_dF():
    q = rand_atoms(7, 11)
    f = rand_table(7, 60, 12)
    g = S.grad_fq(q, f, 0.2, cast_f32=False)
    h = 1e-6
    for a in range(7):
        for c in range(3):
            qp, qm = q.copy(), q.copy()
            qp[a, c] += h
            qm[a, c] -= h
            dF = (S.fq(qp, f, 0.2, cast_f32=False) - S.fq(qm, f, 0.2, cast_f32=False)) / (2 * h)
            np.testing.assert_allclose(-2.0 * g[a, c], dF, rtol=1e-5, atol=1e-7)


def test_grad_invariants_and_nxn():
    q = rand_atoms(14, 13)
    f = rand_table(14, 70, 14)
    g = S.grad_fq(q, f, 0.1, cast_f32=False)
    scale = np.max(np.abs(g))
    # translation invariance: sum_i grad_i = 0
    np.testing.assert_allclose(g.sum(axis=0), 0.0, atol=1e-11 * scale * 14)
    # rotation invariance: sum_i q_i x grad_i = 0
    torque = np.cross(q[:, :, None], g, axisa=1, axisb=1, axisc=1).sum(axis=0)
    np.testing.assert_allclose(torque, 0.0, atol=1e-10 * scale * 14 * 8)
    np.testing.assert_allclose(g, S.grad_fq_nxn(q, f, 0.1, cast_f32=False), rtol=1e-10,
                               atol=1e-12 * scale)
    rows, qi = np.array([1, 5, 13]), np.array([0, 9, 69])
    np.testing.assert_allclose(S.grad_fq(q, f, 0.1, q_index=qi, atom_index=rows,
                                         cast_f32=False), g[rows][:, :, qi], rtol=1e-12)


def test_grad_dimer_closed_form():
    q = np.array([[0.1, 0.2, 0.3], [2.0, -0.4, 1.1]])
    f = np.full((2, 50), 3.0)
    Q = S.q_values(0.3, 50)
    d = q[0] - q[1]
    r = np.linalg.norm(d)
    a = (Q * np.cos(Q * r) - np.sin(Q * r) / r) / r ** 2
    g = S.grad_fq(q, f, 0.3, cast_f32=False)
    # grad[0] = -(1/na) f^2 a d with na = 2 f^2 (N=2)
    np.testing.assert_allclose(g[0], -(a[None, :] * d[:, None]) / 2, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(g[1], -g[0], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("qbin", [0.1, 0.02])
def test_q_grid_uses_float32_qbin(qbin):
    Q = S.q_values(qbin, 10)
    assert Q[1] == float(np.float32(qbin))
