"""Pins of the learned integrator's CPU restatement (oracle/surrogate.py)."""
import numpy as np

from oracle import surrogate as S
from pyiid_amd import structures
from pyiid_amd.surrogate import SurrogateNet

RBF = (6.0, 2.0, 4.0 / 15.0, (15.0 / 4.0) ** 2)


def _setup(n=(20, 13), seed=0):
    pos = [structures.jitter(structures.fcc_nearest(k), 0.08, seed=seed + k) for k in n]
    offsets = np.concatenate([[0], np.cumsum([len(p) for p in pos])])
    return np.concatenate(pos).astype(np.float64), offsets


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 1.0 + 2 ** -8], np.float32)
    # ties go to even: 1 + 2^-9 -> 1, 1 + 3 * 2^-9 -> 1 + 2^-7; 1 + 2^-8 is a tie -> 1
    np.testing.assert_array_equal(S.bf16(x), [1.0, 1.0, 1.0 + 2 ** -7, -2.5, 0.0, 1.0])


def test_descriptor_single_pair_closed_form():
    rc, mu0, dmu, eta = RBF
    r = 3.1
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    d = S.descriptors(pos, [0, 2], rc, mu0, dmu, eta)
    mu = mu0 + dmu * np.arange(16)
    want = np.exp(-eta * (r - mu) ** 2) * 0.5 * (np.cos(np.pi * r / rc) + 1)
    np.testing.assert_allclose(d[0], want, rtol=1e-14)
    np.testing.assert_allclose(d[1], want, rtol=1e-14)
    # beyond the cutoff nothing contributes
    far = S.descriptors(np.array([[0, 0, 0], [rc + 1e-9, 0, 0]], float), [0, 2], rc, mu0, dmu, eta)
    assert np.all(far == 0)


def test_descriptor_invariances_and_chain_separation():
    pos, off = _setup()
    rc, mu0, dmu, eta = RBF
    d = S.descriptors(pos, off, rc, mu0, dmu, eta)
    rs = np.random.RandomState(1)
    q, _ = np.linalg.qr(rs.normal(size=(3, 3)))
    d2 = S.descriptors(pos @ q.T + np.array([3.0, -1.0, 7.0]), off, rc, mu0, dmu, eta)
    np.testing.assert_allclose(d2, d, rtol=1e-10, atol=1e-12)
    # moving chain 1 on top of chain 0 changes nothing: chains do not interact
    moved = pos.copy()
    moved[off[1]:] += pos[0] - pos[off[1]]
    np.testing.assert_allclose(S.descriptors(moved, off, rc, mu0, dmu, eta), d, rtol=1e-12)


def test_forces_are_minus_gradient_of_energy():
    pos, off = _setup((14, 9), seed=3)
    net = SurrogateNet(seed=2).as_dict()
    net["mean"] = np.full(16, 0.5, np.float32)
    net["inv_scale"] = np.full(16, 0.7, np.float32)
    w3 = np.random.RandomState(5).normal(size=(2, 64))
    E, F = S.energy_and_forces(pos, off, RBF, net, w3)
    h = 1e-5
    for a in [0, 5, 13, 14, 22]:
        for c in range(3):
            p1 = pos.copy()
            p1[a, c] += h
            p2 = pos.copy()
            p2[a, c] -= h
            e1 = S.energy_and_forces(p1, off, RBF, net, w3)[0].sum()
            e2 = S.energy_and_forces(p2, off, RBF, net, w3)[0].sum()
            fd = -(e1 - e2) / (2 * h)
            assert abs(fd - F[a, c]) <= 1e-6 * max(1.0, abs(F[a, c]))
    # forces of a chain sum to zero (pairwise antisymmetric terms)
    for b in range(2):
        np.testing.assert_allclose(F[off[b]:off[b + 1]].sum(axis=0), 0.0, atol=1e-9)


def test_ridge_fit_recovers_exact_weights():
    rs = np.random.RandomState(0)
    phi = rs.normal(size=(64, 500))
    w = rs.normal(size=64)
    got, r2 = S.ridge_fit(phi, phi.T @ w, 1e-12)
    np.testing.assert_allclose(got, w, rtol=1e-6, atol=1e-8)
    assert r2 > 1 - 1e-10


def _states(rs, n_states=3, m=120):
    return [(rs.normal(size=(64, m)), rs.normal(size=m), rs.normal(size=64), rs.normal())
            for _ in range(n_states)]


def test_fit_over_states_is_the_weighted_stacked_least_squares():
    # rows of state t carry sqrt(forget^(T-1-t)); energy row t (t >= 1) has
    # the same decay and sqrt(weight); ridge rows sqrt(lam) I: a QR solve of
    # that stacked system must give the normal-equation answer
    rs = np.random.RandomState(1)
    forget, ridge, weight = 0.5, 1e-3, 7.0
    st = _states(rs)
    T = len(st)
    rows, rhs = [], []
    for t, (phi, y, e, E) in enumerate(st):
        s = np.sqrt(forget ** (T - 1 - t))
        rows.append(s * phi.T)
        rhs.append(s * y)
        if t > 0:
            rows.append(s * np.sqrt(weight) * (e - st[t - 1][2])[None, :])
            rhs.append(s * np.sqrt(weight) * np.array([E - st[t - 1][3]]))
    A = np.vstack(rows)
    b = np.concatenate(rhs)
    lam = ridge * np.trace(A.T @ A) / 64
    A = np.vstack([A, np.sqrt(lam) * np.eye(64)])
    b = np.concatenate([b, np.zeros(64)])
    want = np.linalg.lstsq(A, b, rcond=None)[0]
    np.testing.assert_allclose(S.fit_over_states(st, forget, ridge, weight), want, rtol=1e-9,
                               atol=1e-12)
    # weight 0 drops the energy rows; a single state is ridge_fit
    np.testing.assert_allclose(S.fit_over_states(st[:1], forget, ridge, weight),
                               S.ridge_fit(st[0][0], st[0][1], ridge)[0], rtol=1e-10)


def test_energy_rows_recover_a_consistent_network():
    # forces and energies of one linear head w_true: the fit recovers it, and
    # energy rows alone (forces 0 at few columns) pin the energy direction
    rs = np.random.RandomState(2)
    w_true = rs.normal(size=64)
    st = []
    for _ in range(4):
        phi = rs.normal(size=(64, 200))
        e = rs.normal(size=64)
        st.append((phi, phi.T @ w_true, e, float(e @ w_true) + 3.0))  # constant offset is free
    np.testing.assert_allclose(S.fit_over_states(st, 1.0, 1e-12, 5.0), w_true, rtol=1e-6,
                               atol=1e-8)
    # energy changes inconsistent with the forces move the fit towards them
    bent = [(p, y, e, E + 10.0 * k) for k, (p, y, e, E) in enumerate(st)]
    w0 = S.fit_over_states(bent, 1.0, 1e-12, 0.0)
    w1 = S.fit_over_states(bent, 1.0, 1e-12, 1e3)
    de = bent[-1][2] - bent[-2][2]
    dE = bent[-1][3] - bent[-2][3]
    assert abs(de @ w1 - dE) < 0.5 * abs(de @ w0 - dE)
