"""Learned integrator on the device against its CPU restatement
(oracle/surrogate.py) and the sampler invariants."""
import math

import numpy as np
import pytest

from oracle import scatter as S
from oracle import surrogate as OS
from pyiid_amd import _lib, structures
from pyiid_amd.elasticscatter import ElasticScatter
from pyiid_amd.formfactors import element_table
from pyiid_amd.sim import SurrogateHMC

from test_gpu_hmc import chain_energy_forces, normal_key, uniform_key

pytestmark = pytest.mark.gpu

EXP = {"qmin": 0.5, "rmax": 12.0}


def _chains(sizes, seed=0, sigma=0.08):
    out = []
    for k, n in enumerate(sizes):
        base = structures.fcc_nearest(n)
        out.append((base, structures.jitter(base, sigma, seed=seed + 10 + k), np.full(n, 79)))
    return out


def _targets(chains, pq, e):
    r = S.r_grid(e)
    nq = S.qmax_bin(e["qmax"], pq)
    table, _, _ = element_table(np.full(1, 79), pq, nq)
    return np.array([S.get_pdf_at_qmin(S.fq(b.astype(np.float32), table[np.zeros(len(b), int)], pq),
                                       e["rstep"], pq, r, e["qmin"]) for b, _, _ in chains])


def _sampler(chains, conv=50.0, net_seed=0, **kw):
    e, pq = S.make_experiment(dict(EXP))
    scat = ElasticScatter(dict(EXP), seed=0)
    targets = _targets(chains, pq, e)
    hmc = SurrogateHMC(scat, [(c[1], c[2]) for c in chains], targets, conv=conv, masses="unit",
                       net_seed=net_seed, **kw)
    return hmc, targets, e, pq


def _state(hmc):
    pos = np.concatenate(hmc.positions()).astype(np.float64)
    off = hmc.offsets
    chain = np.repeat(np.arange(hmc.B), np.diff(off))
    return pos, off, chain


def test_descriptors_match_oracle():
    chains = _chains([40, 17, 2])
    hmc, _, _, _ = _sampler(chains)
    pos, off, _ = _state(hmc)
    net = hmc.net
    # with the identity standardisation the raw descriptors come back
    hmc.ctx.sur_setup(net.params, net.w1, net.b1, net.w2, net.b2, np.zeros(16), np.ones(16),
                      hmc.max_nbr)
    got = hmc.ctx.sur_descriptors(len(pos))
    want = OS.descriptors(pos, off, *net.rbf)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-5 * want.max())
    hmc.close()


def test_network_and_forces_match_oracle():
    chains = _chains([40, 23], seed=1)
    hmc, _, _, _ = _sampler(chains)
    pos, off, chain = _state(hmc)
    w3 = np.random.RandomState(3).normal(0, 0.3, (hmc.B, 64)).astype(np.float32)
    hmc.ctx.sur_set_output(w3)
    f, e = hmc.ctx.sur_forces(len(pos), energies=True)
    d = hmc.ctx.sur_descriptors(len(pos)).astype(np.float64)
    e_ref, c_ref = OS.network(d, hmc.net.as_dict(), w3, chain, bf16_round=True)
    np.testing.assert_allclose(e, e_ref, rtol=2e-3, atol=2e-3 * np.abs(e_ref).max())
    f_ref = OS.forces(pos, off, hmc.net.rbf[0], *hmc.net.rbf[1:], c_ref)
    assert np.linalg.norm(f - f_ref) <= 2e-3 * np.linalg.norm(f_ref)
    # the unrounded network differs only by BF16 rounding (well below 5%)
    _, f64 = OS.energy_and_forces(pos, off, hmc.net.rbf, hmc.net.as_dict(), w3)
    assert np.linalg.norm(f - f64) <= 5e-2 * np.linalg.norm(f64)
    hmc.close()


def test_feature_forces_are_unit_output_forces():
    chains = _chains([30, 12], seed=2)
    hmc, _, _, _ = _sampler(chains)
    pos, off, chain = _state(hmc)
    n = len(pos)
    phi = hmc.ctx.sur_features(n)
    d = hmc.ctx.sur_descriptors(n).astype(np.float64)
    for k in [0, 17, 63]:
        w3 = np.zeros((hmc.B, 64), np.float32)
        w3[:, k] = 1.0
        _, c = OS.network(d, hmc.net.as_dict(), w3, chain, bf16_round=True)
        want = OS.forces(pos, off, *hmc.net.rbf, c).reshape(-1)
        assert np.linalg.norm(phi[k] - want) <= 2e-3 * np.linalg.norm(want)
    hmc.close()


def test_training_matches_ridge_fit_on_exact_forces():
    chains = _chains([35, 20], seed=4)
    hmc, targets, e, pq = _sampler(chains, conv=50.0)
    pos, off, _ = _state(hmc)
    n = len(pos)
    r2 = hmc.train(forget=0.0, ridge=1e-6)
    phi = hmc.ctx.sur_features(n).astype(np.float64)
    en, fexact = hmc.ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, targets, _lib.POT_RW, 50.0, n)
    fexact = 2.0 * fexact   # exact -dE/dq (the reference convention is half)
    for b in range(hmc.B):
        cols = slice(3 * off[b], 3 * off[b + 1])
        w, r2_ref = OS.ridge_fit(phi[:, cols], fexact[off[b]:off[b + 1]].reshape(-1), 1e-6)
        assert abs(r2[b] - r2_ref) < 2e-3
    hmc.close()


def test_exact_fallback_equals_exact_hmc(fixed_shapes):
    chains = _chains([25, 18], seed=5)
    runs = []
    for kind in ("exact", "fallback"):
        hmc, _, _, _ = _sampler(chains, conv=100.0)
        if kind == "exact":
            out = hmc.run_exact(6, n_leapfrog=4, step=0.01, beta=5.0, seed=11)
            runs.append((out[0], out[1], hmc.positions()))
        else:
            out = hmc.run(6, n_leapfrog=4, step=0.01, beta=5.0, seed=11, train_every=3,
                          min_r2=np.inf)
            assert np.all(out["exact_mode"] == 1)
            runs.append((out["energies"], out["accepted"], hmc.positions()))
        hmc.close()
    (e1, a1, p1), (e2, a2, p2) = runs
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_allclose(e1, e2, rtol=1e-12)
    for x, y in zip(p1, p2):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("use_graph", [False, True])
def test_surrogate_run_energies_are_exact(use_graph):
    chains = _chains([30, 44, 21], seed=6)
    hmc, targets, e, pq = _sampler(chains, conv=100.0)
    # small chains at conv 100 are stiff (force rms ~16): a short step keeps
    # the surrogate's energy error, and so the rejection rate, small
    out = hmc.run(8, n_leapfrog=5, step=0.001, beta=5.0, seed=3, train_every=4,
                  use_graph=use_graph)
    assert np.all(out["exact_mode"] == 0)
    assert np.all(np.diff(out["times_ms"]) > 0)
    n = int(hmc.offsets[-1])
    en, _ = hmc.ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, targets, _lib.POT_RW, 100.0, n,
                                      forces=False)
    np.testing.assert_allclose(out["energies"][-1], en, rtol=1e-9)
    assert 0 < out["accepted"].mean() <= 1
    hmc.close()


def test_graph_and_direct_surrogate_runs_agree(fixed_shapes):
    chains = _chains([26, 33], seed=7)
    res = []
    for g in (False, True):
        hmc, _, _, _ = _sampler(chains, conv=100.0)
        out = hmc.run(10, n_leapfrog=5, step=0.002, beta=5.0, seed=9, train_every=5, use_graph=g)
        res.append((out, hmc.positions()))
        hmc.close()
    (o1, p1), (o2, p2) = res
    np.testing.assert_array_equal(o1["accepted"], o2["accepted"])
    np.testing.assert_allclose(o1["energies"], o2["energies"], rtol=1e-12)
    for x, y in zip(p1, p2):
        np.testing.assert_array_equal(x, y)


def test_one_surrogate_proposal_matches_cpu_replay():
    chains = _chains([19, 27], seed=8)
    conv, beta, eps, n_leap, seed = 60.0, 8.0, 0.02, 3, 4242
    hmc, targets, e, pq = _sampler(chains, conv=conv)
    pos, off, chain = _state(hmc)
    w3 = np.random.RandomState(2).normal(0, 0.05, (hmc.B, 64)).astype(np.float32)
    hmc.ctx.sur_set_output(w3)
    out = hmc.run(1, n_leapfrog=n_leap, step=eps, beta=beta, seed=seed, train_every=0,
                  use_graph=False)
    got = hmc.positions()
    net = hmc.net.as_dict()

    def sur_force(x):
        d = OS.descriptors(x, off, *hmc.net.rbf)
        d = d.astype(np.float32).astype(np.float64)
        _, c = OS.network(d, net, w3, chain, bf16_round=True)
        return OS.forces(x, off, *hmc.net.rbf, c)

    x = pos.copy()
    p = np.array([[normal_key(seed, 0, 3 * i + c) / math.sqrt(beta) for c in range(3)]
                  for i in range(len(x))])
    k0 = 0.5 * np.array([np.sum(p[off[b]:off[b + 1]] ** 2) for b in range(hmc.B)])
    p = p + 0.5 * eps * sur_force(x)
    for l in range(1, n_leap + 1):
        x = x + eps * p
        p = p + (eps if l < n_leap else 0.5 * eps) * sur_force(x)
    k1 = 0.5 * np.array([np.sum(p[off[b]:off[b + 1]] ** 2) for b in range(hmc.B)])
    for b, (base, _, z) in enumerate(chains):
        e0 = chain_energy_forces(pos[off[b]:off[b + 1]], z, targets[b], conv, pq, e)[0]
        e1 = chain_energy_forces(x[off[b]:off[b + 1]], z, targets[b], conv, pq, e)[0]
        acc = uniform_key(seed, 0, 2, b) < math.exp(beta * ((e0 + k0[b]) - (e1 + k1[b])))
        assert out["accepted"][0, b] == int(acc)
        want = x[off[b]:off[b + 1]] if acc else pos[off[b]:off[b + 1]]
        np.testing.assert_allclose(got[b], want, atol=2e-5)
        assert abs(out["energies"][0, b] - (e1 if acc else e0)) < 1e-4 * max(1.0, abs(e0))
    hmc.close()


def test_displacement_guard_rejects_and_counts():
    chains = _chains([20, 15], seed=9)
    hmc, _, _, _ = _sampler(chains, conv=10.0)
    before = hmc.positions()
    w3 = np.zeros((hmc.B, 64), np.float32)
    hmc.ctx.sur_set_output(w3)     # free flight: moves of ~1 A exceed skin / 2
    out = hmc.run(3, n_leapfrog=5, step=0.5, beta=1.0, seed=1, train_every=0)
    assert np.all(out["accepted"] == 0)
    np.testing.assert_array_equal(out["violations"], [3, 3])
    for x, y in zip(before, hmc.positions()):
        np.testing.assert_array_equal(x, y)
    hmc.close()


def test_neighbour_capacity_is_checked():
    chains = _chains([40], seed=10)
    with pytest.raises(ValueError, match="neighbours"):
        _sampler(chains, max_nbr=8)


def test_low_acceptance_switches_to_exact_forces():
    """min_acceptance: a block whose first three learned-force proposals were
    accepted less often than the bar runs the rest of the block with exact
    forces, and the next block tries the learned forces again.  Blocks of
    three or fewer proposals are judged when they end: a failed one is
    followed by a block of exact forces.  A bar above 1 fails every learned
    proposal; a bar of 0 never switches."""
    chains = _chains([24, 30], seed=11)
    hmc, _, _, _ = _sampler(chains, conv=50.0)
    out = hmc.run(30, n_leapfrog=3, step=0.001, beta=5.0, seed=2, train_every=10,
                  min_acceptance=1.01)
    np.testing.assert_array_equal(out["exact_mode"], ([0] * 3 + [1] * 7) * 3)
    out = hmc.run(12, n_leapfrog=3, step=0.001, beta=5.0, seed=4, train_every=3,
                  min_acceptance=1.01)
    np.testing.assert_array_equal(out["exact_mode"], ([0] * 3 + [1] * 3) * 2)
    out = hmc.run(20, n_leapfrog=3, step=0.001, beta=5.0, seed=3, train_every=10,
                  min_acceptance=0.0)
    np.testing.assert_array_equal(out["exact_mode"], [0] * 20)
    hmc.close()


def test_feature_energies_are_unit_output_chain_energies():
    chains = _chains([30, 12], seed=2)
    hmc, _, _, _ = _sampler(chains)
    pos, off, chain = _state(hmc)
    n = len(pos)
    ef = hmc.feature_energies()
    assert ef.shape == (hmc.B, 64)
    d = hmc.ctx.sur_descriptors(n).astype(np.float64)
    for k in [0, 17, 63]:
        w3 = np.zeros((hmc.B, 64), np.float32)
        w3[:, k] = 1.0
        e, _ = OS.network(d, hmc.net.as_dict(), w3, chain, bf16_round=True)
        want = np.array([e[off[b]:off[b + 1]].sum() for b in range(hmc.B)])
        np.testing.assert_allclose(ef[:, k], want, rtol=2e-3, atol=2e-3 * np.abs(want).max())
    hmc.close()


def _two_state_fit(chains, weight, ridge):
    """Train at the start and again after two exact proposals (forget 1);
    returns the device's output layer and each state's (phi, F, e, E)."""
    hmc, targets, e, pq = _sampler(chains, conv=50.0, energy_weight=weight)
    pos, off, _ = _state(hmc)
    n = len(pos)
    samples = []
    for step in range(2):
        if step:
            _, acc = hmc.run_exact(2, n_leapfrog=3, step=0.002, beta=5.0, seed=21)
            assert acc.any()
        hmc.train(forget=1.0, ridge=ridge)
        phi = hmc.ctx.sur_features(n).astype(np.float64)
        ef = hmc.feature_energies()
        en, f = hmc.ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, targets, _lib.POT_RW, 50.0, n)
        samples.append((phi, 2.0 * f, ef, en))   # exact -dE/dq (reference convention is half)
    w = hmc.output_weights().astype(np.float64)
    hmc.close()
    return w, samples, off


def test_energy_rows_enter_the_fit():
    """The output layer after two training states equals the oracle's fit
    over both (each state's exact forces plus, weighted, the exact energy
    change between them), checked through its predictions: the forces at the
    second state and the energy change.  A heavy energy weight moves the
    fitted energy change onto the exact one."""
    chains = _chains([35, 20], seed=4)
    ridge = 1e-3
    gap = {}
    for weight in (0.0, 1e5):
        w_dev, samples, off = _two_state_fit(chains, weight, ridge)
        for b in range(len(chains)):
            cols = slice(3 * off[b], 3 * off[b + 1])
            st = [(phi[:, cols], f[off[b]:off[b + 1]].reshape(-1), ef[b], en[b])
                  for phi, f, ef, en in samples]
            w_ref = OS.fit_over_states(st, 1.0, ridge, weight)
            de = st[1][2] - st[0][2]
            dE = st[1][3] - st[0][3]
            f_dev, f_ref = st[1][0].T @ w_dev[b], st[1][0].T @ w_ref
            print("\nweight %g chain %d: |dF|/|F| %.2e  dE %.4e fit %.4e ref %.4e"
                  % (weight, b, np.linalg.norm(f_dev - f_ref) / np.linalg.norm(f_ref), dE,
                     de @ w_dev[b], de @ w_ref))
            assert np.linalg.norm(f_dev - f_ref) <= 2e-2 * np.linalg.norm(f_ref)
            assert abs(de @ w_dev[b] - de @ w_ref) <= 2e-2 * max(abs(de @ w_ref), abs(dE))
            gap[weight, b] = abs(de @ w_ref - dE)
    for b in range(len(chains)):
        assert gap[1e5, b] < 0.5 * gap[0.0, b]
        assert gap[1e5, b] < 0.1 * abs(samples[1][3][b] - samples[0][3][b])
