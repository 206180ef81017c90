"""Batched energy/forces and the device-resident HMC against the oracle.

One device proposal is replayed on the CPU: momenta from the same
counter-based generator (splitmix64 + Box-Muller, restated below), forces
-dE/dq from the FP64 oracle chain positions -> F -> G -> Rw, velocity-Verlet
steps (oracle.scatter.leapfrog) and the Metropolis rule of
pyiid/sim/hmc.py:66-92.  Positions, energies and the decision must agree.
"""
import math

import numpy as np
import pytest

from oracle import scatter as S
from pyiid_amd import _lib, structures
from pyiid_amd.elasticscatter import ElasticScatter
from pyiid_amd.formfactors import MASSES, element_table
from pyiid_amd.sim import DeviceHMC

pytestmark = pytest.mark.gpu
M64 = (1 << 64) - 1


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def uniform_key(seed, counter, stream, idx):
    k = mix64((mix64(seed ^ mix64((counter * 4 + stream) & M64)) + idx) & M64)
    return ((k >> 11) + 1.0) * (1.0 / 9007199254740992.0)


def normal_key(seed, counter, idx):
    u1 = uniform_key(seed, counter, 0, idx)
    u2 = uniform_key(seed, counter, 1, idx)
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2)


EXP = {"qmin": 0.5, "rmax": 12.0}


def chain_energy_forces(pos, numbers, gobs, conv, pq, e):
    nq = S.qmax_bin(e["qmax"], pq)
    table, idx, _ = element_table(numbers, pq, nq)
    sp = table[idx]
    r = S.r_grid(e)
    g = S.get_pdf_at_qmin(S.fq(pos, sp, pq, cast_f32=False), e["rstep"], pq, r, e["qmin"])
    rw = S.get_rw(gobs, g)[0]
    gfq = S.grad_fq(pos, sp, pq, cast_f32=False)
    gfq[:, :, :int(e["qmin"] / pq)] = 0
    gp = S.grad_pdf(gfq, e["rstep"], pq, r, e["qmin"])
    ref_forces = S.wrap_grad_rw(gp, g, gobs) * conv     # reference convention (half)
    return conv * rw, ref_forces


def make_chains(sizes, seed=0):
    rs = np.random.RandomState(seed)
    out = []
    for k, n in enumerate(sizes):
        base = structures.fcc_nearest(n)
        z = np.full(n, 79) if k % 2 == 0 else structures.random_alloy(n, 78, 28, seed=k)
        out.append((base, structures.jitter(base, 0.1, seed=seed + 10 + k), z))
    return out


def test_batched_energy_forces():
    e, pq = S.make_experiment(dict(EXP))
    nq = S.qmax_bin(e["qmax"], pq)
    r = S.r_grid(e)
    chains = make_chains([13, 40, 2, 19])
    ctx = _lib.Context(0)
    ctx.set_batch([c[1] for c in chains])
    numbers = np.concatenate([c[2] for c in chains])
    table, index, _ = element_table(numbers, pq, nq)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_PDF, table)
    ctx.set_transform(e["rstep"], pq, r, e["qmin"], nq)
    targets, ref = [], []
    for base, pos, z in chains:
        t_idx = np.searchsorted(np.unique(numbers), z)
        gobs = S.get_pdf_at_qmin(S.fq(base.astype(np.float32), table[t_idx], pq), e["rstep"], pq, r,
                                 e["qmin"])
        targets.append(gobs)
    targets = np.array(targets)
    en, forces = ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, targets, _lib.POT_RW, 2.0,
                                       len(numbers))
    fq = ctx.fq(_lib.SLOT_PDF, pq, nq, n_chain=len(chains))
    off = np.cumsum([0] + [len(c[1]) for c in chains])
    for b, (base, pos, z) in enumerate(chains):
        t_idx = np.searchsorted(np.unique(numbers), z)
        np.testing.assert_allclose(fq[b], S.fq(pos, table[t_idx], pq), rtol=1e-5,
                                   atol=1e-5 * max(1e-12, np.max(np.abs(S.fq(pos, table[t_idx], pq)))))
        ew, fw = chain_energy_forces(pos.astype(np.float32).astype(np.float64), z, targets[b], 2.0,
                                     pq, e)
        assert abs(en[b] - ew) < 2e-6
        fb = forces[off[b]:off[b + 1]]
        assert np.max(np.abs(fb - fw)) <= 1e-5 * max(1e-12, np.max(np.abs(fw))) + 1e-12
    ctx.close()


def replay(chains, targets, masses, conv, beta, eps, n_leap, seed, pq, e):
    """CPU replay of proposal 0 of every chain; returns positions, energies, accept."""
    out_pos, out_e, out_a = [], [], []
    off = 0
    for b, (base, pos, z) in enumerate(chains):
        n = len(pos)
        m = masses[off:off + n]
        x = pos.astype(np.float32).astype(np.float64)
        e0, f0r = chain_energy_forces(x, z, targets[b], conv, pq, e)
        p = np.array([[math.sqrt(m[i] / beta) * normal_key(seed, 0, 3 * (off + i) + c)
                       for c in range(3)] for i in range(n)])
        k0 = np.sum(p * p / (2 * m[:, None]))
        forces = lambda y: 2.0 * chain_energy_forces(y, z, targets[b], conv, pq, e)[1]  # noqa
        xs, ps = x.copy(), p.copy()
        # leapfrog over n_leap steps = n_leap velocity-Verlet steps
        for _ in range(n_leap):
            xs, ps = S.leapfrog(xs, ps, m, forces, eps)
        e1, _ = chain_energy_forces(xs, z, targets[b], conv, pq, e)
        k1 = np.sum(ps * ps / (2 * m[:, None]))
        u = uniform_key(seed, 0, 2, b)
        acc = u < math.exp(beta * ((e0 + k0) - (e1 + k1)))
        out_pos.append(xs if acc else x)
        out_e.append(e1 if acc else e0)
        out_a.append(int(acc))
        off += n
    return out_pos, out_e, out_a


@pytest.mark.parametrize("use_graph", [False, True])
def test_one_proposal_matches_cpu_replay(use_graph):
    e, pq = S.make_experiment(dict(EXP))
    scat = ElasticScatter(dict(EXP), seed=0)
    chains = make_chains([13, 19], seed=3)
    r = scat.get_r()
    nq = S.qmax_bin(e["qmax"], pq)
    numbers = np.concatenate([c[2] for c in chains])
    table, index, elems = element_table(numbers, pq, nq)
    targets = []
    for base, pos, z in chains:
        t_idx = np.searchsorted(elems, z)
        targets.append(S.get_pdf_at_qmin(S.fq(base.astype(np.float32), table[t_idx], pq),
                                         e["rstep"], pq, r, e["qmin"]))
    targets = np.array(targets)
    masses = np.array([MASSES[int(zz)] for zz in numbers])
    hmc = DeviceHMC(scat, [(c[1], c[2]) for c in chains], targets, conv=50.0)
    beta, eps, n_leap, seed = 40.0, 0.05, 3, 12345
    energies, accepted = hmc.run(1, n_leapfrog=n_leap, step=eps, beta=beta, seed=seed,
                                 use_graph=use_graph)
    got = hmc.positions()
    hmc.close()
    want_pos, want_e, want_a = replay(chains, targets, masses, 50.0, beta, eps, n_leap, seed, pq, e)
    for b in range(len(chains)):
        assert accepted[0, b] == want_a[b]
        assert abs(energies[0, b] - want_e[b]) < 1e-4 * max(1.0, abs(want_e[b]))
        np.testing.assert_allclose(got[b], want_pos[b], atol=2e-5)


def test_graph_replay_equals_direct_and_lowers_rw(fixed_shapes):
    e, pq = S.make_experiment(dict(EXP))
    scat = ElasticScatter(dict(EXP), seed=0)
    chains = make_chains([30, 55, 20], seed=4)
    r = scat.get_r()
    nq = S.qmax_bin(e["qmax"], pq)
    numbers = np.concatenate([c[2] for c in chains])
    table, index, elems = element_table(numbers, pq, nq)
    targets = np.array([S.get_pdf_at_qmin(S.fq(b.astype(np.float32), table[np.searchsorted(elems, z)],
                                                pq), e["rstep"], pq, r, e["qmin"])
                        for b, _, z in chains])
    runs = []
    for use_graph in (False, True):
        hmc = DeviceHMC(scat, [(c[1], c[2]) for c in chains], targets, conv=200.0, masses="unit")
        runs.append(hmc.run(25, n_leapfrog=5, step=0.01, beta=5.0, seed=7, use_graph=use_graph))
        hmc.close()
    (e1, a1), (e2, a2) = runs
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_allclose(e1, e2, rtol=1e-12)
    assert np.all(e1[-1] < e1[0])         # refinement towards the target lowers Rw
    assert 0 < a1.mean() <= 1


@pytest.mark.parametrize("tiny", [1, 0])
def test_batch_with_degenerate_chain(tiny):
    """A chain of one atom (no pairs) or of none in the middle of a batch:
    its G is zero, so Rw is 0/0 = NaN as in the reference's get_rw
    (master_kernel.py:206-236); the other chains' energies and forces are
    those of the batch without it (to the rounding of launch shapes the
    autotuner may pick differently for the two batches)."""
    e, pq = S.make_experiment(dict(EXP))
    nq = S.qmax_bin(e["qmax"], pq)
    r = S.r_grid(e)
    a = structures.jitter(structures.fcc_nearest(13), 0.1, seed=1).astype(np.float32)
    b = structures.jitter(structures.fcc_nearest(19), 0.1, seed=2).astype(np.float32)
    t = np.zeros((tiny, 3), np.float32)
    table, _, _ = element_table(np.array([79]), pq, nq)
    # targets from rattled copies: Rw well away from 0, so that the forces are
    # not dominated by rounding
    gobs = np.stack([S.get_pdf_at_qmin(S.fq(structures.jitter(p, 0.08, seed=7).astype(np.float32),
                                            table[np.zeros(len(p), int)], pq),
                                       e["rstep"], pq, r, e["qmin"]) for p in (a, b)])

    def run(chains, targets):
        ctx = _lib.Context(0)
        ctx.set_batch(chains)
        n = sum(len(c) for c in chains)
        ctx.set_elements(np.zeros(n, np.int32))
        ctx.set_scatter(_lib.SLOT_PDF, table)
        ctx.set_transform(e["rstep"], pq, r, e["qmin"], nq)
        en, forces = ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, targets, _lib.POT_RW, 2.0, n)
        ctx.close()
        return en, forces

    en2, f2 = run([a, b], gobs)
    en3, f3 = run([a, t, b], np.stack([gobs[0], gobs[0], gobs[1]]))
    assert np.isnan(en3[1])
    assert np.all(np.isfinite(en3[[0, 2]]))
    assert np.all(en2 > 0.02)
    # energy = 2 Rw: the Rw bar of 1e-6
    np.testing.assert_allclose(en3[[0, 2]], en2, rtol=0, atol=2e-6)
    scale = np.max(np.abs(f2))
    np.testing.assert_allclose(f3[:13], f2[:13], rtol=0, atol=1e-6 * scale)
    np.testing.assert_allclose(f3[13 + tiny:], f2[13:], rtol=0, atol=1e-6 * scale)
