"""The Debye and sampler workloads at their full sizes on the device.

Each is checked against the FP64 oracle where that finishes in seconds, and
otherwise by properties that hold at any size.

* Debye F(Q) + gradient: the 6,699-atom Au sphere of radius 30 A (a = 4.08 A,
  seeded sigma 0.05 A jitter, float32 positions) on 1,250 bins of
  0.02 A^-1, through pyiid_grad_fq.  F and the gradient are compared with the
  threaded FP64 oracle on strided bins and atoms, at 1e-5 of max|x*|.  Each
  bin and each gradient row is an independent sum, so a subset checks those
  entries exactly.
* Batched chains: 256 chains of 100-3,000 Au atoms, the batched sampler
  workload.
  - Energies (conv Rw) and forces of the smallest, the median and the
    largest chain are compared with the oracle.  The bars are 1e-6 of Rw and
    1e-5 of max|F|.
  - Over three proposals of a captured proposal graph (the second and third
    are replays), each recorded energy equals the energy recomputed at the
    state the sampler holds, and rejected chains keep their positions bit
    for bit.
* Learned integrator: 64 chains of 5,000 Au atoms, run through a sequence
  with training steps.  The checks:
  - the last energies equal the exact recompute;
  - energies change only on accepted proposals;
  - a chain's guard rejections do not exceed its rejections;
  - the device energy of one chain at float32 positions matches the oracle.
"""
import os

import numpy as np
import pytest

from oracle import native
from oracle import scatter as S
from pyiid_amd import _lib, structures, workloads
from pyiid_amd.elasticscatter import ElasticScatter
from pyiid_amd.formfactors import element_table
from pyiid_amd.sim import DeviceHMC, SurrogateHMC

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def oracle_energy_forces(pos, frow, gobs, conv, pq, e, forces=True):
    """conv Rw and the reference-convention forces of one single-element
    chain from the FP64 oracle: F on every bin, G by the oracle's transform,
    Rw, and the gradient chain rule of the reference's Calc1D."""
    n = len(pos)
    nq = len(frow)
    rows = np.arange(n) if forces else None
    fq, gfq = native.fp64_subset(pos, np.zeros(n, np.int32), frow[None, :], pq, np.arange(nq),
                                 rows, nthreads=NT)
    r = S.r_grid(e)
    g = S.get_pdf_at_qmin(fq, e["rstep"], pq, r, e["qmin"])
    rw = S.get_rw(gobs, g)[0]
    if not forces:
        return conv * rw, None
    gfq[:, :, :int(e["qmin"] / pq)] = 0
    gp = S.grad_pdf(gfq, e["rstep"], pq, r, e["qmin"])
    return conv * rw, S.wrap_grad_rw(gp, g, gobs) * conv


def test_sphere_6699_f_and_grad_vs_fp64_oracle():
    pos = structures.jitter(structures.fcc_sphere(30.0, 4.08), 0.05, 0).astype(np.float32)
    n, qbin, nq = len(pos), 0.02, 1250
    assert n == 6699
    table, index, _ = element_table(np.full(n, 79), qbin, nq)
    ctx = _lib.Context(0)
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_FQ, table)
    fq = np.empty(nq, np.float32)
    grad = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, fq_out=fq)
    ctx.close()
    q_idx = np.arange(2, nq, 83)
    atoms = np.linspace(0, n - 1, 32).astype(int)
    print("oracle: %d bins x %d rows" % (len(q_idx), len(atoms)), flush=True)
    f_ref, g_ref = native.fp64_subset(pos, index, table, qbin, q_idx, atoms, nthreads=NT)
    err_f, err_g = rel(fq[q_idx], f_ref), rel(grad[atoms][:, :, q_idx], g_ref)
    print("au6699_q1250: F %.2e, grad %.2e" % (err_f, err_g))
    assert err_f < 1e-5 and err_g < 1e-5


@pytest.fixture(scope="module")
def batched():
    starts, targs, numbers = workloads.build("batched", 0)
    scat = ElasticScatter(dict(workloads.EXP), seed=0)
    gt = workloads.target_pdfs(scat, targs, numbers)
    e, pq = S.make_experiment(dict(workloads.EXP))
    assert abs(pq - scat.pdf_qbin) < 1e-15
    return starts, numbers, scat, gt, e, pq


def test_batched_256_energies_and_forces_vs_oracle(batched):
    starts, numbers, scat, gt, e, pq = batched
    assert len(starts) == 256
    conv = 100.0
    nq = S.qmax_bin(e["qmax"], pq)
    r = S.r_grid(e)
    ctx = _lib.Context(0)
    ctx.set_batch(starts)
    table, index, _ = element_table(np.concatenate(numbers), pq, nq)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_PDF, table)
    ctx.set_transform(e["rstep"], pq, r, e["qmin"], nq)
    n = int(sum(len(s) for s in starts))
    en, forces = ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, gt, _lib.POT_RW, conv, n)
    ctx.close()
    assert np.all(np.isfinite(en)) and np.all(np.isfinite(forces))
    sizes = np.array([len(s) for s in starts])
    off = np.concatenate([[0], np.cumsum(sizes)])
    order = np.argsort(sizes, kind="stable")
    for b in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        pos = starts[b].astype(np.float32).astype(np.float64)
        want_f = sizes[b] <= 2000
        print("chain %d: %d atoms (oracle forces: %s)" % (b, sizes[b], want_f), flush=True)
        ew, fw = oracle_energy_forces(pos, table[0], gt[b], conv, pq, e, forces=want_f)
        print("  energy %.9f vs %.9f" % (en[b], ew), flush=True)
        assert abs(en[b] - ew) <= 1e-6 * conv
        if want_f:
            fb = forces[off[b]:off[b + 1]]
            assert rel(fb, fw) < 1e-5


def test_batched_256_graph_proposals_keep_their_books(batched):
    starts, numbers, scat, gt, e, pq = batched
    conv = 100.0
    hmc = DeviceHMC(scat, list(zip(starts, numbers)), gt, conv=conv, masses="unit")
    n = int(hmc.offsets[-1])
    n_rejected = n_accepted = 0
    try:
        for k in range(3):
            before = hmc.positions()
            # beta * (H1 - H0) does not depend on beta (momenta ~ N(0, m / beta));
            # 25 times the workload's step makes the larger chains reject
            en, acc = hmc.run(1, n_leapfrog=10, step=5e-2, beta=50.0, seed=5, use_graph=True)
            after = hmc.positions()
            now, _ = hmc.ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, gt, _lib.POT_RW, conv, n,
                                               forces=False)
            print("proposal %d: accepted %d of 256" % (k, int(acc.sum())), flush=True)
            np.testing.assert_allclose(en[0], now, rtol=1e-9, atol=0)
            for b in np.nonzero(acc[0] == 0)[0]:
                np.testing.assert_array_equal(after[b], before[b])
            n_rejected += int((acc[0] == 0).sum())
            n_accepted += int(acc[0].sum())
    finally:
        hmc.close()
    assert n_rejected > 0 and n_accepted > 0


def test_large_64x5000_surrogate_sequence():
    starts, targs, numbers = workloads.build("large", 0)
    scat = ElasticScatter(dict(workloads.EXP), seed=0)
    gt = workloads.target_pdfs(scat, targs, numbers)
    e, pq = S.make_experiment(dict(workloads.EXP))
    conv, n_prop = 100.0, 6
    sur = SurrogateHMC(scat, list(zip(starts, numbers)), gt, conv=conv, masses="unit", net_seed=0)
    n = int(sur.offsets[-1])
    assert n == 64 * 5000
    try:
        e0, _ = sur.ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, gt, _lib.POT_RW, conv, n,
                                          forces=False)
        out = sur.run(n_prop, n_leapfrog=10, step=2e-3, beta=50.0, seed=0, train_every=2,
                      forget=0.5, ridge=1e-6, min_r2=0.5)
        en, acc = out["energies"], out["accepted"]
        print("accepted per proposal", acc.sum(axis=1).tolist(), "exact-mode",
              out["exact_mode"].tolist(), flush=True)
        prev = e0
        for p in range(n_prop):
            rej = acc[p] == 0
            if p == 0:
                np.testing.assert_allclose(en[p][rej], prev[rej], rtol=1e-9, atol=0)
            else:
                np.testing.assert_array_equal(en[p][rej], prev[rej])
            prev = en[p]
        now, _ = sur.ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, gt, _lib.POT_RW, conv, n,
                                           forces=False)
        np.testing.assert_allclose(en[-1], now, rtol=1e-9, atol=0)
        assert np.all(out["violations"] >= 0)
        assert np.all(out["violations"] <= (acc == 0).sum(axis=0))
        assert np.all(np.diff(out["times_ms"]) > 0)
        assert acc.mean() > 0
        final = sur.positions()
    finally:
        sur.close()
    # one chain at float32 positions: device energy against the oracle
    b = 17
    pos = final[b].astype(np.float32)
    nq = S.qmax_bin(e["qmax"], pq)
    table, index, _ = element_table(numbers[b], pq, nq)
    ctx = _lib.Context(0)
    ctx.set_batch([pos])
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_PDF, table)
    ctx.set_transform(e["rstep"], pq, S.r_grid(e), e["qmin"], nq)
    ed, _ = ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, gt[b:b + 1], _lib.POT_RW, conv, len(pos),
                                  forces=False)
    ctx.close()
    print("oracle energy of chain %d" % b, flush=True)
    ew, _ = oracle_energy_forces(pos.astype(np.float64), table[0], gt[b], conv, pq, e,
                                 forces=False)
    print("  energy %.9f vs %.9f" % (ed[0], ew))
    assert abs(ed[0] - ew) <= 1e-6 * conv
