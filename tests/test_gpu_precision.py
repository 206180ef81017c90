"""Precision at scale: a 20,000-atom Pt/Ni random alloy (50/50, seeded), Q
0 ... 29.98 A^-1 in steps of 0.02 (1,500 bins), Cromer-Mann form factors per
element.  The device F(Q) (FP32 pair terms, FP64 cross-tile reduction) and
gradient are compared with the FP64 oracle on a strided subset of Q bins and
atoms (each bin and each gradient row is an independent sum, so a subset is an
exact check of those entries), computed by the oracle's threaded FP64 C
path (pinned against the numpy oracle in test_oracle_native.py).  Bar: 1e-5
of max|x*|."""
import json
import os

import numpy as np
import pytest

from oracle import native
from pyiid_amd import _lib, structures
from pyiid_amd.formfactors import element_table

pytestmark = pytest.mark.gpu

N, QBIN, NQ = 20000, 0.02, 1500


def test_alloy_20k_f_and_grad_vs_fp64_oracle():
    base = structures.fcc_nearest(N, a=3.724)
    pos = structures.jitter(base, 0.05, seed=0).astype(np.float32)
    numbers = structures.random_alloy(N, 78, 28, seed=0)
    assert np.sum(numbers == 78) == N // 2
    table, index, _ = element_table(numbers, QBIN, NQ)
    ctx = _lib.Context(0)
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_FQ, table)
    fq = np.empty(NQ, np.float32)
    grad = ctx.grad_fq(_lib.SLOT_FQ, QBIN, N, NQ, f64=True, fq_out=fq)
    fq_only = ctx.fq(_lib.SLOT_FQ, QBIN, NQ)
    ctx.close()
    q_idx = np.array([25, 400, 925, 1499])          # 0.5, 8, 18.5, 29.98 A^-1
    atoms = np.linspace(0, N - 1, 24).astype(int)
    f_ref, g_ref = native.fp64_subset(pos, index, table, QBIN, q_idx, atoms,
                                      nthreads=min(16, os.cpu_count() or 1))
    err_f = float(np.max(np.abs(fq[q_idx] - f_ref)) / np.max(np.abs(f_ref)))
    err_f_only = float(np.max(np.abs(fq_only[q_idx] - f_ref)) / np.max(np.abs(f_ref)))
    g_dev = grad[atoms][:, :, q_idx]
    err_g = float(np.max(np.abs(g_dev - g_ref)) / np.max(np.abs(g_ref)))
    print(json.dumps({"precision_20k_alloy": {"fq_fused_rel": err_f, "fq_kernel_rel": err_f_only,
                                              "grad_rel": err_g, "q_bins": q_idx.tolist(),
                                              "atoms_checked": len(atoms)}}))
    assert err_f < 1e-5 and err_f_only < 1e-5 and err_g < 1e-5


def test_gradient_rows_beyond_one_grid_column():
    """22,500 atoms: 67,500 gradient rows (atom, component), more than one
    launch column of the split reduction holds (65,535), so its row loop
    wraps; the gradient of atoms on both sides of that boundary against the
    FP64 oracle."""
    n, qbin, nq = 22500, 0.1, 200
    pos = structures.jitter(structures.fcc_nearest(n, a=4.08), 0.05, seed=3).astype(np.float32)
    table, index, _ = element_table(np.full(n, 79), qbin, nq)
    ctx = _lib.Context(0)
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_FQ, table)
    grad = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, f64=True)
    ctx.close()
    q_idx = np.array([5, 60, 199])
    atoms = np.array([0, 7, 21844, 21845, 21846, 22000, n - 1])
    _, g_ref = native.fp64_subset(pos, index, table, qbin, q_idx, atoms,
                                  nthreads=min(16, os.cpu_count() or 1))
    g_dev = grad[atoms][:, :, q_idx]
    assert float(np.max(np.abs(g_dev - g_ref)) / np.max(np.abs(g_ref))) < 1e-5


@pytest.mark.parametrize("case", ["sphere", "alloy"])
def test_compensated_sums_match_oracle(case):
    """Compensated FP32 sums (pyiid_set_precision): F from both kernels and
    the gradient against the FP64 oracle, single-element and weighted, on
    bins and atoms across the grid."""
    if case == "sphere":
        pos = structures.jitter(structures.fcc_sphere(16.0), 0.05, seed=4).astype(np.float32)
        numbers = np.full(len(pos), 79)
        qbin, nq = 0.05, 500
    else:
        pos = structures.jitter(structures.fcc_nearest(3000, a=3.724), 0.05,
                                seed=5).astype(np.float32)
        numbers = structures.random_alloy(len(pos), 78, 28, seed=6)
        qbin, nq = 0.02, 1500
    n = len(pos)
    table, index, _ = element_table(numbers, qbin, nq)
    ctx = _lib.Context(0)
    ctx.set_precision(True)
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_FQ, table)
    fq = np.empty(nq, np.float32)
    grad = ctx.grad_fq(_lib.SLOT_FQ, qbin, n, nq, f64=True, fq_out=fq)
    fq_only = ctx.fq(_lib.SLOT_FQ, qbin, nq)
    ctx.close()
    q_idx = np.unique(np.linspace(1, nq - 1, 12).astype(int))
    atoms = np.linspace(0, n - 1, 16).astype(int)
    f_ref, g_ref = native.fp64_subset(pos, index, table, qbin, q_idx, atoms,
                                      nthreads=min(16, os.cpu_count() or 1))
    rel = lambda a, b: float(np.max(np.abs(a - b)) / np.max(np.abs(b)))  # noqa: E731
    errs = (rel(fq[q_idx], f_ref), rel(fq_only[q_idx], f_ref),
            rel(grad[atoms][:, :, q_idx], g_ref))
    print("compensated %s N=%d: F fused %.2e, F-only %.2e, grad %.2e" % ((case, n) + errs))
    assert max(errs) < 1e-5
