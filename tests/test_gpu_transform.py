"""GPU parity of the F(Q)->G(r) transform (FP32 MFMA) and the FP64 potentials."""
import numpy as np
import pytest

from oracle import scatter as oracle
from pyiid_amd import _lib, structures
from pyiid_amd.formfactors import element_table

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("qmin", [0.0, 0.7])
def test_pdf_matches_fft_path(ctx, qmin):
    exp, pq = oracle.make_experiment({"qmin": qmin})
    nq = oracle.qmax_bin(exp["qmax"], pq)
    r = oracle.r_grid(exp)
    rs = np.random.RandomState(1)
    f = rs.randn(nq).astype(np.float32)
    ctx.set_transform(exp["rstep"], pq, r, qmin, nq)
    got = ctx.pdf(f, len(r))
    want = oracle.get_pdf_at_qmin(f, exp["rstep"], pq, r, qmin)
    assert rel(got, want) < 1e-5


@pytest.mark.parametrize("n_atoms,dtype", [(37, np.float64), (342, np.float64),
                                           (1111, np.float64), (1111, np.float32)])
def test_grad_pdf_rows(ctx, n_atoms, dtype):
    # 111 rows take the 64 x 64 tile product; 1,026 and 3,333 rows (at least
    # 1,024, not multiples of the tile) the 128 x 128 XCD-mapped one
    exp, pq = oracle.make_experiment({"qmin": 0.3, "rmax": 20.0})
    nq = oracle.qmax_bin(exp["qmax"], pq)
    r = oracle.r_grid(exp)
    rs = np.random.RandomState(2)
    g = rs.randn(n_atoms, 3, nq).astype(dtype).astype(np.float64)
    ctx.set_transform(exp["rstep"], pq, r, exp["qmin"], nq)
    got = ctx.grad_pdf(g.astype(dtype), len(r))
    want = oracle.grad_pdf(g.copy(), exp["rstep"], pq, r, exp["qmin"])
    assert got.shape == want.shape
    assert rel(got, want) < 1e-5


def test_potentials_known_answers(ctx):
    x = np.arange(0, 2 * np.pi, .1)
    rw, _ = ctx.potential(_lib.POT_RW, np.sin(x), np.cos(x))
    assert abs(rw - 1) < 1e-12
    rw, _ = ctx.potential(_lib.POT_RW, np.sin(x), np.sin(x))
    assert abs(rw) < 1e-12
    chi, _ = ctx.potential(_lib.POT_CHI_SQ, np.sin(x), np.cos(x))
    assert abs(chi - 63.01399) < 1e-4


@pytest.mark.parametrize("potential", [_lib.POT_RW, _lib.POT_CHI_SQ])
def test_potential_grad(ctx, potential):
    rs = np.random.RandomState(4)
    gobs = rs.randn(400)
    gcalc = gobs * 0.7 + 0.2 * rs.randn(400)
    gp = rs.randn(11, 3, 400)
    got = ctx.potential_grad(potential, gp, gobs, gcalc)
    if potential == _lib.POT_RW:
        want = oracle.wrap_grad_rw(gp, gcalc, gobs)
        v, s = oracle.get_rw(gobs, gcalc)
    else:
        want = oracle.wrap_grad_chi_sq(gp, gcalc, gobs)
        v, s = oracle.get_chi_sq(gobs, gcalc)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12)
    v2, s2 = ctx.potential(potential, gobs, gcalc)
    assert abs(v2 - v) < 1e-12 * max(1, abs(v)) and abs(s2 - s) < 1e-12


@pytest.mark.parametrize("potential", [_lib.POT_RW, _lib.POT_CHI_SQ])
def test_pdf_energy_forces(ctx, potential):
    exp, pq = oracle.make_experiment({"qmin": 0.5})
    nq = oracle.qmax_bin(exp["qmax"], pq)
    r = oracle.r_grid(exp)
    base = structures.cuboctahedron(4.08)
    pos_t = base.astype(np.float32)
    pos = structures.jitter(base, 0.05, seed=11).astype(np.float32)
    table, index, _ = element_table(np.full(len(pos), 79), pq, nq)
    per_atom = table[index]
    gobs = oracle.get_pdf_at_qmin(oracle.fq(pos_t, per_atom, pq), exp["rstep"], pq, r, exp["qmin"])
    ctx.set_positions(pos)
    ctx.set_elements(index)
    ctx.set_scatter(_lib.SLOT_PDF, table)
    ctx.set_transform(exp["rstep"], pq, r, exp["qmin"], nq)
    e, f = ctx.pdf_energy_forces(_lib.SLOT_PDF, pq, gobs, potential, 1.0, len(pos))
    gcalc = oracle.get_pdf_at_qmin(oracle.fq(pos, per_atom, pq), exp["rstep"], pq, r, exp["qmin"])
    gfq = oracle.grad_fq(pos, per_atom, pq)
    gfq[:, :, :int(exp["qmin"] / pq)] = 0
    gpdf = oracle.grad_pdf(gfq, exp["rstep"], pq, r, exp["qmin"])
    if potential == _lib.POT_RW:
        ew = oracle.get_rw(gobs, gcalc)[0]
        fw = oracle.wrap_grad_rw(gpdf, gcalc, gobs)
        assert abs(e - ew) < 1e-6
    else:
        ew = oracle.get_chi_sq(gobs, gcalc)[0]
        fw = oracle.wrap_grad_chi_sq(gpdf, gcalc, gobs)
        assert abs(e - ew) < 1e-6 * max(1.0, abs(ew))
    assert rel(f, fw) < 1e-5
