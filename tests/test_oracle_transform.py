"""Pins of the oracle's F(Q)->G(r) transform, Rw / chi^2, their gradients and
the leapfrog step.

The dense matrix T is pinned by the independent FFT restatement of the
reference (numpy.fft.ifft of the odd extension) and by a continuous closed
form; the potentials by the reference's own known answers
(pyiid/tests/test_master/test_master_kernel.py:19-48) and by finite
differences of the full positions -> F -> G -> Rw chain.
"""
import math

import numpy as np
import pytest

from oracle import scatter as S


@pytest.mark.parametrize("exp_dict", [
    None,
    {"qmin": 0.9},
    {"qmin": 1.2, "qmax": 21.3, "rmax": 33.0, "rstep": 0.013},
    {"qmin": 0.4, "qmax": 23.0, "rmin": 1.7, "rmax": 45.0, "sampling": "ns"},
])
def test_transform_matrix_equals_fft_path(exp_dict):
    exp, pq = S.make_experiment(exp_dict)
    nq = S.qmax_bin(exp["qmax"], pq)
    r = S.r_grid(exp)
    rs = np.random.RandomState(0)
    f = rs.randn(nq)
    want = S.get_pdf_at_qmin(f, exp["rstep"], pq, r, exp["qmin"])
    t = S.transform_matrix(exp["rstep"], pq, r, exp["qmin"], nq)
    np.testing.assert_allclose(t @ f, want, rtol=1e-10, atol=1e-12 * np.max(np.abs(want)))
    # rows of a gradient go through the same transform
    g = rs.randn(2, 3, nq)
    np.testing.assert_allclose(np.einsum("rq,acq->acr", t, g),
                               S.grad_pdf(g.copy(), exp["rstep"], pq, r, exp["qmin"]),
                               rtol=1e-10, atol=1e-11)


def test_transform_grid_numbers():
    # default experiment: pdf_qbin = pi/(40 + 12 pi/25), nfromdr = 4151 -> npad2 = 16384
    exp, pq = S.make_experiment()
    assert abs(pq - 0.07568650242000183) < 1e-15
    npad2, drpad, q_zero = S.transform_grid(exp["rstep"], pq, exp["qmin"], 330)
    assert npad2 == 16384 and q_zero == 0
    assert abs(drpad - math.pi / (16384 * pq)) < 1e-18


def test_textbook_sine_sum_differs():
    # G != (2/pi) dq sum F sin(Q r) at the 1e-4 level: the interpolated grid matters
    exp, pq = S.make_experiment()
    nq = S.qmax_bin(exp["qmax"], pq)
    r = S.r_grid(exp)
    f = np.random.RandomState(3).randn(nq)
    Q = np.arange(nq) * pq
    textbook = (2 / np.pi) * pq * np.sin(np.outer(r, Q)) @ f
    g = S.get_pdf_at_qmin(f, exp["rstep"], pq, r, 0.0)
    assert np.max(np.abs(g - textbook)) / np.max(np.abs(g)) > 1e-4


def test_transform_continuous_closed_form():
    # F(Q) = Q exp(-a Q^2): (2/pi) int_0^inf F sin(Q r) dQ = r exp(-r^2/4a) / (2 sqrt(pi) a^1.5)
    a = 0.05
    exp, pq = S.make_experiment({"qmax": 40.0})
    nq = S.qmax_bin(exp["qmax"], pq)
    Q = np.arange(nq) * pq
    r = S.r_grid(exp)
    g = S.get_pdf_at_qmin(Q * np.exp(-a * Q * Q), exp["rstep"], pq, r, 0.0)
    want = r * np.exp(-r * r / (4 * a)) / (2 * math.sqrt(math.pi) * a ** 1.5)
    assert np.max(np.abs(g - want)) / np.max(np.abs(want)) < 2e-3


def test_potential_known_answers():
    # pyiid/tests/test_master/test_master_kernel.py:19-48
    x = np.arange(0, 2 * np.pi, .1)
    assert abs(S.get_rw(np.sin(x), np.cos(x))[0] - 1) < 1e-12
    assert abs(S.get_rw(np.sin(x), np.sin(x))[0]) < 1e-15
    chi, scale = S.get_chi_sq(np.sin(x), np.cos(x))
    assert abs(chi - 63.01399) < 1e-4 and scale == 1
    assert abs(S.get_chi_sq(np.sin(x), np.sin(x))[0]) < 1e-15


def chain(q, f, exp, pq, r, gobs, potential):
    fq = S.fq(q, f, pq, cast_f32=False)
    g = S.get_pdf_at_qmin(fq, exp["rstep"], pq, r, exp["qmin"])
    return (S.get_rw if potential == "rw" else S.get_chi_sq)(gobs, g)[0], g


@pytest.mark.parametrize("potential", ["rw", "chi_sq"])
def test_potential_gradient_finite_difference(potential):
    exp, pq = S.make_experiment({"qmin": 0.6, "rmax": 12.0})
    nq = S.qmax_bin(exp["qmax"], pq)
    r = S.r_grid(exp)
    rs = np.random.RandomState(5)
    q0 = rs.random_sample((6, 3)) * 5
    q = q0 + rs.normal(0, 0.2, q0.shape)
    f = np.full((6, nq), 79.0)
    gobs = S.get_pdf_at_qmin(S.fq(q0, f, pq, cast_f32=False), exp["rstep"], pq, r, exp["qmin"])
    _, g = chain(q, f, exp, pq, r, gobs, potential)
    gfq = S.grad_fq(q, f, pq, cast_f32=False)
    gfq[:, :, :int(exp["qmin"] / pq)] = 0
    gpdf = S.grad_pdf(gfq, exp["rstep"], pq, r, exp["qmin"])
    wrap = S.wrap_grad_rw if potential == "rw" else S.wrap_grad_chi_sq
    grad = wrap(gpdf, g, gobs)
    h = 1e-6
    for a in range(6):
        for c in range(3):
            qp, qm = q.copy(), q.copy()
            qp[a, c] += h
            qm[a, c] -= h
            d = (chain(qp, f, exp, pq, r, gobs, potential)[0] -
                 chain(qm, f, exp, pq, r, gobs, potential)[0]) / (2 * h)
            # the reference gradient array is -1/2 dG/dq, so grad = -1/2 dP/dq
            np.testing.assert_allclose(-2.0 * grad[a, c], d, rtol=2e-5,
                                       atol=1e-7 * np.max(np.abs(grad)) + 1e-12)


def test_rw_nonpositive_scale():
    x = np.arange(0, 2 * np.pi, .1)
    assert S.get_rw(np.sin(x), -np.sin(x)) == (1, 1)


def harmonic(x):
    return -2.0 * x


def test_leapfrog_invariants():
    # pyiid/tests/test_sim/test_leapfrog.py:25-73 (null force, momentum, reversibility)
    x = np.random.RandomState(0).randn(4, 3)
    m = np.full(4, 2.0)
    zero = np.zeros_like(x)
    x1, p1 = S.leapfrog(x, zero, m, lambda y: np.zeros_like(y), 1.0)
    np.testing.assert_array_equal(x1, x)
    p = np.ones_like(x)
    x1, _ = S.leapfrog(x, p, m, lambda y: np.zeros_like(y), 1.0)
    np.testing.assert_allclose(x1, x + p / m[:, None])
    xa, pa = S.leapfrog(x, p, m, harmonic, 0.1)
    xb, pb = S.leapfrog(xa, -pa, m, harmonic, 0.1)
    np.testing.assert_allclose(xb, x, atol=1e-12)
    np.testing.assert_allclose(-pb, p, atol=1e-12)


@pytest.mark.parametrize("qmin,rmin,rstep", [(0.0, 0.0, 0.01), (0.7, 1.2, 0.013)])
def test_batched_rows_equal_row_by_row(qmin, rmin, rstep):
    """The batched transform (rows through one FFT call, two threads, chunks
    that do not divide the row count) against the row-by-row restatement."""
    rs = np.random.RandomState(4)
    e, pq = S.make_experiment({"qmin": qmin, "rmin": rmin, "rstep": rstep, "rmax": 30.0})
    nq = S.qmax_bin(e["qmax"], pq)
    r = S.r_grid(e)
    g = rs.normal(size=(7, 3, nq))
    want = S.grad_pdf(g.copy(), e["rstep"], pq, r, e["qmin"])
    got = S.grad_pdf_rows(g, e["rstep"], pq, r, e["qmin"], workers=2)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.max(np.abs(want)))
    rows = S.get_pdf_at_qmin_rows(g.reshape(21, nq), e["rstep"], pq, r, e["qmin"], chunk=5)
    np.testing.assert_allclose(rows.reshape(7, 3, -1), want, rtol=0,
                               atol=1e-12 * np.max(np.abs(want)))
