"""The recurrences the Debye kernels run along Q, emulated in float32 on the
CPU (each FMA formed in float64 and rounded once to float32), against the
float64 sequence sin(theta0 + m dtheta).

Pins the numerical choices of the kernels:
* the three-term recurrence y' = k y - y_prev with k = 2 cos(dtheta) rounded
  to float32 drifts in phase by ~m ulp / |sin dtheta|, which is large near
  every multiple of pi (not only near 0);
* with k split into float32 hi + lo and 8 lo applied every eighth step (the
  kernels' form) the error stays bounded for every step, 0 and pi included;
* the Reinsch difference form y' = y + D', D'' = D' - eps y' on the sequence
  (-1)^(k m) y with eps = 4 sin^2(d/2) (the F-only kernel's form for steps
  near a multiple of pi) is accurate there too.
Errors are in units of the float32 ulp of 1 (2^-23) over 32, 64 and 112 steps."""
import numpy as np
import pytest

F32 = np.float32
ULP = 2.0 ** -23


def fma(a, b, c):
    return F32(np.float64(a) * np.float64(b) + np.float64(c))


def truth(theta0, dth, n):
    return np.sin(theta0 + dth * np.arange(n))


def plain(theta0, dth, n):
    k = F32(2.0 * np.cos(dth))
    y0, y1 = F32(np.sin(theta0)), F32(np.sin(theta0 + dth))
    out = [y0, y1]
    for _ in range(n - 2):
        y0, y1 = y1, fma(k, y1, -y0)
        out.append(y1)
    return np.array(out, np.float64)


def two_float(theta0, dth, n):
    k = 2.0 * np.cos(dth)
    hi = F32(k)
    lo8 = F32(8.0 * F32(k - np.float64(hi)))
    # the kernels start from (y_-1, y_0); here from (y_0, y_1) with y_1 exact
    yp, y = F32(np.sin(theta0)), F32(np.sin(theta0 + dth))
    out = [yp, y]
    for m in range(1, n - 1):
        yn = fma(hi, y, -yp)
        if m % 8 == 7:
            yn = fma(lo8, y, yn)
        yp, y = y, yn
        out.append(y)
    return np.array(out, np.float64)


def reinsch(theta0, dth, n):
    kk = np.rint(dth / np.pi)
    d = dth - kk * np.pi
    sgn = -1.0 if int(kk) % 2 else 1.0
    eps = F32(4.0 * np.sin(0.5 * d) ** 2)
    y = F32(np.sin(theta0))
    dd = F32(np.sin(theta0 + d) - np.sin(theta0))
    out = []
    for m in range(n):
        out.append(np.float64(y) * (sgn ** m))
        y = F32(np.float64(y) + np.float64(dd))
        dd = fma(-eps, y, dd)
    return np.array(out, np.float64)


STEPS = [1e-4, 3e-3, 0.05, 0.3, 1.0, np.pi / 2, np.pi - 0.3, np.pi - 3e-3, np.pi + 1e-4,
         2 * np.pi + 2e-3, 3 * np.pi - 0.05]


def worst(fn, dth, n, trials=40):
    rs = np.random.RandomState(int(dth * 1e6) % 1000)
    e = 0.0
    for t0 in rs.uniform(0, 2 * np.pi, trials):
        e = max(e, float(np.max(np.abs(fn(t0, dth, n) - truth(t0, dth, n)))))
    return e / ULP


@pytest.mark.parametrize("n,bound", [(32, 80), (64, 200)])
def test_two_float_three_term_is_bounded_for_every_step(n, bound):
    # measured: at most ~75 ulp after 32 steps (the fused and contracted-force
    # kernels' anchor spacing) and ~180 after 64 for steps near k pi (~65 and
    # ~150 with the low part every fourth step)
    for dth in STEPS:
        assert worst(two_float, dth, n) < bound, dth


def test_rounded_coefficient_drifts_near_multiples_of_pi():
    # far from k pi the plain form is as good; near 0, pi and 2 pi it is
    # several times worse (measured ~250 ulp after 32 steps)
    assert worst(plain, 1.0, 32) < 20
    for dth in (3e-3, np.pi - 3e-3, 2 * np.pi + 2e-3):
        assert worst(plain, dth, 32) > max(200, 3 * worst(two_float, dth, 32)), dth


@pytest.mark.parametrize("dth", [1e-4, 3e-3, 0.05, 0.15, np.pi - 0.15, np.pi + 1e-4,
                                 2 * np.pi + 2e-3])
def test_reinsch_form_near_multiples_of_pi(dth):
    # the F-only kernel's form for steps within 0.2 of k pi: a few ulp after
    # 64 steps, where the two-float three-term form reaches over 100
    assert worst(reinsch, dth, 64) < 10


@pytest.mark.parametrize("dth", [0.5, 1.0, 1.3, np.pi / 2 - 0.01, 2.0, 4.0, 5.5])
def test_reinsch_form_for_every_step(dth):
    # a wave that takes the Reinsch form runs it on all its lanes, whatever
    # their step: with eps rounded once from FP64 it stays within ~15 ulp
    assert worst(reinsch, dth, 64) < 20


def test_f_only_longest_chunk():
    # the F-only kernel runs up to 112 bins from an anchor: its three-term form
    # (only for steps at least 0.2 from k pi) stays within ~20 ulp and its
    # Reinsch form within ~20 for every step
    far = [0.2, 0.25, 0.5, 1.0, 1.3, np.pi - 0.2, np.pi + 0.2, 2 * np.pi - 0.2, 2.5, 3.5, 5.0]
    assert max(worst(two_float, d, 112) for d in far) < 120
    anyd = [1e-3, 0.01, 0.1, 0.19, np.pi - 0.01, np.pi + 0.1, 2 * np.pi - 0.05, 0.5, 1.5, 2.9]
    assert max(worst(reinsch, d, 112) for d in anyd) < 30
