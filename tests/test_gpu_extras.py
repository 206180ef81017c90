"""ElasticScatter's remaining getters on the device: I(Q), the detector
image, and the seeded noise of get_fq / get_sq / get_pdf
(pyiid/experiments/elasticscatter/__init__.py:304-475).

F comes from the device.  The oracle side restates each getter around the
FP64 oracle's F(Q):
  * S = F/Q + 1 with inf -> 0, and I = S <f>^2;
  * the detector image, bin by bin;
  * the noise: widths sigma_I Q / <f>^2 for F and sigma_I Q / <f^2> for the
    PDF's F, with a zero first width raised to 1e-9, drawn from a
    RandomState of the experiment's seed.
The noise widths are made large, so that a wrong width formula moves the
result far beyond the 1e-5 bar.
"""
import numpy as np
import pytest

from oracle import scatter as S
from pyiid_amd import Atoms, structures
from pyiid_amd.elasticscatter import ElasticScatter
from test_host_api import _project_by_bins

pytestmark = pytest.mark.gpu
TOL = 1e-5
# qmin 0.5: at 0.6 the reference's own grids disagree by a bin
# (floor(0.6 / 0.1) = 5, but arange(0.6, 22, 0.1) has 214 points), and its
# get_sq cannot broadcast F against Q
EXP = {"qmin": 0.5, "qmax": 22.0, "qbin": 0.1, "rmin": 0.0, "rmax": 30.0, "rstep": 0.01}


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def _alloy():
    base = structures.fcc_nearest(201, a=3.9)
    pos = structures.jitter(base, 0.05, seed=8)
    z = structures.random_alloy(len(pos), 78, 28, seed=9)
    return Atoms(numbers=z, positions=pos)


def _oracle_f(atoms, name, qbin):
    pos = atoms.get_positions().astype(np.float32)
    return S.fq(pos, atoms.get_array(name), qbin)


def _oracle_iq(atoms, scat, f):
    q = scat.get_scatter_vector()
    first = int(np.floor(scat.exp["qmin"] / scat.exp["qbin"]))
    with np.errstate(all="ignore"):
        sq = f / q + 1.0
    sq[np.isinf(sq)] = 0.0
    mean_f = atoms.get_array("F(Q) scatter").mean(axis=0)
    return sq, sq * (mean_f ** 2)[first:]


def test_iq_and_detector_image():
    atoms = _alloy()
    scat = ElasticScatter(exp_dict=dict(EXP), seed=1)
    iq = scat.get_iq(atoms)
    first = int(np.floor(EXP["qmin"] / EXP["qbin"]))
    f = _oracle_f(atoms, "F(Q) scatter", EXP["qbin"])[first:]
    sq_ref, iq_ref = _oracle_iq(atoms, scat, f)
    assert rel(scat.get_sq(atoms), sq_ref) < TOL
    assert rel(iq, iq_ref) < TOL
    # a flat detector 0.2 m from the sample at 0.5 A: Q = 4 pi sin(theta) / lambda
    yy, xx = np.mgrid[-128:128, -128:128] * 1e-3
    two_theta = np.arctan(np.hypot(xx, yy) / 0.2)
    pix = 4 * np.pi * np.sin(two_theta / 2) / 0.5
    img = scat.get_2d_scatter(atoms, pix)
    want = _project_by_bins(scat.get_scatter_vector(), iq_ref, EXP["qbin"], pix)
    assert img.shape == pix.shape
    assert np.count_nonzero(want) > 0.9 * pix.size
    assert rel(img, want) < TOL


def test_fq_and_sq_noise():
    atoms = _alloy()
    seed, iq_std = 11, 40.0
    scat = ElasticScatter(exp_dict=dict(EXP), seed=seed)
    noisy = scat.get_fq(atoms, iq_std=iq_std)
    first = int(np.floor(EXP["qmin"] / EXP["qbin"]))
    f = _oracle_f(atoms, "F(Q) scatter", EXP["qbin"])[first:]
    mean_f = atoms.get_array("F(Q) scatter").mean(axis=0)[first:]
    width = iq_std * np.abs(scat.get_scatter_vector()) / mean_f ** 2
    rs = np.random.RandomState(seed)
    want = f + rs.normal(0.0, width)
    assert rel(noisy, want) < TOL
    assert rel(noisy, f) > 100 * TOL          # the noise is there
    # S(Q) takes the generator's next draws
    sq = scat.get_sq(atoms, iq_std=iq_std)
    f2 = f + rs.normal(0.0, width)
    with np.errstate(all="ignore"):
        want_sq = f2 / scat.get_scatter_vector() + 1.0
    assert rel(sq, want_sq) < TOL


@pytest.mark.parametrize("kind", ["scalar", "array"])
def test_pdf_noise(kind):
    atoms = _alloy()
    seed = 5
    scat = ElasticScatter(exp_dict=dict(EXP), seed=seed)
    q = np.abs(scat.get_scatter_vector(pdf=True))
    iq_std = 60.0 if kind == "scalar" else np.linspace(20.0, 90.0, len(q))
    gr = scat.get_pdf(atoms, iq_std=iq_std)
    f = _oracle_f(atoms, "PDF scatter", scat.pdf_qbin)
    per = atoms.get_array("PDF scatter").astype(np.float64)
    width = iq_std * q / np.mean(per ** 2, axis=0)
    assert width[0] == 0.0                     # Q = 0: the 1e-9 floor applies
    width[0] = 1e-9
    f = f + np.random.RandomState(seed).normal(0.0, width)
    r = scat.get_r()
    want = S.get_pdf_at_qmin(f, EXP["rstep"], scat.pdf_qbin, r, EXP["qmin"])
    assert rel(gr, want) < TOL
    clean = S.get_pdf_at_qmin(_oracle_f(atoms, "PDF scatter", scat.pdf_qbin), EXP["rstep"],
                              scat.pdf_qbin, r, EXP["qmin"])
    assert rel(gr, clean) > 100 * TOL


def test_pdf_noise_width_array_on_another_grid():
    """A width array of another length is resampled linearly by bin index,
    holding the last value beyond its end.  The reference's resampling call
    raises a TypeError here (np.arange with a tuple stop)."""
    atoms = _alloy()
    scat = ElasticScatter(exp_dict=dict(EXP), seed=2)
    q = np.abs(scat.get_scatter_vector(pdf=True))
    short = np.linspace(30.0, 70.0, 40)
    gr = scat.get_pdf(atoms, iq_std=short)
    idx = np.arange(len(q))
    resampled = np.where(idx < len(short) - 1,
                         short[0] + (short[1] - short[0]) * idx, short[-1])
    per = atoms.get_array("PDF scatter").astype(np.float64)
    width = resampled * q / np.mean(per ** 2, axis=0)
    width[0] = 1e-9
    f = _oracle_f(atoms, "PDF scatter", scat.pdf_qbin) + \
        np.random.RandomState(2).normal(0.0, width)
    want = S.get_pdf_at_qmin(f, EXP["rstep"], scat.pdf_qbin, scat.get_r(), EXP["qmin"])
    assert rel(gr, want) < TOL
