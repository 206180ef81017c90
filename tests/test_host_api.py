"""Host-side logic of the ElasticScatter mirror that needs no device: grids,
per-atom scatter arrays, cache invalidation (pyiid/tests/test_scatter_state.py),
processor selection errors, and the reference's _wrap_atoms quirks."""
import numpy as np
import pytest

from oracle import scatter as S
from pyiid_amd import Atoms
from pyiid_amd import elasticscatter as es
from pyiid_amd import processor


@pytest.fixture
def scat(monkeypatch):
    monkeypatch.setattr(processor, "device_available", lambda: True)
    return es.ElasticScatter({"qmin": 0.7, "qmax": 22.0, "qbin": 0.05, "rmax": 30.0},
                             seed=3)


def test_grids_match_oracle(scat):
    exp, pq = S.make_experiment({"qmin": 0.7, "qmax": 22.0, "qbin": 0.05, "rmax": 30.0})
    assert scat.pdf_qbin == pq
    np.testing.assert_array_equal(scat.get_scatter_vector(), S.scatter_vector(exp))
    np.testing.assert_array_equal(scat.get_scatter_vector(pdf=True),
                                  S.pdf_scatter_vector(exp, pq))
    np.testing.assert_array_equal(scat.get_r(), S.r_grid(exp))


def test_ns_sampling_and_defaults(monkeypatch):
    monkeypatch.setattr(processor, "device_available", lambda: True)
    d = {"sampling": "ns", "qmax": 20.0}
    s = es.ElasticScatter(d)
    assert s.exp is d and d["rstep"] == np.pi / 20.0 and d["qbin"] == .1


def test_processor_selection(scat, monkeypatch):
    assert scat.processor == "MI355X" and scat.fq is processor.wrap_fq
    with pytest.raises(ValueError):
        scat.set_processor("CPU", "flat")
    monkeypatch.setattr(processor, "device_available", lambda: False)
    with pytest.raises(RuntimeError):
        scat.set_processor()
    with pytest.raises(ValueError):
        es.ElasticScatter(seed=1.5)


def test_wrap_atoms_and_state(scat):
    rs = np.random.RandomState(0)
    atoms = Atoms(numbers=[79, 28, 79, 78], positions=rs.random_sample((4, 3)) * 5)
    assert scat._check_wrap_atoms_state(atoms) is False
    assert scat._check_wrap_atoms_state(atoms) is True
    f = atoms.get_array("F(Q) scatter")
    p = atoms.get_array("PDF scatter")
    assert f.shape == (4, int(np.floor(22.0 / 0.05))) and p.shape == (4, int(np.floor(22.0 / scat.pdf_qbin)))
    np.testing.assert_array_equal(f[0], f[2])
    table, index = processor.compact_scatter(atoms, "F(Q) scatter")
    np.testing.assert_array_equal(table[index], f)
    atoms2 = atoms + (79, [1.0, 1.0, 1.0])
    assert scat.check_wrap_atoms_state(atoms2) is False
    del atoms2[0]
    assert scat._check_wrap_atoms_state(atoms2) is False
    scat.update_experiment({"qmax": 20.0})
    assert scat.check_wrap_atoms_state(atoms) is False


def test_compact_form_from_foreign_arrays(scat):
    atoms = Atoms("Au3", np.eye(3))
    atoms.set_array("F(Q) scatter", np.array([[1, 2], [3, 4], [1, 2]], np.float32))
    table, index = processor.compact_scatter(atoms, "F(Q) scatter")
    np.testing.assert_array_equal(table[index], atoms.arrays["F(Q) scatter"])


def test_compat_formfactor_quirk(monkeypatch):
    # reference _wrap_atoms: rows evaluated for numbers[0..E-1], assigned to
    # set-ordered elements, both tables on the qbin grid (__init__.py:127-143)
    monkeypatch.setattr(processor, "device_available", lambda: True)
    s = es.ElasticScatter(compat_formfactors=True)
    atoms = Atoms(numbers=[28, 28, 79], positions=np.eye(3))
    s._wrap_atoms(atoms)
    from pyiid_amd.formfactors import form_factor
    e_list = list(set([28, 28, 79]))
    q = np.arange(250) * 0.1
    rows = [form_factor(28, q), form_factor(28, q)]   # numbers[0], numbers[1]
    for i, z in enumerate(e_list):
        got = atoms.get_array("F(Q) scatter")[np.array([28, 28, 79]) == z][0]
        np.testing.assert_allclose(got, rows[i].astype(np.float32))


def _project_by_bins(q, iq, qbin, pixels):
    """Detector projection bin by bin: every pixel strictly inside a bin's
    window (Q_k - qbin/2, Q_k + qbin/2) takes that bin's I, later bins
    overwriting earlier ones (pyiid/experiments/elasticscatter/__init__.py:
    448-475)."""
    flat = pixels.ravel()
    img = np.zeros(flat.shape)
    for qk, ik in zip(q, iq):
        img[(qk - qbin / 2. < flat) & (qk + qbin / 2. > flat)] = ik
    return img.reshape(pixels.shape)


@pytest.mark.parametrize("qmin,qbin", [(0.0, 0.1), (0.7, 0.0756865), (1.3, 0.12)])
def test_detector_projection_equals_bin_loop(qmin, qbin):
    rs = np.random.RandomState(3)
    q = np.arange(qmin, np.floor(25.0 / qbin) * qbin, qbin)
    iq = rs.random_sample(len(q)) + 0.5
    centres = rs.choice(q, 50)
    edges = np.concatenate([q - qbin / 2., q + qbin / 2.])[rs.randint(0, 2 * len(q), 60)]
    pix = np.concatenate([rs.uniform(-1.0, 27.0, 4096 - 110 - 4), centres, edges,
                          [np.nan, -np.inf, np.inf, q[-1] + qbin / 2.]]).reshape(64, 64)
    got = es.project_to_detector(q, iq, qbin, pix)
    want = _project_by_bins(q, iq, qbin, pix)
    np.testing.assert_array_equal(got, want)
    # pixels just inside a window take that bin; outside every window, 0
    k = len(q) // 2
    inside = np.array([[q[k] - 0.49 * qbin, q[k] + 0.49 * qbin, q[0] - 0.51 * qbin]])
    np.testing.assert_array_equal(es.project_to_detector(q, iq, qbin, inside),
                                  [[iq[k], iq[k], 0.0]])


def test_shared_context_serves_each_thread_its_own_structure(monkeypatch):
    """Two threads share the per-device context.  Each must get F(Q) and the
    gradient of the structure it passed.  The stand-in context answers from
    the positions last uploaded and sleeps between upload and compute, so
    without the lock held across both, the threads' structures would
    interleave."""
    import threading
    import time

    class Recording:
        def __init__(self, device=0):
            self.pos = None

        def set_positions(self, pos):
            self.pos = np.array(pos, np.float32)
            time.sleep(0.001)

        def set_elements(self, elem):
            pass

        def set_scatter(self, slot, table):
            pass

        def fq(self, slot, qbin, nq):
            time.sleep(0.001)
            return np.full(nq, self.pos.sum(), np.float32)

        def grad_fq(self, slot, qbin, n, nq, out=None, f64=False):
            time.sleep(0.001)
            out = np.empty((n, 3, nq)) if out is None else out
            out[...] = self.pos.sum()
            return out

    monkeypatch.setattr(processor._lib, "Context", Recording)
    monkeypatch.setattr(processor, "_ctx", {})
    structs = []
    for k in range(2):
        a = Atoms("Au%d" % (5 + k), np.random.RandomState(k).random_sample((5 + k, 3)) * 5)
        a.set_array("F(Q) scatter", np.ones((len(a), 30), np.float32))
        structs.append(a)
    wrong, errors = [], []

    def work(a):
        want = np.float32(np.asarray(a.get_positions(), np.float32).sum())
        try:
            for _ in range(40):
                if processor.wrap_fq(a)[0] != want or processor.wrap_fq_grad(a)[0, 0, 0] != want:
                    wrong.append(len(a))
        except Exception as exc:  # a failure in a thread must fail the test
            errors.append(repr(exc))

    threads = [threading.Thread(target=work, args=(a,)) for a in structs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert errors == []
    assert wrong == []
