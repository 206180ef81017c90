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
