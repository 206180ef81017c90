"""Cromer-Mann form factors against the reference's carbon fixture.

tests/golden/c60_carbon_row.txt is pyIID's xraylib carbon table
(pyiid/tests/test_master/c60_scat.txt); the reference compares its own
get_scatter_array with it at rtol 1e-2 (test_master_kernel.py:8-16).  Other
elements are "parity unpinned" (xraylib is absent); f(0) = sum a_i + c must
still be the electron count Z.
"""
import os

import numpy as np
import pytest

from pyiid_amd import formfactors

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_carbon_matches_reference_fixture():
    want = np.loadtxt(os.path.join(GOLD, "c60_carbon_row.txt"))
    q = np.arange(250) * 0.1
    got = formfactors.form_factor(6, q)
    # Cromer-Mann is fitted for s = Q/4pi <= 2 A^-1: 1e-2 inside, 1.5e-2 at the edge
    inside = q / (4 * np.pi) < 1.9
    np.testing.assert_allclose(got[inside], want[inside], rtol=1e-2)
    np.testing.assert_allclose(got, want, rtol=1.5e-2)


@pytest.mark.parametrize("z", sorted(formfactors.CROMER_MANN))
def test_forward_scattering_is_electron_count(z):
    assert abs(formfactors.form_factor(z, [0.0])[0] - z) < 0.05


def test_element_table_layout():
    numbers = np.array([79, 28, 79, 78, 28])
    table, index, elems = formfactors.element_table(numbers, 0.1, 7)
    assert list(elems) == [28, 78, 79]
    assert table.shape == (3, 7) and table.dtype == np.float32
    np.testing.assert_array_equal(index, [2, 0, 2, 1, 0])
    np.testing.assert_allclose(table[index[0]], formfactors.form_factor(79, np.arange(7) * 0.1),
                               rtol=1e-6)


def _fake_xraylib(monkeypatch, calls):
    """A stand-in xraylib module whose FF_Rayl(Z, s) = Z exp(-s) records its
    arguments, so the plumbing of the xraylib source is testable without it."""
    import sys
    import types
    mod = types.ModuleType("xraylib")

    def ff_rayl(z, s):
        calls.append((z, s))
        return z * np.exp(-s)

    mod.FF_Rayl = ff_rayl
    monkeypatch.setitem(sys.modules, "xraylib", mod)


def test_source_resolution(monkeypatch):
    import sys
    monkeypatch.setitem(sys.modules, "xraylib", None)   # import raises ImportError
    assert formfactors.resolve_source("auto") == "cromer-mann"
    with pytest.raises(ImportError):
        formfactors.resolve_source("xraylib")
    with pytest.raises(ValueError):
        formfactors.resolve_source("waasmaier")
    _fake_xraylib(monkeypatch, [])
    assert formfactors.resolve_source("auto") == "xraylib"
    assert formfactors.resolve_source("cromer-mann") == "cromer-mann"


def test_xraylib_source_calls_ff_rayl_at_q_over_4pi(monkeypatch):
    # the reference fills f[atom][k] = FF_Rayl(Z, k qbin / 4 / pi)
    # (master_kernel.py:31-36)
    calls = []
    _fake_xraylib(monkeypatch, calls)
    table, index, elems = formfactors.element_table(np.array([79, 28, 79]), 0.1, 5, "xraylib")
    assert list(elems) == [28, 79]
    want = np.array([[z * np.exp(-(k * 0.1 / 4 / np.pi)) for k in range(5)] for z in (28, 79)],
                    np.float32)
    np.testing.assert_array_equal(table, want)
    assert [z for z, _ in calls] == [28] * 5 + [79] * 5
    assert [s for _, s in calls[:5]] == [k * 0.1 / 4 / np.pi for k in range(5)]
    np.testing.assert_array_equal(index, [1, 0, 1])


def test_elasticscatter_uses_the_resolved_source(monkeypatch):
    from pyiid_amd import Atoms, processor
    from pyiid_amd import elasticscatter as es
    monkeypatch.setattr(processor, "device_available", lambda: True)
    _fake_xraylib(monkeypatch, [])
    atoms = Atoms(numbers=[28, 79], positions=np.eye(3)[:2])
    s = es.ElasticScatter({"qmax": 1.0, "qbin": 0.25})
    assert s.form_factor_source == "xraylib"
    s._wrap_atoms(atoms)
    got = atoms.get_array("F(Q) scatter")
    q = np.arange(4) * 0.25
    np.testing.assert_array_equal(got[1], (79 * np.exp(-(q / 4 / np.pi))).astype(np.float32))
    c = es.ElasticScatter({"qmax": 1.0, "qbin": 0.25}, form_factors="cromer-mann")
    c._wrap_atoms(atoms)
    np.testing.assert_array_equal(atoms.get_array("F(Q) scatter")[1],
                                  formfactors.form_factor(79, q).astype(np.float32))
