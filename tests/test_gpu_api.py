"""The ElasticScatter / Calc1D drop-in on the device, against the FP64 oracle
and the committed golden fixtures (the reference's cross-processor checks:
pyiid/tests/test_scatter.py, test_scatter_state.py, test_calc/test_calc_1d*.py).

Tolerance: 1e-5 of max|x*| for F, G and gradients; 1e-6 absolute for Rw.
"""
import os

import numpy as np
import pytest

from oracle import scatter as S
from pyiid_amd import Atoms, structures
from pyiid_amd.calc import Calc1D, wrap_rw
from pyiid_amd.elasticscatter import ElasticScatter

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


def random_experiment(rs):
    # ranges of pyiid/tests/__init__.py:117-128 (qbin drawn in [0.08, 0.12])
    return {"qmin": rs.uniform(0, 1.5), "qmax": rs.uniform(19., 25.),
            "qbin": rs.uniform(.08, .12), "rmin": rs.uniform(0., 2.5),
            "rmax": rs.uniform(30., 50.), "rstep": rs.uniform(.005, .015),
            "sampling": rs.choice(["full", "ns"])}


def test_golden_cuboctahedron():
    d = np.load(os.path.join(GOLD, "cuboctahedron55.npz"))
    atoms = Atoms("Au55", d["positions"].astype(np.float64))
    scat = ElasticScatter(seed=0)
    assert rel(scat.get_fq(atoms), d["fq"]) < TOL
    assert rel(scat.get_grad_fq(atoms), d["grad_fq"]) < TOL
    gr = scat.get_pdf(atoms)
    assert rel(gr, d["gr"]) < TOL
    rw, scale = wrap_rw(gr, d["gr_ideal"])
    assert abs(rw - float(d["rw"])) < 1e-6
    calc = Calc1D(target_data=d["gr_ideal"], exp_function=scat.get_pdf,
                  exp_grad_function=scat.get_grad_pdf, potential="rw")
    atoms.set_calculator(calc)
    assert abs(atoms.get_potential_energy() - float(d["rw"])) < 1e-6
    assert rel(atoms.get_forces(), d["forces_rw"]) < TOL


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_getters_random_experiments(seed):
    rs = np.random.RandomState(seed)
    exp = random_experiment(rs)
    n = [10, 100, 400][seed]
    pos = rs.random_sample((n, 3)) * 10
    atoms = Atoms("Au%d" % n, pos)
    atoms.center()
    scat = ElasticScatter(exp_dict=dict(exp), seed=1)
    e, pq = S.make_experiment(exp)
    fq = scat.get_fq(atoms)
    sf = atoms.get_array("F(Q) scatter")
    sp = atoms.get_array("PDF scatter")
    first = int(np.floor(e["qmin"] / e["qbin"]))
    fq_o = S.fq(atoms.positions, sf, e["qbin"])
    assert rel(fq, fq_o[first:]) < TOL
    assert rel(scat.get_grad_fq(atoms), S.grad_fq(atoms.positions, sf, e["qbin"])[:, :, first:]) < TOL
    r = S.r_grid(e)
    g_o = S.get_pdf_at_qmin(S.fq(atoms.positions, sp, pq), e["rstep"], pq, r, e["qmin"])
    assert rel(scat.get_pdf(atoms), g_o) < TOL
    with np.errstate(all="ignore"):
        sq_o = fq_o[first:] / S.scatter_vector(e) + 1
    sq_o[np.isinf(sq_o)] = 0
    np.testing.assert_allclose(scat.get_sq(atoms), sq_o, rtol=1e-4, atol=TOL * np.max(np.abs(sq_o)))
    if n <= 100:
        gfq = S.grad_fq(atoms.positions, sp, pq)
        gfq[:, :, :int(e["qmin"] / pq)] = 0
        gg_o = S.grad_pdf(gfq, e["rstep"], pq, r, e["qmin"])
        gg = scat.get_grad_pdf(atoms)
        assert gg.shape == gg_o.shape and rel(gg, gg_o) < TOL


def test_results_are_new_arrays():
    atoms = Atoms("Au20", np.random.RandomState(4).random_sample((20, 3)) * 8)
    scat = ElasticScatter(seed=2)
    a, b = scat.get_fq(atoms), scat.get_fq(atoms)
    assert a is not b and np.array_equal(a, b)
    g1, g2 = scat.get_grad_pdf(atoms), scat.get_grad_pdf(atoms)
    assert g1 is not g2


@pytest.mark.parametrize("change", ["add", "delete"])
def test_scatter_state_add_delete(change):
    # pyiid/tests/test_scatter_state.py:11-71
    rs = np.random.RandomState(6)
    atoms = Atoms("Au30", rs.random_sample((30, 3)) * 10)
    scat = ElasticScatter(seed=4)
    assert scat._check_wrap_atoms_state(atoms) is False
    ans1 = scat.get_fq(atoms)
    assert scat._check_wrap_atoms_state(atoms) is True and np.any(ans1)
    if change == "add":
        atoms2 = atoms + (79, [0.0, 0.0, 0.0])
    else:
        atoms2 = atoms.copy()
        del atoms2[rs.choice(len(atoms2))]
    assert scat._check_wrap_atoms_state(atoms2) is False
    ans2 = scat.get_fq(atoms2)
    assert np.any(ans2) and not np.allclose(ans1, ans2)
    assert rel(ans2, S.fq(atoms2.positions, atoms2.get_array("F(Q) scatter"), 0.1)) < TOL


@pytest.mark.parametrize("kind", ["FQ", "PDF"])
@pytest.mark.parametrize("potential", ["rw", "chi_sq"])
def test_calc1d_square(kind, potential):
    # pyiid/tests/test_calc/test_calc_1d_known.py: square vs its 0.75 copy
    sq = np.array([[0, 0, 0], [3, 0, 0], [0, 3, 0], [3, 3, 0]], float)
    atoms1 = Atoms("Au4", sq)
    atoms1.center()
    atoms2 = atoms1.copy()
    atoms2.positions *= .75
    scat = ElasticScatter(seed=5)
    f, g = (scat.get_fq, scat.get_grad_fq) if kind == "FQ" else (scat.get_pdf, scat.get_grad_pdf)
    target = f(atoms1)
    calc = Calc1D(target_data=target, exp_function=f, exp_grad_function=g, potential=potential)
    atoms2.set_calculator(calc)
    e = atoms2.get_potential_energy()
    forces = atoms2.get_forces()
    if potential == "rw":
        assert e >= .9
    # torque free (rotation invariance of the potential)
    com = atoms2.get_center_of_mass()
    torque = np.cross(atoms2.positions - com, forces)
    assert np.max(np.abs(torque)) < 1e-6 * max(1.0, np.max(np.abs(forces)))
    # fused device path == generic callable path (reference formula via getters)
    gen = Calc1D(target_data=target, exp_function=lambda a: f(a),
                 exp_grad_function=lambda a: g(a), potential=potential)
    assert abs(gen.get_potential_energy(atoms2) - e) < 1e-6 * max(1.0, abs(e))
    assert rel(gen.get_forces(atoms2), forces) < TOL


def test_calc1d_alloy_forces_vs_oracle():
    base = structures.fcc_nearest(147, a=3.724)
    pos = structures.jitter(base, 0.05, seed=9)
    numbers = structures.random_alloy(len(pos), 78, 28, seed=2)
    atoms0 = Atoms(numbers=numbers, positions=base)
    atoms = Atoms(numbers=numbers, positions=pos)
    scat = ElasticScatter({"qmin": 0.5, "rmax": 20.0}, seed=1)
    target = scat.get_pdf(atoms0)
    calc = Calc1D(target_data=target, exp_function=scat.get_pdf,
                  exp_grad_function=scat.get_grad_pdf, potential="rw", conv=100.)
    atoms.set_calculator(calc)
    e = atoms.get_potential_energy()
    fo = atoms.get_forces()
    e_, pq = S.make_experiment({"qmin": 0.5, "rmax": 20.0})
    r = S.r_grid(e_)
    scat._wrap_atoms(atoms)   # the calculator worked on a copy of the atoms
    sp = atoms.get_array("PDF scatter")
    gcalc = S.get_pdf_at_qmin(S.fq(atoms.positions, sp, pq), e_["rstep"], pq, r, e_["qmin"])
    gfq = S.grad_fq(atoms.positions, sp, pq)
    gfq[:, :, :int(0.5 / pq)] = 0
    gp = S.grad_pdf(gfq, e_["rstep"], pq, r, e_["qmin"])
    assert abs(e - 100 * S.get_rw(target, gcalc)[0]) < 1e-4
    assert rel(fo, 100 * S.wrap_grad_rw(gp, gcalc, target)) < TOL


def test_nuts_example_runs_on_device():
    # examples/Au_NP_PDF.py: an Au octahedron dilated by 5 % refined against
    # the PDF of the original with the NUTS ensemble and Calc1D(conv=100, 'rw')
    from pyiid_amd.nuts import NUTSCanonicalEnsemble
    atoms = Atoms(numbers=np.full(19, 79), positions=structures.octahedron(3))
    scat = ElasticScatter(seed=0)
    pdf = scat.get_pdf(atoms)
    atoms2 = atoms.copy()
    atoms2.positions *= 1.05
    calc = Calc1D(target_data=pdf, exp_function=scat.get_pdf, exp_grad_function=scat.get_grad_pdf,
                  conv=100, potential="rw")
    atoms2.set_calculator(calc)
    e_start = atoms2.get_potential_energy()
    ens = NUTSCanonicalEnsemble(atoms2, temperature=1000, escape_level=5, seed=3)
    traj, meta = ens.run(4)
    assert len(traj) == 1 + meta["accepted_samples"] and meta["samples_total"] > 0
    last = traj[-1]
    e_last = last.get_potential_energy()
    assert np.isfinite(e_start) and np.isfinite(e_last)
    # the device energy of the final state against the oracle's Rw
    e_, pq = S.make_experiment({})
    r = S.r_grid(e_)
    scat._wrap_atoms(last)
    sp = last.get_array("PDF scatter")
    gcalc = S.get_pdf_at_qmin(S.fq(last.positions, sp, pq), e_["rstep"], pq, r, e_["qmin"])
    assert abs(e_last - 100 * S.get_rw(pdf, gcalc)[0]) < 1e-4
