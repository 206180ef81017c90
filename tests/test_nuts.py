"""NUTS host mirror (pyiid_amd.nuts) with an analytic calculator.

Pins: a free particle conserves energy under velocity Verlet, so every leaf
is in the slice; the U-turn rule stops a harmonic oscillator's tree before
the escape depth; seeded runs are reproducible; the sample counts follow the
doubling (sum of 2^j per tree)."""
import numpy as np

from pyiid_amd import structures
from pyiid_amd.atoms import Atoms
from pyiid_amd.nuts import KB, NUTSCanonicalEnsemble, buildtree, maxwell_boltzmann


class Harmonic(object):
    """E = k/2 |x - x0|^2 (eV, A), forces -k (x - x0)."""

    def __init__(self, x0, k):
        self.x0 = np.array(x0, float)
        self.k = float(k)

    def get_potential_energy(self, atoms=None, force_consistent=False):
        d = atoms.positions - self.x0
        return 0.5 * self.k * float(np.sum(d * d))

    def get_forces(self, atoms=None):
        return -self.k * (atoms.positions - self.x0)


def _atoms(k=1.0, seed=0):
    x0 = structures.octahedron(3)
    a = Atoms(numbers=np.full(len(x0), 79), positions=structures.jitter(x0, 0.05, seed))
    a.set_calculator(Harmonic(x0, k))
    return a


def test_octahedron_magic_numbers():
    assert [len(structures.octahedron(n)) for n in (2, 3, 4, 5)] == [6, 19, 44, 85]
    # nearest-neighbour distance a / sqrt(2)
    p = structures.octahedron(3)
    d = np.linalg.norm(p[:, None] - p[None], axis=-1)
    np.testing.assert_allclose(np.min(d[d > 0]), 4.08 / np.sqrt(2), rtol=1e-12)


def test_maxwell_boltzmann_forces_temperature():
    a = _atoms()
    rs = np.random.RandomState(1)
    maxwell_boltzmann(a, 300 * KB, rs, force_temp=True)
    np.testing.assert_allclose(a.get_kinetic_energy(), 1.5 * len(a) * 300 * KB, rtol=1e-12)


def test_free_particle_leaves_are_in_the_slice():
    a = _atoms(k=0.0)
    maxwell_boltzmann(a, 0.1, np.random.RandomState(2))
    e0 = a.get_total_energy()
    neg, pos, prime, n, s, alpha, na = buildtree(a, 0.999, 1, 3, 0.05, e0,
                                                 np.random.RandomState(3))
    assert n == 8 and na == 8 and abs(alpha - 8.0) < 1e-9
    # a straight flight never turns back
    assert s == 1
    # leapfrog re-centres after each step: compare up to a rigid translation
    want = a.positions + 8 * 0.05 * a.get_velocities()
    got = pos.positions
    np.testing.assert_allclose(got - got.mean(0), want - want.mean(0), atol=1e-9)


def test_nuts_runs_reproducibly_and_stops_on_u_turns():
    runs = []
    for _ in range(2):
        ens = NUTSCanonicalEnsemble(_atoms(k=2.0), temperature=300, escape_level=8, seed=7)
        traj, meta = ens.run(6)
        runs.append((traj, meta, ens))
    (t1, m1, e1), (t2, m2, e2) = runs
    assert len(t1) == len(t2) and len(t1) >= 2
    for a, b in zip(t1, t2):
        np.testing.assert_array_equal(a.positions, b.positions)
    assert m1["seed"] == 7 and m1["planned iterations"] == 6
    assert m1["accepted_samples"] == len(t1) - 1
    # each tree doubles: samples_total is a sum of 2^j terms, all below the escape depth
    assert m1["samples_total"] >= 2 * 6
    assert e1.m == 6 and np.isfinite(e1.step_size) and e1.step_size > 0
    assert len(e1.pe) == len(t1) - 1
