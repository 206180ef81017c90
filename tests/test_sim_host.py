"""Host samplers (pyiid_amd.nuts, pyiid_amd.sim) with analytic calculators.

Pins:
* buildtree (built leaf by leaf with a stack of finished halves) against the
  paper's recursive BuildTree, restated here (Hoffman & Gelman 2014,
  Algorithm 3, with the slice, Delta_max and alpha conventions the module
  documents): for the same random stream both give the same edges,
  proposal, counts and statistics, including trees cut short by a U-turn
  and by the slice;
* simulate_dynamics with unit masses equals n_steps + 1 velocity-Verlet
  steps written out here;
* run_hmc's bookkeeping and step-size rule.
"""
import numpy as np
import pytest

from pyiid_amd import structures
from pyiid_amd.atoms import Atoms
from pyiid_amd.nuts import buildtree, maxwell_boltzmann
from pyiid_amd.sim import leapfrog, run_hmc, simulate_dynamics


class Harmonic(object):
    """E = k/2 |x - x0|^2, F = -k (x - x0)."""

    def __init__(self, x0, k):
        self.x0 = np.array(x0, float)
        self.k = float(k)

    def get_potential_energy(self, atoms=None, force_consistent=False):
        d = atoms.positions - self.x0
        return 0.5 * self.k * float(np.sum(d * d))

    def get_forces(self, atoms=None):
        return -self.k * (atoms.positions - self.x0)


def _atoms(k, seed=0, masses=None):
    x0 = structures.octahedron(3)
    a = Atoms(numbers=np.full(len(x0), 79), positions=structures.jitter(x0, 0.05, seed),
              masses=masses)
    a.set_calculator(Harmonic(x0 - x0.mean(0), k))
    return a


def _recursive_tree(state, u, v, j, e, e0, rs):
    """The paper's recursive BuildTree with the module's conventions:
    leaf in the slice iff u <= exp(e0 - E), alive iff u < exp(200 + e0 - E),
    alpha against the parent state, U-turn on velocities."""
    if j == 0:
        leaf = leapfrog(state, v * e)
        d = e0 - leaf.get_total_energy()
        with np.errstate(over="ignore"):
            n = int(u <= np.exp(d))
            s = int(u < np.exp(200.0 + d))
            a = min(1.0, float(np.exp(state.get_total_energy() - leaf.get_total_energy())))
        return leaf, leaf, leaf, n, s, a, 1
    minus, plus, prop, n, s, a, na = _recursive_tree(state, u, v, j - 1, e, e0, rs)
    if s == 1:
        if v == -1:
            minus, _, prop2, n2, s2, a2, na2 = _recursive_tree(minus, u, v, j - 1, e, e0, rs)
        else:
            _, plus, prop2, n2, s2, a2, na2 = _recursive_tree(plus, u, v, j - 1, e, e0, rs)
        if rs.uniform() < n2 / float(max(n + n2, 1)):
            prop = prop2
        span = (plus.positions - minus.positions).ravel()
        s = s2 * int(span @ minus.get_velocities().ravel() >= 0) * \
            int(span @ plus.get_velocities().ravel() >= 0)
        n, a, na = n + n2, a + a2, na + na2
    return minus, plus, prop, n, s, a, na


@pytest.mark.parametrize("k,step,u,j,v", [
    (0.5, 0.05, 0.3, 4, 1),      # grows to the full depth
    (2.0, 0.6, 0.3, 5, 1),       # turns back before the end
    (2.0, 0.6, 0.3, 5, -1),
    (40.0, 0.2, 0.9, 4, 1),      # energy error drops leaves out of the slice
    (300.0, 0.5, 0.5, 3, -1),    # leaves leave the slice tolerance (tree dies)
    (1.0, 0.1, 0.2, 0, 1),
])
def test_buildtree_equals_recursive_restatement(k, step, u, j, v):
    a = _atoms(k, seed=2, masses=np.full(19, 1.0))
    maxwell_boltzmann(a, 0.5, np.random.RandomState(5))
    e0 = a.get_total_energy()
    got = buildtree(a, u, v, j, step, e0, np.random.RandomState(11))
    want = _recursive_tree(a, u, v, j, step, e0, np.random.RandomState(11))
    for x, y in zip(got[:3], want[:3]):
        np.testing.assert_array_equal(x.positions, y.positions)
        np.testing.assert_array_equal(x.get_momenta(), y.get_momenta())
    assert got[3:5] == want[3:5]
    assert got[5] == pytest.approx(want[5], rel=1e-15, abs=0.0)
    assert got[6] == want[6]


def test_buildtree_cases_cover_each_ending():
    """The parameter sets above reach a full tree, a U-turn and a dead slice."""
    ends = set()
    for k, step, u, j, v in [(0.5, 0.05, 0.3, 4, 1), (2.0, 0.6, 0.3, 5, 1),
                             (300.0, 0.5, 0.5, 3, -1)]:
        a = _atoms(k, seed=2, masses=np.full(19, 1.0))
        maxwell_boltzmann(a, 0.5, np.random.RandomState(5))
        e0 = a.get_total_energy()
        _, _, _, n, s, _, na = buildtree(a, u, v, j, step, e0, np.random.RandomState(11))
        ends.add("full" if s == 1 and na == 2 ** j else ("cut" if na < 2 ** j else "stop"))
    assert {"full", "cut"} <= ends


def test_simulate_dynamics_is_velocity_verlet_with_unit_masses():
    k, h, n_steps = 1.5, 0.07, 6
    a = _atoms(k, seed=3, masses=np.full(19, 1.0))
    rs = np.random.RandomState(4)
    a.set_momenta(rs.normal(0, 0.3, (19, 3)))
    x = a.positions.copy()
    p = a.get_momenta().copy()
    x0 = a.calc.x0
    for _ in range(n_steps + 1):
        p = p - 0.5 * h * k * (x - x0)
        x = x + h * p
        p = p - 0.5 * h * k * (x - x0)
    simulate_dynamics(a, h, n_steps)
    np.testing.assert_allclose(a.positions, x, rtol=0, atol=1e-12)
    np.testing.assert_allclose(a.get_momenta(), p, rtol=0, atol=1e-12)


def test_run_hmc_bookkeeping_and_step_rule():
    a = _atoms(0.8, seed=6, masses=np.full(19, 1.0))
    rs = np.random.RandomState(8)
    traj, decisions, accepted_at = run_hmc(a, 12, 0.05, 4, 0.5, 0.5, 0.6, 1.1, 0.9, 0.01, 0.2,
                                           temp=10.0, rs=rs)
    assert len(decisions) == 12
    assert len(traj) == 1 + sum(bool(d) for d in decisions)
    assert accepted_at == [i for i, d in enumerate(decisions) if d]
    assert traj[0] is a
    # the step rule alone: running rate relaxed towards the mean acceptance
    from pyiid_amd.sim import _AcceptanceStep
    ad = _AcceptanceStep(0.1, 0.5, 0.5, 0.6, 2.0, 0.5, 0.01, 1.0)
    assert ad.record(True) == pytest.approx(0.2)       # rate 0.75 > 0.6: grow
    assert ad.record(False) == pytest.approx(0.4)      # rate 0.625 > 0.6: grow
    assert ad.record(False) == pytest.approx(0.2)      # rate 0.479 < 0.6: shrink
    for _ in range(10):
        ad.record(False)
    assert ad.step == pytest.approx(0.01)              # clamped at the minimum
