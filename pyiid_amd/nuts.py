"""No-U-Turn sampler over the MI355X Calc1D (host side).

Written from the published algorithm, Hoffman & Gelman, "The No-U-Turn
Sampler", JMLR 15 (2014): Algorithm 3 (efficient NUTS: slice variable,
trajectory doubling in a random direction, progressive sampling of the
proposal, U-turn stop), Algorithm 4 (reasonable first step size) and
Algorithm 6 / section 3.2 (dual averaging of the step size).  The interface
is that of pyIID's ensembles so that callers switch without edits:
``Ensemble.run`` returns ``(traj, metadata)`` with the initial configuration
first (pyiid/sim/__init__.py:41-82), ``NUTSCanonicalEnsemble(atoms,
temperature=, escape_level=, accept_target=, momentum=, seed=)`` and
``buildtree(input_atoms, u, v, j, e, e0, rs, beta)`` returning
``(minus, plus, proposal, n, s, alpha_sum, n_alpha)`` (pyiid/sim/nuts_hmc.py).

Where pyIID's sampler departs from the paper this module follows pyIID
(each item is behaviour its users see):

* energies enter the slice and U-turn tests in eV as they are (the paper's
  log-density is -H with kT = 1), while momenta are thermal at the ensemble's
  temperature; the slice variable is drawn as u ~ U(0, 1) against
  exp(E0 - E(leaf)) (the paper's u ~ U(0, exp(-H0)) rescaled);
* a leaf ends the tree when u >= exp(200 + E0 - E(leaf)) (the paper's
  Delta_max is 1000);
* the acceptance statistic of a leaf is min(1, exp(E(parent) - E(leaf)))
  against the state it was integrated from, not against the trajectory's
  start;
* a tree stops doubling at depth ``escape_level``;
* ``metadata['samples_total']`` grows by 2**depth after each doubling, with
  the depth counted after it;
* dual averaging weighs the m-th update by 1 / (m + t0) and sets
  log eps = mu - sqrt(m) / gamma * H with m counted from 0 before the
  update; there is no averaged iterate (the paper's kappa);
* the first step size search starts at 0.5 and gives up (step 1.0) outside
  [1e-7, 1e7];
* the U-turn test projects the span on velocities (M^-1 p).

Deltas from pyIID itself: momenta come from the ensemble's RandomState
(pyIID draws them through ASE from numpy's global generator), so a seeded
run repeats exactly; accepted samples' potential energies are kept in
``pe`` (pyIID leaves it empty), which the ``eq_steps`` test of ``run`` reads,
and that test reports "not settled" while fewer than 2 eq_steps + 2 energies
exist (pyIID's raises there and ``run`` swallows the error).
"""
from copy import deepcopy

import numpy as np

from .atoms import Atoms
from .sim import leapfrog

__all__ = ["Ensemble", "NUTSCanonicalEnsemble", "buildtree", "maxwell_boltzmann", "KB", "FS"]

KB = 8.617330337217213e-05       # Boltzmann constant, eV / K
FS = 0.09822694788464063          # 1 fs in A sqrt(amu / eV)
DELTA_MAX = 200.0                 # slice tolerance that ends a tree (pyIID's Emax)


def _copy(atoms):
    """This package's Atoms share their calculator (and its device context)
    through copy(); other atoms objects are deep-copied."""
    return atoms.copy() if isinstance(atoms, Atoms) else deepcopy(atoms)


def maxwell_boltzmann(atoms, thermal_energy, rs, force_temp=True):
    """Momenta p ~ N(0, m kT) per component; with ``force_temp`` rescaled so
    that the kinetic energy is exactly 3/2 N kT."""
    sd = np.sqrt(atoms.get_masses() * thermal_energy)
    p = rs.normal(0.0, 1.0, (len(atoms), 3)) * sd[:, None]
    atoms.set_momenta(p)
    if force_temp and len(atoms):
        ke = atoms.get_kinetic_energy()
        if ke > 0:
            atoms.set_momenta(p * np.sqrt(1.5 * len(atoms) * thermal_energy / ke))


def _turned_back(minus, plus):
    """True when the trajectory between the two edge states has started to
    come back on itself: the span (plus - minus) points against the velocity
    at either end (Hoffman & Gelman eq. 9)."""
    span = (plus.positions - minus.positions).ravel()
    return (span @ minus.get_velocities().ravel() < 0.0 or
            span @ plus.get_velocities().ravel() < 0.0)


class _Subtree(object):
    """A run of 2**level consecutive leapfrog states in one direction: its
    edge states, the state it proposes, how many of its leaves lie in the
    slice, whether it may still grow, and its acceptance statistics."""
    __slots__ = ("minus", "plus", "proposal", "n_slice", "alive", "alpha_sum", "n_alpha",
                 "level")

    def __init__(self, minus, plus, proposal, n_slice, alive, alpha_sum, n_alpha, level):
        self.minus, self.plus, self.proposal = minus, plus, proposal
        self.n_slice, self.alive = n_slice, alive
        self.alpha_sum, self.n_alpha, self.level = alpha_sum, n_alpha, level

    def front(self, v):
        return self.plus if v > 0 else self.minus

    def as_tuple(self):
        return (self.minus, self.plus, self.proposal, self.n_slice, int(self.alive),
                self.alpha_sum, self.n_alpha)


def _leaf(parent, u, v, e, e0):
    """One leapfrog step of size v e from ``parent`` (a subtree of level 0)."""
    state = leapfrog(parent, v * e)
    e_new = state.get_total_energy()
    d = e0 - e_new                                       # -dH against the start
    with np.errstate(over="ignore"):
        in_slice = int(u <= np.exp(d))
        alive = bool(u < np.exp(DELTA_MAX + d))
        alpha = min(1.0, float(np.exp(parent.get_total_energy() - e_new)))
    return _Subtree(state, state, state, in_slice, alive, alpha, 1, 0)


def _join(older, newer, v, rs):
    """The subtree made of ``older`` (built first) and ``newer`` (its
    continuation in direction v): the proposal moves to the newer half's with
    probability n_newer / (n_older + n_newer) (progressive sampling), the
    counts add, and the joined run may grow on only if the newer half could
    and the whole run has not turned back."""
    if rs.uniform() < float(newer.n_slice) / max(older.n_slice + newer.n_slice, 1):
        proposal = newer.proposal
    else:
        proposal = older.proposal
    minus, plus = (newer.minus, older.plus) if v < 0 else (older.minus, newer.plus)
    alive = newer.alive and not _turned_back(minus, plus)
    return _Subtree(minus, plus, proposal, older.n_slice + newer.n_slice, alive,
                    older.alpha_sum + newer.alpha_sum, older.n_alpha + newer.n_alpha,
                    older.level + 1)


def buildtree(input_atoms, u, v, j, e, e0, rs, beta=1):
    """Subtree of depth j (2**j leapfrog steps of size e in direction v) from
    ``input_atoms``, for slice variable u and start energy e0 (Hoffman &
    Gelman's BuildTree).  ``beta`` is accepted for interface compatibility
    and not used (energies enter in eV, see the module notes).

    Built leaf by leaf: completed halves wait on a stack until their sibling
    of the same size is complete (a binary counter), then the two join.  The
    first leaf or join that ends growth stops the walk; the halves still on
    the stack then join the failed part innermost first, which is the order
    in which the paper's recursion returns.  Returns (minus edge, plus edge,
    proposal, n in slice, s, sum of alpha, number of alpha)."""
    del beta
    waiting = []
    edge = input_atoms
    for _ in range(1 << j):
        node = _leaf(edge, u, v, e, e0)
        edge = node.front(v)
        while node.alive and waiting and waiting[-1].level == node.level:
            node = _join(waiting.pop(), node, v, rs)
        if not node.alive:
            while waiting:
                node = _join(waiting.pop(), node, v, rs)
            return node.as_tuple()
        waiting.append(node)
    return waiting[-1].as_tuple()


class _DualAveraging(object):
    """Step-size adaptation towards a target mean acceptance statistic
    (Hoffman & Gelman section 3.2), with pyIID's indexing (module notes)."""

    def __init__(self, eps0, target, gamma=0.05, t0=10):
        self.mu = np.log(10.0 * eps0)
        self.target, self.gamma, self.t0 = target, gamma, t0
        self.h_bar = 0.0
        self.m = 0

    def update(self, mean_alpha):
        w = 1.0 / (self.m + self.t0)
        self.h_bar = (1.0 - w) * self.h_bar + w * (self.target - mean_alpha)
        eps = float(np.exp(self.mu - np.sqrt(self.m) / self.gamma * self.h_bar))
        self.m += 1
        return eps


class Ensemble(object):
    """Sampler base: keeps the trajectory (initial state first), the seed and
    the run's metadata; ``step`` is the sampler's move."""

    def __init__(self, atoms, restart=None, logfile=None, trajectory=None, seed=None,
                 verbose=False):
        atoms.get_forces()
        atoms.get_potential_energy()
        if seed is None:
            seed = np.random.randint(0, 2 ** 31)
        self.atoms = atoms
        self.verbose = verbose
        self.random_state = np.random.RandomState(seed)
        self.starting_atoms = _copy(atoms)
        self.traj = [_copy(atoms)]
        self.pe = []
        self.metadata = {"seed": seed}
        self.observers = []

    def attach(self, function, *args, **kwargs):
        self.observers.append((function, args, kwargs))

    def call_observers(self):
        for function, args, kwargs in self.observers:
            function(*args, **kwargs)

    def check_eq(self, eq_steps, tol):
        """Whether the recorded potential energies have settled: the moving
        average over ``eq_steps`` samples, past its first ``eq_steps`` values,
        has a summed gradient below ``tol``.  False while too few samples
        exist for that."""
        k = int(eq_steps)
        pe = np.asarray(self.pe, dtype=float)
        if k < 1 or len(pe) < 2 * k + 2:
            return False
        moving = np.convolve(pe, np.full(k, 1.0 / k), mode="valid")
        return bool(np.sum(np.gradient(moving[k:])) < tol)

    def run(self, steps=100000000, eq_steps=None, eq_tol=None, **kwargs):
        self.metadata["planned iterations"] = steps
        try:
            for i in range(steps):
                if eq_steps is not None and self.check_eq(eq_steps, eq_tol):
                    break
                if self.verbose:
                    print("iteration number", i)
                self.step()
        except KeyboardInterrupt:
            print("Interrupted, returning data")
        return self.traj, self.metadata

    def step(self):
        pass


class NUTSCanonicalEnsemble(Ensemble):
    """NUTS at a fixed temperature (K) over a calculator's potential energy
    (interface of pyiid/sim/nuts_hmc.py:91-244)."""

    def __init__(self, atoms, restart=None, logfile=None, trajectory=None, temperature=100,
                 escape_level=13, accept_target=.65, momentum=None, seed=None, verbose=False):
        Ensemble.__init__(self, atoms, restart, logfile, trajectory, seed, verbose)
        self.accept_target = accept_target
        self.temp = temperature
        self.thermal_nrg = self.temp * KB
        self.momentum = momentum
        self.escape_level = escape_level
        self.metadata.update({"samples_total": 0, "accepted_samples": 0})
        self.step_size = self._find_step_size(atoms, self.thermal_nrg)
        self._adapt = _DualAveraging(self.step_size, accept_target)

    # pyIID exposes the adaptation state as attributes
    @property
    def m(self):
        return self._adapt.m

    @property
    def mu(self):
        return self._adapt.mu

    def _draw_momenta(self, atoms, thermal_nrg):
        if self.momentum is None:
            maxwell_boltzmann(atoms, thermal_nrg, self.random_state, force_temp=True)
        else:
            atoms.set_momenta(self.random_state.normal(0, 1, (len(atoms), 3)))

    def _find_step_size(self, input_atoms, thermal_nrg=None):
        """Hoffman & Gelman Algorithm 4: double or halve the step until one
        leapfrog step's acceptance ratio exp(-dH) crosses 1/2."""
        atoms = _copy(input_atoms)
        self._draw_momenta(atoms, thermal_nrg)
        h0 = atoms.get_total_energy()

        def ratio(step):
            with np.errstate(over="ignore"):
                return float(np.exp(h0 - leapfrog(atoms, step).get_total_energy()))

        step = 0.5
        r = ratio(step)
        direction = 1 if r > 0.5 else -1
        while r ** direction > 2.0 ** -direction:
            step *= 2.0 ** direction
            if self.verbose:
                print("trying step size", step)
            if not 1e-7 <= step <= 1e7:
                step = 1.0
                break
            r = ratio(step)
        if self.verbose:
            print("first step size", step)
        return step

    def step(self):
        """One NUTS iteration (Hoffman & Gelman Algorithm 3 with progressive
        sampling): fresh momenta and slice variable, then doublings in random
        directions until a subtree stops or the trajectory turns back; every
        doubling may move the sample to its subtree's proposal, which is then
        appended to ``traj``.  Returns the accepted states, or None."""
        rs = self.random_state
        start = self.traj[-1]
        if self.verbose:
            print("\tstep size", self.step_size / FS, "fs")
        self._draw_momenta(start, self.thermal_nrg)
        u = rs.uniform(0, 1)
        e0 = start.get_total_energy()
        minus, plus = _copy(start), _copy(start)
        n_total, depth, growing = 1, 0, True
        accepted = []
        alpha_sum, n_alpha = 0.0, 1
        while growing:
            v = rs.choice([-1, 1])
            edge = minus if v == -1 else plus
            sub_minus, sub_plus, proposal, n_sub, s_sub, alpha_sum, n_alpha = buildtree(
                edge, u, v, depth, self.step_size, e0, rs, 1.0 / self.thermal_nrg)
            if v == -1:
                minus = sub_minus
            else:
                plus = sub_plus
            if s_sub == 1 and rs.uniform() < min(1.0, n_sub / float(n_total)):
                self._accept(proposal)
                accepted.append(proposal)
            n_total += n_sub
            growing = s_sub == 1 and not _turned_back(minus, plus)
            depth += 1
            if self.verbose:
                print("\t\tdepth", depth, "samples", 2 ** depth)
            self.metadata["samples_total"] += 2 ** depth
            if depth >= self.escape_level:
                if self.verbose:
                    print("\t\t\tdepth limit", depth)
                growing = False
        self.step_size = self._adapt.update(alpha_sum / n_alpha)
        return accepted or None

    def _accept(self, state):
        self.traj.append(state)
        self.metadata["accepted_samples"] += 1
        state.get_forces()
        self.pe.append(state.get_potential_energy())
        self.call_observers()
