"""No-U-Turn sampler driving the MI355X Calc1D (host side).

Mirrors the ensemble interface of pyiid/sim/__init__.py:41-82 (``Ensemble``:
``run`` returns ``(traj, metadata)``, ``traj`` starts with the initial
configuration) and the NUTS variant of pyiid/sim/nuts_hmc.py:15-232:

* the tree is built recursively from velocity-Verlet steps (``leapfrog``,
  pyiid/sim/__init__.py:10-38) of size +-step; a leaf is inside the slice when
  u <= exp(E0 - E_total(leaf)) and the tree stops growing when
  u >= exp(Emax + E0 - E_total) (Emax = 200) or the trajectory turns back on
  itself (span . v < 0 at either end);
* the slice variable u ~ U(0, 1) is in post-exponential units, momenta are
  Maxwell-Boltzmann at the ensemble's thermal energy (kB T) with the kinetic
  energy forced to exactly 3/2 N kB T;
* the step size starts from the doubling / halving search of
  ``_find_step_size`` (nuts_hmc.py:116-158) and adapts by dual averaging
  towards ``accept_target`` (gamma = 0.05, t0 = 10, mu = log(10 step)).

Random numbers come from one ``numpy.random.RandomState(seed)`` in the order
the reference draws them; the reference draws its Maxwell-Boltzmann momenta
from numpy's global generator (through ASE), this module from the same
RandomState, so a seeded run is reproducible.  Accepted samples' potential
energies are appended to ``pe`` (the reference leaves it empty), which the
``eq_steps`` stopping test of ``run`` reads.
"""
from copy import deepcopy as dc

import numpy as np

from .atoms import Atoms
from .sim import leapfrog

__all__ = ["Ensemble", "NUTSCanonicalEnsemble", "buildtree", "maxwell_boltzmann", "KB", "FS"]

KB = 8.617330337217213e-05      # Boltzmann constant, eV / K
FS = 0.09822694788464063         # 1 fs in Angstrom sqrt(amu / eV) time units
EMAX = 200.0


def _copy(atoms):
    # this package's Atoms share their calculator (and its device context)
    # through copy(); other atoms objects are deep-copied as the reference does
    return atoms.copy() if isinstance(atoms, Atoms) else dc(atoms)


def maxwell_boltzmann(atoms, thermal_energy, rs, force_temp=True):
    """Momenta p ~ N(0, sqrt(m kT)) per component; with ``force_temp`` the
    kinetic energy is rescaled to exactly 3/2 N kT."""
    m = atoms.get_masses()
    p = rs.normal(0.0, 1.0, (len(atoms), 3)) * np.sqrt(m * thermal_energy)[:, None]
    atoms.set_momenta(p)
    if force_temp and len(atoms):
        ke = atoms.get_kinetic_energy()
        if ke > 0:
            atoms.set_momenta(p * np.sqrt(1.5 * len(atoms) * thermal_energy / ke))


def buildtree(input_atoms, u, v, j, e, e0, rs, beta=1):
    """One subtree of depth j in direction v (nuts_hmc.py:15-88).

    Returns (leftmost, rightmost, proposal, n, s, alpha sum, alpha count)."""
    if j == 0:
        atoms_prime = leapfrog(input_atoms, v * e)
        neg_delta = e0 - atoms_prime.get_total_energy()
        with np.errstate(over="ignore"):
            exp1 = np.exp(neg_delta)
            exp2 = np.exp(EMAX + neg_delta)
            alpha = min(1.0, np.exp(-atoms_prime.get_total_energy()
                                    + input_atoms.get_total_energy()))
        n_prime = int(u <= exp1)
        s_prime = int(u < exp2)
        return atoms_prime, atoms_prime, atoms_prime, n_prime, s_prime, alpha, 1
    neg, pos, prime, n_p, s_p, a_p, na_p = buildtree(input_atoms, u, v, j - 1, e, e0, rs, beta)
    if s_p == 1:
        if v == -1:
            neg, _, pp, n_pp, s_pp, a_pp, na_pp = buildtree(neg, u, v, j - 1, e, e0, rs, beta)
        else:
            _, pos, pp, n_pp, s_pp, a_pp, na_pp = buildtree(pos, u, v, j - 1, e, e0, rs, beta)
        if rs.uniform() < float(n_pp / max(n_p + n_pp, 1)):
            prime = pp
        a_p += a_pp
        na_p += na_pp
        span = (pos.positions - neg.positions).ravel()
        s_p = s_pp * int(span.dot(neg.get_velocities().ravel()) >= 0) * \
            int(span.dot(pos.get_velocities().ravel()) >= 0)
        n_p = n_p + n_pp
    return neg, pos, prime, n_p, s_p, a_p, na_p


class Ensemble(object):
    """Sampler base (pyiid/sim/__init__.py:41-82)."""

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
        ret = np.cumsum(self.pe, dtype=float)
        ret[eq_steps:] = ret[eq_steps:] - ret[:-eq_steps]
        ret = ret[eq_steps - 1:] / eq_steps
        return np.sum(np.gradient(ret[eq_steps:])) < tol

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
    """nuts_hmc.py:91-244."""

    def __init__(self, atoms, restart=None, logfile=None, trajectory=None, temperature=100,
                 escape_level=13, accept_target=.65, momentum=None, seed=None, verbose=False):
        Ensemble.__init__(self, atoms, restart, logfile, trajectory, seed, verbose)
        self.accept_target = accept_target
        self.temp = temperature
        self.thermal_nrg = self.temp * KB
        self.momentum = momentum
        self.step_size = self._find_step_size(atoms, self.thermal_nrg)
        self.mu = np.log(10 * self.step_size)
        self.sim_hbar = 0
        self.gamma = 0.05
        self.t0 = 10
        self.metadata.update({"samples_total": 0, "accepted_samples": 0})
        self.escape_level = escape_level
        self.m = 0

    def _draw_momenta(self, atoms, thermal_nrg):
        if self.momentum is None:
            maxwell_boltzmann(atoms, thermal_nrg, self.random_state, force_temp=True)
        else:
            atoms.set_momenta(self.random_state.normal(0, 1, (len(atoms), 3)))

    def _find_step_size(self, input_atoms, thermal_nrg=None):
        atoms = _copy(input_atoms)
        step_size = .5
        self._draw_momenta(atoms, thermal_nrg)

        def ratio(step):
            with np.errstate(over="ignore"):
                return np.exp(-leapfrog(atoms, step).get_total_energy()
                              + atoms.get_total_energy())

        r = ratio(step_size)
        a = 2 * int(r > 0.5) - 1
        while r ** a > 2.0 ** -a:
            step_size *= 2.0 ** a
            if self.verbose:
                print("trying step size", step_size)
            r = ratio(step_size)
            if step_size < 1e-7 or step_size > 1e7:
                step_size = 1.
                break
        if self.verbose:
            print("optimal step size", step_size)
        return step_size

    def step(self):
        new_configurations = []
        if self.verbose:
            print("\ttime step size", self.step_size / FS, "fs")
        self._draw_momenta(self.traj[-1], self.thermal_nrg)
        u = self.random_state.uniform(0, 1)
        e0 = self.traj[-1].get_total_energy()
        e = self.step_size
        n, s, j = 1, 1, 0
        neg = _copy(self.traj[-1])
        pos = _copy(self.traj[-1])
        a, na = 0.0, 1
        while s == 1:
            v = self.random_state.choice([-1, 1])
            if v == -1:
                neg, _, prime, n_p, s_p, a, na = buildtree(neg, u, v, j, e, e0,
                                                           self.random_state,
                                                           1 / self.thermal_nrg)
            else:
                _, pos, prime, n_p, s_p, a, na = buildtree(pos, u, v, j, e, e0,
                                                           self.random_state,
                                                           1 / self.thermal_nrg)
            if s_p == 1 and self.random_state.uniform() < min(1, n_p * 1. / n):
                self.traj += [prime]
                self.metadata["accepted_samples"] += 1
                new_configurations.append(prime)
                prime.get_forces()
                self.pe.append(prime.get_potential_energy())
                self.call_observers()
            n = n + n_p
            span = (pos.positions - neg.positions).ravel()
            s = s_p * int(span.dot(neg.get_velocities().ravel()) >= 0) * \
                int(span.dot(pos.get_velocities().ravel()) >= 0)
            j += 1
            if self.verbose:
                print("\t \tdepth", j, "samples", 2 ** j)
            self.metadata["samples_total"] += 2 ** j
            if j >= self.escape_level:
                if self.verbose:
                    print("\t \t \tjmax emergency escape at {}".format(j))
                s = 0
        w = 1. / (self.m + self.t0)
        self.sim_hbar = (1 - w) * self.sim_hbar + w * (self.accept_target - a / na)
        self.step_size = np.exp(self.mu - (self.m ** .5 / self.gamma) * self.sim_hbar)
        self.m += 1
        return new_configurations or None
