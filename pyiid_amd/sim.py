"""Samplers driven by the MI355X Calc1D.

* :func:`leapfrog` (velocity Verlet on an atoms object with an attached
  calculator) and :func:`simulate_dynamics`, :func:`mh_accept`,
  :func:`hmc_move`, :func:`run_hmc` keep the interfaces and behaviour of
  ``pyiid/sim/__init__.py:10-38`` and ``pyiid/sim/hmc.py:33-209``;
* :class:`DeviceHMC` runs the same Hamiltonian Monte Carlo for a batch of
  independent structures entirely on the device (``pyiid_hmc_setup`` /
  ``pyiid_hmc_run``): every force evaluation, leapfrog update and Metropolis
  decision stays in HBM, and one proposal is replayed as a HIP graph.
"""
from copy import deepcopy as dc

import numpy as np

from . import _lib
from .formfactors import MASSES, element_table

__all__ = ["leapfrog", "simulate_dynamics", "mh_accept", "hmc_move", "run_hmc", "DeviceHMC"]


def _kick(atoms, h, velocities=False):
    """Momenta (or, with ``velocities``, velocities) += h F."""
    f = atoms.get_forces()
    if velocities:
        atoms.set_velocities(atoms.get_velocities() + h * f)
    else:
        atoms.set_momenta(atoms.get_momenta() + h * f)


def _drift(atoms, h, by_momenta=False):
    """Positions += h v (or h p with ``by_momenta``)."""
    rate = atoms.get_momenta() if by_momenta else atoms.get_velocities()
    atoms.set_positions(atoms.get_positions() + h * rate)


def leapfrog(atoms, step, center=True):
    """One velocity-Verlet step of size ``step`` on a copy of ``atoms``: half
    kick, drift by the velocities, half kick at the new positions, then
    optionally re-centre (interface of pyiid/sim/__init__.py:10-38).  The
    input is left unchanged."""
    moved = atoms.copy() if hasattr(atoms, "copy") else dc(atoms)
    _kick(moved, 0.5 * step)
    _drift(moved, step)
    _kick(moved, 0.5 * step)
    if center:
        moved.center()
    return moved


def simulate_dynamics(atoms, stepsize, n_steps):
    """Move ``atoms`` in place along the Hamiltonian flow for n_steps + 1
    position updates (pyiid/sim/hmc.py:33-63).  pyIID's schedule is kept as
    it is: the opening and closing half steps add F to the velocities (the
    force taken as an acceleration) and the first drift moves by the
    velocities, while the n_steps inner steps add F to the momenta and move
    by the momenta.  With unit masses this is velocity Verlet."""
    h = float(stepsize)
    _kick(atoms, 0.5 * h, velocities=True)
    _drift(atoms, h)
    for _ in range(int(n_steps)):
        _kick(atoms, h)
        _drift(atoms, h, by_momenta=True)
    _kick(atoms, 0.5 * h, velocities=True)


def mh_accept(initial_energy, next_energy, temp=1, rs=np.random):
    """Metropolis decision on total energies (pyiid/sim/hmc.py:66-92): True
    with probability min(1, exp((H0 - H1) temp)); ``temp`` multiplies the
    energy difference as in pyIID.  Returns a 1-element boolean array."""
    with np.errstate(over="ignore"):
        return rs.random_sample((1,)) < np.exp((initial_energy - next_energy) * temp)


def hmc_move(atoms, stepsize, n_steps, temp, rs=np.random):
    """One HMC proposal (pyiid/sim/hmc.py:95-130): velocities drawn as
    N(0, 1) / (3 N) per component (pyIID's choice), dynamics on a copy, then
    the Metropolis decision.  Returns (accepted, state to continue from)."""
    atoms.set_velocities(rs.normal(0, 1, (len(atoms), 3)) / 3 / len(atoms))
    trial = atoms.copy()
    simulate_dynamics(trial, stepsize, n_steps)
    if mh_accept(atoms.get_total_energy(), trial.get_total_energy(), temp, rs):
        return True, trial
    return False, atoms


class _AcceptanceStep(object):
    """pyIID's step-size rule for run_hmc: a running acceptance rate, relaxed
    towards the mean of every decision so far with time constant
    ``slowness``, grows the step by ``inc`` above ``target`` and shrinks it by
    ``dec`` otherwise, within [lo, hi]."""

    def __init__(self, step, rate, slowness, target, inc, dec, lo, hi):
        self.step, self.rate, self.slowness, self.target = step, rate, slowness, target
        self.inc, self.dec, self.lo, self.hi = inc, dec, lo, hi
        self.n_seen = 0
        self.n_accepted = 0

    def record(self, accepted):
        self.n_seen += 1
        self.n_accepted += int(bool(accepted))
        mean = self.n_accepted / float(self.n_seen)
        self.rate = self.slowness * self.rate + (1.0 - self.slowness) * mean
        self.step *= self.inc if self.rate > self.target else self.dec
        self.step = min(max(self.step, self.lo), self.hi)
        return self.step


def run_hmc(atoms, iterations, stepsize, n_steps, avg_acceptance_slowness, avg_acceptance_rate,
            target_acceptance_rate, stepsize_inc, stepsize_dec, stepsize_min, stepsize_max,
            temp=1, wtraj=None, rs=np.random):
    """HMC with an adapted step size (interface of pyiid/sim/hmc.py:133-209).
    Returns (trajectory: the start then every accepted state, the decision of
    every iteration, the iterations that accepted)."""
    adapt = _AcceptanceStep(stepsize, avg_acceptance_rate, avg_acceptance_slowness,
                            target_acceptance_rate, stepsize_inc, stepsize_dec, stepsize_min,
                            stepsize_max)
    traj, decisions, accepted_at = [atoms], [], []
    state = atoms
    try:
        for it in range(int(iterations)):
            ok, state = hmc_move(state, adapt.step, n_steps, temp, rs)
            decisions.append(ok)
            if ok:
                traj.append(state)
                accepted_at.append(it)
                if wtraj is not None:
                    wtraj.write(state)
            adapt.record(ok)
    except KeyboardInterrupt:
        pass
    return traj, decisions, accepted_at


class DeviceHMC(object):
    """Batched device-resident HMC against PDF targets.

    Parameters
    ----------
    scat: ElasticScatter
        Supplies the experiment (grids, transform).
    structures: list of (positions (N_b, 3), atomic numbers (N_b,))
        Independent chains (uneven sizes allowed).
    targets: (B, R) array
        Target G(r) of each chain on ``scat.get_r()``.
    potential: 'rw' or 'chi_sq'; conv: energy = conv * potential.
    masses: None (ASE atomic masses in amu) or 'unit'.

    Energies are conv * Rw; forces are the exact -dE/dq.
    """

    def __init__(self, scat, structures, targets, potential="rw", conv=1.0, masses=None,
                 device=0):
        self.scat = scat
        self.ctx = _lib.Context(device)
        self.B = len(structures)
        pos = [np.asarray(p, np.float64) for p, _ in structures]
        numbers = np.concatenate([np.asarray(z) for _, z in structures])
        self.sizes = np.array([len(p) for p in pos])
        self.offsets = self.ctx.set_batch(pos)
        nq = int(np.floor(scat.exp["qmax"] / scat.pdf_qbin))
        table, index, self.elements = element_table(numbers, scat.pdf_qbin, nq,
                                                     scat.form_factor_source)
        self.ctx.set_elements(index)
        self.ctx.set_scatter(_lib.SLOT_PDF, table)
        r = scat.get_r()
        self.ctx.set_transform(scat.exp["rstep"], scat.pdf_qbin, r, scat.exp["qmin"], nq)
        if masses is None:
            m = np.array([MASSES[int(z)] for z in numbers], np.float64)
        elif masses == "unit":
            m = np.ones(len(numbers))
        else:
            m = np.asarray(masses, np.float64)
        self.masses = m
        self.pot = _lib.POT_RW if potential == "rw" else _lib.POT_CHI_SQ
        self.conv = conv
        targets = np.ascontiguousarray(targets, np.float64).reshape(self.B, len(r))
        self.ctx.hmc_setup(_lib.SLOT_PDF, scat.pdf_qbin, targets, self.pot, conv, m)

    def run(self, n_proposals, n_leapfrog=10, step=2e-3, beta=1.0, seed=0, use_graph=True,
            times=False):
        """Returns (energies, accepted), each (n_proposals, B), plus the
        proposals' completion times in ms when ``times``."""
        return self.ctx.hmc_run(n_proposals, n_leapfrog, step, beta, seed, self.B, use_graph,
                                times)

    def positions(self):
        """Current positions of every chain, as a list of (N_b, 3) float32 arrays."""
        allp = self.ctx.get_positions(int(self.offsets[-1]))
        return [allp[self.offsets[b]:self.offsets[b + 1]] for b in range(self.B)]

    def close(self):
        self.ctx.close()


class SurrogateHMC(DeviceHMC):
    """DeviceHMC whose proposals follow the learned integrator's forces.

    Metropolis still uses the exact energy (conv * Rw from the Debye path), so
    the stationary distribution is that of exact-force HMC.  The network's
    output layer is refitted every ``train_every`` proposals on the exact
    forces at the chains' current states and on the exact energy change since
    the previous fit, weighted by ``energy_weight`` (A^-2; the default 0
    fits forces only; pyiid_amd.surrogate)."""

    def __init__(self, scat, structures, targets, potential="rw", conv=1.0, masses=None,
                 device=0, net=None, net_seed=0, max_nbr=None, energy_weight=0.0):
        from . import surrogate
        DeviceHMC.__init__(self, scat, structures, targets, potential, conv, masses, device)
        self.net = net if net is not None else surrogate.SurrogateNet(net_seed)
        self.max_nbr = surrogate.MAX_NBR if max_nbr is None else int(max_nbr)
        self.net.normalise(self.ctx, int(self.offsets[-1]), self.max_nbr)
        self.ctx.sur_set_energy_weight(energy_weight)

    def output_weights(self):
        """The output layer in use, (B, 64) float32."""
        return self.ctx.sur_get_output(self.B)

    def feature_energies(self):
        """Chain energies of the 64 output units at the current state, (B, 64)."""
        return self.ctx.sur_feature_energies(self.B)

    def train(self, forget=0.0, ridge=1e-6):
        """One fit at the current state; returns the per-chain r2."""
        return self.ctx.sur_train(self.B, forget, ridge)

    def run(self, n_proposals, n_leapfrog=10, step=2e-3, beta=1.0, seed=0, use_graph=True,
            train_every=10, forget=0.5, ridge=1e-6, min_r2=-np.inf, min_acceptance=0.0):
        """Returns a dict of energies / accepted (n_proposals, B), exact_mode
        (proposals run with exact forces), times_ms, violations.  A block of
        train_every proposals runs with exact forces when the fit before it
        has mean r2 < min_r2 or the learned-force block before it accepted
        less often than min_acceptance; with train_every > 3 a block whose
        first three learned-force proposals fall short of min_acceptance
        finishes with exact forces."""
        return self.ctx.sur_hmc_run(n_proposals, n_leapfrog, step, beta, seed, self.B, use_graph,
                                    train_every, forget, ridge, min_r2, min_acceptance)

    def run_exact(self, n_proposals, n_leapfrog=10, step=2e-3, beta=1.0, seed=0, use_graph=True,
                  times=False):
        return DeviceHMC.run(self, n_proposals, n_leapfrog, step, beta, seed, use_graph, times)

    def forces(self):
        return self.ctx.sur_forces(int(self.offsets[-1]))
