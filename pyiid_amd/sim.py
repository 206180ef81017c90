"""Samplers driven by the MI355X Calc1D.

* :func:`leapfrog` mirrors ``pyiid/sim/__init__.py:10-38`` (velocity Verlet on
  an atoms object with an attached calculator);
* :func:`simulate_dynamics`, :func:`mh_accept`, :func:`hmc_move`,
  :func:`run_hmc` mirror ``pyiid/sim/hmc.py:10-209``;
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


def leapfrog(atoms, step, center=True):
    """One velocity-Verlet step on a copy of ``atoms`` (sim/__init__.py:10-38):
    p += step/2 F; x += step v; p += step/2 F(x'); optional centring."""
    latoms = dc(atoms) if not hasattr(atoms, "copy") else atoms.copy()
    latoms.set_momenta(latoms.get_momenta() + 0.5 * step * latoms.get_forces())
    latoms.positions += step * latoms.get_velocities()
    latoms.set_momenta(latoms.get_momenta() + 0.5 * step * latoms.get_forces())
    if center:
        latoms.center()
    return latoms


def _lf(atoms, step):
    """The inner step of hmc.py:10-30 (kick with the full step, then drift by
    the momenta)."""
    f = atoms.get_forces()
    atoms.set_momenta(atoms.get_momenta() + step * f)
    atoms.set_positions(atoms.get_positions() + step * atoms.get_momenta())


def simulate_dynamics(atoms, stepsize, n_steps):
    """hmc.py:33-63."""
    grad = atoms.get_forces()
    atoms.set_velocities(atoms.get_velocities() + 0.5 * stepsize * grad)
    atoms.set_positions(atoms.get_positions() + stepsize * atoms.get_velocities())
    for _ in range(n_steps):
        _lf(atoms, stepsize)
    atoms.set_velocities(atoms.get_velocities() + 0.5 * stepsize * atoms.get_forces())


def mh_accept(initial_energy, next_energy, temp=1, rs=np.random):
    """Metropolis rule of hmc.py:66-92: accept when u < exp((H0 - H1) temp)."""
    diff = initial_energy - next_energy
    return rs.random_sample((1,)) < np.exp(diff * temp)


def hmc_move(atoms, stepsize, n_steps, temp, rs=np.random):
    """hmc.py:95-130 (velocities N(0,1)/3/N, as the reference draws them)."""
    atoms.set_velocities(rs.normal(0, 1, (len(atoms), 3)) / 3 / len(atoms))
    prop = atoms.copy()
    simulate_dynamics(prop, stepsize, n_steps)
    accept = mh_accept(atoms.get_total_energy(), prop.get_total_energy(), temp, rs)
    if accept:
        return True, prop
    return False, atoms


def run_hmc(atoms, iterations, stepsize, n_steps, avg_acceptance_slowness, avg_acceptance_rate,
            target_acceptance_rate, stepsize_inc, stepsize_dec, stepsize_min, stepsize_max,
            temp=1, wtraj=None, rs=np.random):
    """Adaptive-step HMC of hmc.py:133-209; returns (traj, accept_list, accept_num)."""
    traj = [atoms]
    accept_list, accept_num = [], []
    i = 0
    while i < iterations:
        accept, atoms = hmc_move(atoms, stepsize, n_steps, temp, rs)
        if accept is True:
            traj += [atoms]
            accept_num.append(i)
            if wtraj is not None:
                wtraj.write(atoms)
        accept_list.append(accept)
        avg_acceptance_rate = avg_acceptance_slowness * avg_acceptance_rate + \
            (1.0 - avg_acceptance_slowness) * np.average(np.array(accept_list))
        stepsize *= stepsize_inc if avg_acceptance_rate > target_acceptance_rate else stepsize_dec
        stepsize = min(max(stepsize, stepsize_min), stepsize_max)
        i += 1
    return traj, accept_list, accept_num


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
        table, index, self.elements = element_table(numbers, scat.pdf_qbin, nq)
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
    output layer is refitted on exact forces at the chains' current states
    every ``train_every`` proposals (pyiid_amd.surrogate)."""

    def __init__(self, scat, structures, targets, potential="rw", conv=1.0, masses=None,
                 device=0, net=None, net_seed=0, max_nbr=None):
        from . import surrogate
        DeviceHMC.__init__(self, scat, structures, targets, potential, conv, masses, device)
        self.net = net if net is not None else surrogate.SurrogateNet(net_seed)
        self.max_nbr = surrogate.MAX_NBR if max_nbr is None else int(max_nbr)
        self.net.normalise(self.ctx, int(self.offsets[-1]), self.max_nbr)

    def train(self, forget=0.0, ridge=1e-6):
        """One fit at the current state; returns the per-chain r2."""
        return self.ctx.sur_train(self.B, forget, ridge)

    def run(self, n_proposals, n_leapfrog=10, step=2e-3, beta=1.0, seed=0, use_graph=True,
            train_every=10, forget=0.5, ridge=1e-6, min_r2=-np.inf):
        """Returns a dict of energies / accepted (n_proposals, B), exact_mode
        (proposals run with exact forces), times_ms, violations."""
        return self.ctx.sur_hmc_run(n_proposals, n_leapfrog, step, beta, seed, self.B, use_graph,
                                    train_every, forget, ridge, min_r2)

    def run_exact(self, n_proposals, n_leapfrog=10, step=2e-3, beta=1.0, seed=0, use_graph=True,
                  times=False):
        return DeviceHMC.run(self, n_proposals, n_leapfrog, step, beta, seed, use_graph, times)

    def forces(self):
        return self.ctx.sur_forces(int(self.offsets[-1]))
