"""Rw / chi^2 potentials and the ``Calc1D`` calculator on the MI355X backend.

Mirrors ``pyiid/calc/__init__.py:10-105`` and ``pyiid/calc/calc_1d.py:9-101``.
The scale, Rw / chi^2 sums and the gradient contraction run on the device in
FP64.  When ``Calc1D`` is given the getters of an :class:`ElasticScatter`
(``get_pdf``/``get_grad_pdf`` or ``get_fq``/``get_grad_fq``), energy and forces
come from one fused device pass that never materialises grad G (the adjoint
u = T^T dP/dG is contracted inside the Debye gradient kernel); the values
are the reference's, including its force convention (forces = conv * grad P
built from the -1/2 dF/dq array, i.e. half of -dE/dq).
"""
import numpy as np

from . import _lib
from . import processor as _proc

__all__ = ["wrap_rw", "wrap_chi_sq", "wrap_grad_rw", "wrap_grad_chi_sq", "Calc1D"]


def _ctx():
    return _proc.state().ctx


def wrap_rw(gcalc, gobs):
    """(Rw, scale); scale <= 0 gives (1, 1) (master_kernel.py:206-236)."""
    st = _proc.state()
    with st.lock:
        return st.ctx.potential(_lib.POT_RW, gobs, gcalc)


def wrap_chi_sq(gcalc, gobs):
    """(chi^2, scale); scale <= 0 is replaced by 1 (master_kernel.py:239-266)."""
    st = _proc.state()
    with st.lock:
        return st.ctx.potential(_lib.POT_CHI_SQ, gobs, gcalc)


def wrap_grad_rw(grad_gcalc, gcalc, gobs):
    """d Rw / d q for every row of grad_gcalc (calc/__init__.py:56-79)."""
    st = _proc.state()
    with st.lock:
        return st.ctx.potential_grad(_lib.POT_RW, grad_gcalc, gobs, gcalc)


def wrap_grad_chi_sq(grad_gcalc, gcalc, gobs):
    """d chi^2 / d q for every row of grad_gcalc (calc/__init__.py:82-105)."""
    st = _proc.state()
    with st.lock:
        return st.ctx.potential_grad(_lib.POT_CHI_SQ, grad_gcalc, gobs, gcalc)


class Calculator(object):
    """The part of ``ase.calculators.calculator.Calculator`` that Calc1D uses:
    a results cache keyed on the atoms' positions and numbers."""

    implemented_properties = []

    def __init__(self, restart=None, ignore_bad_restart_file=False, label=None, atoms=None,
                 **kwargs):
        self.results = {}
        self.atoms = None
        self.parameters = dict(kwargs)

    def check_state(self, atoms):
        if self.atoms is None:
            return ["positions", "numbers"]
        changes = []
        if len(atoms) != len(self.atoms) or not np.array_equal(atoms.numbers, self.atoms.numbers):
            changes.append("numbers")
        if len(atoms) != len(self.atoms) or not np.array_equal(atoms.positions,
                                                               self.atoms.positions):
            changes.append("positions")
        return changes

    def calculate(self, atoms=None, properties=["energy"], system_changes=["positions"]):
        if atoms is not None:
            self.atoms = atoms.copy()
            self.atoms.calc = None

    def get_property(self, name, atoms):
        changes = self.check_state(atoms)
        if changes:
            self.results = {}
        if name not in self.results:
            self.calculate(atoms, [name], changes)
        return self.results[name]

    def get_potential_energy(self, atoms=None, force_consistent=False):
        return self.get_property("energy", atoms)

    def get_forces(self, atoms=None):
        return self.get_property("forces", atoms).copy()


class Calc1D(Calculator):
    """PDF / F(Q) based Rw or chi^2 calculator (calc_1d.py:9-101)."""

    implemented_properties = ["energy", "forces"]

    def __init__(self, restart=None, ignore_bad_restart_file=False, label=None, atoms=None,
                 target_data=None, exp_function=None, exp_grad_function=None, conv=1.,
                 potential="rw", **kwargs):
        Calculator.__init__(self, restart, ignore_bad_restart_file, label, atoms, **kwargs)
        if target_data is None or len(target_data.shape) != 1:
            raise NotImplementedError("Need a 1d array target data set")
        if exp_function is None or exp_grad_function is None:
            raise NotImplementedError("Need functions which return the simulated data "
                                      "associated with the experiment and its gradient")
        self.target_data = target_data
        self.exp_function = exp_function
        self.exp_grad_function = exp_grad_function
        self.scale = 1
        self.rw_to_eV = conv
        if potential == "chi_sq":
            self.potential = wrap_chi_sq
            self.grad = wrap_grad_chi_sq
            self._pot = _lib.POT_CHI_SQ
        elif potential == "rw":
            self.potential = wrap_rw
            self.grad = wrap_grad_rw
            self._pot = _lib.POT_RW
        else:
            raise NotImplementedError("Potential not implemented")

    def calculate(self, atoms=None, properties=["energy"],
                  system_changes=["positions", "numbers", "cell", "pbc", "charges", "magmoms"]):
        Calculator.calculate(self, atoms, properties, system_changes)
        if len(system_changes) > 0:
            if "energy" in properties:
                self.calculate_energy(self.atoms)
            if "forces" in properties:
                self.calculate_forces(self.atoms)
        for prop in properties:
            if prop not in self.results:
                if prop == "energy":
                    self.calculate_energy(self.atoms)
                if prop == "forces":
                    self.calculate_forces(self.atoms)

    # fused device path -------------------------------------------------
    def _fused(self, atoms, want_forces):
        scat = getattr(self.exp_function, "__self__", None)
        gscat = getattr(self.exp_grad_function, "__self__", None)
        if scat is None or scat is not gscat or not hasattr(scat, "pdf_qbin"):
            return None
        fname = getattr(self.exp_function, "__name__", "")
        gname = getattr(self.exp_grad_function, "__name__", "")
        if (fname, gname) not in (("get_pdf", "get_grad_pdf"), ("get_fq", "get_grad_fq")):
            return None
        if getattr(scat, "fq", None) is not _proc.wrap_fq:
            return None
        scat._ensure(atoms)
        gobs = np.asarray(self.target_data, dtype=np.float64)
        if fname == "get_pdf":
            if len(gobs) != len(scat.get_r()):
                return None
            with _proc.loaded(atoms, "PDF") as (st, slot, nq):
                st.load_transform(scat.exp["rstep"], scat.pdf_qbin, scat.get_r(), scat.exp["qmin"],
                                  nq)
                return st.ctx.pdf_energy_forces(slot, scat.pdf_qbin, gobs, self._pot,
                                                self.rw_to_eV, len(atoms), want_forces)
        with _proc.loaded(atoms, "fq") as (st, slot, nq):
            first = int(np.floor(scat.exp["qmin"] / scat.exp["qbin"]))
            if len(gobs) != nq - first:
                return None
            return st.ctx.fq_energy_forces(slot, scat.exp["qbin"], first, gobs, self._pot,
                                           self.rw_to_eV, len(atoms), want_forces)

    def calculate_energy(self, atoms):
        """energy = P(exp_function(atoms), target) * conv (calc_1d.py:78-87)."""
        fused = self._fused(atoms, False)
        if fused is not None:
            self.results["energy"] = fused[0]
            return
        energy, scale = self.potential(self.exp_function(atoms), self.target_data)
        self.scale = scale
        self.results["energy"] = energy * self.rw_to_eV

    def calculate_forces(self, atoms):
        """forces = grad P(exp_grad_function, exp_function, target) * conv
        (calc_1d.py:89-95)."""
        fused = self._fused(atoms, True)
        if fused is not None:
            self.results["forces"] = fused[1]
            return
        forces = self.grad(self.exp_grad_function(atoms), self.exp_function(atoms),
                           self.target_data) * self.rw_to_eV
        self.results["forces"] = forces
