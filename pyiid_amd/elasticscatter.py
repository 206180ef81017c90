"""``ElasticScatter``: the experiment object of pyIID's scattering path, bound
to the MI355X processor.

Mirrors ``pyiid/experiments/elasticscatter/__init__.py:52-558`` (same
constructor, experiment dictionary, grids, caching of per-atom scatter arrays
and getters).  Differences, by design:

* the only processor is ``'MI355X'`` (HIP kernels).  ``set_processor`` raises
  if it is unavailable or another processor is requested, instead of silently
  keeping the previous one (``__init__.py:228-292``);
* per-element form factors come from ``xraylib.FF_Rayl`` as in the
  reference when xraylib is installed, and from the Cromer-Mann table
  otherwise (``form_factors=`` "auto", "xraylib" or "cromer-mann";
  :mod:`pyiid_amd.formfactors`);
* with ``compat_formfactors=False`` (default) each element's row is evaluated
  for that element on its own grid.  ``compat_formfactors=True`` reproduces
  the reference's ``_wrap_atoms`` indexing (rows taken for
  ``numbers[0..E-1]``, assigned to set-ordered elements, both tables on the
  ``qbin`` grid: ``__init__.py:127-143``).
"""
import math

import numpy as np

from . import processor as _proc
from .formfactors import element_table, form_factor, resolve_source

all_changes = ["positions", "numbers", "cell", "pbc", "charges", "magmoms", "exp"]


def project_to_detector(q, iq, qbin, pixel_array):
    """Detector image of I(Q): each pixel holds the Q it sees and receives
    I(Q_k) of the bin whose open window (Q_k - qbin/2, Q_k + qbin/2) contains
    it.  Where rounding makes two neighbouring windows overlap, the later bin
    wins.  Pixels outside every window are 0.  These are the semantics of
    ``__init__.py:448-475``, found per pixel by binary search over the
    window edges rather than by a pass over the bins."""
    q = np.asarray(q, dtype=np.float64)
    pixels = np.asarray(pixel_array)
    flat = pixels.ravel()
    half = qbin / 2.
    last = np.searchsorted(q - half, flat, side="left") - 1   # last k with Q_k - qbin/2 < p
    first = np.searchsorted(q + half, flat, side="right")     # first k with Q_k + qbin/2 > p
    hit = (last >= first) & (last >= 0)
    img = np.zeros(flat.shape)
    img[hit] = np.asarray(iq)[last[hit]]
    return img.reshape(pixels.shape)


class ElasticScatter(object):
    """Elastic powder scattering of a set of atoms: F(Q), S(Q), I(Q), G(r) and
    the gradients of F(Q) and G(r)."""

    def __init__(self, exp_dict=None, verbose=False, seed=None, compat_formfactors=False,
                 form_factors="auto"):
        self.verbose = verbose
        self.form_factor_source = resolve_source(form_factors)
        self.wrap_atoms_state = None
        if seed is None:
            self.seed = int(np.random.random() * 2 ** 32)
        elif isinstance(seed, (int, np.integer)):
            self.seed = int(seed)
        else:
            raise ValueError("Expected an integer!")
        self.rs = np.random.RandomState(self.seed)
        self.compat_formfactors = compat_formfactors
        self.avail_pro = ["MI355X"]
        self.exp_dict_keys = ["qmin", "qmax", "qbin", "rmin", "rmax", "rstep", "sampling"]
        self.default_values = [0.0, 25, .1, 0.0, 40.0, .01, "full"]
        self.alg = None
        self.processor = None
        self.exp = None
        self.pdf_qbin = None
        self.update_experiment(exp_dict)
        self.fq = None
        self.grad = None
        self.grad_pdf = None
        self.set_processor()

    # ------------------------------------------------------------------
    def _wrap_atoms(self, atoms):
        """Attach the 'F(Q) scatter' and 'PDF scatter' per-atom arrays
        (``__init__.py:110-151``) and their compact (table, index) form."""
        if "qbin" not in self.exp.keys():
            self.exp["qbin"] = .1
        n = len(atoms)
        numbers = np.asarray(atoms.get_atomic_numbers())
        compact = {}
        for qbin, name in zip([self.exp["qbin"], self.pdf_qbin], ["F(Q) scatter", "PDF scatter"]):
            nq = int(math.floor(self.exp["qmax"] / qbin))
            if self.compat_formfactors:
                e_list = list(set(numbers.tolist()))
                q = np.arange(nq) * self.exp["qbin"]
                rows = [form_factor(numbers[i], q, self.form_factor_source)
                        for i in range(len(e_list))]
                table = np.asarray(rows, dtype=np.float32).reshape(len(e_list), nq)
                index = np.zeros(n, dtype=np.int32)
                for i, z in enumerate(e_list):
                    index[numbers == z] = i
            else:
                table, index, _ = element_table(numbers, qbin, nq, self.form_factor_source)
            per_atom = table[index] if n else np.zeros((0, nq), np.float32)
            if name in atoms.arrays:
                del atoms.arrays[name]
            atoms.set_array(name, per_atom)
            compact[name] = (table, index, atoms.arrays[name])
        atoms.info["exp"] = self.exp
        atoms.info["scatter_atoms"] = n
        atoms.info["pyiid_compact"] = compact

    def _check_wrap_atoms_state(self, atoms):
        """``__init__.py:153-178``."""
        t_value = True
        if self.wrap_atoms_state is None:
            t_value = False
        elif "F(Q) scatter" not in atoms.arrays.keys():
            t_value = False
        elif atoms.info["exp"] != self.exp or atoms.info["scatter_atoms"] != len(atoms):
            t_value = False
        if not t_value:
            if self.verbose:
                print("calculating new scatter factors")
            self._wrap_atoms(atoms)
            self.wrap_atoms_state = atoms
        return t_value

    def update_experiment(self, exp_dict):
        """``__init__.py:180-204`` (fills defaults into the caller's dict)."""
        if exp_dict is None or bool(exp_dict) is False:
            exp_dict = {}
        for key, dv in zip(self.exp_dict_keys, self.default_values):
            if key not in exp_dict.keys():
                exp_dict[key] = dv
        if exp_dict["sampling"] == "ns":
            exp_dict["rstep"] = np.pi / exp_dict["qmax"]
        self.exp = exp_dict
        self.pdf_qbin = np.pi / (self.exp["rmax"] + 6 * 2 * np.pi / self.exp["qmax"])

    def set_processor(self, processor=None, kernel_type="flat"):
        """Bind the MI355X processor (``__init__.py:206-292``).  Returns True;
        raises RuntimeError when no HIP device / library is available and
        ValueError for processors this backend does not provide."""
        if processor is None:
            processor = self.avail_pro[0]
        if processor != "MI355X":
            raise ValueError("processor %r is not provided by this backend (available: %s)"
                             % (processor, self.avail_pro))
        if not _proc.device_available():
            raise RuntimeError("MI355X processor unavailable: no HIP device or "
                               "libpyiid_hip.so not loadable")
        self.fq = _proc.wrap_fq
        self.grad = _proc.wrap_fq_grad
        self.grad_pdf = _proc.grad_pdf
        self.processor = processor
        self.alg = "flat"
        return True

    def check_wrap_atoms_state(self, atoms):
        """``__init__.py:294-302``."""
        if self.wrap_atoms_state is None:
            return False
        if "F(Q) scatter" not in atoms.arrays.keys():
            return False
        if atoms.info["exp"] != self.exp or atoms.info["scatter_atoms"] != len(atoms):
            return False
        return True

    def _ensure(self, atoms):
        if self.check_wrap_atoms_state(atoms) is False:
            if self.verbose:
                print("calculating new scatter factors")
            self._wrap_atoms(atoms)
            self.wrap_atoms_state = atoms

    # ------------------------------------------------------------------
    def get_fq(self, atoms, iq_std=None, noise_distribution=None):
        """Reduced structure factor F(Q) from qmin (``__init__.py:304-341``)."""
        self._ensure(atoms)
        fq = self.fq(atoms, self.exp["qbin"])
        first = int(np.floor(self.exp["qmin"] / self.exp["qbin"]))
        fq = fq[first:]
        if iq_std is not None:
            # the I(Q) width carried to F = Q (I / <f>^2 - 1): sigma_F = sigma_I Q / <f>^2
            mean_f = atoms.get_array("F(Q) scatter").mean(axis=0)
            width = iq_std * np.abs(self.get_scatter_vector()) / (mean_f * mean_f)[first:]
            fq += self._noise(width)
        return fq

    def _noise(self, width):
        """One Gaussian draw per bin with the given widths, from the
        experiment's seeded generator.  A zero width in the first bin (Q = 0)
        becomes 1e-9, as in the reference (``__init__.py:337-338,378-379``)."""
        width = np.array(width, dtype=np.float64)
        if width[0] == 0.0:
            width[0] = 1e-9
        return self.rs.normal(0.0, width)

    def get_pdf(self, atoms, iq_std=None, noise_distribution=np.random.normal):
        """Atomic pair distribution function G(r) (``__init__.py:343-391``)."""
        self._ensure(atoms)
        fq = self.fq(atoms, self.pdf_qbin, "PDF")
        if iq_std is not None:
            q = np.abs(self.get_scatter_vector(pdf=True))
            # here the reference divides by <f^2>, not <f>^2 (__init__.py:371-372)
            f2 = atoms.get_array("PDF scatter")
            mean_f2 = (f2 * f2).mean(axis=0)
            if hasattr(iq_std, "shape") and iq_std.shape != q.shape:
                # a width array on another grid is resampled linearly by bin
                # index (the reference uses griddata, which gives NaN past the
                # last point; np.interp holds the end value)
                iq_std = np.interp(np.arange(len(q)), np.arange(len(iq_std)), iq_std)
            fq += self._noise(iq_std * q / mean_f2)
        return _proc.pdf_at_qmin(fq, self.exp["rstep"], self.pdf_qbin, self.get_r(),
                                 self.exp["qmin"])

    def get_sq(self, atoms, iq_std=None, noise_distribution=np.random.normal):
        """Structure factor S(Q) = F/Q + 1, inf -> 0 (``__init__.py:393-420``)."""
        fq = self.get_fq(atoms, iq_std, noise_distribution)
        with np.errstate(all="ignore"):
            sq = (fq / self.get_scatter_vector()) + np.ones(self.get_scatter_vector().shape)
        sq[np.isinf(sq)] = 0.
        return sq

    def get_iq(self, atoms, iq_std=None, noise_distribution=np.random.normal):
        """Scattering intensity I(Q) = S(Q) <f>^2 (``__init__.py:422-446``)."""
        sq = self.get_sq(atoms, iq_std, noise_distribution)
        f2 = np.average(atoms.get_array("F(Q) scatter"), axis=0) ** 2
        return sq * f2[int(np.floor(self.exp["qmin"] / self.exp["qbin"])):]

    def get_2d_scatter(self, atoms, pixel_array):
        """I(Q) projected onto a detector map of Q values (``__init__.py:448-475``)."""
        return project_to_detector(self.get_scatter_vector(), self.get_iq(atoms),
                                   self.exp["qbin"], pixel_array)

    def get_grad_fq(self, atoms):
        """Gradient of F(Q), reference convention -1/2 dF/dq (``__init__.py:477-496``)."""
        self._ensure(atoms)
        g = self.grad(atoms, self.exp["qbin"])
        return g[:, :, int(np.floor(self.exp["qmin"] / self.exp["qbin"])):]

    def get_grad_pdf(self, atoms):
        """Gradient of G(r) (``__init__.py:498-524``)."""
        self._ensure(atoms)
        fq_grad = self.grad(atoms, self.pdf_qbin, "PDF")
        qmin_bin = int(self.exp["qmin"] / self.pdf_qbin)
        fq_grad[:, :, :qmin_bin] = 0.
        return self.grad_pdf(fq_grad, self.exp["rstep"], self.pdf_qbin, self.get_r(),
                             self.exp["qmin"])

    def get_scatter_vector(self, pdf=False):
        """``__init__.py:526-547``."""
        if pdf:
            return np.arange(0., math.floor(self.exp["qmax"] / self.pdf_qbin) * self.pdf_qbin,
                             self.pdf_qbin)
        return np.arange(self.exp["qmin"], math.floor(self.exp["qmax"] / self.exp["qbin"]) *
                         self.exp["qbin"], self.exp["qbin"])

    def get_r(self):
        """``__init__.py:549-558``."""
        return np.arange(self.exp["rmin"], self.exp["rmax"], self.exp["rstep"])
