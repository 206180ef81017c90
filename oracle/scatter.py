"""FP64 NumPy restatement of the reference scattering path (test infrastructure).

Every routine here is a restatement of the reference's algorithm written from
its source text, evaluated in float64.  Inputs that the reference casts to
float32 (positions, form-factor tables, the Q step passed to the numba
kernels) are cast the same way first and then widened to float64, so the
oracle answers "what the reference computes, without float32 rounding in the
arithmetic".

Citations are ``path:line`` relative to the reference repository root.
"""
import math

import numpy as np

__all__ = [
    "DEFAULT_EXP", "make_experiment", "qmax_bin", "scatter_vector", "r_grid",
    "pdf_scatter_vector", "q_values", "ij_to_k", "k_to_ij", "pair_geometry",
    "pair_geometry_nxn", "normalization", "normalization_bruteforce", "fq",
    "fq_nxn", "grad_fq", "grad_fq_nxn", "fft_gr_to_fq", "fft_fq_to_gr",
    "get_pdf_at_qmin", "transform_grid", "transform_matrix", "grad_pdf",
    "get_scale", "get_rw", "get_chi_sq", "get_grad_scale", "get_grad_rw",
    "get_grad_chi_sq", "wrap_grad_rw", "wrap_grad_chi_sq", "leapfrog",
]

# ---------------------------------------------------------------------------
# Experiment grids: pyiid/experiments/elasticscatter/__init__.py:86-89,180-204
# ---------------------------------------------------------------------------
DEFAULT_EXP = {"qmin": 0.0, "qmax": 25, "qbin": .1, "rmin": 0.0,
               "rmax": 40.0, "rstep": .01, "sampling": "full"}


def make_experiment(exp_dict=None):
    """Fill defaults and derive ``pdf_qbin`` (``__init__.py:180-204``).

    Returns ``(exp, pdf_qbin)``; ``'ns'`` sampling sets ``rstep = pi/qmax``.
    """
    exp = {} if not exp_dict else dict(exp_dict)
    for key, value in DEFAULT_EXP.items():
        exp.setdefault(key, value)
    if exp["sampling"] == "ns":
        exp["rstep"] = np.pi / exp["qmax"]
    pdf_qbin = np.pi / (exp["rmax"] + 6 * 2 * np.pi / exp["qmax"])
    return exp, pdf_qbin


def qmax_bin(qmax, qbin):
    """Number of Q bins of a scatter table (``__init__.py:132``)."""
    return int(math.floor(qmax / qbin))


def scatter_vector(exp):
    """``get_scatter_vector(pdf=False)`` (``__init__.py:544-547``)."""
    return np.arange(exp["qmin"], math.floor(exp["qmax"] / exp["qbin"]) *
                     exp["qbin"], exp["qbin"])


def pdf_scatter_vector(exp, pdf_qbin):
    """``get_scatter_vector(pdf=True)`` (``__init__.py:540-543``)."""
    return np.arange(0., math.floor(exp["qmax"] / pdf_qbin) * pdf_qbin,
                     pdf_qbin)


def r_grid(exp):
    """``get_r`` (``__init__.py:549-558``)."""
    return np.arange(exp["rmin"], exp["rmax"], exp["rstep"])


def q_values(qbin, nq):
    """Q_m = m * qbin with qbin rounded to float32 as the numba kernels receive
    it (``cpu_flat.py:128-147``: ``get_omega(..., qbin: f4)``), evaluated
    exactly in float64."""
    return float(np.float32(qbin)) * np.arange(nq, dtype=np.float64)


# ---------------------------------------------------------------------------
# Flat pair indexing: kernels/__init__.py:10-19
# ---------------------------------------------------------------------------
def ij_to_k(i, j):
    """k = j + i(i-1)/2 for i > j (``kernels/__init__.py:10-12``)."""
    return j + i * (i - 1) // 2


def k_to_ij(k):
    """Inverse of :func:`ij_to_k` in exact integer arithmetic
    (``kernels/__init__.py:15-19`` evaluates it in float64)."""
    k = np.asarray(k, dtype=np.int64)
    i = ((1 + np.sqrt(1 + 8.0 * k)) // 2).astype(np.int64)
    # repair float rounding at the triangle boundaries
    i = np.where(i * (i - 1) // 2 > k, i - 1, i)
    i = np.where((i + 1) * i // 2 <= k, i + 1, i)
    return i, k - i * (i - 1) // 2


def _as_f32_f64(x, cast=True):
    """float32-rounded (as the reference casts positions and tables), widened
    to float64; ``cast=False`` keeps float64 input exact (finite differences)."""
    if not cast:
        return np.asarray(x, dtype=np.float64)
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def pair_geometry(q, k_offset=0, k_count=None):
    """Flat pair displacement ``d[k] = q_i - q_j`` and ``r[k] = |d[k]|`` for
    pairs ``k = j + i(i-1)/2`` (``cpu_flat.py:14-52``).  FP64 from float32
    positions."""
    q = _as_f32_f64(q)
    n = len(q)
    kmax = n * (n - 1) // 2
    if k_count is None:
        k_count = kmax - k_offset
    i, j = k_to_ij(np.arange(k_offset, k_offset + k_count))
    d = q[i] - q[j]
    return d, np.sqrt(np.sum(d * d, axis=1))


def pair_geometry_nxn(q, cast_f32=True):
    """Full ``d[i, j] = q_j - q_i`` and ``r[i, j]`` (``cpu_nxn.py:17-54``)."""
    q = _as_f32_f64(q, cast_f32)
    d = q[None, :, :] - q[:, None, :]
    return d, np.sqrt(np.sum(d * d, axis=-1))


# ---------------------------------------------------------------------------
# Normalisation: cpu_flat.py:55-74, flat_multi_cpu_wrap.py:52-54
# ---------------------------------------------------------------------------
def normalization(scatter):
    """na(Q) = N * mean_{i>j} f_i(Q) f_j(Q), in closed form.

    The reference materialises the K x Q product array and takes its mean
    (``flat_multi_cpu_wrap.py:52-54``); the mean over unordered pairs equals
    ((sum f)^2 - sum f^2) / (N (N-1)).
    """
    f = _as_f32_f64(scatter)
    n = f.shape[0]
    if n < 2:
        return np.full(f.shape[1], np.nan)
    s1 = f.sum(axis=0)
    s2 = (f * f).sum(axis=0)
    return n * (s1 * s1 - s2) / (n * (n - 1))


def normalization_bruteforce(scatter):
    """Explicit mean over the flat pair list (small N only)."""
    f = _as_f32_f64(scatter)
    n = f.shape[0]
    i, j = k_to_ij(np.arange(n * (n - 1) // 2))
    return np.mean(f[i] * f[j], axis=0) * n


def _nan_to_num_div(num, den):
    with np.errstate(all="ignore"):
        return np.nan_to_num(num / den)


def _pair_chunks(n, chunk):
    """Yield (i, j) index arrays over the flat pair list in chunks."""
    kmax = n * (n - 1) // 2
    for k0 in range(0, kmax, chunk):
        yield k_to_ij(np.arange(k0, min(k0 + chunk, kmax)))


# ---------------------------------------------------------------------------
# Debye F(Q): cpu_flat.py:77-114, cpu_atomics.py:60-78,
# flat_multi_cpu_wrap.py:22-60
# ---------------------------------------------------------------------------
def fq(q, scatter, qbin, q_index=None, chunk_elems=1 << 22, cast_f32=True):
    """Reduced structure factor F(Q) on the full grid from Q = 0.

    F(Q) = 2 / na(Q) * sum_{i>j} f_i f_j sin(Q r_ij) / r_ij with
    ``nan_to_num`` (``flat_multi_cpu_wrap.py:49-60``).  ``q_index`` restricts
    the evaluation to a subset of Q bins (the sum is independent per bin).
    """
    q = _as_f32_f64(q, cast_f32)
    f = _as_f32_f64(scatter)
    n, nq = f.shape
    idx = np.arange(nq) if q_index is None else np.asarray(q_index)
    Q = q_values(qbin, nq)[idx]
    fs = f[:, idx]
    acc = np.zeros(len(idx))
    chunk = max(1, chunk_elems // max(1, len(idx)))
    for i, j in _pair_chunks(n, chunk):
        r = np.sqrt(np.sum((q[i] - q[j]) ** 2, axis=1))
        acc += np.sum(fs[i] * fs[j] * np.sin(np.outer(r, Q)) / r[:, None],
                      axis=0)
    na = normalization(scatter)[idx]
    return 2.0 * _nan_to_num_div(acc, na)


def fq_nxn(q, scatter, qbin, cast_f32=True):
    """F(Q) by the full N x N double sum (``cpu_nxn.py:57-111`` +
    ``nxn_cpu_wrap.py:206-266``): every ordered pair i != j once, no factor 2.
    An independent code path used to pin :func:`fq` (small N)."""
    f = _as_f32_f64(scatter)
    n, nq = f.shape
    Q = q_values(qbin, nq)
    _, r = pair_geometry_nxn(q, cast_f32)
    off = ~np.eye(n, dtype=bool)
    rr = r[off]
    ii, jj = np.nonzero(off)
    s = np.sum(f[ii] * f[jj] * np.sin(np.outer(rr, Q)) / rr[:, None], axis=0)
    return _nan_to_num_div(s, normalization(scatter))


# ---------------------------------------------------------------------------
# grad F(Q): cpu_flat.py:117-153, cpu_experimental.py:8-15,
# flat_multi_cpu_wrap.py:63-102
# ---------------------------------------------------------------------------
def grad_fq(q, scatter, qbin, q_index=None, atom_index=None,
            chunk_elems=1 << 22, cast_f32=True):
    """What ``get_grad_fq`` returns: (N, 3, Qb) array equal to -1/2 dF/dq.

    Per flat pair k=(i>j), d = q_i - q_j, g = f_i f_j (Q cos(Q r) - sin(Q r)/r)
    / r^2 * d (``cpu_flat.py:120-153``), scattered as ``grad[i] -= g;
    grad[j] += g`` (``cpu_experimental.py:8-15``) and divided by na without the
    factor 2 used for F (``flat_multi_cpu_wrap.py:93-102``).  Restricting to
    ``atom_index`` rows evaluates those rows only (the sum is per row).
    """
    q = _as_f32_f64(q, cast_f32)
    f = _as_f32_f64(scatter)
    n, nq = f.shape
    qi_idx = np.arange(nq) if q_index is None else np.asarray(q_index)
    rows = np.arange(n) if atom_index is None else np.asarray(atom_index)
    Q = q_values(qbin, nq)[qi_idx]
    fs = f[:, qi_idx]
    out = np.zeros((len(rows), 3, len(qi_idx)))
    if n < 2:
        return out
    chunk_rows = max(1, chunk_elems // max(1, n * len(qi_idx)))
    for r0 in range(0, len(rows), chunk_rows):
        rr_idx = rows[r0:r0 + chunk_rows]
        d = q[rr_idx][:, None, :] - q[None, :, :]           # (b, n, 3)
        r = np.sqrt(np.sum(d * d, axis=-1))                 # (b, n)
        self_mask = r == 0.0
        r_safe = np.where(self_mask, 1.0, r)
        th = r_safe[:, :, None] * Q[None, None, :]          # (b, n, nq)
        a = (Q * np.cos(th) - np.sin(th) / r_safe[:, :, None]) / \
            (r_safe * r_safe)[:, :, None]
        a = a * fs[rr_idx][:, None, :] * fs[None, :, :]
        a[self_mask] = 0.0
        # grad[i] = -sum_{j != i} a_ij (q_i - q_j)
        out[r0:r0 + len(rr_idx)] = -np.einsum("bjq,bjc->bcq", a, d)
    na = normalization(scatter)[qi_idx]
    return _nan_to_num_div(out, na)


def grad_fq_nxn(q, scatter, qbin, cast_f32=True):
    """-1/2 dF/dq by the N x N path (``cpu_nxn.py:137-197`` +
    ``nxn_cpu_wrap.py:269-334``): sum over j of a_ij (q_j - q_i) / na."""
    f = _as_f32_f64(scatter)
    n, nq = f.shape
    Q = q_values(qbin, nq)
    d, r = pair_geometry_nxn(q, cast_f32)
    out = np.zeros((n, 3, nq))
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            th = Q * r[i, j]
            a = (Q * np.cos(th) - np.sin(th) / r[i, j]) / r[i, j] ** 2
            out[i] += (f[i] * f[j] * a)[None, :] * d[i, j][:, None]
    return _nan_to_num_div(out, normalization(scatter))


# ---------------------------------------------------------------------------
# F(Q) -> G(r): kernels/master_kernel.py:39-203
# ---------------------------------------------------------------------------
def fft_gr_to_fq(g, rstep, rmin):
    """Odd-extended, zero-padded inverse FFT sine transform
    (``master_kernel.py:131-203``); returns the imaginary part, length npad2."""
    g = np.asarray(g, dtype=np.float64)
    padrmin = int(round(rmin / rstep))
    npad1 = padrmin + len(g)
    npad2 = (1 << int(math.ceil(math.log(npad1, 2)))) * 2
    npad4 = 4 * npad2
    ext = np.zeros(npad4)
    ext[0:2 * len(g):2] = g                      # even slots carry g
    m = np.arange(1, len(g))
    ext[npad4 - 2 * m] = -g[1:]                  # odd (negative) image
    spec = np.fft.ifft(ext)
    return (spec[0:2 * npad2:2] * npad2 * rstep).imag


def fft_fq_to_gr(f, qbin, qmin):
    """``fft_fq_to_gr`` = (2/pi) * ``fft_gr_to_fq`` (``master_kernel.py:108-128``)."""
    return fft_gr_to_fq(f, qbin, qmin) * (2.0 / math.pi)


def get_pdf_at_qmin(fpad, rstep, qstep, rgrid, qmin):
    """G(r) from F(Q) by the FFT path (``master_kernel.py:39-104``).

    Works on a copy: the reference zeroes ``fpad[:ceil(qmin/qstep)]`` in place.
    """
    fpad = np.array(fpad, dtype=np.float64)
    fpad[:int(math.ceil(qmin / qstep))] = 0.0
    nfromdr = int(math.ceil(math.pi / rstep / qstep))
    if nfromdr > len(fpad):
        grown = np.zeros(nfromdr)
        grown[:len(fpad)] = fpad
        fpad = grown
    gpad = fft_fq_to_gr(fpad, qstep, qmin)
    drpad = math.pi / (len(gpad) * qstep)
    x = np.asarray(rgrid, dtype=np.float64) / drpad / 2
    lo = x.astype(np.int64)
    w = x - lo
    return 2.0 * ((1.0 - w) * gpad[lo] + w * gpad[lo + 1])


def transform_grid(rstep, qstep, qmin, nq):
    """The FFT grid parameters of :func:`get_pdf_at_qmin` for an F(Q) of
    length ``nq``: ``(npad2, drpad, q_zero)`` with ``q_zero = ceil(qmin/qstep)``
    the number of leading bins zeroed."""
    nfromdr = int(math.ceil(math.pi / rstep / qstep))
    length = max(nq, nfromdr)
    npad1 = int(round(qmin / qstep)) + length
    npad2 = (1 << int(math.ceil(math.log(npad1, 2)))) * 2
    drpad = math.pi / (npad2 * qstep)
    return npad2, drpad, int(math.ceil(qmin / qstep))


def transform_matrix(rstep, qstep, rgrid, qmin, nq):
    """Dense T (R x nq) with G = T @ F, equal to :func:`get_pdf_at_qmin`.

    T[t, m] = (2 dq / pi) [(1 - w_t) sin(2 pi l_t m / npad2)
                           + w_t sin(2 pi (l_t + 1) m / npad2)]
    for m >= ceil(qmin/dq), 0 below; l_t = floor(r_t / (2 drpad)).
    The integer products l*m are reduced mod npad2 before the sine so the
    phase is exact.
    """
    npad2, drpad, q_zero = transform_grid(rstep, qstep, qmin, nq)
    x = np.asarray(rgrid, dtype=np.float64) / drpad / 2
    lo = x.astype(np.int64)
    w = x - lo
    m = np.arange(nq, dtype=np.int64)
    ph_lo = (np.outer(lo, m) % npad2) * (2.0 * math.pi / npad2)
    ph_hi = (np.outer(lo + 1, m) % npad2) * (2.0 * math.pi / npad2)
    t = (2.0 * qstep / math.pi) * ((1.0 - w)[:, None] * np.sin(ph_lo) +
                                   w[:, None] * np.sin(ph_hi))
    t[:, :q_zero] = 0.0
    return t


def grad_pdf(grad_fq_arr, rstep, qstep, rgrid, qmin):
    """Row-wise :func:`get_pdf_at_qmin` over an (N, 3, Qp) gradient
    (``master_kernel.py:272-290``)."""
    g = np.asarray(grad_fq_arr, dtype=np.float64)
    n = g.shape[0]
    out = np.empty((n, 3, len(rgrid)))
    for a in range(n):
        for c in range(3):
            out[a, c] = get_pdf_at_qmin(g[a, c], rstep, qstep, rgrid, qmin)
    return out


def get_pdf_at_qmin_rows(f_rows, rstep, qstep, rgrid, qmin, workers=1, chunk=128):
    """:func:`get_pdf_at_qmin` of every row of ``f_rows`` (rows x nq): the
    same zeroing, padding, odd-extended inverse FFT and interpolation, run as
    batched FFTs over ``chunk`` rows at a time on ``workers`` threads
    (scipy.fft).  The reference spreads the rows of its gradient transform
    over a process pool instead (``master_kernel.py:272-290``)."""
    import scipy.fft
    f_rows = np.array(f_rows, dtype=np.float64, ndmin=2)
    rows = f_rows.shape[0]
    f_rows[:, :int(math.ceil(qmin / qstep))] = 0.0
    length = max(f_rows.shape[1], int(math.ceil(math.pi / rstep / qstep)))
    npad1 = int(round(qmin / qstep)) + length
    npad2 = (1 << int(math.ceil(math.log(npad1, 2)))) * 2
    npad4 = 4 * npad2
    drpad = math.pi / (npad2 * qstep)
    x = np.asarray(rgrid, dtype=np.float64) / drpad / 2
    lo = x.astype(np.int64)
    w = x - lo
    m = np.arange(1, f_rows.shape[1])
    out = np.empty((rows, len(x)))
    for r0 in range(0, rows, chunk):
        blk = f_rows[r0:r0 + chunk]
        ext = np.zeros((len(blk), npad4))
        ext[:, 0:2 * blk.shape[1]:2] = blk
        ext[:, npad4 - 2 * m] = -blk[:, 1:]
        spec = scipy.fft.ifft(ext, axis=1, workers=workers)
        gpad = (spec[:, 0:2 * npad2:2] * npad2 * qstep).imag * (2.0 / math.pi)
        out[r0:r0 + chunk] = 2.0 * ((1.0 - w) * gpad[:, lo] + w * gpad[:, lo + 1])
    return out


def grad_pdf_rows(grad_fq_arr, rstep, qstep, rgrid, qmin, workers=1):
    """:func:`grad_pdf` through :func:`get_pdf_at_qmin_rows`."""
    g = np.asarray(grad_fq_arr, dtype=np.float64)
    n = g.shape[0]
    flat = get_pdf_at_qmin_rows(g.reshape(n * 3, -1), rstep, qstep, rgrid, qmin, workers)
    return flat.reshape(n, 3, len(rgrid))


# ---------------------------------------------------------------------------
# Rw / chi^2 and gradients: master_kernel.py:206-266,293-396,
# calc/__init__.py:10-105
# ---------------------------------------------------------------------------
def get_scale(target, calculated):
    """Least-squares scale (``master_kernel.py:388-389``)."""
    calculated = np.asarray(calculated, dtype=np.float64)
    with np.errstate(all="ignore"):
        return np.dot(calculated, target) / np.dot(calculated, calculated)


def get_rw(gobs, gcalc, weight=None):
    """Rw and scale; ``scale <= 0`` returns ``(1, 1)`` (``master_kernel.py:206-236``)."""
    gobs = np.asarray(gobs, dtype=np.float64)
    gcalc = np.asarray(gcalc, dtype=np.float64)
    if weight is None:
        weight = np.ones(gcalc.shape)
    scale = get_scale(gobs, gcalc)
    if scale <= 0:
        return 1, 1
    top = np.sum(weight * (gobs - scale * gcalc) ** 2)
    bottom = np.sum(weight * gobs ** 2)
    return float(np.sqrt(top / bottom)), float(scale)


def get_chi_sq(gobs, gcalc):
    """chi^2 and scale, scale <= 0 replaced by 1 (``master_kernel.py:239-266``)."""
    gobs = np.asarray(gobs, dtype=np.float64)
    gcalc = np.asarray(gcalc, dtype=np.float64)
    scale = get_scale(gobs, gcalc)
    if scale <= 0:
        scale = 1
    return float(np.sum((gobs - scale * gcalc) ** 2)), float(scale)


def get_grad_scale(target, calculated, grad_calculated):
    """d(scale)/dq for every row (``master_kernel.py:392-396``), vectorised."""
    a = get_scale(target, calculated)
    gdot_c = np.tensordot(grad_calculated, calculated, axes=([-1], [0]))
    gdot_t = np.tensordot(grad_calculated, target, axes=([-1], [0]))
    return (-2 * a * gdot_c + gdot_t) / np.dot(calculated, calculated)


def get_grad_rw(grad_pdf_arr, gcalc, gobs, rw, scale):
    """Rw gradient (``master_kernel.py:293-347``), for ``scale > 0`` as every
    caller passes it (``wrap_grad_rw`` takes scale from :func:`get_rw`, which
    never returns a non-positive scale)."""
    gp = np.asarray(grad_pdf_arr, dtype=np.float64)
    gcalc = np.asarray(gcalc, dtype=np.float64)
    gobs = np.asarray(gobs, dtype=np.float64)
    grad_a = get_grad_scale(gobs, gcalc, gp)
    res = gobs - scale * gcalc
    proj = np.tensordot(gp, res, axes=([-1], [0]))
    return -1 * rw / np.dot(res, res) * (scale * proj +
                                         grad_a * np.dot(gcalc, res))


def get_grad_chi_sq(grad_pdf_arr, gcalc, gobs, scale):
    """chi^2 gradient (``master_kernel.py:350-375``) for ``scale > 0``."""
    gp = np.asarray(grad_pdf_arr, dtype=np.float64)
    gcalc = np.asarray(gcalc, dtype=np.float64)
    gobs = np.asarray(gobs, dtype=np.float64)
    grad_a = get_grad_scale(gobs, gcalc, gp)
    res = gobs - scale * gcalc
    proj = np.tensordot(gp, res, axes=([-1], [0]))
    return -2 * (scale * proj + grad_a * np.dot(gcalc, res))


def wrap_grad_rw(grad_gcalc, gcalc, gobs):
    """``calc/__init__.py:56-79``."""
    rw, scale = get_rw(gobs, gcalc)
    return get_grad_rw(grad_gcalc, gcalc, gobs, rw, scale)


def wrap_grad_chi_sq(grad_gcalc, gcalc, gobs):
    """``calc/__init__.py:82-105``."""
    _, scale = get_chi_sq(gobs, gcalc)
    return get_grad_chi_sq(grad_gcalc, gcalc, gobs, scale)


# ---------------------------------------------------------------------------
# Leapfrog: pyiid/sim/__init__.py:10-38
# ---------------------------------------------------------------------------
def leapfrog(positions, momenta, masses, forces_fn, step):
    """One velocity-Verlet step: p += eps/2 F(x); x += eps p/m;
    p += eps/2 F(x') (``sim/__init__.py:10-38``, without ``center()``, which
    translates the whole structure and leaves every pair distance unchanged)."""
    x = np.array(positions, dtype=np.float64)
    p = np.array(momenta, dtype=np.float64)
    m = np.asarray(masses, dtype=np.float64)[:, None]
    p = p + 0.5 * step * forces_fn(x)
    x = x + step * p / m
    p = p + 0.5 * step * forces_fn(x)
    return x, p
