"""CPU restatement of the learned HMC integrator (test infrastructure only).

The learned integrator has no counterpart in pyIID: its sampler
(pyiid/sim/hmc.py:33-130) integrates with exact forces.  This module states
the surrogate's mathematics in float64 numpy so that tests can check the
device kernels:

* ``descriptors``: D[i][m] = sum_{j != i, r_ij < rc} exp(-eta (r_ij - mu_m)^2)
  * (cos(pi r_ij / rc) + 1) / 2 over the atoms of i's chain (brute force).
* ``network``: x = (D - mean) * inv_scale -> SiLU(x W1 + b1) -> SiLU(. W2 + b2)
  -> . w3, and the backward pass dE/dD.  With ``bf16=True`` operands are
  rounded to bfloat16 where the device kernel rounds them (inputs, weights,
  hidden activations and backward seeds); with ``bf16=False`` it is the exact
  function, pinned by finite differences in the tests.
* ``forces``: F_a = -sum_j sum_m (c_a,m + c_j,m) g'_m(r_aj) (x_a - x_j) / r_aj.
* ``ridge_fit``: the output-layer fit (normal equations plus ridge).
* ``fit_over_states``: the fit over successive training states with
  forgetting and the energy-difference rows (e_t - e_{t-1}) . w = E_t - E_{t-1}.

Pins: the force formula against central differences of the float64 energy;
descriptor invariances (translation, rotation, permutation) and the
single-pair closed form; ``fit_over_states`` against a QR least-squares
solve of the stacked, weighted rows and against recovery of a known w.
"""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _pairs(pos, offsets, rc):
    """(i, j, d_ij = x_i - x_j, r_ij) for ordered pairs of one chain within rc."""
    out_i, out_j = [], []
    for b in range(len(offsets) - 1):
        o, e = int(offsets[b]), int(offsets[b + 1])
        if e - o < 2:
            continue
        p = pos[o:e]
        d = p[:, None, :] - p[None, :, :]
        r = np.sqrt(np.sum(d * d, axis=-1))
        ii, jj = np.nonzero((r < rc) & ~np.eye(e - o, dtype=bool))
        out_i.append(ii + o)
        out_j.append(jj + o)
    if not out_i:
        z = np.zeros(0, np.int64)
        return z, z, np.zeros((0, 3)), np.zeros(0)
    i = np.concatenate(out_i)
    j = np.concatenate(out_j)
    d = pos[i] - pos[j]
    return i, j, d, np.sqrt(np.sum(d * d, axis=1))


def _rbf(r, rc, mu0, dmu, eta):
    mu = mu0 + dmu * np.arange(16)
    x = r[:, None] - mu[None, :]
    ex = np.exp(-eta * x * x)
    fc = 0.5 * (np.cos(np.pi * r / rc) + 1.0)
    dfc = -0.5 * (np.pi / rc) * np.sin(np.pi * r / rc)
    g = ex * fc[:, None]
    dg = ex * (-2.0 * eta * x * fc[:, None] + dfc[:, None])
    return g, dg


def descriptors(pos, offsets, rc, mu0, dmu, eta):
    pos = np.asarray(pos, np.float64)
    i, j, d, r = _pairs(pos, offsets, rc)
    g, _ = _rbf(r, rc, mu0, dmu, eta)
    out = np.zeros((len(pos), 16))
    np.add.at(out, i, g)
    return out


def _silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s))


def network(D, net, w3_rows, chain_of_atom, bf16_round=True):
    """Per-atom outputs e (n,) and c = dE/dD (n, 16).

    net: dict w1 (16, 64), b1 (64,), w2 (64, 64), b2 (64,), mean (16,),
    inv_scale (16,); w3_rows (B, 64); chain_of_atom (n,)."""
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    f32 = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if bf16_round else rnd
    mean = np.asarray(net["mean"], np.float64)
    isc = np.asarray(net["inv_scale"], np.float64)
    w1 = rnd(net["w1"])
    w2 = rnd(net["w2"])
    b1 = np.asarray(net["b1"], np.float64)
    b2 = np.asarray(net["b2"], np.float64)
    w3 = rnd(w3_rows)[chain_of_atom]
    x = rnd(f32((np.asarray(D, np.float64) - mean) * isc))
    h1, s1 = _silu(x @ w1 + b1)
    h1q = rnd(h1)
    h2, s2 = _silu(h1q @ w2 + b2)
    e = np.sum(h2 * w3, axis=1)
    g2 = rnd(w3 * s2)
    g1 = rnd((g2 @ w2.T) * s1)
    c = (g1 @ w1.T) * isc
    return e, c


def forces(pos, offsets, rc, mu0, dmu, eta, c):
    pos = np.asarray(pos, np.float64)
    i, j, d, r = _pairs(pos, offsets, rc)
    _, dg = _rbf(r, rc, mu0, dmu, eta)
    s = np.sum((c[i] + c[j]) * dg, axis=1)
    f = np.zeros((len(pos), 3))
    np.add.at(f, i, -(s / r)[:, None] * d)
    return f


def energy_and_forces(pos, offsets, rbf, net, w3_rows, bf16_round=False):
    """Surrogate energy per chain and forces (float64 unless bf16_round)."""
    rc, mu0, dmu, eta = rbf
    offsets = np.asarray(offsets)
    chain = np.repeat(np.arange(len(offsets) - 1), np.diff(offsets))
    D = descriptors(pos, offsets, rc, mu0, dmu, eta)
    e, c = network(D, net, w3_rows, chain, bf16_round)
    E = np.array([np.sum(e[offsets[b]:offsets[b + 1]]) for b in range(len(offsets) - 1)])
    return E, forces(pos, offsets, rc, mu0, dmu, eta, c)


def ridge_fit(phi, y, ridge):
    """w minimising |phi^T w - y|^2 + lam |w|^2, lam = ridge * trace(phi phi^T) / 64
    (phi: 64 x m feature rows, y: m targets)."""
    a = phi @ phi.T
    lam = max(ridge, 1e-12) * np.trace(a) / a.shape[0]
    w = np.linalg.solve(a + lam * np.eye(a.shape[0]), phi @ y)
    res = y - phi.T @ w
    return w, 1.0 - np.dot(res, res) / np.dot(y, y)


def fit_over_states(samples, forget, ridge, weight):
    """Output-layer weights after training steps at successive states.

    samples: per state, in order, (phi (64 x m), y (m), e (64), E): feature
    forces, exact forces, feature energies and exact energy.  The normal
    equations are scaled by ``forget`` at every state, then take the state's
    force rows and, from the second state on, the energy row
    sqrt(weight) (e_t - e_{t-1}) . w = sqrt(weight) (E_t - E_{t-1}).  Ridge
    lam = ridge * trace / 64 as in ``ridge_fit``."""
    a = np.zeros((64, 64))
    r = np.zeros(64)
    prev = None
    for phi, y, e, E in samples:
        phi = np.asarray(phi, np.float64)
        a *= forget
        r *= forget
        a += phi @ phi.T
        r += phi @ np.asarray(y, np.float64)
        e = np.asarray(e, np.float64)
        if prev is not None and weight > 0:
            de = e - prev[0]
            a += weight * np.outer(de, de)
            r += weight * de * (E - prev[1])
        prev = (e, E)
    lam = max(ridge, 1e-12) * np.trace(a) / 64
    return np.linalg.solve(a + lam * np.eye(64), r)
