"""Learned HMC integrator (device side in csrc/surrogate.hip).

A per-atom network on radial neighbour descriptors sums to a surrogate energy
E_s; its forces -dE_s/dq drive the leapfrog, while the exact Rw of the Debye
path decides every Metropolis step, so the chains sample the same
distribution as exact-force HMC (pyiid/sim/hmc.py:33-130) and only the
proposal changes.

Network: 16 Gaussian radial basis functions (centres 2.0 .. 6.0 A, cutoff
6 A) -> 64 SiLU -> 64 SiLU -> 1, BF16 weights and activations, FP32
accumulation on the matrix cores.  The hidden layers are fixed seeded random
features; the output layer w3 is fitted per chain, online, by ridge regression
of the feature forces on exact forces that the Debye kernels produce at the
chains' current states during sampling, optionally with one row per fit that
matches the network's energy change since the previous fit to the exact
change in conv * Rw (forces and energies are linear in w3, so each fit is a
64 x 64 solve).  When the fit is poor the sampler falls back to exact forces
until the next fit.
"""
import numpy as np

RC = 6.0
SKIN = 0.5
MU0 = 2.0
DMU = (RC - MU0) / 15.0
ETA = 1.0 / DMU ** 2
MAX_NBR = 192


class SurrogateNet(object):
    """Seeded network parameters (float32; rounded to BF16 on the device)."""

    def __init__(self, seed=0, rc=RC, skin=SKIN, mu0=MU0, dmu=DMU, eta=ETA):
        rs = np.random.RandomState(seed)
        self.rbf = (float(rc), float(mu0), float(dmu), float(eta))
        self.skin = float(skin)
        self.w1 = (rs.normal(0.0, 1.0, (16, 64)) / 4.0).astype(np.float32)
        self.b1 = rs.uniform(-1.0, 1.0, 64).astype(np.float32)
        self.w2 = (rs.normal(0.0, 1.0, (64, 64)) * np.sqrt(2.0 / 64)).astype(np.float32)
        self.b2 = rs.uniform(-1.0, 1.0, 64).astype(np.float32)
        self.mean = np.zeros(16, np.float32)
        self.inv_scale = np.ones(16, np.float32)

    @property
    def params(self):
        rc, mu0, dmu, eta = self.rbf
        return np.array([rc, self.skin, mu0, dmu, eta], np.float32)

    def as_dict(self):
        return {"w1": self.w1, "b1": self.b1, "w2": self.w2, "b2": self.b2,
                "mean": self.mean, "inv_scale": self.inv_scale}

    def upload(self, ctx, max_nbr=MAX_NBR):
        ctx.sur_setup(self.params, self.w1, self.b1, self.w2, self.b2, self.mean,
                      self.inv_scale, max_nbr)

    def normalise(self, ctx, n_atoms, max_nbr=MAX_NBR):
        """Set the input standardisation from the descriptors of the structure
        loaded in ctx (per-channel mean; scale 1/std with a floor)."""
        self.mean = np.zeros(16, np.float32)
        self.inv_scale = np.ones(16, np.float32)
        self.upload(ctx, max_nbr)
        d = ctx.sur_descriptors(n_atoms).astype(np.float64)
        sd = d.std(axis=0)
        floor = max(1e-3, 0.05 * float(sd.max())) if sd.size else 1e-3
        self.mean = d.mean(axis=0).astype(np.float32)
        self.inv_scale = (1.0 / np.maximum(sd, floor)).astype(np.float32)
        self.upload(ctx, max_nbr)
