"""FP64 CPU restatement of the learned per-atom energy surrogate.

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py
may import this module, as the checker of the HIP kernels.

The surrogate has no counterpart in the reference (moldyn has LJ only).  Its
definition, which the HIP kernels implement:

  descriptor  G_ik = sum_{j != i, r_ij < r_c} exp(-(r_ij - mu_k)^2 / (2 w^2)) f_c(r_ij)
              k = 0..31, mu_k = sigma (0.8 + 1.7 k / 31), w = 0.1 sigma,
              r_c = 2.5 sigma, f_c(r) = (cos(pi r / r_c) + 1) / 2; r_ij under the
              reference minimum image (potential.rs:186-200); sigma = 0.3418
  network     x = (G - g_mean) * g_inv_std
              a1 = W1 x + b1, h1 = silu(a1)        (128)
              a2 = W2 h1 + b2, h2 = silu(a2)       (128)
              E_i = e_scale (w3 . h2 + b3) + e_shift
  forces      F_i = sum_j (c_i(r_ij) + c_j(r_ij)) d_ij / r_ij,
              c_i(r) = sum_k (dE_i/dG_ik) dg_k/dr, d_ij = x_j - x_i
  loss        L = w_e mean_i (E_i - E_i^ref)^2 + w_f mean_{i,a} (F_ia - F_ia^ref)^2
              E_i^ref = Particle::potential / 2 (energy.rs:40-49 sums potential / 2)

Pinning: forces against central differences of the total energy, parameter
gradients against central differences of the loss (tests/test_mlp_oracle.py),
Newton-3 (sum F = 0) and translation invariance.
"""
from __future__ import annotations

import numpy as np

SIGMA = 0.3418
RC = 2.5 * SIGMA
NK = 32
HIDDEN = 128
MU = SIGMA * (0.8 + 1.7 * np.arange(NK) / (NK - 1))
WIDTH = 0.1 * SIGMA


def min_image(d, box):
    """potential.rs:186-200: one correction per axis against b / 2."""
    d = d.copy()
    for k in range(3):
        b = box[k]
        lo = d[..., k] < -b / 2.0
        hi = d[..., k] > b / 2.0
        d[..., k] = np.where(lo, d[..., k] + b, np.where(hi, d[..., k] - b, d[..., k]))
    return d


def pair_list(pos, box, rc=RC):
    """Full neighbour list within rc: (i, j, d_ij = x_j - x_i, r_ij), i != j."""
    pos = np.asarray(pos, dtype=np.float64)
    n = len(pos)
    d = min_image(pos[None, :, :] - pos[:, None, :], np.asarray(box, dtype=np.float64))
    r = np.sqrt((d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2)
    mask = (r < rc) & ~np.eye(n, dtype=bool)
    i, j = np.nonzero(mask)
    return i, j, d[i, j], r[i, j]


def basis(r):
    """g_k(r) and dg_k/dr for pair distances r (< r_c): shapes (P, 32)."""
    r = np.asarray(r, dtype=np.float64)[:, None]
    e = np.exp(-((r - MU) ** 2) / (2.0 * WIDTH * WIDTH))
    fc = 0.5 * (np.cos(np.pi * r / RC) + 1.0)
    dfc = -0.5 * np.pi / RC * np.sin(np.pi * r / RC)
    g = e * fc
    dg = e * (-(r - MU) / (WIDTH * WIDTH) * fc + dfc)
    return g, dg


def descriptors(pos, box):
    i, j, d, r = pair_list(pos, box)
    g, dg = basis(r)
    G = np.zeros((len(pos), NK))
    np.add.at(G, i, g)
    return G, (i, j, d, r, dg)


def silu(a):
    s = 1.0 / (1.0 + np.exp(-a))
    return a * s


def dsilu(a):
    s = 1.0 / (1.0 + np.exp(-a))
    return s * (1.0 + a * (1.0 - s))


def d2silu(a):
    s = 1.0 / (1.0 + np.exp(-a))
    return s * (1.0 - s) * (2.0 + a * (1.0 - 2.0 * s))


class Params:
    """Network parameters (trained) and normalisation constants (fixed)."""

    NAMES = ("W1", "b1", "W2", "b2", "w3", "b3")

    def __init__(self, seed=0, g_mean=None, g_inv_std=None, e_scale=1.0, e_shift=0.0):
        rng = np.random.default_rng(seed)
        self.W1 = rng.normal(0.0, 1.0 / np.sqrt(NK), (HIDDEN, NK))
        self.b1 = np.zeros(HIDDEN)
        self.W2 = rng.normal(0.0, 1.0 / np.sqrt(HIDDEN), (HIDDEN, HIDDEN))
        self.b2 = np.zeros(HIDDEN)
        self.w3 = rng.normal(0.0, 1.0 / np.sqrt(HIDDEN), HIDDEN)
        self.b3 = np.zeros(1)
        self.g_mean = np.zeros(NK) if g_mean is None else np.asarray(g_mean, dtype=np.float64)
        self.g_inv_std = np.ones(NK) if g_inv_std is None else np.asarray(g_inv_std, dtype=np.float64)
        self.e_scale = float(e_scale)
        self.e_shift = float(e_shift)

    def copy(self):
        p = Params.__new__(Params)
        for k in self.NAMES + ("g_mean", "g_inv_std"):
            setattr(p, k, getattr(self, k).copy())
        p.e_scale, p.e_shift = self.e_scale, self.e_shift
        return p

    def flat(self):
        return np.concatenate([getattr(self, k).ravel() for k in self.NAMES])

    def set_flat(self, v):
        o = 0
        for k in self.NAMES:
            a = getattr(self, k)
            setattr(self, k, v[o:o + a.size].reshape(a.shape).copy())
            o += a.size


def forward(p, G):
    x = (G - p.g_mean) * p.g_inv_std
    a1 = x @ p.W1.T + p.b1
    h1 = silu(a1)
    a2 = h1 @ p.W2.T + p.b2
    h2 = silu(a2)
    y = h2 @ p.w3 + p.b3[0]
    E = p.e_scale * y + p.e_shift
    return E, (x, a1, h1, a2, h2)


def input_grad(p, cache):
    """dE_i/dG_i (N, 32) and the backward intermediates."""
    x, a1, h1, a2, h2 = cache
    d2 = p.w3 * dsilu(a2)                    # dy/da2
    d1 = (d2 @ p.W2) * dsilu(a1)             # dy/da1
    gx = d1 @ p.W1                           # dy/dx
    return p.e_scale * p.g_inv_std * gx, (d1, d2)


def energy_forces(p, pos, box):
    """Per-atom energies and forces of the surrogate."""
    G, (i, j, d, r, dg) = descriptors(pos, box)
    E, cache = forward(p, G)
    gam, _ = input_grad(p, cache)
    c = np.einsum("pk,pk->p", gam[i], dg) + np.einsum("pk,pk->p", gam[j], dg)
    F = np.zeros_like(np.asarray(pos, dtype=np.float64))
    np.add.at(F, i, (c / r)[:, None] * d)
    return E, F, G


def loss(p, frames, w_e=1.0, w_f=1.0):
    """frames: [(pos, box, E_ref (N,), F_ref (N, 3))]; mean over all atoms."""
    le = lf = 0.0
    n = 0
    for pos, box, er, fr in frames:
        E, F, _ = energy_forces(p, pos, box)
        le += np.sum((E - er) ** 2)
        lf += np.sum((F - fr) ** 2)
        n += len(pos)
    return w_e * le / n + w_f * lf / (3 * n)


def loss_grad(p, frames, w_e=1.0, w_f=1.0):
    """dL/dtheta for the parameters in Params.NAMES (double backward through
    the force term)."""
    gr = {k: np.zeros_like(getattr(p, k)) for k in Params.NAMES}
    n = sum(len(f[0]) for f in frames)
    for pos, box, er, fr in frames:
        G, (i, j, d, r, dg) = descriptors(pos, box)
        E, cache = forward(p, G)
        x, a1, h1, a2, h2 = cache
        gam, (d1, d2) = input_grad(p, cache)
        c = np.einsum("pk,pk->p", gam[i], dg) + np.einsum("pk,pk->p", gam[j], dg)
        F = np.zeros((len(pos), 3))
        np.add.at(F, i, (c / r)[:, None] * d)
        # energy term: dL/dy_i
        dy = w_e * 2.0 * (E - er) / n * p.e_scale
        gr["w3"] += h2.T @ dy
        gr["b3"] += np.sum(dy)
        da2 = dy[:, None] * (p.w3 * dsilu(a2))
        gr["W2"] += da2.T @ h1
        gr["b2"] += da2.sum(0)
        da1 = (da2 @ p.W2) * dsilu(a1)
        gr["W1"] += da1.T @ x
        gr["b1"] += da1.sum(0)
        # force term: phi = dL/dF, u_i = dL/d(dE_i/dG_i)
        phi = w_f * 2.0 * (F - fr) / (3 * n)
        proj = np.einsum("pa,pa->p", phi[i] - phi[j], d) / r
        u = np.zeros((len(pos), NK))
        np.add.at(u, i, proj[:, None] * dg)
        ut = u * (p.e_scale * p.g_inv_std)     # in x-space: L' = ut . dy/dx
        v1 = ut @ p.W1.T                       # W1 ut
        pp = v1 * dsilu(a1)
        v2 = pp @ p.W2.T                       # W2 p
        gr["w3"] += np.sum(v2 * dsilu(a2), axis=0)
        q2 = v2 * p.w3 * d2silu(a2)
        gr["b2"] += q2.sum(0)
        gr["W2"] += q2.T @ h1 + d2.T @ pp
        r1 = (q2 @ p.W2) * dsilu(a1) + (d2 @ p.W2) * v1 * d2silu(a1)
        gr["b1"] += r1.sum(0)
        gr["W1"] += r1.T @ x + d1.T @ ut
    return gr


def adam_step(p, grads, state, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update in place; state = {"t": int, "m": {...}, "v": {...}}."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in Params.NAMES:
        m = state.setdefault("m", {}).setdefault(k, np.zeros_like(getattr(p, k)))
        v = state.setdefault("v", {}).setdefault(k, np.zeros_like(getattr(p, k)))
        g = grads[k]
        m[:] = beta1 * m + (1.0 - beta1) * g
        v[:] = beta2 * v + (1.0 - beta2) * g * g
        mh = m / (1.0 - beta1 ** t)
        vh = v / (1.0 - beta2 ** t)
        setattr(p, k, getattr(p, k) - lr * mh / (np.sqrt(vh) + eps))
