"""Pins the FP64 surrogate oracle (oracle/mlp_oracle.py) by the mathematics:
forces against central differences of the total energy, loss gradients
against central differences of the loss, Newton-3 and invariances."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO


def dense_frame(n=24, box=1.1, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, box, (n, 3))
    return pos, np.array([box, box, box])


def params(seed=1):
    p = MO.Params(seed=seed, g_mean=np.full(MO.NK, 0.3), g_inv_std=np.full(MO.NK, 1.7),
                  e_scale=0.8, e_shift=-0.2)
    rng = np.random.default_rng(seed + 10)
    p.b1 = rng.normal(0, 0.3, MO.HIDDEN)
    p.b2 = rng.normal(0, 0.3, MO.HIDDEN)
    p.b3 = rng.normal(0, 0.3, 1)
    return p


def test_basis_derivative_matches_central_difference():
    r = np.linspace(0.2, MO.RC * 0.999, 57)
    h = 1e-7
    g, dg = MO.basis(r)
    gp, _ = MO.basis(r + h)
    gm, _ = MO.basis(r - h)
    assert np.allclose(dg, (gp - gm) / (2 * h), rtol=1e-6, atol=1e-8)
    # the cutoff function takes the basis smoothly to zero at r_c
    gc, dgc = MO.basis(np.array([MO.RC * (1 - 1e-12)]))
    assert np.abs(gc).max() < 1e-12 and np.abs(dgc).max() < 1e-9


def test_forces_are_minus_energy_gradient():
    pos, box = dense_frame()
    p = params()
    E, F, _ = MO.energy_forces(p, pos, box)
    h = 1e-6
    fd = np.zeros_like(F)
    for a in range(len(pos)):
        for k in range(3):
            pp = pos.copy()
            pp[a, k] += h
            pm = pos.copy()
            pm[a, k] -= h
            fd[a, k] = -(MO.energy_forces(p, pp, box)[0].sum() - MO.energy_forces(p, pm, box)[0].sum()) / (2 * h)
    assert np.abs(F - fd).max() < 1e-6 * max(1.0, np.abs(F).max())
    assert np.abs(F).max() > 1e-3  # not trivially zero
    # Newton-3: the pair form sums to zero
    assert np.abs(F.sum(0)).max() < 1e-10 * np.abs(F).max()


def test_descriptor_invariances():
    pos, box = dense_frame(seed=3)
    p = params()
    E, F, G = MO.energy_forces(p, pos, box)
    # translation (periodic wrap) and permutation
    shift = np.array([0.37, -0.21, 0.55])
    pt = np.mod(pos + shift, box)
    E2, F2, G2 = MO.energy_forces(p, pt, box)
    assert np.allclose(G, G2, atol=1e-12) and np.allclose(F, F2, atol=1e-10)
    perm = np.random.default_rng(0).permutation(len(pos))
    E3, F3, G3 = MO.energy_forces(p, pos[perm], box)
    assert np.allclose(G[perm], G3, atol=1e-12) and np.allclose(F[perm], F3, atol=1e-10)


@pytest.mark.parametrize("w_e,w_f", [(1.0, 0.0), (0.0, 1.0), (0.7, 2.0)])
def test_loss_gradient_matches_central_difference(w_e, w_f):
    frames = []
    for s in range(2):
        pos, box = dense_frame(n=16, box=0.95, seed=20 + s)
        rng = np.random.default_rng(30 + s)
        frames.append((pos, box, rng.normal(0, 0.5, len(pos)), rng.normal(0, 2.0, (len(pos), 3))))
    p = params(seed=4)
    gr = MO.loss_grad(p, frames, w_e, w_f)
    flat = p.flat()
    gflat = np.concatenate([gr[k].ravel() for k in MO.Params.NAMES])
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(len(flat), 40, replace=False), [len(flat) - 1]])
    h = 1e-5
    for q in idx:
        v = flat.copy()
        v[q] += h
        p.set_flat(v)
        lp = MO.loss(p, frames, w_e, w_f)
        v[q] -= 2 * h
        p.set_flat(v)
        lm = MO.loss(p, frames, w_e, w_f)
        p.set_flat(flat)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gflat[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, gflat[q])


def test_adam_step_decreases_loss():
    frames = []
    for s in range(2):
        pos, box = dense_frame(n=16, box=0.95, seed=40 + s)
        E, F, _ = MO.energy_forces(params(seed=9), pos, box)  # a different network as target
        frames.append((pos, box, E, F))
    p = params(seed=4)
    st = {}
    l0 = MO.loss(p, frames)
    for _ in range(30):
        MO.adam_step(p, MO.loss_grad(p, frames), st, lr=3e-3)
    assert MO.loss(p, frames) < 0.5 * l0
