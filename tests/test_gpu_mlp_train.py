"""Surrogate training on the GPU (md_trainer_*, FP32 kernels) against the FP64
oracle (oracle/mlp_oracle.py): loss, energies and forces, the gradient
through the forces (double backward), and what training gains.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from tests.common import rms

pytestmark = pytest.mark.gpu


def small_frames(n_frames=3, n=64, box=1.8, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, box, (n_frames, n, 3))
    boxes = np.full((n_frames, 3), box)
    return pos, boxes


def params_for(pos, boxes, seed=2):
    p = MO.Params(seed=seed)
    G = np.concatenate([MO.descriptors(x, b)[0] for x, b in zip(pos, boxes)])
    p.g_mean = G.mean(0)
    p.g_inv_std = 1.0 / (G.std(0) + 1e-3)
    rng = np.random.default_rng(seed)
    p.b1 = rng.normal(0, 0.2, MO.HIDDEN)
    p.b2 = rng.normal(0, 0.2, MO.HIDDEN)
    p.b3 = np.array([0.1])
    p.e_scale, p.e_shift = 0.9, -0.4
    return p


def flat(p):
    return np.concatenate([p.W1.ravel(), p.b1, p.W2.ravel(), p.b2, p.w3, p.b3])


def trainer_for(pos, boxes, er, fr, p):
    from moldyn_amd.mlp import Frames, Trainer
    t = Trainer(Frames(pos, boxes, er, fr))
    t.set_params(flat(p).astype(np.float32), p.g_mean, p.g_inv_std, p.e_scale, p.e_shift)
    return t


def test_trainer_features_predict_loss_and_gradient():
    pos, boxes = small_frames()
    p = params_for(pos, boxes)
    rng = np.random.default_rng(7)
    er = rng.normal(-1.0, 0.5, pos.shape[:2])
    fr = rng.normal(0.0, 3.0, pos.shape)
    t = trainer_for(pos, boxes, er, fr, p)
    # features: data-set statistics against the oracle's
    G = np.concatenate([MO.descriptors(x, b)[0] for x, b in zip(pos, boxes)])
    fs = t.feature_stats()
    assert np.allclose(fs["mean"], G.mean(0), rtol=1e-4, atol=1e-6)
    assert np.allclose(fs["max"], G.max(0), rtol=1e-4, atol=1e-6)
    # predictions of one frame (FP32)
    e1, f1 = t.predict(1)
    E, F, _ = MO.energy_forces(p, pos[1], boxes[1])
    assert rms(e1 - E) <= 1e-4 * rms(E) and rms(f1 - F) <= 1e-4 * rms(F)
    # loss and its gradient over the data set, both weights on
    frames = list(zip(pos, boxes, er, fr))
    w_e, w_f = 0.7, 0.05
    out, g = t.loss(w_e, w_f, batch_frames=2, grad=True)
    L = MO.loss(p, frames, w_e, w_f)
    assert abs(out[0] - L) <= 1e-4 * L
    gr = MO.loss_grad(p, frames, w_e, w_f)
    gref = np.concatenate([gr[k].ravel() for k in MO.Params.NAMES])
    assert np.linalg.norm(g - gref) <= 2e-3 * np.linalg.norm(gref)
    # every parameter block, not just the largest entries
    o = 0
    for k in MO.Params.NAMES:
        nk = gr[k].size
        assert np.linalg.norm(g[o:o + nk] - gref[o:o + nk]) <= 5e-3 * np.linalg.norm(gref[o:o + nk]) + 1e-7, k
        o += nk
    t.close()


def test_adam_epochs_follow_the_oracle_update():
    """Three full-batch epochs: each is one Adam step (beta 0.9, 0.999, eps
    1e-8) on the gradient the trainer computes, as the oracle's update."""
    pos, boxes = small_frames(n_frames=2, seed=3)
    p = params_for(pos, boxes, seed=4)
    rng = np.random.default_rng(8)
    er = rng.normal(-1.0, 0.5, pos.shape[:2])
    fr = rng.normal(0.0, 3.0, pos.shape)
    t = trainer_for(pos, boxes, er, fr, p)
    st = {}
    q = p.copy()
    q.set_flat(t.params().astype(np.float64))
    for _ in range(3):
        _, g = t.loss(1.0, 0.1, batch_frames=2, grad=True)
        t.epoch(batch_frames=2, lr=1e-3, w_e=1.0, w_f=0.1)
        grads, o = {}, 0
        for k in MO.Params.NAMES:
            a = getattr(q, k)
            grads[k] = g[o:o + a.size].astype(np.float64).reshape(a.shape)
            o += a.size
        MO.adam_step(q, grads, st, lr=1e-3)
        assert np.allclose(t.params(), q.flat(), rtol=1e-5, atol=2e-6)
    t.close()


def test_training_gains_over_the_untrained_network():
    """LJ-labelled liquid and dense-gas frames made by the engine: the trained
    network's forces on held-out frames are far closer to LJ than the
    untrained network's, in the trainer (FP32) and deployed (BF16 MFMA)."""
    import moldyn_amd as M
    from moldyn_amd import mlp
    fr = mlp.make_frames(n_side=6, lattices=(0.40, 0.45, 0.52), temps=(80.0, 300.0), n_frames=48, runs=6,
                         equil=200, every=10, seed=1)
    train_idx = np.arange(40)
    test_idx = np.arange(40, 48)
    sur = mlp.train(fr.subset(train_idx), epochs=150, batch_frames=8, lr=3e-3, lr_final=2e-4, seed=0)
    assert sur.history[-1][0] < 0.2 * sur.history[0][0]
    # deployed on held-out frames (force mode "mlp", BF16 MFMA inference)
    rel_trained, rel_untrained = [], []
    for k in test_idx:
        x, box = fr.pos[k], fr.box[k]
        for theta, out in ((sur.theta, rel_trained), (mlp.init_params(0), rel_untrained)):
            e = M.Engine([len(x)], [mlp.AR_MASS], box, force_mode="mlp")
            e.set_mlp(theta, sur.g_mean, sur.g_inv_std, None, None, sur.e_scale, sur.e_shift, 0)
            e.upload(x, np.zeros_like(x), box)
            e.update_force()
            f = e.download(("force",))["force"]
            e.close()
            out.append(rms(f - fr.f_ref[k]) / rms(fr.f_ref[k]))
    assert np.mean(rel_trained) < 0.25 * np.mean(rel_untrained)
    assert np.mean(rel_trained) < 0.15
