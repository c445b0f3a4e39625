"""Learned per-atom energy surrogate: training frames from the Lennard-Jones
engine, training on the GPU (md_trainer_* kernels), deployment into an Engine
running force mode "mlp".

The surrogate has no counterpart in the reference; its labels are the
reference's own Lennard-Jones energies and forces (update_force,
potential.rs:158-216), computed by the FP64 list path, which matches the
reference bitwise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .engine import Engine, lattice, mb_velocities, thermostat

SIGMA = 0.3418
AR_MASS = 66.335
NK, NH = 32, 128
NP = L.MD_MLP_NUM_PARAMS


def init_params(seed=0):
    """W ~ N(0, 1/fan_in), zero biases; flat W1|b1|W2|b2|w3|b3 (float32)."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0.0, 1.0 / np.sqrt(NK), (NH, NK))
    w2 = rng.normal(0.0, 1.0 / np.sqrt(NH), (NH, NH))
    w3 = rng.normal(0.0, 1.0 / np.sqrt(NH), NH)
    return np.concatenate([w1.ravel(), np.zeros(NH), w2.ravel(), np.zeros(NH), w3, [0.0]]).astype(np.float32)


@dataclass
class Surrogate:
    theta: np.ndarray
    g_mean: np.ndarray
    g_inv_std: np.ndarray
    g_lo: np.ndarray
    g_hi: np.ndarray
    e_scale: float
    e_shift: float
    history: list = field(default_factory=list)

    def apply(self, engine, judge_every=100, range_check=True):
        engine.set_mlp(self.theta, self.g_mean, self.g_inv_std,
                       self.g_lo if range_check else None, self.g_hi if range_check else None,
                       self.e_scale, self.e_shift, judge_every)


@dataclass
class Frames:
    pos: np.ndarray    # (F, N, 3)
    box: np.ndarray    # (F, 3)
    e_ref: np.ndarray  # (F, N) per-atom energy (Particle::potential / 2)
    f_ref: np.ndarray  # (F, N, 3)

    def subset(self, idx):
        return Frames(self.pos[idx], self.box[idx], self.e_ref[idx], self.f_ref[idx])


def lj_labels(pos, box):
    """FP64 LJ per-atom energies and forces of one frame (FP64 list path)."""
    n = len(pos)
    e = Engine([n], [AR_MASS], box, force_mode="fp64_list")
    try:
        e.upload(pos, np.zeros_like(pos), box)
        e.update_force()
        d = e.download(("force", "pot"))
    finally:
        e.close()
    return d["pot"] / 2.0, d["force"]


def make_frames(n_side=16, lattices=(0.38, 0.42, 0.47, 0.55, 0.7), temps=(80.0, 400.0), n_frames=2000,
                runs=20, equil=400, every=20, dt=0.002, seed=0, log=None):
    """Argon frames of n_side^3 atoms: `runs` thermostatted trajectories from
    simple-cubic lattices (position.rs:34-51) at the given lattice edges and
    target temperatures uniform in `temps`, sampled every `every` steps after
    `equil` steps, labelled with FP64 LJ energies and forces."""
    rng = np.random.default_rng(seed)
    n = n_side ** 3
    per = int(np.ceil(n_frames / runs))
    pos_l, box_l, e_l, f_l = [], [], [], []
    for r in range(runs):
        a = lattices[r % len(lattices)]
        t0 = float(rng.uniform(*temps))
        box = np.full(3, n_side * a)
        x = lattice("u", [n_side] * 3, a)
        x = x + rng.uniform(-0.05, 0.05, x.shape) * a
        x = np.mod(x, box)
        v = mb_velocities(n, t0, AR_MASS, int(rng.integers(1 << 30)))
        eng = Engine([n], [AR_MASS], box, force_mode="fp32_list")
        lab = Engine([n], [AR_MASS], box, force_mode="fp64_list")  # labels (box fixed: no barostat)
        try:
            eng.upload(x, v, box)
            eng.update_force()
            th = thermostat("berendsen", tau=0.5, target_t=t0)
            eng.step(dt, equil, th)
            for k in range(per):
                if len(pos_l) >= n_frames:
                    break
                eng.step(dt, every, th)
                d = eng.download(("x", "box"))
                lab.upload(d["x"], np.zeros_like(d["x"]), d["box"])
                lab.update_force()
                lb = lab.download(("force", "pot"))
                pos_l.append(d["x"])
                box_l.append(np.array(d["box"]))
                e_l.append(lb["pot"] / 2.0)
                f_l.append(lb["force"])
        finally:
            eng.close()
            lab.close()
        if log:
            log(f"[mlp] run {r}: a {a} T {t0:.0f} K, {len(pos_l)} frames")
    return Frames(np.array(pos_l), np.array(box_l), np.array(e_l), np.array(f_l))


class Trainer:
    """GPU trainer over a resident data set of frames (one atom count)."""

    def __init__(self, frames: Frames, sigma=SIGMA):
        self._lib = L.load()
        f, n = frames.e_ref.shape
        self.n_frames, self.n_atoms = f, n
        pos = np.ascontiguousarray(frames.pos, dtype=np.float64)
        box = np.ascontiguousarray(frames.box, dtype=np.float64)
        er = np.ascontiguousarray(frames.e_ref, dtype=np.float64)
        fr = np.ascontiguousarray(frames.f_ref, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        h = C.c_void_p()
        rc = self._lib.md_trainer_create(C.byref(h), f, n, pos.ctypes.data_as(dp), box.ctypes.data_as(dp),
                                         er.ctypes.data_as(dp), fr.ctypes.data_as(dp), sigma)
        if rc != 0:
            raise L.MdError(rc, self._lib.md_trainer_last_error(None).decode())
        self._h = h

    def _chk(self, rc):
        if rc != 0:
            raise L.MdError(rc, self._lib.md_trainer_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.md_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def feature_stats(self):
        out = np.zeros(4 * NK)
        self._chk(self._lib.md_trainer_feature_stats(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return {"mean": out[:NK], "std": out[NK:2 * NK], "min": out[2 * NK:3 * NK], "max": out[3 * NK:]}

    def set_params(self, theta, g_mean, g_inv_std, e_scale, e_shift):
        fp = C.POINTER(C.c_float)
        th = np.ascontiguousarray(theta, dtype=np.float32)
        gm = np.ascontiguousarray(g_mean, dtype=np.float32)
        gs = np.ascontiguousarray(g_inv_std, dtype=np.float32)
        self._chk(self._lib.md_trainer_set_params(self._h, th.ctypes.data_as(fp), gm.ctypes.data_as(fp),
                                                  gs.ctypes.data_as(fp), float(e_scale), float(e_shift)))

    def params(self):
        th = np.zeros(NP, dtype=np.float32)
        self._chk(self._lib.md_trainer_get_params(self._h, th.ctypes.data_as(C.POINTER(C.c_float))))
        return th

    def loss(self, w_e=1.0, w_f=1.0, batch_frames=64, grad=False):
        out = np.zeros(3)
        g = np.zeros(NP, dtype=np.float32) if grad else None
        self._chk(self._lib.md_trainer_loss(self._h, w_e, w_f, batch_frames, out.ctypes.data_as(C.POINTER(C.c_double)),
                                            g.ctypes.data_as(C.POINTER(C.c_float)) if grad else None))
        return (out, g) if grad else out

    def epoch(self, batch_frames, lr, w_e=1.0, w_f=1.0, order=None):
        out = np.zeros(3)
        o = np.ascontiguousarray(order, dtype=np.int32) if order is not None else None
        self._chk(self._lib.md_trainer_epoch(self._h, batch_frames, lr, w_e, w_f,
                                             o.ctypes.data_as(C.POINTER(C.c_int32)) if o is not None else None,
                                             out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def predict(self, frame):
        e = np.zeros(self.n_atoms)
        f = np.zeros((self.n_atoms, 3))
        dp = C.POINTER(C.c_double)
        self._chk(self._lib.md_trainer_predict(self._h, frame, e.ctypes.data_as(dp), f.ctypes.data_as(dp)))
        return e, f


def train(frames: Frames, epochs=60, batch_frames=16, lr=2e-3, lr_final=1e-4, seed=0, w_e=None, w_f=None,
          range_margin=0.05, log=None):
    """Train a surrogate on `frames`; returns a Surrogate (range = feature
    min/max of the data widened by `range_margin` of the span)."""
    tr = Trainer(frames)
    try:
        fs = tr.feature_stats()
        g_mean = fs["mean"].astype(np.float32)
        g_inv_std = (1.0 / np.maximum(fs["std"], 1e-3)).astype(np.float32)
        e_shift = float(np.mean(frames.e_ref))
        e_scale = float(np.std(frames.e_ref)) or 1.0
        w_e = w_e if w_e is not None else 1.0 / max(float(np.var(frames.e_ref)), 1e-12)
        w_f = w_f if w_f is not None else 1.0 / max(float(np.var(frames.f_ref)), 1e-12)
        tr.set_params(init_params(seed), g_mean, g_inv_std, e_scale, e_shift)
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            lr_ep = lr * (lr_final / lr) ** (ep / max(epochs - 1, 1))
            out = tr.epoch(batch_frames, lr_ep, w_e, w_f, order=rng.permutation(frames.e_ref.shape[0]))
            hist.append(out.tolist())
            if log and (ep % 10 == 0 or ep == epochs - 1):
                log(f"[mlp] epoch {ep}: loss {out[0]:.4e} (E mse {out[1]:.3e}, F mse {out[2]:.3e}) lr {lr_ep:.2e}")
        span = fs["max"] - fs["min"]
        return Surrogate(tr.params(), g_mean, g_inv_std,
                         (fs["min"] - range_margin * span).astype(np.float32),
                         (fs["max"] + range_margin * span).astype(np.float32), e_scale, e_shift, hist)
    finally:
        tr.close()
