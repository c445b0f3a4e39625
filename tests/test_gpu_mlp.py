"""Learned surrogate on the GPU (force mode "mlp") against the FP64 oracle
(oracle/mlp_oracle.py), through the C ABI.

Inference runs in BF16 on MFMA with FP32 accumulation, so energies and forces
are compared to the FP64 oracle within a relative RMSE written in each test;
the LJ fallback path is compared bitwise with the FP32 Lennard-Jones path.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from oracle import oracle as O
from tests.common import AR_MASS, engine_for, liquid_state, rms, sc_state

pytestmark = pytest.mark.gpu


def oracle_params(state, seed=3):
    p = MO.Params(seed=seed)
    G, _ = MO.descriptors(state.pos, state.box)
    p.g_mean = G.mean(0)
    p.g_inv_std = 1.0 / (G.std(0) + 1e-3)
    rng = np.random.default_rng(seed)
    p.b1 = rng.normal(0, 0.2, MO.HIDDEN)
    p.b2 = rng.normal(0, 0.2, MO.HIDDEN)
    p.b3 = np.array([0.3])
    p.e_scale, p.e_shift = 0.7, -1.5
    return p


def theta(p):
    return np.concatenate([p.W1.ravel(), p.b1, p.W2.ravel(), p.b2, p.w3, p.b3]).astype(np.float32)


def mlp_engine(state, p, lo=None, hi=None, judge_every=100, **kw):
    e = engine_for(state, "mlp", **kw)
    e.set_mlp(theta(p), p.g_mean, p.g_inv_std, lo, hi, p.e_scale, p.e_shift, judge_every)
    return e


def test_mlp_energy_and_forces_match_oracle():
    s = liquid_state(4, temperature=100.0)  # 256 atoms, ~78 neighbours within r_c
    p = oracle_params(s)
    e = mlp_engine(s, p)
    e.update_force()
    d = e.download(("force", "pot"))
    E, F, _ = MO.energy_forces(p, s.pos, s.box)
    # Particle::potential holds twice the per-atom energy (PE = sum / 2)
    assert rms(d["pot"] / 2.0 - E) <= 0.02 * rms(E)
    assert rms(d["force"] - F) <= 0.03 * rms(F)
    assert np.abs(F).max() > 1.0


def test_mlp_step_is_velocity_verlet_with_surrogate_forces():
    """One NVE step: positions and velocities from the surrogate forces at the
    old and the new positions (integrator.rs:28-53)."""
    s = liquid_state(4, temperature=100.0)
    p = oracle_params(s, seed=5)
    e = mlp_engine(s, p)
    e.update_force()
    dt = 0.002
    e.step(dt, 1)
    d = e.download(("x", "v"))
    _, F0, _ = MO.energy_forces(p, s.pos, s.box)
    c = dt / (2.0 * AR_MASS)
    v1 = s.vel + F0 * c
    x1 = np.mod(s.pos + v1 * dt, s.box)
    _, F1, _ = MO.energy_forces(p, x1, s.box)
    v2 = v1 + F1 * c
    dx = d["x"] - x1
    dx -= s.box * np.round(dx / s.box)
    assert np.abs(dx).max() < 0.03 * np.abs(F0).max() * c * dt + 1e-12
    assert rms(d["v"] - v2) <= 0.03 * rms(F0 + F1) * c + 1e-15


@pytest.mark.parametrize("thermo", [False, True])
def test_mlp_fallback_is_the_lj_path(thermo):
    """A training range no feature fits in: every launch falls back to the
    LJ forces, and the trajectory is bitwise that of the FP32 LJ path."""
    import moldyn_amd as M
    s = sc_state(8, a=0.55, temperature=120.0, seed=7, jitter=0.04)
    p = oracle_params(s)
    lo = np.full(32, 1e9, dtype=np.float32)
    hi = np.full(32, 2e9, dtype=np.float32)
    a = mlp_engine(s, p, lo, hi)
    b = engine_for(s, "fp32_list")
    for e in (a, b):
        e.update_force()
    assert np.array_equal(a.download(("force",))["force"], b.download(("force",))["force"])
    for e in (a, b):
        th = M.thermostat("berendsen", tau=1.0, target_t=150.0) if thermo else None
        ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=5.0) if thermo else None
        e.step(0.002, 300, th, ba)
    da, db = a.download(), b.download()
    assert np.array_equal(da["x"], db["x"]) and np.array_equal(da["v"], db["v"])
    assert np.array_equal(da["box"], db["box"])
    st = a.mlp_stats()
    assert st["fallbacks"] == 300


@pytest.mark.parametrize("npt", [False, True])
def test_mlp_schedule_invariance(npt):
    """Surrogate trajectories do not depend on the skin, on split steps (a
    skin thinner than one step's drift: the second launch copies its drifted
    positions for the features), on capacity-overflow replays, on the block
    size or on graph capture: features and forces sum the pairs within the
    cutoff in slot order, whatever list holds them."""
    import moldyn_amd as M
    s = liquid_state(4, temperature=150.0)
    p = oracle_params(s, seed=11)
    outs = []
    for kw in [dict(), dict(skin=0.002), dict(skin=0.4), dict(list_capacity=2, cell_capacity=1),
               dict(graph_chunk=1), dict(block_size=64)]:
        e = mlp_engine(s, p, judge_every=0, **kw)
        e.update_force()
        th = M.thermostat("berendsen", tau=1.0, target_t=200.0) if npt else None
        ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=5.0) if npt else None
        e.step(0.002, 60, th, ba)
        outs.append((e.download(), e.stats()))
        e.close()
    assert outs[1][1]["split_steps"] >= 10 and outs[0][1]["split_steps"] == 0
    assert outs[3][1]["grows"] >= 1
    for o, st in outs[1:]:
        for key in ("x", "v", "force", "box"):
            assert np.array_equal(o[key], outs[0][0][key]), (key, st)


def test_mlp_periodic_lj_check():
    s = liquid_state(4, temperature=100.0)
    p = oracle_params(s, seed=8)
    e = mlp_engine(s, p, judge_every=10)
    e.update_force()
    e.step(0.001, 50)
    st = e.mlp_stats()
    assert st["checks"] == 5 and st["fallbacks"] == 0
    # an untrained network is far from LJ, and the check says so
    assert 0.1 < st["last_rmse_rel"] < 1e3 and st["max_rmse_rel"] >= st["last_rmse_rel"]
    assert st["last_rms_force"] > 1.0


def test_mlp_kernel_timing_leaves_results_unchanged():
    """Device timestamps of the three surrogate kernels (features, network,
    forces) cover the launches of a graph-replayed run and change nothing."""
    s = liquid_state(4, temperature=100.0)
    p = oracle_params(s, seed=7)
    outs = []
    for prof in (False, True):
        e = mlp_engine(s, p, judge_every=0)
        e.update_force()
        e.set_profiling(prof)
        e.step(0.002, 600)
        outs.append(e.download(("x", "v")))
        if prof:
            launches = e.stats()["launches"]
            for key in ("drift", "mlp_feat", "mlp_net", "step"):
                ms, n = e.kernel_time(key, span=True)
                assert 0.95 * launches <= n <= launches, key
                assert 0.0 < ms / n < 5.0, key
            # the network kernel starts after the feature kernel ends
            ms_net, n_net = e.kernel_time("mlp_net")
            assert 0.8 * launches <= n_net <= launches
        e.close()
    assert np.array_equal(outs[0]["x"], outs[1]["x"]) and np.array_equal(outs[0]["v"], outs[1]["v"])
