"""The HIP path at the benchmark workloads, against the CPU oracle (needs an
MI355X).

Each workload runs at its full size.  Forces of the exact FP64 list path are
compared bit for bit with the oracle's cell-grid update_force (bitwise its
all-pairs restatement of potential.rs:158-216), FP32 forces by relative RMSE
(bound 1e-5 of the RMS force), and trajectories either bit for bit (NVE) or
to FP64 reduction-order tolerance (rtol 1e-10) where Berendsen feedback reads
global sums (thermostat.rs:24-44, barostat.rs:21-36).  Long runs are checked
through properties that do not depend on the run length: energy drift, the
force error at the final state, the oracle's macro folds of that state.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.common import AR_MASS, engine_for, rms, sc_state

pytestmark = pytest.mark.gpu

DT = 0.002
SIGMA = 0.3418


def oracle_forces(x, v, box):
    s = O.State([len(x)], [AR_MASS], box, pos=x, vel=v)
    O.update_force_cells(s)
    return s


def close_pos(a, b, box, tol):
    d = a - b
    d -= box * np.round(d / box)
    return float(np.abs(d).max()) <= tol


def test_nve_64k_fp64_list_100_steps_bitwise():
    """64,000-atom argon (SC 40^3, a = 3.338339, +-0.02 a jitter, 120 K), NVE,
    skin 0.3 sigma: from the state after 5,000 steps (the lattice has
    dissolved and pairs interact), 100 steps of the FP64 list path equal the
    oracle's (integrator.rs:14-64 with potential.rs:158-216) bit for bit."""
    s = sc_state(40, temperature=120.0, seed=2, jitter=0.02, jitter_seed=2)
    e = engine_for(s, "fp64_list", skin=0.3 * SIGMA)
    e.update_force()
    e.step(DT, 5000)
    mid = e.download(("x", "v", "box"))
    e.upload(mid["x"], mid["v"], mid["box"])
    e.update_force()
    e.step(DT, 100)
    d = e.download(("x", "v", "force", "pot", "virial"))
    e.close()
    ref = O.State([s.n], [AR_MASS], mid["box"], pos=mid["x"], vel=mid["v"])
    tab = O.potential_table(1)
    O.update_force_cells(ref, tab)
    for _ in range(100):
        O.verlet_step(ref, DT, tab, cells=True)
    print("atoms with force", np.count_nonzero(ref.force.any(axis=1)))
    assert np.count_nonzero(ref.force.any(axis=1)) > 300
    for k, r in (("x", ref.pos), ("v", ref.vel), ("force", ref.force), ("pot", ref.pot), ("virial", ref.temp)):
        assert np.array_equal(d[k], r), k


def test_nve_64k_fp32_10000_steps():
    """The same workload in FP32 for its 10,000 steps (20 ps): the total
    energy drifts less than 1e-5 of itself per ns, and the final state's
    FP32 forces are within 1e-5 RMS of the oracle's."""
    s = sc_state(40, temperature=120.0, seed=2, jitter=0.02, jitter_seed=2)
    e = engine_for(s, "fp32_list", skin=0.3 * SIGMA)
    e.update_force()
    m0 = e.macro()
    e.step(DT, 10000)
    m1 = e.macro()
    d = e.download(("x", "v", "box"))
    e.close()
    e0 = m0["kinetic"] + m0["potential"]
    e1 = m1["kinetic"] + m1["potential"]
    ns = 10000 * DT * 1e-3
    drift = abs(e1 - e0) / abs(e0) / ns
    print("relative energy drift per ns", drift)
    assert drift < 1e-5
    ref = oracle_forces(d["x"], d["v"], d["box"])
    f = engine_for(ref, "fp32_list")
    f.update_force()
    got = f.download(("force",))["force"]
    f.close()
    assert np.count_nonzero(ref.force.any(axis=1)) > 300
    rel = rms(got - ref.force) / rms(ref.force)
    print("fp32 force rmse / rms", rel, "max abs", np.abs(got - ref.force).max())
    assert rel <= 1e-5
    assert np.array_equal(got.any(axis=1), ref.force.any(axis=1))


@pytest.fixture(scope="module")
def npt_1m_state():
    """1,000,000-atom argon (SC 100^3, a = 3.338339, box 333.8339, 273.15 K,
    seed 3) after 1,500 FP32 steps under Berendsen T (tau 10, 300 K) and P
    (beta 1, tau 5, 1.01325): the barostat has compressed the gas and pairs
    interact."""
    import moldyn_amd as M
    s = sc_state(100, temperature=273.15, seed=3)
    e = engine_for(s, "fp32_list", skin=0.5)
    e.update_force()
    th = M.thermostat("berendsen", tau=10.0, target_t=300.0)
    ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=1.01325)
    e.step(DT, 1500, th, ba)
    d = e.download(("x", "v", "box"))
    e.close()
    return d


def test_npt_1m_forces_and_macro_vs_oracle(npt_1m_state):
    """At the 1M-atom state: FP64 list forces, potentials and virials equal
    the oracle's bit for bit; FP32 forces are within 1e-5 RMS; the device's
    temperature, pressure and energies equal the oracle's folds of the same
    state (macro_parameters/*.rs) to reduction-order tolerance."""
    d = npt_1m_state
    assert d["box"][0] < 333.8339
    ref = oracle_forces(d["x"], d["v"], d["box"])
    n_with = int(np.count_nonzero(ref.force.any(axis=1)))
    print("atoms with force", n_with, "box", d["box"])
    assert n_with > 10000
    e = engine_for(ref, "fp64_list")
    e.update_force()
    g = e.download(("force", "pot", "virial"))
    m = e.macro()
    e.close()
    assert np.array_equal(g["force"], ref.force)
    assert np.array_equal(g["pot"], ref.pot)
    assert np.array_equal(g["virial"], ref.temp)
    mv = O.center_of_mass_velocity(ref, 0)
    th = O.thermal_energy(ref, 0, mv)
    assert abs(m["temperature"] - O.temperature(th, ref.n)) <= 1e-12 * O.temperature(th, ref.n)
    p = O.pressure(ref, 0, mv)
    assert abs(m["pressure"] - p) <= 1e-12 * abs(p)
    assert abs(m["kinetic"] - O.kinetic_energy(ref, 0)) <= 1e-12 * O.kinetic_energy(ref, 0)
    pe = O.potential_energy(ref, 0)
    assert abs(m["potential"] - pe) <= 1e-12 * abs(pe)
    f = engine_for(ref, "fp32_list")
    f.update_force()
    got = f.download(("force",))["force"]
    f.close()
    rel = rms(got - ref.force) / rms(ref.force)
    print("fp32 force rmse / rms", rel)
    assert rel <= 1e-5
    assert np.array_equal(got.any(axis=1), ref.force.any(axis=1))


def test_npt_1m_fp64_list_20_steps_vs_oracle(npt_1m_state):
    """20 Berendsen NPT steps of the FP64 list path from the 1M-atom state
    track the oracle's (cell-grid forces) to reduction-order tolerance."""
    import moldyn_amd as M
    d = npt_1m_state
    ref = O.State([len(d["x"])], [AR_MASS], d["box"], pos=d["x"], vel=d["v"])
    e = engine_for(ref, "fp64_list", skin=0.5)
    e.update_force()
    th = M.thermostat("berendsen", tau=10.0, target_t=300.0)
    ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=1.01325)
    e.step(DT, 20, th, ba)
    g = e.download(("x", "v", "box"))
    e.close()
    tab = O.potential_table(1)
    O.update_force_cells(ref, tab)
    oth, oba = O.Thermostat.berendsen(10.0), O.Barostat.berendsen(1.0, 5.0)
    for _ in range(20):
        O.verlet_step(ref, DT, tab, oba, 1.01325, oth, 300.0, cells=True)
    assert np.allclose(g["box"], ref.box, rtol=1e-10, atol=0)
    assert close_pos(g["x"], ref.pos, ref.box, 1e-9)
    assert np.allclose(g["v"], ref.vel, rtol=1e-10, atol=1e-12)
    assert abs(th.lambda_ - oth.lambda_) < 1e-12 and abs(ba.myu - oba.myu) < 1e-12


def replica_batch(n_rep=256, seed=4):
    """N_r = round(exp(U[ln 500, ln 32768])) (seed 4) on ceil(N^(1/3))^3
    simple-cubic sites filled in index order, a_r ~ U[3.0, 4.2], Berendsen
    T0_r ~ U[80, 400] K and P0_r ~ U[0.5, 5.0]; MB velocities at T0_r."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(500), np.log(32768), n_rep))).astype(np.int64)
    a = rng.uniform(3.0, 4.2, n_rep)
    t0 = rng.uniform(80.0, 400.0, n_rep)
    p0 = rng.uniform(0.5, 5.0, n_rep)
    states = []
    for r in range(n_rep):
        side = int(np.ceil(counts[r] ** (1.0 / 3.0) - 1e-9))
        pos = O.lattice_u([side] * 3, a[r])[:counts[r]]
        vel = O.mb_velocities(int(counts[r]), t0[r], AR_MASS, 1000 + r)
        states.append(O.State([int(counts[r])], [AR_MASS], [side * a[r]] * 3, pos=pos, vel=vel))
    return counts, t0, p0, states


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_replicas_256_vs_oracle(mode):
    """The 256-replica batch (about 2.3M atoms) in one engine, 100 steps
    under each replica's own Berendsen T and P targets (integrator.rs:18-27):
    the smallest, the median and the largest replica each track the oracle
    run alone (FP64 to reduction-order tolerance; FP32 within 1e-9 nm)."""
    import moldyn_amd as M
    from moldyn_amd import _lib as L
    counts, t0, p0, states = replica_batch()
    R = len(states)
    boxes = np.array([s.box for s in states])
    e = M.Engine(counts[:, None], [AR_MASS], boxes, force_mode=mode, skin=0.2)
    e.upload(np.concatenate([s.pos for s in states]), np.concatenate([s.vel for s in states]), boxes)
    e.update_force()
    th = (L.MdThermostat * R)(*[M.thermostat("berendsen", 10.0, float(t0[r])) for r in range(R)])
    ba = (L.MdBarostat * R)(*[M.barostat("berendsen", 1.0, 5.0, float(p0[r])) for r in range(R)])
    e.step(DT, 100, th, ba)
    d = e.download(("x", "v", "box"))
    e.close()
    order = np.argsort(counts, kind="stable")
    picks = [int(order[0]), int(order[R // 2]), int(order[-1])]
    offs = np.concatenate([[0], np.cumsum(counts)])
    tab = O.potential_table(1)
    for r in picks:
        ref = states[r].copy()
        O.update_force_cells(ref, tab)
        oth, oba = O.Thermostat.berendsen(10.0), O.Barostat.berendsen(1.0, 5.0)
        for _ in range(100):
            O.verlet_step(ref, DT, tab, oba, float(p0[r]), oth, float(t0[r]), cells=True)
        sl = slice(offs[r], offs[r + 1])
        print("replica", r, "atoms", counts[r], "box", d["box"][r], ref.box)
        assert np.allclose(d["box"][r], ref.box, rtol=1e-10, atol=0), r
        tol = 1e-9 if mode == "fp32_list" else 1e-10
        assert close_pos(d["x"][sl], ref.pos, ref.box, tol), r
        assert np.allclose(d["v"][sl], ref.vel, rtol=1e-10, atol=1e-12 if mode != "fp32_list" else 1e-9), r
        assert abs(th[r].lambda_ - oth.lambda_) < 1e-12 and abs(ba[r].myu - oba.myu) < 1e-12, r


def test_surrogate_262k():
    """Learned surrogate at its inference workload: 64^3 SC argon
    (262,144 atoms) at a = 0.47 nm equilibrated at 150 K, trained on the GPU
    from engine-generated frames.  Forces of a sample of atoms match the
    FP64 surrogate oracle within 3 % RMS (BF16 network); a 200-step NPT run
    stays in the training range and its periodic Lennard-Jones check reads
    a force error below 2 % of the LJ RMS force."""
    import moldyn_amd as M
    from moldyn_amd import mlp
    from oracle import mlp_oracle as MO
    fr = mlp.make_frames(n_side=16, n_frames=2000, runs=20, equil=400, every=20, seed=5)
    sur = mlp.train(fr, epochs=40, batch_frames=16, lr=2e-3, lr_final=1e-4)
    n_side, a = 64, 0.47
    n = n_side ** 3
    box = np.full(3, n_side * a)
    rng = np.random.default_rng(6)
    pos = np.mod(O.lattice_u([n_side] * 3, a) + rng.uniform(-0.02, 0.02, (n, 3)) * a, box)
    vel = O.mb_velocities(n, 150.0, AR_MASS, 6)
    probe = M.Engine([n], [AR_MASS], box, force_mode="fp32_list")
    probe.upload(pos, vel, box)
    probe.update_force()
    probe.step(DT, 500, M.thermostat("berendsen", tau=0.1, target_t=150.0))
    p_eq = probe.macro()["pressure"]
    st = probe.download(("x", "v", "box"))
    probe.close()
    e = M.Engine([n], [AR_MASS], st["box"], force_mode="mlp")
    sur.apply(e, judge_every=100, range_check=True)
    e.upload(st["x"], st["v"], st["box"])
    e.update_force()
    f_gpu = e.download(("force",))["force"]
    # the surrogate oracle on the local environment of sampled atoms: every
    # atom within 2 r_c of atom i, unwrapped around it, so that the features
    # of i and of its neighbours are complete
    p = MO.Params()
    th = sur.theta.astype(np.float64)
    off = 0
    for name, shape in (("W1", (MO.HIDDEN, MO.NK)), ("b1", (MO.HIDDEN,)), ("W2", (MO.HIDDEN, MO.HIDDEN)),
                        ("b2", (MO.HIDDEN,)), ("w3", (MO.HIDDEN,)), ("b3", (1,))):
        k = int(np.prod(shape))
        setattr(p, name, th[off:off + k].reshape(shape))
        off += k
    p.g_mean = np.asarray(sur.g_mean, dtype=np.float64)
    p.g_inv_std = np.asarray(sur.g_inv_std, dtype=np.float64)
    p.e_scale, p.e_shift = float(sur.e_scale), float(sur.e_shift)
    x, b = st["x"], st["box"]
    errs, refs = [], []
    for i in np.random.default_rng(9).choice(n, 48, replace=False):
        dd = x - x[i]
        dd -= b * np.round(dd / b)
        near = np.nonzero((dd ** 2).sum(1) <= (2.0 * MO.RC + 0.05) ** 2)[0]
        loc = dd[near] + 50.0  # one open cluster in a box far larger than it
        k = int(np.nonzero(near == i)[0][0])
        _, F, _ = MO.energy_forces(p, loc, np.full(3, 1000.0))
        errs.append(f_gpu[i] - F[k])
        refs.append(F[k])
    rel = rms(np.array(errs)) / rms(np.array(refs))
    print("sampled surrogate force rmse / rms vs FP64 oracle", rel)
    assert rel <= 0.03
    th2 = M.thermostat("berendsen", tau=1.0, target_t=150.0)
    ba2 = M.barostat("berendsen", beta=1.0, tau=50.0, target_p=p_eq)
    e.step(DT, 200, th2, ba2)
    ms = e.mlp_stats()
    e.close()
    print("surrogate stats", ms)
    assert ms["checks"] == 2 and ms["fallbacks"] == 0
    assert ms["max_rmse_rel"] <= 0.02
