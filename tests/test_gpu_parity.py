"""HIP path vs the CPU oracle, through the C ABI (needs an MI355X).

Bit-exact where the arithmetic is the reference's FP64 in the reference's
order (all-pairs and FP64-list update_force, NVE trajectories); FP64
reduction-order tolerance where Berendsen / Nose-Hoover feedback reads global
sums (rtol 1e-10); FP32 forces against the north-star bound (force RMSE
<= 1e-5 of the RMS force).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.common import AR_MASS, engine_for, gas_state, liquid_state, rms, sc_state

pytestmark = pytest.mark.gpu

FP64_MODES = ["fp64_allpairs", "fp64_list"]
ALL_MODES = FP64_MODES + ["fp32_list"]


def f8(x):
    return "%.8f" % x


def two_body(golden, key):
    c = golden[key]
    return O.State([2], [c["mass"]], c["box"], pos=c["pos"], vel=c["vel"])


@pytest.mark.parametrize("mode", ALL_MODES)
def test_update_force_golden(golden, mode):
    c = golden["update_force_lennard_jones"]
    s = O.State([2], [c["mass"]], c["box"], pos=c["pos"])
    e = engine_for(s, mode)
    e.update_force()
    f = e.download()["force"]
    O.update_force(s)
    if mode != "fp32_list":
        assert [f8(x) for x in f[0]] == c["force_p1"]
        assert np.array_equal(f, s.force)
    else:
        assert np.allclose(f, s.force, rtol=1e-6, atol=0)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_verlet_with_lennard_jones(golden, mode):
    """solver/src/lib.rs:109-262 through md_step one step at a time."""
    c = golden["verlet_with_lennard_jones"]
    s = two_body(golden, "verlet_with_lennard_jones")
    e = engine_for(s, mode)
    e.update_force()
    O.update_force(s)
    for k, ref in enumerate(c["steps"]):
        if k > 0:
            e.step(c["dt"], 1)
            O.verlet_step(s, c["dt"])
        d = e.download()
        if mode != "fp32_list":
            assert [[f8(x) for x in row] for row in d["x"]] == ref["pos"], k
            assert [[f8(x) for x in row] for row in d["v"]] == ref["vel"], k
            assert [[f8(x) for x in row] for row in d["force"]] == ref["force"], k
            assert np.array_equal(d["x"], s.pos) and np.array_equal(d["v"], s.vel)
            assert np.array_equal(d["force"], s.force)
        else:
            # FP32 pair forces: 1e-6 relative on F, far below the 8th decimal on x, v
            assert np.allclose(d["force"], s.force, rtol=1e-6, atol=1e-12), k
            assert np.allclose(d["x"], s.pos, rtol=0, atol=1e-11), k
            assert np.allclose(d["v"], s.vel, rtol=0, atol=1e-9), k


@pytest.mark.parametrize("mode", ALL_MODES)
def test_verlet_lj_1000_iterations(golden, mode):
    c = golden["verlet_lj_1000_iterations"]
    s = two_body(golden, "verlet_lj_1000_iterations")
    e = engine_for(s, mode)
    e.update_force()
    e.step(c["dt"], c["n_steps"])
    d = e.download()
    m = e.macro()
    if mode == "fp32_list":
        gx = np.array([[float(x) for x in row] for row in c["final_pos"]])
        gv = np.array([[float(x) for x in row] for row in c["final_vel"]])
        assert np.allclose(d["x"], gx, atol=1e-7) and np.allclose(d["v"], gv, atol=1e-7)
        assert abs(m["pressure"] - float(c["pressure"])) < 1e-6
        return
    assert [[f8(x) for x in row] for row in d["x"]] == c["final_pos"]
    assert [[f8(x) for x in row] for row in d["v"]] == c["final_vel"]
    assert [[f8(x) for x in row] for row in d["force"]] == c["final_force"]
    assert f8(m["kinetic"]) == c["kinetic"]
    assert f8(m["thermal"]) == c["thermal"]
    assert f8(m["potential"]) == c["potential"]
    assert f8(m["temperature"] / 100.0) == c["temperature_over_100"]
    assert f8(m["pressure"]) == c["pressure"]
    if mode != "fp32_list":
        tab = O.potential_table(1)
        O.update_force(s, tab)
        for _ in range(c["n_steps"]):
            O.verlet_step(s, c["dt"], tab)
        assert np.array_equal(d["x"], s.pos) and np.array_equal(d["v"], s.vel)


def test_energies_temperature_pressure(golden):
    c = golden["energies"]
    s = two_body(golden, "energies")
    e = engine_for(s, "fp64_allpairs")
    e.update_force()
    m = e.macro()
    assert list(m["com_velocity"]) == c["com_velocity"]
    assert f8(m["kinetic"]) == c["kinetic"]
    assert f8(m["thermal"]) == c["thermal"]
    assert f8(m["potential"]) == c["potential"]
    assert f8(m["temperature"]) == golden["temperature"]["temperature"]
    assert f8(m["pressure"]) == golden["pressure"]["pressure"]


@pytest.mark.parametrize("mode", ALL_MODES)
def test_dense_liquid_forces(mode):
    s = liquid_state(5)
    e = engine_for(s, mode)
    e.update_force()
    d = e.download()
    O.update_force(s)
    if mode == "fp32_list":
        err = d["force"] - s.force
        rel_rmse = rms(err) / rms(s.force)
        print("fp32 force rmse/rms", rel_rmse, "max", np.abs(err).max() / rms(s.force))
        assert rel_rmse <= 1e-5
        assert np.allclose(d["pot"], s.pot, rtol=1e-4, atol=1e-4 * np.abs(s.pot).max())
        assert abs(d["virial"].sum() - s.temp.sum()) <= 1e-5 * np.abs(s.temp).sum()
    else:
        assert np.array_equal(d["force"], s.force)
        assert np.array_equal(d["pot"], s.pot)
        assert np.array_equal(d["virial"], s.temp)


def adversarial_gas(seed=7):
    """Gas with planted pairs at r_cut exactly, one ulp either side, across
    the periodic face and at |dx| = L / 2."""
    sigma, eps, rc, uc = O.new_lennard_jones(0.3418, 1.712)
    box = 6.0
    s = gas_state(300, box, seed)
    p = s.pos
    p[0] = [1.0, 1.0, 1.0]
    p[1] = [1.0 + rc, 1.0, 1.0]
    p[2] = [3.0, 3.0, 3.0]
    p[3] = [3.0 + np.nextafter(rc, 0), 3.0, 3.0]
    p[4] = [2.0, 4.5, 2.0]
    p[5] = [2.0 + np.nextafter(rc, 1), 4.5, 2.0]
    p[6] = [0.1, 5.0, 5.0]
    p[7] = [box - (rc - 0.1), 5.0, 5.0]
    p[8] = [0.0, 0.5, 0.5]
    p[9] = [box / 2, 0.5, 0.5]
    return s


@pytest.mark.parametrize("mode", FP64_MODES)
def test_pair_set_exact_adversarial(mode):
    s = adversarial_gas()
    e = engine_for(s, mode)
    e.update_force()
    d = e.download()
    O.update_force(s)
    assert np.array_equal(d["force"], s.force)
    assert np.array_equal(d["virial"], s.temp)


def test_fp32_pair_set_exact_adversarial():
    """FP32 forces, but the pair set is decided in FP64 by the reference
    criterion: atoms with/without an interaction agree exactly."""
    s = adversarial_gas()
    e = engine_for(s, "fp32_list")
    e.update_force()
    d = e.download()
    O.update_force(s)
    assert np.array_equal(d["force"] != 0, s.force != 0)
    assert np.array_equal(d["virial"] != 0, s.temp != 0)


def run_oracle(s, n, dt, th=None, t0=0.0, ba=None, p0=0.0):
    tab = O.potential_table(s.n_types)
    O.update_force(s, tab)
    for _ in range(n):
        O.verlet_step(s, dt, tab, ba, p0, th, t0)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_config1_npt(mode):
    """1,000-atom argon gas, Berendsen T (tau 10, 300 K) and P (1, 5; 1.01325),
    dt 0.002, 1,000 steps: x, v, box, lambda, mu vs the oracle."""
    import moldyn_amd as M
    s = sc_state(10, seed=1)
    e = engine_for(s, mode)
    e.update_force()
    th = M.thermostat("berendsen", tau=10.0, target_t=300.0)
    ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=1.01325)
    e.step(0.002, 1000, th, ba)
    d = e.download()
    oth = O.Thermostat.berendsen(10.0)
    oba = O.Barostat.berendsen(1.0, 5.0)
    run_oracle(s, 1000, 0.002, oth, 300.0, oba, 1.01325)
    assert np.allclose(d["box"], s.box, rtol=1e-10, atol=0)
    assert np.allclose(d["x"], s.pos, rtol=1e-10, atol=1e-10)
    assert np.allclose(d["v"], s.vel, rtol=1e-10, atol=1e-12)
    assert abs(th.lambda_ - oth.lambda_) < 1e-12
    assert abs(ba.myu - oba.myu) < 1e-12
    m = e.macro()
    mv = O.center_of_mass_velocity(s, 0)
    te = O.temperature(O.thermal_energy(s, 0, mv), 1000)
    assert abs(m["temperature"] - te) < 1e-9 * te
    assert abs(m["pressure"] - O.pressure(s, 0, mv)) < 1e-9 * abs(O.pressure(s, 0, mv))


@pytest.mark.parametrize("mode", ALL_MODES)
def test_dense_nvt_nose_hoover(mode):
    """Liquid argon under Nose-Hoover, 100 steps: FP64 paths track the oracle
    to reduction-order tolerance; FP32 within a trajectory bound."""
    import moldyn_amd as M
    s = liquid_state(4, temperature=100.0)
    e = engine_for(s, mode)
    e.update_force()
    th = M.thermostat("nose_hoover", tau=0.5, target_t=90.0, psi=0.01)
    e.step(0.002, 100, th)
    d = e.download()
    oth = O.Thermostat.nose_hoover(0.5, psi=0.01)
    run_oracle(s, 100, 0.002, oth, 90.0)
    tol = 1e-8 if mode != "fp32_list" else 1e-3
    dx = d["x"] - s.pos
    dx -= s.box * np.round(dx / s.box)
    assert np.abs(dx).max() < tol
    assert abs(th.psi - oth.psi) < tol
    assert abs(th.lambda_ - oth.lambda_) < tol


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_list_schedule_invariance(mode):
    """Trajectories do not depend on when lists are rebuilt, on the cell grid,
    on capacity-overflow replays or on graph capture."""
    s = liquid_state(4, temperature=150.0)
    outs = []
    # list_capacity=600: more LDS list columns than the pool holds at once,
    # so dense bricks build their centre particles in passes
    for kw in [dict(), dict(skin=0.05), dict(skin=0.4, cells_per_atom=0.2),
               dict(list_capacity=2, cell_capacity=1), dict(graph_chunk=1), dict(list_capacity=600)]:
        e = engine_for(s, mode, **kw)
        e.update_force()
        e.step(0.002, 200)
        outs.append(e.download())
        st = e.stats()
        print(kw, st)
        e.close()
    for o in outs[1:]:
        assert np.array_equal(o["x"], outs[0]["x"])
        assert np.array_equal(o["v"], outs[0]["v"])
        assert np.array_equal(o["force"], outs[0]["force"])


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_overflow_replay_under_thermostat_and_barostat(mode):
    """A capacity overflow replays the chunk from its snapshot; the replay's
    first controls come from the sums at the chunk start, so an NPT
    trajectory with replays equals one without."""
    import moldyn_amd as M
    s = sc_state(8, a=0.55, temperature=120.0, seed=7, jitter=0.04)
    outs = []
    for kw in [dict(), dict(list_capacity=2, cell_capacity=1)]:
        e = engine_for(s, mode, **kw)
        e.update_force()
        th = M.thermostat("berendsen", tau=1.0, target_t=150.0)
        ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=5.0)
        e.step(0.002, 300, th, ba)
        outs.append((e.download(), th.lambda_, ba.myu, e.stats()))
        e.close()
    (a, la, ma, sa), (b, lb, mb, sb) = outs
    assert sb["retries"] >= 1
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["v"], b["v"])
    assert np.array_equal(a["box"], b["box"]) and la == lb and ma == mb


@pytest.mark.parametrize("mode,npt", [("fp64_list", False), ("fp32_list", False), ("fp32_list", True)])
def test_chunk_boundaries(mode, npt):
    """md_step runs in chunks of at least 512 steps; the list path steps on
    position + kicked-velocity records packed at a chunk's start and written
    back to the full records by its last step.  1100 steps in one call
    (chunks of 512, 512, 76) equal 600 + 500 in two calls, and on the FP64
    path the oracle's 1100 steps, bit for bit."""
    import moldyn_amd as M
    s = liquid_state(4, temperature=150.0)
    outs = []
    for parts in ([1100], [600, 500]):
        e = engine_for(s, mode)
        e.update_force()
        for k in parts:
            th = M.thermostat("berendsen", tau=1.0, target_t=150.0) if npt else None
            ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=5.0) if npt else None
            e.step(0.002, k, th, ba)
        outs.append(e.download())
        e.close()
    a, b = outs
    for key in ("x", "v", "force", "pot", "virial", "box"):
        assert np.array_equal(a[key], b[key]), key
    if mode == "fp64_list":
        ref = s.copy()
        tab = O.potential_table(1)
        O.update_force(ref, tab)
        for _ in range(1100):
            O.verlet_step(ref, 0.002, tab)
        assert np.array_equal(a["x"], ref.pos) and np.array_equal(a["v"], ref.vel)
        assert np.array_equal(a["force"], ref.force)


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_reupload_after_capacity_growth(mode):
    """Capacity overflows reallocate and clear the cell grid; the grid epochs
    keep counting, so a second run from a re-upload into the same engine
    meets no stale cell of the first and equals it bit for bit, and a run
    in a fresh engine with roomy capacities."""
    import moldyn_amd as M
    s = sc_state(8, a=0.55, temperature=120.0, seed=7, jitter=0.04)

    def run(e):
        e.upload(s.pos, s.vel, s.box)
        e.update_force()
        th = M.thermostat("berendsen", tau=1.0, target_t=150.0)
        ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=5.0)
        e.step(0.002, 300, th, ba)
        return e.download(), e.stats()

    e = engine_for(s, mode, skin=0.05, list_capacity=2, cell_capacity=1)
    a, sa = run(e)
    b, sb = run(e)
    e.close()
    f = engine_for(s, mode, skin=0.05, list_capacity=64, cell_capacity=64)
    c, _ = run(f)
    f.close()
    assert sa["list_capacity"] > 2 and sa["cell_capacity"] > 1  # the grid was reallocated
    for key in ("x", "v", "force", "box"):
        assert np.array_equal(a[key], b[key]), key
        assert np.array_equal(a[key], c[key]), key


@pytest.mark.parametrize("mode,npt", [("fp64_list", False), ("fp64_list", True), ("fp32_list", True)])
def test_split_steps(mode, npt):
    """A skin thinner than one step's drift: each step drifts first, rebuilds
    its lists at the drifted positions, then computes forces and the second
    half kick (three launches).  The trajectory equals the one with a roomy
    skin bit for bit, and the oracle's on the FP64 NVE path."""
    import moldyn_amd as M
    s = liquid_state(4, temperature=150.0)
    outs = []
    for skin in (0.002, 0.3):
        e = engine_for(s, mode, skin=skin)
        e.update_force()
        th = M.thermostat("berendsen", tau=1.0, target_t=200.0) if npt else None
        ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=5.0) if npt else None
        e.step(0.002, 60, th, ba)
        outs.append((e.download(), e.stats()))
        e.close()
    (a, sa), (b, sb) = outs
    assert sa["split_steps"] >= 10 and sb["split_steps"] == 0
    for key in ("x", "v", "force", "box"):
        assert np.array_equal(a[key], b[key]), key
    if not npt:
        ref = s.copy()
        tab = O.potential_table(1)
        O.update_force(ref, tab)
        for _ in range(60):
            O.verlet_step(ref, 0.002, tab)
        assert np.array_equal(a["x"], ref.pos) and np.array_equal(a["v"], ref.vel)
        assert np.array_equal(a["force"], ref.force)


def test_fp64_list_equals_allpairs_trajectory():
    s = liquid_state(4, temperature=150.0)
    a = engine_for(s, "fp64_allpairs")
    b = engine_for(s, "fp64_list")
    for e in (a, b):
        e.update_force()
        e.step(0.002, 300)
    da, db = a.download(), b.download()
    assert np.array_equal(da["x"], db["x"]) and np.array_equal(da["v"], db["v"])


@pytest.mark.parametrize("mode", ALL_MODES)
def test_replicas_match_single(mode):
    """A batch of uneven replicas equals each replica run alone."""
    import moldyn_amd as M
    sizes = [(3, 3.0), (5, 3.6), (2, 4.1), (4, 0.6)]
    states = []
    for k, (n, a) in enumerate(sizes):
        st = sc_state(n, a=a, temperature=150.0 + 40 * k, seed=10 + k, jitter=0.05, jitter_seed=20 + k)
        states.append(st)
    counts = [[st.n] for st in states]
    boxes = [st.box for st in states]
    e = M.Engine(counts, [AR_MASS], boxes, force_mode=mode)
    e.upload(np.concatenate([st.pos for st in states]), np.concatenate([st.vel for st in states]), boxes)
    e.update_force()
    th = (M.engine.L.MdThermostat * 4)(*[M.thermostat("berendsen", 10.0, 100.0 + 50 * k) for k in range(4)])
    ba = (M.engine.L.MdBarostat * 4)(*[M.barostat("berendsen", 1.0, 5.0, 0.5 + k) for k in range(4)])
    e.step(0.002, 150, th, ba)
    d = e.download()
    off = 0
    for k, st in enumerate(states):
        one = engine_for(st, mode)
        one.update_force()
        th1 = M.thermostat("berendsen", 10.0, 100.0 + 50 * k)
        ba1 = M.barostat("berendsen", 1.0, 5.0, 0.5 + k)
        one.step(0.002, 150, th1, ba1)
        d1 = one.download()
        sl = slice(off, off + st.n)
        assert np.array_equal(d["x"][sl], d1["x"]), k
        assert np.array_equal(d["v"][sl], d1["v"]), k
        assert np.array_equal(d["box"][k], d1["box"]), k
        assert th[k].lambda_ == th1.lambda_ and ba[k].myu == ba1.myu
        # each replica alone through the oracle under its own targets
        # (integrator.rs:18-27, thermostat.rs:24-44, barostat.rs:21-36)
        ref = st.copy()
        oth, oba = O.Thermostat.berendsen(10.0), O.Barostat.berendsen(1.0, 5.0)
        run_oracle(ref, 150, 0.002, oth, 100.0 + 50 * k, oba, 0.5 + k)
        tol = 1e-10 if mode != "fp32_list" else 1e-6
        assert np.allclose(d["box"][k], ref.box, rtol=1e-10 if mode != "fp32_list" else 1e-8, atol=0), k
        dx = d["x"][sl] - ref.pos
        dx -= ref.box * np.round(dx / ref.box)
        assert np.abs(dx).max() < tol, k
        assert np.allclose(d["v"][sl], ref.vel, rtol=tol, atol=tol), k
        assert abs(th[k].lambda_ - oth.lambda_) < (1e-12 if mode != "fp32_list" else 1e-8), k
        assert abs(ba[k].myu - oba.myu) < (1e-12 if mode != "fp32_list" else 1e-8), k
        off += st.n
        one.close()


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_replica_buffers_after_split_steps(mode):
    """A hot replica whose every step is split (drift, rebuild, forces: the
    step ends in the buffer it started from) beside a cold one whose steps
    alternate buffers: downloads, force uploads and device initialisers of
    the batch address each replica's own current buffer."""
    import moldyn_amd as M
    hot = liquid_state(3, temperature=2000.0, seed=3)
    cold = liquid_state(3, temperature=20.0, seed=4)
    states = [hot, cold]
    counts = [[st.n] for st in states]
    boxes = [st.box for st in states]
    e = M.Engine(counts, [AR_MASS], boxes, force_mode=mode, skin=0.005)
    e.upload(np.concatenate([st.pos for st in states]), np.concatenate([st.vel for st in states]), boxes)
    e.update_force()
    singles = []
    for st in states:
        one = engine_for(st, mode, skin=0.005)
        one.update_force()
        singles.append(one)
    splits = 0
    for call in range(7):
        e.step(0.002, 1)
        for one in singles:
            one.step(0.002, 1)
        splits += e.stats()["split_steps"]
        d = e.download()
        off = 0
        for k, one in enumerate(singles):
            d1 = one.download()
            sl = slice(off, off + states[k].n)
            for key in ("x", "v", "force"):
                assert np.array_equal(d[key][sl], d1[key]), (call, k, key)
            off += states[k].n
    assert splits >= 5
    # forces uploaded into the batch land in each replica's current buffer:
    # the next step's first half kick reads them (integrator.rs:28-34)
    d = e.download()
    f_new = np.random.default_rng(1).normal(0.0, 5.0, d["force"].shape)
    e.upload_forces(f_new, d["pot"], d["virial"])
    got = e.download()
    assert np.array_equal(got["x"], d["x"]) and np.array_equal(got["v"], d["v"])
    if mode == "fp64_list":
        assert np.array_equal(got["force"], f_new)
    else:
        assert np.array_equal(got["force"], f_new.astype(np.float32).astype(np.float64))
    # a device initialiser edits one replica and leaves the other as it is
    e.init_velocities(50.0, seed=9, replica=1)
    after = e.download()
    n0 = hot.n
    assert np.array_equal(after["x"], d["x"]) and np.array_equal(after["v"][:n0], d["v"][:n0])
    ref_v = M.mb_velocities(cold.n, 50.0, AR_MASS, 9)
    assert np.allclose(after["v"][n0:], ref_v, rtol=1e-12, atol=1e-13)
    for one in singles:
        one.close()
    e.close()


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_thermal_sums_with_fast_centre_of_mass(mode):
    """A liquid whose centre of mass moves a thousand times faster than its
    thermal speed: temperature and pressure take v - v_com before squaring
    (energy.rs:25-37, pressure.rs:11-14), so a one-pass sum of m |v|^2
    minus M |v_com|^2 would lose about 1e-10 of them.  Macro values equal the
    oracle's folds to 1e-12, and an NPT run tracks the oracle."""
    import moldyn_amd as M
    s = liquid_state(4, temperature=100.0, seed=6)
    v_th = np.sqrt(O.K_B * 100.0 * 0.01 / AR_MASS)
    s.vel = s.vel + np.array([1000.0 * v_th, -400.0 * v_th, 700.0 * v_th])
    e = engine_for(s, mode)
    e.update_force()
    m = e.macro()
    ref = s.copy()
    O.update_force(ref)
    mv = O.center_of_mass_velocity(ref, 0)
    te = O.temperature(O.thermal_energy(ref, 0, mv), ref.n)
    pr = O.pressure(ref, 0, mv)
    tol = 1e-12 if mode != "fp32_list" else 1e-5
    assert abs(m["temperature"] - te) <= 1e-12 * te
    assert abs(m["pressure"] - pr) <= tol * abs(pr)
    assert abs(m["thermal"] - O.thermal_energy(ref, 0, mv)) <= 1e-12 * m["thermal"]
    assert abs(m["kinetic"] - O.kinetic_energy(ref, 0)) <= 1e-13 * m["kinetic"]
    th = M.thermostat("berendsen", tau=0.5, target_t=150.0)
    ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=50.0)
    e.step(0.002, 100, th, ba)
    d = e.download()
    oth, oba = O.Thermostat.berendsen(0.5), O.Barostat.berendsen(1.0, 5.0)
    run_oracle(s, 100, 0.002, oth, 150.0, oba, 50.0)
    ttol = 1e-10 if mode != "fp32_list" else 1e-4
    assert abs(th.lambda_ - oth.lambda_) < ttol and abs(ba.myu - oba.myu) < ttol
    assert np.allclose(d["box"], s.box, rtol=ttol, atol=0)
    assert np.allclose(d["v"], s.vel, rtol=ttol, atol=ttol)


def big_box_adversarial(seed=12):
    """A 333.8339 nm box (the 1M-atom workload's), where an FP32 coordinate
    has an ulp of 3e-5 nm: a background gas plus planted pairs at
    |dx| = r_cut exactly and one ulp beyond, straddling the periodic faces
    with coordinates next to L, along the diagonal, and at |dx| = L / 2."""
    sigma, eps, rc, uc = O.new_lennard_jones(0.3418, 1.712)
    L = 100 * 3.338339
    rng = np.random.default_rng(seed)
    pos = [rng.uniform(0.0, L, (20000, 3))]
    top = np.nextafter(L, 0.0)

    def pair_at(xi, along, inside):
        # x_j = x_i + delta along one axis, moved ulp by ulp to the last
        # separation the reference criterion keeps (r_abs <= r_cut) or the
        # first it drops; r_abs is |dx| when the other components are 0
        xi = np.array(xi, dtype=np.float64)
        xj = xi.copy()
        xj[along] = xi[along] + rc
        if xj[along] >= L:
            xj[along] -= L

        def r_of(xj):
            d = xj[along] - xi[along]
            if d < -L / 2.0:
                d += L
            elif d > L / 2.0:
                d -= L
            return np.sqrt(d * d)
        while r_of(xj) > rc:
            xj[along] = np.nextafter(xj[along], -np.inf)
        while r_of(np.where(np.arange(3) == along, np.nextafter(xj[along], np.inf), xj)) <= rc:
            xj[along] = np.nextafter(xj[along], np.inf)
        if not inside:
            xj[along] = np.nextafter(xj[along], np.inf)
        return [xi, xj]

    planted = []
    for k, (xi, ax) in enumerate([([10.0, 20.0, 30.0], 0), ([top, 50.0, 60.0], 0), ([70.0, top, 80.0], 1),
                                  ([90.0, 100.0, L - 0.5], 2), ([L - 0.3, 120.0, 130.0], 0),
                                  ([140.0, 150.0, top - 0.6], 2)]):
        planted += pair_at(xi, ax, inside=True)
        q = list(xi)
        q[(ax + 1) % 3] += 7.0
        planted += pair_at(q, ax, inside=False)
    # diagonal pairs straddling the corner at r_cut (FP32 rounding on every axis)
    h = rc / np.sqrt(3.0)
    planted += [[top, top, top], [h - 1e-6 - (L - top), h - 1e-6 - (L - top), h - 1e-6 - (L - top)]]
    planted += [[top - 1.0, top - 2.0, 3.0], [h - 1.0, h - 2.0, 3.0 + h]]
    # |dx| = L / 2 exactly (no minimum-image correction at equality)
    planted += [[0.0, 200.0, 200.0], [L / 2, 200.0, 200.0]]
    pos.append(np.array(planted))
    pos = np.concatenate(pos)
    s = O.State([len(pos)], [AR_MASS], [L] * 3, pos=pos, vel=np.zeros_like(pos))
    O.apply_boundary_conditions(s)
    return s, len(planted)


def test_adversarial_pairs_in_the_1m_atom_box():
    """FP64 list forces equal the oracle's bit for bit at the planted pairs
    of the 333.8 nm box, and the FP32 path (FP32-prefiltered list build)
    keeps exactly the reference's pair set (potential.rs:186-204)."""
    s, n_planted = big_box_adversarial()
    ref = s.copy()
    O.update_force_cells(ref)
    kept = ref.force[-n_planted:].any(axis=1)
    assert kept.sum() >= 12 and (~kept).sum() >= 12
    e = engine_for(s, "fp64_list")
    e.update_force()
    d = e.download(("force", "virial"))
    e.close()
    assert np.array_equal(d["force"], ref.force) and np.array_equal(d["virial"], ref.temp)
    f = engine_for(s, "fp32_list")
    f.update_force()
    g = f.download(("force", "virial"))
    f.close()
    assert np.array_equal(g["force"].any(axis=1), ref.force.any(axis=1))
    assert np.array_equal(g["virial"] != 0, ref.temp != 0)
    assert rms(g["force"] - ref.force) <= 1e-5 * rms(ref.force)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_apply_boundary_conditions_entry(mode):
    """md_apply_boundary_conditions: one conditional wrap per axis
    (core/src/particle.rs:120-141), including coordinates more than a box
    away, which one wrap leaves outside."""
    s = liquid_state(3)
    rng = np.random.default_rng(2)
    s.pos = s.pos + rng.choice([-1.0, 0.0, 1.0, 1.7, -2.2], size=s.pos.shape) * s.box
    e = engine_for(s, mode)
    e.apply_boundary_conditions()
    d = e.download(("x",))
    O.apply_boundary_conditions(s)
    assert np.array_equal(d["x"], s.pos)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_upload_forces_and_momentum(mode):
    """md_upload_forces hands the first half kick its forces
    (integrator.rs:28-34), and md_macro's momentum is the reference's
    sum of m v (macro_parameters/mod.rs:28-34)."""
    s = liquid_state(3, temperature=130.0, seed=8)
    e = engine_for(s, mode)
    ref = s.copy()
    O.update_force(ref)
    f = ref.force * (1.0 + np.random.default_rng(3).uniform(-0.1, 0.1, ref.force.shape))
    if mode == "fp32_list":
        f = f.astype(np.float32).astype(np.float64)
    e.upload_forces(f, ref.pot, ref.temp)
    ref.force[:] = f
    m = e.macro()
    assert np.allclose(m["momentum"], O.momentum(ref, 0), rtol=0, atol=1e-12 * np.abs(ref.vel).sum() * AR_MASS)
    e.step(0.002, 1)
    d = e.download()
    O.verlet_step(ref, 0.002)
    if mode == "fp32_list":
        assert np.allclose(d["v"], ref.vel, rtol=1e-6, atol=1e-9)
    else:
        assert np.array_equal(d["x"], ref.pos) and np.array_equal(d["v"], ref.vel)
    m = e.macro()
    # the device's sum of m v of its own state; the FP64 paths' state is the oracle's
    scale = np.abs(d["v"]).sum() * AR_MASS
    assert np.allclose(m["momentum"], (d["v"] * AR_MASS).sum(0), rtol=0, atol=1e-13 * scale)
    if mode != "fp32_list":
        assert np.allclose(m["momentum"], O.momentum(ref, 0), rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_momentum_invariant(mode):
    """solver/src/lib.rs:59-75: 8 argon atoms on a 2x2x2 grid of edge
    3.338339 in a box of twice that, MB at 273.15 K: every component of the
    summed velocity stays below 1e-12 over 100,000 steps (checked after each
    of the first 200 steps, then every 1000).  The FP32 path sums each
    atom's pair forces in FP32, so where three or more atoms interact at
    once the per-atom sums are antisymmetric only to FP32 rounding: its
    bound is 1e-8."""
    n = 8
    pos = O.lattice_u([2, 2, 2], 3.338339)
    vel = O.mb_velocities(n, 273.15, AR_MASS, 21)
    s = O.State([n], [AR_MASS], [2 * 3.338339] * 3, pos=pos, vel=vel)
    e = engine_for(s, mode)
    e.update_force()
    assert np.abs(e.download(("v",))["v"].sum(0)).max() < 1e-12
    done = 0
    while done < 100000:
        k = 1 if done < 200 else 1000
        e.step(0.002, k)
        done += k
        bound = 1e-12 if mode != "fp32_list" else 1e-8
        assert np.abs(e.download(("v",))["v"].sum(0)).max() < bound, done


@pytest.mark.parametrize("shape", ["liquid", "gas"])
def test_bricks_with_unequal_tiled_edges(shape):
    """Grids whose bricks are a cell longer along one or two axes that are
    tiled, not whole (brick_dims in md_device.h, restated in tests/common.py
    and asserted here): FP64 list forces equal the oracle's; a thin-skin run
    of 200 steps equals the all-pairs path bit for bit."""
    from tests.common import brick_dims, grid_dims
    if shape == "liquid":
        # SC 12^3 at a = 0.48: a 6^3 grid, bricks of 2 x 3 x 3 cells
        s = sc_state(12, a=0.48, temperature=120.0, seed=3, jitter=0.05, jitter_seed=4)
    else:
        # 1,600 atoms in a 20 nm box: a 14 x 14 x 16 grid, bricks of 9 x 9 x 10;
        # the first 1,600 sites of a jittered 12^3 lattice (no overlapping pairs)
        a = 20.0 / 12
        pos = O.lattice_u([12, 12, 12], a)[:1600]
        pos = pos + np.random.default_rng(23).uniform(-0.4, 0.4, pos.shape) * a
        s = O.State([1600], [AR_MASS], [20.0] * 3, pos=pos, vel=O.mb_velocities(1600, 273.15, AR_MASS, 23))
        O.apply_boundary_conditions(s)
    rlist = 2.5 * 0.3418 + 0.3 * 0.3418
    gd = grid_dims(s.box, rlist, int(np.ceil(2.0 * s.n)))
    edge, whole = brick_dims(gd, s.n)
    print("grid", gd, "brick", edge, "whole", whole)
    assert len(set(edge)) > 1 and not any(whole)
    e = engine_for(s, "fp64_list")
    e.update_force()
    f = e.download(("force",))["force"]
    ref = s.copy()
    O.update_force_cells(ref)
    assert np.count_nonzero(ref.force.any(axis=1)) > 100
    assert np.array_equal(f, ref.force)
    e.close()
    thin = engine_for(s, "fp64_list", skin=0.02)
    allp = engine_for(s, "fp64_allpairs")
    for eng in (thin, allp):
        eng.update_force()
        eng.step(0.002, 200)
    da, db = allp.download(), thin.download()
    assert thin.stats()["rebuilds"] >= 5
    assert np.array_equal(da["x"], db["x"]) and np.array_equal(da["v"], db["v"])


@pytest.mark.parametrize("mode", ["fp64_list", "fp64_allpairs"])
def test_multitype_with_several_tiles_per_block(mode):
    """Two types in 64-slot tiles, three tiles per block: every slot of a
    group carries its type (type 1 holds 106 atoms, two tiles, one block),
    so the pair parameters and the t1 <= t2 rule see the right type."""
    import moldyn_amd as M
    rng = np.random.default_rng(5)
    n0, n1 = 150, 106
    box = 2.4
    pos = rng.uniform(0, box, size=(n0 + n1, 3))
    vel = rng.normal(0, 0.2, size=(n0 + n1, 3))
    s = O.State([n0, n1], [AR_MASS, 39.948], [box] * 3, pos=pos, vel=vel)
    tab = O.potential_table(2, {(0, 1): O.new_lennard_jones(0.30, 1.2), (1, 1): O.new_lennard_jones(0.36, 0.9)})
    e = M.Engine([n0, n1], [AR_MASS, 39.948], [box] * 3, force_mode=mode, block_size=64, tiles_per_block=3)
    e.set_pair_potential(0, 1, 0.30, 1.2)
    e.set_pair_potential(1, 1, 0.36, 0.9)
    e.upload(pos, vel, [box] * 3)
    e.update_force()
    d = e.download()
    e.close()
    O.update_force(s, tab)
    assert np.array_equal(d["force"], s.force)
    assert np.array_equal(d["pot"], s.pot) and np.array_equal(d["virial"], s.temp)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_multitype_update_force_and_nve(mode):
    """Two types with distinct potentials: the reference's type-pair rule
    (potential.rs:171-180: only t1 <= t2 accumulates) and per-type masses."""
    import moldyn_amd as M
    rng = np.random.default_rng(3)
    n0, n1 = 150, 106
    box = 2.4
    pos = rng.uniform(0, box, size=(n0 + n1, 3))
    vel = rng.normal(0, 0.2, size=(n0 + n1, 3))
    s = O.State([n0, n1], [AR_MASS, 39.948], [box] * 3, pos=pos, vel=vel)
    p01 = O.new_lennard_jones(0.30, 1.2)
    p11 = O.new_lennard_jones(0.36, 0.9)
    tab = O.potential_table(2, {(0, 1): p01, (1, 1): p11})
    e = M.Engine([n0, n1], [AR_MASS, 39.948], [box] * 3, force_mode=mode)
    e.set_pair_potential(0, 1, 0.30, 1.2)
    e.set_pair_potential(1, 1, 0.36, 0.9)
    e.upload(pos, vel, [box] * 3)
    e.update_force()
    d = e.download()
    O.update_force(s, tab)
    if mode == "fp32_list":
        assert rms(d["force"] - s.force) <= 1e-5 * rms(s.force)
    else:
        assert np.array_equal(d["force"], s.force)
        e.step(0.001, 50)
        for _ in range(50):
            O.verlet_step(s, 0.001, tab)
        d = e.download()
        assert np.array_equal(d["x"], s.pos) and np.array_equal(d["v"], s.vel)


def test_fp32_force_rmse_gas_and_compressed():
    """North-star bound on FP32 pair forces: RMSE <= 1e-5 of the RMS force,
    on a 10x-density gas and on the liquid fixture."""
    for s in (gas_state(4000, 5.0, 11), liquid_state(6)):
        e = engine_for(s, "fp32_list")
        e.update_force()
        f = e.download()["force"]
        O.update_force_cells(s)
        assert rms(f - s.force) <= 1e-5 * rms(s.force)
        e.close()


def test_unimplemented_custom_variants():
    import moldyn_amd as M
    s = sc_state(3)
    e = engine_for(s, "fp32_list")
    with pytest.raises(M.MdError) as ei:
        e.step(0.002, 1, M.thermostat("custom"))
    assert ei.value.code == -4
    with pytest.raises(M.MdError) as ei:
        e.step(0.002, 1, None, M.barostat("custom"))
    assert ei.value.code == -4


def test_standalone_thermostat_barostat_api():
    """md_thermostat_* / md_barostat_* mirror thermostat.rs / barostat.rs."""
    import moldyn_amd as M
    from moldyn_amd import _lib as L
    import ctypes as C
    s = liquid_state(3, temperature=130.0)
    e = engine_for(s, "fp64_allpairs")
    e.update_force()
    O.update_force(s)
    lib = L.load()
    th = M.thermostat("berendsen", tau=0.5, target_t=100.0)
    e._chk(lib.md_thermostat_calculate_lambda(e._h, C.byref(th), 0.002, 0))
    oth = O.Thermostat.berendsen(0.5)
    O.thermostat_calculate_lambda(oth, s, 0.002, 0, 100.0)
    assert abs(th.lambda_ - oth.lambda_) < 1e-13
    e._chk(lib.md_thermostat_update(e._h, C.byref(th), 0.002, 0))
    O.thermostat_update(oth, s, 0.002, 0, 100.0)
    assert np.array_equal(e.download()["v"], s.vel)
    ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=1.0)
    e._chk(lib.md_barostat_calculate_myu(e._h, C.byref(ba), 0.002, 0))
    oba = O.Barostat.berendsen(1.0, 5.0)
    O.barostat_calculate_myu(oba, s, 0.002, 0, 1.0)
    assert abs(ba.myu - oba.myu) < 1e-13
    ba.myu = oba.myu
    e._chk(lib.md_barostat_update(e._h, C.byref(ba), 0))
    O.barostat_update(oba, s, 0)
    d = e.download()
    assert np.array_equal(d["x"], s.pos) and np.array_equal(d["box"], s.box)


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_large_gas_interior_bricks(mode):
    """A box much wider than r_list: cell bricks with halos that wrap the
    periodic faces (the small fixtures have single-brick grids)."""
    s = gas_state(6000, 30.0, 21)
    e = engine_for(s, mode)
    e.update_force()
    d = e.download()
    O.update_force_cells(s)
    assert np.count_nonzero(s.force) > 1000
    if mode == "fp64_list":
        assert np.array_equal(d["force"], s.force)
        assert np.array_equal(d["virial"], s.temp)
    else:
        assert rms(d["force"] - s.force) <= 1e-5 * rms(s.force)
        assert np.array_equal(d["virial"] != 0, s.temp != 0)


@pytest.mark.parametrize("shape", ["liquid", "gas"])
def test_elongated_boxes(shape):
    """Boxes much longer along one axis: grids whose short axes are whole
    bricks while the long one is tiled.  FP64 list forces equal the
    oracle's; with a thin skin (frequent rebuilds) 200 steps equal the
    all-pairs path bitwise."""
    rng = np.random.default_rng(17)
    if shape == "liquid":
        a = 0.5256
        pos = O.lattice_fcc([3, 4, 24], a)
        pos = pos + rng.uniform(-0.02, 0.02, pos.shape)
        box = [3 * a, 4 * a, 24 * a]
    else:
        box = [6.0, 9.0, 40.0]
        pos = rng.uniform(0.0, 1.0, (6000, 3)) * np.array(box)
    n = len(pos)
    s = O.State([n], [AR_MASS], box, pos=pos, vel=O.mb_velocities(n, 150.0, AR_MASS, 3))
    O.apply_boundary_conditions(s)
    e = engine_for(s, "fp64_list", skin=0.05)
    e.update_force()
    f = e.download(("force",))["force"]
    ref = s.copy()
    O.update_force(ref)  # all pairs, one minimum image (the short axes are below 2 r_cut)
    assert np.count_nonzero(ref.force) > n
    assert np.array_equal(f, ref.force)
    a_ = engine_for(s, "fp64_allpairs")
    a_.update_force()
    for eng in (a_, e):
        eng.step(0.002, 200)
    da, db = a_.download(), e.download()
    assert e.stats()["rebuilds"] >= 5
    assert np.array_equal(da["x"], db["x"]) and np.array_equal(da["v"], db["v"])


def droplet_in_gas(seed=41):
    """A dense FCC droplet (8^3 cells, a = 0.5256 nm, 2048 atoms) in a
    dilute gas (a jittered SC lattice of 14 / 13 nm, 14 nm box): the grid's
    mean occupancy sizes the bricks for the gas, so bricks over the droplet
    hold more particles than LDS stages."""
    rng = np.random.default_rng(seed)
    box = 14.0
    a = 0.5256
    off = (box - 8 * a) / 2
    drop = O.lattice_fcc([8, 8, 8], a) + off
    g = O.lattice_u([13, 13, 13], box / 13) + rng.uniform(-0.25, 0.25, (13 ** 3, 3))
    lo, hi = off - 0.6, off + 8 * a + 0.6
    keep = ~np.all((g > lo) & (g < hi), axis=1)
    pos = np.concatenate([drop + rng.uniform(-0.01, 0.01, drop.shape), g[keep]])
    n = len(pos)
    s = O.State([n], [AR_MASS], [box] * 3, pos=pos, vel=O.mb_velocities(n, 100.0, AR_MASS, seed))
    O.apply_boundary_conditions(s)
    return s


def test_dense_droplet_splits_bricks():
    """Bricks over the droplet overflow the LDS staging and are built as
    sub-bricks, and sub-bricks that still overflow cell by cell (restated in
    tests/common.py and asserted here): FP64 list forces equal the oracle's,
    FP32 ones are within 1e-5, and a thin-skin run of 100 steps (rebuilds
    through the split regions) equals the all-pairs path bit for bit."""
    from tests.common import brick_dims, brick_regions, cell_counts, grid_dims
    s = droplet_in_gas()
    rlist = 2.5 * 0.3418 + 0.3 * 0.3418
    gd = grid_dims(s.box, rlist, int(np.ceil(2.0 * s.n)))
    edge, whole = brick_dims(gd, s.n)
    over = brick_regions(cell_counts(s.pos, s.box, gd), edge, whole)
    print("grid", gd, "brick", edge, "overflowing regions", over)
    assert over[0] > 0 and over[1] > 0 and over[2] == 0
    ref = s.copy()
    O.update_force_cells(ref)
    assert np.count_nonzero(ref.force.any(axis=1)) > 800
    for mode in ("fp64_list", "fp32_list"):
        e = engine_for(s, mode)
        e.update_force()
        f = e.download(("force",))["force"]
        e.close()
        if mode == "fp64_list":
            assert np.array_equal(f, ref.force)
        else:
            assert rms(f - ref.force) <= 1e-5 * rms(ref.force)
    thin = engine_for(s, "fp64_list", skin=0.02)
    allp = engine_for(s, "fp64_allpairs")
    for eng in (thin, allp):
        eng.update_force()
        eng.step(0.002, 100)
    da, db = allp.download(), thin.download()
    assert thin.stats()["rebuilds"] >= 3
    assert np.array_equal(da["x"], db["x"]) and np.array_equal(da["v"], db["v"])


def test_dense_bricks_fall_back_to_grid_scan():
    """Liquid argon with ~32 atoms per cell: brick halos exceed the LDS
    staging capacity and lists are built straight from the grid."""
    s = liquid_state(10)
    e = engine_for(s, "fp64_list")
    e.update_force()
    d = e.download()
    O.update_force_cells(s)
    assert np.array_equal(d["force"], s.force)


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_reupload_into_used_engine(mode):
    """Grid buffers keep the headers of earlier fills: a new state uploaded
    into a used engine must not see any of them."""
    a = gas_state(3000, 12.0, 31)
    b = gas_state(3000, 12.0, 32)
    e = engine_for(a, mode)
    e.update_force()
    e.step(0.002, 300)
    e.upload(b.pos, b.vel, b.box)
    e.update_force()
    d = e.download()
    O.update_force_cells(b)
    if mode == "fp64_list":
        assert np.array_equal(d["force"], b.force)
    else:
        assert rms(d["force"] - b.force) <= 1e-5 * rms(b.force)
    # and a trajectory from the new state equals a fresh engine's
    e.step(0.002, 200)
    f = engine_for(b, mode)
    f.update_force()
    f.step(0.002, 200)
    assert np.array_equal(e.download()["x"], f.download()["x"])


@pytest.mark.parametrize("mode", FP64_MODES)
def test_multitype_npt(mode):
    """Two types under the Berendsen thermostat and barostat: lambda is
    overwritten per type and the box is scaled once per type
    (integrator.rs:16-27,54-58, barostat.rs:45-48), so positions and box
    scale inconsistently and the list paths rebuild before every step."""
    import moldyn_amd as M
    rng = np.random.default_rng(11)
    n0, n1 = 90, 70
    box = 3.2
    pos = rng.uniform(0, box, size=(n0 + n1, 3))
    vel = rng.normal(0, 0.3, size=(n0 + n1, 3))
    masses = [AR_MASS, 39.948]
    s = O.State([n0, n1], masses, [box] * 3, pos=pos, vel=vel)
    e = M.Engine([n0, n1], masses, [box] * 3, force_mode=mode)
    e.upload(pos, vel, [box] * 3)
    e.update_force()
    th = M.thermostat("berendsen", tau=2.0, target_t=200.0)
    ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=20.0)
    e.step(0.002, 60, th, ba)
    d = e.download()
    oth = O.Thermostat.berendsen(2.0)
    oba = O.Barostat.berendsen(1.0, 5.0)
    run_oracle(s, 60, 0.002, oth, 200.0, oba, 20.0)
    assert np.allclose(d["box"], s.box, rtol=1e-10, atol=0)
    dx = d["x"] - s.pos
    dx -= s.box * np.round(dx / s.box)
    assert np.abs(dx).max() < 1e-9
    assert np.allclose(d["v"], s.vel, rtol=1e-9, atol=1e-10)
    assert abs(th.lambda_ - oth.lambda_) < 1e-12 and abs(ba.myu - oba.myu) < 1e-12
    if mode == "fp64_list":
        assert e.stats()["rebuilds"] >= 60


@pytest.mark.parametrize("mode", ["fp64_list", "fp32_list"])
def test_tiles_per_block(mode):
    """Blocks that own several particle tiles (software-pipelined loads):
    forces and NVE trajectories are bitwise those of one tile per block;
    under a thermostat only the reduction grouping differs."""
    import moldyn_amd as M
    s = liquid_state(6, temperature=150.0)  # 864 atoms, 14 tiles of 64
    runs = []
    for tpb in (1, 3, 16):
        e = engine_for(s, mode, block_size=64, tiles_per_block=tpb)
        e.update_force()
        f0 = e.download(("force",))["force"]
        e.step(0.002, 200)
        d = e.download()
        th = M.thermostat("berendsen", tau=1.0, target_t=100.0)
        e.step(0.002, 100, th)
        runs.append((f0, d, e.download(), th.lambda_))
        e.close()
    f0, d0, t0, l0 = runs[0]
    for f, d, t, lam in runs[1:]:
        assert np.array_equal(f, f0)
        assert np.array_equal(d["x"], d0["x"]) and np.array_equal(d["v"], d0["v"])
        assert np.allclose(t["v"], t0["v"], rtol=1e-9, atol=1e-12)
        assert abs(lam - l0) < 1e-12


def test_device_kernel_timing():
    """Per-launch device timestamps cover every launch of a graph-replayed
    run and leave its results unchanged."""
    import moldyn_amd as M
    s = sc_state(12, seed=4)
    outs = []
    for prof in (False, True):
        e = engine_for(s, "fp32_list")
        e.update_force()
        e.set_profiling(prof)
        th = M.thermostat("berendsen", tau=10.0, target_t=300.0)
        e.step(0.002, 700, th)
        outs.append(e.download(("x", "v")))
        if prof:
            ms, n = e.kernel_time("step", span=True)
            fms, fn = e.kernel_time("finalize", span=True)
            st = e.stats()
            assert n >= 700 - st["rebuilds"] - 1 and n <= st["launches"]
            assert fn == st["launches"]
            assert 0.0 < ms / n < 5.0 and 0.0 < fms / fn < 5.0
            # stream time: back-to-back launches only, never shorter than the span
            sms, sn = e.kernel_time("step")
            assert 0.8 * n <= sn <= n and sms / sn >= 0.99 * ms / n
        e.close()
    assert np.array_equal(outs[0]["x"], outs[1]["x"]) and np.array_equal(outs[0]["v"], outs[1]["v"])
