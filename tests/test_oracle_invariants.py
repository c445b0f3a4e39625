"""Pin the oracle where the reference's golden digits say nothing.

Closed forms evaluated independently (numpy long double), invariants, and
brute force on small inputs, chosen so that a dropped term, a wrong sign or
index, or a transposed operand in oracle/md_oracle.c fails at least one.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

K_B = O.K_B
AR_M = 66.335


def ld_lj(sigma, eps, r, ucut):
    s = np.longdouble(sigma) / np.longdouble(r)
    s6 = s ** 6
    u = 4 * np.longdouble(eps) * (s6 * s6 - s6) - np.longdouble(ucut)
    f = (24 * np.longdouble(eps) / np.longdouble(r)) * (s6 - 2 * s6 * s6)
    return u, f


def test_lj_closed_form_and_cut():
    sigma, eps, rc, ucut = O.new_lennard_jones(0.3418, 1.712)
    assert rc == 0.3418 * 2.5
    ucut_cf = 4 * 1.712 * (2.5 ** -12 - 2.5 ** -6)
    assert abs(ucut - ucut_cf) < 1e-15
    for r in np.linspace(0.30, rc, 57):
        u, f = O.lj((sigma, eps, rc, ucut), r)
        ul, fl = ld_lj(sigma, eps, r, ucut)
        assert abs(u - float(ul)) <= 1e-13 * max(1.0, abs(float(ul)))
        assert abs(f - float(fl)) <= 1e-13 * max(1.0, abs(float(fl)))
        # f is dU/dr: compare with a central difference of U
        h = 1e-6
        up, _ = O.lj((sigma, eps, rc + 1, ucut), r + h)
        um, _ = O.lj((sigma, eps, rc + 1, ucut), r - h)
        assert abs((up - um) / (2 * h) - f) <= 1e-5 * max(1.0, abs(f))
    # beyond the cut: zero; at the cut: included, U(r_cut) - u_cut = 0
    assert O.lj((sigma, eps, rc, ucut), rc * (1 + 1e-15)) == (0.0, 0.0)
    u, f = O.lj((sigma, eps, rc, ucut), rc)
    assert u == 0.0 and f > 0.0


def random_state(n, box, seed, n_types=1, masses=None):
    rng = np.random.default_rng(seed)
    counts = [n // n_types] * n_types
    counts[-1] += n - sum(counts)
    masses = masses or [AR_M] * n_types
    pos = rng.uniform(0, box, size=(n, 3))
    vel = rng.normal(0, 0.3, size=(n, 3))
    return O.State(counts, masses, [box] * 3, pos=pos, vel=vel)


def minimum_image_counts(pos, box, rc):
    d = pos[None, :, :] - pos[:, None, :]
    d -= box * np.round(d / box)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    return (r <= rc).sum(1)


def test_pair_count_brute_force():
    for seed, (n, box) in enumerate([(300, 3.0), (200, 1.9), (64, 5.0)]):
        s = random_state(n, box, seed)
        rc = 0.3418 * 2.5
        got = O.pair_counts(s)
        assert np.array_equal(got, minimum_image_counts(s.pos, box, rc))


def test_newton_third_law_bitwise_two_body():
    s = O.State([2], [AR_M], [2.0, 2.0, 2.0], pos=[[0.3, 0.7, 1.1], [0.61, 0.52, 1.33]])
    O.update_force(s)
    assert np.array_equal(s.force[0], -s.force[1])
    assert s.pot[0] == s.pot[1] and s.temp[0] == s.temp[1]


def test_total_force_vanishes_dense():
    pos = O.lattice_fcc([4, 4, 4], 0.5256)
    rng = np.random.default_rng(5)
    pos = pos + rng.uniform(-0.02, 0.02, pos.shape)
    s = O.State([len(pos)], [AR_M], [4 * 0.5256] * 3, pos=pos)
    O.apply_boundary_conditions(s)
    O.update_force(s)
    fmax = np.abs(s.force).max()
    assert np.abs(s.force.sum(0)).max() < 1e-11 * fmax * len(pos)
    # virial sum and pair counts are symmetric in i, j
    counts = O.pair_counts(s)
    assert counts.sum() % 2 == 0 and counts.mean() > 40


@pytest.mark.parametrize("case", ["liquid", "gas", "small_box", "two_types"])
def test_cell_variant_bitwise(case):
    if case == "liquid":
        pos = O.lattice_fcc([5, 5, 5], 0.5256)
        s = O.State([len(pos)], [AR_M], [5 * 0.5256] * 3, pos=pos)
        s.pos += np.random.default_rng(1).uniform(-0.03, 0.03, s.pos.shape)
        O.apply_boundary_conditions(s)
    elif case == "gas":
        s = random_state(2000, 12.0, 2)
    elif case == "small_box":
        s = random_state(90, 1.7, 3)
    else:
        s = random_state(400, 4.0, 4, n_types=2, masses=[AR_M, 39.948])
    tab = O.potential_table(s.n_types, {(0, 1): O.new_lennard_jones(0.33, 1.2)})
    a = s.copy()
    b = s.copy()
    O.update_force(a, tab)
    O.update_force_cells(b, tab)
    assert np.array_equal(a.force, b.force)
    assert np.array_equal(a.pot, b.pot)
    assert np.array_equal(a.temp, b.temp)


@pytest.mark.parametrize("case", ["liquid_nve", "gas_npt", "dilute_npt", "two_types_npt"])
def test_cell_step_bitwise(case):
    """orc_verlet_step_cells (forces through the cell grid) follows
    orc_verlet_step bit for bit; the dilute gas has more cells than the
    grid's budget of about two per particle, so the grid is coarsened."""
    th = ba = None
    t0 = p0 = 0.0
    if case == "liquid_nve":
        pos = O.lattice_fcc([4, 4, 4], 0.5256)
        s = O.State([len(pos)], [AR_M], [4 * 0.5256] * 3, pos=pos,
                    vel=O.mb_velocities(len(pos), 150.0, AR_M, 3))
        s.pos += np.random.default_rng(1).uniform(-0.03, 0.03, s.pos.shape)
        O.apply_boundary_conditions(s)
    elif case in ("gas_npt", "dilute_npt"):
        # jittered simple-cubic lattices (no overlapping pairs)
        side, a, jit = (11, 0.5, 0.2) if case == "gas_npt" else (10, 1.25, 0.25)
        pos = O.lattice_u([side] * 3, a)
        pos += np.random.default_rng(5).uniform(-jit, jit, pos.shape) * a
        s = O.State([len(pos)], [AR_M], [side * a] * 3, pos=pos,
                    vel=O.mb_velocities(len(pos), 300.0, AR_M, 4))
        O.apply_boundary_conditions(s)
    else:
        pos = O.lattice_u([8, 8, 8], 0.5)
        pos += np.random.default_rng(7).uniform(-0.1, 0.1, pos.shape)
        s = O.State([300, 212], [AR_M, 39.948], [4.0] * 3, pos=pos,
                    vel=np.random.default_rng(8).normal(0, 0.3, pos.shape))
        O.apply_boundary_conditions(s)
    if case.endswith("npt"):
        th, t0, ba, p0 = "b", 200.0, "b", 5.0
    tab = O.potential_table(s.n_types, {(0, 1): O.new_lennard_jones(0.33, 1.2)})
    a, b = s.copy(), s.copy()
    for st in (a, b):
        O.update_force(st, tab)
    tha = thb = bea = beb = None
    if th:
        tha, thb = O.Thermostat.berendsen(1.0), O.Thermostat.berendsen(1.0)
        bea, beb = O.Barostat.berendsen(1.0, 5.0), O.Barostat.berendsen(1.0, 5.0)
    for _ in range(40):
        O.verlet_step(a, 0.002, tab, bea, p0, tha, t0)
        O.verlet_step(b, 0.002, tab, beb, p0, thb, t0, cells=True)
    for k in ("pos", "vel", "force", "pot", "temp", "box"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert np.count_nonzero(a.force) > 0


def np_temperature(vel, mass):
    m = np.longdouble(mass)
    v = vel.astype(np.longdouble)
    com = (v * m).sum(0) / (m * len(v))
    e = (m * ((v - com) ** 2).sum(1) / 2).sum()
    return float(2 * e / (3 * len(v) * np.longdouble(K_B)) * 100), com, e


def test_ideal_gas_pressure_and_temperature():
    # config-1 shape: no pair inside r_cut, so P = N k T / V exactly in theory
    pos = O.lattice_u([6, 6, 6], 3.338339)
    vel = O.mb_velocities(216, 273.15, AR_M, seed=1)
    s = O.State([216], [AR_M], [6 * 3.338339] * 3, pos=pos, vel=vel)
    O.update_force(s)
    assert not s.force.any() and not s.temp.any()
    mv = O.center_of_mass_velocity(s, 0)
    e = O.thermal_energy(s, 0, mv)
    t = O.temperature(e, 216)
    p = O.pressure(s, 0, mv)
    tl, _, el = np_temperature(s.vel, AR_M)
    assert abs(t - tl) < 1e-12 * tl
    vol = (6 * 3.338339) ** 3
    assert abs(p - 216 * K_B * (t / 100.0) / vol) < 1e-12 * p
    assert abs(p - 2 * float(el) / (3 * vol)) < 1e-12 * p


def test_berendsen_thermostat_closed_form():
    s = random_state(50, 10.0, 7)
    t_now, _, _ = np_temperature(s.vel, AR_M)
    th = O.Thermostat.berendsen(tau=10.0)
    O.thermostat_calculate_lambda(th, s, 0.002, 0, 300.0)
    lam = math.sqrt(1.0 + 0.002 / 10.0 * (300.0 / t_now - 1.0))
    assert abs(th.lambda_ - lam) < 1e-14
    th2 = O.Thermostat.berendsen(tau=10.0)
    O.thermostat_calculate_lambda(th2, s, 0.002, 0, t_now)
    assert abs(th2.lambda_ - 1.0) < 1e-14
    v0 = s.vel.copy()
    O.thermostat_update(th, s, 0.002, 0, 300.0)
    assert np.array_equal(s.vel, v0 * th.lambda_)


def test_nose_hoover_closed_form():
    s = random_state(50, 10.0, 8)
    t_now, _, _ = np_temperature(s.vel, AR_M)
    th = O.Thermostat.nose_hoover(tau=0.5, psi=0.1)
    O.thermostat_calculate_lambda(th, s, 0.002, 0, 200.0)
    psi = 0.1 + (-((200.0 / t_now) - 1.0) / 0.5) * 0.001
    assert abs(th.psi - psi) < 1e-14
    assert abs(th.lambda_ - math.exp(-psi * 0.002 / 2)) < 1e-15
    # update scales by lambda and advances psi with the pre-scaling temperature
    O.thermostat_update(th, s, 0.002, 0, 200.0)
    assert abs(th.psi - (psi + (-((200.0 / t_now) - 1.0) / 0.5) * 0.001)) < 1e-14


def test_berendsen_barostat_closed_form():
    s = random_state(60, 3.0, 9)
    O.update_force(s)
    m = np.longdouble(AR_M)
    v = s.vel.astype(np.longdouble)
    com = (v * m).sum(0) / (m * len(v))
    r1 = (m * ((v - com) ** 2)).sum()
    r2 = -s.temp.astype(np.longdouble).sum()
    p_cf = float((r1 + r2 / 2) / np.longdouble(27.0) / 3)
    mv = O.center_of_mass_velocity(s, 0)
    p = O.pressure(s, 0, mv)
    assert abs(p - p_cf) < 1e-12 * abs(p_cf)
    ba = O.Barostat.berendsen(beta=1.0, tau=5.0)
    O.barostat_calculate_myu(ba, s, 0.002, 0, 1.01325)
    mu = (1.0 + 0.002 * 1.0 / 5.0 * (p - 1.01325)) ** (1.0 / 3.0)
    assert abs(ba.myu - mu) < 1e-15
    ba2 = O.Barostat.berendsen(beta=1.0, tau=5.0)
    O.barostat_calculate_myu(ba2, s, 0.002, 0, p)
    assert ba2.myu == 1.0
    x0 = s.pos.copy()
    O.barostat_update(ba, s, 0)
    assert np.array_equal(s.pos, x0 * ba.myu)
    assert np.array_equal(s.box, np.full(3, 3.0 * ba.myu))


def test_wrap_range():
    rng = np.random.default_rng(3)
    box = 5.0
    pos = rng.uniform(-box + 1e-9, 2 * box - 1e-9, size=(1000, 3))
    s = O.State([1000], [1.0], [box] * 3, pos=pos)
    O.apply_boundary_conditions(s)
    assert (s.pos >= 0).all() and (s.pos < box).all()
    # one conditional correction, not a modulo (core/src/particle.rs:124-138)
    expect = np.where(pos < 0, pos + box, np.where(pos >= box, pos - box, pos))
    assert np.array_equal(s.pos, expect)


def test_mb_velocities_invariants():
    v = O.mb_velocities(20001, 273.15, AR_M, seed=3)
    assert np.array_equal(v[:10000], -v[10000:20000])
    assert not v[-1].any()
    sig = math.sqrt(K_B * 2.7315 / AR_M)
    # sample std of 30000 normal draws: relative error ~ 1/sqrt(2*30000)
    assert abs(v[:10000].std() / sig - 1) < 0.02
    assert np.array_equal(v, O.mb_velocities(20001, 273.15, AR_M, seed=3))
    assert not np.array_equal(v, O.mb_velocities(20001, 273.15, AR_M, seed=4))


def test_fcc_lattice():
    pos = O.lattice_fcc([3, 3, 3], 0.5)
    assert len(pos) == 108
    d = pos[None] - pos[:, None]
    d -= 1.5 * np.round(d / 1.5)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    assert abs(r.min() - 0.5 / math.sqrt(2)) < 1e-12
    assert ((np.abs(r - 0.5 / math.sqrt(2)) < 1e-9).sum(1) == 12).all()


def py_update_force(pos, counts, box, table):
    """Independent pure-Python reading of potential.rs:158-216 (tiny inputs)."""
    n = len(pos)
    types = np.repeat(np.arange(len(counts)), counts)
    force = np.zeros((n, 3)); pot = np.zeros(n); temp = np.zeros(n)
    for i in range(n):
        for j in range(n):
            if types[j] < types[i] or i == j:
                continue
            p = table[(types[i], types[j])]
            r = pos[j] - pos[i]
            r = np.where(r < -box / 2, r + box, np.where(r > box / 2, r - box, r))
            ra = math.sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2])
            if ra > p[2]:
                continue
            sr = p[0] / ra
            s6 = (sr * sr) * ((sr * sr) * (sr * sr))
            u = 4 * p[1] * (s6 * s6 - s6) - p[3]
            du = (24 * p[1] / ra) * (s6 - 2 * s6 * s6)
            fv = (r / ra) * du
            force[i] += fv; pot[i] += u; temp[i] += (fv[0] * r[0] + fv[1] * r[1]) + fv[2] * r[2]
    return force, pot, temp


def test_multitype_asymmetry_matches_independent_reading():
    s = random_state(40, 1.6, 12, n_types=3, masses=[1.0, 2.0, 3.0])
    pots = {(0, 0): O.new_lennard_jones(0.3418, 1.712), (0, 1): O.new_lennard_jones(0.30, 1.0),
            (1, 1): O.new_lennard_jones(0.32, 1.5), (0, 2): O.new_lennard_jones(0.28, 0.9),
            (1, 2): O.new_lennard_jones(0.31, 1.1), (2, 2): O.new_lennard_jones(0.35, 2.0)}
    tab = O.potential_table(3, pots)
    O.update_force(s, tab)
    full = {}
    for (a, b), p in pots.items():
        full[(a, b)] = p
        full[(b, a)] = p
    f, pot, temp = py_update_force(s.pos, s.counts, s.box, full)
    assert np.allclose(s.force, f, rtol=1e-13, atol=1e-13)
    assert np.allclose(s.pot, pot, rtol=1e-13, atol=1e-13)
    assert np.allclose(s.temp, temp, rtol=1e-13, atol=1e-13)
