"""Pin the CPU oracle against the reference's own known answers.

Each test restates one of the reference's unit tests (AndrewChe7/moldyn
solver/src/lib.rs, cli/src/tests.rs) against oracle/md_oracle.c, using the
digits in tests/golden/reference_golden.json.
"""
import numpy as np
import pytest

from oracle import oracle as O

AR = (0.3418, 1.712)


def f8(x):
    return "%.8f" % x


def two_body(g, key):
    c = g[key]
    return O.State([2], [c["mass"]], c["box"], pos=c["pos"], vel=c["vel"])


def test_lennard_jones(golden):
    c = golden["lennard_jones"]
    p = O.new_lennard_jones(c["sigma"], c["eps"])
    u, f = O.lj(p, c["r"])
    assert f8(u) == c["potential"]
    assert f8(f) == c["force"]


def test_update_force_lennard_jones(golden):
    c = golden["update_force_lennard_jones"]
    s = O.State([2], [c["mass"]], c["box"], pos=c["pos"])
    O.update_force(s)
    assert [f8(x) for x in s.force[0]] == c["force_p1"]


def test_verlet_with_lennard_jones(golden):
    c = golden["verlet_with_lennard_jones"]
    s = two_body(golden, "verlet_with_lennard_jones")
    O.update_force(s)
    for k, ref in enumerate(c["steps"]):
        if k > 0:
            O.verlet_step(s, c["dt"])
        assert [[f8(x) for x in row] for row in s.pos] == ref["pos"], k
        assert [[f8(x) for x in row] for row in s.vel] == ref["vel"], k
        assert [[f8(x) for x in row] for row in s.force] == ref["force"], k


def test_verlet_lj_1000_iterations(golden):
    c = golden["verlet_lj_1000_iterations"]
    s = two_body(golden, "verlet_lj_1000_iterations")
    tab = O.potential_table(1)
    O.update_force(s, tab)
    for _ in range(c["n_steps"]):
        O.verlet_step(s, c["dt"], tab)
    assert [[f8(x) for x in row] for row in s.pos] == c["final_pos"]
    assert [[f8(x) for x in row] for row in s.vel] == c["final_vel"]
    assert [[f8(x) for x in row] for row in s.force] == c["final_force"]
    mv = O.center_of_mass_velocity(s, 0)
    kinetic = O.kinetic_energy(s, 0)
    thermal = O.thermal_energy(s, 0, mv)
    potential = O.potential_energy(s, 0)
    assert f8(kinetic) == c["kinetic"]
    assert f8(thermal) == c["thermal"]
    assert f8(potential) == c["potential"]
    assert f8(thermal + potential) == c["internal"]
    assert f8(kinetic + potential) == c["full"]
    assert f8(O.temperature(thermal, 2) / 100.0) == c["temperature_over_100"]
    assert f8(O.pressure(s, 0, mv)) == c["pressure"]


def test_energies_temperature_pressure(golden):
    c = golden["energies"]
    s = two_body(golden, "energies")
    O.update_force(s)
    mv = O.center_of_mass_velocity(s, 0)
    assert list(mv) == c["com_velocity"]
    ek = O.kinetic_energy(s, 0)
    et = O.thermal_energy(s, 0, mv)
    ep = O.potential_energy(s, 0)
    assert f8(ek) == c["kinetic"]
    assert f8(et) == c["thermal"]
    assert f8(ep) == c["potential"]
    assert f8(et + ep) == c["internal"]
    assert f8(ek + ep) == c["full"]
    assert f8(O.temperature(et, 2)) == golden["temperature"]["temperature"]
    assert f8(O.pressure(s, 0, mv)) == golden["pressure"]["pressure"]


def test_initialize_uniform_grid(golden):
    c = golden["initialize_uniform_grid"]
    pos = O.lattice_u(c["grid"], c["a"])
    assert list(pos[c["index"]]) == c["pos"]


def test_initialization_box(golden):
    c = golden["initialization"]
    pos = O.lattice_u(c["grid"], c["a"])
    assert len(pos) == c["n"]
    # cli/src/commands.rs:64-67: box = lattice_cell * size
    assert c["a"] * float(c["grid"][0]) == c["box"]
    assert pos.max() < c["box"]


def test_momentum_invariant(golden):
    """solver/src/lib.rs:49-75, with this build's seeded MB draws, 20,000 steps."""
    c = golden["momentum"]
    pos = O.lattice_u(c["grid"], c["a"])
    vel = O.mb_velocities(8, c["temperature"], c["mass"], seed=11)
    s = O.State([8], [c["mass"]], np.array([2.0, 2.0, 2.0]) * c["a"], pos=pos, vel=vel)
    tab = O.potential_table(1)
    O.update_force(s, tab)
    assert np.all(np.abs(s.vel.sum(axis=0)) < c["bound"])
    for _ in range(20000):
        O.verlet_step(s, 0.002, tab, nthreads=1)
    assert np.all(np.abs(s.vel.sum(axis=0)) < c["bound"])


def test_golden_digits_have_pinning_power(golden):
    """Plausible slips in the LJ restatement change the asserted digits."""
    c = golden["lennard_jones"]
    sigma, eps, rc, ucut = O.new_lennard_jones(c["sigma"], c["eps"])
    r = c["r"]
    s6 = (sigma / r) ** 6
    good = 4 * eps * (s6 * s6 - s6) - ucut
    assert f8(good) == c["potential"]
    assert f8(4 * eps * (s6 * s6 - s6)) != c["potential"]        # u_cut dropped
    assert f8(4 * eps * (s6 * s6 - s6) + ucut) != c["potential"]  # wrong sign
    assert f8((24 * eps / r) * (s6 - s6 * s6)) != c["force"]     # factor 2 dropped
    assert f8((24 * eps / r) * (s6 - 2 * s6 * s6)) == c["force"]
