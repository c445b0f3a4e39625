"""State files through the GPU (md_frame_writer_*, md_text_*) against the
oracle's text (oracle/frame_text.py) byte for byte, the reference's CLI tests
through the files (cli/src/tests.rs), solve_macro against the FP64 oracle,
and a 1M-atom round trip."""
import math
import os

import numpy as np
import pytest

from oracle import frame_text as T
from oracle import oracle as O
from tests.common import AR_MASS, liquid_state, sc_state

pytestmark = pytest.mark.gpu


def read(path):
    with open(path, newline="") as f:
        return f.read()


def rand_doubles(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 2 ** 64, n, dtype=np.uint64).view(np.float64)
    return x[np.isfinite(x)]


def edge_doubles():
    out = [0.0, -0.0, 1.0, -1.0, 0.1, 0.3, 1e15, 1e16, 1e-5, 1e-6, 5e-324, 1.7976931348623157e308,
           2.2250738585072014e-308, 2.0 ** 53, 9007199254740993.0, 1.23e40, 1.23e-40, 33.38339, 0.00012]
    for e in range(-1074, 1024, 9):
        p = math.ldexp(1.0, e)
        out += [p, math.nextafter(p, 0.0), math.nextafter(p, math.inf)]
    return np.array(out)


def test_gpu_text_matches_the_exact_oracle_on_edge_values():
    from moldyn_amd.frames import text_formatter
    vals = edge_doubles()
    got = text_formatter().f64(vals)
    assert got == [T.f64_text(v, exact=(v != 0.0)) for v in vals]
    assert text_formatter().f64([np.nan, np.inf, -np.inf]) == ["NaN", "inf", "-inf"]


def test_gpu_text_matches_the_oracle_on_random_doubles():
    from moldyn_amd.frames import text_formatter
    rng = np.random.default_rng(3)
    vals = np.concatenate([rand_doubles(300000, 1), rng.uniform(0, 400, 100000), rng.normal(0, 3, 100000),
                           np.round(rng.uniform(0, 100, 50000), 3)])
    got = text_formatter().f64(vals)
    want = [T.f64_text(v) for v in vals]
    bad = [i for i in range(len(vals)) if got[i] != want[i]]
    assert not bad, [(repr(vals[i]), got[i], want[i]) for i in bad[:5]]


def test_gpu_table_with_keys():
    from moldyn_amd.frames import MACRO_HEADER, text_formatter
    rng = np.random.default_rng(4)
    vals = rng.normal(0, 1e3, (1000, 9))
    keys = np.arange(1000, dtype=np.uint64) * 977 + 18446744073709551000 // 1000
    text = text_formatter().table(vals, keys=keys)
    assert MACRO_HEADER + "\n" + text == T.table_csv(MACRO_HEADER, [[int(k)] + list(r) for k, r in zip(keys, vals)])


def two_type_engine(force_mode="fp64_list"):
    import moldyn_amd as M
    s = liquid_state(n_cells=4)
    n = s.n
    counts = [n - 100, 100]
    e = M.Engine(counts, [AR_MASS, 39.948], s.box, force_mode=force_mode)
    x = s.pos.copy()
    x[0] = [0.0, 1e-7, s.box[2] / 3]
    v = s.vel.copy()
    v[0] = [-0.0, 0.0, 1e-6]
    e.upload(x, v, s.box)
    e.update_force()
    return e, counts


def oracle_frame(e, counts):
    d = e.download(("x", "v", "box"))
    ids = np.repeat(np.arange(len(counts)), counts)
    return T.frame_csv(ids, d["x"], d["v"]), [float(b) for b in np.asarray(d["box"]).ravel()[:3]]


def test_frame_writer_files_are_the_oracle_text(tmp_path):
    """Frames saved while stepping hold the state of their save call; bb.csv
    rows are replaced or appended as StateToSave::save_bb does."""
    import moldyn_amd as M
    from moldyn_amd.frames import FrameWriter
    e, counts = two_type_engine()
    th = M.thermostat("berendsen", tau=0.5, target_t=120.0)
    ba = M.barostat("berendsen", beta=1.0, tau=5.0, target_p=50.0)
    want, rows = {}, []
    with FrameWriter(e, tmp_path) as w:
        for k in range(4):
            w.save(k)
            want[k], box = oracle_frame(e, counts)
            rows = T.bb_save(rows, k, box)
            e.step(0.002, 3, th, ba)
        w.save(1)  # replace row 1
        want[1], box = oracle_frame(e, counts)
        rows = T.bb_save(rows, 1, box)
        w.save(9)  # past the end: appended
        want[9], box = oracle_frame(e, counts)
        rows = T.bb_save(rows, 9, box)
        w.flush()
        st = w.stats()
    for k, text in want.items():
        assert read(tmp_path / "data" / f"{k}.csv") == text, k
    assert read(tmp_path / "bb.csv") == T.bb_csv(rows)
    assert st["frames"] == 6 and st["bytes"] > 0
    e.close()


def test_frame_writer_rewrites_a_foreign_box_file(tmp_path):
    import moldyn_amd as M
    from moldyn_amd.frames import FrameWriter
    with open(tmp_path / "bb.csv", "w") as f:
        f.write("x,y,z\n1e1,+2,3.\n")
    e = M.Engine([8], [AR_MASS], [4.0, 4.0, 4.0], force_mode="fp64_list")
    e.upload(O.lattice_u([2, 2, 2], 2.0), np.zeros((8, 3)), [4.0, 4.0, 4.0])
    with FrameWriter(e, tmp_path) as w:
        w.save(1)
    assert read(tmp_path / "bb.csv") == "x,y,z\n10.0,2.0,3.0\n4.0,4.0,4.0\n"
    e.close()


def test_state_files_round_trip_at_1m_atoms(tmp_path):
    """1,000,000-atom SC state: load(save(state)) returns every double bit for
    bit; sampled rows are the oracle's text."""
    import moldyn_amd as M
    from moldyn_amd.frames import FrameWriter, StateToSave
    s = sc_state(100)
    e = M.Engine([s.n], [AR_MASS], s.box, force_mode="fp32_list")
    e.upload(s.pos, s.vel, s.box)
    e.update_force()
    e.step(0.002, 10, M.thermostat("berendsen", tau=10.0, target_t=300.0))
    d = e.download(("x", "v", "box"))
    with FrameWriter(e, tmp_path) as w:
        w.save(0)
        w.flush()
        st = w.stats()
    back = StateToSave.load_from_file(tmp_path, 0)
    assert back.x.tobytes() == d["x"].tobytes() and back.v.tobytes() == d["v"].tobytes()
    assert back.boundary_box.tobytes() == np.asarray(d["box"], dtype=np.float64).ravel()[:3].tobytes()
    assert not back.ids.any()
    lines = read(tmp_path / "data" / "0.csv").split("\n")
    assert lines[0] == T.FRAME_HEADER and lines[-1] == "" and len(lines) == s.n + 2
    for i in np.random.default_rng(0).integers(0, s.n, 2000):
        assert lines[i + 1] == "0," + ",".join(T.f64_text(a) for a in list(d["x"][i]) + list(d["v"][i]))
    print("1M frame:", st)
    e.close()


def test_solvation_through_the_files(tmp_path, golden):
    """cli/src/tests.rs:35-98: two argon atoms saved as state 0, three solve
    steps through the frame directory, state 3 loaded, update_force; the
    reference's 8-decimal digits."""
    import moldyn_amd as M
    from moldyn_amd import commands
    from moldyn_amd.frames import StateToSave, load_boxes, save_particles_data
    g = golden["verlet_with_lennard_jones"]
    StateToSave(np.zeros(2, dtype=np.uint16), np.array(g["pos"]), np.array(g["vel"]),
                np.array(g["box"])).save_to_file(tmp_path, 0)
    save_particles_data(tmp_path, {0: ("Argon", 66.335, 0.071)})
    commands.solve(tmp_path, 0, 3, 0.002)
    assert sorted(os.listdir(tmp_path / "data")) == ["0.csv", "1.csv", "2.csv", "3.csv"]
    assert load_boxes(str(tmp_path)).tolist() == [[2.0, 2.0, 2.0]] * 4
    s = StateToSave.load_from_file(tmp_path, 3)
    e = M.Engine([2], [66.335], s.boundary_box, force_mode="fp64_list")
    e.upload(s.x, s.v, s.boundary_box)
    e.update_force()
    d = e.download(("x", "v", "force"))
    e.close()
    step3 = g["steps"][3]
    for key, arr in (("pos", d["x"]), ("vel", d["v"]), ("force", d["force"])):
        assert [[f"{a:.8f}" for a in row] for row in arr] == step3[key], key
    assert read(tmp_path / "data" / "0.csv") == (
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n"
        "0,0.75,0.75,0.5,1.0,1.0,0.0\n0,1.25,0.75,0.5,-1.0,1.0,0.0\n")


def test_initialization_through_the_files(tmp_path, golden):
    """cli/src/tests.rs:9-33: initialize a 10x10x10 U lattice, read state 0
    and the particle database back."""
    from moldyn_amd import commands
    from moldyn_amd.frames import StateToSave, load_particles_data
    g = golden["initialization"]
    commands.initialize(tmp_path, "U", g["grid"], "Argon", 66.335, 0.071, g["a"], 273.15, seed=3)
    s = StateToSave.load_from_file(tmp_path, 0)
    db = load_particles_data(tmp_path)
    assert db == {0: ("Argon", 66.335, 0.071)}
    counts, masses, x, v = s.into_state({0: 66.335})
    assert list(counts) == [g["n"]] and list(s.boundary_box) == [g["box"]] * 3
    assert read(tmp_path / "db.csv") == "id,name,mass,radius\n0,Argon,66.335,0.071\n"
    # the file holds the library's seeded draws (host libm; the oracle's
    # build of the same draws agrees to the last bits of log, sin and cos)
    import moldyn_amd as M
    v_lib = M.mb_velocities(g["n"], 273.15, 66.335, 3)
    assert np.abs(v_lib - O.mb_velocities(g["n"], 273.15, 66.335, 3)).max() <= 1e-15
    assert read(tmp_path / "data" / "0.csv") == T.frame_csv(np.zeros(g["n"]), O.lattice_u(g["grid"], g["a"]), v_lib)


def test_solve_macro_matches_the_oracle(tmp_path):
    """commands.rs:204-274 over six stored states of a two-type liquid: every
    macro.csv row holds the FP64 oracle's macro parameters of that state
    (the sums are reduced on the device in another order than the
    reference's sequential fold: relative 1e-12), in the file's layout."""
    from moldyn_amd import commands
    from moldyn_amd.frames import FrameWriter, save_particles_data
    import moldyn_amd as M
    e, counts = two_type_engine()
    save_particles_data(tmp_path, {0: ("Argon", AR_MASS, 0.071), 1: ("Ar40", 39.948, 0.071)})
    th = M.thermostat("berendsen", tau=0.5, target_t=150.0)
    with FrameWriter(e, tmp_path) as w:
        for k in range(6):
            w.save(k)
            e.step(0.002, 5, th)
    e.close()
    commands.solve_macro(tmp_path, kinetic_energy=True, potential_energy=True, thermal_energy=True,
                         temperature=True, pressure=True, batch=4)
    rows = []
    n_total = int(sum(counts))
    for k in range(6):
        ids, x, v = T.read_frame(read(tmp_path / "data" / f"{k}.csv"))
        box = T.read_csv(read(tmp_path / "bb.csv"))[1][k]
        s = O.State(counts, [AR_MASS, 39.948], [float(b) for b in box], pos=x, vel=v)
        O.update_force(s)
        com = O.center_of_mass_velocity(s, 0)
        ke, pe, th_e = O.kinetic_energy(s, 0), O.potential_energy(s, 0), O.thermal_energy(s, 0, com)
        rows.append([k, ke, pe, th_e, ke / n_total, pe / n_total, th_e / n_total,
                     O.temperature(th_e, n_total), O.pressure(s, 0, com), 0.0])
    head, got = T.read_csv(read(tmp_path / "macro.csv"))
    assert ",".join(head) == T.MACRO_HEADER and [int(r[0]) for r in got] == list(range(6))
    for r, w in zip(got, rows):
        vals = [float(a) for a in r[1:]]
        assert all(T.f64_text(float(a)) == a for a in r[1:])  # the file's numbers are in ryu's form
        assert np.allclose(vals, w[1:], rtol=1e-12, atol=0), (vals, w)


def test_potentials_and_particle_files(tmp_path):
    """potentials.json in serde_json's pretty layout, read back and applied:
    forces of a two-type system with its own (1, 1) pair equal the oracle's."""
    import moldyn_amd as M
    from moldyn_amd import frames as F
    lj0 = M.new_lennard_jones(0.3418, 1.712)
    F.save_potentials_to_file(tmp_path, {(0, 0): ("LennardJones", 0.3418, 1.712) + tuple(lj0)})
    assert read(tmp_path / "potentials.json") == (
        '{\n  "0,0": {\n    "LennardJones": {\n      "sigma": 0.3418,\n      "eps": 1.712,\n'
        '      "r_cut": 0.8545,\n      "u_cut": -0.027934517624831987\n    }\n  }\n}')
    rc, uc = M.new_lennard_jones(0.3, 1.2)
    F.save_potentials_to_file(tmp_path, {(1, 1): ("LennardJones", 0.3, 1.2, rc, uc),
                                         (0, 1): ("LennardJones", 0.32, 1.5, 0.8, -0.01)})
    pots = F.load_potentials_from_file(tmp_path)
    s = liquid_state(n_cells=3)
    counts = [s.n - 40, 40]
    e = M.Engine(counts, [AR_MASS, AR_MASS], s.box, force_mode="fp64_list")
    F.apply_potentials(e, pots)
    e.upload(s.pos, s.vel, s.box)
    e.update_force()
    f = e.download(("force",))["force"]
    e.close()
    ref = O.State(counts, [AR_MASS, AR_MASS], s.box, pos=s.pos, vel=s.vel)
    table = O.potential_table(2, {(1, 1): (0.3, 1.2, rc, uc), (0, 1): (0.32, 1.5, 0.8, -0.01)})
    O.update_force(ref, table)
    assert f.tobytes() == ref.force.tobytes()
    F.save_particles_data(tmp_path, {3: ("Ne, light", 20.18, 0.0), 0: ("Argon", 66.335, 0.071)})
    assert read(tmp_path / "db.csv") == 'id,name,mass,radius\n0,Argon,66.335,0.071\n3,"Ne, light",20.18,0.0\n'
