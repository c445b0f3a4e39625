"""The state-file oracle (oracle/frame_text.py) against what pins it, and the
host side of the file boundary (md_frame_load, the Python mirrors) against
the oracle; no GPU needed."""
import math
import os

import numpy as np
import pytest

from oracle import frame_text as T

# digit strings of the ryu crate's own regression tests and published examples,
# with the layout the crate gives them
RYU_KNOWN = [
    (1.234, "1.234"), (1.23e40, "1.23e40"), (1.23e-40, "1.23e-40"),
    (0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (-1.0, "-1.0"),
    (float("nan"), "NaN"), (float("inf"), "inf"), (float("-inf"), "-inf"),
    (5e-324, "5e-324"), (1.7976931348623157e308, "1.7976931348623157e308"),
    (2.2250738585072014e-308, "2.2250738585072014e-308"),
    (2.9802322387695312e-8, "2.9802322387695312e-8"),
    (-2.109808898695963e16, "-2.109808898695963e16"), (4.940656e-318, "4.940656e-318"),
    (1.18575755e-316, "1.18575755e-316"), (2.989102097996e-312, "2.989102097996e-312"),
    (9.0608011534336e15, "9060801153433600.0"), (4.708356024711512e18, "4.708356024711512e18"),
    (9.409340012568248e18, "9.409340012568248e18"), (1.2345678, "1.2345678"),
    (5.764607523034235e39, "5.764607523034235e39"), (1.152921504606847e40, "1.152921504606847e40"),
    (2.305843009213694e40, "2.305843009213694e40"), (1.2345678901234567, "1.2345678901234567"),
]

# the reference's own values that go through its CSV files (cli/src/tests.rs:35-63,
# core/src/lib.rs tests, cli/src/tests.rs:9-33)
REFERENCE_VALUES = [(0.75, "0.75"), (1.25, "1.25"), (0.5, "0.5"), (2.0, "2.0"), (33.38339, "33.38339"),
                    (66.335, "66.335"), (0.071, "0.071"), (0.1, "0.1"), (0.2, "0.2"), (0.3, "0.3")]


def random_doubles(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 2 ** 64, n, dtype=np.uint64).view(np.float64)
    return [float(v) for v in x if math.isfinite(v)]


def boundary_doubles():
    out = []
    for e in list(range(-1074, -1000, 7)) + list(range(-60, 70, 3)) + list(range(960, 1024, 7)):
        p = math.ldexp(1.0, e)
        out += [p, math.nextafter(p, 0.0), math.nextafter(p, math.inf)]
    out += [2.0 ** 53, 2.0 ** 53 + 2, 2.0 ** 54 + 4, 9007199254740993.0, 1e15, 1e16, 1e17, 1e-5, 1e-6,
            0.1 + 0.2, 1 / 3, 2 / 3, 123456789012345678.0]
    return out


@pytest.mark.parametrize("x,text", RYU_KNOWN + REFERENCE_VALUES)
def test_known_texts(x, text):
    assert T.f64_text(x) == text
    if math.isfinite(x) and x != 0.0:
        assert T.f64_text(x, exact=True) == text


def test_repr_digits_are_the_exact_shortest():
    """Python's repr (Gay) gives the shortest, closest digits that the exact
    search defines, on random bit patterns and at binade edges."""
    for x in random_doubles(1500, 1) + boundary_doubles():
        if x == 0.0:
            continue
        assert T.shortest_repr(x) == T.shortest_exact(x), repr(x)


def test_exact_search_catches_planted_mistakes():
    """A wrong shortest-digit rule fails against the exact search: keeping 17
    digits, or rounding the 17th instead of choosing the shortest."""
    for x in [0.1, 0.3, 1 / 3, 2.0 ** -20]:
        exact = T.shortest_exact(x)
        seventeen = T._strip(("%.16e" % x).replace(".", "").split("e")[0], int(("%.16e" % x).split("e")[1]) - 16)
        assert exact == T.shortest_repr(x)
        if x in (0.1, 0.3):
            assert seventeen != exact


def test_layout_thresholds():
    assert [T.f64_text(v) for v in (1e15, 1e16, 1e-5, 1e-6, 123456.0, 1.5e16, 0.00012)] == \
        ["1000000000000000.0", "1e16", "0.00001", "1e-6", "123456.0", "1.5e16", "0.00012"]


def test_frame_and_bb_text():
    ids = np.array([0, 0, 1])
    x = np.array([[0.75, 0.75, 0.5], [1.25, 0.75, 0.5], [0.0, 1e-7, 2.0]])
    v = np.array([[1.0, 1.0, 0.0], [-1.0, 1.0, 0.0], [-0.0, 3.5e20, 0.1]])
    assert T.frame_csv(ids, x, v) == (
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n"
        "0,0.75,0.75,0.5,1.0,1.0,0.0\n"
        "0,1.25,0.75,0.5,-1.0,1.0,0.0\n"
        "1,0.0,1e-7,2.0,-0.0,3.5e20,0.1\n")
    assert T.frame_csv([], np.zeros((0, 3)), np.zeros((0, 3))) == ""
    rows = T.bb_save([], 0, [2.0, 2.0, 2.0])
    rows = T.bb_save(rows, 1, [3.0, 3.0, 3.0])
    rows = T.bb_save(rows, 0, [1.5, 2.0, 2.0])
    rows = T.bb_save(rows, 7, [4.0, 4.0, 4.0])  # save_bb appends past the end
    assert T.bb_csv(rows) == "x,y,z\n1.5,2.0,2.0\n3.0,3.0,3.0\n4.0,4.0,4.0\n"
    ids2, x2, v2 = T.read_frame(T.frame_csv(ids, x, v))
    assert np.array_equal(ids2, ids) and np.array_equal(x2, x) and np.array_equal(v2, v)


def _write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)


def _load(tmp, n):
    from moldyn_amd.frames import StateToSave
    return StateToSave.load_from_file(str(tmp), n)


def test_frame_load_reads_the_oracle_text_exactly(tmp_path):
    """md_frame_load (host) reads back every double bit for bit, by column
    name, on a frame of random and boundary values."""
    os.makedirs(tmp_path / "data")
    vals = np.array(random_doubles(6000, 2) + boundary_doubles() + [0.0, -0.0, 1e-7, 0.1])
    n = len(vals) // 6
    vals = vals[:6 * n].reshape(n, 6)
    ids = np.arange(n) % 3
    _write(tmp_path / "data" / "0.csv", T.frame_csv(ids, vals[:, :3], vals[:, 3:]))
    _write(tmp_path / "bb.csv", T.bb_csv([[2.0, 3.0, 4.5], [1.0, 1.0, 1.0]]))
    s = _load(tmp_path, 0)
    assert np.array_equal(s.ids, ids)
    assert s.x.tobytes() == np.ascontiguousarray(vals[:, :3]).tobytes()
    assert s.v.tobytes() == np.ascontiguousarray(vals[:, 3:]).tobytes()
    assert list(s.boundary_box) == [2.0, 3.0, 4.5]
    # columns in another order, CRLF line ends, a '+' sign and a blank line
    _write(tmp_path / "data" / "1.csv", "velocity_z,id,position_x,position_y,position_z,velocity_x,velocity_y\r\n"
                                        "0.5,2,+1.0,2e0,3.,.25,1E-3\r\n\r\n")
    s = _load(tmp_path, 1)
    assert list(s.ids) == [2] and list(s.x[0]) == [1.0, 2.0, 3.0] and list(s.v[0]) == [0.25, 1e-3, 0.5]
    assert list(s.boundary_box) == [1.0, 1.0, 1.0]


@pytest.mark.parametrize("text,what", [
    ("id,position_x,position_y,position_z,velocity_x,velocity_y\n0,1,2,3,4,5\n", "missing column"),
    ("id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n0,1,2,3,4,5\n", "fields"),
    ("id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n0,1,2,x,4,5,6\n", "bad number"),
    ("id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n70000,1,2,3,4,5,6\n", "bad integer"),
    ("id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n0, 1,2,3,4,5,6\n", "bad number"),
])
def test_frame_load_errors(tmp_path, text, what):
    import moldyn_amd as M
    os.makedirs(tmp_path / "data")
    _write(tmp_path / "data" / "0.csv", text)
    _write(tmp_path / "bb.csv", "x,y,z\n1.0,1.0,1.0\n")
    with pytest.raises(M.MdError, match=what):
        _load(tmp_path, 0)


def test_frame_load_missing_box_row(tmp_path):
    import moldyn_amd as M
    os.makedirs(tmp_path / "data")
    _write(tmp_path / "data" / "3.csv", "")
    _write(tmp_path / "bb.csv", "x,y,z\n1.0,1.0,1.0\n")
    with pytest.raises(M.MdError, match="no row 3"):
        _load(tmp_path, 3)


def test_state_into_groups_by_id():
    """Into<State> (save_data.rs:129-151): types 0..=max id, file order kept."""
    from moldyn_amd.frames import StateToSave
    ids = np.array([1, 0, 1, 0, 0], dtype=np.uint16)
    x = np.arange(15.0).reshape(5, 3)
    s = StateToSave(ids, x, -x, np.ones(3))
    counts, masses, xs, vs = s.into_state({0: 2.0, 1: 3.0})
    assert list(counts) == [3, 2] and masses == [2.0, 3.0]
    assert np.array_equal(xs, x[[1, 3, 4, 0, 2]]) and np.array_equal(vs, -xs)
    with pytest.raises(KeyError):
        s.into_state({0: 2.0})


def test_macro_file_mirror():
    """DataFileMacro (save_data.rs:249-288, 317-337): unit values per
    particle, frame counting, and load's start/count."""
    from moldyn_amd.frames import DataFileMacro
    d = DataFileMacro()
    d.add_macro_params(2, [("kinetic_energy", 10.0), ("temperature", 300.0)], 4)
    d.add_macro_params(0, [("potential_energy", -8.0), ("pressure", 1.5)], 4)
    assert d.macro_parameters[2]["unit_kinetic_energy"] == 2.5
    assert d.macro_parameters[0]["unit_potential_energy"] == -2.0
    assert d.frame_count == 2
    with pytest.raises(NotImplementedError):
        d.add_macro_params(1, [("custom", 1.0)], 4)


def test_particles_db_and_potentials_read(tmp_path):
    """db.csv and potentials.json as the reference writes them are read back."""
    from moldyn_amd import frames as F
    _write(tmp_path / "db.csv", 'id,name,mass,radius\n0,Argon,66.335,0.071\n3,"Ne, light",20.18,0.0\n0,Dup,1.0,1.0\n')
    db = F.load_particles_data(str(tmp_path))
    assert db == {0: ("Argon", 66.335, 0.071), 3: ("Ne, light", 20.18, 0.0)}
    _write(tmp_path / "potentials.json", '{\n  "0,0": {\n    "LennardJones": {\n      "sigma": 0.3418,\n'
                                         '      "eps": 1.712,\n      "r_cut": 0.8545,\n'
                                         '      "u_cut": -0.027934517624831987\n    }\n  }\n}')
    p = F.load_potentials_from_file(str(tmp_path))
    assert p == {(0, 0): ("LennardJones", 0.3418, 1.712, 0.8545, -0.027934517624831987)}
