"""The reference's on-disk state, written and read through libmoldyn_hip.

Mirrors core/src/save_data.rs (StateToSave, DataFileMacro), the db.csv of
core/src/particles_database.rs:157-228 and the potentials.json of
solver/src/solver/potential.rs:104-139.  Every f64 is printed on the GPU
(md_frame_writer_*, md_text_table) in the text the reference's csv 1.3 and
serde_json 1.0 crates give it (the ryu crate's shortest round-trip form), so
the files are the reference's byte for byte.  Two deliberate differences:
files are truncated when rewritten (the reference's open_file_or_create,
core/src/lib.rs:17-23, leaves stale bytes after a shorter rewrite), and
db.csv / potentials.json list their entries sorted by id (the reference
iterates a HashMap, in no fixed order).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

FRAME_HEADER = "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z"
BB_HEADER = "x,y,z"
MACRO_FIELDS = ("kinetic_energy", "potential_energy", "thermal_energy", "unit_kinetic_energy",
                "unit_potential_energy", "unit_thermal_energy", "temperature", "pressure", "custom")
MACRO_HEADER = "iteration," + ",".join(MACRO_FIELDS)
DB_HEADER = "id,name,mass,radius"


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class TextFormatter:
    """GPU formatter of numeric tables (md_text_*)."""

    def __init__(self):
        self._lib = L.load()
        h = C.c_void_p()
        rc = self._lib.md_text_create(C.byref(h))
        if rc != 0:
            raise L.MdError(rc, self._lib.md_global_error().decode())
        self._h = h

    def _chk(self, rc):
        if rc != 0:
            raise L.MdError(rc, self._lib.md_text_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.md_text_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _args(vals, keys):
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        if vals.ndim == 1:
            vals = vals[:, None]
        k = None
        if keys is not None:
            k = np.ascontiguousarray(keys, dtype=np.uint64)
            if len(k) != len(vals):
                raise ValueError("one key per row")
        return vals, k

    def table(self, vals, keys=None) -> str:
        """Rows of `keys` (optional integer column) and the columns of `vals`."""
        vals, k = self._args(vals, keys)
        kp = k.ctypes.data_as(C.POINTER(C.c_uint64)) if k is not None else None
        n = C.c_int64()
        self._chk(self._lib.md_text_table(self._h, len(vals), vals.shape[1], kp, _dp(vals), None, 0, C.byref(n)))
        buf = C.create_string_buffer(max(n.value, 1))
        self._chk(self._lib.md_text_table(self._h, len(vals), vals.shape[1], kp, _dp(vals), buf, n.value,
                                          C.byref(n)))
        return buf.raw[:n.value].decode()

    def f64(self, values) -> list:
        """The text of each value."""
        values = np.atleast_1d(np.asarray(values, dtype=np.float64))
        if len(values) == 0:
            return []
        return self.table(values).split("\n")[:-1]

    def write_csv(self, path, header, vals, keys=None):
        vals, k = self._args(vals, keys)
        kp = k.ctypes.data_as(C.POINTER(C.c_uint64)) if k is not None else None
        self._chk(self._lib.md_csv_write_table(self._h, os.fsencode(path), header.encode(), len(vals),
                                               vals.shape[1], kp, _dp(vals)))


_text = None


def text_formatter() -> TextFormatter:
    global _text
    if _text is None:
        _text = TextFormatter()
    return _text


class FrameWriter:
    """Asynchronous writer of data/<n>.csv and bb.csv from an Engine's
    device state (md_frame_writer_*).  save(n) stands for
    StateToSave::from(&state).save_to_file(path, n) (save_data.rs:192-213)."""

    def __init__(self, engine, path, replica=0):
        self._lib = L.load()
        self.engine = engine  # the context must outlive the writer
        h = C.c_void_p()
        rc = self._lib.md_frame_writer_create(C.byref(h), engine._h, os.fsencode(str(path)), int(replica))
        if rc != 0:
            raise L.MdError(rc, self._lib.md_global_error().decode())
        self._h = h

    def _chk(self, rc):
        if rc != 0:
            raise L.MdError(rc, self._lib.md_frame_writer_last_error(self._h).decode())

    def save(self, state_number):
        self._chk(self._lib.md_frame_writer_save(self._h, int(state_number)))

    def flush(self):
        self._chk(self._lib.md_frame_writer_flush(self._h))

    def stats(self):
        out = np.zeros(6)
        self._chk(self._lib.md_frame_writer_stats(self._h, _dp(out)))
        return dict(zip(("frames", "bytes", "gpu_ms", "copy_ms", "write_ms", "wait_ms"), out.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            try:
                self.flush()
            finally:
                self._lib.md_frame_writer_destroy(self._h)
                self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _read_csv(path):
    with open(path, newline="") as f:
        lines = [ln for ln in f.read().replace("\r\n", "\n").split("\n") if ln != ""]
    if not lines:
        return [], []
    return lines[0].split(","), [ln.split(",") for ln in lines[1:]]


def _f64(s: str) -> float:
    # Rust's str::parse::<f64>: optional sign, decimal or exponent form, inf/nan
    t = s[1:] if s.startswith("+") and not s.startswith(("++", "+-")) else s
    if t.lower() in ("inf", "+inf", "-inf", "infinity", "-infinity", "nan", "-nan"):
        return float(t)
    if not t or any(c not in "0123456789.eE+-" for c in t):
        raise ValueError(f"invalid float literal {s!r}")
    return float(t)


def load_boxes(path):
    """bb.csv rows (StateToSave::get_bbs, save_data.rs:155-163)."""
    head, rows = _read_csv(os.path.join(path, "bb.csv"))
    if not rows:
        return np.zeros((0, 3))
    col = [head.index(c) for c in ("x", "y", "z")]
    return np.array([[_f64(r[c]) for c in col] for r in rows])


@dataclass
class StateToSave:
    """Particles in file order (id = type) and the box (save_data.rs:22-27)."""
    ids: np.ndarray
    x: np.ndarray
    v: np.ndarray
    boundary_box: np.ndarray

    @classmethod
    def from_engine(cls, engine, replica=0):
        d = engine.download(("x", "v", "box"))
        counts = engine.counts[replica]
        first = int(engine.counts[:replica].sum())
        n = int(counts.sum())
        ids = np.repeat(np.arange(len(counts), dtype=np.uint16), counts)
        return cls(ids, d["x"][first:first + n].copy(), d["v"][first:first + n].copy(),
                   np.asarray(d["box"]).reshape(-1, 3)[replica].copy())

    def save_to_file(self, path, state_number):
        """save_data.rs:192-213, with the text made on the GPU."""
        path = str(path)
        os.makedirs(os.path.join(path, "data"), exist_ok=True)
        boxes = load_boxes(path) if os.path.exists(os.path.join(path, "bb.csv")) else np.zeros((0, 3))
        rows = [list(b) for b in boxes]
        if len(rows) > state_number:
            rows[state_number] = list(map(float, self.boundary_box))
        else:
            rows.append(list(map(float, self.boundary_box)))
        tf = text_formatter()
        tf.write_csv(os.path.join(path, "bb.csv"), BB_HEADER, np.array(rows, dtype=np.float64).reshape(-1, 3))
        vals = np.hstack([np.asarray(self.x, np.float64).reshape(-1, 3), np.asarray(self.v, np.float64).reshape(-1, 3)])
        tf.write_csv(os.path.join(path, "data", f"{state_number}.csv"), FRAME_HEADER, vals,
                     keys=np.asarray(self.ids, dtype=np.uint64))

    @classmethod
    def load_from_file(cls, path, state_number):
        """save_data.rs:215-229 (md_frame_load)."""
        lib = L.load()
        p = os.fsencode(str(path))
        n = C.c_int64()
        if lib.md_frame_load(p, int(state_number), 0, C.byref(n), None, None, None, None) != 0:
            raise L.MdError(L.MD_ERR_ARG, lib.md_global_error().decode())
        ids = np.zeros(n.value, dtype=np.uint16)
        x = np.zeros((n.value, 3))
        v = np.zeros((n.value, 3))
        box = np.zeros(3)
        rc = lib.md_frame_load(p, int(state_number), n.value, C.byref(n),
                               ids.ctypes.data_as(C.POINTER(C.c_uint16)), _dp(x), _dp(v), _dp(box))
        if rc != 0:
            raise L.MdError(rc, lib.md_global_error().decode())
        return cls(ids, x, v, box)

    def into_state(self, masses):
        """Into<State> (save_data.rs:129-151): particles grouped by id into
        max_id + 1 types, file order kept within a type.  masses: id -> mass
        (ParticleDatabase); returns counts, per-type masses, x, v type-major."""
        n_types = int(self.ids.max()) + 1 if len(self.ids) else 1
        for t in np.unique(self.ids):
            if int(t) not in masses:
                raise KeyError(f"Can't convert particle: no mass for id {int(t)}")
        order = np.argsort(self.ids, kind="stable")
        counts = np.bincount(self.ids.astype(np.int64), minlength=n_types)
        mass = [float(masses[t]) if t in masses else 0.0 for t in range(n_types)]
        return counts, mass, self.x[order], self.v[order]


# -------------------------------------------------------------- macro.csv

@dataclass
class DataFileMacro:
    """save_data.rs:53-60, 232-338: macro parameters by frame."""
    macro_parameters: dict = field(default_factory=dict)
    start_frame: int = 0
    frame_count: int = 0

    def add_macro_params(self, frame, parameters, particles_count):
        """parameters: (kind, value) pairs, kind in kinetic_energy,
        potential_energy, thermal_energy, temperature, pressure (:249-288)."""
        row = self.macro_parameters.setdefault(int(frame), dict.fromkeys(MACRO_FIELDS, 0.0))
        for kind, value in parameters:
            value = float(value)
            if kind in ("kinetic_energy", "potential_energy", "thermal_energy"):
                row[kind] = value
                row["unit_" + kind] = value / float(particles_count)
            elif kind in ("pressure", "temperature"):
                row[kind] = value
            elif kind == "custom":
                raise NotImplementedError("MacroParameterType::Custom is todo!() in the reference")
            else:
                raise ValueError(kind)
        self.frame_count += 1

    def save_to_file(self, path):
        its = sorted(self.macro_parameters)
        vals = np.array([[self.macro_parameters[i][k] for k in MACRO_FIELDS] for i in its],
                        dtype=np.float64).reshape(-1, len(MACRO_FIELDS))
        text_formatter().write_csv(str(path), MACRO_HEADER, vals, keys=np.array(its, dtype=np.uint64))

    def reset_old(self):
        self.macro_parameters.clear()
        self.start_frame += self.frame_count
        self.frame_count = 0

    def append_data(self, other):
        self.macro_parameters.update({k: dict(v) for k, v in other.macro_parameters.items()})

    @classmethod
    def load_from_file(cls, path):
        head, rows = _read_csv(str(path))
        col = {h: k for k, h in enumerate(head)}
        mp = {}
        for r in rows:
            it = int(r[col["iteration"]])
            mp[it] = {k: _f64(r[col[k]]) for k in MACRO_FIELDS}
        lo = min(mp) if mp else 2 ** 64 - 1
        hi = max(mp) if mp else 0
        return cls(mp, lo, hi - lo)


# -------------------------------------------------------------- db.csv

def _csv_field(s: str) -> str:
    # csv QuoteStyle::Necessary: quote fields holding a delimiter, quote or line break
    if any(c in s for c in ',"\r\n'):
        return '"' + s.replace('"', '""') + '"'
    return s


def save_particles_data(path, db):
    """db.csv (ParticleDatabase::save_particles_data, particles_database.rs:
    157-199); db: id -> (name, mass, radius)."""
    os.makedirs(str(path), exist_ok=True)
    ids = sorted(db)
    nums = text_formatter().f64([x for i in ids for x in (db[i][1], db[i][2])])
    lines = [DB_HEADER] + [f"{i},{_csv_field(db[i][0])},{nums[2 * k]},{nums[2 * k + 1]}" for k, i in enumerate(ids)]
    with open(os.path.join(str(path), "db.csv"), "w", newline="") as f:
        f.write("\n".join(lines) + "\n" if ids else "")


def _csv_split(line: str):
    out, cur, q, i = [], "", False, 0
    while i < len(line):
        ch = line[i]
        if q:
            if ch == '"' and i + 1 < len(line) and line[i + 1] == '"':
                cur += '"'
                i += 1
            elif ch == '"':
                q = False
            else:
                cur += ch
        elif ch == '"':
            q = True
        elif ch == ",":
            out.append(cur)
            cur = ""
        else:
            cur += ch
        i += 1
    out.append(cur)
    return out


def load_particles_data(path):
    """db.csv rows: id -> (name, mass, radius); the first row of an id wins
    (particles_database.rs:209-228, entry().or_insert)."""
    with open(os.path.join(str(path), "db.csv"), newline="") as f:
        lines = [ln for ln in f.read().replace("\r\n", "\n").split("\n") if ln != ""]
    db = {}
    if not lines:
        return db
    head = _csv_split(lines[0])
    col = {h: k for k, h in enumerate(head)}
    for ln in lines[1:]:
        r = _csv_split(ln)
        i = int(r[col["id"]])
        if not 0 <= i <= 65535:
            raise ValueError("id out of u16 range")
        db.setdefault(i, (r[col["name"]], _f64(r[col["mass"]]), _f64(r[col["radius"]])))
    return db


# -------------------------------------------------------------- potentials.json

def _json_str(s: str) -> str:
    return json.dumps(s, ensure_ascii=False)


def _json_pretty(v, nums, ind=0) -> str:
    # serde_json::ser::to_writer_pretty: two-space indent, "key": value
    pad, pad2 = "  " * ind, "  " * (ind + 1)
    if isinstance(v, dict):
        if not v:
            return "{}"
        items = [pad2 + _json_str(k) + ": " + _json_pretty(x, nums, ind + 1) for k, x in v.items()]
        return "{\n" + ",\n".join(items) + "\n" + pad + "}"
    if isinstance(v, list):
        if not v:
            return "[]"
        return "[\n" + ",\n".join(pad2 + _json_pretty(x, nums, ind + 1) for x in v) + "\n" + pad + "]"
    if isinstance(v, str):
        return _json_str(v)
    if isinstance(v, float):
        return nums[id(v)] if math.isfinite(v) else "null"
    return str(v)


def save_potentials_to_file(path, potentials):
    """potentials.json (PotentialsDatabase::save_potentials_to_file,
    potential.rs:104-122); potentials: (i, j) -> ("LennardJones", sigma, eps,
    r_cut, u_cut) or ("Custom", name, [values])."""
    doc, floats = {}, []
    for (i, j) in sorted(potentials):
        p = potentials[(i, j)]
        if p[0] == "LennardJones":
            body = {"sigma": float(p[1]), "eps": float(p[2]), "r_cut": float(p[3]), "u_cut": float(p[4])}
            floats += list(body.values())
        elif p[0] == "Custom":
            body = {"name": str(p[1]), "custom_data": [float(x) for x in p[2]]}
            floats += body["custom_data"]
        else:
            raise ValueError(p[0])
        doc[f"{i},{j}"] = {p[0]: body}
    texts = text_formatter().f64(floats) if floats else []
    nums = {id(f): t for f, t in zip(floats, texts)}
    os.makedirs(str(path), exist_ok=True)
    with open(os.path.join(str(path), "potentials.json"), "w") as f:
        f.write(_json_pretty(doc, nums))


def load_potentials_from_file(path):
    """potential.rs:124-139: keys "i,j" -> potential tuples as above."""
    with open(os.path.join(str(path), "potentials.json")) as f:
        doc = json.load(f)
    out = {}
    for key, pot in doc.items():
        i, j = (int(x) for x in key.split(","))
        if not (0 <= i <= 65535 and 0 <= j <= 65535):
            raise ValueError(f"Can't convert {key} to u16")
        (kind, body), = pot.items()
        if kind == "LennardJones":
            out[(i, j)] = ("LennardJones", float(body["sigma"]), float(body["eps"]), float(body["r_cut"]),
                           float(body["u_cut"]))
        elif kind == "Custom":
            out[(i, j)] = ("Custom", body["name"], [float(x) for x in body["custom_data"]])
        else:
            raise ValueError(kind)
    return out


def apply_potentials(engine, potentials):
    """PotentialsDatabase::set_potential for every entry (the key order
    (min, max), potential.rs:141-144); types without an entry keep the
    default Ar LJ(0.3418, 1.712).  Custom potentials are todo!() in the
    reference and rejected."""
    for (i, j), p in potentials.items():
        a, b = min(i, j), max(i, j)
        if b >= engine.n_types:
            continue  # a pair of types this system does not have
        if p[0] != "LennardJones":
            raise L.MdError(L.MD_ERR_UNIMPLEMENTED, "Potential::Custom is todo!() in the reference")
        engine.set_pair_potential(a, b, p[1], p[2], r_cut=p[3], u_cut=p[4])
