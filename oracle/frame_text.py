"""CPU oracle of the reference's state/frame files -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker of the GPU frame writer (moldyn_amd/csrc/md_frames.hip);
no product code imports it.

What it restates
----------------
* The files (core/src/save_data.rs):
  - data/<n>.csv: header "id,position_x,position_y,position_z,velocity_x,
    velocity_y,velocity_z", one row per particle, types in order
    (ParticleToSave, StateToSave::from, :11-27, :113-127, :192-213);
  - bb.csv: header "x,y,z", row n = box of state n; a save replaces row n
    when it exists, else appends a row (:155-190);
  - macro.csv: header "iteration,kinetic_energy,...,custom", rows sorted by
    iteration, unit_* = value / particle count (:39-51, :249-300).
  The `csv` 1.3 crate writes the header with the first record (so an empty
  record list gives an empty file), "," between fields, "\\n" after each row,
  and no quotes around numbers.
* The text of an f64 field: csv and serde_json print f64 with the `ryu` crate
  (third-party, absent from the reference; version unpinned, the reference
  has no Cargo.lock; ryu's output format is fixed across its 1.x releases):
  the shortest decimal in the rounding interval of the double, the closest
  to it among equally short ones, ties to an even last digit; laid out by
  `layout()` below (ryu's pretty::format64).

How it is pinned
----------------
* shortest_exact(): the definition itself, computed with exact decimal
  arithmetic (the double's exact value, the exact midpoints to its
  neighbours, every candidate length in turn).
* shortest_repr(): Python's repr (David Gay's shortest round-trip digits, an
  implementation independent of Ryu); tests check it against
  shortest_exact() on random and boundary doubles, then use it for bulk data.
* layout(): ryu's published examples (README: 1.23e40, 1.23e-40; docs:
  1.234) and the digit strings of ryu's own regression tests
  (2.109808898695963E16, 2.9802322387695312E-8, 5E-324,
  1.7976931348623157E308, ...).  The two thresholds of the layout (plain
  digits while kk <= 16, "0.000..." while kk > -5) are restated from ryu's
  source and fixed by none of these examples at the boundary itself: parity
  of the layout at exactly those boundaries is unpinned beyond this
  restatement.
* Reading back: Python's float() is correctly rounded, as Rust's
  f64::from_str the reference's csv reader uses.
"""
from __future__ import annotations

import math
from decimal import ROUND_CEILING, ROUND_FLOOR, Decimal, localcontext

import numpy as np

FRAME_HEADER = "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z"
BB_HEADER = "x,y,z"
MACRO_HEADER = ("iteration,kinetic_energy,potential_energy,thermal_energy,unit_kinetic_energy,"
                "unit_potential_energy,unit_thermal_energy,temperature,pressure,custom")


def _strip(digits: str, exp: int):
    digits = digits.lstrip("0") or "0"
    while len(digits) > 1 and digits.endswith("0"):
        digits = digits[:-1]
        exp += 1
    return digits, exp


def shortest_exact(x: float):
    """(digits, exp10) of finite nonzero |x| by exhaustive exact search."""
    x = abs(float(x))
    assert math.isfinite(x) and x != 0.0
    with localcontext() as ctx:
        ctx.prec = 1200
        v = Decimal(x)
        below = math.nextafter(x, 0.0)
        above = math.nextafter(x, math.inf)
        lo = (v + Decimal(below)) / 2
        if math.isinf(above):
            hi = v + (v - Decimal(below)) / 2
        else:
            hi = (v + Decimal(above)) / 2
        mant = int.from_bytes(np.float64(x).tobytes(), "little") & ((1 << 52) - 1)
        if x >= 2.0 ** -1022:
            mant |= 1 << 52
        inclusive = mant % 2 == 0

        def inside(val):
            return (lo <= val <= hi) if inclusive else (lo < val < hi)

        for p in range(1, 18):
            e = v.adjusted() - p + 1
            q = Decimal(1).scaleb(e)
            cands = []
            for rounding in (ROUND_FLOOR, ROUND_CEILING):
                d = (v / q).to_integral_value(rounding=rounding)
                if inside(d * q) and d not in [c for c in cands]:
                    cands.append(d)
            if cands:
                best = min(cands, key=lambda d: (abs(d * q - v), int(d) % 2))
                return _strip(str(int(best)), e)
    raise AssertionError("no candidate of 17 digits")


def shortest_repr(x: float):
    """(digits, exp10) of finite nonzero |x| from Python's repr."""
    s = repr(abs(float(x)))
    m, _, e = s.partition("e")
    exp = int(e) if e else 0
    ip, _, fp = m.partition(".")
    return _strip(ip + fp, exp - len(fp))


def layout(neg: bool, digits: str, exp10: int) -> str:
    """ryu pretty::format64 layout of digits * 10^exp10."""
    n = len(digits)
    k = exp10
    kk = n + k
    s = "-" if neg else ""
    if 0 <= k and kk <= 16:
        return s + digits + "0" * (kk - n) + ".0"
    if 0 < kk <= 16:
        return s + digits[:kk] + "." + digits[kk:]
    if -5 < kk <= 0:
        return s + "0." + "0" * (-kk) + digits
    e = str(kk - 1)
    if n == 1:
        return s + digits + "e" + e
    return s + digits[0] + "." + digits[1:] + "e" + e


def f64_text(x, exact=False) -> str:
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "-inf" if x < 0 else "inf"
    neg = math.copysign(1.0, x) < 0
    if x == 0.0:
        return "-0.0" if neg else "0.0"
    d, e = (shortest_exact if exact else shortest_repr)(x)
    return layout(neg, d, e)


def frame_csv(ids, x, v) -> str:
    """data/<n>.csv text of a state (save_data.rs:192-213)."""
    ids = np.asarray(ids)
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    if len(ids) == 0:
        return ""
    out = [FRAME_HEADER]
    for i in range(len(ids)):
        out.append(",".join([str(int(ids[i]))] + [f64_text(a) for a in x[i]] + [f64_text(a) for a in v[i]]))
    return "\n".join(out) + "\n"


def table_csv(header: str, rows) -> str:
    """CSV text of rows of numbers (ints printed as integers, floats by ryu)."""
    rows = list(rows)
    if not rows:
        return ""
    out = [header]
    for r in rows:
        out.append(",".join(str(int(a)) if isinstance(a, (int, np.integer)) else f64_text(a) for a in r))
    return "\n".join(out) + "\n"


def bb_csv(boxes) -> str:
    return table_csv(BB_HEADER, [list(map(float, b)) for b in boxes])


def bb_save(rows: list, state_number: int, box) -> list:
    """The row list after StateToSave::save_bb(path, state_number)
    (save_data.rs:165-184): replace row n when it exists, else append."""
    rows = [list(r) for r in rows]
    if len(rows) > state_number:
        rows[state_number] = list(map(float, box))
    else:
        rows.append(list(map(float, box)))
    return rows


def read_csv(text: str):
    """(header fields, rows of fields) of csv text."""
    lines = [ln for ln in text.replace("\r\n", "\n").split("\n") if ln != ""]
    if not lines:
        return [], []
    return lines[0].split(","), [ln.split(",") for ln in lines[1:]]


def read_frame(text: str):
    """(ids, x, v) of data/<n>.csv text, columns matched by header name as
    serde does (StateToSave::load_from_file, save_data.rs:215-229)."""
    head, rows = read_csv(text)
    col = {h: k for k, h in enumerate(head)}
    ids = np.array([int(r[col["id"]]) for r in rows], dtype=np.int64)
    x = np.array([[float(r[col[f"position_{a}"]]) for a in "xyz"] for r in rows]).reshape(-1, 3)
    v = np.array([[float(r[col[f"velocity_{a}"]]) for a in "xyz"] for r in rows]).reshape(-1, 3)
    return ids, x, v
