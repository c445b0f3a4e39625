"""ctypes front end of the CPU FP64 oracle (oracle/md_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker.  The HIP product path
never imports it.

The functions mirror moldyn's solver API (AndrewChe7/moldyn
solver/src/solver/potential.rs, integrator.rs, initializer/*.rs,
macro_parameters/*.rs) over a flat type-major state.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmd_oracle.so")

K_B = 1.380648528  # core/src/lib.rs:15


class OrcLJ(C.Structure):
    _fields_ = [("sigma", C.c_double), ("eps", C.c_double), ("r_cut", C.c_double),
                ("u_cut", C.c_double)]


class OrcState(C.Structure):
    _fields_ = [("n_types", C.c_int), ("counts", C.POINTER(C.c_int64)),
                ("pos", C.POINTER(C.c_double)), ("vel", C.POINTER(C.c_double)),
                ("force", C.POINTER(C.c_double)), ("pot", C.POINTER(C.c_double)),
                ("temp", C.POINTER(C.c_double)), ("mass", C.POINTER(C.c_double)),
                ("box", C.c_double * 3)]


class Thermostat(C.Structure):
    """initializer/thermostat.rs:4-20: kind 1 Berendsen{tau, lambda}, 2 NoseHoover{tau, psi, lambda}."""
    _fields_ = [("kind", C.c_int), ("tau", C.c_double), ("lambda_", C.c_double),
                ("psi", C.c_double)]

    @classmethod
    def berendsen(cls, tau, lambda_=0.0):
        return cls(1, tau, lambda_, 0.0)

    @classmethod
    def nose_hoover(cls, tau, psi=0.0, lambda_=0.0):
        return cls(2, tau, lambda_, psi)


class Barostat(C.Structure):
    """initializer/barostat.rs:4-16: kind 1 Berendsen{beta, tau, myu}."""
    _fields_ = [("kind", C.c_int), ("beta", C.c_double), ("tau", C.c_double),
                ("myu", C.c_double)]

    @classmethod
    def berendsen(cls, beta, tau, myu=0.0):
        return cls(1, beta, tau, myu)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        sp = C.POINTER(OrcState)
        lp = C.POINTER(OrcLJ)
        L.orc_new_lennard_jones.argtypes = [C.c_double, C.c_double, lp]
        L.orc_lj_potential_and_force.argtypes = [lp, C.c_double, dp, dp]
        L.orc_update_force.argtypes = [sp, lp, C.c_int]
        L.orc_update_force_rows.argtypes = [sp, lp, C.c_int64, C.c_int64, C.c_int]
        L.orc_update_force_cells.argtypes = [sp, lp, C.c_int]
        L.orc_update_force_cells.restype = C.c_int
        L.orc_pair_counts.argtypes = [sp, lp, C.c_double, C.POINTER(C.c_int32), C.c_int]
        L.orc_apply_boundary_conditions.argtypes = [sp]
        for name in ("orc_center_of_mass_velocity", "orc_momentum"):
            getattr(L, name).argtypes = [sp, C.c_int, dp]
        for name in ("orc_kinetic_energy", "orc_potential_energy"):
            getattr(L, name).argtypes = [sp, C.c_int]
            getattr(L, name).restype = C.c_double
        for name in ("orc_thermal_energy", "orc_pressure"):
            getattr(L, name).argtypes = [sp, C.c_int, dp]
            getattr(L, name).restype = C.c_double
        L.orc_temperature.argtypes = [C.c_double, C.c_int64]
        L.orc_temperature.restype = C.c_double
        tp = C.POINTER(Thermostat)
        bp = C.POINTER(Barostat)
        L.orc_thermostat_calculate_lambda.argtypes = [tp, sp, C.c_double, C.c_int, C.c_double]
        L.orc_thermostat_update.argtypes = [tp, sp, C.c_double, C.c_int, C.c_double]
        L.orc_barostat_calculate_myu.argtypes = [bp, sp, C.c_double, C.c_int, C.c_double]
        L.orc_barostat_update.argtypes = [bp, sp, C.c_int]
        L.orc_verlet_step.argtypes = [sp, lp, C.c_double, bp, C.c_double, tp, C.c_double, C.c_int]
        L.orc_verlet_step_cells.argtypes = [sp, lp, C.c_double, bp, C.c_double, tp, C.c_double, C.c_int]
        L.orc_verlet_step_cells.restype = C.c_int
        i64p = C.POINTER(C.c_int64)
        L.orc_lattice_u.argtypes = [dp, dp, i64p, C.c_double]
        L.orc_lattice_fcc.argtypes = [dp, dp, i64p, C.c_double]
        L.orc_mb_velocities.argtypes = [dp, C.c_int64, C.c_double, C.c_double, C.c_uint64]
        L.orc_uniform.argtypes = [dp, C.c_int64, C.c_uint64]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_threads():
    return max(1, min(os.cpu_count() or 1, 16))


class State:
    """Flat type-major particle state (core/src/particle.rs:5-32)."""

    def __init__(self, counts, masses, box, pos=None, vel=None):
        self.counts = np.ascontiguousarray(counts, dtype=np.int64)
        n = int(self.counts.sum())
        self.n = n
        self.pos = np.zeros((n, 3)) if pos is None else np.array(pos, dtype=np.float64).reshape(n, 3)
        self.vel = np.zeros((n, 3)) if vel is None else np.array(vel, dtype=np.float64).reshape(n, 3)
        self.force = np.zeros((n, 3))
        self.pot = np.zeros(n)
        self.temp = np.zeros(n)
        self.mass = np.repeat(np.asarray(masses, dtype=np.float64), self.counts)
        self.box = np.array(box, dtype=np.float64).reshape(3)
        self._cs = None

    @property
    def n_types(self):
        return len(self.counts)

    def copy(self):
        s = State.__new__(State)
        s.counts = self.counts.copy()
        s.n = self.n
        for k in ("pos", "vel", "force", "pot", "temp", "mass", "box"):
            setattr(s, k, getattr(self, k).copy())
        s._cs = None
        return s

    def _sync_in(self):
        for k in ("pos", "vel", "force", "pot", "temp", "mass"):
            a = getattr(self, k)
            if not (a.flags.c_contiguous and a.dtype == np.float64):
                setattr(self, k, np.ascontiguousarray(a, dtype=np.float64))
        cs = OrcState()
        cs.n_types = self.n_types
        cs.counts = self.counts.ctypes.data_as(C.POINTER(C.c_int64))
        cs.pos, cs.vel, cs.force = _dp(self.pos), _dp(self.vel), _dp(self.force)
        cs.pot, cs.temp, cs.mass = _dp(self.pot), _dp(self.temp), _dp(self.mass)
        for d in range(3):
            cs.box[d] = float(self.box[d])
        self._cs = cs
        return C.byref(cs)

    def _sync_out(self):
        self.box[:] = [self._cs.box[d] for d in range(3)]


def new_lennard_jones(sigma, eps):
    """potential.rs:27-55 -> (sigma, eps, r_cut, u_cut)."""
    p = OrcLJ()
    lib().orc_new_lennard_jones(sigma, eps, C.byref(p))
    return (p.sigma, p.eps, p.r_cut, p.u_cut)


def lj(p, r):
    """potential.rs:57-75 -> (U - u_cut, dU/dr)."""
    u, f = C.c_double(), C.c_double()
    lib().orc_lj_potential_and_force(C.byref(OrcLJ(*p)), r, C.byref(u), C.byref(f))
    return u.value, f.value


def potential_table(n_types, overrides=None):
    """potential.rs:89-155: every unordered pair defaults to Ar LJ(0.3418, 1.712)."""
    default = new_lennard_jones(0.3418, 1.712)
    tab = (OrcLJ * (n_types * n_types))()
    for a in range(n_types):
        for b in range(n_types):
            key = (min(a, b), max(a, b))
            p = (overrides or {}).get(key, default)
            tab[a * n_types + b] = OrcLJ(*p)
    return tab


def update_force(state, table=None, nthreads=None):
    table = table if table is not None else potential_table(state.n_types)
    lib().orc_update_force(state._sync_in(), table, nthreads or default_threads())


def update_force_rows(state, i0, i1, table=None, nthreads=None):
    table = table if table is not None else potential_table(state.n_types)
    lib().orc_update_force_rows(state._sync_in(), table, i0, i1, nthreads or default_threads())


def update_force_cells(state, table=None, nthreads=None):
    table = table if table is not None else potential_table(state.n_types)
    rc = lib().orc_update_force_cells(state._sync_in(), table, nthreads or default_threads())
    if rc != 0:
        raise MemoryError("orc_update_force_cells failed")


def pair_counts(state, table=None, radius=0.0, nthreads=None):
    table = table if table is not None else potential_table(state.n_types)
    out = np.zeros(state.n, dtype=np.int32)
    lib().orc_pair_counts(state._sync_in(), table, radius,
                          out.ctypes.data_as(C.POINTER(C.c_int32)), nthreads or default_threads())
    return out


def apply_boundary_conditions(state):
    lib().orc_apply_boundary_conditions(state._sync_in())


def center_of_mass_velocity(state, t=0):
    out = np.zeros(3)
    lib().orc_center_of_mass_velocity(state._sync_in(), t, _dp(out))
    return out


def momentum(state, t=0):
    out = np.zeros(3)
    lib().orc_momentum(state._sync_in(), t, _dp(out))
    return out


def kinetic_energy(state, t=0):
    return lib().orc_kinetic_energy(state._sync_in(), t)


def thermal_energy(state, t, com):
    com = np.ascontiguousarray(com, dtype=np.float64)
    return lib().orc_thermal_energy(state._sync_in(), t, _dp(com))


def potential_energy(state, t=0):
    return lib().orc_potential_energy(state._sync_in(), t)


def temperature(e, n):
    return lib().orc_temperature(e, n)


def pressure(state, t, com):
    com = np.ascontiguousarray(com, dtype=np.float64)
    return lib().orc_pressure(state._sync_in(), t, _dp(com))


def thermostat_calculate_lambda(th, state, dt, t, target):
    lib().orc_thermostat_calculate_lambda(C.byref(th), state._sync_in(), dt, t, target)


def thermostat_update(th, state, dt, t, target):
    lib().orc_thermostat_update(C.byref(th), state._sync_in(), dt, t, target)


def barostat_calculate_myu(ba, state, dt, t, target):
    lib().orc_barostat_calculate_myu(C.byref(ba), state._sync_in(), dt, t, target)


def barostat_update(ba, state, t):
    lib().orc_barostat_update(C.byref(ba), state._sync_in(), t)
    state._sync_out()


def verlet_step(state, dt, table=None, barostat=None, target_p=0.0, thermostat=None,
                target_t=0.0, nthreads=None, cells=False):
    """integrator.rs:14-64 (one step, in place).  cells: forces through the
    cell grid (bitwise the same), for states too large for all pairs."""
    table = table if table is not None else potential_table(state.n_types)
    args = (state._sync_in(), table, dt, C.byref(barostat) if barostat is not None else None, target_p,
            C.byref(thermostat) if thermostat is not None else None, target_t, nthreads or default_threads())
    if cells:
        if lib().orc_verlet_step_cells(*args) != 0:
            raise MemoryError("orc_verlet_step_cells failed")
    else:
        lib().orc_verlet_step(*args)
    state._sync_out()


def lattice_u(grid, a, start=(0.0, 0.0, 0.0)):
    g = np.ascontiguousarray(grid, dtype=np.int64)
    pos = np.zeros((int(np.prod(g)), 3))
    st = np.ascontiguousarray(start, dtype=np.float64)
    lib().orc_lattice_u(_dp(pos), _dp(st), g.ctypes.data_as(C.POINTER(C.c_int64)), a)
    return pos


def lattice_fcc(grid, a, start=(0.0, 0.0, 0.0)):
    g = np.ascontiguousarray(grid, dtype=np.int64)
    pos = np.zeros((4 * int(np.prod(g)), 3))
    st = np.ascontiguousarray(start, dtype=np.float64)
    lib().orc_lattice_fcc(_dp(pos), _dp(st), g.ctypes.data_as(C.POINTER(C.c_int64)), a)
    return pos


def mb_velocities(n, temperature_k, mass, seed):
    vel = np.zeros((n, 3))
    lib().orc_mb_velocities(_dp(vel), n, temperature_k, mass, seed)
    return vel


def uniform(n, seed):
    out = np.zeros(n)
    lib().orc_uniform(_dp(out), n, seed)
    return out
