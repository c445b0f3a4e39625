"""Device-resident MD system: a thin Python owner of one md_ctx.

Arrays cross the boundary as N x 3 float64 numpy arrays in the reference's
particle order (type-major; replica-major first for batches).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

FORCE_MODES = {"fp64_allpairs": L.MD_FORCE_FP64_ALLPAIRS, "fp64_list": L.MD_FORCE_FP64_LIST,
               "fp32_list": L.MD_FORCE_FP32_LIST, "mlp": L.MD_FORCE_MLP}
KERNELS = {"step": 0, "finalize": 1, "allpairs": 2, "drift": 3, "mlp_net": 5, "mlp_feat": 6}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Engine:
    """One system, or a batch of independent replicas, on the GPU.

    counts: per type (single system) or n_replicas x n_types (batch).
    box: 3 values, or n_replicas x 3.
    """

    def __init__(self, counts, masses, box, force_mode="fp32_list", skin=None, block_size=None,
                 cells_per_atom=2.0, graph_chunk=32, list_capacity=0, cell_capacity=0,
                 tiles_per_block=0):
        self._lib = L.load()
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        if counts.ndim == 1:
            counts = counts[None, :]
        self.counts = counts
        self.n_replicas, self.n_types = counts.shape
        self.n = int(counts.sum())
        self.masses = _f64(masses)
        if len(self.masses) != self.n_types:
            raise ValueError("one mass per type")
        boxes = _f64(box, (self.n_replicas, 3))
        h = C.c_void_p()
        rc = self._lib.md_create_replicas(C.byref(h), self.n_replicas, self.n_types,
                                          counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                          _dp(self.masses), _dp(boxes))
        if rc != 0:
            raise L.MdError(rc, self._lib.md_global_error().decode())
        self._h = h
        opt = L.MdOptions()
        self._chk(self._lib.md_get_options(self._h, C.byref(opt)))
        opt.force_mode = FORCE_MODES[force_mode]
        opt.skin = float(skin) if skin is not None else 0.0
        opt.block_size = int(block_size) if block_size is not None else 0  # 0: the library's choice
        opt.cells_per_atom = float(cells_per_atom)
        opt.graph_chunk = int(graph_chunk)
        opt.list_capacity = int(list_capacity)
        opt.cell_capacity = int(cell_capacity)
        opt.tiles_per_block = int(tiles_per_block)
        self._chk(self._lib.md_set_options(self._h, C.byref(opt)))
        self.force_mode = force_mode

    # -- plumbing
    def _chk(self, rc):
        if rc != 0:
            raise L.MdError(rc, self._lib.md_last_error(self._h).decode())
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._lib.md_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- configuration
    def options(self):
        opt = L.MdOptions()
        self._chk(self._lib.md_get_options(self._h, C.byref(opt)))
        return {k: getattr(opt, k) for k, _ in opt._fields_}

    def set_pair_potential(self, t0, t1, sigma, eps, r_cut=None, u_cut=None):
        if r_cut is None:
            self._chk(self._lib.md_set_pair_potential(self._h, t0, t1, sigma, eps))
        else:
            self._chk(self._lib.md_set_pair_potential_lj(self._h, t0, t1, sigma, eps, r_cut, u_cut))

    # -- state
    def upload(self, x, v, box):
        x = _f64(x, (self.n, 3))
        v = _f64(v, (self.n, 3))
        b = _f64(box, (self.n_replicas, 3))
        self._chk(self._lib.md_upload_state(self._h, _dp(x), _dp(v), _dp(b)))

    def init_lattice(self, kind, grid, a, start=(0.0, 0.0, 0.0), replica=0, type_id=0):
        """UnitCell::initialize_particles_position on the device (md_init_positions)."""
        st = _f64(start, (3,))
        g = np.ascontiguousarray(grid, dtype=np.int64)
        k = {"u": 0, "fcc": 1}[kind.lower()]
        self._chk(self._lib.md_init_positions(self._h, int(replica), int(type_id), k, _dp(st),
                                              g.ctypes.data_as(C.POINTER(C.c_int64)), float(a)))

    def init_velocities(self, temperature, seed, replica=0, type_id=0):
        """Maxwell-Boltzmann velocities on the device (md_init_velocities)."""
        self._chk(self._lib.md_init_velocities(self._h, int(replica), int(type_id), float(temperature), int(seed)))

    def upload_forces(self, force=None, pot=None, virial=None):
        f = _f64(force, (self.n, 3)) if force is not None else None
        p = _f64(pot, (self.n,)) if pot is not None else None
        w = _f64(virial, (self.n,)) if virial is not None else None
        self._chk(self._lib.md_upload_forces(self._h, _dp(f) if f is not None else None,
                                             _dp(p) if p is not None else None,
                                             _dp(w) if w is not None else None))

    def download(self, what=("x", "v", "force", "pot", "virial", "box")):
        out = {}
        bufs = {}
        for k in ("x", "v", "force"):
            bufs[k] = np.zeros((self.n, 3)) if k in what else None
        for k in ("pot", "virial"):
            bufs[k] = np.zeros(self.n) if k in what else None
        bufs["box"] = np.zeros((self.n_replicas, 3))
        ptr = {k: (_dp(a) if a is not None else None) for k, a in bufs.items()}
        self._chk(self._lib.md_download_state(self._h, ptr["x"], ptr["v"], ptr["force"], ptr["pot"],
                                              ptr["virial"], ptr["box"]))
        for k in what:
            out[k] = bufs[k]
        if "box" in out and self.n_replicas == 1:
            out["box"] = out["box"][0]
        return out

    # -- hot path
    def update_force(self):
        self._chk(self._lib.md_update_force(self._h))

    def apply_boundary_conditions(self):
        self._chk(self._lib.md_apply_boundary_conditions(self._h))

    def step(self, dt, n_steps=1, thermostat=None, barostat=None):
        """n_steps velocity-Verlet steps.  thermostat / barostat: MdThermostat /
        MdBarostat (one, or a ctypes array of n_replicas for a batch); updated in place."""
        th = C.byref(thermostat) if isinstance(thermostat, L.MdThermostat) else thermostat
        ba = C.byref(barostat) if isinstance(barostat, L.MdBarostat) else barostat
        fn = self._lib.md_step if self.n_replicas == 1 else self._lib.md_step_replicas
        if self.n_replicas > 1 and isinstance(thermostat, L.MdThermostat):
            raise ValueError("a batch needs one thermostat per replica (ctypes array)")
        if self.n_replicas > 1 and isinstance(barostat, L.MdBarostat):
            raise ValueError("a batch needs one barostat per replica (ctypes array)")
        self._chk(fn(self._h, float(dt), th, ba, int(n_steps)))

    def macro(self, replica=0, type_id=0):
        out = np.zeros(11)
        self._chk(self._lib.md_macro(self._h, replica, type_id, _dp(out)))
        return {"kinetic": out[0], "potential": out[1], "thermal": out[2], "temperature": out[3],
                "pressure": out[4], "momentum": out[5:8].copy(), "com_velocity": out[8:11].copy()}

    # -- learned surrogate (force_mode "mlp")
    def set_mlp(self, theta, g_mean, g_inv_std, g_lo=None, g_hi=None, e_scale=1.0, e_shift=0.0,
                judge_every=100):
        """Network parameters (flat, W1|b1|W2|b2|w3|b3), input normalisation,
        training range of the features (None: no fallback) and the LJ check interval."""
        def f32(a, n):
            a = np.ascontiguousarray(a, dtype=np.float32).ravel()
            if a.size != n:
                raise ValueError(f"expected {n} values, got {a.size}")
            return a
        fp = C.POINTER(C.c_float)
        th = f32(theta, L.MD_MLP_NUM_PARAMS)
        gm, gs = f32(g_mean, 32), f32(g_inv_std, 32)
        lo = f32(g_lo, 32) if g_lo is not None else None
        hi = f32(g_hi, 32) if g_hi is not None else None
        self._chk(self._lib.md_mlp_set_params(
            self._h, th.ctypes.data_as(fp), gm.ctypes.data_as(fp), gs.ctypes.data_as(fp),
            lo.ctypes.data_as(fp) if lo is not None else None,
            hi.ctypes.data_as(fp) if hi is not None else None,
            float(e_scale), float(e_shift), int(judge_every)))

    def mlp_stats(self, replica=0):
        out = np.zeros(5)
        self._chk(self._lib.md_mlp_get_stats(self._h, replica, _dp(out)))
        return {"last_rmse_rel": out[0], "max_rmse_rel": out[1], "checks": int(out[2]),
                "fallbacks": int(out[3]), "last_rms_force": out[4]}

    # -- observability
    def stats(self):
        s = L.MdStats()
        self._chk(self._lib.md_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def set_profiling(self, on):
        self._chk(self._lib.md_set_profiling(self._h, 1 if on else 0))

    def kernel_time(self, kernel, span=False):
        """(total ms, launches) of a kernel while profiling: stream time (previous
        kernel end to end), or with span=True first block start to last block end."""
        ms, cnt = C.c_double(), C.c_int64()
        sms, scnt = C.c_double(), C.c_int64()
        self._chk(self._lib.md_kernel_time(self._h, KERNELS[kernel], C.byref(ms), C.byref(cnt),
                                           C.byref(sms), C.byref(scnt)))
        return (sms.value, scnt.value) if span else (ms.value, cnt.value)


def thermostat(kind, tau=0.0, target_t=0.0, lambda_=0.0, psi=0.0):
    kinds = {"none": 0, "berendsen": 1, "nose_hoover": 2, "custom": 3}
    return L.MdThermostat(kinds[kind], tau, target_t, lambda_, psi)


def barostat(kind, beta=0.0, tau=0.0, target_p=0.0, myu=0.0):
    kinds = {"none": 0, "berendsen": 1, "custom": 2}
    return L.MdBarostat(kinds[kind], beta, tau, target_p, myu)


def lattice(kind, grid, a, start=(0.0, 0.0, 0.0)):
    """initializer/position.rs:34-99 (kind 'u' or 'fcc')."""
    lib = L.load()
    g = np.ascontiguousarray(grid, dtype=np.int64)
    per = 1 if kind == "u" else 4
    pos = np.zeros((per * int(np.prod(g)), 3))
    st = _f64(start)
    rc = lib.md_init_lattice(0 if kind == "u" else 1, _dp(st), g.ctypes.data_as(C.POINTER(C.c_int64)),
                             float(a), _dp(pos))
    if rc != 0:
        raise L.MdError(rc, lib.md_global_error().decode())
    return pos


def mb_velocities(n, temperature, mass, seed):
    """Seeded Maxwell-Boltzmann with pair negation (initializer/velocity.rs:6-29 structure)."""
    lib = L.load()
    v = np.zeros((n, 3))
    rc = lib.md_init_mb_velocities(_dp(v), int(n), float(temperature), float(mass), int(seed))
    if rc != 0:
        raise L.MdError(rc, lib.md_global_error().decode())
    return v


def lj_potential_and_force(sigma, eps, r_cut, u_cut, r):
    lib = L.load()
    u, f = C.c_double(), C.c_double()
    lib.md_lj_potential_and_force(sigma, eps, r_cut, u_cut, r, C.byref(u), C.byref(f))
    return u.value, f.value


def new_lennard_jones(sigma, eps):
    lib = L.load()
    rc, uc = C.c_double(), C.c_double()
    lib.md_new_lennard_jones(sigma, eps, C.byref(rc), C.byref(uc))
    return rc.value, uc.value
