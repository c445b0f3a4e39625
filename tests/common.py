"""Shared inputs for the parity tests: the benchmark configurations and dense
fixtures, built with the oracle's seeded initialisers (test side only)."""
import numpy as np

from oracle import oracle as O

AR_MASS = 66.335
A_LATTICE = 3.338339


def sc_state(n_side, a=A_LATTICE, temperature=273.15, seed=1, jitter=0.0, jitter_seed=2):
    """Simple-cubic argon: position.rs:34-51 grid, seeded MB velocities."""
    n = n_side ** 3
    pos = O.lattice_u([n_side] * 3, a)
    if jitter:
        u = O.uniform(3 * n, jitter_seed).reshape(n, 3)
        pos = pos + (2.0 * u - 1.0) * jitter * a
    vel = O.mb_velocities(n, temperature, AR_MASS, seed)
    s = O.State([n], [AR_MASS], [n_side * a] * 3, pos=pos, vel=vel)
    O.apply_boundary_conditions(s)
    return s


def liquid_state(n_cells=4, a=0.5256, jitter=0.02, seed=5, temperature=120.0):
    """FCC liquid argon, rho sigma^3 ~ 0.8 (dense force-parity fixture)."""
    pos = O.lattice_fcc([n_cells] * 3, a)
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter, jitter, pos.shape)
    n = len(pos)
    vel = O.mb_velocities(n, temperature, AR_MASS, seed)
    s = O.State([n], [AR_MASS], [n_cells * a] * 3, pos=pos, vel=vel)
    O.apply_boundary_conditions(s)
    return s


def gas_state(n, box, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, box, size=(n, 3))
    vel = O.mb_velocities(n, 273.15, AR_MASS, seed)
    return O.State([n], [AR_MASS], [box] * 3, pos=pos, vel=vel)


def engine_for(state, force_mode, masses=None, **kw):
    import moldyn_amd as M
    masses = masses if masses is not None else [state.mass[int(np.sum(state.counts[:t]))]
                                                for t in range(state.n_types)]
    e = M.Engine(state.counts, masses, state.box, force_mode=force_mode, **kw)
    e.upload(state.pos, state.vel, state.box)
    return e


def rms(a):
    return float(np.sqrt(np.mean(np.square(a))))
