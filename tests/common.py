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


def grid_dims(box, rlist, budget):
    """Cells per axis of a grid fill (md_device.h grid_dims, restated)."""
    g = []
    for b in box:
        q = np.floor(b / (rlist * (1.0 + 1e-9)))
        g.append(1 if q < 1.0 else (1024 if q > 1024.0 else int(q)))

    def coarsen(x):
        dec = x // 8
        n = x - (1 if dec < 1 else dec)
        return 1 if n < 1 else n
    while g[0] * g[1] * g[2] > budget:
        if g[1] > g[0] and g[1] >= g[2]:
            g[1] = coarsen(g[1])
        elif g[2] > g[0] and g[2] > g[1]:
            g[2] = coarsen(g[2])
        else:
            g[0] = coarsen(g[0])
    return g


def brick_dims(gd, n_atoms, edge_max=14, cells_max=4096, target=900, own_max=440):
    """Brick edge per axis and the mask of whole axes (md_device.h brick_dims,
    restated): edges (e, e, e), (e, e, e + 1) or (e, e + 1, e + 1) with the
    most centre particles within the staging limits, ties to the smaller halo."""
    per_cell = n_atoms / float(gd[0] * gd[1] * gd[2])
    best = (1, 1, 1)
    best_own, best_halo = -1.0, 0.0
    for e in range(1, edge_max + 1):
        for x in range(3):
            es = (e, e + (1 if x >= 2 else 0), e + (1 if x >= 1 else 0))
            halo = np.prod([float(g if g <= b + 2 else b + 2) for g, b in zip(gd, es)])
            own = np.prod([float(g if g <= b + 2 else b) for g, b in zip(gd, es)])
            if (halo <= cells_max and halo * per_cell <= target and own * per_cell <= own_max
                    and (own > best_own or (own == best_own and halo < best_halo))):
                best_own, best_halo, best = own, halo, es
    edge = [g if g <= b + 2 else b for g, b in zip(gd, best)]
    whole = [g <= b + 2 for g, b in zip(gd, best)]
    return edge, whole


def cell_counts(pos, box, gd):
    """Particles per cell of a grid fill (md_common.h cell_coord, restated)."""
    q = []
    for k in range(3):
        u = pos[:, k] / box[k]
        u = u - np.floor(u)
        q.append(np.clip((u * gd[k]).astype(np.int64), 0, gd[k] - 1))
    counts = np.zeros(gd, dtype=np.int64)
    np.add.at(counts, (q[0], q[1], q[2]), 1)
    return counts


def halo_total(counts, org, hd):
    """Particles staged for a region: its halo of hd cells from grid cell org
    (periodic)."""
    gd = counts.shape
    idx = [np.arange(org[k], org[k] + hd[k]) % gd[k] for k in range(3)]
    return int(counts[np.ix_(idx[0], idx[1], idx[2])].sum())


def brick_regions(counts, edge, whole, capacity=1024):
    """The regions md_kernels.hip brick_build stages, restated for tiled axes
    (no whole axis): {level: number of regions whose halo overflowed}, level
    0 bricks, 1 sub-bricks, 2 cells."""
    gd = counts.shape
    assert not any(whole)
    over = {0: 0, 1: 0, 2: 0}
    nb = [(gd[k] + edge[k] - 1) // edge[k] for k in range(3)]
    for b0 in range(nb[0]):
        for b1 in range(nb[1]):
            for b2 in range(nb[2]):
                bc = (b0, b1, b2)
                borg = [bc[k] * edge[k] - 1 for k in range(3)]
                bcen = [min(gd[k] - bc[k] * edge[k], edge[k]) for k in range(3)]
                tot = halo_total(counts, borg, [c + 2 for c in bcen])
                if tot <= capacity:
                    continue
                over[0] += 1
                per = tot / float(np.prod([c + 2 for c in bcen]))
                m = list(bcen)
                for _ in range(32):
                    hv = float(np.prod([x + 2 for x in m]))
                    if max(m) <= 1 or hv * per <= 0.75 * capacity:
                        break
                    k = int(np.argmax(m))  # the first of the largest
                    m[k] = (m[k] + 1) // 2
                sp = [(bcen[k] + m[k] - 1) // m[k] for k in range(3)]
                for q in np.ndindex(*sp):
                    org = [borg[k] + q[k] * m[k] for k in range(3)]
                    cen = [min(m[k], bcen[k] - q[k] * m[k]) for k in range(3)]
                    if halo_total(counts, org, [c + 2 for c in cen]) <= capacity:
                        continue
                    over[1] += 1
                    for c in np.ndindex(*cen):
                        corg = [org[k] + c[k] for k in range(3)]
                        if halo_total(counts, corg, [3, 3, 3]) > capacity:
                            over[2] += 1
    return over
