"""Initialisers on the device (md_init_positions, md_init_velocities) against
the oracle's lattices and Maxwell-Boltzmann draws (oracle/md_oracle.c,
position.rs:24-104, velocity.rs:6-29)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.common import AR_MASS

pytestmark = pytest.mark.gpu

K_B = 1.380648528


def test_lattices_on_a_fresh_context():
    import moldyn_amd as M
    e = M.Engine([1000], [AR_MASS], [33.38339] * 3, force_mode="fp64_list")
    e.init_lattice("u", [10, 10, 10], 3.338339)
    d = e.download(("x", "v", "box"))
    assert np.array_equal(d["x"], O.lattice_u([10, 10, 10], 3.338339))
    assert not d["v"].any() and list(d["box"]) == [33.38339] * 3
    e.close()


def test_fcc_into_one_type_keeps_the_rest():
    """UnitCell::FCC for type 1 of a two-type system: its first 4 x 3 x 4 x 5
    particles on the lattice, its last ones and type 0 where they were."""
    import moldyn_amd as M
    rng = np.random.default_rng(1)
    counts = [500, 300]
    box = [2.0, 2.5, 3.0]
    x0 = rng.uniform(0, 2.0, (800, 3))
    v0 = rng.normal(0, 1, (800, 3))
    e = M.Engine(counts, [AR_MASS, 39.948], box, force_mode="fp32_list")
    e.upload(x0, v0, box)
    e.init_lattice("fcc", [3, 4, 5], 0.5, start=(0.1, 0.2, 0.05), type_id=1)
    d = e.download(("x", "v"))
    want = x0.copy()
    want[500:740] = O.lattice_fcc([3, 4, 5], 0.5, start=(0.1, 0.2, 0.05))
    assert np.array_equal(d["x"], want) and np.array_equal(d["v"], v0)
    with pytest.raises(M.MdError, match="TooBig"):
        e.init_lattice("fcc", [4, 4, 5], 0.5, type_id=1)
    with pytest.raises(M.MdError, match="OutOfBoundary"):
        e.init_lattice("u", [5, 1, 1], 0.5, type_id=1)
    e.close()


@pytest.mark.parametrize("n", [1000, 1001])
def test_maxwell_boltzmann_draws(n):
    """The seeded draws of md_init_mb_velocities, one thread per pair: equal
    up to the last bits of the device's log, sin and cos (|dv| <= 1e-13
    sigma), pairs negated bit for bit, an odd last particle left at 0."""
    import moldyn_amd as M
    T, seed = 273.15, 17
    e = M.Engine([n], [AR_MASS], [40.0] * 3, force_mode="fp64_list")
    e.init_lattice("u", [10, 10, 10], 3.338339)
    e.init_velocities(T, seed)
    v = e.download(("v",))["v"]
    e.close()
    ref = O.mb_velocities(n, T, AR_MASS, seed)
    sigma = np.sqrt(K_B * (T * 0.01) / AR_MASS)
    assert np.abs(v - ref).max() <= 1e-13 * sigma
    h = n // 2
    assert np.array_equal(v[h:2 * h], -v[:h])
    if n % 2:
        assert not v[-1].any()
    assert np.abs(v.sum(0)).max() <= 1e-12 * sigma


def test_device_initialised_state_steps_like_an_upload():
    """Replica 1 of a batch set up on the device steps exactly as the same
    state uploaded from the host."""
    import moldyn_amd as M
    counts = [[512], [729]]
    boxes = [[8 * 0.9] * 3, [9 * 0.9] * 3]
    e = M.Engine(counts, [AR_MASS], boxes, force_mode="fp64_list")
    e.init_lattice("u", [8, 8, 8], 0.9, replica=0)
    e.init_lattice("u", [9, 9, 9], 0.9, replica=1)
    e.init_velocities(150.0, 3, replica=0)
    e.init_velocities(300.0, 4, replica=1)
    st = e.download(("x", "v", "box"))
    f = M.Engine(counts, [AR_MASS], boxes, force_mode="fp64_list")
    f.upload(st["x"], st["v"], st["box"])
    import moldyn_amd._lib as L
    outs = []
    for g in (e, f):
        th = (L.MdThermostat * 2)(*[M.thermostat("berendsen", tau=1.0, target_t=200.0) for _ in range(2)])
        g.update_force()
        g.step(0.002, 50, th)
        outs.append(g.download(("x", "v", "box")))
        g.close()
    assert np.array_equal(outs[0]["x"], outs[1]["x"]) and np.array_equal(outs[0]["v"], outs[1]["v"])
