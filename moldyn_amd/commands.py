"""The reference CLI's commands around the hot path (cli/src/commands.rs),
on the GPU engine and the GPU frame writer.

initialize   commands.rs:43-80    lattice + Maxwell-Boltzmann state 0, db.csv
solve        commands.rs:82-191   state n loaded, `iteration_count` velocity
                                  Verlet steps, every state saved
solve_macro  commands.rs:204-274  macro parameters of every stored state into
                                  macro.csv; the states are evaluated in
                                  batches, one replica per state, by one
                                  update_force launch per batch
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib as L
from .engine import Engine, barostat as mk_barostat, lattice, mb_velocities, thermostat as mk_thermostat
from .frames import (DataFileMacro, FrameWriter, StateToSave, apply_potentials, load_particles_data,
                     load_potentials_from_file, save_particles_data)

K_B = 1.380648528  # core/src/lib.rs:15


def get_temperature(thermal_energy, number_particles):
    """macro_parameters/temperature.rs:4-7."""
    t = (2.0 * thermal_energy) / (3.0 * float(number_particles) * K_B)
    return t * 100.0


def initialize(path, crystal_cell_type, size, particle_name, particle_mass, particle_radius, lattice_cell,
               temperature, seed=0):
    """commands.rs:43-80: particle 0 in db.csv, a U (simple cubic) or FCC
    lattice from the origin, Maxwell-Boltzmann velocities (seeded: the
    reference's thread_rng is not), saved as state 0."""
    kind = {"u": "u", "fcc": "fcc"}[crystal_cell_type.lower()]
    size = [int(s) for s in size]
    count = size[0] * size[1] * size[2] * (4 if kind == "fcc" else 1)
    box = np.array([lattice_cell * float(size[0]), lattice_cell * float(size[1]), lattice_cell * float(size[2])])
    x = lattice(kind, size, lattice_cell)
    v = mb_velocities(count, temperature, particle_mass, seed)
    StateToSave(np.zeros(count, dtype=np.uint16), x, v, box).save_to_file(path, 0)
    save_particles_data(path, {0: (particle_name, float(particle_mass), float(particle_radius))})


def _engine_for(state: StateToSave, db, force_mode, potentials):
    counts, masses, x, v = state.into_state({i: m for i, (_, m, _) in db.items()})
    e = Engine(counts, masses, state.boundary_box, force_mode=force_mode)
    if potentials:
        apply_potentials(e, potentials)
    e.upload(x, v, state.boundary_box)
    return e


def solve(path, state_number, iteration_count, delta_time, integrator="verlet", use_potentials=False,
          thermostat=None, thermostat_params=None, temperature=None, barostat=None, barostat_params=None,
          pressure=None, force_mode="fp64_list", save_every=1):
    """commands.rs:82-191.  thermostat: None, "berendsen" or "nose_hoover"
    (params [tau]); barostat: None or "berendsen" (params [beta, tau]).
    save_every = 1 saves every state as the reference does; k > 1 saves
    every k-th (and the last) and keeps the state in HBM in between."""
    if integrator != "verlet":
        raise L.MdError(L.MD_ERR_UNIMPLEMENTED, "only the Verlet method exists in the reference")
    path = str(path)
    data = StateToSave.load_from_file(path, state_number)
    db = load_particles_data(path)
    potentials = load_potentials_from_file(path) if use_potentials else None
    e = _engine_for(data, db, force_mode, potentials)
    th = ba = None
    if thermostat is not None:
        if temperature is None:
            raise ValueError("No temperature was passed")
        kind = {"berendsen": "berendsen", "nose_hoover": "nose_hoover"}[thermostat]
        th = mk_thermostat(kind, tau=float(thermostat_params[0]), target_t=float(temperature), lambda_=0.0, psi=0.0)
    if barostat is not None:
        if pressure is None:
            raise ValueError("No pressure was passed")
        ba = mk_barostat("berendsen", beta=float(barostat_params[0]), tau=float(barostat_params[1]),
                         target_p=float(pressure), myu=0.0)
    try:
        e.update_force()
        with FrameWriter(e, path) as w:
            i = 0
            while i < iteration_count:
                w.save(state_number + i)
                k = min(save_every, iteration_count - i)
                e.step(delta_time, k, th, ba)
                i += k
            w.save(state_number + iteration_count)
            stats = w.stats()
    finally:
        e.close()
    return stats


def _last_frame(path):
    """get_last_path (commands.rs:193-202): the largest file number in data/."""
    nums = [int(os.path.splitext(f)[0]) for f in os.listdir(os.path.join(path, "data"))]
    if not nums:
        raise ValueError("no states in data/")
    return max(nums)


def solve_macro(path, kinetic_energy=False, potential_energy=False, thermal_energy=False, temperature=False,
                pressure=False, custom=False, use_potentials=False, force_mode="fp64_list", batch=64):
    """commands.rs:204-274: update_force of every stored state 0..=last, the
    requested macro parameters of type 0 (counts over all particles), and
    macro.csv."""
    if custom:
        raise NotImplementedError("custom macro parameters are todo!() in the reference")
    path = str(path)
    potentials = load_potentials_from_file(path) if use_potentials else None
    end = _last_frame(path)
    db = load_particles_data(path)
    masses = {i: m for i, (_, m, _) in db.items()}
    macro = DataFileMacro()
    frames = list(range(0, end + 1))
    for b0 in range(0, len(frames), batch):
        states = {i: StateToSave.load_from_file(path, i) for i in frames[b0:b0 + batch]}
        # one engine per distinct composition in the batch, one replica per state
        groups = {}
        for i, s in states.items():
            counts, mass, x, v = s.into_state(masses)
            groups.setdefault((tuple(counts), tuple(mass)), []).append((i, x, v, s.boundary_box))
        for (counts, mass), items in groups.items():
            e = Engine([list(counts)] * len(items), list(mass), np.array([it[3] for it in items]),
                       force_mode=force_mode)
            try:
                if potentials:
                    apply_potentials(e, potentials)
                e.upload(np.concatenate([it[1] for it in items]), np.concatenate([it[2] for it in items]),
                         np.array([it[3] for it in items]))
                e.update_force()
                particle_count = int(sum(counts))
                for r, it in enumerate(items):
                    m = e.macro(replica=r, type_id=0)
                    params = []
                    if kinetic_energy:
                        params.append(("kinetic_energy", m["kinetic"]))
                    if potential_energy:
                        params.append(("potential_energy", m["potential"]))
                    thermal = 0.0
                    if thermal_energy or temperature:
                        thermal = m["thermal"]
                        params.append(("thermal_energy", thermal))
                    if temperature:
                        params.append(("temperature", get_temperature(thermal, particle_count)))
                    if pressure:
                        params.append(("pressure", m["pressure"]))
                    macro.add_macro_params(it[0], params, particle_count)
            finally:
                e.close()
    macro.save_to_file(os.path.join(path, "macro.csv"))
    macro.reset_old()
    return macro
