"""Algorithmic HBM bytes per phase of one step (minimum compulsory traffic,
one pass per phase) and the MI355X peaks they are priced against.

R is the byte size of a particle-state real (4 for FP32 state, 8 for FP64).
Per simulator record: x, y, z, vx, vy, vz, E_rot (7 R) + cell + species
(2 x 4 B) = 36 B (FP32) or 64 B (FP64).
"""
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s (1 GB = 1e9 B)
N_BASE_ROWS = 8                # sampler rows 0-7 (N, M, M v, M |v|^2, E_rot, rotDoF)
N_MOM_ROWS = 28                # per-cell sums kept for the BGK family
BGK_PARAMS = 184               # per-cell BGK parameter record


def record_bytes(R):
    return 7 * R + 8


def sampled_rows(sampler_full=False, n_species=1):
    """FP64 sampler rows accumulated per cell and sampled step."""
    return (20 if sampler_full else N_BASE_ROWS) + 2 * n_species


def phase_bytes(phase, R, n, n_cells, cand=0, coll=0, n_bgk=0, rows=N_BASE_ROWS + 2, bgk=False, hybrid=False):
    """Algorithmic bytes of one launch of `phase` over n simulators.

    cand, coll: NTC candidates and accepted collisions of the step;
    n_bgk: simulators redrawn by the BGK family; rows: sampler rows;
    bgk / hybrid: the case keeps BGK moments / hybrid accumulators."""
    if phase == "move":        # read x, v, cell; write x, cell
        return n * (3 * R + 3 * R + 4 + 3 * R + 4)
    if phase == "scatter":     # read cell and rank, write slot
        return n * 12
    if phase == "scan":        # count: read cell, write rank; scan: counts read twice, offsets written
        return n * 8 + n_cells * 12
    if phase == "cells":       # slot list, gather of the record, sorted record written;
        #                        per cell: offsets, tile table, sampler r/w, hybrid r/w, BGK moments w
        per_cell = 8 + 4 + 2 * 8 * rows + (2 * 8 * 8 if hybrid else 0) + (8 * N_MOM_ROWS if bgk else 0)
        return n * (4 + 2 * record_bytes(R)) + n_cells * per_cell
    if phase == "collide":     # NTC: sigma_max r/w, volume, offsets; pair reads, collided pair writes.
        #                        BGK: moments read, parameters w + r, USP state r/w, selection list
        #                        (one slot per redrawn simulator, w + r), redrawn velocities,
        #                        conservation (read v + species, write v)
        b = n_cells * (16 + 8 + 8) + cand * 2 * (3 * R + 4) + coll * 2 * (4 * R)
        if bgk:
            b += n_cells * (8 * N_MOM_ROWS + 2 * BGK_PARAMS + 2 * 8 * 10 + 8)
            b += n_bgk * (4 + 4 + 3 * R + 4) + n * ((3 * R + 4) + 3 * R)
        return b
    return 0


def step_bytes(R, n, n_cells, cand=0, coll=0, n_bgk=0, rows=N_BASE_ROWS + 2, bgk=False, hybrid=False):
    return sum(phase_bytes(p, R, n, n_cells, cand, coll, n_bgk, rows, bgk, hybrid)
               for p in ("move", "scatter", "scan", "cells", "collide"))
