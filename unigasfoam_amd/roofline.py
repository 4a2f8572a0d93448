"""Algorithmic HBM bytes per phase of one step (minimum compulsory traffic,
one pass per phase) and the MI355X peaks they are priced against.

R is the byte size of a particle-state real (4 for FP32 state, 8 for FP64).
Per simulator record: x, y, z, vx, vy, vz, E_rot (7 R) + cell + species
(2 x 4 B) = 36 B (FP32) or 64 B (FP64).
"""
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s (1 GB = 1e9 B)
N_SAMPLER_ROWS = 36            # FP64 sampler accumulators per cell (ugx_sample_index)


def record_bytes(R):
    return 7 * R + 8


def phase_bytes(phase, R, n, n_cells, cand=0, coll=0, n_bgk=0):
    """Algorithmic bytes of one launch of `phase` over n simulators."""
    if phase == "move":        # read x, v, cell; write x, cell
        return n * (3 * R + 3 * R + 4 + 3 * R + 4)
    if phase == "scatter":     # read cell, write slot
        return n * 8
    if phase == "segsort":     # read slot + cell of slot, write slot
        return n * 12
    if phase == "gather":      # read slot, read record, write record
        return n * (4 + 2 * record_bytes(R))
    if phase == "sample":      # read v, E_rot, species; sampler r/w + offsets per cell
        return n * (4 * R + 4) + n_cells * (8 + 2 * 8 * N_SAMPLER_ROWS)
    if phase == "ntc":         # offsets, sigma_max r/w, volume per cell; pairs read, collided pairs written
        return n_cells * 32 + cand * 2 * (3 * R + 4) + coll * 2 * (4 * R)
    if phase == "bgk":         # two reads of v, E_rot, species and two writes of v over BGK cells
        return n_bgk * (2 * (4 * R + 4) + 2 * 3 * R)
    if phase == "scan":        # counts read, offsets written (two passes over counts)
        return n_cells * 12
    if phase == "cells":       # slot list, gather of the record, sorted record written; per-cell sampler,
        #                        sigma_max, offsets; collided pairs touch LDS only
        return n * (4 + 2 * record_bytes(R)) + n_cells * (8 + 2 * 8 * N_SAMPLER_ROWS + 16 + 8)
    return 0


def step_bytes(R, n, n_cells, cand=0, coll=0, n_bgk=0):
    return sum(phase_bytes(p, R, n, n_cells, cand, coll, n_bgk) for p in ("move", "scatter", "cells", "scan"))
