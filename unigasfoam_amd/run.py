"""Run a uniGasFoam case directory on the GPU: the time loop of
applications/solvers/discreteMethods/uniGasFoam/uniGasFoam.C:70-110 around
the library's evolve().

    python -m unigasfoam_amd.run CASE [--precision fp32|fp64] [--end-time T]

Reads the case (unigasfoam_amd.case_dir), applies its uniGasMeshFill
configurations, then advances deltaT steps to endTime.  At every
writeInterval the uniGasVolFields of the sampler are written to
CASE/<time>/ (foam_io.write_fields, with the wall measurements on wall
patches) and, while resetAtOutput holds, the
sampler is reset (uniGasVolFields.C:1432-1502).  Every nTerminalOutputs steps
it prints the reference's log lines: the time, the number of particles
(uniGasCloud.C:898) and ExecutionTime / ClockTime (uniGasFoam.C:97-101).
"""
from __future__ import annotations

import argparse
import math
import sys
import time

from . import case_dir, foam_io
from .core import Case


def time_name(t):
    """OpenFOAM's general time format at six significant digits."""
    return f"{t:.6g}"


def run_case(path, precision="fp64", end_time=None, seed=12345, out=sys.stdout, write=True, use_graphs=False):
    spec, fills, run = case_dir.read_case(path, seed=seed)
    case = Case(spec, precision=precision, use_graphs=use_graphs)
    for f in fills:
        case.fill(**f)
    dt = run.delta_t
    t_end = run.end_time if end_time is None else end_time
    n_steps = int(math.floor((t_end - run.start_time) / dt + 0.5))
    w_steps = int(math.floor(run.write_interval / dt + 0.5)) if run.write_interval > 0 else 0
    out_steps = max(1, run.n_terminal_outputs)
    t0_wall = time.perf_counter()
    t0_cpu = time.process_time()
    step = 0
    written = []
    while step < n_steps:
        nxt = n_steps
        if w_steps:
            nxt = min(nxt, (step // w_steps + 1) * w_steps)
        nxt = min(nxt, (step // out_steps + 1) * out_steps)
        case.step(nxt - step)
        step = nxt
        t = run.start_time + step * dt
        if step % out_steps == 0 or step == n_steps:
            case.sync()
            print(f"Time = {time_name(t)}", file=out)
            print(f"    Number of particles = {case.count()}", file=out)
            print(f"ExecutionTime = {time.process_time() - t0_cpu:.2f} s  ClockTime = "
                  f"{time.perf_counter() - t0_wall:.0f} s", file=out, flush=True)
        if write and w_steps and step % w_steps == 0:
            fields = case.fields()
            patch, face = case.wall_faces()
            wall = {"patch": patch, "face": face, "fields": case.wall_fields()}
            foam_io.write_fields(path, time_name(t), spec.mesh, fields, run.field_name, wall=wall)
            written.append(time_name(t))
            if run.reset_at_output and t < run.reset_at_output_until_time:
                case.reset_samples()
    case.sync()
    return case, written


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("case")
    ap.add_argument("--precision", default="fp64", choices=["fp32", "fp64"])
    ap.add_argument("--end-time", type=float, default=None)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--graphs", action="store_true", help="replay each step as a captured HIP graph")
    a = ap.parse_args(argv)
    try:
        case, written = run_case(a.case, a.precision, a.end_time, a.seed, use_graphs=a.graphs)
    except case_dir.Unsupported as e:
        print(str(e), file=sys.stderr)
        return e.code
    except case_dir.CaseError as e:
        print(str(e), file=sys.stderr)
        return e.code
    print(f"End: {case.counters()['steps']} steps, written {' '.join(written) or 'nothing'}")
    case.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
