"""Benchmark of the FDTD time-stepping loop on one MI355X.

Workload (the configuration the headline metric is quoted on): 256^3 cells,
dx = 5 nm, 12-cell CFS-PML, oblique Gaussian-beam TF/SF (theta 30, phi 45),
4x4 Drude-Lorentz spheres, reflected and transmitted plane ports with a
64-frequency FP64 running DFT (stride 8).  A step is one full time step of
that grid.  FP32 fields with FP64 accumulators is the headline; the plain FP64
form runs beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  Per-kernel durations come from HIP events recorded on
the engine's stream around each launch (cfd_kernel_times).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(eng_case_model, S, nd_vars):
    """Algorithmic HBM bytes per launch of the fused H and E kernels: each
    updated cell reads its own 6 field values and ID byte and writes 3 field
    values (neighbour values are reuse, not traffic); CPML slab cells read and
    write 2 psi values per slab they lie in; Drude-Lorentz cells read their
    aux slot (4 B) and read/write 3 x nd_vars auxiliaries."""
    m = eng_case_model
    Nx, Ny, Nz = m["N"]
    n = m["npml"]
    cells_h = Nx * Ny * Nz
    cells_e = (Nx - 1) * (Ny - 1) * (Nz - 1)
    per_cell = 9 * S + 1
    slab_h = (n[0][0] + n[0][1]) * Ny * Nz + Nx * (n[1][0] + n[1][1]) * Nz + Nx * Ny * (n[2][0] + n[2][1])
    slab_e = (n[0][0] + n[0][1]) * (Ny - 1) * (Nz - 1) + (Nx - 1) * (n[1][0] + n[1][1]) * (Nz - 1) + \
        (Nx - 1) * (Ny - 1) * (n[2][0] + n[2][1])
    psi = 2 * 2 * S
    bh = cells_h * per_cell + slab_h * psi
    be = cells_e * per_cell + slab_e * psi + m["ndl"] * (4 + 3 * nd_vars * 2 * S)
    return {"update_h": bh, "update_e": be, "cells_h": cells_h, "cells_e": cells_e, "slab_h": slab_h, "slab_e": slab_e}


def run_engine(case, precision, steps, warmup, timing_steps):
    import cpsfdtd_amd as eng
    e = eng.Engine(case, precision)
    e.step(warmup)
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)
    e.sync()
    wall = time.perf_counter() - t0
    e.set_timing(True)
    e.step(timing_steps)
    e.sync()
    kt = e.kernel_times()
    e.set_timing(False)
    nbytes = e.device_bytes()
    e.close()
    return wall, kt, nbytes


def cpu_baseline(scale_steps=20):
    """The reference's own OpenMP loop (oracle/_ref/refdrv, compiled from the
    reference sources) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from cpsfdtd_amd import configs
    threads = len(os.sched_getaffinity(0))
    if pyoracle.ref_available():
        d = tempfile.mkdtemp(prefix="bench_ref_")
        case = configs.c2_reference_core(d, nt=scale_steps + 2)
        r = pyoracle.ref_bench(case, scale_steps, threads=threads)
        return {"value": round(r["mcells_per_s"], 3), "unit": "Mcell-updates/s", "cores": r["threads"],
                "kind": "reference",
                "sample": "%d steps of the same 256^3 grid (CPML + TF/SF; the reference has no 3-D Drude-Lorentz, its "
                          "ID-2 cells are skipped), FP64, reference sources at -O2 -fopenmp, %.1f s" %
                          (scale_steps, r["seconds"])}
    d = tempfile.mkdtemp(prefix="bench_port_")
    case = configs.c2(d, nt=4, nfreq=0)
    case.ports = []
    o = pyoracle.OracleSim(case)
    t0 = time.perf_counter()
    o.step(3)
    sec = time.perf_counter() - t0
    return {"value": round(256 ** 3 * 3 / sec / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": "3 steps of the same workload on the FP64 oracle, %.1f s" % sec}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--timing-steps", type=int, default=40)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported (multi-GPU sharding is a design "
              "note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import cpsfdtd_amd as eng
    from cpsfdtd_amd import configs
    info = eng.device_info()
    log("bench: device", info)
    nt = args.warmup + args.steps + args.timing_steps
    d = tempfile.mkdtemp(prefix="bench_c2_")
    case = configs.c2(d, scale=args.scale, nt=nt)
    model = eng.Model(case)
    N = model.N
    cells = N[0] * N[1] * N[2]
    npml = [[len(model.table("pml_Hb%s%d" % (ax, e + 1))) for e in range(2)] for ax in "xyz"]
    ndl = len(model.cells(False))
    nd_vars = 1 + 2 * 2 + 1  # 1 Drude + 2 Lorentz poles: JD, JL, JL1, E1
    geo = {"N": N, "npml": npml, "ndl": ndl}
    del model

    res = {}
    for prec in ([32] if args.no_fp64 else [32, 64]):
        S = prec // 8
        log("bench: FP%d run, %d warmup + %d timed steps" % (prec, args.warmup, args.steps))
        wall, kt, nbytes = run_engine(case, prec, args.steps, args.warmup, args.timing_steps)
        ab = algorithmic_bytes(geo, S, nd_vars)
        kstat = {}
        for name, (ms, cnt) in kt.items():
            kstat[name] = {"avg_ms": ms / cnt, "launches": cnt}
        dom = max(("update_h", "update_e"), key=lambda k: kt.get(k, (0, 1))[0])
        avg_s = kstat[dom]["avg_ms"] / 1e3
        step_ms = sum(v["avg_ms"] * v["launches"] for v in kstat.values()) / args.timing_steps
        res[prec] = {
            "wall": wall, "mcells": cells * args.steps / wall / 1e6, "ms_per_step": wall / args.steps * 1e3,
            "kernels": kstat, "dominant": dom, "dom_gbs": ab[dom] / avg_s / 1e9, "dom_bytes": ab[dom],
            "step_gbs": (ab["update_h"] + ab["update_e"]) / (wall / args.steps) / 1e9, "device_bytes": nbytes,
            "kernel_ms_per_step": step_ms,
        }
        log("bench: FP%d %.1f Mcell-upd/s, %.3f ms/step, %s %.1f GB/s" % (
            prec, res[prec]["mcells"], res[prec]["ms_per_step"], dom, res[prec]["dom_gbs"]))

    r32 = res[32]
    out = {
        "metric": "Mcell-updates/s & HBM GB/s vs 8 TB/s; p50/p99 solve latency, FP32 vs FP64",
        "value": round(r32["mcells"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (generated 256^3 geometry in the reference's input formats)",
        "config": {"workload": "c2: 256^3 Drude-Lorentz spheres + 12-cell CFS-PML + oblique Gaussian-beam TF/SF, "
                               "2 plane ports x 64-frequency FP64 DFT (stride 8), FP32 fields",
                   "cells": cells, "steps_in_run": nt, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(r32["dom_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r32["dom_gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": r32["dominant"], "bytes_per_launch": r32["dom_bytes"],
                     "avg_ms": round(r32["kernels"][r32["dominant"]]["avg_ms"], 5)},
        "kernels_f32": {k: round(v["avg_ms"], 5) for k, v in r32["kernels"].items()},
        "step_gbs_f32": round(r32["step_gbs"], 1),
        "device": info["name"],
    }
    if 64 in res:
        r64 = res[64]
        out["fp64"] = {"value": round(r64["mcells"], 2), "ms_per_step": round(r64["ms_per_step"], 5),
                       "dominant_gbs": round(r64["dom_gbs"], 1),
                       "kernels": {k: round(v["avg_ms"], 5) for k, v in r64["kernels"].items()}}
        out["fp32_over_fp64"] = round(r32["mcells"] / r64["mcells"], 3)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as ex:  # reported, never fatal for the GPU figure
            out["cpu_baseline"] = {"error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
