"""Benchmark of the FDTD time-stepping loop on one MI355X.

Workload (the configuration the headline metric is quoted on): 256^3 cells,
dx = 5 nm, 12-cell CFS-PML, oblique Gaussian-beam TF/SF (theta 30, phi 45),
4x4 Drude-Lorentz spheres, reflected and transmitted plane ports with a
64-frequency FP64 running DFT (stride 8).  A step is one full time step of
that grid.  FP32 fields with FP64 accumulators is the headline; the plain FP64
form runs beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  Warmup, timed and per-kernel timing windows are each
rounded up to whole 128-step blocks (at least one): the engine replays the
step sequence from a captured HIP graph per 128-step block, and each block
holds 16 sampled DFT steps and one DFT flush, so every timed window runs the
workload's whole step, DFT included.  The line's "steps" is the number of
steps actually timed ("steps_requested" is K).  Per-kernel durations come
from HIP events recorded on the engine's stream around each launch
(cfd_kernel_times) over a separate 128-step window.  The FP32-vs-FP64 quality
figures (relative L2 of the port |E|^2 series, DFTs and flux spectra,
spectral peak and Q) come from two whole 4000-step runs, whatever K is.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BLOCK = 128  # steps per replayed graph block of the 256^3 workload (DFT stride 8 x 16 staged samples)


def whole_blocks(n):
    """n rounded up to whole graph blocks, at least one."""
    return max(BLOCK, -(-int(n) // BLOCK) * BLOCK)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pass_geometry(N, npml, periodic=0):
    """Cell classes of the fused step (same rule as the engine's Sim::classify):
    per half-step the lean interior box and the CPML shell boxes."""
    out = {}
    for ef in (False, True):
        lo, a, b = [0] * 3, [0] * 3, [0] * 3
        for ax in range(3):
            n1, n2 = npml[ax]
            lo[ax] = ((0 if periodic & (1 << ax) else 1) if ax < 2 else 1) if ef else 0
            a[ax] = (n1 + 1 if ef else n1) if n1 else lo[ax]
            b[ax] = N[ax] - n2
        lean = (a[0], b[0], a[1], b[1], a[2], b[2])
        total = (N[0] - lo[0]) * (N[1] - lo[1]) * (N[2] - lo[2])
        nlean = max(0, b[0] - a[0]) * max(0, b[1] - a[1]) * max(0, b[2] - a[2])
        slab = 0  # slab memberships (a corner cell counts once per slab)
        for ax in range(3):
            n1, n2 = npml[ax]
            plane = 1
            for o in range(3):
                if o != ax:
                    plane *= N[o] - lo[o]
            slab += (n1 + n2) * plane
        out["e" if ef else "h"] = {"lean_box": lean, "lean": nlean, "shell": total - nlean, "slab": slab}
    return out


def workload_geometry(model, periodic=0):
    """pass_geometry of a workload's host model, plus the special-pass cells
    (Drude-Lorentz and gain) inside and outside the E interior box."""
    N = model.N
    npml = [[len(model.table("pml_Hb%s%d" % (ax, e + 1))) for e in range(2)] for ax in "xyz"]
    geo = pass_geometry(N, npml, periodic)
    lb = geo["e"]["lean_box"]
    inside = total = 0
    for gain in (False, True):
        f = np.asarray(model.cells(gain), np.int64)
        if len(f) == 0:
            continue
        i, j, k = f // ((N[1] + 1) * (N[2] + 1)), (f // (N[2] + 1)) % (N[1] + 1), f % (N[2] + 1)
        inside += int(((lb[0] <= i) & (i < lb[1]) & (lb[2] <= j) & (j < lb[3]) & (lb[4] <= k) & (k < lb[5])).sum())
        total += len(f)
    geo["dl_interior"], geo["dl_shell"] = inside, total - inside
    return geo


def update_gbs(geo, S, kt):
    """Algorithmic GB/s of the H and E update kernels together (whichever
    form ran), from per-kernel times over the same steps."""
    ab = algorithmic_bytes(geo, S, 0)
    names = [k for k in ("full_h", "full_e", "lean_h", "shell_h", "lean_e", "shell_e") if k in kt]
    t = sum(kt[k][0] for k in names) / 1e3
    nbytes = sum(ab[k] * kt[k][1] for k in names)
    return round(nbytes / t / 1e9, 1) if t > 0 else None


def algorithmic_bytes(geo, S, nd_vars):
    """Algorithmic HBM bytes per launch of each update pass: an updated cell
    reads its own 6 field values and its ID byte and writes 3 field values
    (neighbour values are reuse, not traffic); each slab a CPML cell lies in
    adds 2 psi values read and written.  Drude-Lorentz cells are counted in
    the special pass (cell index 8 B, slot 4 B, 3 x nd_vars auxiliaries read
    and written); the pass that leaves them alone is charged their ID byte."""
    per_cell = 9 * S + 1
    dl_bytes = 8 + 4 + 3 * nd_vars * 2 * S
    h, e = geo["h"], geo["e"]
    nsp, nsh = geo["dl_interior"], geo["dl_shell"]
    return {
        # merged form: one launch per half-step over every cell
        "full_h": (h["lean"] + h["shell"]) * per_cell + h["slab"] * 4 * S,
        "full_e": (e["lean"] + e["shell"] - nsp - nsh) * per_cell + nsp + nsh + e["slab"] * 4 * S,
        # split form: interior box, CPML shell boxes
        "lean_h": h["lean"] * per_cell,
        "shell_h": h["shell"] * per_cell + h["slab"] * 4 * S,
        "lean_e": (e["lean"] - nsp) * per_cell + nsp,
        "shell_e": (e["shell"] - nsh) * per_cell + nsh + e["slab"] * 4 * S,
        "special_e": (nsp + nsh) * (per_cell + dl_bytes),
    }


def profiled_traffic(kernel, path=os.path.join(ROOT, "profiles", "hbm_traffic_c2_256_fp32.txt")):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy), or None."""
    try:
        with open(path) as fh:
            for line in fh:
                f = line.split()
                if len(f) == 3 and f[0] == "k_" + kernel:
                    return int(f[1]) + int(f[2])
    except OSError:
        pass
    return None


def run_engine(case, precision, steps, warmup, timing_steps):
    """Warmup, then ``steps`` timed steps (wall clock between two device
    synchronisations), then ``timing_steps`` with per-kernel HIP events."""
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


def run_results(case, precision):
    """The whole run of a case; per port: |E|^2 series, DFT, and (with DFT
    frequencies) the spectral energy through the plane (the workload's ports
    are z-normal planes)."""
    import cpsfdtd_amd as eng
    e = eng.Engine(case, precision)
    e.step(case.nt)
    e.sync()
    results = [(e.port_intensity(q), e.port_dft(q), e.port_flux(q, 2) if case.freqs else None)
               for q in range(len(case.ports))]
    e.close()
    return results


def spectrum_peak(freqs, spectrum):
    """Peak frequency and Q = f0 / FWHM of a spectrum (the transmitted
    port's spectral energy; half-maximum crossings by linear interpolation;
    None where the peak is not bracketed)."""
    f = np.asarray(freqs)
    p = np.asarray(spectrum, np.float64)
    order = np.argsort(f)
    f, p = f[order], p[order]
    k = int(np.argmax(p))
    half = p[k] / 2

    def cross(rng):
        for a in rng:
            b = a + 1 if rng.step > 0 else a - 1
            if 0 <= b < len(p) and (p[a] - half) * (p[b] - half) <= 0 and p[a] != p[b]:
                return f[a] + (half - p[a]) * (f[b] - f[a]) / (p[b] - p[a])
        return None
    lo, hi = cross(range(k, 0, -1)), cross(range(k, len(p) - 1))
    q = float(f[k] / (hi - lo)) if lo is not None and hi is not None and hi > lo else None
    return float(f[k]), q


def lorentz_fit(freqs, spectrum):
    """Peak frequency and Q = f0 / FWHM of a Lorentzian A / (1 + ((f - f0)/g)^2)
    fitted by least squares to the spectrum's samples above half maximum
    around its peak (FWHM = 2g); None, None when the half-maximum window is
    not closed on both sides, has fewer than four samples, or the fit does
    not converge."""
    from scipy.optimize import curve_fit
    f = np.asarray(freqs, np.float64)
    p = np.asarray(spectrum, np.float64)
    order = np.argsort(f)
    f, p = f[order], p[order]
    k = int(np.argmax(p))
    lo = k
    while lo > 0 and p[lo - 1] >= p[k] / 2:
        lo -= 1
    hi = k
    while hi < len(p) - 1 and p[hi + 1] >= p[k] / 2:
        hi += 1
    if lo == 0 or hi == len(p) - 1 or hi - lo + 1 < 4:  # not bracketed, or too few samples
        return None, None
    x, y = f[lo:hi + 1] / f[k], p[lo:hi + 1] / p[k]  # unit scale for the solver
    g0 = max((x[-1] - x[0]) / 2, 1e-3)
    try:
        (a, f0, g), _ = curve_fit(lambda v, a, f0, g: a / (1 + ((v - f0) / g) ** 2), x, y, p0=(1.0, 1.0, g0),
                                  maxfev=2000)
    except (RuntimeError, ValueError):
        return None, None
    if not (np.isfinite(f0) and np.isfinite(g) and g != 0):
        return None, None
    return float(f0 * f[k]), float(f0 / (2 * abs(g)))


def quality(res32, res64, freqs):
    """FP32 against FP64 over the same run: relative L2 of every port's |E|^2
    time series, DFT and flux spectrum, and the transmitted port's flux
    spectrum peak and Q (read at half maximum, and from a Lorentzian fit)."""

    def rel(a, b):
        den = np.linalg.norm(np.ravel(b))
        return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / den) if den > 0 else 0.0
    out = {"port_series_rel_l2": [round(rel(a[0], b[0]), 9) for a, b in zip(res32, res64)],
           "port_dft_rel_l2": [round(rel(a[1], b[1]), 9) for a, b in zip(res32, res64)]}
    if freqs and res64:
        out["port_flux_rel_l2"] = [round(rel(a[2], b[2]), 9) for a, b in zip(res32, res64)]
        f32, q32 = spectrum_peak(freqs, res32[-1][2])
        f64, q64 = spectrum_peak(freqs, res64[-1][2])
        out.update({"peak_thz_f32": round(f32 / 1e12, 4), "peak_thz_f64": round(f64 / 1e12, 4),
                    "q_f32": q32 and round(q32, 4), "q_f64": q64 and round(q64, 4)})
        for prec, res in ((32, res32), (64, res64)):
            f0, q = lorentz_fit(freqs, res[-1][2])
            out["lorentz_f%d" % prec] = {"peak_thz": f0 and round(f0 / 1e12, 4), "q": q and round(q, 4)}
    return out


def cpu_baseline(scale_steps=200):
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
                          "ID-2 cells are skipped), FP64, reference sources at -O2 -fopenmp, %.1f s; the reference "
                          "runs its own CPML-H sign and TF/SF face forms, the GPU the fixed ones (same work per "
                          "cell: a rate comparison)" % (scale_steps, r["seconds"])}
    d = tempfile.mkdtemp(prefix="bench_port_")
    case = configs.c2(d, nt=4, nfreq=0)
    case.ports = []
    o = pyoracle.OracleSim(case)
    t0 = time.perf_counter()
    o.step(3)
    sec = time.perf_counter() - t0
    return {"value": round(256 ** 3 * 3 / sec / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": "3 steps of the same workload on the FP64 oracle, %.1f s" % sec}


def cpu_variants():
    """The reference loop's rate on the same 256^3 grid at the reference
    makefile's optimisation (-O0) and at -O2, on one thread and on all of
    this process's cores; each a bounded sample (10-200 steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from cpsfdtd_amd import configs
    threads = len(os.sched_getaffinity(0))
    out = []
    for opt, nth, steps in (("O2", threads, 200), ("O2", 1, 20), ("O0", threads, 40), ("O0", 1, 10)):
        d = tempfile.mkdtemp(prefix="bench_ref_")
        case = configs.c2_reference_core(d, nt=steps + 2)
        r = pyoracle.ref_bench(case, steps, threads=nth, opt=opt)
        out.append({"opt": "-" + opt, "threads": r["threads"], "steps": steps, "seconds": round(r["seconds"], 2),
                    "mcells_per_s": round(r["mcells_per_s"], 3)})
        log("bench: reference %s, %d threads: %.1f Mcell-upd/s" % (opt, r["threads"], r["mcells_per_s"]))
    return out


def batched_c4(steps=200, warmup=10, n=512):
    """The batched unit-cell workload: n periodic 64x64x96 cells (seeded
    pillars and Drude disks, normal-incidence TF/SF, R/T ports with a
    32-frequency DFT), FP32, one launch per pass over all cells."""
    import cpsfdtd_amd as eng
    from cpsfdtd_amd import configs
    root = tempfile.mkdtemp(prefix="bench_c4_")
    ps = configs.c4_params(n)
    cases = []
    for i in range(n):
        d = os.path.join(root, "i%d" % i)
        os.makedirs(d)
        cases.append(configs.c4(d, ps[i], nt=steps + warmup + 20))
    t0 = time.perf_counter()
    b = eng.Batch(cases, 32)
    setup = time.perf_counter() - t0
    b.step(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    wall = time.perf_counter() - t0
    b.set_timing(True)
    b.step(20)
    b.sync()
    kt = b.kernel_times()
    nbytes = b.device_bytes()
    b.close()
    cells = n * 64 * 64 * 96
    return {"workload": "c4: %d periodic 64x64x96 cells, FP32" % n, "instances": n, "steps": steps,
            "cells_per_step": cells, "value": round(cells * steps / wall / 1e6, 2), "unit": "Mcell-updates/s",
            "ms_per_step": round(wall / steps * 1e3, 4), "setup_s": round(setup, 2), "device_gb": round(nbytes / 1e9, 2),
            "kernels_ms": {k: round(v[0] / v[1], 4) for k, v in kt.items()}}


def serving_c5(nreq=120, precision=32, quantum=256, policy=0, max_active=0, pack_cells=0, step_cells=0):
    """The first ``nreq`` requests of the serving workload's stream (Poisson
    20/s, log-uniform 32-320 edges, 500-8000 steps, sphere / Drude-Lorentz /
    gain mix), submitted at their arrival times to the server: latency from
    arrival to port results on the host, p50 and p99."""
    import numpy as np
    import cpsfdtd_amd as eng
    from cpsfdtd_amd import configs
    root = tempfile.mkdtemp(prefix="bench_c5_")
    rq = configs.c5_requests(nreq)
    cases = []
    for k, (t, p) in enumerate(rq):
        d = os.path.join(root, "r%d" % k)
        os.makedirs(d)
        cases.append((t, configs.c5(d, p)))
    work = sum(p["N"][0] * p["N"][1] * p["N"][2] * p["steps"] for _, p in rq)
    srv = eng.Server(quantum=quantum, policy=policy, max_active=max_active, pack_cells=pack_cells, step_cells=step_cells)
    try:
        t0 = time.perf_counter()
        ids = []
        for t, case in cases:
            dt = t0 + t - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(case, precision))
        lat, first, last = [], None, 0.0
        for k, rid in enumerate(ids):
            if k % 100 == 0:
                log("bench: serving, %d of %d requests done (%.0f s)" % (k, len(ids), time.perf_counter() - t0))
            st = srv.wait(rid, 1800.0)
            if st != eng.REQ_DONE:
                raise RuntimeError("request %d: %s" % (rid, srv.error(rid)))
            _, (ts, tb, td) = srv.poll(rid)
            lat.append(td - ts)
            first = ts if first is None else min(first, ts)
            last = max(last, td)
            srv.release(rid)
        stats = srv.stats()
    finally:
        srv.close()
    lat = np.array(lat)
    span = last - first
    return {"workload": "c5: first %d of the 2000 requests of the seeded stream (Poisson 20/s), FP%d" % (nreq, precision),
            "requests": nreq, "p50_s": round(float(np.percentile(lat, 50)), 3),
            "p99_s": round(float(np.percentile(lat, 99)), 3), "mean_s": round(float(lat.mean()), 3),
            "value": round(work / span / 1e6, 2), "unit": "Mcell-updates/s (served)",
            "arrival_span_s": round(rq[-1][0], 2), "makespan_s": round(span, 2),
            "demand_mcells": round(work / rq[-1][0] / 1e6, 1), "peak_active": stats["peak_active"],
            "peak_device_gb": round(stats["peak_bytes"] / 1e9, 2), "quantum": quantum,
            "policy": "shortest work first" if policy == eng.POLICY_SHORTEST else "first come, first served",
            "stepping": ("requests under %g cells packed (one launch per pass over all of them), larger ones on their "
                         "own streams" % pack_cells if pack_cells else "one stream per request")
                        + ("; at most %g cells stepping at once" % (step_cells if step_cells > 0 else 2e7)
                           if step_cells >= 0 else "")}


def serving_c5_sequential(nreq=40, precision=32):
    """The same first ``nreq`` requests solved one after another on the
    engine alone (set-up, every step, port DFT read back, close): the rate a
    single stream of work gets, beside the server's concurrent streams, and
    a run whose kernel mix rocprofv3 can trace from one thread."""
    import cpsfdtd_amd as eng
    from cpsfdtd_amd import configs
    root = tempfile.mkdtemp(prefix="bench_c5s_")
    rq = configs.c5_requests(nreq)
    work, t0 = 0, time.perf_counter()
    for k, (_, p) in enumerate(rq):
        d = os.path.join(root, "r%d" % k)
        os.makedirs(d)
        case = configs.c5(d, p)
        e = eng.Engine(case, precision)
        e.step(case.nt)
        e.port_dft(0)
        e.close()
        work += p["N"][0] * p["N"][1] * p["N"][2] * p["steps"]
        if k % 10 == 9:
            log("bench: sequential, %d of %d requests (%.0f s)" % (k + 1, nreq, time.perf_counter() - t0))
    wall = time.perf_counter() - t0
    return {"workload": "c5: first %d of the 2000 requests, solved one after another, FP%d" % (nreq, precision),
            "requests": nreq, "value": round(work / wall / 1e6, 2), "unit": "Mcell-updates/s", "wall_s": round(wall, 2)}


def serving_c5_kernels(nreq=40, precision=32):
    """Per-kernel time mix of the first ``nreq`` requests, each solved alone
    with HIP events around every launch (cfd_kernel_times; eager launches,
    no graphs), summed over the requests: total ms, launches, mean us and
    share of the kernel time, and the requests' cells x steps."""
    import cpsfdtd_amd as eng
    from cpsfdtd_amd import configs
    root = tempfile.mkdtemp(prefix="bench_c5k_")
    tot, work = {}, 0
    for k, (_, p) in enumerate(configs.c5_requests(nreq)):
        d = os.path.join(root, "r%d" % k)
        os.makedirs(d)
        case = configs.c5(d, p)
        e = eng.Engine(case, precision)
        e.set_timing(True)
        e.step(case.nt)
        for name, (ms, cnt) in e.kernel_times().items():
            a = tot.setdefault(name, [0.0, 0])
            a[0] += ms
            a[1] += cnt
        e.close()
        work += p["N"][0] * p["N"][1] * p["N"][2] * p["steps"]
    all_ms = sum(v[0] for v in tot.values())
    return {"workload": "c5: first %d of the 2000 requests, each alone, FP%d, HIP events per launch" % (nreq, precision),
            "cell_updates": work, "kernel_ms": round(all_ms, 2),
            "kernel_rate_mcells": round(work / (all_ms / 1e3) / 1e6, 1),
            "kernels": {k: {"total_ms": round(v[0], 2), "launches": v[1], "avg_us": round(v[0] / v[1] * 1e3, 2),
                            "share": round(v[0] / all_ms, 4)}
                        for k, v in sorted(tot.items(), key=lambda kv: -kv[1][0])}}


def serving_c2(nreq=4, precision=32, steps=4000, step_cells=0):
    """``nreq`` copies of the 256^3 workload (the whole 4000-step run each)
    submitted at once to the server with its defaults: served rate over the
    makespan and latency from submission to port results on the host, beside
    the same requests solved one after another on the engine alone."""
    import cpsfdtd_amd as eng
    from cpsfdtd_amd import configs
    cases = [configs.c2(tempfile.mkdtemp(prefix="bench_c2s_"), nt=steps) for _ in range(nreq)]
    cells = 256 ** 3
    srv = eng.Server(step_cells=step_cells)
    try:
        t0 = time.perf_counter()
        ids = [srv.submit(c, precision) for c in cases]
        lat = []
        for rid in ids:
            if srv.wait(rid, 900.0) != eng.REQ_DONE:
                raise RuntimeError("request %d: %s" % (rid, srv.error(rid)))
            _, (ts, tb, td) = srv.poll(rid)
            lat.append(td - ts)
            srv.release(rid)
        span = time.perf_counter() - t0
        stats = srv.stats()
    finally:
        srv.close()
    t0 = time.perf_counter()
    for c in cases:
        e = eng.Engine(c, precision)
        e.step(c.nt)
        e.port_dft(0)
        e.close()
    solo = time.perf_counter() - t0
    return {"workload": "c2: %d copies of the 256^3 workload, %d steps each, FP%d, submitted at once" % (nreq, steps, precision),
            "requests": nreq, "value": round(nreq * cells * steps / span / 1e6, 2), "unit": "Mcell-updates/s (served)",
            "makespan_s": round(span, 2), "p50_s": round(float(np.percentile(lat, 50)), 3),
            "max_s": round(max(lat), 3), "peak_active": stats["peak_active"],
            "solo_sequential": round(nreq * cells * steps / solo / 1e6, 2),
            "stepping": "one stream per request; at most %g cells stepping at once" % (step_cells if step_cells > 0 else 2e7)
            if step_cells >= 0 else "one stream per request",
            "basis": "server: setup, stepping, results to host and release inside the makespan; solo: the same "
                     "per request (Engine set-up, steps, DFT read back, close)"}


def cpu_fifo_model(nreq, mcells_per_s):
    """Latencies of the same request stream solved one after another, first
    come first served, by a CPU at a fixed rate (the measured reference rate
    on the same kind of grid): service time = Nx Ny Nz steps / rate."""
    from cpsfdtd_amd import configs
    rq = configs.c5_requests(nreq)
    free, lat = 0.0, []
    for t, p in rq:
        done = max(t, free) + p["N"][0] * p["N"][1] * p["N"][2] * p["steps"] / (mcells_per_s * 1e6)
        lat.append(done - t)
        free = done
    lat = np.array(lat)
    return {"p50_s": round(float(np.percentile(lat, 50)), 1), "p99_s": round(float(np.percentile(lat, 99)), 1),
            "makespan_s": round(free - rq[0][0], 1), "rate_mcells": mcells_per_s,
            "basis": "sequential FIFO solve at the CPU baseline's measured rate (a model, not a run)"}


def other_configs(steps=512, warmup=128, names=("c1", "c3")):
    """The single-grid workloads beside c2, each in FP32 and FP64: c1 (128^3
    sphere, Ez source, line port recording every step) and c3 (192x192x64
    rod slab with four-level gain, periodic x/y, sphere port), ``steps``
    timed steps after ``warmup`` (whole 128-step graph blocks, as the
    engine replays them), Mcell-updates/s (Nx Ny Nz steps / wall)."""
    import cpsfdtd_amd as eng
    from cpsfdtd_amd import configs
    out = {}
    for name, make in (("c1", configs.c1), ("c3", configs.c3)):
        if name not in names:
            continue
        case = make(tempfile.mkdtemp(prefix="bench_%s_" % name), nt=warmup + steps + BLOCK)
        model = eng.Model(case)
        N = model.N
        geo = workload_geometry(model, case.periodic)
        del model
        cells = N[0] * N[1] * N[2]
        row = {"cells": cells, "steps": steps}
        for prec in (32, 64):
            e = eng.Engine(case, prec)
            e.step(warmup)
            e.sync()
            t0 = time.perf_counter()
            e.step(steps)
            e.sync()
            wall = time.perf_counter() - t0
            e.set_timing(True)
            e.step(BLOCK)
            e.sync()
            kt = e.kernel_times()
            e.close()
            row["f%d" % prec] = round(cells * steps / wall / 1e6, 1)
            row["f%d_update_gbs" % prec] = update_gbs(geo, prec // 8, kt)
        out[name] = row
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3584)
    ap.add_argument("--warmup", type=int, default=128)
    ap.add_argument("--timing-steps", type=int, default=128)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched c4 measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the served c5 measurement")
    ap.add_argument("--no-configs", action="store_true", help="skip the c1 / c3 measurements")
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--only", choices=["c1", "c3", "c4", "serve", "serve-seq", "serve-kernels", "serve-c2"],
                    help="only this figure (for per-workload profiles): print its JSON and exit")
    ap.add_argument("--no-quality", action="store_true", help="skip the two whole 4000-step quality runs")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--cpu-variants", action="store_true",
                    help="only time the reference loop at -O0 / -O2 on 1 / all threads (no GPU) and print it")
    args = ap.parse_args()
    if args.cpu_variants:
        print(json.dumps({"cpu_variants": cpu_variants(), "workload": "256^3 grid of the c2 workload, CPML + TF/SF, "
                          "FP64, the reference's own sources"}), flush=True)
        return
    if args.copy_calib:
        import torch
        n = 1 << 28
        a = torch.ones(n, device="cuda:0", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": n * 4, "dispatches": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not supported (multi-GPU sharding is a design "
              "note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import cpsfdtd_amd as eng
    from cpsfdtd_amd import configs
    info = eng.device_info()
    log("bench: device", info)
    if args.only:
        log("bench: only %s" % args.only)
        fig = {"c1": lambda: other_configs(names=("c1",)), "c3": lambda: other_configs(names=("c3",)),
               "c4": batched_c4, "serve": lambda: serving_c5(args.serve_requests),
               "serve-seq": lambda: serving_c5_sequential(args.serve_requests),
               "serve-kernels": lambda: serving_c5_kernels(args.serve_requests), "serve-c2": serving_c2}[args.only]()
        print(json.dumps({"only": args.only, "result": fig, "device": info["name"]}), flush=True)
        return
    warmup, steps, tsteps = whole_blocks(args.warmup), whole_blocks(args.steps), whole_blocks(args.timing_steps)
    nt = warmup + steps + tsteps
    d = tempfile.mkdtemp(prefix="bench_c2_")
    case = configs.c2(d, scale=args.scale, nt=nt)
    model = eng.Model(case)
    N = model.N
    cells = N[0] * N[1] * N[2]
    nd_vars = 1 + 2 * 2 + 1  # 1 Drude + 2 Lorentz poles: JD, JL, JL1, E1
    geo = workload_geometry(model)
    del model

    res = {}
    for prec in ([32] if args.no_fp64 else [32, 64]):
        S = prec // 8
        log("bench: FP%d run, %d warmup + %d timed steps" % (prec, warmup, steps))
        wall, kt, nbytes = run_engine(case, prec, steps, warmup, tsteps)
        ab = algorithmic_bytes(geo, S, nd_vars)
        kstat = {}
        for name, (ms, cnt) in kt.items():
            kstat[name] = {"avg_ms": ms / cnt, "launches": cnt}
        dom = max((k for k in ab if k in kt), key=lambda k: kt[k][0])
        avg_s = kstat[dom]["avg_ms"] / 1e3
        step_ms = sum(v["avg_ms"] * v["launches"] for v in kstat.values()) / tsteps
        res[prec] = {
            "wall": wall, "mcells": cells * steps / wall / 1e6, "ms_per_step": wall / steps * 1e3,
            "kernels": kstat, "dominant": dom, "dom_gbs": ab[dom] / avg_s / 1e9, "dom_bytes": ab[dom],
            "step_gbs": sum(v for k, v in ab.items() if k in kt) / (wall / steps) / 1e9, "device_bytes": nbytes,
            "kernel_ms_per_step": step_ms,
            "pass_gbs": {k: round(ab[k] / (kstat[k]["avg_ms"] / 1e3) / 1e9, 1) for k in ab if k in kstat},
        }
        log("bench: FP%d %.1f Mcell-upd/s, %.3f ms/step, %s %.1f GB/s" % (
            prec, res[prec]["mcells"], res[prec]["ms_per_step"], dom, res[prec]["dom_gbs"]))

    r32 = res[32]
    out = {
        "metric": "Mcell-updates/s & HBM GB/s vs 8 TB/s; p50/p99 solve latency, FP32 vs FP64",
        "value": round(r32["mcells"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "steps_requested": args.steps,
        "ms_per_step": round(r32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (generated 256^3 geometry in the reference's input formats)",
        "config": {"workload": "c2: 256^3 Drude-Lorentz spheres + 12-cell CFS-PML + oblique Gaussian-beam TF/SF, "
                               "2 plane ports x 64-frequency FP64 DFT (stride 8), FP32 fields",
                   "cells": cells, "steps_in_run": nt, "parallelism": "single GPU",
                   "timed_path": "replayed 128-step graph blocks (every pass of the step, port DFT flush included)"},
        "roofline": {"bound": "hbm", "achieved": round(r32["dom_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r32["dom_gbs"] / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(r32["dominant"]),
                     "traffic_source": "profiles/hbm_traffic_c2_256_fp32.txt",
                     "kernel": r32["dominant"], "bytes_per_launch": r32["dom_bytes"],
                     "avg_ms": round(r32["kernels"][r32["dominant"]]["avg_ms"], 5)},
        "kernels_f32": {k: round(v["avg_ms"], 5) for k, v in r32["kernels"].items()},
        "step_gbs_f32": round(r32["step_gbs"], 1),
        "pass_gbs_f32": r32["pass_gbs"],
        "device": info["name"],
    }
    if 64 in res:
        r64 = res[64]
        out["fp64"] = {"value": round(r64["mcells"], 2), "ms_per_step": round(r64["ms_per_step"], 5),
                       "dominant_gbs": round(r64["dom_gbs"], 1),
                       "kernels": {k: round(v["avg_ms"], 5) for k, v in r64["kernels"].items()}}
        out["fp32_over_fp64"] = round(r32["mcells"] / r64["mcells"], 3)
    if 64 in res and not args.no_quality:
        log("bench: whole 4000-step runs in FP32 and FP64 for the quality figures")
        qcase = configs.c2(tempfile.mkdtemp(prefix="bench_q_"), scale=args.scale, nt=4000)
        try:
            q = quality(run_results(qcase, 32), run_results(qcase, 64), qcase.freqs)
            q["run"] = "whole 4000-step workload in each precision"
            out["quality_f32_vs_f64"] = q
        except Exception as ex:
            out["quality_f32_vs_f64"] = {"error": str(ex)}
    # the other workloads' figures are reported beside the main one; a
    # failure there is recorded in the line instead of losing it
    if not args.no_configs:
        log("bench: c1 and c3, FP32 and FP64")
        try:
            out["configs"] = other_configs()
            log("bench: %s" % out["configs"])
        except Exception as ex:
            out["configs"] = {"error": str(ex)}
            log("bench: c1/c3 failed: %s" % ex)
    if not args.no_batch:
        log("bench: batched c4 (512 unit cells)")
        try:
            out["batched"] = batched_c4()
            log("bench: batched c4 %.1f Mcell-upd/s" % out["batched"]["value"])
        except Exception as ex:
            out["batched"] = {"error": str(ex)}
            log("bench: batched c4 failed: %s" % ex)
    if not args.no_serve:
        log("bench: serving four 256^3 c2 requests")
        try:
            out["serving_c2"] = serving_c2()
            log("bench: served c2 %.1f Mcell-upd/s" % out["serving_c2"]["value"])
        except Exception as ex:
            out["serving_c2"] = {"error": str(ex)}
            log("bench: served c2 failed: %s" % ex)
        log("bench: serving the first %d c5 requests" % args.serve_requests)
        # the server's default admission (first come, first served), then
        # shortest work first on the same stream
        for key, pol in (("serving", eng.POLICY_FIFO), ("serving_shortest_first", eng.POLICY_SHORTEST)):
            try:
                out[key] = serving_c5(args.serve_requests, policy=pol)
                log("bench: %s p50 %.2f s, p99 %.2f s" % (key, out[key]["p50_s"], out[key]["p99_s"]))
            except Exception as ex:
                out[key] = {"error": str(ex)}
                log("bench: %s failed: %s" % (key, ex))
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as ex:  # reported, never fatal for the GPU figure
            out["cpu_baseline"] = {"error": str(ex)}
        rate = out["cpu_baseline"].get("value")
        if rate and "p50_s" in out.get("serving", {}):
            out["serving"]["cpu_fifo_model"] = cpu_fifo_model(args.serve_requests, rate)
        if rate and "value" in out.get("batched", {}):
            # the reference has no periodic boundaries (PBC.cpp is a stub), so
            # the 512 unit cells cannot run on it: its measured rate on the
            # 256^3 grid, applied to their cells
            b = out["batched"]
            b["cpu_rate_model"] = {"ms_per_step": round(b["cells_per_step"] / (rate * 1e6) * 1e3, 1),
                                   "rate_mcells": rate,
                                   "basis": "the CPU baseline's measured rate applied to the batch's cells (a model: "
                                            "the reference has no periodic boundaries)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
