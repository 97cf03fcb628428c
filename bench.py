"""Benchmark of the FDTD hot path on one MI355X.

Workload (N=1): 3-D 256^3 FP32, 12-cell CPML on all faces, 48 seeded
dielectric blocks (eps_r U[1,12]), x-propagating Ez-polarised TFSF plane wave
(Gaussian-modulated sinusoid) on a box inset 20 cells, running DFT of all 6
components on 2 z-planes at 8 frequencies.  One bench "step" is one FDTD time
step over the whole grid.  value = interior cells x timed steps / seconds.

Prints ONE JSON line.  Extra objects:
  roofline      algorithmic HBM bytes per step / measured time per step vs 8 TB/s
  cpu_baseline  the reference's own update functors (oracle/_ref/harness) timed
                on this host on a bounded sample of the same grid
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s + %HBM roofline (fwd, adjoint), batched sims/s vs plain"


def algorithmic_bytes(sc, esz):
    """Minimum HBM bytes of one step: every E and H component read and written
    once, one material byte per cell, CPML psi (4 per slab cell) read+written,
    DFT accumulators read+written per monitor cell."""
    nx, ny, nz = sc["nx"], sc["ny"], sc["nz"]
    interior = (nx - 2) * (ny - 2) * (nz - 2)
    b = interior * (12 * esz + 1)
    n = sc["pml"]["n"]
    # slab = indices whose C coefficient is non-zero: 1..n at the low end, L-n-1..L-2 at the high end
    for L, (o1, o2) in ((nx, (ny, nz)), (ny, (nx, nz)), (nz, (nx, ny))):
        slab = 2 * n
        b += slab * (o1 - 2) * (o2 - 2) * 4 * 2 * esz
    for region, mask, pe, _ in sc["dfts"]:
        cells = (region[1] - region[0]) * (region[3] - region[2]) * (region[5] - region[4])
        b += cells * bin(mask).count("1") * pe.shape[1] * 2 * 2 * esz
    return b, interior


def cpu_baseline(size, sample_steps):
    harness = os.path.join(ROOT, "oracle", "_ref", "harness")
    if os.path.exists(harness):
        t0 = time.time()
        r = subprocess.run([harness, "bench3d", str(size), str(sample_steps), "12"], capture_output=True, text=True,
                           timeout=900)
        if r.returncode == 0:
            js = json.loads(r.stdout.strip().splitlines()[-1])
            return {"value": js["cell_updates_per_s"] / 1e9, "unit": "Gcell-updates/s", "cores": 1,
                    "kind": "reference",
                    "sample": f"{size}^3, 12-cell CPML, 48 eps blocks, Dz soft source; {sample_steps} timed steps "
                              f"after 1 untimed; the reference's YeeUpdate/UpdatePML/ConstantUpdate/VacuumUpdate "
                              f"functors, one sweep each (TFSF/DFT have no 3-D form in the reference); "
                              f"wall incl. set-up {time.time() - t0:.1f} s"}
    # fallback: the FP64 oracle restatement with OpenMP
    from oracle import refcpu
    import synth
    from scenarios import apply
    sc = synth.grid3d_cpml_tfsf_dft(n=size, nsteps=sample_steps + 1)
    cores = refcpu.nthreads(0)
    o = apply(refcpu.Sim("ThreeD", size, size, size, sc["dx"], sc["dt"]), sc)
    o.step(1)
    t0 = time.perf_counter()
    o.step(sample_steps)
    dt = time.perf_counter() - t0
    return {"value": (size - 2) ** 3 * sample_steps / dt / 1e9, "unit": "Gcell-updates/s", "cores": cores,
            "kind": "port", "sample": f"{size}^3 full workload, {sample_steps} timed steps, FP64 oracle"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--cpu-sample-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (z-slab sharding is a design "
              "note, not code); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from fdtd_amd import Grid
    import synth
    from scenarios import apply

    torch.cuda.set_device(0)
    total = args.warmup + args.steps
    sc = synth.grid3d_cpml_tfsf_dft(n=args.size, nsteps=total)
    esz = 4 if args.precision == "f32" else 8
    g = apply(Grid("ThreeD", args.size, args.size, args.size, sc["dx"], sc["dt"], precision=args.precision), sc)
    tile = g.autotune(trial_steps=10)
    g.step(args.warmup)
    g.synchronize()
    torch.cuda.synchronize()
    g.set_profiling(True)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    g.step(args.steps)
    ev1.record()
    torch.cuda.synchronize()
    g.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    ke, ne = g.kernel_time(0)
    kh, nh = g.kernel_time(1)
    ka, na = g.kernel_time(2)
    ms_step = ms_total / args.steps
    abytes, interior = algorithmic_bytes(sc, esz)
    gcells = interior * args.steps / (ms_total / 1e3) / 1e9
    kern_ms = (ke / max(ne, 1)) + (kh / max(nh, 1))
    achieved = abytes / (kern_ms / 1e3) / 1e9
    out = {
        "metric": METRIC, "value": round(gcells, 4), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"3-D {args.size}^3 {args.precision.upper()} forward: 12-cell CPML + 48 eps blocks "
                               f"+ TFSF box inset 20 (x-propagating Ez plane wave) + running DFT of 6 comps on "
                               f"2 z-planes x 8 freqs",
                   "cells": args.size ** 3, "interior_cells": interior, "tile": list(tile)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "e_pass_3d + h_pass_3d (one step)",
                     "algorithmic_bytes_per_step": abytes,
                     "avg_ms": {"e_pass": round(ke / max(ne, 1), 5), "h_pass": round(kh / max(nh, 1), 5),
                                "aux_line": round(ka / max(na, 1), 5)}},
    }
    g.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.size, args.cpu_sample_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
