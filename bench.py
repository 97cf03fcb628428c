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
DFT_BLOCK = 16  # steps of samples folded into the DFT accumulators at once
METRIC = "Gcell-updates/s + %HBM roofline (fwd, adjoint), batched sims/s vs plain"


def algorithmic_bytes(sc, esz):
    """Minimum HBM bytes of one step: every E and H component read and written
    once, one material byte per cell, CPML psi (4 per slab cell) read+written,
    and the time-blocked running DFT at its true minimum: per monitor cell and
    component one sample written and read back, plus the FP64 accumulators
    (re, im per frequency) read and written once per DFT_BLOCK steps."""
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
        b += cells * bin(mask).count("1") * (2 * esz + pe.shape[1] * 2 * 2 * 8 / DFT_BLOCK)
    return int(b), interior


def host_cpu():
    """CPU model and logical core count of this host."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(size, sample_steps):
    """The FP64 oracle (oracle/refcpu.c: structure of arrays, OpenMP over cells, the reference's per-functor
    sweeps and operation order) on every host core, on `sample_steps` steps of the full workload; beside it,
    where it was built, the reference's own update functors (oracle/_ref/harness, single-threaded like the
    reference) on 2 steps of the same grid without TFSF and DFT (which the reference has no 3-D form of)."""
    from oracle import refcpu
    import synth
    from scenarios import apply
    model, ncpu = host_cpu()
    sc = synth.grid3d_cpml_tfsf_dft(n=size, nsteps=sample_steps + 1)
    cores = refcpu.nthreads(0)
    o = apply(refcpu.Sim("ThreeD", size, size, size, sc["dx"], sc["dt"]), sc)
    o.step(1)
    t0 = time.perf_counter()
    o.step(sample_steps)
    dt = time.perf_counter() - t0
    del o
    out = {"value": round((size - 2) ** 3 * sample_steps / dt / 1e9, 6), "unit": "Gcell-updates/s", "cores": cores,
           "kind": "port",
           "sample": f"{size}^3 full workload (CPML, 48 blocks, TFSF, 2 DFT planes x 8 freqs), {sample_steps} timed "
                     f"steps after 1 untimed, FP64 oracle with OpenMP on {cores} threads ({dt:.1f} s)",
           "host": {"cpu_model": model, "logical_cores": ncpu}}
    harness = os.path.join(ROOT, "oracle", "_ref", "harness")
    if os.path.exists(harness):
        t0 = time.time()
        r = subprocess.run([harness, "bench3d", str(size), "2", "12"], capture_output=True, text=True, timeout=900)
        if r.returncode == 0:
            js = json.loads(r.stdout.strip().splitlines()[-1])
            out["reference_functors_1core"] = {
                "value": round(js["cell_updates_per_s"] / 1e9, 6), "unit": "Gcell-updates/s", "cores": 1,
                "kind": "reference",
                "sample": f"{size}^3, 12-cell CPML, 48 eps blocks, Dz soft source; 2 timed steps after 1 untimed; "
                          f"YeeUpdate/UpdatePML/ConstantUpdate/VacuumUpdate, one sweep each; wall incl. set-up "
                          f"{time.time() - t0:.1f} s"}
    return out


def copy_rate(seconds=0.3):
    """Measured device-copy bandwidth (1 GiB FP32 -> 1 GiB), GB/s of bytes read + written."""
    import torch
    n = 1 << 28
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    gbs = 2 * n * 4 * reps / (time.perf_counter() - t0) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def timed_steps(g, steps):
    """Times `steps` steps with no per-launch events (the measured value), then the same number of steps again
    with a hipEvent pair around every kernel launch on the grid's stream; returns (ms_total of the first run,
    per-kernel ms per launch, per-kernel ms per step).  The DFT entry covers the per-step sampling and the fold
    every 16 steps."""
    import torch
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    ms_total = (time.perf_counter() - t0) * 1e3
    g.set_profiling(True)
    g.step(steps)
    g.synchronize()
    avg, per_step = {}, {}
    for kid, name in ((0, "e_pass"), (1, "h_pass"), (2, "aux_line"), (3, "fused_pass"), (4, "dft")):
        ms, n = g.kernel_time(kid)
        if n:
            avg[name] = round(ms / n, 5)
            per_step[name] = round(ms / steps, 5)
    g.set_profiling(False)
    return ms_total, avg, per_step


def base_line(args, metric_value, unit, ms_step, dtype, config, higher=True):
    return {"metric": METRIC, "value": metric_value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def run_c1(args):
    """2-D TM 256^2 FP64, 10-cell CPML, centre Gaussian pulse, 4 probes (cache-resident: Gcell/s only)."""
    from fdtd_amd import Grid
    import synth
    from scenarios import apply
    import torch
    sc = synth.tm2d_c1(nsteps=3 * (args.warmup + args.steps))
    g = apply(Grid("TM", 256, 256, 1, sc["dx"], sc["dt"], precision="f64"), sc)
    tile = g.autotune(trial_steps=20)
    g.step(args.warmup)
    ms_plain, avg, _ = timed_steps(g, args.steps)
    # the same grid continued with captured chunks: kernel launches dominate this small grid
    g.set_graph(True)
    g.step(args.warmup)
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.synchronize()
    ms_graph = (time.perf_counter() - t0) * 1e3
    interior = 254 * 254
    # the value is the faster of the two launch modes (both reported)
    ms = min(ms_plain, ms_graph)
    out = base_line(args, round(interior * args.steps / (ms / 1e3) / 1e9, 4), "Gcell-updates/s",
                    round(ms / args.steps, 5), "f64",
                    {"workload": "2-D TM 256^2 FP64, 10-cell CPML, Dz Gaussian soft source, 4 Ez probes",
                     "tile": list(tile), "kernel_avg_ms": avg, "graph_chunks": ms_graph < ms_plain,
                     "plain_launches_gcell_per_s": round(interior * args.steps / (ms_plain / 1e3) / 1e9, 4),
                     "graph_chunks_gcell_per_s": round(interior * args.steps / (ms_graph / 1e3) / 1e9, 4),
                     "note": "state (~3.5 MB) is cache-resident: no HBM roofline; tile (0, TB, C) is the "
                             "time-blocked pass, TB steps per launch on C x C cores"})
    return out


def run_c3(args):
    """3-D 192^3 FP32 Drude + 2-Lorentz sphere (r=40), 12-cell CPML, dipole, DFT box surface x 16 freqs."""
    from fdtd_amd import Grid
    import synth
    from scenarios import apply
    n = args.size if args.size != 256 else 192
    sc = synth.grid3d_dispersive_sphere(n=n, radius=40 if n >= 160 else n // 5, nsteps=args.warmup + 2 * args.steps)
    g = apply(Grid("ThreeD", n, n, n, sc["dx"], sc["dt"], precision=args.precision), sc)
    if args.tile:
        tile = tuple(int(v) for v in args.tile.split(","))
        g.set_tile(*tile)
    else:
        tile = g.autotune(trial_steps=30)
    g.step(args.warmup)
    ms, avg, per_step = timed_steps(g, args.steps)
    esz = 4 if args.precision == "f32" else 8
    abytes, interior = algorithmic_bytes(sc, esz)
    disp = int((sc["matid"] == 1).sum())
    abytes += disp * 3 * (1 + 2 * 3) * 2 * esz  # D and 3 poles x (P, J), read + written
    ms_step = ms / args.steps
    kern_ms = sum(v for k, v in per_step.items() if k != "aux_line")
    achieved = abytes / (kern_ms / 1e3) / 1e9
    return base_line(args, round(interior * args.steps / (ms / 1e3) / 1e9, 4), "Gcell-updates/s", round(ms_step, 5),
                     args.precision,
                     {"workload": f"3-D {n}^3 {args.precision.upper()}: Drude + 2 Lorentz sphere r=40, 12-cell CPML, "
                                  f"z-dipole soft source, DFT of 6 comps on the sphere's box surface x 16 freqs",
                      "tile": list(tile), "dispersive_cells": disp, "kernel_avg_ms": avg,
                      "kernel_ms_per_step": per_step,
                      "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)}})


def run_unit_cell(args):
    """3-D 256^3 FP32 periodic unit cell: x and y periodic (boundary ghost writes after each pass), 12-cell CPML
    along z, 4 dielectric blocks, two point sources, DFT of 6 comps on one z-plane x 4 freqs."""
    from fdtd_amd import Grid
    import synth
    from scenarios import apply
    n = args.size
    sc = synth.grid3d_boundaries(n=n, nsteps=args.warmup + 2 * args.steps, variant="unit_cell", npml=12)
    esz = 4 if args.precision == "f32" else 8
    g = apply(Grid("ThreeD", n, n, n, sc["dx"], sc["dt"], precision=args.precision), sc)
    if args.tile:
        tile = tuple(int(v) for v in args.tile.split(","))
        g.set_tile(*tile)
    else:
        tile = g.autotune(trial_steps=30)
    g.step(args.warmup)
    ms, avg, per_step = timed_steps(g, args.steps)
    # surface charge density and currents of one z-plane of the cell (a synchronous read-back call)
    plane = [1, n - 1, 1, n - 1, n // 2, n // 2 + 1]
    g.surface_quantities(plane)
    t0 = time.perf_counter()
    for _ in range(10):
        g.surface_quantities(plane)
    sq_ms = (time.perf_counter() - t0) / 10 * 1e3
    abytes, interior = algorithmic_bytes(dict(sc, pml=dict(sc["pml"], n=0)), esz)
    # CPML along z only, and the ghost writes: 2 periodic axes x 2 faces x (E and H) x 3 comps, read + written
    abytes += 2 * 12 * (n - 2) * (n - 2) * 4 * 2 * esz
    abytes += 2 * 2 * 2 * 3 * (n - 2) * (n - 2) * 2 * esz
    kern_ms = sum(v for k, v in per_step.items() if k != "aux_line")
    achieved = abytes / (kern_ms / 1e3) / 1e9
    return base_line(args, round(interior * args.steps / (ms / 1e3) / 1e9, 4), "Gcell-updates/s",
                     round(ms / args.steps, 5), args.precision,
                     {"workload": f"3-D {n}^3 {args.precision.upper()} periodic unit cell: x and y periodic, 12-cell "
                                  f"CPML along z, 4 eps blocks, Dz + Bx point sources, DFT of 6 comps on one z-plane "
                                  f"x 4 freqs", "tile": list(tile), "kernel_avg_ms": avg,
                      "surface_quantities_ms_per_plane_call": round(sq_ms, 4),
                      "kernel_ms_per_step": per_step,
                      "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                                   "algorithmic_bytes_per_step": abytes}})


def run_c4(args):
    """64 uneven 3-D FP32 grids: one batch (tile table, hipGraph chunks) vs the same grids stepped one by one."""
    import torch
    from fdtd_amd import Batch, Grid
    import synth
    from scenarios import apply
    nsteps = args.steps
    scs = synth.batch_c4(nsteps=nsteps + args.warmup)
    mk = lambda: [apply(Grid("ThreeD", sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision="f32"), sc)
                  for sc in scs]
    seq = mk()
    for g in seq:
        g.step(args.warmup)
        g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g in seq:
        g.step(nsteps)
    for g in seq:
        g.synchronize()
    t_seq = time.perf_counter() - t0
    for g in seq:
        g.close()
    bat = mk()
    b = Batch(bat)
    btile = b.autotune(trial_steps=10)
    b.step(args.warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(nsteps)
    b.synchronize()
    t_bat = time.perf_counter() - t0
    cells = sum((sc["nx"] - 2) * (sc["ny"] - 2) * (sc["nz"] - 2) for sc in scs)
    sims_bat = len(scs) / t_bat
    return base_line(args, round(sims_bat, 4), "sims/s", round(t_bat / nsteps * 1e3, 5), "f32",
                     {"workload": f"{len(scs)} independent 3-D FP32 grids, edges U[64,160], 8-cell CPML, 10 eps "
                                  f"blocks, Dz point source, 4-freq DFT plane each, {nsteps} steps",
                      "total_interior_cells": cells, "batch_tile": list(btile), "batched_sims_per_s": round(sims_bat, 4),
                      "sequential_sims_per_s": round(len(scs) / t_seq, 4), "speedup": round(t_seq / t_bat, 3),
                      "batched_gcell_per_s": round(cells * nsteps / t_bat / 1e9, 3),
                      "sequential_gcell_per_s": round(cells * nsteps / t_seq / 1e9, 3)})


def run_c5(args):
    """Adjoint eps gradient, 3-D 128^3 FP32: 10-cell CPML, TFSF continuous wave at one frequency, 64x64x16
    design box (eps_r U[1,12]), objective sum |F[Ez]|^2 on a 64x64 plane, checkpoint every 100 steps.
    value = gradient time / plain forward time (both from step 0 over the same steps)."""
    import numpy as np
    import torch
    from fdtd_amd import Grid
    import synth
    from scenarios import apply
    n = args.size if args.size != 256 else 128
    nsteps = args.steps
    sc = synth.grid3d_adjoint(n=n, nsteps=nsteps)
    box, eps = sc["design"]
    mk = lambda prec="f32": apply(Grid("ThreeD", n, n, n, sc["dx"], sc["dt"], precision=prec), sc)

    def wall(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    # plain forward at its autotuned tile
    g = mk()
    tile = g.autotune(trial_steps=10)
    g.step(args.warmup)
    g.synchronize()
    g.close()
    g = mk()
    g.set_tile(*tile)
    t_fwd, _ = wall(lambda: (g.step(nsteps), g.synchronize()))
    g.close()
    two_pass = tile if tile[0] >= 8 else (64, 4, 2)  # the gradient's recording forward runs two passes
    # the gradient's forward pass runs at the forward's two-pass tile; its
    # reverse sweep and recompute at their own tile: with the recompute running
    # beside the reverse steps, other shapes share the GPU better; the best
    # candidate by short gradient runs (an untimed first run loads the adjoint
    # kernels; each candidate keeps the faster of two runs)
    trial = min(600, nsteps)
    best = None
    g = mk()
    g.adjoint_eps_gradient(trial, 0, 4, checkpoint_every=100, history_mb=0.0)
    g.close()
    for cand in dict.fromkeys([two_pass, (128, 2, 4), (64, 4, 4), (64, 2, 4), (64, 4, 6), (64, 4, 8), (128, 2, 8)]):
        for _ in range(2):
            g = mk()
            g.set_tile(*two_pass)
            g.set_adjoint_tile(*cand)
            t, _ = wall(lambda: g.adjoint_eps_gradient(trial, 0, 4, checkpoint_every=100, history_mb=0.0))
            g.close()
            if best is None or t < best[0]:
                best = (t, cand)
    adj_tile = best[1]
    res, grads = {}, {}
    for name, K, mb in (("checkpointed", 100, 0.0), ("history_resident", 100, 1e5)):
        # timed run (no per-kernel events), then a profiled run for the kernel averages
        g = mk()
        g.set_tile(*two_pass)
        g.set_adjoint_tile(*adj_tile)
        t, (J, grad) = wall(lambda: g.adjoint_eps_gradient(nsteps, 0, 4, checkpoint_every=K, history_mb=mb))
        g.close()
        g = mk()
        g.set_tile(*two_pass)
        g.set_adjoint_tile(*adj_tile)
        g.set_profiling(True)
        g.adjoint_eps_gradient(nsteps, 0, 4, checkpoint_every=K, history_mb=mb)
        avg = {}
        for kid, kn in ((0, "e_pass"), (1, "h_pass"), (2, "aux_line"), (4, "dft"), (5, "adj_e_pass"),
                        (6, "adj_h_pass")):
            ms, cnt = g.kernel_time(kid)
            if cnt:
                avg[kn] = round(ms / cnt, 5)
        g.close()
        res[name] = {"seconds": round(t, 4), "ratio_to_forward": round(t / t_fwd, 3), "checkpoint_every": K,
                     "history_mb": mb, "kernel_avg_ms": avg, "objective": J}
        grads[name] = (J, grad)
    # the checkpointed gradient (recomputed segments) is the one the headline ratio times: it must equal the
    # gradient from recorded history bit for bit, and it is the one checked against finite differences below
    Jc, grad = grads["checkpointed"]
    bitwise = Jc == grads["history_resident"][0] and bool(np.array_equal(grad, grads["history_resident"][1]))
    # a batch of 8 designs: one gradient call for all (concurrent streams) vs 8 calls one after another
    from fdtd_amd.grid import adjoint_eps_gradient_batch
    rng8 = np.random.default_rng(8)
    designs = [1.0 + 11.0 * rng8.random(eps.shape) for _ in range(8)]

    def mkd(e):
        q = apply(Grid("ThreeD", n, n, n, sc["dx"], sc["dt"], precision="f32"), dict(sc, design=(box, e)))
        q.set_tile(*two_pass)
        q.set_adjoint_tile(*adj_tile)
        return q

    t_seq = 0.0
    for e in designs:
        q = mkd(e)
        t, _ = wall(lambda: q.adjoint_eps_gradient(nsteps, 0, 4, checkpoint_every=100, history_mb=0.0))
        t_seq += t
        q.close()
    qs = [mkd(e) for e in designs]
    t_bat, _ = wall(lambda: adjoint_eps_gradient_batch(qs, nsteps, 0, 4, checkpoint_every=100, history_mb=0.0))
    for q in qs:
        q.close()
    batch8 = {"designs": 8, "checkpoint_every": 100, "history_mb": 0.0, "batched_seconds": round(t_bat, 4),
              "sequential_seconds": round(t_seq, 4), "batched_designs_per_s": round(8 / t_bat, 3),
              "sequential_designs_per_s": round(8 / t_seq, 3), "speedup": round(t_seq / t_bat, 3),
              "batched_ratio_to_forward_per_design": round(t_bat / 8 / t_fwd, 3)}
    # central finite differences (h = 1e-3) of FP64 forward runs on 16 random design cells
    rng = np.random.default_rng(16)
    cells = rng.choice(eps.size, 16, replace=False)

    def objective(e):
        q = apply(Grid("ThreeD", n, n, n, sc["dx"], sc["dt"], precision="f64"), dict(sc, design=(box, e)))
        q.set_tile(*two_pass)
        q.step(nsteps)
        F = q.get_dft(0)
        q.close()
        return float(np.sum(F * F))

    fd = []
    for c in cells:
        ep = eps.reshape(-1).copy()
        em = eps.reshape(-1).copy()
        ep[c] += 1e-3
        em[c] -= 1e-3
        fd.append((objective(ep.reshape(eps.shape)) - objective(em.reshape(eps.shape))) / 2e-3)
    fd = np.array(fd)
    got = grad.reshape(-1)[cells]
    fd_err = float(np.max(np.abs(got - fd)) / np.max(np.abs(fd)))
    # roofline of the adjoint pair: the forward's per-step bytes (fields, material, psi) plus, per design
    # cell, the recorded E increment (3 reads) and the gradient read-modify-write (FP64)
    esz = 4
    abytes, interior = algorithmic_bytes(dict(sc, dfts=[]), esz)
    ndes = int(eps.size)
    abytes += ndes * (3 * esz + 16)
    # adjoint kernels timed where nothing runs beside them (with history resident there is no recompute)
    adj = res["history_resident"]["kernel_avg_ms"]
    adj_ms = adj.get("adj_e_pass", 0) + adj.get("adj_h_pass", 0)
    achieved = abytes / (adj_ms / 1e3) / 1e9 if adj_ms else None
    out = base_line(args, res["checkpointed"]["ratio_to_forward"], "x forward time", round(t_fwd / nsteps * 1e3, 5),
                    "f32",
                    {"workload": f"adjoint eps gradient, 3-D {n}^3 FP32, 10-cell CPML, TFSF continuous wave at one "
                                 f"frequency, design box {list(eps.shape[::-1])} eps_r U[1,12], objective sum |F[Ez]|^2 "
                                 f"on a {sc['dfts'][0][0][3] - sc['dfts'][0][0][2]}^2 plane, {nsteps} steps",
                     "forward_seconds": round(t_fwd, 4), "forward_tile": list(tile), "gradient_tile": list(adj_tile),
                     "forward_gcell_per_s": round(interior * nsteps / t_fwd / 1e9, 3),
                     "gradient": res, "checkpointed_equals_history_resident_bitwise": bitwise,
                     "fd_check": {"cells": 16, "h": 1e-3, "reference": "FP64 GPU forward runs",
                                  "gradient": "checkpointed (K = 100, recomputed segments)",
                                  "max_abs_err_over_max_abs_fd": fd_err},
                     "batch_of_designs": batch8},
                    higher=False)
    if achieved:
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                           "kernel": "adjoint step: adj_e_pass + adj_h_pass", "algorithmic_bytes_per_step": abytes}
    return out


def run_copy(args):
    """Plain device copy of a 1 GiB FP32 buffer into another (torch): the calibration kernel for the HBM
    counters (FETCH_SIZE / WRITE_SIZE of a known 1 GiB read + 1 GiB write) and the measured copy bandwidth."""
    import torch
    n = 1 << 28
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    gbs = 2 * n * 4 * args.steps / (ms / 1e3) / 1e9
    return base_line(args, round(gbs, 1), "GB/s", round(ms / args.steps, 5), "f32",
                     {"workload": "device copy, 1 GiB FP32 -> 1 GiB FP32", "bytes_per_copy": 2 * n * 4})


# committed FETCH_SIZE / WRITE_SIZE summaries of the headline workload, per schedule (the bytes a schedule moves
# depend on it): the two passes at tile 128,2,4 and the time-blocked pass at 2,12,32
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", "hbm_traffic_3d256_f32.json"),
                 os.path.join(ROOT, "profiles", "hbm_traffic_3d256_f32_time_blocked.json")]
TRAFFIC_NOTE = (": rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this workload at the same tile, each scaled by "
                "the counter's reading for a known 1 GiB device copy; not measured inside this run")
PASS_KERNELS = ("fused3d_lds", "e_pass_3d", "h_pass_3d", "dft_sample", "dft_fold", "fusedz_3d", "tm_block", "tb_3d")


def _short(name):
    for k in PASS_KERNELS + ("dft_accumulate", "aux_line", "copyBuffer"):
        if k in name:
            return k
    return name.split("(")[0][:60]


def _counter_rows(d, counter):
    import csv
    agg = {}
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                agg.setdefault(_short(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return agg


def traffic_summary(args):
    """HBM bytes per launch of each kernel from rocprofv3 --pmc runs of this bench: FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch) of the workload, each scaled by known bytes / counter for the 1 GiB device copy
    (`--workload copy`, the runtime's copyBuffer kernel: 1 GiB read + 1 GiB written).  Launches per step are
    the kernel's dispatches over the step kernel's (h_pass_3d, or the single fused pass)."""
    fetch_dir, write_dir, cfetch_dir, cwrite_dir = args.traffic_summary
    known = float(1 << 30)
    mean = lambda v: sum(v) / len(v)
    cf = mean(_counter_rows(cfetch_dir, "FETCH_SIZE")["copyBuffer"]) * 1024.0
    cw = mean(_counter_rows(cwrite_dir, "WRITE_SIZE")["copyBuffer"]) * 1024.0
    ff, wf = known / cf, known / cw
    fetch = _counter_rows(fetch_dir, "FETCH_SIZE")
    write = _counter_rows(write_dir, "WRITE_SIZE")
    if "tb_3d" in fetch:
        # the time-blocked pass advances two steps per launch (one at odd starts and ends)
        if not args.traffic_steps:
            sys.exit("--traffic-summary of the time-blocked pass needs --traffic-steps (the profiled run's steps)")
        nsteps = args.traffic_steps
    else:
        step_kernel = "fused3d_lds" if "fused3d_lds" in fetch else ("h_pass_3d" if "h_pass_3d" in fetch else "fusedz_3d")
        nsteps = len(fetch[step_kernel])
    kern = {k: {"fetch_bytes": round(mean(fetch[k]) * 1024.0 * ff),
                "write_bytes": round(mean(write.get(k, [0.0])) * 1024.0 * wf),
                "launches_per_step": round(len(fetch[k]) / nsteps, 5)}
            for k in fetch if k in PASS_KERNELS}
    per_step = sum(v["launches_per_step"] * (v["fetch_bytes"] + v["write_bytes"]) for v in kern.values())
    out = {"workload": "3-D 256^3 FP32 forward (bench.py default workload)", "tile": args.tile,
           "fetch_factor": round(ff, 4), "write_factor": round(wf, 4),
           "copy_counters_kib": {"FETCH_SIZE": cf / 1024.0, "WRITE_SIZE": cw / 1024.0}, "kernels": kern,
           "bytes_per_step": round(per_step)}
    with open(args.traffic_out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)


def step_traffic(tile, per_step=None, steps=None):
    """HBM bytes of one step and their source from the committed counter summary taken at this tile (the bytes
    a schedule moves depend on it); (None, None) when there is none."""
    for path in TRAFFIC_FILES:
        try:
            with open(path) as f:
                t = json.load(f)
        except OSError:
            continue
        if t.get("tile") and [int(v) for v in str(t["tile"]).split(",")] == list(tile):
            return t.get("bytes_per_step"), os.path.relpath(path, ROOT) + TRAFFIC_NOTE
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 2000; c4 1500; c5 3000, the workload's own step count)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--cpu-sample-steps", type=int, default=150)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "unit_cell", "copy"],
                    help="c2 (default): the headline 3-D 256^3 forward; c1: 2-D TM 256^2; c3: dispersive sphere; "
                         "c4: 64 batched grids vs sequential; c5: adjoint eps gradient 128^3; unit_cell: 256^3 with periodic "
                         "x/y boundaries; copy: 1 GiB device copy")
    ap.add_argument("--tile", default=None,
                    help="bx,by,kc: fixed tile instead of the autotuner (c2; bx = 0 selects the fused pass)")
    ap.add_argument("--graph", action="store_true", help="capture chunks of steps in hipGraphs (c2)")
    ap.add_argument("--traffic-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="turn rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this bench and of --workload copy into "
                         "HBM bytes per launch (no GPU work)")
    ap.add_argument("--traffic-out", default="hbm_traffic.json")
    ap.add_argument("--traffic-steps", type=int, default=None,
                    help="steps the profiled run took (warm-up + 2 x steps): per-step bytes of a pass that advances "
                         "several steps per launch")
    args = ap.parse_args()
    if args.traffic_summary:
        traffic_summary(args)
        return
    if args.steps is None:
        args.steps = {"c4": 1500, "c5": 3000, "copy": 20}.get(args.workload, 2000)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (z-slab sharding is a design "
              "note, not code); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from fdtd_amd import Grid
    import synth
    from scenarios import apply

    torch.cuda.set_device(0)
    if args.workload != "c2":
        fn = {"c1": run_c1, "c3": run_c3, "c4": run_c4, "c5": run_c5, "unit_cell": run_unit_cell,
              "copy": run_copy}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return
    total = args.warmup + 2 * args.steps  # timed run + profiled rerun
    sc = synth.grid3d_cpml_tfsf_dft(n=args.size, nsteps=total)
    esz = 4 if args.precision == "f32" else 8
    g = apply(Grid("ThreeD", args.size, args.size, args.size, sc["dx"], sc["dt"], precision=args.precision), sc)
    if args.tile:
        tile = tuple(int(v) for v in args.tile.split(","))
        g.set_tile(*tile)
    else:
        tile = g.autotune(trial_steps=30)
    if args.graph:
        g.set_graph(True)
    g.step(args.warmup)
    g.synchronize()
    torch.cuda.synchronize()
    ms_total, avg, per_step = timed_steps(g, args.steps)
    ms_step = ms_total / args.steps
    abytes, interior = algorithmic_bytes(sc, esz)
    gcells = interior * args.steps / (ms_total / 1e3) / 1e9
    # device time of the kernels of one step (the aux line runs once per 256-step chunk)
    kern_ms = sum(v for k, v in per_step.items() if k != "aux_line")
    achieved = abytes / (kern_ms / 1e3) / 1e9
    traffic, traffic_src = (step_traffic(tile, per_step, steps=args.steps) if (args.size == 256 and args.precision == "f32")
                            else (None, None))
    out = {
        "metric": METRIC, "value": round(gcells, 4), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"3-D {args.size}^3 {args.precision.upper()} forward: 12-cell CPML + 48 eps blocks "
                               f"+ TFSF box inset 20 (x-propagating Ez plane wave) + running DFT of 6 comps on "
                               f"2 z-planes x 8 freqs",
                   "cells": args.size ** 3, "interior_cells": interior, "tile": list(tile),
                   "graph_chunks": bool(args.graph)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "one step: " + " + ".join(k for k in avg if k != "aux_line"),
                     "algorithmic_bytes_per_step": abytes, "avg_ms": avg, "ms_per_step_by_kernel": per_step},
    }
    g.close()
    copy_gbs = copy_rate()
    out["roofline"]["copy_GBs"] = round(copy_gbs, 1)
    out["roofline"]["frac_of_copy"] = round(achieved / copy_gbs, 4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.size, args.cpu_sample_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
