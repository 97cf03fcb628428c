"""Seeded synthetic scenarios of the benchmark shapes, scaled by size.

grid3d_cpml_tfsf_dft     3-D CPML + dielectric blocks + TFSF plane wave + DFT planes
                         (256^3 x 2000 steps at full size)
grid3d_dispersive_sphere 3-D Drude + 2-Lorentz sphere, z-dipole source (Gaussian moment), DFT box surface
                         (192^3 x 3000 steps at full size)
tm2d_c1                  2-D TM 256^2, 10-cell CPML, centre Gaussian pulse, 4 probes
batch_c4                 64 uneven 3-D grids (edges 64..160), z-dipole point source, 4-frequency plane each
grid3d_adjoint           3-D CPML + TFSF continuous wave, per-cell eps design box, Ez DFT plane
                         at one frequency (128^3, design 64x64x16, 3000 steps at full size)
"""
import math

import numpy as np

from scenarios import B, C0, D, E, EPS0, IMP0, H, SplitMix64, cw_signal, dft_phases, dipole_table, gauss_table, gmsin_signal

DRUDE, LORENTZ = 1, 2


def _pml(n, dx):
    return dict(n=n, m=3.0, ma=1.0, smax=0.8 * (3.0 + 1) / (IMP0 * dx), kmax=5.0, amax=0.05)


def _dims(n):
    return (n, n, n) if isinstance(n, int) else tuple(n)


def seed_blocks(matid, nblocks, lo_margin, rng, side_lo, side_hi, eps_lo=1.0, eps_hi=12.0, first_id=1):
    """Axis-aligned blocks inside the non-PML interior; later blocks override earlier."""
    nz, ny, nx = matid.shape
    mats = []
    for b in range(nblocks):
        s = [rng.randint(side_lo, side_hi) for _ in range(3)]
        lo = []
        for ax, L in enumerate((nx, ny, nz)):
            hi = max(lo_margin, L - lo_margin - s[ax])
            lo.append(rng.randint(lo_margin, hi))
        eps = eps_lo + (eps_hi - eps_lo) * rng.uniform()
        mid = first_id + b
        mats.append((eps, 1.0, []))
        x1 = min(nx - lo_margin, lo[0] + s[0])
        y1 = min(ny - lo_margin, lo[1] + s[1])
        z1 = min(nz - lo_margin, lo[2] + s[2])
        matid[lo[2]:z1, lo[1]:y1, lo[0]:x1] = mid
    return mats


def grid3d_cpml_tfsf_dft(n=256, nsteps=2000, npml=None, nblocks=None, nfreq=None, inset=None, seed=2,
                         tfsf=True, dft=True):
    nx, ny, nz = _dims(n)
    small = min(nx, ny, nz) < 128
    npml = npml if npml is not None else (5 if small else 12)
    nblocks = nblocks if nblocks is not None else (6 if small else 48)
    nfreq = nfreq if nfreq is not None else (4 if small else 8)
    inset = inset if inset is not None else (max(npml + 3, min(nx, ny, nz) // 6) if small else 20)
    dx = 1e-6
    dt = 0.5 * dx / (C0 * math.sqrt(3.0))
    rng = SplitMix64(seed)
    matid = np.zeros((nz, ny, nx), np.uint8)
    side_hi = max(3, min(nx, ny, nz) // 4)
    blocks = seed_blocks(matid, nblocks, npml + 2, rng, max(2, side_hi // 8), side_hi)
    sc = {
        "name": f"grid3d_cpml_tfsf_dft_{nx}x{ny}x{nz}", "mode": "ThreeD", "nx": nx, "ny": ny, "nz": nz,
        "dx": dx, "dt": dt, "nsteps": nsteps, "pml": _pml(npml, dx), "pml_axes": [True, True, True],
        "materials": [(1.0, 1.0, [])] + blocks, "matid": matid, "sources": [], "probes": [], "dfts": [],
        "tfsf": None,
    }
    if tfsf:
        box = [inset, nx - 1 - inset, inset, ny - 1 - inset, inset, nz - 1 - inset]
        f0 = C0 / (20 * dx)
        tw = 2.0 / f0
        t0 = 3.0 * tw
        origin = [math.floor(0.5 * (box[0] + box[1])) + 0.5, math.floor(0.5 * (box[2] + box[3])) + 0.5,
                  math.floor(0.5 * (box[4] + box[5])) + 0.5]
        ext = max(box[1] - box[0], box[3] - box[2], box[5] - box[4]) + 12
        ext += ext % 2
        sc["tfsf"] = dict(box=box, kvec=[1.0, 0.0, 0.0], pol=[0.0, 0.0, 1.0],
                          polH=[0.0, 1.0, 0.0],  # pol x k: consistent with the aux line's (Ez, Hy)
                          origin=origin, aux_ncells=ext, signal=gmsin_signal(nsteps, dt, f0, tw, t0))
    else:
        sc["sources"].append((D, 2, nx // 2, ny // 2, nz // 2, gauss_table(nsteps, 40, 12)))
    if dft:
        freqs = np.linspace(0.6, 1.4, nfreq) * C0 / (20 * dx)
        pe, ph = dft_phases(nsteps, dt, freqs)
        for kz in (nz // 3, (2 * nz) // 3):
            sc["dfts"].append(([1, nx - 1, 1, ny - 1, kz, kz + 1], 63, pe, ph))
    sc["probes"] = [(E, 2, nx // 2, ny // 2, nz // 2), (H, 1, nx // 2 + 1, ny // 2, nz // 2)]
    return sc


def grid3d_dispersive_sphere(n=192, radius=40, nsteps=3000, npml=None, nfreq=None, seed=3):
    nx, ny, nz = _dims(n)
    small = min(nx, ny, nz) < 96
    npml = npml if npml is not None else (5 if small else 12)
    nfreq = nfreq if nfreq is not None else (4 if small else 16)
    dx = 1e-8  # 10 nm cells: omega_p dt ~ 0.13
    dt = 0.5 * dx / (C0 * math.sqrt(3.0))
    rng = SplitMix64(seed)
    w0 = [2e15 + 4e15 * rng.uniform() for _ in range(2)]
    poles = [(DRUDE, 1.37e16, 1e14, 1.0), (LORENTZ, w0[0], 1e14, 1.5), (LORENTZ, w0[1], 1e14, 0.8)]
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    inside = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= radius * radius
    matid = np.where(inside, 1, 0).astype(np.uint8)
    src_x = min(nx - npml - 3, cx + radius + 3)
    sc = {
        "name": f"grid3d_dispersive_sphere_{nx}x{ny}x{nz}", "mode": "ThreeD", "nx": nx, "ny": ny, "nz": nz,
        "dx": dx, "dt": dt, "nsteps": nsteps, "pml": _pml(npml, dx), "pml_axes": [True, True, True],
        "materials": [(1.0, 1.0, []), (1.0, 1.0, poles)], "matid": matid,
        # a z-dipole whose Gaussian moment (spread 49 steps) puts its spectrum across the monitored band
        "sources": [(D, 2, src_x, cy, cz, dipole_table(nsteps, 200, 49))], "probes": [], "dfts": [], "tfsf": None,
    }
    freqs = np.linspace(1e14, 1.2e15, nfreq)
    pe, ph = dft_phases(nsteps, dt, freqs)
    r = radius + 2
    b = [max(1, cx - r), min(nx - 1, cx + r + 1), max(1, cy - r), min(ny - 1, cy + r + 1), max(1, cz - r),
         min(nz - 1, cz + r + 1)]
    faces = [
        [b[0], b[0] + 1, b[2], b[3], b[4], b[5]], [b[1] - 1, b[1], b[2], b[3], b[4], b[5]],
        [b[0], b[1], b[2], b[2] + 1, b[4], b[5]], [b[0], b[1], b[3] - 1, b[3], b[4], b[5]],
        [b[0], b[1], b[2], b[3], b[4], b[4] + 1], [b[0], b[1], b[2], b[3], b[5] - 1, b[5]],
    ]
    for f in faces:
        sc["dfts"].append((f, 63, pe, ph))
    sc["probes"] = [(E, 2, cx, cy, cz), (E, 2, src_x, cy, cz)]
    return sc


def tm2d_c1(nsteps=1000):
    """2-D TM 256^2, FP64 in the reference, dx = 1 m, c0 dt/dx = 0.5, 10-cell CPML,
    Dz[128,128] += eps0 exp(-((n-40)/12)^2), 4 seeded probes in the non-PML interior."""
    nx = ny = 256
    dx = 1.0
    dt = 0.5 * dx / C0
    rng = SplitMix64(1)
    probes = [(rng.randint(12, 243), rng.randint(12, 243)) for _ in range(4)]
    return {
        "name": "tm2d_c1", "mode": "TM", "nx": nx, "ny": ny, "nz": 1, "dx": dx, "dt": dt, "nsteps": nsteps,
        "pml": _pml(10, dx), "pml_axes": [True, True, False], "materials": [(1.0, 1.0, [])],
        "matid": np.zeros((1, ny, nx), np.uint8), "sources": [(D, 2, 128, 128, 0, gauss_table(nsteps, 40, 12))],
        "probes": [(E, 2, i, j, 0) for i, j in probes], "dfts": [], "tfsf": None,
    }


def batch_c4(nsims=64, nsteps=1500, lo=64, hi=160, seed=4):
    """64 independent 3-D FP32 grids, edges seeded uniform in [64,160] per axis,
    8-cell CPML, 10 eps blocks U[1,12], own z-dipole point source, 4-frequency DFT
    of all components on the mid z-plane."""
    rng = SplitMix64(seed)
    out = []
    dx = 1e-6
    dt = 0.5 * dx / (C0 * math.sqrt(3.0))
    freqs = np.linspace(0.7, 1.3, 4) * C0 / (20 * dx)
    pe, ph = dft_phases(nsteps, dt, freqs)
    for q in range(nsims):
        nx, ny, nz = (rng.randint(lo, hi) for _ in range(3))
        matid = np.zeros((nz, ny, nx), np.uint8)
        side_hi = max(3, min(nx, ny, nz) // 4)
        blocks = seed_blocks(matid, 10, 8 + 2, rng, 2, side_hi)
        sx, sy, sz = rng.randint(10, nx - 11), rng.randint(10, ny - 11), rng.randint(10, nz - 11)
        out.append({
            "name": f"batch{q}_{nx}x{ny}x{nz}", "mode": "ThreeD", "nx": nx, "ny": ny, "nz": nz, "dx": dx, "dt": dt,
            "nsteps": nsteps, "pml": _pml(8, dx), "pml_axes": [True, True, True],
            "materials": [(1.0, 1.0, [])] + blocks, "matid": matid,
            "sources": [(D, 2, sx, sy, sz, dipole_table(nsteps, 40 + q % 7, 10))], "probes": [],
            "dfts": [([1, nx - 1, 1, ny - 1, nz // 2, nz // 2 + 1], 63, pe, ph)], "tfsf": None,
        })
    return out


def grid3d_adjoint(n=128, nsteps=3000, design=(64, 64, 16), npml=10, monitor=64, seed=5, poles=False):
    """Inverse-design gradient workload: x-propagating, Ez-polarised continuous
    wave (TFSF, one frequency f0 = c0/(20 dx)) through a design box of per-cell
    eps_r seeded U[1,12] centred in the grid; objective sum |F[Ez](f0)|^2 over a
    monitor x-plane of monitor x monitor cells behind the design box.
    poles: pole materials on the wave's path: a Lorentz block before the design
    box, a Drude slab between it and the monitor, a conductive + Lorentz block
    beside it."""
    nx = ny = nz = n
    dx = 1e-6
    dt = 0.5 * dx / (C0 * math.sqrt(3.0))
    rng = SplitMix64(seed)
    wx, wy, wz = design
    x0, y0, z0 = (nx - wx) // 2, (ny - wy) // 2, (nz - wz) // 2
    dbox = [x0, x0 + wx, y0, y0 + wy, z0, z0 + wz]
    eps = np.array([1.0 + 11.0 * rng.uniform() for _ in range(wx * wy * wz)]).reshape(wz, wy, wx)
    inset = npml + 2
    box = [inset, nx - 1 - inset, inset, ny - 1 - inset, inset, nz - 1 - inset]
    f0 = C0 / (20 * dx)
    origin = [math.floor(0.5 * (box[0] + box[1])) + 0.5, math.floor(0.5 * (box[2] + box[3])) + 0.5,
              math.floor(0.5 * (box[4] + box[5])) + 0.5]
    ext = max(box[1] - box[0], box[3] - box[2], box[5] - box[4]) + 12
    ext += ext % 2
    mx = min(box[1] - 1, x0 + wx + max(1, (box[1] - x0 - wx) // 2))
    m = min(monitor, ny - 2 * inset - 2, nz - 2 * inset - 2)
    my0, mz0 = (ny - m) // 2, (nz - m) // 2
    pe, ph = dft_phases(nsteps, dt, [f0])
    mats = [(1.0, 1.0, [])]
    matid = np.zeros((nz, ny, nx), np.uint8)
    if poles:
        w0 = 2 * math.pi * f0
        mats += [(1.0, 1.0, [(DRUDE, 1.2 * w0, 0.1 * w0, 1.0)]), (2.0, 1.0, [(LORENTZ, 0.8 * w0, 0.05 * w0, 1.0)]),
                 (1.5, 1.0, [(3, 0.05 * EPS0 / dt, 0.0, 1.0), (LORENTZ, 1.3 * w0, 0.08 * w0, 1.0)])]
        matid[z0 - 1:z0 + wz + 1, y0 - 1:y0 + wy + 1, x0 + wx:mx] = 1       # Drude slab before the monitor
        matid[z0:z0 + wz, y0:y0 + wy, max(inset, x0 - 3):x0 - 1] = 2       # Lorentz block before the design box
        matid[z0:z0 + wz, y0 + wy + 1:y0 + wy + 3, x0:x0 + wx] = 3         # conductive + Lorentz beside it
    return {
        "name": f"grid3d_adjoint_{n}", "mode": "ThreeD", "nx": nx, "ny": ny, "nz": nz, "dx": dx, "dt": dt,
        "nsteps": nsteps, "pml": _pml(npml, dx), "pml_axes": [True, True, True], "materials": mats,
        "matid": matid, "sources": [], "probes": [],
        "tfsf": dict(box=box, kvec=[1.0, 0.0, 0.0], pol=[0.0, 0.0, 1.0], polH=[0.0, 1.0, 0.0], origin=origin,
                     aux_ncells=ext, signal=cw_signal(nsteps, dt, f0)),
        "dfts": [([mx, mx + 1, my0, my0 + m, mz0, mz0 + m], 4, pe, ph)],
        "design": (dbox, eps), "f0": f0,
    }


def grid3d_boundaries(n=(30, 26, 22), nsteps=120, variant="closed", nfreq=4, seed=6, npml=5):
    """3-D grids with boundary conditions (fdtd_set_boundary).

    variant "closed": x PEC (min) / PMC (max), y Symmetric / Antisymmetric, z
    periodic; no CPML.  variant "unit_cell": x and y periodic (a metasurface
    unit cell), `npml`-cell CPML along z only, a DFT plane above the blocks.  Both:
    seeded dielectric blocks, a Dz and a Bx point source off the symmetry
    planes, probes next to every face."""
    nx, ny, nz = _dims(n)
    dx = 1e-6
    dt = 0.5 * dx / (C0 * math.sqrt(3.0))
    rng = SplitMix64(seed)
    matid = np.zeros((nz, ny, nx), np.uint8)
    blocks = seed_blocks(matid, 4, 3, rng, 2, max(3, min(nx, ny, nz) // 4))
    if variant == "closed":
        bcs = [(0, 0, "PEC"), (0, 1, "PMC"), (1, 0, "Symmetric"), (1, 1, "Antisymmetric"), (2, 0, "Periodic"),
               (2, 1, "Periodic")]
        pml_axes = [False, False, False]
        npml = 0
    else:
        bcs = [(0, 0, "Periodic"), (0, 1, "Periodic"), (1, 0, "Periodic"), (1, 1, "Periodic")]
        pml_axes = [False, False, True]
    sc = {
        "name": f"grid3d_boundaries_{variant}", "mode": "ThreeD", "nx": nx, "ny": ny, "nz": nz, "dx": dx, "dt": dt,
        "nsteps": nsteps, "pml": _pml(max(npml, 1), dx), "pml_axes": pml_axes,
        "materials": [(1.0, 1.0, [])] + blocks, "matid": matid,
        "sources": [(D, 2, nx // 3, ny // 2 + 1, nz // 2, gauss_table(nsteps, 20, 6)),
                    (B, 0, 2 * nx // 3, ny // 3, nz // 3, gauss_table(nsteps, 30, 8, scale=4e-7 * math.pi))],
        "probes": [(E, 2, 2, ny // 2, nz // 2), (E, 1, nx - 2, ny // 2, nz // 2), (H, 0, nx // 2, 1, nz // 2),
                   (H, 2, nx // 2, ny - 2, nz // 2), (E, 0, nx // 2, ny // 2, 1), (H, 1, nx // 2, ny // 2, nz - 2)],
        "dfts": [], "tfsf": None, "boundaries": bcs,
    }
    if variant != "closed":
        freqs = np.linspace(0.7, 1.3, nfreq) * C0 / (20 * dx)
        pe, ph = dft_phases(nsteps, dt, freqs)
        sc["dfts"].append(([1, nx - 1, 1, ny - 1, nz - npml - 3, nz - npml - 2], 63, pe, ph))
    return sc


def tm2d_boundaries(n=(48, 40), nsteps=200):
    """2-D TM with x periodic and y PEC (min) / Symmetric (max), 8-cell CPML
    nowhere (closed in y by the conditions), a dielectric block and a Dz source."""
    nx, ny = n
    dx = 1e-6
    dt = 0.5 * dx / (C0 * math.sqrt(2.0))
    matid = np.zeros((1, ny, nx), np.uint8)
    matid[0, 12:20, 10:18] = 1
    return {
        "name": "tm2d_boundaries", "mode": "TM", "nx": nx, "ny": ny, "nz": 1, "dx": dx, "dt": dt, "nsteps": nsteps,
        "pml": _pml(1, dx), "pml_axes": [False, False, False], "materials": [(1.0, 1.0, []), (4.0, 1.0, [])],
        "matid": matid, "sources": [(D, 2, nx // 3, ny // 2, 0, gauss_table(nsteps, 25, 8))],
        "probes": [(E, 2, 1, ny // 2, 0), (E, 2, nx - 2, ny // 3, 0), (H, 0, nx // 2, 1, 0), (H, 1, nx // 2, ny - 2, 0)],
        "dfts": [], "tfsf": None,
        "boundaries": [(0, 0, "Periodic"), (0, 1, "Periodic"), (1, 0, "PEC"), (1, 1, "Symmetric")],
    }


def tm2d_bloch(n=(40, 44), nsteps=150, theta=1.1):
    """2-D TM unit cell: x Bloch-periodic (e^{-/+ i theta} at the min / max face),
    10-cell CPML along y, a dielectric block, a Dz source and a 2-frequency DFT
    line across the cell."""
    nx, ny = n
    dx = 1e-6
    dt = 0.5 * dx / (C0 * math.sqrt(2.0))
    matid = np.zeros((1, ny, nx), np.uint8)
    matid[0, 14:22, 6:19] = 1
    f = complex(math.cos(theta), math.sin(theta))
    return {
        "name": "tm2d_bloch", "mode": "TM", "nx": nx, "ny": ny, "nz": 1, "dx": dx, "dt": dt, "nsteps": nsteps,
        "pml": _pml(10, dx), "pml_axes": [False, True, False], "materials": [(1.0, 1.0, []), (2.25, 1.0, [])],
        "matid": matid, "sources": [(D, 2, nx // 3, ny // 2, 0, gauss_table(nsteps, 25, 8))],
        "probes": [(E, 2, 1, ny // 2, 0), (E, 2, nx - 2, ny // 3, 0), (H, 0, nx // 2, 12, 0), (H, 1, 1, ny - 13, 0)],
        "dfts": [_dft_line(nx, ny // 2 + 5, dt, nsteps)], "tfsf": None,
        "bloch": [(0, 0, f.conjugate()), (0, 1, f)],
    }


def _dft_line(nx, j, dt, nsteps):
    """DFT monitor of Ez, Hx, Hy on row j at two frequencies (region, mask, phE, phH)."""
    return dft_monitor([1, nx - 1, j, j + 1, 0, 1], (1 << 2) | (1 << 3) | (1 << 4), dt, nsteps)


def dft_monitor(region, mask, dt, nsteps, freqs=(0.02, 0.035)):
    """(region, mask, phE, phH) of a running DFT at the given frequencies (units of
    1/dt): E sampled at (n+1) dt, H at (n+1.5) dt, weights e^{-i w t} dt."""
    f = np.asarray(freqs) / dt
    n = np.arange(nsteps)
    phE = np.stack([np.cos(2 * np.pi * f[None, :] * (n[:, None] + 1) * dt) * dt,
                    -np.sin(2 * np.pi * f[None, :] * (n[:, None] + 1) * dt) * dt], -1)
    phH = np.stack([np.cos(2 * np.pi * f[None, :] * (n[:, None] + 1.5) * dt) * dt,
                    -np.sin(2 * np.pi * f[None, :] * (n[:, None] + 1.5) * dt) * dt], -1)
    return (list(region), mask, phE, phH)


def _plane(axis, side, cells, kvec, origin, aux_ncells, signal, pol=(0.0, 0.0, 1.0)):
    """A single TFSF plane's arguments; polH = pol x k (the aux line's Ez, Hy pair)."""
    k = np.asarray(kvec, float)
    k = k / np.linalg.norm(k)
    return dict(axis=axis, side=side, cells=list(cells), kvec=list(kvec), pol=list(pol),
                polH=list(np.cross(np.asarray(pol, float), k)), origin=list(origin), aux_ncells=aux_ncells,
                signal=signal)


def grid3d_sources_planes(n=(40, 30, 26), nsteps=80, npml=5):
    """3-D: soft sources on Ez (inside a Drude block), Ex and Hy (added after the
    material steps) beside one on Dz, and two single TFSF planes (x-MIN, y-MAX
    faces) under an oblique wave, each plane with its own line and pulse
    delay; CPML on every face, a dielectric block, a DFT monitor and probes."""
    nx, ny, nz = n
    dx = 1e-8
    dt = 0.5 * dx / (C0 * math.sqrt(3.0))
    matid = np.zeros((nz, ny, nx), np.uint8)
    matid[10:16, 17:23, 20:27] = 1
    matid[14:20, 8:13, 12:17] = 2
    mats = [(1.0, 1.0, []), (3.0, 1.0, []), (1.5, 1.0, [(1, 1.37e16, 1e14, 1.0)])]
    f0 = C0 / (20 * dx)
    tw = 2.0 / f0
    kvec = [math.cos(0.5), math.sin(0.5) * math.cos(0.3), math.sin(0.5) * math.sin(0.3)]
    origin = [20.5, 15.5, 13.5]
    planes = [_plane(0, 0, [8, 9, 8, 22, 7, 19], kvec, origin, 90, gmsin_signal(nsteps, dt, f0, tw, 3.0 * tw)),
              _plane(1, 1, [9, 33, 22, 23, 7, 19], kvec, origin, 90, gmsin_signal(nsteps, dt, f0, tw, 4.0 * tw))]
    sources = [(E, 2, 14, 10, 16, gauss_table(nsteps, 20, 6, scale=1.0)),
               (E, 0, 28, 20, 9, gauss_table(nsteps, 30, 7, scale=1.0)),
               (H, 1, 17, 24, 12, gauss_table(nsteps, 25, 6, scale=1.0 / IMP0)),
               (D, 2, 30, 8, 18, gauss_table(nsteps, 15, 5))]
    probes = [(E, 2, 14, 10, 16), (E, 0, 28, 20, 9), (H, 1, 17, 24, 12), (E, 2, 8, 15, 13), (E, 2, 7, 15, 13),
              (H, 2, 20, 22, 12), (E, 1, 20, 23, 12)]
    pe, ph = dft_phases(nsteps, dt, np.array([0.8, 1.1]) * f0)
    return {
        "name": "grid3d_sources_planes", "mode": "ThreeD", "nx": nx, "ny": ny, "nz": nz, "dx": dx, "dt": dt,
        "nsteps": nsteps, "pml": _pml(npml, dx), "pml_axes": [True, True, True], "materials": mats,
        "matid": matid, "sources": sources, "probes": probes, "tfsf": None, "planes": planes,
        "dfts": [([6, 34, 5, 25, 13, 14], 63, pe, ph)],
    }


def grid3d_plane_slab(n=(48, 12, 10), nsteps=140, npml=8, faces=(0, 1), layers=(14, 33)):
    """Vacuum, CPML along x only, periodic y and z; single TFSF planes with
    normal x on the given sides (0 MIN at layers[0], 1 MAX at layers[1]),
    each covering every interior y, z cell, k = x: an x-propagating plane
    wave confined to the total-field slab between them."""
    nx, ny, nz = n
    dx = 1e-6
    dt = 0.5 * dx / (C0 * math.sqrt(3.0))
    f0 = C0 / (20 * dx)
    tw = 2.0 / f0
    sig = gmsin_signal(nsteps, dt, f0, tw, 3.0 * tw)
    origin = [0.5 * (layers[0] + layers[1]), ny / 2, nz / 2]
    planes = [_plane(0, side, [layers[side], layers[side] + 1, 1, ny - 1, 1, nz - 1], [1.0, 0.0, 0.0], origin,
                     nx + 20, sig) for side in faces]
    return {
        "name": "grid3d_plane_slab", "mode": "ThreeD", "nx": nx, "ny": ny, "nz": nz, "dx": dx, "dt": dt,
        "nsteps": nsteps, "pml": _pml(npml, dx), "pml_axes": [True, False, False],
        "materials": [(1.0, 1.0, [])], "matid": np.zeros((nz, ny, nx), np.uint8), "sources": [], "probes": [],
        "dfts": [], "tfsf": None, "planes": planes,
        "boundaries": [(ax, sd, "Periodic") for ax in (1, 2) for sd in (0, 1)],
    }


def te2d_features(n=(52, 44), nsteps=160, npml=8):
    """2-D TE (Ex, Ey, Hz): CPML along x, y closed by PEC (min) and PMC (max);
    a dielectric, a Drude and a Lorentz block; soft sources on Dx, on Ey
    inside the Lorentz block and on Hz (after the material steps) beside one
    on Bz; a single TFSF plane on an x-MIN face under an oblique TE wave
    (pol = (k_y, -k_x, 0), polH = pol x k = z); a DFT monitor of Ex, Ey, Hz."""
    nx, ny = n
    dx = 1e-7
    dt = 0.5 * dx / (C0 * math.sqrt(2.0))
    matid = np.zeros((1, ny, nx), np.uint8)
    matid[0, 10:17, 24:31] = 1
    matid[0, 26:33, 14:21] = 2
    matid[0, 22:28, 32:38] = 3
    mats = [(1.0, 1.0, []), (4.0, 1.0, []), (1.5, 1.0, [(1, 0.3 / dt, 0.02 / dt, 1.0)]),
            (2.0, 1.0, [(2, 0.2 / dt, 0.01 / dt, 0.7)])]
    f0 = C0 / (20 * dx)
    tw = 2.0 / f0
    kvec = [math.cos(0.4), math.sin(0.4), 0.0]
    pol = (kvec[1], -kvec[0], 0.0)
    planes = [_plane(0, 0, [12, 13, 6, 38, 0, 1], kvec, [26.5, 22.5, 0.0], 90,
                     gmsin_signal(nsteps, dt, f0, tw, 3.0 * tw), pol=pol)]
    sources = [(D, 0, 20, 14, 0, gauss_table(nsteps, 20, 6)),
               (E, 1, 34, 24, 0, gauss_table(nsteps, 30, 7, scale=1.0)),
               (H, 2, 17, 29, 0, gauss_table(nsteps, 25, 6, scale=1.0 / IMP0)),
               (B, 2, 40, 12, 0, gauss_table(nsteps, 15, 5, scale=1e-9))]
    probes = [(E, 0, 20, 14, 0), (E, 1, 34, 24, 0), (H, 2, 17, 29, 0), (E, 0, 30, 1, 0), (H, 2, 30, ny - 2, 0),
              (E, 1, 11, 20, 0), (E, 1, 12, 20, 0)]
    pe, ph = dft_phases(nsteps, dt, np.array([0.8, 1.1]) * f0)
    return {
        "name": "te2d_features", "mode": "TE", "nx": nx, "ny": ny, "nz": 1, "dx": dx, "dt": dt, "nsteps": nsteps,
        "pml": _pml(npml, dx), "pml_axes": [True, False, False], "materials": mats, "matid": matid,
        "sources": sources, "probes": probes, "tfsf": None, "planes": planes,
        "dfts": [([4, nx - 4, 3, ny - 3, 0, 1], 1 | 2 | 32, pe, ph)],
        "boundaries": [(1, 0, "PEC"), (1, 1, "PMC")],
    }


def te2d_bloch(n=(40, 44), nsteps=150, theta=0.9):
    """2-D TE unit cell: x Bloch-periodic (e^{-/+ i theta} at the min / max face),
    10-cell CPML along y, a dielectric block, a Dy source and a 2-frequency DFT
    line of Ex, Ey, Hz across the cell."""
    nx, ny = n
    dx = 1e-6
    dt = 0.5 * dx / (C0 * math.sqrt(2.0))
    matid = np.zeros((1, ny, nx), np.uint8)
    matid[0, 14:22, 6:19] = 1
    f = complex(math.cos(theta), math.sin(theta))
    return {
        "name": "te2d_bloch", "mode": "TE", "nx": nx, "ny": ny, "nz": 1, "dx": dx, "dt": dt, "nsteps": nsteps,
        "pml": _pml(10, dx), "pml_axes": [False, True, False], "materials": [(1.0, 1.0, []), (2.25, 1.0, [])],
        "matid": matid, "sources": [(D, 1, nx // 3, ny // 2, 0, gauss_table(nsteps, 25, 8))],
        "probes": [(E, 1, 1, ny // 2, 0), (E, 0, nx - 2, ny // 3, 0), (H, 2, nx // 2, 12, 0), (H, 2, 1, ny - 13, 0)],
        "dfts": [dft_monitor([1, nx - 1, ny // 2 + 5, ny // 2 + 6, 0, 1], 1 | 2 | 32, dt, nsteps)], "tfsf": None,
        "bloch": [(0, 0, f.conjugate()), (0, 1, f)],
    }
