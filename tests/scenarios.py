"""Scenario descriptions shared by the oracle tests and the GPU parity tests.

A scenario is a plain dict.  ``apply(sim, sc)`` drives any object exposing the
common set-up methods (the oracle's ``Sim`` and the product's ``Grid`` both
do): set_pml_idiom, set_material, set_matid, add_source, set_tfsf, add_probe,
add_dft.
"""
import math

import numpy as np

from golden_io import load

EPS0 = 8.854e-12
C0 = 2.99792458e8
MU0 = math.pi * 4.0e-7
IMP0 = math.sqrt(MU0 / EPS0)

# field codes used by both sims
D, E, B, H, P, J = 0, 1, 2, 3, 4, 5


class SplitMix64:
    """Same generator as the fixture harness (oracle/ref_harness/harness.cpp)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo, hi):
        return lo + int(self.next() % (hi - lo + 1))


def ref_material(kind, eps, freq, gamma, delta, bias=None):
    """Map a reference material (kind 0 vacuum, 1 constant, 2 Drude, 3 Lorentz,
    4 conductive with freq = the conductivity, 5 magnetized Drude with `bias`,
    6 self-magnetized Drude, 7 / 8 / 9 fluid Drude without magnetization / in
    `bias` / self-magnetized) to (eps_r, mu_r, poles), for kinds 5 and 6 to
    (eps_r, mu_r, poles, bias) with bias None for the cell's own B, for kinds
    7-9 to {"fluid": magnetization, "mu": 1, "bias": ...}.  The reference's
    Lorentz update never reads its delta (include/DispersiveMaterials.hpp:711-735),
    so its pole weight is 1."""
    kind = int(kind)
    if kind in (7, 8, 9):
        return {"fluid": kind - 7, "mu": 1.0, "bias": list(bias) if kind == 8 else None}
    if kind in (5, 6):
        return (eps, 1.0, [(1, freq, gamma, 1.0)], list(bias) if kind == 5 else None)
    if kind == 4:
        return (eps, 1.0, [(3, freq, 0.0, 1.0)])
    if kind == 0:
        return (1.0, 1.0, [])
    if kind == 1:
        return (eps, 1.0, [])
    if kind == 2:
        return (eps, 1.0, [(1, freq, gamma, 1.0)])
    return (eps, 1.0, [(2, freq, gamma, 1.0)])


def ref_magnetic(kind, mu, freq, gamma, delta):
    """Map the magnetic side of a reference material (kind 0 vacuum, 1 constant,
    2 Drude, 3 Lorentz, 4 conductive with freq = the magnetic conductivity) to
    (mu_r, poles)."""
    kind = int(kind)
    if kind == 0:
        return 1.0, []
    if kind == 1:
        return mu, []
    if kind == 4:
        return mu, [(3, freq, 0.0, 1.0)]
    return mu, [(1 if kind == 2 else 2, freq, gamma, 1.0)]


def set_ref_material(sim, mid, spec):
    """Apply a ref_material spec to a Grid or an oracle Sim."""
    if isinstance(spec, dict):
        sim.set_fluid_drude(mid, spec["mu"], spec["fluid"], spec["bias"])
    elif len(spec) == 4:
        eps, mu, poles, bias = spec
        sim.set_magnetized_drude(mid, eps, mu, poles[0][1], poles[0][2], bias)
    else:
        sim.set_material(mid, *spec)


def from_golden(name):
    p, a = load(name)
    mode = p["mode"]
    nx, ny = int(p["nx"]), int(p["ny"])
    nz = int(p.get("nz", 1))
    sc = {
        "name": name, "mode": mode, "nx": nx, "ny": ny, "nz": nz,
        "dx": p["dx"], "dt": p["dt"], "nsteps": int(p["nsteps"]),
        "pml": dict(n=int(p["pml_n"]), m=p["pml_m"], ma=p["pml_ma"], smax=p["pml_smax"],
                    kmax=p["pml_kmax"], amax=p["pml_amax"]),
        "pml_axes": [bool(p.get("pmlx", 1)), bool(p.get("pmly", 1)), mode == "ThreeD"],
        "materials": [ref_material(*row, bias=a["mats_bias"][m] if "mats_bias" in a else None)
                      for m, row in enumerate(a["mats_kind_eps_freq_gamma_delta"])],
        "magnetic": [],
        "sources": [], "probes": [], "dfts": [], "tfsf": None,
    }
    shape = (nz if mode == "ThreeD" else 1, ny, nx)
    if "matid" in a:
        sc["matid"] = a["matid"].reshape(shape).astype(np.uint8)
    else:
        sc["matid"] = np.zeros(shape, np.uint8)
    if "theta_x" in p:  # Bloch-periodic x and y: e^{-/+ i theta} at the min / max face
        sc["bloch"] = [(ax, side, complex(math.cos(t), math.sin(t)) if side else complex(math.cos(t), -math.sin(t)))
                       for ax, t in ((0, p["theta_x"]), (1, p["theta_y"])) for side in (0, 1)]
    if "fluid_ne" in a:
        sc["fluid"] = (a["fluid_ne"].reshape(shape), a["fluid_nu"].reshape(shape))
    if int(p.get("src_i", -1)) >= 0:
        sc["sources"].append((D, 2, int(p["src_i"]), int(p["src_j"]), int(p.get("src_k", 0)), a["src"]))
    if "bsrc" in a:
        c, i, j, k = (int(v) for v in a["bsrc_comp_i_j_k"])
        sc["sources"].append((B, c, i, j, k, a["bsrc"]))
    if "mmats_kind_mu_freq_gamma_delta" in a:  # magnetic sides: mu_r into the material, poles beside it
        for mid, row in enumerate(a["mmats_kind_mu_freq_gamma_delta"]):
            mu, poles = ref_magnetic(*row)
            eps, _, epoles = sc["materials"][mid]
            sc["materials"][mid] = (eps, mu, epoles)
            if poles:
                sc["magnetic"].append((mid, poles))
    if "src_field_comp_i_j" in a:  # D / B point sources (TE fixtures): field 0 D, 2 B
        for row, tab in zip(a["src_field_comp_i_j"], a["src_tables"]):
            f, c, i, j = (int(v) for v in row)
            sc["sources"].append((f, c, i, j, 0, tab))
    if "soft_field_comp_i_j" in a:
        for row, tab in zip(a["soft_field_comp_i_j"], a["soft_tables"]):
            f, c, i, j = (int(v) for v in row)
            sc["sources"].append((f, c, i, j, 0, tab))
    if "planes_axis_side_layer_lo_hi_t0" in a:
        kx, ky = p["kx"], p["ky"]
        k = np.array([kx, ky]) / math.hypot(kx, ky)
        sc["planes"] = []
        for q, row in enumerate(a["planes_axis_side_layer_lo_hi_t0"]):
            axis, side, layer, lo, hi = (int(v) for v in row[:5])
            cells = [layer, layer + 1, lo, hi + 1, 0, 1] if axis == 0 else [lo, hi + 1, layer, layer + 1, 0, 1]
            sc["planes"].append(dict(axis=axis, side=side, cells=cells, kvec=[kx, ky, 0.0], pol=[0.0, 0.0, 1.0],
                                     polH=[k[1], -k[0], 0.0], origin=[p["cx"], p["cy"], 0.0],
                                     aux_ncells=int(p["aux_ncells"]), signal=a[f"plane{q}_signal"]))
    if "probes" in a:
        for pr in a["probes"]:
            ijk = [int(v) for v in pr] + [0] * (3 - len(pr))
            if mode == "ThreeD":
                for f, c in [(E, 0), (E, 1), (E, 2), (H, 0), (H, 1), (H, 2)]:
                    sc["probes"].append((f, c, *ijk))
            elif mode == "TE":
                for f, c in [(E, 0), (E, 1), (H, 2)]:
                    sc["probes"].append((f, c, *ijk))
            elif mode == "TEM":
                for f, c in [(E, 2), (H, 1)]:
                    sc["probes"].append((f, c, *ijk))
            else:
                for f, c in [(E, 2), (H, 0), (H, 1)]:
                    sc["probes"].append((f, c, *ijk))
    if int(p.get("tfsf", 0)) and mode == "TEM":  # a pair of planes along x; the reference's TEM defaults
        sc["tfsf"] = dict(box=[int(p["bx0"]), int(p["bx1"]), 0, 0, 0, 0], kvec=[1.0, 0.0, 0.0], pol=[0.0, 0.0, 1.0],
                          polH=[0.0, 1.0, 0.0], origin=[p["cx"], 0.0, 0.0], aux_ncells=int(p["aux_ncells"]),
                          signal=a["aux_signal"])
    elif int(p.get("tfsf", 0)):
        kx, ky = p["kx"], p["ky"]
        sc["tfsf"] = dict(
            box=[int(p["bx0"]), int(p["bx1"]), int(p["by0"]), int(p["by1"]), 0, 0],
            kvec=[kx, ky, 0.0],
            # the reference's TM defaults, include/Source.hpp:577-580
            pol=[0.0, 0.0, 1.0], polH=None,
            origin=[p["cx"], p["cy"], 0.0], aux_ncells=int(p["aux_ncells"]), signal=a["aux_signal"])
        k = np.array([kx, ky])
        k = k / math.sqrt(k[0] * k[0] + k[1] * k[1])
        sc["tfsf"]["polH"] = [k[1], -k[0], 0.0]
        if mode == "TE":  # the reference's TE defaults, include/Source.hpp:573-576
            sc["tfsf"]["pol"], sc["tfsf"]["polH"] = [k[1], -k[0], 0.0], [0.0, 0.0, 1.0]
    return sc, p, a


def apply(sim, sc):
    pm = sc["pml"]
    for ax, on in enumerate(sc["pml_axes"]):
        if on:
            sim.set_pml_idiom(ax, pm["n"], pm["m"], pm["ma"], pm["smax"], pm["kmax"], pm["amax"])
    for mid, spec in enumerate(sc["materials"]):
        set_ref_material(sim, mid, spec)
    for mid, poles in sc.get("magnetic", ()):
        sim.set_material_magnetic(mid, poles)
    sim.set_matid(sc["matid"])
    if sc.get("fluid") is not None:
        sim.set_fluid_state(*sc["fluid"])
    for s in sc["sources"]:
        sim.add_source(*s)
    if sc.get("tfsf"):
        t = sc["tfsf"]
        sim.set_tfsf(t["box"], t["kvec"], t["pol"], t["polH"], t["origin"], t["aux_ncells"], t["signal"])
    for t in sc.get("planes", ()):
        sim.add_tfsf_plane(t["axis"], t["side"], t["cells"], t["kvec"], t["pol"], t["polH"], t["origin"],
                           t["aux_ncells"], t["signal"])
    for pr in sc["probes"]:
        sim.add_probe(*pr)
    for d in sc.get("dfts", []):
        sim.add_dft(*d)
    if sc.get("design"):
        sim.set_design(*sc["design"])
    for axis, side, kind in sc.get("boundaries", ()):
        sim.set_boundary(axis, side, kind)
    for axis, side, factor in sc.get("bloch", ()):
        sim.set_bloch_boundary(axis, side, factor)
    return sim


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.reshape(-1))
    num = np.linalg.norm((a - b).reshape(-1))
    if den == 0:
        return num
    return num / den


def gauss_table(nsteps, t0, spread, scale=EPS0):
    n = np.arange(nsteps, dtype=np.float64)
    return scale * np.exp(-((n - t0) / spread) ** 2)


def dipole_table(nsteps, t0, spread, scale=EPS0):
    """Soft-source table of a point dipole whose moment is a Gaussian pulse: the impressed current is the
    moment's time derivative, so the values sum to zero and the source leaves no static charge behind
    (a Gaussian added to D would leave a net charge and its static Coulomb field)."""
    n = np.arange(nsteps, dtype=np.float64)
    u = (n - t0) / spread
    return -scale * math.sqrt(2.0 * math.e) * u * np.exp(-u * u)


def gmsin_signal(nsteps, dt, f0, tw, t0):
    """Gaussian-modulated sinusoid sampled at the aux line's accumulated time
    (t += dt before each use, include/Source.hpp:414)."""
    out = np.zeros(nsteps)
    t = 0.0
    for n in range(nsteps):
        t += dt
        a = (t - t0) / tw
        out[n] = math.exp(-a * a) * math.sin(2 * math.pi * f0 * (t - t0))
    return out


def cw_signal(nsteps, dt, f0, ramp_periods=3.0):
    """sin(2 pi f0 t) switched on by 1 - exp(-(t/tau)^2), tau = ramp_periods / f0,
    at the aux line's accumulated time (t += dt before each use)."""
    out = np.zeros(nsteps)
    tau = ramp_periods / f0
    t = 0.0
    for n in range(nsteps):
        t += dt
        out[n] = (1.0 - math.exp(-(t / tau) ** 2)) * math.sin(2 * math.pi * f0 * t)
    return out


def dft_phases(nsteps, dt, freqs):
    """[nsteps][nfreq][2] phase tables (cos, -sin)(w t) * dt for E (t=(n+1)dt) and H (t=(n+1.5)dt)."""
    n = np.arange(nsteps, dtype=np.float64)[:, None]
    w = 2 * np.pi * np.asarray(freqs, dtype=np.float64)[None, :]
    tE = (n + 1.0) * dt
    tH = (n + 1.5) * dt
    pe = np.stack([np.cos(w * tE) * dt, -np.sin(w * tE) * dt], axis=-1)
    ph = np.stack([np.cos(w * tH) * dt, -np.sin(w * tH) * dt], axis=-1)
    return pe, ph
