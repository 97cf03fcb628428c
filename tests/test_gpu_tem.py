"""1-D TEM (Ez, Hy along x) on the HIP path, through the C ABI.

- The reference's own TEM runs (tem1d_small: CPML at both ends, dielectric,
  Drude, Lorentz and conductive segments, D and B sources; tem1d_tfsf: a
  TFSF plane pair with the reference's TEM polarisation): every field, D and
  B, and the probe traces against the oracle and the reference's arrays.
- PEC / PMC ends, E / H soft sources, a single TFSF plane, a magnetic-pole
  segment and a DFT monitor against the oracle; plain launches and captured
  chunks give the same bits.
- Refusals: Bloch faces, surface quantities, the adjoint gradient, batches.

Tolerances, relative L2 per component, probes and DFT: FP64 1e-12, FP32 1e-5.
"""
import math

import numpy as np
import pytest

from oracle import refcpu
from scenarios import B, C0, D, E, EPS0, H, IMP0, apply, from_golden, gauss_table, gmsin_signal, dft_phases, rel_l2
import synth

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 1e-5}
COMPS = [("Ez", E, 2), ("Hy", H, 1)]


def _grid(sc, precision):
    from fdtd_amd import Grid
    return apply(Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision), sc)


def _oracle(sc):
    return apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", ["tem1d_small", "tem1d_tfsf"])
def test_tem_reference_runs(name, precision):
    sc, p, a = from_golden(name)
    print(f"[gpu-case] {name} {precision}", flush=True)
    g, o = _grid(sc, precision), _oracle(sc)
    g.step(sc["nsteps"])
    o.step(sc["nsteps"])
    for nm, f, c in COMPS + [("Dz", D, 2), ("By", B, 1)]:
        e = rel_l2(g.get(f, c), o.get(f, c))
        print(nm, f"{e:.2e}", flush=True)
        assert e <= TOL[precision], (nm, e)
        assert rel_l2(g.get(f, c).reshape(a[nm].shape), a[nm]) <= TOL[precision], nm
    assert np.abs(g.get(E, 2)).max() > 0
    pr = g.probes()
    assert rel_l2(pr, o.probes()) <= TOL[precision]
    assert rel_l2(pr, a["probe_trace_Ez_Hy"].reshape(pr.shape)) <= TOL[precision]


def _features(nsteps=300):
    """A 160-cell line: CPML at the max end only (PEC at the min end, then PMC
    on a second run), a dielectric, a Drude and a magnetic Drude segment, soft
    sources on Ez and Hy beside one on Dz, a single TFSF plane (MIN side) and a
    DFT monitor of Ez and Hy."""
    nx = 160
    dx = 1e-6
    dt = 0.9 * dx / C0
    matid = np.zeros((1, 1, nx), np.uint8)
    matid[0, 0, 50:60] = 1
    matid[0, 0, 70:80] = 2
    matid[0, 0, 90:100] = 3
    mats = [(1.0, 1.0, []), (2.25, 1.0, []), (1.5, 1.0, [(1, 0.3 / dt, 0.02 / dt, 1.0)]), (1.0, 1.4, [])]
    f0 = C0 / (20 * dx)
    tw = 2.0 / f0
    plane = dict(axis=0, side=0, cells=[40, 41, 0, 1, 0, 1], kvec=[1.0, 0.0, 0.0], pol=[0.0, 0.0, 1.0],
                 polH=[0.0, 1.0, 0.0], origin=[80.5, 0.0, 0.0], aux_ncells=100,
                 signal=gmsin_signal(nsteps, dt, f0, tw, 3.0 * tw))
    pe, ph = dft_phases(nsteps, dt, np.array([0.8, 1.1]) * f0)
    return {
        "name": "tem1d_features", "mode": "TEM", "nx": nx, "ny": 1, "nz": 1, "dx": dx, "dt": dt, "nsteps": nsteps,
        "pml": synth._pml(10, dx), "pml_axes": [True, False, False], "materials": mats, "matid": matid,
        "magnetic": [(3, [(1, 0.25 / dt, 0.03 / dt, 1.0)])],
        "sources": [(D, 2, 30, 0, 0, gauss_table(nsteps, 30, 8)), (E, 2, 65, 0, 0, gauss_table(nsteps, 40, 9, scale=1.0)),
                    (H, 1, 85, 0, 0, gauss_table(nsteps, 50, 9, scale=1.0 / IMP0))],
        "probes": [(E, 2, 30, 0, 0), (E, 2, 65, 0, 0), (H, 1, 85, 0, 0), (E, 2, 1, 0, 0), (H, 1, 95, 0, 0)],
        "planes": [plane], "tfsf": None,
        "dfts": [([1, nx - 1, 0, 1, 0, 1], (1 << 2) | (1 << 4), pe, ph)],
        "boundaries": [(0, 0, "PEC")],
    }


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("wall", ["PEC", "PMC"])
def test_tem_sources_plane_boundaries_dft(wall, precision):
    sc = _features()
    sc["boundaries"] = [(0, 0, wall)]
    o = _oracle(sc)
    o.step(sc["nsteps"])
    ref = None
    for graph in (False, True):
        print(f"[gpu-case] tem1d_features {wall} {precision} graph {graph}", flush=True)
        g = _grid(sc, precision)
        g.set_graph(graph)
        g.step(77)
        g.step(sc["nsteps"] - 77)
        if ref is None:
            ref = g
            for nm, f, c in COMPS:
                e = rel_l2(g.get(f, c), o.get(f, c))
                assert e <= TOL[precision], (nm, e)
            assert rel_l2(g.probes(), o.probes()) <= TOL[precision]
            assert rel_l2(g.get_dft(0), o.get_dft(0)) <= TOL[precision]
            # the magnetic pole's M on its cells, measured against the flux B it is part of (B = mu_r mu0 H + M)
            m = np.isin(sc["matid"], [3]).reshape(-1)
            dm = g.get(6, 1).reshape(-1)[m] - o.get(8, 1).reshape(-1)[m]
            assert np.linalg.norm(dm) <= TOL[precision] * np.linalg.norm(o.get(B, 1).reshape(-1)[m])
            assert np.abs(o.get(8, 1)).max() > 0
            ez = g.get(E, 2)[0, 0]  # the ghost mirrors Ez, its sign flipped at a PEC wall
            assert ez[0] == (-ez[1] if wall == "PEC" else ez[1])
            # at a PEC wall the mirrored Hy ghost leaves the wall cell's curl at zero: Ez there stays 0
            wall_trace = np.abs(g.probes()[:, 3])
            assert (wall_trace.max() == 0.0) if wall == "PEC" else (wall_trace.max() > 0.0)
        else:
            for nm, f, c in COMPS:
                np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=nm)
            np.testing.assert_array_equal(g.get_dft(0), ref.get_dft(0))


def test_tem_refusals():
    from fdtd_amd import Batch, FdtdError, Grid
    sc, p, a = from_golden("tem1d_small")
    g = _grid(sc, "f64")
    with pytest.raises(FdtdError):
        g.set_bloch_boundary(0, 0, 1j)
    with pytest.raises(FdtdError):
        g.surface_quantities([1, 10, 0, 1, 0, 1])
    with pytest.raises(FdtdError):
        Batch([g])
    g.set_design([20, 30, 0, 1, 0, 1], np.full(10, 2.0))
    g.add_dft([1, 50, 0, 1, 0, 1], 4, np.zeros((10, 1, 2)), np.zeros((10, 1, 2)))
    with pytest.raises(FdtdError):
        g.adjoint_eps_gradient(10, 0, 4)
    g.step(3)
    assert Grid("TEM", 12, 7, 5, 1e-6, 1e-15).shape == (1, 1, 12)
