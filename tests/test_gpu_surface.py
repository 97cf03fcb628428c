"""GPU surface charge density and surface currents (fdtd_surface_quantities).

- On the seeded 3-D and TM field states of the `surface_quantities` fixture
  (the reference's own UpdateSurfaceChargeDensity / UpdateSurfaceCurrentE /
  UpdateSurfaceCurrentH, include/SurfaceQuantities.hpp), an FP64 grid gives
  the reference's values bit for bit; an FP32 grid gives the oracle's values
  on the FP32-rounded fields bit for bit (the kernel works in FP64).
- After a CPML + TFSF + dielectric-block run, the values on a box region match
  the oracle's (FP64 1e-12, FP32 1e-4 relative L2: the quantities are
  differences of neighbouring fields).
"""
import numpy as np
import pytest

from oracle import refcpu
from scenarios import E, H, apply, load, rel_l2
import synth

pytestmark = pytest.mark.gpu


def _grid(mode, nx, ny, nz, dx, precision):
    from fdtd_amd import Grid
    return Grid(mode, nx, ny, nz, dx, 1e-15, precision=precision)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_surface_quantities_on_reference_states(precision):
    p, a = load("surface_quantities")
    nx, ny, nz, dx = int(p["nx"]), int(p["ny"]), int(p["nz"]), p["dx"]
    print(f"[gpu-case] surface quantities fixture {precision}", flush=True)
    init = a["init3_E_H"].reshape(6, nz, ny, nx)
    if precision == "f32":
        init = init.astype(np.float32).astype(np.float64)
    g = _grid("ThreeD", nx, ny, nz, dx, precision)
    o = refcpu.Sim("ThreeD", nx, ny, nz, dx, 1e-15)
    for c in range(3):
        g.set(E, c, init[c])
        g.set(H, c, init[3 + c])
        o.set(E, c, init[c])
        o.set(H, c, init[3 + c])
    region = [1, nx - 1, 1, ny - 1, 1, nz - 1]
    got = g.surface_quantities(region)
    np.testing.assert_array_equal(got, o.surface_quantities(region))
    if precision == "f64":
        want = np.moveaxis(a["sq3_rho_Je_Jh"].reshape(nz - 2, ny - 2, nx - 2, 7), -1, 0)
        np.testing.assert_array_equal(got, want)
    sub = [2, 6, 3, 5, 2, 4]  # a sub-box: indexing of a region that is not the whole interior
    np.testing.assert_array_equal(g.surface_quantities(sub), o.surface_quantities(sub))

    tnx, tny = int(p["tm_nx"]), int(p["tm_ny"])
    init = a["initTM_Ez_Hx_Hy"].reshape(3, 1, tny, tnx)
    if precision == "f32":
        init = init.astype(np.float32).astype(np.float64)
    g = _grid("TM", tnx, tny, 1, dx, precision)
    o = refcpu.Sim("TM", tnx, tny, 1, dx, 1e-15)
    for sim in (g, o):
        sim.set(E, 2, init[0])
        sim.set(H, 0, init[1])
        sim.set(H, 1, init[2])
    region = [1, tnx - 1, 1, tny - 1, 0, 1]
    got = g.surface_quantities(region)
    np.testing.assert_array_equal(got, o.surface_quantities(region))
    if precision == "f64":
        want = np.moveaxis(a["sqTM_Jez_Jhx_Jhy"].reshape(1, tny - 2, tnx - 2, 3), -1, 0)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_surface_quantities_after_a_run(precision):
    sc = synth.grid3d_cpml_tfsf_dft(n=(36, 32, 28), nsteps=90, npml=4, inset=8)
    print(f"[gpu-case] surface quantities run {precision}", flush=True)
    from fdtd_amd import Grid
    g = apply(Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision), sc)
    o = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    g.step(sc["nsteps"])
    o.step(sc["nsteps"])
    region = [6, 30, 5, 27, 7, 21]  # crosses the TFSF faces (inset 8)
    got, want = g.surface_quantities(region), o.surface_quantities(region)
    tol = 1e-12 if precision == "f64" else 1e-4
    for q in range(got.shape[0]):
        assert np.abs(want[q]).max() > 0
        e = rel_l2(got[q], want[q])
        assert e <= tol, (q, e)
    from fdtd_amd import FdtdError
    with pytest.raises(FdtdError):
        g.surface_quantities([0, 4, 1, 4, 1, 4])  # the ghost layer has no backward neighbour
