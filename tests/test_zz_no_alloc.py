"""fdtd_surface_quantities_static (the entry point that allocates nothing)
against fdtd_surface_quantities, through the C ABI: the same values bit for
bit on two calls in a row, the second on a smaller region, in 3-D (FP32 and
FP64) and TM; regions beyond FDTD_SURFACE_STATIC_MAX_CELLS and grids not yet
set up on the device are refused."""
import numpy as np
import pytest

from scenarios import apply
import synth

pytestmark = pytest.mark.gpu

MAX_CELLS = 65536  # FDTD_SURFACE_STATIC_MAX_CELLS


def _grid(sc, precision):
    from fdtd_amd import Grid
    return apply(Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision), sc)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_static_surface_equals_allocating_3d(precision):
    sc = synth.grid3d_cpml_tfsf_dft(n=(40, 34, 30), nsteps=60)
    print(f"[gpu-case] no-alloc surface 3-D {precision}", flush=True)
    g = _grid(sc, precision)
    g.step(60)
    big = [1, 39, 1, 33, 10, 20]    # 38 x 32 x 10 = 12160 cells
    small = [5, 17, 6, 13, 14, 16]  # 12 x 7 x 2 = 168 cells
    a1 = g.surface_quantities(big, static=True)
    b1 = g.surface_quantities(big)
    np.testing.assert_array_equal(a1, b1)
    assert np.abs(b1).max() > 0
    a2 = g.surface_quantities(small, static=True)
    b2 = g.surface_quantities(small)
    np.testing.assert_array_equal(a2, b2)
    # the second, smaller call reads only its own cells of the shared array
    np.testing.assert_array_equal(a2, b1[:, 4:6, 5:12, 4:16])
    g.step(5)
    np.testing.assert_array_equal(g.surface_quantities(small, static=True), g.surface_quantities(small))
    g.close()


def test_static_surface_equals_allocating_tm():
    sc = synth.tm2d_c1(nsteps=120)
    print("[gpu-case] no-alloc surface TM", flush=True)
    g = _grid(sc, "f64")
    g.step(120)
    for region in ([1, 255, 1, 255, 0, 1], [100, 140, 90, 170, 0, 1]):
        np.testing.assert_array_equal(g.surface_quantities(region, static=True), g.surface_quantities(region))
    g.close()


def test_static_surface_limits():
    from fdtd_amd import FdtdError
    sc = synth.grid3d_cpml_tfsf_dft(n=(70, 70, 30), nsteps=4)
    g = _grid(sc, "f32")
    with pytest.raises(FdtdError, match="not set up"):
        g.surface_quantities([1, 10, 1, 10, 1, 10], static=True)
    g.step(2)
    with pytest.raises(FdtdError, match="exceeds"):
        g.surface_quantities([1, 69, 1, 69, 1, 16], static=True)  # 68 x 68 x 15 > 65536
    ok = [1, 69, 1, 69, 1, 15]  # 68 x 68 x 14 = 64736 cells, within the limit
    assert 68 * 68 * 14 <= MAX_CELLS
    np.testing.assert_array_equal(g.surface_quantities(ok, static=True), g.surface_quantities(ok))
    g.close()
