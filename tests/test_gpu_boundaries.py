"""GPU boundary conditions (fdtd_set_boundary) against the FP64 oracle.

The oracle's ghost-layer writes are pinned against the reference's own
boundary updaters (tests/test_oracle_boundaries.py); here the HIP path runs
whole simulations with PEC, PMC, Symmetric, Antisymmetric and Periodic faces
and must match the oracle to 1e-12 (FP64) / 1e-5 (FP32) relative L2 per
field, keep the PEC / PMC walls exactly zero, and give the same bits for every
tile and with captured chunks.
"""
import numpy as np
import pytest

from oracle import refcpu
from scenarios import E, H, apply, rel_l2
import synth

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}
SCEN = {
    "closed": lambda: synth.grid3d_boundaries(variant="closed"),
    "unit_cell": lambda: synth.grid3d_boundaries(variant="unit_cell"),
    "tm": synth.tm2d_boundaries,
}


def _grid(sc, precision):
    from fdtd_amd import Grid
    return Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision)


def _comps(sc):
    if sc["mode"] == "TM":
        return [("Ez", E, 2), ("Hx", H, 0), ("Hy", H, 1)]
    return [("Ex", E, 0), ("Ey", E, 1), ("Ez", E, 2), ("Hx", H, 0), ("Hy", H, 1), ("Hz", H, 2)]


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", list(SCEN))
def test_boundaries_match_oracle(name, precision):
    sc = SCEN[name]()
    print(f"[gpu-case] boundaries {name} {precision}", flush=True)
    g = apply(_grid(sc, precision), sc)
    o = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    g.step(sc["nsteps"])
    o.step(sc["nsteps"])
    for nm, f, c in _comps(sc):
        e = rel_l2(g.get(f, c), o.get(f, c))
        assert e <= TOL[precision], (nm, e)
    assert rel_l2(g.probes(), o.probes()) <= TOL[precision]
    for idx in range(len(sc["dfts"])):
        assert rel_l2(g.get_dft(idx), o.get_dft(idx)) <= TOL[precision]
    if name == "closed":  # x: PEC at min, PMC at max
        for f, c in ((E, 1), (E, 2), (H, 0)):
            assert np.all(g.get(f, c)[:, :, 1] == 0.0)
        for f, c in ((E, 0), (H, 1), (H, 2)):
            assert np.all(g.get(f, c)[:, :, -2] == 0.0)


@pytest.mark.parametrize("name", ["closed", "unit_cell"])
def test_boundaries_schedule_invariance(name):
    """Every tile, the fused tiles (which fall back to two passes when a face has
    a condition) and captured chunks give the same bits."""
    sc = SCEN[name]()
    ref = apply(_grid(sc, "f32"), sc)
    ref.step(sc["nsteps"])
    for tile, graph in (((64, 4, 2), False), ((256, 1, 4), False), ((0, 5, 2), False), ((128, 2, 1), True)):
        print(f"[gpu-case] boundaries {name} tile {tile} graph {graph}", flush=True)
        g = apply(_grid(sc, "f32"), sc)
        g.set_tile(*tile)
        g.set_graph(graph)
        g.step(37)
        g.step(sc["nsteps"] - 37)
        for nm, f, c in _comps(sc):
            np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{tile} {graph} {nm}")
        np.testing.assert_array_equal(g.probes(), ref.probes())
        g.close()


def test_boundary_errors():
    from fdtd_amd import Batch, FdtdError
    sc = SCEN["unit_cell"]()
    print("[gpu-case] boundary errors", flush=True)
    g = apply(_grid(sc, "f32"), sc)
    with pytest.raises(FdtdError, match="complex"):
        g.set_boundary(2, 0, "BlochPeriodic")
    with pytest.raises(FdtdError):
        g.set_boundary(3, 0, "PEC")
    other = apply(_grid(sc, "f32"), sc)
    with pytest.raises(FdtdError, match="boundary"):
        Batch([g, other])
    nx, ny, nz = sc["nx"], sc["ny"], sc["nz"]
    g.set_design([4, 8, 4, 8, 4, 8], np.full((4, 4, 4), 2.0))
    with pytest.raises(FdtdError, match="boundary"):
        g.adjoint_eps_gradient(10, 0, comp_mask=4)
    g.step(2)
    with pytest.raises(FdtdError, match="fixed"):
        g.set_boundary(0, 0, "PEC")


def _bloch_scenario(name):
    from scenarios import from_golden
    if name == "tm2d_bloch":
        return synth.tm2d_bloch(), None
    sc, p, a = from_golden("grid3d_bloch")
    sc["dfts"] = [synth.dft_monitor([1, sc["nx"] - 1, 1, sc["ny"] - 1, 10, 11], 63, sc["dt"], sc["nsteps"])]
    return sc, a


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", ["grid3d_bloch", "tm2d_bloch"])
def test_bloch_complex_fields(name, precision):
    """Bloch-periodic faces on complex fields: real and imaginary parts of every
    field, the probes and the complex DFT against the oracle (1e-12 / 1e-5), and
    for the 3-D case against the reference's own complex run."""
    sc, a = _bloch_scenario(name)
    print(f"[gpu-case] bloch {name} {precision}", flush=True)
    g = apply(_grid(sc, precision), sc)
    o = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    assert np.all(g.get_dft(0) == 0) and np.all(g.get(E, 2, imag=True) == 0)  # before the first step
    g.step(sc["nsteps"])
    o.step(sc["nsteps"])
    for nm, f, c in _comps(sc):
        for imag in (False, True):
            e = rel_l2(g.get(f, c, imag=imag), o.get(f, c, imag=imag))
            assert e <= TOL[precision], (nm, imag, e)
            if a is not None:
                ref = a[nm + ("_im" if imag else "_re")]
                assert rel_l2(g.get(f, c, imag=imag).reshape(ref.shape), ref) <= TOL[precision], (nm, imag)
    assert np.abs(o.get(*_comps(sc)[0][1:], imag=True)).max() > 0
    for imag in (False, True):
        assert rel_l2(g.probes(imag=imag), o.probes(imag=imag)) <= TOL[precision]
    for idx in range(len(sc["dfts"])):
        assert rel_l2(g.get_dft(idx), o.get_dft(idx)) <= TOL[precision]


def test_bloch_schedule_invariance_and_errors():
    """Every tile (fused requests fall back to two passes) and captured chunks
    give the same bits for both parts; batches and the adjoint refuse complex
    grids; the phase is fixed once stepping has started."""
    from fdtd_amd import Batch, FdtdError
    sc, _ = _bloch_scenario("grid3d_bloch")
    ref = apply(_grid(sc, "f32"), sc)
    ref.step(sc["nsteps"])
    for tile, graph in (((64, 4, 2), False), ((256, 1, 4), False), ((0, 5, 2), False), ((128, 2, 1), True)):
        print(f"[gpu-case] bloch tile {tile} graph {graph}", flush=True)
        g = apply(_grid(sc, "f32"), sc)
        g.set_tile(*tile)
        g.set_graph(graph)
        g.step(37)
        g.step(sc["nsteps"] - 37)
        for nm, f, c in _comps(sc):
            for imag in (False, True):
                np.testing.assert_array_equal(g.get(f, c, imag=imag), ref.get(f, c, imag=imag), err_msg=f"{tile} {nm}")
        np.testing.assert_array_equal(g.probes(imag=True), ref.probes(imag=True))
        np.testing.assert_array_equal(g.get_dft(0), ref.get_dft(0))
        g.close()
    with pytest.raises(FdtdError, match="fixed"):
        ref.set_bloch_boundary(0, 0, 1j)
    other = apply(_grid(sc, "f32"), sc)
    with pytest.raises(FdtdError, match="boundary"):
        Batch([ref, other])
    real = apply(_grid(synth.tm2d_boundaries(), "f32"), synth.tm2d_boundaries())
    with pytest.raises(FdtdError, match="real fields"):
        real.probes(imag=True)


def test_bloch_refuses_self_magnetized_material():
    """The real and imaginary parts are stepped apart, which holds only for
    updates linear with real coefficients: a self-magnetized plasma (its bias is
    the complex field itself) on a complex grid is refused."""
    from fdtd_amd import FdtdError
    sc, _ = _bloch_scenario("grid3d_bloch")
    print("[gpu-case] bloch self-magnetized refusal", flush=True)
    g = apply(_grid(sc, "f32"), sc)
    g.set_magnetized_drude(2, 1.0, 1.0, 1.37e16, 1e14, None)
    with pytest.raises(FdtdError, match="self-magnetized"):
        g.step(1)
