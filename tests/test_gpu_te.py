"""2-D TE (Ex, Ey, Hz) on the HIP path, through the C ABI.

- The reference's own TE runs (te2d_small: CPML, dielectric, Drude and
  Lorentz blocks, a Dx and a Bz point source; te2d_tfsf: an oblique TE TFSF
  box around a block): every field, D and B, the probe traces against the
  oracle and the reference's arrays.
- Boundaries (PEC, PMC), E / H soft sources beside D / B ones, a single TFSF
  plane and a DFT monitor (te2d_features), against the oracle; every tile,
  chunked steps and captured chunks give the same bits.
- Bloch-periodic complex fields (te2d_bloch): both parts, probes and DFT.
- Surface charge density and currents on the reference's seeded TE state
  (bit for bit in FP64) and after a run.
- Refusals: the adjoint gradient and batches are 3-D only.

Tolerances, relative L2 per component (DFT: per bin relative to the monitor's
field at that frequency; probes): FP64 1e-12, FP32 1e-5.
"""
import numpy as np
import pytest

from oracle import refcpu
from parity import bin_errors, worst_bin
from scenarios import B, D, E, H, apply, from_golden, rel_l2
from golden_io import load
import synth

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 1e-5}
COMPS = [("Ex", E, 0), ("Ey", E, 1), ("Hz", H, 2)]
FLUX = [("Dx", D, 0), ("Dy", D, 1), ("Bz", B, 2)]


def _grid(sc, precision):
    from fdtd_amd import Grid
    return apply(Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision), sc)


def _oracle(sc):
    return apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", ["te2d_small", "te2d_tfsf"])
def test_te_reference_runs(name, precision):
    sc, p, a = from_golden(name)
    print(f"[gpu-case] {name} {precision}", flush=True)
    g, o = _grid(sc, precision), _oracle(sc)
    g.step(sc["nsteps"])
    o.step(sc["nsteps"])
    # te2d_small's Bz source is the pulse of the reference's tests/testTE.cpp,
    # nearly constant over the run (1e-9 T per step): it builds a static
    # circulating E whose curl cancels the source at its cell to 0.5%, and each
    # of the two cancelling terms, as H, is 3x the norm of the whole final Hz
    # (computed from the reference's own arrays).  FP32 rounding of those terms
    # lands in Hz and Bz, so FP32 is held to 1e-4 there; every other field and
    # every FP64 comparison keeps the common bound.
    ill = name == "te2d_small" and precision == "f32"
    for nm, f, c in COMPS + FLUX:
        tol = 1e-4 if ill and f in (H, B) else TOL[precision]
        e = rel_l2(g.get(f, c), o.get(f, c))
        print(nm, f"{e:.2e}", flush=True)
        assert e <= tol, (nm, e)
        assert rel_l2(g.get(f, c).reshape(a[nm].shape), a[nm]) <= tol, nm
    assert np.abs(g.get(H, 2)).max() > 0 and np.abs(g.get(E, 0)).max() > 0
    pr = g.probes()
    ref = a["probe_trace_Ex_Ey_Hz"].reshape(sc["nsteps"], -1)
    ptol = 1e-4 if ill else TOL[precision]
    assert rel_l2(pr, o.probes()) <= ptol
    assert rel_l2(pr, ref) <= ptol


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_te_sources_planes_boundaries_dft(precision):
    sc = synth.te2d_features()
    n = sc["nsteps"]
    o = _oracle(sc)
    o.step(n)
    ref = None
    for tile, graph in (((128, 2, 1), False), ((64, 4, 1), False), ((32, 8, 1), False), ((0, 4, 16), False),
                        ((128, 2, 1), True)):
        print(f"[gpu-case] te2d_features {precision} tile {tile} graph {graph}", flush=True)
        g = _grid(sc, precision)
        g.set_tile(*tile)
        g.set_graph(graph)
        g.step(37)
        g.step(n - 37)
        if ref is None:
            ref = g
            for nm, f, c in COMPS:
                e = rel_l2(g.get(f, c), o.get(f, c))
                assert e <= TOL[precision], (nm, e)
            for c in (0, 1):  # D of the interior: the boundary conditions write E and H of the ghosts only
                assert rel_l2(g.get(D, c)[:, 1:-1, 1:-1], o.get(D, c)[:, 1:-1, 1:-1]) <= TOL[precision]
            assert rel_l2(g.probes(), o.probes()) <= TOL[precision]
            berr = bin_errors(g, o, sc)
            assert worst_bin(berr) <= TOL[precision], berr
            # PEC wall at y = 1 (Ex tangential), PMC at y = ny-2 (Hz tangential)
            assert np.all(g.get(E, 0)[:, 1, :] == 0.0)
            assert np.all(g.get(H, 2)[:, -2, :] == 0.0)
            assert np.abs(g.get(E, 1)).max() > 0
        else:
            for nm, f, c in COMPS:
                np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{nm} {tile} {graph}")
            np.testing.assert_array_equal(g.get_dft(0), ref.get_dft(0))
            np.testing.assert_array_equal(g.probes(), ref.probes())


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_te_bloch_complex_fields(precision):
    sc = synth.te2d_bloch()
    print(f"[gpu-case] te2d_bloch {precision}", flush=True)
    g, o = _grid(sc, precision), _oracle(sc)
    g.step(sc["nsteps"])
    o.step(sc["nsteps"])
    for nm, f, c in COMPS:
        for imag in (False, True):
            e = rel_l2(g.get(f, c, imag=imag), o.get(f, c, imag=imag))
            assert e <= TOL[precision], (nm, imag, e)
    assert np.abs(o.get(E, 1, imag=True)).max() > 0
    for imag in (False, True):
        assert rel_l2(g.probes(imag=imag), o.probes(imag=imag)) <= TOL[precision]
    assert rel_l2(g.get_dft(0), o.get_dft(0)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_te_surface_quantities(precision):
    from fdtd_amd import Grid
    p, a = load("surface_quantities")
    nx, ny, dx = int(p["te_nx"]), int(p["te_ny"]), p["dx"]
    init = a["initTE_Ex_Ey_Hz"].reshape(3, 1, ny, nx)
    print(f"[gpu-case] TE surface quantities {precision}", flush=True)
    g = Grid("TE", nx, ny, 1, dx, 1e-15, precision=precision)
    o = refcpu.Sim("TE", nx, ny, 1, dx, 1e-15)
    for sim in (g, o):
        sim.set(E, 0, init[0])
        sim.set(E, 1, init[1])
        sim.set(H, 2, init[2])
    region = [1, nx - 1, 1, ny - 1, 0, 1]
    got = g.surface_quantities(region)
    want = a["sqTE_rho_Jex_Jey_Jhz"].reshape(1, ny - 2, nx - 2, 4)
    if precision == "f64":
        np.testing.assert_array_equal(np.moveaxis(got, 0, -1), want)
    else:
        assert rel_l2(np.moveaxis(got, 0, -1), want) <= TOL["f32"]
    np.testing.assert_array_equal(got, g.surface_quantities(region, static=True))
    # after a run: the oracle's quantities of the oracle's state
    sc = synth.te2d_bloch(nsteps=60)
    sc.pop("bloch")
    gg, oo = _grid(sc, precision), _oracle(sc)
    gg.step(sc["nsteps"])
    oo.step(sc["nsteps"])
    r = [2, sc["nx"] - 2, 3, sc["ny"] - 3, 0, 1]
    got, want = gg.surface_quantities(r), oo.surface_quantities(r)
    for q in range(4):
        assert rel_l2(got[q], want[q]) <= TOL[precision], q


def test_te_refusals():
    from fdtd_amd import Batch, FdtdError, Grid
    sc = synth.te2d_bloch(nsteps=4)
    sc.pop("bloch")
    g = _grid(sc, "f64")
    g.set_design([10, 14, 10, 14, 0, 1], np.full(16, 2.0))
    with pytest.raises(FdtdError):
        g.adjoint_eps_gradient(2, 0, 1)
    with pytest.raises(FdtdError):
        Batch([g])
    g.step(2)
    with pytest.raises(ValueError):
        Grid("TX", 8, 8)
