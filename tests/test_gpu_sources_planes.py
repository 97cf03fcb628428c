"""E / H soft sources and single TFSF planes on the HIP path (the general
instances of the passes), through the C ABI.

- 3-D against the oracle: soft sources on Ez inside a Drude block (FP32 takes
  the stored source value back off before its E-form ADE update), on Ex and
  on Hy, beside a D source; two single planes (x-MIN, y-MAX faces, their own
  lines and pulse delays) under an oblique wave; CPML, DFT, probes.  Relative
  L2 per field component, probe traces and DFT bins (per bin, relative to the
  monitor's field at that frequency): FP64 1e-12, FP32 1e-5.
- Every schedule gives the same bits: two-pass tiles; the fused tiles, which
  fall back to the two passes for these grids; chunked steps; graph capture.
- The plane-pair slab of a y/z-periodic vacuum grid (k = x): outside the
  total-field slab the HIP fields stay at round-off, as the oracle's do.
- Refusals: a plane more than one layer thick or outside the interior, a
  seventh plane, the adjoint gradient and batches of grids with either.
(The 2-D TM case runs against the reference's own fixture tm2d_soft_planes in
test_gpu_parity.py.)
"""
import numpy as np
import pytest

from oracle import refcpu
from parity import bin_errors, worst_bin
from scenarios import D, E, H, apply, gauss_table, rel_l2
import synth

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 1e-5}
COMPS = [(E, 0), (E, 1), (E, 2), (H, 0), (H, 1), (H, 2)]


def _grid(sc, precision):
    from fdtd_amd import Grid
    return apply(Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision), sc)


def _oracle(sc):
    return apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_sources_and_planes_against_oracle_every_schedule(precision):
    sc = synth.grid3d_sources_planes()
    n = sc["nsteps"]
    o = _oracle(sc)
    o.step(n)
    ref = None
    for tile, graph in (((128, 2, 4), False), ((64, 4, 2), False), ((256, 1, 4), False), ((0, 5, 1), False),
                        ((1, 8, 8), False), ((2, 12, 8), False), ((128, 2, 4), True)):
        print(f"[gpu-case] sources+planes {precision} tile {tile} graph {graph}", flush=True)
        g = _grid(sc, precision)
        g.set_tile(*tile)
        g.set_graph(graph)
        g.step(29)
        g.step(n - 29)
        if ref is None:
            ref = g
            errs = {f"{f}{c}": rel_l2(g.get(f, c), o.get(f, c)) for f, c in COMPS}
            berr = bin_errors(g, o, sc)
            perr = rel_l2(g.probes(), o.probes())
            print(precision, {k: f"{v:.2e}" for k, v in errs.items()}, f"probes {perr:.2e}",
                  f"worst bin {worst_bin(berr):.2e}", flush=True)
            for k, v in errs.items():
                assert v <= TOL[precision], (k, v)
            assert perr <= TOL[precision]
            assert worst_bin(berr) <= TOL[precision]
            # the sources and the planes all act: the source cells carry their tables' pulses
            tr = o.probes()
            assert np.abs(tr[:, 0]).max() > 0.1 and np.abs(tr[:, 2]).max() > 0
            continue
        for f, c in COMPS:
            np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{tile} {graph} {f}{c}")
        np.testing.assert_array_equal(g.probes(), ref.probes())
        np.testing.assert_array_equal(g.get_dft(0), ref.get_dft(0))
        g.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_plane_pair_slab_leakage(precision):
    sc = synth.grid3d_plane_slab()
    print(f"[gpu-case] plane slab {precision}", flush=True)
    g = _grid(sc, precision)
    g.step(sc["nsteps"])
    a, b = 14, 33
    npml = sc["pml"]["n"]
    x = np.arange(sc["nx"])
    out = ((x < a - 1) | (x > b + 1)) & (x > npml + 1) & (x < sc["nx"] - npml - 2)
    inside = (x > a) & (x < b)
    # FP32: the two sides' corrections cancel to its rounding of the wave
    bound = 1e-12 if precision == "f64" else 1e-5
    for f in (E, H):
        tot = max(np.max(np.abs(g.get(f, c)[:, :, inside])) for c in range(3))
        leak = max(np.max(np.abs(g.get(f, c)[:, :, out])) for c in range(3))
        print(precision, f, f"leak/peak {leak / tot:.2e}", flush=True)
        assert tot > 0
        assert leak <= bound * tot, (f, leak / tot)


def test_refusals():
    from fdtd_amd import FdtdError, Grid
    sc = synth.grid3d_sources_planes(n=(24, 20, 18), nsteps=10)
    print("[gpu-case] plane / source refusals", flush=True)
    g = Grid("ThreeD", 24, 20, 18, sc["dx"], sc["dt"], precision="f32")
    p = sc["planes"][0]
    args = (p["kvec"], p["pol"], p["polH"], p["origin"], 40, p["signal"])
    with pytest.raises(FdtdError, match="one cell layer"):
        g.add_tfsf_plane(0, 0, [8, 10, 2, 10, 2, 10], *args)
    with pytest.raises(FdtdError, match="interior"):
        g.add_tfsf_plane(0, 0, [8, 9, 0, 10, 2, 10], *args)
    for q in range(6):
        g.add_tfsf_plane(q % 3, q % 2, [[5, 6, 2, 10, 2, 10], [2, 10, 5, 6, 2, 10], [2, 10, 2, 10, 5, 6]][q % 3], *args)
    with pytest.raises(FdtdError, match="too many TFSF planes"):
        g.add_tfsf_plane(0, 0, [5, 6, 2, 10, 2, 10], *args)
    g.step(3)  # six planes step
    # the adjoint gradient and batches refuse E / H sources
    a = Grid("ThreeD", 24, 20, 18, sc["dx"], sc["dt"], precision="f32")
    a.add_source(E, 2, 10, 10, 9, gauss_table(10, 3, 2, scale=1.0))
    a.set_design([8, 12, 8, 12, 8, 10], np.full(32, 2.0))
    a.add_dft([8, 12, 8, 12, 9, 10], 4, *synth.dft_monitor([8, 12, 8, 12, 9, 10], 4, sc["dt"], 10)[2:])
    with pytest.raises(FdtdError, match="not E / H sources"):
        a.adjoint_eps_gradient(10, dft=0)
    from fdtd_amd import Batch
    with pytest.raises(FdtdError, match="E / H sources"):
        Batch([a])


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_signals_written_between_steps(precision):
    """Sources and plane signals given as empty tables and written run by run
    with fdtd_set_signal (the tables grow as they go) step exactly like the
    same grid given whole tables up front."""
    sc = synth.grid3d_sources_planes(nsteps=70)
    n = sc["nsteps"]
    print(f"[gpu-case] signals between steps {precision}", flush=True)
    whole = _grid(sc, precision)
    whole.step(n)
    sc2 = dict(sc, sources=[s[:5] + (np.zeros(0),) for s in sc["sources"]],
               planes=[dict(p, signal=np.zeros(0)) for p in sc["planes"]])
    g = _grid(sc2, precision)
    for s0, s1 in ((0, 5), (5, 33), (33, 34), (34, 70)):
        for q, s in enumerate(sc["sources"]):
            g.set_signal("source", q, s0, s[5][s0:s1])
        for q, p in enumerate(sc["planes"]):
            g.set_signal("plane", q, s0, p["signal"][s0:s1])
        g.step(s1 - s0)
    for f, c in COMPS:
        np.testing.assert_array_equal(g.get(f, c), whole.get(f, c), err_msg=f"{f}{c}")
    np.testing.assert_array_equal(g.probes(), whole.probes())
