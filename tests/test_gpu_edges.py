"""Edge cases of the HIP path against the oracle.

- The smallest grids (one interior cell: 3x3x3 and 3x3 TM) under the
  two-pass tiles, whose blocks are far larger than the grid.
- A ragged 3-D grid (37 x 29 x 23, no size a multiple of any tile) under every
  schedule: two-pass tiles, the streamed and z-marching fused passes, the
  time-blocked tile (which runs the two passes here: the grid is dispersive),
  captured chunks; all give the same bits and match the oracle.
- Every table at its limit (16 sources, 32 probes, 8 DFT monitors, 256
  material ids, 3 poles) and one past it (refused).
- DFT monitors of a single cell and monitors whose phase tables end before the
  run does (they stop accumulating), zero steps, spectra read before stepping.
"""
import numpy as np
import pytest

from oracle import refcpu
from scenarios import D, E, H, B, apply, gauss_table, rel_l2
import synth

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-12, "f32": 1e-5}
C0 = refcpu.constants()["c0"]


def _pair(mode, n, dx, dt, precision, setup):
    from fdtd_amd import Grid
    g = Grid(mode, *n, dx, dt, precision=precision)
    o = refcpu.Sim(mode, *n, dx, dt)
    for sim in (g, o):
        setup(sim)
    return g, o


def _comps(mode):
    if mode == "TM":
        return [(E, 2), (H, 0), (H, 1)]
    if mode == "TE":
        return [(E, 0), (E, 1), (H, 2)]
    return [(E, 0), (E, 1), (E, 2), (H, 0), (H, 1), (H, 2)]


@pytest.mark.parametrize("mode", ["ThreeD", "TM", "TE"])
@pytest.mark.parametrize("tile", [(64, 4, 2), (256, 1, 4), (128, 2, 1)])
def test_one_interior_cell(mode, tile):
    n = (3, 3, 3) if mode == "ThreeD" else (3, 3, 1)
    dx = 1e-6
    dt = 0.5 * dx / (C0 * np.sqrt(3.0 if mode == "ThreeD" else 2.0))
    print(f"[gpu-case] one interior cell {mode} tile {tile}", flush=True)

    def setup(s):
        s.set_material(1, 2.5)
        s.set_matid(np.ones((n[2], n[1], n[0]), np.uint8))
        ec = 0 if mode == "TE" else 2  # an E component of the mode
        s.add_source(D, ec, 1, 1, 1 if mode == "ThreeD" else 0, gauss_table(40, 10, 4))
        if mode == "ThreeD":
            s.add_source(B, 0, 1, 1, 1, 1e-9 * gauss_table(40, 14, 4))
        if mode == "TE":
            s.add_source(B, 2, 1, 1, 0, 1e-9 * gauss_table(40, 14, 4))
        s.add_probe(E, ec, 1, 1, 1 if mode == "ThreeD" else 0)
        s.add_dft(*synth.dft_monitor([1, 2, 1, 2, 1, 2] if mode == "ThreeD" else [1, 2, 1, 2, 0, 1], 1 << ec, dt, 40))

    g, o = _pair(mode, n, dx, dt, "f64", setup)
    if mode == "ThreeD" or tile[2] == 1:
        g.set_tile(*tile)
    g.step(40)
    o.step(40)
    # components that only cancellation feeds (Ex, Ey in the lone cell) sit at
    # round-off level: compare against each field type's own scale
    for ft in (E, H):
        scale = max(np.abs(o.get(f, c)).max() for f, c in _comps(mode) if f == ft)
        assert scale > 0
        for f, c in _comps(mode):
            if f == ft:
                assert np.abs(g.get(f, c) - o.get(f, c)).max() <= 1e-12 * scale, (f, c)
    assert rel_l2(g.probes(), o.probes()) <= 1e-12
    assert rel_l2(g.get_dft(0), o.get_dft(0)) <= 1e-12


def _ragged_setup(dt, nsteps):
    n = (37, 29, 23)
    dx = 1e-6
    rng = np.random.default_rng(12)
    ids = rng.integers(0, 4, size=(n[2], n[1], n[0])).astype(np.uint8)

    def setup(s):
        for ax in range(3):
            s.set_pml_idiom(ax, 3, 3.0, 1.0, 0.8 * 4 / (refcpu.constants()["imp0"] * dx), 5.0, 0.05)
        s.set_material(1, 2.0)
        s.set_material(2, 5.5)
        s.set_material(3, 1.5, 1.0, [(1, 2e14, 1e13, 1.0)])
        s.set_matid(ids)
        s.add_source(D, 2, 18, 14, 11, gauss_table(nsteps, 15, 5))
        s.add_source(D, 0, 9, 20, 6, gauss_table(nsteps, 22, 6))
        s.add_probe(E, 2, 18, 14, 11)
        s.add_probe(H, 1, 30, 5, 17)
        s.add_dft(*synth.dft_monitor([2, 35, 3, 26, 12, 13], 63, dt, nsteps))

    return n, dx, setup


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_ragged_grid_every_schedule(precision):
    nsteps = 70
    dt = 0.5 * 1e-6 / (C0 * np.sqrt(3.0))
    n, dx, setup = _ragged_setup(dt, nsteps)
    ref = None
    o = None
    for tile, graph in (((64, 4, 2), False), ((256, 1, 4), False), ((0, 5, 1), False), ((1, 8, 5), False),
                        ((2, 12, 6), False), ((128, 2, 4), True)):
        print(f"[gpu-case] ragged {precision} tile {tile} graph {graph}", flush=True)
        g, oo = _pair("ThreeD", n, dx, dt, precision, setup)
        g.set_tile(*tile)
        g.set_graph(graph)
        g.step(29)
        g.step(nsteps - 29)
        if ref is None:
            ref, o = g, oo
            o.step(nsteps)
            for f, c in _comps("ThreeD"):
                assert rel_l2(g.get(f, c), o.get(f, c)) <= TOL[precision], (f, c)
            assert rel_l2(g.get_dft(0), o.get_dft(0)) <= TOL[precision]
            assert rel_l2(g.probes(), o.probes()) <= TOL[precision]
            continue
        for f, c in _comps("ThreeD"):
            np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{tile} {graph}")
        np.testing.assert_array_equal(g.get_dft(0), ref.get_dft(0))
        g.close()


def test_tables_at_their_limits():
    from fdtd_amd import FdtdError, Grid
    n = (24, 22, 20)
    dx = 1e-8
    dt = 0.5 * dx / (C0 * np.sqrt(3.0))
    nsteps = 60
    rng = np.random.default_rng(13)
    ids = rng.integers(0, 256, size=(n[2], n[1], n[0])).astype(np.uint8)
    eps = 1.0 + 11.0 * rng.random(256)
    cells = [(int(rng.integers(2, n[0] - 2)), int(rng.integers(2, n[1] - 2)), int(rng.integers(2, n[2] - 2)))
             for _ in range(32)]
    print("[gpu-case] tables at their limits", flush=True)

    def setup(s):
        for m in range(256):
            if m == 7:  # three poles: Drude + Lorentz + conductive
                s.set_material(m, 1.5, 1.0, [(1, 1.37e16, 1e14, 1.0), (2, 4e15, 1e14, 1.0), (3, 1e5, 0.0, 1.0)])
            else:
                s.set_material(m, eps[m], 1.0 + (m % 3) * 0.5)
        s.set_matid(ids)
        for q in range(16):
            i, j, k = cells[q]
            s.add_source(D, q % 3, i, j, k, gauss_table(nsteps, 10 + q, 4))
        for q in range(32):
            i, j, k = cells[q]
            s.add_probe(E if q % 2 else H, q % 3, i, j, k)
        for q in range(8):
            s.add_dft(*synth.dft_monitor([2, n[0] - 2, 2, n[1] - 2, 2 + 2 * q, 3 + 2 * q], 1 << (q % 6), dt, nsteps))

    g, o = _pair("ThreeD", n, dx, dt, "f64", setup)
    g.step(nsteps)
    o.step(nsteps)
    for f, c in _comps("ThreeD"):
        assert rel_l2(g.get(f, c), o.get(f, c)) <= 1e-12, (f, c)
    assert rel_l2(g.probes(), o.probes()) <= 1e-12
    for q in range(8):
        assert rel_l2(g.get_dft(q), o.get_dft(q)) <= 1e-12, q
    over = Grid("ThreeD", *n, dx, dt, precision="f32")
    for q in range(16):
        over.add_source(D, 2, 5, 5, 5, gauss_table(4, 1, 1))
    with pytest.raises(FdtdError, match="too many sources"):
        over.add_source(D, 2, 5, 5, 5, gauss_table(4, 1, 1))
    for q in range(32):
        over.add_probe(E, 2, 5, 5, 5)
    with pytest.raises(FdtdError, match="too many probes"):
        over.add_probe(E, 2, 5, 5, 5)
    for q in range(8):
        over.add_dft(*synth.dft_monitor([2, 4, 2, 4, 2, 4], 4, dt, 4))
    with pytest.raises(FdtdError, match="too many DFT"):
        over.add_dft(*synth.dft_monitor([2, 4, 2, 4, 2, 4], 4, dt, 4))
    with pytest.raises(FdtdError):
        over.set_material(0, 1.0, 1.0, [(1, 1.0, 1.0, 1.0)] * 4)  # four poles


def test_dft_edges_and_zero_steps():
    """A single-cell monitor, a monitor whose phase table ends at step 23 of 50
    (it keeps what it had), spectra read before any step (zero), step(0)."""
    from fdtd_amd import Grid
    n = (20, 18, 16)
    dx = 1e-6
    dt = 0.5 * dx / (C0 * np.sqrt(3.0))
    print("[gpu-case] dft edges", flush=True)

    def setup(s):
        s.add_source(D, 2, 10, 9, 8, gauss_table(50, 8, 3))
        s.add_dft(*synth.dft_monitor([11, 12, 9, 10, 8, 9], 63, dt, 50))
        s.add_dft(*synth.dft_monitor([3, 17, 3, 15, 10, 11], 4 | 8, dt, 23))

    g, o = _pair("ThreeD", n, dx, dt, "f32", setup)
    assert np.all(g.get_dft(0) == 0) and np.all(g.get_dft(1) == 0)
    g.step(0)
    assert g.step_count() == 0
    g.step(31)
    g.step(19)
    o.step(50)
    for q in range(2):
        assert rel_l2(g.get_dft(q), o.get_dft(q)) <= 1e-5, q
    assert np.abs(o.get_dft(1)).max() > 0
