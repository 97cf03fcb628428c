"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle and the
reference's golden vectors, on the same seeded inputs.

Tolerances (relative L2 per field component; per DFT bin relative to the
monitor's E or H field at that frequency; probe traces; P and J):
  FP64 GPU vs FP64 oracle: 1e-12
  FP32 GPU vs FP64 oracle: 1e-5
"""
import sys

import numpy as np
import pytest

from oracle import refcpu
from parity import bin_errors, worst_bin
from scenarios import D, E, B, H, P, J, apply, from_golden, gauss_table, rel_l2
import synth

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


def _grid(sc, precision):
    from fdtd_amd import Grid
    return Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision)


def _oracle(sc):
    return refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"])


def _comps(sc):
    if sc["mode"] == "TM":
        return [("Ez", E, 2), ("Hx", H, 0), ("Hy", H, 1)]
    return [("Ex", E, 0), ("Ey", E, 1), ("Ez", E, 2), ("Hx", H, 0), ("Hy", H, 1), ("Hz", H, 2)]


def run_pair(sc, precision, nsteps=None):
    print(f"[gpu-case] {sc['name']} {precision}", flush=True)
    n = sc["nsteps"] if nsteps is None else nsteps
    g = apply(_grid(sc, precision), sc)
    o = apply(_oracle(sc), sc)
    g.step(n)
    o.step(n)
    g.synchronize()
    return g, o


def field_errors(g, o, sc):
    return {nm: rel_l2(g.get(f, c), o.get(f, c)) for nm, f, c in _comps(sc)}


GOLDEN = ["tm2d_small", "tm2d_tfsf_x", "tm2d_tfsf_oblique", "grid3d_small", "grid3d_disp", "tm2d_c1",
          "tm2d_conductive", "grid3d_magnetized", "grid3d_fluid", "tm2d_fluid", "tm2d_soft_planes", "grid3d_magnetic"]
M_FIELD, K_FIELD = 6, 7  # magnetization and magnetic current of magnetic poles (FDTD_M, FDTD_K)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", GOLDEN)
def test_golden_scenarios(name, precision):
    sc, p, a = from_golden(name)
    g, o = run_pair(sc, precision)
    errs = field_errors(g, o, sc)
    print(name, precision, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= TOL[precision], (k, v)
    # probes against the oracle's and, for FP64, the reference's own traces
    pr_g, pr_o = g.probes(), o.probes()
    assert pr_g.shape == pr_o.shape
    assert rel_l2(pr_g, pr_o) <= TOL[precision]
    # fields against the reference's golden arrays directly
    for nm, f, c in _comps(sc):
        if nm in a:
            assert rel_l2(g.get(f, c).reshape(a[nm].shape), a[nm]) <= TOL[precision], nm
    if "dispersive_q_D_P_J" in a:
        dq = a["dispersive_q_D_P_J"]
        q = dq[:, 0].astype(int)
        for c in range(3):
            assert rel_l2(g.get(D, c).reshape(-1)[q], dq[:, 1 + c]) <= TOL[precision]
            assert rel_l2(g.get(P, c).reshape(-1)[q], dq[:, 4 + c]) <= TOL[precision]
            assert rel_l2(g.get(J, c).reshape(-1)[q], dq[:, 7 + c]) <= TOL[precision]
    if "magnetic_q_B_M_K" in a:  # magnetic poles: B, M, K of their cells against the reference's
        mq = a["magnetic_q_B_M_K"]
        q = mq[:, 0].astype(int)
        for c in range(3):
            assert rel_l2(g.get(B, c).reshape(-1)[q], mq[:, 1 + c]) <= TOL[precision], ("B", c)
            assert rel_l2(g.get(M_FIELD, c).reshape(-1)[q], mq[:, 4 + c]) <= TOL[precision], ("M", c)
            assert rel_l2(g.get(K_FIELD, c).reshape(-1)[q], mq[:, 7 + c]) <= TOL[precision], ("K", c)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_3d_cpml_tfsf_dft(precision):
    sc = synth.grid3d_cpml_tfsf_dft(n=40, nsteps=120)
    g, o = run_pair(sc, precision)
    errs = field_errors(g, o, sc)
    print(precision, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= TOL[precision], (k, v)
    berr = bin_errors(g, o, sc)
    print(precision, f"worst DFT bin {worst_bin(berr):.2e}", flush=True)
    assert worst_bin(berr) <= TOL[precision], berr


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_3d_multipole_sphere(precision):
    sc = synth.grid3d_dispersive_sphere(n=32, radius=8, nsteps=150)
    g, o = run_pair(sc, precision)
    errs = field_errors(g, o, sc)
    print(precision, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= TOL[precision], (k, v)
    for pole in range(3):
        for c in range(3):
            assert rel_l2(g.get(P, c, pole), o.get(P, c, pole)) <= TOL[precision]


def test_anisotropic_3d_f64_chunked_steps():
    """Stepping in uneven chunks equals one long call (incident table chunking)."""
    sc = synth.grid3d_cpml_tfsf_dft(n=(30, 26, 22), nsteps=300)
    from fdtd_amd import Grid
    g1 = apply(_grid(sc, "f64"), sc)
    g2 = apply(_grid(sc, "f64"), sc)
    g1.step(300)
    for n in (1, 255, 44):
        g2.step(n)
    for nm, f, c in _comps(sc):
        np.testing.assert_array_equal(g1.get(f, c), g2.get(f, c))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_tile_independence_and_autotune_restores_state(precision):
    """Every tile gives bitwise-identical fields (no halo/ordering races), and
    the autotuner's trial steps leave the simulation state untouched."""
    sc = synth.grid3d_cpml_tfsf_dft(n=(40, 34, 30), nsteps=80)
    ref = apply(_grid(sc, precision), sc)
    ref.step(80)
    outs = {nm: ref.get(f, c) for nm, f, c in _comps(sc)}
    for tile in [(64, 4, 1), (128, 2, 4), (64, 1, 8), (2, 12, 8)]:
        g = apply(_grid(sc, precision), sc)
        g.set_tile(*tile)
        g.step(30)
        g.set_tile(256, 1, 2)
        g.step(50)
        for nm, f, c in _comps(sc):
            np.testing.assert_array_equal(g.get(f, c), outs[nm], err_msg=f"{tile} {nm}")
        g.close()
    g = apply(_grid(sc, precision), sc)
    g.step(10)
    tile = g.autotune(trial_steps=3)
    assert g.step_count() == 10
    g.step(70)
    for nm, f, c in _comps(sc):
        np.testing.assert_array_equal(g.get(f, c), outs[nm], err_msg=f"autotune {tile} {nm}")
    np.testing.assert_array_equal(g.probes(), ref.probes())
    for idx in range(len(sc["dfts"])):
        np.testing.assert_array_equal(g.get_dft(idx), ref.get_dft(idx))


def test_tm2d_c1_probes_match_reference_trace():
    """2-D TM 256^2 x 1000 steps (the reference's CPU-runnable case): the four
    probe traces on the GPU in FP64 against the reference's own trace."""
    sc, p, a = from_golden("tm2d_c1")
    g = apply(_grid(sc, "f64"), sc)
    g.step(sc["nsteps"])
    ref = a["probe_trace_Ez_Hx_Hy"].reshape(sc["nsteps"], -1)
    assert rel_l2(g.probes(), ref) <= 1e-12


def test_autotune_as_first_call():
    """Regression: autotune before any step must leave the material and CPML
    tables uploaded (the snapshot is taken after pending uploads)."""
    sc = synth.grid3d_cpml_tfsf_dft(n=(36, 30, 28), nsteps=60)
    ref = apply(_grid(sc, "f32"), sc)
    ref.step(60)
    g = apply(_grid(sc, "f32"), sc)
    g.autotune(trial_steps=2)
    assert g.step_count() == 0
    g.step(60)
    for nm, f, c in _comps(sc):
        np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=nm)


# (0, 5, kc): streamed fused pass (128 x 5 E regions, kc persistent workgroups
# per compute unit; FP64 grids run the two-pass default); (1, TY, kc):
# z-marching fused pass (64 x TY workgroups owning 63 x (TY - 1) columns:
# nx=300 gives five tiles, the last 47 columns wide, ny=23 a partly empty last
# row of tiles, and the streamed pass three tiles, the last 44 columns wide);
# (2, TY, kc): time-blocked pass, two steps per launch (cores of 60 x (TY - 4)
# columns: five tiles along x, the last 58 wide; 17 steps end on a one-step
# launch)
FUSED = [(0, 5, 1), (0, 5, 2), (1, 4, 7), (1, 8, 16), (1, 16, 5), (2, 8, 16), (2, 12, 7), (2, 16, 5)]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_fused_pass_equals_two_pass(precision):
    """The fused single pass (ping-pong state, halo recomputation) gives the
    same fields as the two in-place passes on an anisotropic grid whose x
    extent spans two fused tiles, with CPML, blocks, TFSF, DFT and probes;
    schedules may be switched between steps."""
    sc = synth.grid3d_cpml_tfsf_dft(n=(300, 23, 21), nsteps=40, npml=4, inset=6)
    ref = apply(_grid(sc, precision), sc)
    ref.set_tile(256, 1, 1)
    ref.step(40)
    for tile in FUSED:
        print(f"[gpu-case] fused {tile} {precision}", flush=True)
        g = apply(_grid(sc, precision), sc)
        g.set_tile(*tile)
        g.step(17)
        g.set_tile(128, 2, 2)
        g.step(3)
        g.set_tile(*tile)
        g.step(20)
        for nm, f, c in _comps(sc):
            np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{tile} {nm}")
        np.testing.assert_array_equal(g.probes(), ref.probes())
        for idx in range(len(sc["dfts"])):
            np.testing.assert_array_equal(g.get_dft(idx), ref.get_dft(idx))
        g.close()


@pytest.mark.parametrize("kc", [1, 2])
def test_streamed_pass_runs_cross_tiles(kc):
    """Streamed fused pass (FP32) on grids whose (tile, plane) units give each persistent workgroup a run of
    several planes that ends in one tile and continues in the next: bitwise equal to the two passes, on a CPML +
    blocks + TFSF + DFT + probes grid and on a dispersive sphere with a dipole source."""
    for sc in (synth.grid3d_cpml_tfsf_dft(n=(260, 40, 70), nsteps=30, npml=6, inset=10),
               synth.grid3d_dispersive_sphere(n=(150, 44, 60), radius=12, nsteps=30)):
        print(f"[gpu-case] streamed runs kc={kc} {sc['name']}", flush=True)
        ref = apply(_grid(sc, "f32"), sc)
        ref.set_tile(128, 2, 2)
        ref.step(30)
        g = apply(_grid(sc, "f32"), sc)
        g.set_tile(0, 5, kc)
        g.step(30)
        for nm, f, c in _comps(sc):
            np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{kc} {nm}")
        np.testing.assert_array_equal(g.probes(), ref.probes())
        for idx in range(len(sc["dfts"])):
            np.testing.assert_array_equal(g.get_dft(idx), ref.get_dft(idx))
        if sc["name"].startswith("grid3d_dispersive"):
            for p in range(3):
                for c in range(3):
                    np.testing.assert_array_equal(g.get(P, c, p), ref.get(P, c, p))
                    np.testing.assert_array_equal(g.get(J, c, p), ref.get(J, c, p))
        g.close()
        ref.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_fused_pass_against_oracle(precision):
    """Fused schedule on the dispersive sphere and on a CPML+TFSF+DFT grid,
    against the FP64 oracle."""
    for sc, tile in ((synth.grid3d_dispersive_sphere(n=30, radius=7, nsteps=90), (0, 5, 1)),
                     (synth.grid3d_cpml_tfsf_dft(n=(37, 29, 33), nsteps=90), (0, 5, 2)),
                     (synth.grid3d_dispersive_sphere(n=30, radius=7, nsteps=90), (1, 8, 8)),
                     (synth.grid3d_cpml_tfsf_dft(n=(37, 29, 33), nsteps=90), (1, 4, 16)),
                     (synth.grid3d_cpml_tfsf_dft(n=(37, 29, 33), nsteps=90), (2, 12, 9)),
                     (synth.grid3d_cpml_tfsf_dft(n=(37, 29, 33), nsteps=90), (2, 8, 40))):
        print(f"[gpu-case] fused-vs-oracle {sc['name']} {tile} {precision}", flush=True)
        g = apply(_grid(sc, precision), sc)
        g.set_tile(*tile)
        o = apply(_oracle(sc), sc)
        g.step(sc["nsteps"])
        o.step(sc["nsteps"])
        errs = field_errors(g, o, sc)
        for k, v in errs.items():
            assert v <= TOL[precision], (sc["name"], k, v)
        if sc["name"].startswith("grid3d_dispersive"):
            for c in range(3):
                assert rel_l2(g.get(P, c, 1), o.get(P, c, 1)) <= TOL[precision]


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_standalone(graph):
    """Grids of uneven sizes stepped as one batch (tile table, hipGraph chunks)
    give the same fields, DFTs and probes as each grid stepped on its own,
    whatever tiles the batch uses, one workgroup per tile or persistent
    workgroups walking the table, and after its autotuner ran."""
    from fdtd_amd import Batch
    scs = synth.batch_c4(nsims=5, nsteps=300, lo=28, hi=44, seed=11)
    tf = synth.grid3d_cpml_tfsf_dft(n=(34, 30, 26), nsteps=300)  # one member with TFSF + probes
    scs.append(tf)
    alone = []
    for sc in scs:
        g = apply(_grid(sc, "f32"), sc)
        g.step(300)
        alone.append(g)
    batched = [apply(_grid(sc, "f32"), sc) for sc in scs]
    print(f"[gpu-case] batch graph={graph}", flush=True)
    b = Batch(batched, graph=graph)
    b.autotune(trial_steps=3)  # trial steps must leave every grid as it was
    b.step(7)
    b.set_tile(128, 2, 4)  # the tile changes speed only
    b.step(200)
    b.set_persistent(4)  # so do persistent workgroups walking the tile table
    b.step(56)
    b.set_tile(64, 4, 2)
    b.set_persistent(1)
    b.step(20)
    b.set_persistent(0)
    b.step(17)
    b.synchronize()
    b.close()
    for sc, g, r in zip(scs, batched, alone):
        assert g.step_count() == 300
        for nm, f, c in _comps(sc):
            np.testing.assert_array_equal(g.get(f, c), r.get(f, c), err_msg=f"{sc['name']} {nm}")
        for idx in range(len(sc["dfts"])):
            np.testing.assert_array_equal(g.get_dft(idx), r.get_dft(idx))
    np.testing.assert_array_equal(batched[-1].probes(), alone[-1].probes())


@pytest.mark.parametrize("tile", [(256, 1, 2), (1, 8, 8)])
def test_graph_chunks_equal_plain_launches(tile):
    """Chunks captured in hipGraphs and replayed give the fields, DFTs and
    probes of plain launches, across uneven step calls and a DFT read-back
    in the middle of a DFT block."""
    sc = synth.grid3d_cpml_tfsf_dft(n=(40, 33, 29), nsteps=700, npml=4, inset=8)
    print(f"[gpu-case] graph chunks {tile}", flush=True)
    ref = apply(_grid(sc, "f32"), sc)
    ref.set_tile(*tile)
    ref.step(700)
    g = apply(_grid(sc, "f32"), sc)
    g.set_tile(*tile)
    g.set_graph(True)
    g.step(300)
    g.step(13)
    g.get_dft(0)
    g.step(387)
    for nm, f, c in _comps(sc):
        np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=nm)
    for idx in range(len(sc["dfts"])):
        np.testing.assert_array_equal(g.get_dft(idx), ref.get_dft(idx))
    np.testing.assert_array_equal(g.probes(), ref.probes())


def test_graph_chunks_tm():
    sc = synth.tm2d_c1(nsteps=600)
    print("[gpu-case] graph chunks TM", flush=True)
    ref = apply(_grid(sc, "f64"), sc)
    ref.step(600)
    g = apply(_grid(sc, "f64"), sc)
    g.set_graph(True)
    g.step(600)
    np.testing.assert_array_equal(g.probes(), ref.probes())
    np.testing.assert_array_equal(g.get(E, 2), ref.get(E, 2))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_3d_conductive_block(precision):
    """A conductive block (ConductiveUpdate, one FDTD_CONDUCTIVE pole) next to a
    Drude block in a 3-D CPML + TFSF + DFT grid, against the oracle."""
    sc = synth.grid3d_cpml_tfsf_dft(n=(34, 30, 26), nsteps=120, npml=4, inset=7)
    sc["materials"][1] = (2.0, 1.0, [(3, 1e6, 0.0, 1.0)])
    sc["materials"][2] = (1.5, 1.0, [(1, 2e15, 1e14, 1.0)])
    sc["name"] = "grid3d_conductive"
    assert (sc["matid"] == 1).any() and (sc["matid"] == 2).any()
    g, o = run_pair(sc, precision)
    errs = field_errors(g, o, sc)
    print(precision, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= TOL[precision], (k, v)
    for c in range(3):
        assert rel_l2(g.get(P, c), o.get(P, c)) <= TOL[precision]


def test_3d_magnetized_schedules_and_errors():
    """Magnetized Drude grids run the two-pass schedule with the E-pass instance
    that carries the coupled update: a fused tile request and every two-pass
    tile give the same bits, captured chunks too; batches refuse such grids."""
    from fdtd_amd import Batch, FdtdError, Grid
    sc, p, a = from_golden("grid3d_magnetized")

    def run(tile, graph):
        print(f"[gpu-case] magnetized tile {tile} graph {graph}", flush=True)
        g = apply(Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision="f32"), sc)
        if tile:
            g.set_tile(*tile)
        g.set_graph(graph)
        g.step(31)
        g.step(sc["nsteps"] - 31)
        return g

    ref = run(None, False)
    for tile, graph in (((0, 5, 2), False), ((1, 4, 2), False), ((64, 4, 2), False), ((256, 1, 4), True)):
        g = run(tile, graph)
        for f in (E, H, P, J):
            for c in range(3):
                np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{tile} {graph} {f} {c}")
        g.close()
    other = apply(Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision="f32"), sc)
    with pytest.raises(FdtdError, match="magnetized"):
        Batch([ref, other])


def test_fluid_state_changes_between_steps():
    """A fluid Drude grid whose density and collision rate change between steps
    (a coupled fluid model) follows the oracle doing the same (FP64 1e-12,
    FP32 1e-5); a fluid material without a fluid state is refused."""
    from fdtd_amd import FdtdError, Grid
    sc, p, a = from_golden("grid3d_fluid")
    ne, nu = sc["fluid"]
    for precision in ("f64", "f32"):
        print(f"[gpu-case] fluid state update {precision}", flush=True)
        g = apply(Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision), sc)
        o = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
        for sim in (g, o):
            sim.step(30)
            sim.set_fluid_state(1.5 * ne, 0.5 * nu)
            sim.step(40)
        for f in (E, H, P, J):
            for c in range(3):
                assert rel_l2(g.get(f, c), o.get(f, c)) <= TOL[precision], (f, c)
        g.close()
    g = Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision="f32")
    g.set_fluid_drude(1, 1.0, 0, None)
    g.set_matid(sc["matid"])
    with pytest.raises(FdtdError, match="fluid"):
        g.step(1)


TM_BLOCK_TILES = [(0, 2, 16), (0, 4, 16), (0, 8, 16), (0, 4, 24), (0, 4, 8), (0, 6, 20)]


@pytest.mark.parametrize("name", ["tm2d_c1", "tm2d_tfsf_oblique", "tm2d_small"])
def test_tm_time_blocked_pass(name):
    """The time-blocked TM pass (TB steps per launch on kc x kc cores with a TB
    halo, ping-pong state) gives the two-pass bits on the reference's TM
    scenarios (CPML, D source, probes; the TFSF box; tm2d_small has dispersive
    blocks, which run the two passes): chunks of odd length, graph capture."""
    sc, p, a = from_golden(name)
    n = sc["nsteps"]
    ref = apply(_grid(sc, "f64"), sc)
    ref.step(n)
    o = apply(_oracle(sc), sc)
    o.step(n)
    for nm, f, c in _comps(sc):
        assert rel_l2(ref.get(f, c), o.get(f, c)) <= TOL["f64"], nm
    for tile in TM_BLOCK_TILES:
        for graph in (False, True):
            print(f"[gpu-case] tm block {name} tile {tile} graph {graph}", flush=True)
            g = apply(_grid(sc, "f64"), sc)
            g.set_tile(*tile)
            g.set_graph(graph)
            g.step(37)
            g.step(n - 37)
            for nm, f, c in _comps(sc):
                np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{name} {tile} {graph} {nm}")
            np.testing.assert_array_equal(g.probes(), ref.probes())
            g.close()


def test_tm_time_blocked_f32_and_autotune():
    sc, p, a = from_golden("tm2d_c1")
    n = 300
    print("[gpu-case] tm block f32 + autotune", flush=True)
    ref = apply(_grid(sc, "f32"), sc)
    ref.step(n)
    g = apply(_grid(sc, "f32"), sc)
    g.set_tile(0, 4, 16)
    g.step(n)
    for nm, f, c in _comps(sc):
        np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=nm)
    t = apply(_grid(sc, "f64"), sc)
    tile = t.autotune(trial_steps=20)
    print("autotuned TM tile", tile, flush=True)
    t.step(n)
    o = apply(_oracle(sc), sc)
    o.step(n)
    for nm, f, c in _comps(sc):
        assert rel_l2(t.get(f, c), o.get(f, c)) <= TOL["f64"], nm


def test_batch_dft_blocks_and_phases():
    """The batch's DFT: time-blocked when every grid is at the same phase of
    its block (a spectrum read in the middle of a block folds what is
    pending), summed step by step when the phases differ (one grid stepped
    alone first); the spectra equal those of the grids stepped alone."""
    from fdtd_amd import Batch
    scs = synth.batch_c4(nsims=3, nsteps=200, lo=26, hi=36, seed=12)
    alone = []
    for sc in scs:
        g = apply(_grid(sc, "f32"), sc)
        g.step(200)
        alone.append(g)
    print("[gpu-case] batch dft phases", flush=True)
    gs = [apply(_grid(sc, "f32"), sc) for sc in scs]
    b = Batch(gs)
    b.step(21)                    # blocked, ends mid-block
    spec = gs[1].get_dft(0)       # folds grid 1's pending samples
    assert np.abs(spec).max() > 0
    b.step(30)                    # grid 1 is now folded further than the others: still the same phase
    b.close()
    gs[0].step(5)                 # phases differ from here on
    b = Batch(gs)
    b.step(60)                    # summed step by step
    b.close()
    gs[1].step(5)
    gs[2].step(5)
    b = Batch(gs)                 # the same phase again: blocked
    b.step(200 - 116)
    b.close()
    for sc, g, r in zip(scs, gs, alone):
        assert g.step_count() == 200
        np.testing.assert_array_equal(g.get(E, 2), r.get(E, 2))
        for idx in range(len(sc["dfts"])):
            np.testing.assert_array_equal(g.get_dft(idx), r.get_dft(idx), err_msg=sc["name"])


def _tb_setup(nsteps, design):
    """A ragged non-dispersive grid for the time-blocked pass: CPML on every
    axis, four materials, D and B point sources (one in a CPML slab), probes of
    E and H, a DFT monitor of all components and, optionally, a design box of
    per-cell permittivities."""
    n = (131, 37, 29)
    dx = 1e-6
    dt = 0.5 * dx / (refcpu.constants()["c0"] * np.sqrt(3.0))
    rng = np.random.default_rng(21)
    ids = rng.integers(0, 4, size=(n[2], n[1], n[0])).astype(np.uint8)

    def setup(s):
        for ax in range(3):
            s.set_pml_idiom(ax, 4, 3.0, 1.0, 0.8 * 4 / (refcpu.constants()["imp0"] * dx), 5.0, 0.05)
        s.set_material(1, 2.0)
        s.set_material(2, 5.5)
        s.set_material(3, 1.5, 2.0)
        s.set_matid(ids)
        s.add_source(D, 2, 64, 18, 14, gauss_table(nsteps, 15, 5))
        s.add_source(B, 0, 2, 20, 6, 1e-9 * gauss_table(nsteps, 22, 6))
        s.add_probe(E, 2, 64, 18, 14)
        s.add_probe(H, 1, 61, 3, 25)
        s.add_probe(E, 0, 1, 1, 1)
        s.add_dft(*synth.dft_monitor([2, 129, 3, 34, 12, 14], 63, dt, nsteps))
        if design:
            s.set_design([50, 90, 5, 30, 8, 20], 1.0 + 11.0 * np.random.default_rng(5).random((12, 25, 40)))

    return n, dx, dt, setup


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("design", [False, True])
def test_time_blocked_pass(precision, design):
    """Time-blocked pass (two steps per launch) against the two passes, bitwise
    (fields, probe traces, DFT bins), for odd step counts and chunk lengths,
    schedule switches and captured chunks; the first run against the oracle."""
    from fdtd_amd import Grid
    nsteps = 75
    n, dx, dt, setup = _tb_setup(nsteps, design)
    ref = Grid("ThreeD", *n, dx, dt, precision=precision)
    setup(ref)
    ref.set_tile(128, 2, 2)
    ref.step(nsteps)
    o = refcpu.Sim("ThreeD", *n, dx, dt)
    setup(o)
    o.step(nsteps)
    comps = [(E, c) for c in range(3)] + [(H, c) for c in range(3)]
    for f, c in comps:
        assert rel_l2(ref.get(f, c), o.get(f, c)) <= TOL[precision], (f, c)
    for tile, graph, split in (((2, 12, 7), False, (75,)), ((2, 8, 5), False, (1, 30, 44)),
                               ((2, 16, 32), True, (33, 42)), ((2, 12, 64), True, (16, 17, 42))):
        print(f"[gpu-case] time-blocked {precision} design={design} {tile} graph={graph} {split}", flush=True)
        g = Grid("ThreeD", *n, dx, dt, precision=precision)
        setup(g)
        g.set_graph(graph)
        for q, m in enumerate(split):
            g.set_tile(*(tile if q % 2 == 0 else (64, 4, 2)))
            g.step(m)
        assert g.get_tile() == tuple(tile if (len(split) - 1) % 2 == 0 else (64, 4, 2))
        for f, c in comps:
            np.testing.assert_array_equal(g.get(f, c), ref.get(f, c), err_msg=f"{tile} {f}{c}")
        np.testing.assert_array_equal(g.probes(), ref.probes())
        np.testing.assert_array_equal(g.get_dft(0), ref.get_dft(0))
        g.close()


def _with_magnetic_poles(name):
    """A 2-D reference scenario whose blocks also get magnetic sides: a Drude,
    a Lorentz and a conductive pole and a constant mu_r (beside their
    electric sides)."""
    sc, p, a = from_golden(name)
    dt = sc["dt"]
    mats = list(sc["materials"])
    for mid, mu in ((1, 1.3), (2, 1.0), (3, 1.6), (4, 2.0)):
        eps, _, poles = mats[mid]
        mats[mid] = (eps, mu, poles)
    sc["materials"] = mats
    sc["magnetic"] = [(1, [(1, 0.3 / dt, 0.02 / dt, 1.0)]), (2, [(2, 0.2 / dt, 0.01 / dt, 1.0)]),
                      (3, [(3, 0.05 * 1.6 * 4e-7 * np.pi / dt, 0.0, 1.0), (2, 0.15 / dt, 0.02 / dt, 1.0)])]
    return sc


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", ["tm2d_small", "te2d_small", "te2d_tfsf"])
def test_magnetic_poles_2d(name, precision):
    """Magnetic poles in TM and TE grids (H passes of both modes) against the oracle."""
    sc = _with_magnetic_poles(name)
    g, o = run_pair(sc, precision)
    comps = _comps(sc) if sc["mode"] != "TE" else [("Ex", E, 0), ("Ey", E, 1), ("Hz", H, 2)]
    for nm, f, c in comps:
        e = rel_l2(g.get(f, c), o.get(f, c))
        print(name, precision, nm, f"{e:.2e}", flush=True)
        tol = 1e-4 if (name == "te2d_small" and precision == "f32" and f == H) else TOL[precision]
        assert e <= tol, (nm, e)
    hc = [c for _, f, c in comps if f == H]
    for c in hc:  # the magnetic state of the pole cells
        mask = np.isin(sc["matid"], [1, 2, 3]).reshape(-1)
        assert rel_l2(g.get(M_FIELD, c).reshape(-1)[mask], o.get(8, c).reshape(-1)[mask]) <= TOL[precision]
        assert rel_l2(g.get(K_FIELD, c).reshape(-1)[mask], o.get(9, c).reshape(-1)[mask]) <= TOL[precision]
        assert np.abs(o.get(8, c)).max() > 0


def test_magnetic_poles_refusals():
    from fdtd_amd import Batch, FdtdError
    sc, p, a = from_golden("grid3d_magnetic")
    g = apply(_grid(sc, "f32"), sc)
    with pytest.raises(FdtdError):
        Batch([g])
    g.step(2)
    bl = apply(_grid(sc, "f32"), sc)
    bl.set_bloch_boundary(0, 0, 1j)
    bl.set_bloch_boundary(0, 1, -1j)
    with pytest.raises(FdtdError):
        bl.step(1)
