"""The C++ host surface (fdtd_amd/include/fdtd_hip.hpp) driven like the
reference's own user loop (tests/testTM.cpp), checked against the reference's
golden probe trace of the 2-D TM 256^2 x 1000-step case."""
import os
import subprocess

import numpy as np
import pytest

from golden_io import load
from scenarios import rel_l2

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "tm_c1_host_api")


def test_tm_c1_through_cpp_host_surface():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/bin/tm_c1_host_api not built (run __graft_entry__.build())")
    p, a = load("tm2d_c1")
    coords = [str(int(v)) for pr in a["probes"] for v in pr]
    print("[gpu-case] tm_c1_host_api", flush=True)
    r = subprocess.run([BIN, *coords], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().splitlines()])
    ref = a["probe_trace_Ez_Hx_Hy"][:, :, 0]
    assert got.shape == ref.shape
    assert rel_l2(got, ref) <= 1e-12


def test_boundaries_through_cpp_host_surface():
    """Boundary conditions set through BoundaryOptions / make_symmetric_boundary
    and the Stepper: probe traces equal the oracle's run of the same scenario."""
    import synth
    from oracle import refcpu
    from scenarios import apply
    exe = os.path.join(os.path.dirname(BIN), "tm_boundaries_host_api")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/bin/tm_boundaries_host_api not built (run __graft_entry__.build())")
    print("[gpu-case] tm_boundaries_host_api", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().splitlines()])
    sc = synth.tm2d_boundaries()
    o = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    o.step(sc["nsteps"])
    ref = o.probes()
    assert got.shape == ref.shape
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) <= 1e-12


@pytest.mark.parametrize("name", ["grid3d_magnetized", "grid3d_fluid"])
def test_plasmas_through_cpp_host_surface(name):
    """MagnetizedDrudeUpdate / SelfMagnetizedDrudeUpdate, and the fluid forms
    with YeeGrid::setFluidState, in a MaterialMap on a 3-D Stepper: probe traces
    equal the reference's own run of the scenario (the golden fixture of the
    same name) to 1e-12."""
    exe = os.path.join(os.path.dirname(BIN), name + "_host_api")
    if not os.path.exists(exe):
        pytest.fail(f"tests/cpp/bin/{name}_host_api not built (run __graft_entry__.build())")
    print(f"[gpu-case] {name}_host_api", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().splitlines()])
    p, a = load(name)
    ref = a["probe_trace_E_H"].reshape(got.shape)
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) <= 1e-12


def test_bloch_through_cpp_host_surface():
    """make_bloch_periodic_boundary on the C++ Stepper (complex fields): real and
    imaginary probe traces equal the reference's own complex run (golden fixture
    grid3d_bloch) to 1e-12."""
    exe = os.path.join(os.path.dirname(BIN), "grid3d_bloch_host_api")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/bin/grid3d_bloch_host_api not built (run __graft_entry__.build())")
    print("[gpu-case] grid3d_bloch_host_api", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().splitlines()])
    p, a = load("grid3d_bloch")
    tr = a["probe_trace_E_H_reim"].reshape(got.shape[0], -1, 2)
    assert rel_l2(got[:, :24], tr[:, :, 0]) <= 1e-12
    assert rel_l2(got[:, 24:], tr[:, :, 1]) <= 1e-12


def test_cell_proxies_through_cpp_host_surface():
    """The reference's per-cell idiom on the C++ host surface: CPML placed with
    cell.setPMLParameters, the initial field written through cell.Ez() on the host
    mirror, the result read by walking getNeighborMax(0): equal to the oracle's
    run of the same scenario (1e-12), and Dz = eps_r eps0 Ez in the block."""
    import math
    from oracle import refcpu
    from scenarios import D, E, H, rel_l2
    exe = os.path.join(os.path.dirname(BIN), "tm_cell_proxy_host_api")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/bin/tm_cell_proxy_host_api not built (run __graft_entry__.build())")
    print("[gpu-case] tm_cell_proxy_host_api", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    row = np.array([float(x) for x in lines[0].split()])
    hx, hy = (float(x) for x in lines[1].split())
    dz, ez = (float(x) for x in lines[2].split())
    nx, ny, dx = 40, 36, 1e-6
    c0 = refcpu.constants()["c0"]
    dt = 0.5 * dx / (c0 * math.sqrt(2.0))
    eps0 = refcpu.constants()["eps0"]
    o = refcpu.Sim("TM", nx, ny, 1, dx, dt)
    smax = 0.8 * (3.0 + 1) / (refcpu.constants()["imp0"] * dx)  # PMLParameterModel(dx) defaults
    for ax in (0, 1):
        o.set_pml_idiom(ax, 8, 3.0, 1.0, smax, 5.0, 0.05)
    o.set_material(1, 2.0)
    ids = np.zeros((1, ny, nx), np.uint8)
    ids[0, 20:26, 24:30] = 1
    o.set_matid(ids)
    jj, ii = np.mgrid[0:ny, 0:nx]
    e0 = np.exp(-((ii - 20.0) ** 2 + (jj - 18.0) ** 2) / 16.0)
    e0[[0, -1], :] = 0.0
    e0[:, [0, -1]] = 0.0
    o.set(E, 2, e0[None])
    o.set(D, 2, (e0 * eps0 * np.where(ids[0] == 1, 2.0, 1.0))[None])
    o.step(60)
    ez_o = o.get(E, 2)[0]
    assert rel_l2(row, ez_o[18, 1:nx - 1]) <= 1e-12
    assert abs(hx - o.get(H, 0)[0, 18, 20]) <= 1e-12 * np.abs(o.get(H, 0)).max()
    assert abs(hy - o.get(H, 1)[0, 18, 20]) <= 1e-12 * np.abs(o.get(H, 1)).max()
    assert abs(ez - ez_o[22, 26]) <= 1e-12 * np.abs(ez_o).max()
    assert abs(dz - 2.0 * eps0 * ez) <= 1e-12 * abs(dz)


def test_for_each_time_loop_against_reference_fixture():
    """The reference's own loop shape through for_each: SoftSource on Ez and Hy
    applied inside the loop after their material steps, two single TFSF planes
    stepped by updateD / updateB, material functors per id, 170 steps flushed
    to the device in runs of 64: the probe traces equal the reference's own
    (golden fixture tm2d_soft_planes, from the reference's templates)."""
    exe = os.path.join(os.path.dirname(BIN), "tm_loop_host_api")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/bin/tm_loop_host_api not built (run __graft_entry__.build())")
    p, a = load("tm2d_soft_planes")
    print("[gpu-case] tm_loop_host_api", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().splitlines()])
    ref = a["probe_trace_Ez_Hx_Hy"].reshape(got.shape)
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) <= 1e-12


@pytest.mark.parametrize("name", ["small", "tfsf"])
def test_te_time_loop_through_cpp_host_surface(name):
    """The reference's TE loop shape (tests/testTE.cpp) through for_each on a
    YeeGrid<TE>: Dx and Bz soft sources inside the loop, dielectric / Drude /
    Lorentz functors per material id, or a TFSF box with the reference's TE
    polarisation stepped by updateD / updateB: the probe traces equal the
    reference's own (golden fixtures te2d_small, te2d_tfsf) to 1e-12."""
    exe = os.path.join(os.path.dirname(BIN), "te_loop_host_api")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/bin/te_loop_host_api not built (run __graft_entry__.build())")
    p, a = load("te2d_" + name)
    print(f"[gpu-case] te_loop_host_api {name}", flush=True)
    r = subprocess.run([exe, name], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().splitlines()])
    ref = a["probe_trace_Ex_Ey_Hz"].reshape(got.shape)
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) <= 1e-12


def test_magnetic_poles_through_cpp_host_surface():
    """Material functors on FieldType::Magnetic (Drude, Lorentz, conductive) in a
    reference-style for_each loop: the probe traces equal the reference's own run
    (golden fixture grid3d_magnetic) to 1e-12."""
    exe = os.path.join(os.path.dirname(BIN), "grid3d_magnetic_host_api")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/bin/grid3d_magnetic_host_api not built (run __graft_entry__.build())")
    p, a = load("grid3d_magnetic")
    print("[gpu-case] grid3d_magnetic_host_api", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().splitlines()])
    ref = a["probe_trace_E_H"].reshape(got.shape)
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) <= 1e-12


def test_tem_time_loop_through_cpp_host_surface():
    """The reference's TEM loop (tests/testTEM.cpp) through for_each on a
    YeeGrid<TEM> line: the probe traces equal the reference's own (golden
    fixture tem1d_small) to 1e-12."""
    exe = os.path.join(os.path.dirname(BIN), "tem_loop_host_api")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/bin/tem_loop_host_api not built (run __graft_entry__.build())")
    p, a = load("tem1d_small")
    print("[gpu-case] tem_loop_host_api", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in line.split()] for line in r.stdout.strip().splitlines()])
    ref = a["probe_trace_Ez_Hy"].reshape(got.shape)
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) <= 1e-12
