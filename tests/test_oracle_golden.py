"""The FP64 oracle (oracle/refcpu.c) against golden vectors written by the
reference's own templates (tests/golden/, made by oracle/ref_harness)."""
import numpy as np
import pytest

from golden_io import load
from oracle import refcpu
from scenarios import D, E, B, H, apply, from_golden, rel_l2


def test_constants():
    p, _ = load("pml_coeffs")
    c = refcpu.constants()
    for k in ("eps0", "mu0", "c0", "imp0"):
        assert c[k] == p[k], k


def test_pml_coefficients_and_model():
    p, a = load("pml_coeffs")
    L = int(p["L"])
    for t in range(3):
        dx, dt, n, smax = p[f"set{t}_dx"], p[f"set{t}_dt"], int(p[f"set{t}_n"]), p[f"set{t}_smax"]
        e, h = refcpu.pml_profile(L, n, 3.0, 1.0, smax, 5.0, 0.05, dt)
        ref = a[f"set{t}_KSABCFG_EH"]
        np.testing.assert_array_equal(e, ref[:, :7])
        np.testing.assert_array_equal(h, ref[:, 7:])
    for x, K, S, A in a["model_x_K_S_A"]:
        np.testing.assert_array_equal(refcpu.pml_model(x, 3.0, 1.0, 123.0, 5.0, 0.05), [K, S, A])


def test_aux1d_matches_reference():
    p, a = load("aux1d")
    aux = refcpu.Aux(int(p["ncells"]), p["dx"], p["dt"])
    assert aux.L == int(p["L"])
    locs = a["locs"]
    tr = np.zeros((int(p["nsteps"]), len(locs), 2))
    for n in range(int(p["nsteps"])):
        aux.step(a["signal_v"][n])
        for q, l in enumerate(locs):
            tr[n, q] = aux.interp_E(l * p["dx"]), aux.interp_H(l * p["dx"])
    np.testing.assert_array_equal(tr, a["trace_EH"])


@pytest.mark.parametrize("name", ["cell_ade", "cell_conductive", "cell_magnetized_tm"])
def test_cell_ade_recursions(name):
    """cell_magnetized_tm: a magnetized Drude material in TM is the Drude update
    per component (the reference's MagnetizedDrudeCall, DispersiveMaterials.hpp:1508-1537)."""
    p, a = load(name)
    dt = p["dt"]
    mats = a["mats_kind_eps_freq_gamma_delta"]
    tr = a["trace_D_E_P_J"]
    from scenarios import ref_material, set_ref_material
    for m, row in enumerate(mats):
        spec = ref_material(*row, bias=(0.0, 0.0, 1.0))
        poles = spec[2]
        sim = refcpu.Sim("TM", 3, 3, 1, 1.0, dt)
        set_ref_material(sim, 0, spec)
        out = []
        for n in range(tr.shape[1]):
            d = np.zeros((1, 3, 3))
            d[0, 1, 1] = tr[m, n, 0]
            sim.set(D, 2, d)
            for c in (0, 1):  # zero magnetic state: the D curl adds exactly 0, only the material acts
                sim.set(B, c, np.zeros((1, 3, 3)))
                sim.set(H, c, np.zeros((1, 3, 3)))
            sim.step(1)
            e = sim.get(E, 2)[0, 1, 1]
            pz = sim.get(4, 2)[0, 1, 1] if poles else 0.0
            jz = sim.get(5, 2)[0, 1, 1] if poles else 0.0
            out.append((e, pz, jz))
        np.testing.assert_array_equal(np.array(out), tr[m, :, 1:])


def test_cell_magnetized_3d():
    """MagnetizedDrudeUpdate3D (DispersiveMaterials.hpp:1283-1340) on one 3-D
    cell with prescribed D_n, for fixed biases along z, oblique and along x, and
    self-magnetized by prescribed B_n: E, P, J bit for bit."""
    p, a = load("cell_magnetized")
    dt = p["dt"]
    mats, tr = a["mats_kind_eps_freq_gamma_bx_by_bz"], a["trace_D3_E3_P3_J3_B3"]
    for m, (kind, eps, freq, gamma, bx, by, bz) in enumerate(mats):
        sim = refcpu.Sim("ThreeD", 3, 3, 3, 1.0, dt)
        sim.set_magnetized_drude(0, eps, 1.0, freq, gamma, (bx, by, bz) if kind == 5 else None)
        out = []
        for n in range(tr.shape[1]):
            row = tr[m, n]
            for c in range(3):
                d = np.zeros((3, 3, 3))
                d[1, 1, 1] = row[c]
                sim.set(D, c, d)
                b = np.zeros((3, 3, 3))
                b[1, 1, 1] = row[12 + c]
                sim.set(B, c, b)
                sim.set(H, c, np.zeros((3, 3, 3)))  # the D curl adds exactly 0
            sim.step(1)
            out.append([sim.get(f, c)[1, 1, 1] for f in (E, 4, 5) for c in range(3)])
        np.testing.assert_array_equal(np.array(out), tr[m, :, 3:12], err_msg=f"material {m}")
    # the bias matters: a zero-bias material differs from the biased trace
    assert np.abs(tr[0, :, 9] - tr[2, :, 9]).max() > 0


def test_cell_fluid():
    """The fluid Drude forms (FluidDrudeUpdate, MagnetizedFluidDrudeUpdate,
    SelfMagnetizedFluidDrudeUpdate; DispersiveMaterials.hpp:1160-1214, 1728-2028)
    on one 3-D cell and one TM cell with prescribed D_n, B_n and a density and
    collision rate that change every step: E, P, J bit for bit."""
    p, a = load("cell_fluid")
    dt = p["dt"]
    mats, tr = a["mats_kind_bx_by_bz"], a["trace_D3_E3_P3_J3_B3_ne_nu"]
    for m, (kind, bx, by, bz) in enumerate(mats):
        sim = refcpu.Sim("ThreeD", 3, 3, 3, 1.0, dt)
        sim.set_fluid_drude(0, 1.0, int(kind) - 7, (bx, by, bz) if kind == 8 else None)
        out = []
        for n in range(tr.shape[1]):
            row = tr[m, n]
            ne, nu = np.zeros((3, 3, 3)), np.zeros((3, 3, 3))
            ne[1, 1, 1], nu[1, 1, 1] = row[15], row[16]
            sim.set_fluid_state(ne, nu)
            for c in range(3):
                d = np.zeros((3, 3, 3))
                d[1, 1, 1] = row[c]
                sim.set(D, c, d)
                b = np.zeros((3, 3, 3))
                b[1, 1, 1] = row[12 + c]
                sim.set(B, c, b)
                sim.set(H, c, np.zeros((3, 3, 3)))
            sim.step(1)
            out.append([sim.get(f, c)[1, 1, 1] for f in (E, 4, 5) for c in range(3)])
        np.testing.assert_array_equal(np.array(out), tr[m, :, 3:12], err_msg=f"material {m}")
    p, a = load("cell_fluid_tm")
    mats, tr = a["mats_kind_bx_by_bz"], a["trace_D_E_P_J_ne_nu"]
    for m, (kind, bx, by, bz) in enumerate(mats):
        sim = refcpu.Sim("TM", 3, 3, 1, 1.0, p["dt"])
        sim.set_fluid_drude(0, 1.0, int(kind) - 7, (bx, by, bz) if kind == 8 else None)
        out = []
        for n in range(tr.shape[1]):
            ne, nu, d = np.zeros((1, 3, 3)), np.zeros((1, 3, 3)), np.zeros((1, 3, 3))
            ne[0, 1, 1], nu[0, 1, 1], d[0, 1, 1] = tr[m, n, 4], tr[m, n, 5], tr[m, n, 0]
            sim.set_fluid_state(ne, nu)
            sim.set(D, 2, d)
            for c in (0, 1):
                sim.set(B, c, np.zeros((1, 3, 3)))
                sim.set(H, c, np.zeros((1, 3, 3)))
            sim.step(1)
            out.append((sim.get(E, 2)[0, 1, 1], sim.get(4, 2)[0, 1, 1], sim.get(5, 2)[0, 1, 1]))
        np.testing.assert_array_equal(np.array(out), tr[m, :, 1:4], err_msg=f"TM material {m}")


@pytest.mark.parametrize("name", ["tm2d_small", "tm2d_tfsf_x", "tm2d_tfsf_oblique", "grid3d_small",
                                  "grid3d_disp", "tm2d_c1", "tm2d_conductive", "grid3d_magnetized",
                                  "grid3d_fluid", "tm2d_fluid", "tm2d_soft_planes", "te2d_small", "te2d_tfsf",
                                  "grid3d_magnetic", "tem1d_small", "tem1d_tfsf"])
def test_grid_scenarios_bit_exact(name):
    sc, p, a = from_golden(name)
    sim = refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"])
    apply(sim, sc)
    sim.step(sc["nsteps"])
    if sc["mode"] == "TM":
        comps = [("Ez", E, 2), ("Dz", D, 2), ("Hx", H, 0), ("Hy", H, 1), ("Bx", B, 0), ("By", B, 1)]
    elif sc["mode"] == "TE":
        comps = [("Ex", E, 0), ("Dx", D, 0), ("Ey", E, 1), ("Dy", D, 1), ("Hz", H, 2), ("Bz", B, 2)]
    elif sc["mode"] == "TEM":
        comps = [("Ez", E, 2), ("Dz", D, 2), ("Hy", H, 1), ("By", B, 1)]
    else:
        comps = [("Ex", E, 0), ("Ey", E, 1), ("Ez", E, 2), ("Hx", H, 0), ("Hy", H, 1), ("Hz", H, 2)]
    checked = 0
    for nm, f, c in comps:
        if nm in a:
            np.testing.assert_array_equal(sim.get(f, c).reshape(a[nm].shape), a[nm], err_msg=nm)
            checked += 1
    if "Ez_row_mid" in a:
        ez = sim.get(E, 2)[0]
        np.testing.assert_array_equal(ez[sc["ny"] // 2, :], a["Ez_row_mid"])
        np.testing.assert_array_equal(ez[:, sc["nx"] // 2], a["Ez_col_mid"])
        checked += 1
    key = [k for k in a if k.startswith("probe_trace")][0]
    pr = sim.probes()
    np.testing.assert_array_equal(pr, a[key].reshape(pr.shape))
    assert checked > 0
    if "dispersive_q_D_P_J" in a:
        dq = a["dispersive_q_D_P_J"]
        q = dq[:, 0].astype(int)
        for c in range(3):
            np.testing.assert_array_equal(sim.get(D, c).reshape(-1)[q], dq[:, 1 + c])
            np.testing.assert_array_equal(sim.get(4, c).reshape(-1)[q], dq[:, 4 + c])
            np.testing.assert_array_equal(sim.get(5, c).reshape(-1)[q], dq[:, 7 + c])
    if "magnetic_q_B_M_K" in a:  # the magnetic side's B, M, K (oracle fields 2, 8, 9)
        mq = a["magnetic_q_B_M_K"]
        q = mq[:, 0].astype(int)
        for c in range(3):
            np.testing.assert_array_equal(sim.get(B, c).reshape(-1)[q], mq[:, 1 + c])
            np.testing.assert_array_equal(sim.get(8, c).reshape(-1)[q], mq[:, 4 + c])
            np.testing.assert_array_equal(sim.get(9, c).reshape(-1)[q], mq[:, 7 + c])


def test_surface_quantities_match_reference_functors():
    """rc_surface_quantities against the reference's UpdateSurfaceChargeDensity /
    UpdateSurfaceCurrentE / UpdateSurfaceCurrentH (include/SurfaceQuantities.hpp)
    on seeded 3-D, TM and TE field states, bit for bit."""
    p, a = load("surface_quantities")
    nx, ny, nz, dx = int(p["nx"]), int(p["ny"]), int(p["nz"]), p["dx"]
    init = a["init3_E_H"].reshape(6, nz, ny, nx)
    sim = refcpu.Sim("ThreeD", nx, ny, nz, dx, 1e-15)
    for c in range(3):
        sim.set(E, c, init[c])
        sim.set(H, c, init[3 + c])
    got = sim.surface_quantities([1, nx - 1, 1, ny - 1, 1, nz - 1])
    want = a["sq3_rho_Je_Jh"].reshape(nz - 2, ny - 2, nx - 2, 7)
    np.testing.assert_array_equal(np.moveaxis(got, 0, -1), want)
    tnx, tny = int(p["tm_nx"]), int(p["tm_ny"])
    init = a["initTM_Ez_Hx_Hy"].reshape(3, 1, tny, tnx)
    sim = refcpu.Sim("TM", tnx, tny, 1, dx, 1e-15)
    sim.set(E, 2, init[0])
    sim.set(H, 0, init[1])
    sim.set(H, 1, init[2])
    got = sim.surface_quantities([1, tnx - 1, 1, tny - 1, 0, 1])
    want = a["sqTM_Jez_Jhx_Jhy"].reshape(1, tny - 2, tnx - 2, 3)
    np.testing.assert_array_equal(np.moveaxis(got, 0, -1), want)
    enx, eny = int(p["te_nx"]), int(p["te_ny"])
    init = a["initTE_Ex_Ey_Hz"].reshape(3, 1, eny, enx)
    sim = refcpu.Sim("TE", enx, eny, 1, dx, 1e-15)
    sim.set(E, 0, init[0])
    sim.set(E, 1, init[1])
    sim.set(H, 2, init[2])
    got = sim.surface_quantities([1, enx - 1, 1, eny - 1, 0, 1])
    want = a["sqTE_rho_Jex_Jey_Jhz"].reshape(1, eny - 2, enx - 2, 4)
    np.testing.assert_array_equal(np.moveaxis(got, 0, -1), want)
