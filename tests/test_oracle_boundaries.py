"""Oracle boundary conditions (BoundaryUpdates.hpp) pinned against the
reference's own boundary updaters and against translation invariance.

- Masks and placement: the golden fixture `boundaries` holds what the
  reference's make_{pec,pmc,symmetric,antisymmetric,periodic}_boundary
  (include/BoundaryUpdates.hpp:527-832), compiled by oracle/ref_harness,
  write into the ghost planes of a seeded 7x6x5 grid for every face and field
  type; the oracle must write the same values, bit for bit.
- Periodic wrap: with periodic x and y and nothing else that depends on x or
  y, moving the source by (dx, dy) cells moves the whole solution cyclically by
  the same amount, bit for bit; a wrong periodic image breaks that.
"""
import numpy as np
import pytest

from oracle import refcpu
from scenarios import D, E, B, H, load, rel_l2
import synth

KINDS = {1: "PEC", 2: "PMC", 3: "Symmetric", 4: "Antisymmetric", 5: "Periodic"}


def test_boundary_updates_match_reference_functors():
    p, a = load("boundaries")
    nx, ny, nz = int(p["nx"]), int(p["ny"]), int(p["nz"])
    init = a["init_E_H"].reshape(6, nz, ny, nx)
    cases = a["cases_type_axis_side_ft"].reshape(-1, 4).astype(int)
    ghosts = a["ghosts"].reshape(len(cases), -1)
    n = (nx, ny, nz)
    seen = set()
    for (kind, axis, side, ft), want in zip(cases, ghosts):
        seen.add(kind)
        s = refcpu.Sim("ThreeD", nx, ny, nz, 1e-6, 1e-15)
        for c in range(3):
            s.set(E, c, init[c])
            s.set(H, c, init[3 + c])
        s.set_boundary(axis, side, KINDS[kind])
        s.apply_boundaries(ft)
        f = [s.get(E if ft == 0 else H, c) for c in range(3)]  # (nz, ny, nx)
        dst = n[axis] - 1 if side else 0
        t1 = 1 if axis == 0 else 0
        t2 = 1 if axis == 2 else 2
        got = []
        for v in range(1, n[t2] - 1):
            for u in range(1, n[t1] - 1):
                ci = [0, 0, 0]
                ci[axis], ci[t1], ci[t2] = dst, u, v
                got += [f[c][ci[2], ci[1], ci[0]] for c in range(3)]
        got = np.array(got)
        np.testing.assert_array_equal(got, want[: len(got)], err_msg=f"{KINDS[kind]} axis {axis} side {side} ft {ft}")
        # the write touched only its ghost plane
        for c in range(3):
            ref = init[(0 if ft == 0 else 3) + c].copy()
            idx = [slice(None)] * 3
            idx[2 - axis] = dst
            got_other = f[c].copy()
            got_other[tuple(idx)] = ref[tuple(idx)]
            np.testing.assert_array_equal(got_other, ref)
    assert seen == set(KINDS)


def _periodic_run(shift, nsteps=40):
    nx, ny, nz = 20, 17, 14
    dx = 1e-6
    dt = 0.5 * dx / (refcpu.constants()["c0"] * np.sqrt(3.0))
    s = refcpu.Sim("ThreeD", nx, ny, nz, dx, dt)
    s.set_pml_idiom(2, 4, 3.0, 1.0, 0.8 * 4 / (refcpu.constants()["imp0"] * dx), 5.0, 0.05)
    for side in (0, 1):
        s.set_boundary(0, side, "Periodic")
        s.set_boundary(1, side, "Periodic")
    s.set_material(1, 3.0, 1.0, [])
    ids = np.zeros((nz, ny, nx), np.uint8)
    ids[5:9, :, :] = 1  # a slab, uniform in x and y
    s.set_matid(ids)
    i0, j0 = 1 + (6 + shift[0]) % (nx - 2), 1 + (5 + shift[1]) % (ny - 2)
    s.add_source(D, 2, i0, j0, 7, synth.gauss_table(nsteps, 12, 4))
    s.step(nsteps)
    return s, (nx, ny, nz)


def test_periodic_translation_invariance():
    a, (nx, ny, nz) = _periodic_run((0, 0))
    b, _ = _periodic_run((5, 3))
    for f in (E, H):
        for c in range(3):
            fa, fb = a.get(f, c), b.get(f, c)
            ia = fa[1:-1, 1:-1, 1:-1]
            ib = fb[1:-1, 1:-1, 1:-1]
            np.testing.assert_array_equal(np.roll(ia, (3, 5), axis=(1, 2)), ib, err_msg=f"field {f} comp {c}")
    assert rel_l2(a.get(E, 2), np.zeros_like(a.get(E, 2))) > 0


def test_pec_and_pmc_walls_are_exact():
    """The reference's PEC masks make layer 1 a PEC wall (tangential E and normal
    H stay exactly zero there for all time), its PMC masks make layer L-2 a PMC
    wall (tangential H and normal E zero); the interior next to them is not zero."""
    from scenarios import apply
    sc = synth.grid3d_boundaries(variant="closed")  # x: PEC at min, PMC at max
    s = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    s.step(sc["nsteps"])
    Ex, Ey, Ez = (s.get(E, c) for c in range(3))
    Hx, Hy, Hz = (s.get(H, c) for c in range(3))
    for f in (Ey, Ez, Hx):
        assert np.all(f[:, :, 1] == 0.0)
        assert np.abs(f[:, :, 2]).max() > 0
    for f in (Ex, Hy, Hz):
        assert np.all(f[:, :, -2] == 0.0)
        assert np.abs(f[:, :, -3]).max() > 0


def test_bloch_complex_grid_matches_reference():
    """Bloch-periodic x and y faces on complex fields (the reference's own
    functors on std::complex<double> cells: make_bloch_periodic_boundary,
    BoundaryUpdates.hpp:845-885, with curl, CPML, a dielectric and a Drude
    block): the oracle's real and imaginary parts equal the reference's bit for
    bit, fields and probe traces."""
    from scenarios import from_golden, apply, E, H
    sc, p, a = from_golden("grid3d_bloch")
    o = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    o.step(sc["nsteps"])
    for nm, f, c in (("Ex", E, 0), ("Ey", E, 1), ("Ez", E, 2), ("Hx", H, 0), ("Hy", H, 1), ("Hz", H, 2)):
        np.testing.assert_array_equal(o.get(f, c), a[nm + "_re"], err_msg=nm)
        np.testing.assert_array_equal(o.get(f, c, imag=True), a[nm + "_im"], err_msg=nm)
    tr = a["probe_trace_E_H_reim"].reshape(sc["nsteps"], -1, 2)
    np.testing.assert_array_equal(o.probes(), tr[:, :, 0])
    np.testing.assert_array_equal(o.probes(imag=True), tr[:, :, 1])
    assert np.abs(tr[:, :, 1]).max() > 1e-3 * np.abs(tr[:, :, 0]).max()  # the phase does reach the fields


def test_bloch_tm_invariants():
    """TM Bloch cell (synth.tm2d_bloch): with theta = 0 the complex grid is the
    periodic real grid (imaginary part exactly zero); with theta = pi the
    imaginary part stays zero too (the factor is -1); otherwise the phase
    couples the parts, and the complex DFT is F_re + i F_im of the parts."""
    import synth
    from scenarios import apply, E
    base = synth.tm2d_bloch(nsteps=80)
    per = dict(base, bloch=[], boundaries=[(0, 0, "Periodic"), (0, 1, "Periodic")])
    o0 = apply(refcpu.Sim("TM", base["nx"], base["ny"], 1, base["dx"], base["dt"]), dict(base, bloch=[(0, 0, 1.0), (0, 1, 1.0)]))
    op = apply(refcpu.Sim("TM", base["nx"], base["ny"], 1, base["dx"], base["dt"]), per)
    o0.step(80)
    op.step(80)
    np.testing.assert_array_equal(o0.get(E, 2), op.get(E, 2))
    assert np.all(o0.get(E, 2, imag=True) == 0)
    np.testing.assert_array_equal(o0.get_dft(0), op.get_dft(0))
    om = apply(refcpu.Sim("TM", base["nx"], base["ny"], 1, base["dx"], base["dt"]), dict(base, bloch=[(0, 0, -1.0), (0, 1, -1.0)]))
    om.step(80)
    assert np.all(om.get(E, 2, imag=True) == 0) and np.abs(om.get(E, 2) - op.get(E, 2)).max() > 0
    o = apply(refcpu.Sim("TM", base["nx"], base["ny"], 1, base["dx"], base["dt"]), base)
    o.step(80)
    assert np.abs(o.get(E, 2, imag=True)).max() > 1e-3 * np.abs(o.get(E, 2)).max()
