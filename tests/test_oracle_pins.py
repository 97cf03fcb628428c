"""CPU pins of the oracle's build-defined pieces (the reference has no 3-D TFSF
polarisation, no running DFT and no multi-pole weights), each against what the
mathematics fixes rather than a retyped formula:

* 3-D TFSF: with polH = pol x k (the consistent Yee plane wave of the aux
  line's (Ez, Hy) pair) and an axis-aligned k, the scattered-field region of
  an empty grid stays at round-off (leakage invariant), for +-x, y and z;
* running DFT: every accumulator equals the brute-force sum of the stored
  probe trace of that cell times the phase table;
* multi-pole ADE: two identical poles of half weight step like one pole of
  full weight (the polarisation equations are linear in the weight);
* single TFSF planes: a pair bounding a slab of a y/z-periodic vacuum grid
  confines an x-propagating wave to the slab (round-off outside), one plane
  alone injects forward only, and in TM the four face planes of a box step
  bit for bit like the box.
"""
import numpy as np
import pytest

import synth
from oracle import refcpu
from scenarios import E, H, apply, dft_phases, from_golden, rel_l2


def _sim(sc):
    return apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)


@pytest.mark.parametrize("kvec,pol", [
    ((1, 0, 0), (0, 0, 1)), ((-1, 0, 0), (0, 0, 1)), ((0, 1, 0), (0, 0, 1)), ((0, 0, 1), (1, 0, 0)),
])
def test_tfsf_3d_leakage_invariant(kvec, pol):
    sc = synth.grid3d_cpml_tfsf_dft(n=32, nsteps=90, nblocks=0, dft=False, inset=10)
    sc["materials"] = [(1.0, 1.0, [])]
    sc["matid"][:] = 0
    sc["probes"] = []
    k = np.array(kvec, float)
    p = np.array(pol, float)
    sc["tfsf"] = dict(sc["tfsf"], kvec=list(k), pol=list(p), polH=list(np.cross(p, k)))
    sim = _sim(sc)
    sim.step(sc["nsteps"])
    box = sc["tfsf"]["box"]
    n = sc["nx"]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    inside = ((x >= box[0]) & (x <= box[1]) & (y >= box[2]) & (y <= box[3]) & (z >= box[4]) & (z <= box[5]))
    # scattered-field cells two or more cells outside the box and outside the CPML
    lo, hi = sc["pml"]["n"] + 2, n - sc["pml"]["n"] - 3
    far = ((x < box[0] - 1) | (x > box[1] + 1) | (y < box[2] - 1) | (y > box[3] + 1) | (z < box[4] - 1) |
           (z > box[5] + 1))
    body = (x >= lo) & (x <= hi) & (y >= lo) & (y <= hi) & (z >= lo) & (z <= hi)
    for f in (E, H):
        tot = max(np.max(np.abs(sim.get(f, c)[inside])) for c in range(3))
        leak = max(np.max(np.abs(sim.get(f, c)[far & body])) for c in range(3))
        assert tot > 0
        assert leak <= 1e-12 * tot, (f, leak / tot)


def test_running_dft_equals_bruteforce_of_probe_traces():
    sc = synth.grid3d_cpml_tfsf_dft(n=20, nsteps=70, dft=False)
    nsteps = sc["nsteps"]
    freqs = np.array([0.7, 1.0, 1.3]) * 2.99792458e8 / (20 * sc["dx"])
    pe, ph = dft_phases(nsteps, sc["dt"], freqs)
    region = [7, 10, 8, 10, 9, 10]
    sc["dfts"] = [(region, 0b100100, pe, ph)]  # Ez and Hz
    cells = [(i, j, k) for k in range(region[4], region[5]) for j in range(region[2], region[3])
             for i in range(region[0], region[1])]
    sc["probes"] = [(E, 2, i, j, k) for i, j, k in cells] + [(H, 2, i, j, k) for i, j, k in cells]
    sim = _sim(sc)
    sim.step(nsteps)
    tr = sim.probes()
    assert tr.shape == (nsteps, 2 * len(cells))
    F = sim.get_dft(0).reshape(2, len(freqs), 2, len(cells))
    for slot, table in ((0, pe), (1, ph)):
        trace = tr[:, slot * len(cells):(slot + 1) * len(cells)]
        brute = np.einsum("nc,nfr->frc", trace, table)
        assert rel_l2(F[slot], brute) <= 1e-13


@pytest.mark.parametrize("kind", [1, 2])
def test_split_pole_equals_single_pole(kind):
    sc = synth.grid3d_dispersive_sphere(n=22, radius=5, nsteps=60, npml=4, nfreq=2)
    w0 = 4e15 if kind == 2 else 1.37e16
    one = dict(sc, materials=[(1.0, 1.0, []), (2.0, 1.0, [(kind, w0, 1e14, 1.4)])])
    two = dict(sc, materials=[(1.0, 1.0, []), (2.0, 1.0, [(kind, w0, 1e14, 0.7), (kind, w0, 1e14, 0.7)])])
    a, b = _sim(one), _sim(two)
    a.step(60)
    b.step(60)
    for f in (E, H):
        for c in range(3):
            ref = a.get(f, c)
            assert np.max(np.abs(ref)) > 0
            assert rel_l2(b.get(f, c), ref) <= 1e-13


def _slab_fields(faces):
    sc = synth.grid3d_plane_slab(faces=faces)
    sim = _sim(sc)
    sim.step(sc["nsteps"])
    return sc, sim


def test_tfsf_plane_pair_confines_the_wave_to_the_slab():
    """Two single planes (x-MIN at layer a, x-MAX at layer b, each with its own
    line and every interior y, z cell) in a y/z-periodic vacuum grid, k = x:
    outside the total-field slab [a, b] the fields stay at round-off, inside
    the incident wave has passed.  A plausible slip (a normal sign, the wrong
    aux field or offset, a plane corrected one layer off) leaks at O(1)."""
    sc, sim = _slab_fields((0, 1))
    a, b = 14, 33
    npml = sc["pml"]["n"]
    x = np.arange(sc["nx"])
    out = ((x < a - 1) | (x > b + 1)) & (x > npml + 1) & (x < sc["nx"] - npml - 2)
    inside = (x > a) & (x < b)
    for f in (E, H):
        tot = max(np.max(np.abs(sim.get(f, c)[:, :, inside])) for c in range(3))
        leak = max(np.max(np.abs(sim.get(f, c)[:, :, out])) for c in range(3))
        assert tot > 0
        assert leak <= 1e-12 * tot, (f, leak / tot)
    # the wave is uniform in y and z (periodic faces, full-width planes)
    ez = sim.get(E, 2)[1:-1, 1:-1, :]
    assert np.max(np.abs(ez - ez[:1, :1, :])) <= 1e-12 * np.max(np.abs(ez))


def test_tfsf_single_plane_injects_forward_only():
    """One x-MIN plane alone: the scattered-field side (x < a) stays at
    round-off while the wave runs on through x > a into the CPML."""
    sc, sim = _slab_fields((0,))
    a = 14
    npml = sc["pml"]["n"]
    x = np.arange(sc["nx"])
    behind = (x < a - 1) & (x > npml + 1)
    ahead = x > a
    for f in (E, H):
        tot = max(np.max(np.abs(sim.get(f, c)[:, :, ahead])) for c in range(3))
        leak = max(np.max(np.abs(sim.get(f, c)[:, :, behind])) for c in range(3))
        assert tot > 0
        assert leak <= 1e-12 * tot, (f, leak / tot)


def test_plane_faces_equal_the_box_in_tm():
    """In TM a box face's component ranges coincide with a plane over the
    face's cells (x faces: rows by0+1..by1; y faces: columns bx0+1..bx1), so
    the box and four single planes sharing its signal step alike."""
    sc, p, a = from_golden("tm2d_tfsf_oblique")
    t = sc["tfsf"]
    box = t["box"]
    sim_box = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], 1, sc["dx"], sc["dt"]), sc)
    sc2 = dict(sc, tfsf=None, planes=[])
    for axis in (0, 1):
        for side in (0, 1):
            layer = box[2 * axis + side]
            o = 1 - axis
            cells = [0, 0, 0, 0, 0, 1]
            cells[2 * axis], cells[2 * axis + 1] = layer, layer + 1
            cells[2 * o], cells[2 * o + 1] = box[2 * o] + 1, box[2 * o + 1] + 1
            sc2["planes"].append(dict(axis=axis, side=side, cells=cells, kvec=t["kvec"], pol=t["pol"], polH=t["polH"],
                                      origin=t["origin"], aux_ncells=t["aux_ncells"], signal=t["signal"]))
    sim_pl = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], 1, sc["dx"], sc["dt"]), sc2)
    sim_box.step(sc["nsteps"])
    sim_pl.step(sc["nsteps"])
    for f, c in ((E, 2), (H, 0), (H, 1)):
        np.testing.assert_array_equal(sim_pl.get(f, c), sim_box.get(f, c))


def test_tm_reference_default_polarisation_leaks_and_pol_cross_k_does_not():
    """The reference's TM default polH = (k_y, -k_x, 0) (include/Source.hpp:
    577-580), reproduced bit for bit by the golden fixtures, is the negative of
    pol x k: the faces then remove the opposite of the line's wave and the
    scattered-field region carries O(1) of the incident amplitude.  With
    polH = pol x k the same box leaves it at round-off."""
    sc, p, a = from_golden("tm2d_tfsf_x")
    box = sc["tfsf"]["box"]
    res = {}
    for name, polH in (("reference", [0.0, -1.0, 0.0]), ("pol x k", [0.0, 1.0, 0.0])):
        sc["tfsf"]["polH"] = polH
        sim = apply(refcpu.Sim("TM", sc["nx"], sc["ny"], 1, sc["dx"], sc["dt"]), sc)
        sim.step(sc["nsteps"])
        ez = sim.get(E, 2)[0]
        inside = np.abs(ez[box[2] + 1:box[3], box[0] + 1:box[1]]).max()
        far = np.concatenate([ez[11:box[2] - 1, 11:-11].ravel(), ez[box[3] + 2:-11, 11:-11].ravel()])
        res[name] = np.abs(far).max() / inside
    assert res["reference"] > 0.1, res
    assert res["pol x k"] < 1e-12, res
