"""Pins of the oracle's adjoint (oracle/refcpu.c, build-defined: the reference
has no adjoint) without any GPU:

* dot-product test <A v, w> = <v, A^T w> of one homogeneous step, block by
  block (D, B, psiE, psiH), so that no block's scale hides another's error;
* the same test with pole materials (Drude, Lorentz, conductive + Lorentz),
  their P and J among the blocks;
* central finite differences (FP64, h = 1e-3) of the eps gradient of a
  CPML + TFSF continuous-wave scenario with a design box, without and with
  pole materials on the wave's path;
* a design box holding the material table's values steps bitwise like the table;
* the light-cone argument of the checkpointed recompute: updating step m only on
  the design box grown by L - m leaves the box's E at step L bitwise exact, and
  one cell less does not.
"""
import numpy as np
import pytest

import synth
from oracle import refcpu
from scenarios import apply

PSI_E, PSI_H = 6, 7


def _dot_sim():
    sc = synth.grid3d_adjoint(n=14, nsteps=4, design=(4, 4, 3), npml=3, monitor=6)
    sc["tfsf"] = None
    sc["dfts"] = []
    rng = np.random.default_rng(7)
    # non-uniform materials (eps and mu) so both material transposes matter
    sc["materials"] = [(1.0, 1.0, [])] + [(1.0 + 11.0 * rng.random(), 0.5 + rng.random(), []) for _ in range(3)]
    matid = np.zeros((14, 14, 14), np.uint8)
    matid[2:9, 3:10, 4:12] = 1
    matid[6:12, 1:7, 2:8] = 2
    matid[1:5, 8:13, 7:13] = 3
    sc["matid"] = matid
    return sc


def _blocks(sim):
    """name -> (getter, setter, list of component codes) of the stepped state."""
    out = {"D": [(0, c) for c in range(3)], "B": [(2, c) for c in range(3)]}
    for f, name in ((PSI_E, "psiE"), (PSI_H, "psiH")):
        out[name] = [(f, 3 * I + J) for I in range(3) for J in range(3) if I != J and sim.get(f, 3 * I + J) is not None]
    return out


def _interior(shape):
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def _dot_sim_poles():
    """_dot_sim with pole materials: a Drude, a Lorentz and a conductive + Lorentz
    block beside a dielectric one (coefficients of order one at this dt)."""
    sc = _dot_sim()
    dt = sc["dt"]
    sc["materials"] = [(1.0, 1.0, []), (2.0, 1.0, [(1, 0.3 / dt, 0.05 / dt, 1.0)]),
                       (1.5, 1.2, [(2, 0.2 / dt, 0.02 / dt, 1.0)]),
                       (3.0, 0.8, [(3, 0.4 * 3.0 * 8.854e-12 / dt, 0.0, 1.0), (2, 0.25 / dt, 0.03 / dt, 1.0)])]
    return sc


POLE_P, POLE_J, P_BAR, J_BAR = 4, 5, 8, 9


def _run_dot(sc, src, dst):
    sim = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    blocks = _blocks(sim)
    inner = _interior(sim.shape)
    # pole state: P and J of every pole of every component, on the cells of pole materials only
    npole = max(len(m[2]) for m in sc["materials"])
    has = np.array([len(m[2]) > 0 for m in sc["materials"]])[sc["matid"]]
    masks = {}
    for name, f in (("P", POLE_P), ("J", POLE_J)):
        blocks[name] = [(f, c, p) for p in range(npole) for c in range(3)]
    def mask(name):
        return inner & has if name in ("P", "J") else inner
    rng = np.random.default_rng(hash((src, dst, len(sc["materials"]))) % (1 << 32))

    def fset(key, a):
        f, c = key[0], key[1]
        sim.set(f, c, a, *key[2:])

    def fget(key):
        return sim.get(key[0], key[1], *key[2:])

    def adj_code(key):
        if len(key) == 3:  # pole P / J: adjoint fields 8 / 9, comp 3 p + I
            return (P_BAR if key[0] == POLE_P else J_BAR), 3 * key[2] + key[1]
        return key

    # v: random in block `src`, zero elsewhere; H, E derived
    v = {}
    for key in blocks[src]:
        a = np.where(mask(src), rng.standard_normal(sim.shape), 0.0)
        v[key] = a
        fset(key, a)
    sim.sync_derived()
    sim.step(1)
    w = {}
    lhs = 0.0
    for key in blocks[dst]:
        a = np.where(mask(dst), rng.standard_normal(sim.shape), 0.0)
        w[key] = a
        lhs += float(np.sum(fget(key) * a))
    sim.adjoint_reset()
    for key, a in w.items():
        sim.set_adjoint(*adj_code(key), a)
    sim.adjoint_step(0)
    sim.adjoint_finish()
    rhs = sum(float(np.sum(a * sim.get_adjoint(*adj_code(key)))) for key, a in v.items())
    # scale of the terms summed: a mistake in a transposed coefficient moves rhs by one of them
    terms = sum(float(np.sum(np.abs(a * sim.get_adjoint(*adj_code(key))))) for key, a in v.items())
    assert terms > 0 or lhs == 0
    assert abs(lhs - rhs) <= 1e-12 * max(terms, 1e-300), (src, dst, lhs, rhs, terms)
    return terms


@pytest.mark.parametrize("src", ["D", "B", "psiE", "psiH"])
@pytest.mark.parametrize("dst", ["D", "B", "psiE", "psiH"])
def test_adjoint_step_dot_product(src, dst):
    _run_dot(_dot_sim(), src, dst)


@pytest.mark.parametrize("src", ["D", "B", "psiE", "P", "J"])
@pytest.mark.parametrize("dst", ["D", "B", "psiH", "P", "J"])
def test_adjoint_step_dot_product_pole_materials(src, dst):
    """The same test on a grid with Drude, Lorentz and conductive + Lorentz
    cells, with their P and J as blocks of the state (on those cells)."""
    terms = _run_dot(_dot_sim_poles(), src, dst)
    if src in ("P", "J") and dst in ("B", "P", "J"):
        assert terms > 0  # P and J reach the new E, hence B, P and J (not D): a dropped term would show


def _grad_scenario():
    return synth.grid3d_adjoint(n=20, nsteps=140, design=(4, 4, 3), npml=4, monitor=8)


def _objective(sc, eps):
    sim = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    sim.set_design(sc["design"][0], eps)
    sim.step(sc["nsteps"])
    F = sim.get_dft(0)
    return float(np.sum(F * F))


@pytest.mark.parametrize("poles", [False, True])
def test_oracle_gradient_matches_central_differences(poles):
    sc = synth.grid3d_adjoint(n=24, nsteps=140, design=(4, 4, 3), npml=4, monitor=8, poles=True) if poles \
        else _grad_scenario()
    box, eps = sc["design"]
    sim = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    J, g = sim.adjoint_eps_gradient(sc["nsteps"], 0, 4)
    assert J > 0
    assert J == pytest.approx(_objective(sc, eps), rel=1e-13)
    rng = np.random.default_rng(11)
    flat = g.reshape(-1)
    cells = rng.choice(flat.size, 8, replace=False)
    h = 1e-3
    fd = []
    for q in cells:
        ep = eps.copy().reshape(-1)
        em = eps.copy().reshape(-1)
        ep[q] += h
        em[q] -= h
        fd.append((_objective(sc, ep.reshape(eps.shape)) - _objective(sc, em.reshape(eps.shape))) / (2 * h))
    fd = np.array(fd)
    err = np.abs(flat[cells] - fd)
    # central differences carry an O(h^2) error; the adjoint is exact to roundoff
    assert np.max(err) <= 1e-5 * np.max(np.abs(fd)), (flat[cells], fd)


def test_design_box_equals_material_table():
    sc = synth.grid3d_cpml_tfsf_dft(n=18, nsteps=30)
    box = [6, 12, 5, 11, 6, 10]
    m = sc["matid"]
    eps_tab = np.array([e for e, _, _ in sc["materials"]])
    eps_box = eps_tab[m[box[4]:box[5], box[2]:box[3], box[0]:box[1]]]
    a = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    b = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    b.set_design(box, eps_box)
    a.step(30)
    b.step(30)
    for c in range(3):
        assert np.array_equal(a.get(1, c), b.get(1, c))
        assert np.array_equal(a.get(3, c), b.get(3, c))


def _cone_run(sc, L, shrink):
    """Step the oracle L + 1 steps from random fields; after step m roll every
    E and H cell outside the design box grown by L - m - shrink (at least 0 on
    the last step) back to its value before the step, as a recompute that only
    updates that region leaves it.  Returns E of the box after step L.

    The CPML psi outside the region is left updated here, where the GPU's
    recompute leaves it stale.  That difference cannot reach the box: a
    cell's psi is read only by that cell's own update, so psi outside the
    region feeds only E and H outside it, which are rolled back (stale) in
    both; inside the region psi was updated on every earlier step in both."""
    from scenarios import E, H
    o = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    rng = np.random.default_rng(11)
    for f in (E, H):
        for c in range(3):
            o.set(f, c, rng.standard_normal(o.shape))
    x0, x1, y0, y1, z0, z1 = sc["design"][0]
    kk, jj, ii = np.meshgrid(np.arange(sc["nz"]), np.arange(sc["ny"]), np.arange(sc["nx"]), indexing="ij")
    for m in range(L + 1):
        d = max(L - m - shrink, 0) if m < L else 0
        inside = ((ii >= x0 - d) & (ii < x1 + d) & (jj >= y0 - d) & (jj < y1 + d) & (kk >= z0 - d) & (kk < z1 + d))
        before = {(f, c): o.get(f, c) for f in (E, H) for c in range(3)}
        o.step(1)
        for (f, c), old in before.items():
            new = o.get(f, c)
            o.set(f, c, np.where(inside, new, old))
    box = (slice(z0, z1), slice(y0, y1), slice(x0, x1))
    return [o.get(E, c)[box] for c in range(3)]


def test_light_cone_recompute_argument():
    """The adjoint's recompute updates step m of a segment ending at L only on
    the design box grown by L - m: E(i) reads H(i-1), H(i) and H(i) reads
    E(i), E(i+1), so stale cells outside never reach the box.  Checked on the
    oracle (CPML, TFSF, random initial fields): the box's E at L equals a
    whole-grid run bit for bit, and a cone one cell narrower does not."""
    sc = synth.grid3d_adjoint(n=20, nsteps=16, design=(4, 4, 3), npml=4, monitor=8)
    L = 12
    full = _cone_run(sc, L, shrink=-100)  # every cell updated every step
    cone = _cone_run(sc, L, shrink=0)
    narrow = _cone_run(sc, L, shrink=1)
    for a, b in zip(cone, full):
        assert np.array_equal(a, b)
    assert any(not np.array_equal(a, b) for a, b in zip(narrow, full))
