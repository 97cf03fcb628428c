"""GPU adjoint eps gradient (fdtd_adjoint_eps_gradient through the C ABI).

* FP64 GPU vs the FP64 oracle adjoint (an independent D/B-form restatement):
  objective rel 1e-12, gradient rel-L2 1e-9;
* FP32 GPU vs the FP64 oracle adjoint: gradient rel-L2 1e-3;
* checkpoint + recompute gives bitwise the gradient of fully recorded history,
  and the grid ends in bitwise the state of a plain forward run;
* FP32 gradient vs central finite differences (h = 1e-3) of FP64 GPU
  forward runs on 16 random design cells: 1e-3 of the largest derivative;
* with pole materials on the wave's path (Drude, Lorentz, conductive +
  Lorentz): the same oracle bounds, recompute bitwise equal to recorded
  history (the checkpoints carry the pole cells' D, P, J), FP32 vs central
  differences; magnetized and fluid materials are refused;
* the torch.autograd binding returns the same gradient.
"""
import numpy as np
import pytest

import synth
from oracle import refcpu
from scenarios import E, H, apply, rel_l2

pytestmark = pytest.mark.gpu


def _grid(sc, precision):
    from fdtd_amd import Grid
    return Grid(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"], precision=precision)


def _small():
    return synth.grid3d_adjoint(n=20, nsteps=140, design=(4, 4, 3), npml=4, monitor=8)


def _oracle_grad(sc):
    o = apply(refcpu.Sim(sc["mode"], sc["nx"], sc["ny"], sc["nz"], sc["dx"], sc["dt"]), sc)
    return o.adjoint_eps_gradient(sc["nsteps"], 0, 4)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32", 1e-3)])
def test_adjoint_matches_oracle(precision, tol):
    sc = _small()
    print(f"[gpu-case] adjoint {sc['name']} {precision}", flush=True)
    J0, g0 = _oracle_grad(sc)
    g = apply(_grid(sc, precision), sc)
    J, grad = g.adjoint_eps_gradient(sc["nsteps"], 0, 4, checkpoint_every=32, history_mb=0.0)
    g.close()
    assert J == pytest.approx(J0, rel=1e-12 if precision == "f64" else 1e-5)
    err = rel_l2(grad, g0)
    print(f"adjoint {precision}: J {J:.6e} vs {J0:.6e}, gradient rel-L2 {err:.3e}", flush=True)
    assert err <= tol


def _poles():
    return synth.grid3d_adjoint(n=24, nsteps=140, design=(4, 4, 3), npml=4, monitor=8, poles=True)


@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32", 1e-3)])
def test_adjoint_with_pole_materials_matches_oracle(precision, tol):
    sc = _poles()
    print(f"[gpu-case] adjoint poles {sc['name']} {precision}", flush=True)
    J0, g0 = _oracle_grad(sc)
    g = apply(_grid(sc, precision), sc)
    J, grad = g.adjoint_eps_gradient(sc["nsteps"], 0, 4, checkpoint_every=32, history_mb=0.0)
    g.close()
    assert J == pytest.approx(J0, rel=1e-12 if precision == "f64" else 1e-5)
    err = rel_l2(grad, g0)
    print(f"adjoint poles {precision}: J {J:.6e} vs {J0:.6e}, gradient rel-L2 {err:.3e}", flush=True)
    assert err <= tol
    # the pole cells matter: the gradient of the same grid without them differs
    sv = synth.grid3d_adjoint(n=24, nsteps=140, design=(4, 4, 3), npml=4, monitor=8)
    _, gv = _oracle_grad(sv)
    assert rel_l2(gv, g0) > 1e-3


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_adjoint_with_pole_materials_recompute_is_bitwise(precision):
    sc = _poles()
    print(f"[gpu-case] adjoint poles recompute {precision}", flush=True)
    runs = []
    for K, mb, rtile in [(140, 1e3, None), (13, 0.0, None), (40, 0.0, (64, 4, 4)), (13, 0.05, None)]:
        g = apply(_grid(sc, precision), sc)
        if rtile:
            g.set_adjoint_tile(*rtile)
        runs.append(g.adjoint_eps_gradient(sc["nsteps"], 0, 4, checkpoint_every=K, history_mb=mb))
        g.close()
    for J, grad in runs[1:]:
        assert J == runs[0][0]
        assert np.array_equal(grad, runs[0][1])


def test_adjoint_refuses_magnetized_and_fluid():
    from fdtd_amd import FdtdError
    sc = _poles()
    for kind in ("magnetized", "fluid"):
        g = apply(_grid(sc, "f64"), sc)
        if kind == "magnetized":
            g.set_magnetized_drude(1, 1.0, 1.0, 1e14, 1e12, [0.0, 0.0, 1.0])
        else:
            g.set_fluid_drude(1, 1.0, 0, None)
        with pytest.raises(FdtdError):
            g.adjoint_eps_gradient(sc["nsteps"], 0, 4)
        g.close()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_checkpoint_recompute_is_bitwise(precision):
    sc = synth.grid3d_adjoint(n=24, nsteps=150, design=(6, 5, 3), npml=4, monitor=8)
    print(f"[gpu-case] adjoint recompute {sc['name']} {precision}", flush=True)
    runs = []
    # (K, history MB): all recorded, 11 of 12 segments recomputed, none, 2 of 4, and a few resident
    # (the last two: the reverse sweep and recompute at tiles of their own)
    for K, mb, rtile in [(150, 1e3, None), (13, 0.0, None), (13, 1e3, None), (40, 0.0, None), (13, 0.08, None),
                         (13, 0.0, (64, 4, 4)), (40, 0.08, (256, 1, 2))]:
        g = apply(_grid(sc, precision), sc)
        if rtile:
            g.set_adjoint_tile(*rtile)
        J, grad = g.adjoint_eps_gradient(sc["nsteps"], 0, 4, checkpoint_every=K, history_mb=mb)
        fields = [g.get(E, c) for c in range(3)] + [g.get(H, c) for c in range(3)]
        runs.append((J, grad, fields, g.get_dft(0), g.step_count()))
        g.close()
    p = apply(_grid(sc, precision), sc)
    p.step(sc["nsteps"])
    plain = [p.get(E, c) for c in range(3)] + [p.get(H, c) for c in range(3)]
    plain_dft = p.get_dft(0)
    p.close()
    for J, grad, fields, dft, steps in runs:
        assert steps == sc["nsteps"]
        assert J == runs[0][0]
        assert np.array_equal(grad, runs[0][1])
        for a, b in zip(fields, plain):
            assert np.array_equal(a, b)
        assert np.array_equal(dft, plain_dft)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_light_cone_recompute_is_bitwise(precision):
    """A grid much larger than the design box: the recompute of a segment
    runs each step's passes only over the box grown by the steps left in the
    segment, and must still give bitwise the gradient of fully recorded
    history, at the grid's tile and at tiles of its own."""
    sc = synth.grid3d_adjoint(n=48, nsteps=120, design=(6, 5, 3), npml=6, monitor=12)
    print(f"[gpu-case] adjoint light cone {sc['name']} {precision}", flush=True)
    ref = apply(_grid(sc, precision), sc)
    J0, g0 = ref.adjoint_eps_gradient(sc["nsteps"], 0, 4, checkpoint_every=120, history_mb=1e3)
    ref.close()
    assert np.any(g0 != 0)
    for K, rtile in [(40, None), (30, (64, 4, 4)), (40, (256, 1, 2)), (7, (64, 2, 4))]:
        g = apply(_grid(sc, precision), sc)
        if rtile:
            g.set_adjoint_tile(*rtile)
        J, grad = g.adjoint_eps_gradient(sc["nsteps"], 0, 4, checkpoint_every=K, history_mb=0.0)
        g.close()
        assert J == J0
        assert np.array_equal(grad, g0), f"K {K} tile {rtile}"


@pytest.mark.parametrize("poles", [False, True])
def test_adjoint_f32_matches_finite_differences(poles):
    sc = synth.grid3d_adjoint(n=40, nsteps=500, design=(16, 16, 6), npml=6, monitor=20, poles=poles)
    box, eps = sc["design"]
    print(f"[gpu-case] adjoint FD {sc['name']}", flush=True)
    g = apply(_grid(sc, "f32"), sc)
    J, grad = g.adjoint_eps_gradient(sc["nsteps"], 0, 4, checkpoint_every=100, history_mb=0.0)
    g.close()

    def objective(e):
        q = apply(_grid(sc, "f64"), dict(sc, design=(box, e)))
        q.step(sc["nsteps"])
        F = q.get_dft(0)
        q.close()
        return float(np.sum(F * F))

    rng = np.random.default_rng(16)
    cells = rng.choice(eps.size, 16, replace=False)
    h = 1e-3
    fd = []
    for c in cells:
        ep = eps.reshape(-1).copy()
        em = eps.reshape(-1).copy()
        ep[c] += h
        em[c] -= h
        fd.append((objective(ep.reshape(eps.shape)) - objective(em.reshape(eps.shape))) / (2 * h))
    fd = np.array(fd)
    got = grad.reshape(-1)[cells]
    rel = np.abs(got - fd) / np.max(np.abs(fd))
    print("FD check: max |adjoint - FD| / max|FD| =", float(rel.max()), flush=True)
    assert rel.max() <= 1e-3


def test_torch_binding_gradient():
    import torch
    from fdtd_amd.adjoint import eps_objective

    sc = _small()
    box, eps = sc["design"]
    print(f"[gpu-case] adjoint torch {sc['name']}", flush=True)

    def make():
        s = dict(sc)
        s.pop("design")
        return apply(_grid(sc, "f64"), s)

    e = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    J = eps_objective(e, make, box, sc["nsteps"], 0)
    J.backward()
    g = apply(_grid(sc, "f64"), sc)
    J2, grad = g.adjoint_eps_gradient(sc["nsteps"], 0, 4)
    g.close()
    assert float(J) == J2
    assert np.array_equal(e.grad.numpy(), grad)


def test_design_batch_equals_single_calls():
    """Several designs at once (one thread and stream each) give each design's
    single-call objective and gradient bitwise."""
    from fdtd_amd.grid import adjoint_eps_gradient_batch
    base = synth.grid3d_adjoint(n=24, nsteps=120, design=(6, 5, 3), npml=4, monitor=8)
    rng = np.random.default_rng(8)
    designs = [1.0 + 11.0 * rng.random(base["design"][1].shape) for _ in range(4)]
    print("[gpu-case] adjoint design batch x4", flush=True)
    singles = []
    for e in designs:
        g = apply(_grid(base, "f32"), dict(base, design=(base["design"][0], e)))
        singles.append(g.adjoint_eps_gradient(base["nsteps"], 0, 4, checkpoint_every=25, history_mb=0.0))
        g.close()
    grids = [apply(_grid(base, "f32"), dict(base, design=(base["design"][0], e))) for e in designs]
    got = adjoint_eps_gradient_batch(grids, base["nsteps"], 0, 4, checkpoint_every=25, history_mb=0.0)
    for g in grids:
        g.close()
    for (J, grad), (J1, grad1) in zip(got, singles):
        assert J == J1
        assert np.array_equal(grad, grad1)
