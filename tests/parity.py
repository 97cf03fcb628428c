"""Error measures shared by the GPU parity tests."""
import numpy as np

COMP = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]


def bin_errors(a, b, sc):
    """Per DFT bin (monitor, component, frequency): the error over the monitor's cells relative to the norm of
    the monitor's whole field of that type (E or H components together) at that frequency, which is the quantity
    the bin measures; beside it the error relative to the component's own norm and the component's share of the
    field norm (a component that the geometry leaves at round-off level, e.g. Ex of an Ez plane wave, has an
    own-norm error of order 1 that says nothing about the solver).
    Returns {monitor: {component: {"err": [...], "own": [...], "share": [...]}}}."""
    out = {}
    for idx, (_, mask, _, _) in enumerate(sc["dfts"]):
        fa, fb = a.get_dft(idx), b.get_dft(idx)
        assert fa.shape == fb.shape
        comps = [c for c in range(6) if (mask >> c) & 1]
        nf = fa.shape[1]
        field_norm = {}
        for ft in (0, 1):
            rows = [q for q, c in enumerate(comps) if (c >= 3) == bool(ft)]
            field_norm[ft] = [float(np.sqrt(sum(np.sum(fb[q, f] ** 2) for q in rows))) for f in range(nf)]
        mon = {}
        for q, c in enumerate(comps):
            fn = field_norm[int(c >= 3)]
            diff = [float(np.linalg.norm((fa[q, f] - fb[q, f]).reshape(-1))) for f in range(nf)]
            own = [float(np.linalg.norm(fb[q, f].reshape(-1))) for f in range(nf)]
            mon[COMP[c]] = {"err": [d / n if n else d for d, n in zip(diff, fn)],
                            "own": [d / o if o else d for d, o in zip(diff, own)],
                            "share": [o / n if n else 0.0 for o, n in zip(own, fn)]}
        out[f"monitor{idx}"] = mon
    return out


def worst_bin(berr):
    return max((e for m in berr.values() for v in m.values() for e in v["err"]), default=0.0)
