"""Python binding of the C ABI (include/fdtd_hip.h) over ctypes.

``Grid`` owns one ``fdtd_grid`` handle.  Method names follow the set-up steps
of the reference's user time loop (PML placement, materials, soft sources,
TFSF planes, probes) so the same scenario description drives this class and
the CPU oracle.  Loading fails loudly when the HIP library is missing: there
is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfdtd_hip.so")

TEM, TM, THREED, TE = 1, 2, 3, 4
MODES = {"TEM": TEM, "TM": TM, "TE": TE, "ThreeD": THREED}
F32, F64 = 0, 1
D, E, B, H, P, J, M, K = 0, 1, 2, 3, 4, 5, 6, 7
DRUDE, LORENTZ, CONDUCTIVE = 1, 2, 3
IMAG = 16  # field-code flag: the imaginary part of a complex grid (FDTD_IMAG)
# boundary kinds (fdtd_boundary in include/fdtd_hip.h; names of the reference's fdtd::Boundary)
BOUNDARY = {"NONE": 0, "PEC": 1, "PMC": 2, "Symmetric": 3, "Antisymmetric": 4, "Periodic": 5, "BlochPeriodic": 6}

# every symbol declared in include/fdtd_hip.h
ABI_SYMBOLS = [
    "fdtd_grid_create", "fdtd_grid_destroy", "fdtd_set_stream", "fdtd_set_pml_axis", "fdtd_set_pml_idiom",
    "fdtd_set_material", "fdtd_set_material_magnetic", "fdtd_set_magnetized_drude", "fdtd_set_fluid_drude", "fdtd_set_fluid_state", "fdtd_set_matid", "fdtd_add_source", "fdtd_set_tfsf", "fdtd_add_tfsf_plane", "fdtd_set_signal", "fdtd_add_dft", "fdtd_get_dft",
    "fdtd_add_probe", "fdtd_get_probes", "fdtd_upload_field", "fdtd_download_field", "fdtd_step",
    "fdtd_synchronize", "fdtd_step_count", "fdtd_set_profiling", "fdtd_kernel_time", "fdtd_last_error",
    "fdtd_set_tile", "fdtd_get_tile", "fdtd_autotune", "fdtd_set_adjoint_tile",
    "fdtd_batch_create", "fdtd_batch_destroy", "fdtd_batch_set_graph", "fdtd_batch_step", "fdtd_batch_synchronize",
    "fdtd_set_design", "fdtd_adjoint_eps_gradient", "fdtd_adjoint_eps_gradient_batch",
    "fdtd_batch_set_tile", "fdtd_batch_set_persistent", "fdtd_batch_autotune", "fdtd_set_graph", "fdtd_set_boundary",
    "fdtd_set_bloch_boundary", "fdtd_get_probes_imag",
    "fdtd_surface_quantities",
    "fdtd_surface_quantities_static",
]


class GridDesc(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("precision", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("nz", ctypes.c_int), ("dx", ctypes.c_double), ("dt", ctypes.c_double)]


_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_LIB = None


def load_library():
    """Load libfdtd_hip.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m fdtd_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "fdtd_grid_create": (_i, [ctypes.POINTER(GridDesc), ctypes.POINTER(_p)]),
        "fdtd_grid_destroy": (_i, [_p]),
        "fdtd_set_stream": (_i, [_p, _p]),
        "fdtd_set_pml_axis": (_i, [_p, _i, _i, _dp, _dp, _dp]),
        "fdtd_set_pml_idiom": (_i, [_p, _i, _i, _d, _d, _d, _d, _d]),
        "fdtd_set_material": (_i, [_p, _i, _d, _d, _i, _dp]),
        "fdtd_set_material_magnetic": (_i, [_p, _i, _i, _dp]),
        "fdtd_set_magnetized_drude": (_i, [_p, _i, _d, _d, _d, _d, _dp]),
        "fdtd_set_fluid_drude": (_i, [_p, _i, _d, _i, _dp]),
        "fdtd_set_fluid_state": (_i, [_p, _dp, _dp]),
        "fdtd_set_matid": (_i, [_p, ctypes.c_char_p]),
        "fdtd_add_source": (_i, [_p, _i, _i, _i, _i, _i, _dp, _i]),
        "fdtd_set_tfsf": (_i, [_p, _ip, _dp, _dp, _dp, _dp, _i, _dp, _i]),
        "fdtd_add_tfsf_plane": (_i, [_p, _i, _i, _ip, _dp, _dp, _dp, _dp, _i, _dp, _i]),
        "fdtd_set_signal": (_i, [_p, _i, _i, ctypes.c_long, _dp, _i]),
        "fdtd_add_dft": (_i, [_p, _ip, _i, _i, _dp, _dp, _i, _ip]),
        "fdtd_get_dft": (_i, [_p, _i, _dp]),
        "fdtd_add_probe": (_i, [_p, _i, _i, _i, _i, _i]),
        "fdtd_get_probes": (_i, [_p, _dp, _i, _ip]),
        "fdtd_get_probes_imag": (_i, [_p, _dp, _i, _ip]),
        "fdtd_set_bloch_boundary": (_i, [_p, _i, _i, _d, _d]),
        "fdtd_upload_field": (_i, [_p, _i, _i, _i, _dp]),
        "fdtd_download_field": (_i, [_p, _i, _i, _i, _dp]),
        "fdtd_step": (_i, [_p, _i]),
        "fdtd_synchronize": (_i, [_p]),
        "fdtd_step_count": (ctypes.c_long, [_p]),
        "fdtd_set_profiling": (_i, [_p, _i]),
        "fdtd_kernel_time": (_i, [_p, _i, _dp, ctypes.POINTER(ctypes.c_long)]),
        "fdtd_last_error": (ctypes.c_char_p, []),
        "fdtd_set_tile": (_i, [_p, _i, _i, _i]),
        "fdtd_set_graph": (_i, [_p, _i]),
        "fdtd_set_boundary": (_i, [_p, _i, _i, _i]),
        "fdtd_surface_quantities": (_i, [_p, _ip, _dp]),
        "fdtd_surface_quantities_static": (_i, [_p, _ip, _dp]),
        "fdtd_get_tile": (_i, [_p, _ip, _ip, _ip]),
        "fdtd_autotune": (_i, [_p, _i, _ip, _ip, _ip]),
        "fdtd_set_adjoint_tile": (_i, [_p, _i, _i, _i]),
        "fdtd_batch_create": (_i, [ctypes.POINTER(_p), _i, ctypes.POINTER(_p)]),
        "fdtd_batch_destroy": (_i, [_p]),
        "fdtd_batch_set_graph": (_i, [_p, _i]),
        "fdtd_batch_step": (_i, [_p, _i]),
        "fdtd_batch_synchronize": (_i, [_p]),
        "fdtd_batch_set_tile": (_i, [_p, _i, _i, _i]),
        "fdtd_batch_set_persistent": (_i, [_p, _i]),
        "fdtd_batch_autotune": (_i, [_p, _i, _ip, _ip, _ip]),
        "fdtd_set_design": (_i, [_p, _ip, _dp]),
        "fdtd_adjoint_eps_gradient": (_i, [_p, _i, _i, _d, _i, _i, _dp, _dp]),
        "fdtd_adjoint_eps_gradient_batch": (_i, [ctypes.POINTER(_p), _i, _i, _i, _d, _i, _i, _dp,
                                                 ctypes.POINTER(_dp)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


class FdtdError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        msg = load_library().fdtd_last_error().decode()
        raise FdtdError(f"fdtd error {rc}: {msg}")


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Grid:
    """One simulation on the GPU.  ``mode`` is "TEM" (a line along x: ny = nz = 1), "TM", "TE" or "ThreeD";
    ``precision`` "f32" or "f64"."""

    def __init__(self, mode, nx, ny, nz=1, dx=1.0, dt=1.0, precision="f64", stream=None):
        L = load_library()
        self.mode = mode
        ny = 1 if mode == "TEM" else ny
        self.shape = (nz if mode == "ThreeD" else 1, ny, nx)
        if mode not in MODES:
            raise ValueError("mode must be one of %s" % ", ".join(MODES))
        desc = GridDesc(MODES[mode], F64 if precision == "f64" else F32, nx, ny, nz if mode == "ThreeD" else 1, dx, dt)
        h = _p()
        _check(L.fdtd_grid_create(ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        self.dft_meta = []
        self._nprobes = 0
        if stream is not None:
            _check(L.fdtd_set_stream(self.h, stream))

    def close(self):
        if getattr(self, "h", None):
            load_library().fdtd_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up
    def set_pml_idiom(self, axis, n, m, ma, smax, kmax, amax):
        _check(load_library().fdtd_set_pml_idiom(self.h, axis, n, m, ma, smax, kmax, amax))

    def set_pml_axis(self, axis, ft, K, S, A):
        K, S, A = _f64(K), _f64(S), _f64(A)
        _check(load_library().fdtd_set_pml_axis(self.h, axis, ft, _dptr(K), _dptr(S), _dptr(A)))

    def set_material(self, mid, eps_r, mu_r=1.0, poles=()):
        arr = _f64(np.array(poles, dtype=np.float64).reshape(-1)) if len(poles) else np.zeros(1)
        _check(load_library().fdtd_set_material(self.h, mid, eps_r, mu_r, len(poles), _dptr(arr)))

    def set_material_magnetic(self, mid, poles=()):
        """Magnetic side of material mid: poles (kind DRUDE / LORENTZ / CONDUCTIVE, freq, gamma,
        weight) on B, H, M, K; mu_r is set_material's.  get(M, c, pole) / get(K, c, pole) read them."""
        arr = _f64(np.array(poles, dtype=np.float64).reshape(-1)) if len(poles) else np.zeros(1)
        _check(load_library().fdtd_set_material_magnetic(self.h, mid, len(poles), _dptr(arr)))

    def set_magnetized_drude(self, mid, eps_r, mu_r, freq, gamma, bias=None):
        """Magnetized Drude material: bias = the fixed field [T] (MagnetizedDrudeUpdate),
        None = the cell's own B (SelfMagnetizedDrudeUpdate)."""
        b = None if bias is None else _f64(np.asarray(bias, dtype=np.float64).reshape(3))
        _check(load_library().fdtd_set_magnetized_drude(self.h, mid, eps_r, mu_r, freq, gamma,
                                                        None if b is None else _dptr(b)))

    def set_fluid_drude(self, mid, mu_r=1.0, magnetization=0, bias=None):
        """Fluid Drude material (omega_p, gamma per cell from set_fluid_state): magnetization
        0 FluidDrudeUpdate, 1 MagnetizedFluidDrudeUpdate in `bias` [T], 2 SelfMagnetizedFluidDrudeUpdate."""
        b = None if bias is None else _f64(np.asarray(bias, dtype=np.float64).reshape(3))
        _check(load_library().fdtd_set_fluid_drude(self.h, mid, mu_r, int(magnetization),
                                                   None if b is None else _dptr(b)))

    def set_fluid_state(self, ne, nu_m):
        """Per-cell electron density [1/m^3] and collision rate [1/s], shape of the grid."""
        ne, nu_m = _f64(np.asarray(ne).reshape(-1)), _f64(np.asarray(nu_m).reshape(-1))
        if ne.size != int(np.prod(self.shape)) or nu_m.size != ne.size:
            raise ValueError("fluid state size mismatch")
        _check(load_library().fdtd_set_fluid_state(self.h, _dptr(ne), _dptr(nu_m)))

    def set_matid(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.uint8).reshape(-1)
        if ids.size != int(np.prod(self.shape)):
            raise ValueError("matid size mismatch")
        _check(load_library().fdtd_set_matid(self.h, ids.tobytes()))

    def add_source(self, field, comp, i, j, k, table):
        t = _f64(table)
        _check(load_library().fdtd_add_source(self.h, field, comp, i, j, k, _dptr(t), t.size))

    def set_tfsf(self, box, kvec, pol, polH, origin, aux_ncells, signal):
        b = (ctypes.c_int * 6)(*box)
        k, p, ph, o, s = (_f64(x) for x in (kvec, pol, polH, origin, signal))
        _check(load_library().fdtd_set_tfsf(self.h, b, _dptr(k), _dptr(p), _dptr(ph), _dptr(o), aux_ncells,
                                            _dptr(s), s.size))

    def add_tfsf_plane(self, axis, side, cells, kvec, pol, polH, origin, aux_ncells, signal):
        """One TFSF plane with its own incident line (fdtd_add_tfsf_plane): normal
        along `axis` toward `side` (0 MIN, 1 MAX), cells [x0,x1)x[y0,y1)x[z0,z1)."""
        c = (ctypes.c_int * 6)(*cells)
        k, p, ph, o, s = (_f64(x) for x in (kvec, pol, polH, origin, signal))
        _check(load_library().fdtd_add_tfsf_plane(self.h, axis, side, c, _dptr(k), _dptr(p), _dptr(ph), _dptr(o),
                                                  aux_ncells, _dptr(s), s.size))

    SIGNAL = {"source": 0, "tfsf": 1, "plane": 2}

    def set_signal(self, kind, index, step0, values):
        """Write values as the entries of steps step0.. of a source's table
        (kind "source"), the TFSF box's signal ("tfsf") or a plane's ("plane")."""
        v = _f64(values)
        _check(load_library().fdtd_set_signal(self.h, self.SIGNAL[kind], index, step0, _dptr(v), v.size))

    def add_dft(self, region, comp_mask, phaseE, phaseH):
        pe, ph = _f64(phaseE), _f64(phaseH)
        nphase, nfreq = pe.shape[0], pe.shape[1]
        r = (ctypes.c_int * 6)(*region)
        idx = ctypes.c_int()
        _check(load_library().fdtd_add_dft(self.h, r, comp_mask, nfreq, _dptr(pe), _dptr(ph), nphase,
                                           ctypes.byref(idx)))
        region = list(region)
        if self.mode != "ThreeD":
            region[4], region[5] = 0, 1
        ncomp = bin(comp_mask).count("1")
        ncell = (region[1] - region[0]) * (region[3] - region[2]) * (region[5] - region[4])
        self.dft_meta.append((ncomp, nfreq, ncell, region))
        return idx.value

    def get_dft(self, idx):
        ncomp, nfreq, ncell, region = self.dft_meta[idx]
        out = np.zeros(ncomp * nfreq * 2 * ncell)
        _check(load_library().fdtd_get_dft(self.h, idx, _dptr(out)))
        shp = (region[5] - region[4], region[3] - region[2], region[1] - region[0])
        return out.reshape(ncomp, nfreq, 2, *shp)

    def add_probe(self, field, comp, i, j, k=0):
        _check(load_library().fdtd_add_probe(self.h, field, comp, i, j, k))
        self._nprobes += 1

    def probes(self, imag=False):
        """[steps][probes]; imag: the imaginary part of a complex grid's probes."""
        n = self.step_count()
        out = np.zeros(max(1, n * self._nprobes))
        got = ctypes.c_int()
        fn = load_library().fdtd_get_probes_imag if imag else load_library().fdtd_get_probes
        _check(fn(self.h, _dptr(out), n, ctypes.byref(got)))
        return out[: got.value * self._nprobes].reshape(got.value, self._nprobes)

    # ---- fields
    def get(self, field, comp, pole=0, imag=False):
        """Field array (z, y, x); imag: the imaginary part of E or H of a complex grid."""
        out = np.zeros(int(np.prod(self.shape)))
        _check(load_library().fdtd_download_field(self.h, field | (IMAG if imag else 0), comp, pole, _dptr(out)))
        return out.reshape(self.shape)

    def set(self, field, comp, arr, pole=0, imag=False):
        a = _f64(arr).reshape(-1)
        _check(load_library().fdtd_upload_field(self.h, field | (IMAG if imag else 0), comp, pole, _dptr(a)))

    # ---- stepping
    def step(self, n=1):
        _check(load_library().fdtd_step(self.h, n))

    def synchronize(self):
        _check(load_library().fdtd_synchronize(self.h))

    def step_count(self):
        return load_library().fdtd_step_count(self.h)

    def set_tile(self, bx, by, kc):
        _check(load_library().fdtd_set_tile(self.h, bx, by, kc))

    def set_boundary(self, axis, side, kind):
        """Boundary condition of one face: kind is a name of BOUNDARY (the reference's
        fdtd::Boundary kinds) or its number; axis 0..2, side 0 (min) / 1 (max)."""
        k = BOUNDARY[kind] if isinstance(kind, str) else int(kind)
        _check(load_library().fdtd_set_boundary(self.h, int(axis), int(side), k))

    def set_bloch_boundary(self, axis, side, factor):
        """Bloch-periodic face with the complex phase factor `factor`: the grid's
        fields become complex (make_bloch_periodic_boundary)."""
        f = complex(factor)
        _check(load_library().fdtd_set_bloch_boundary(self.h, int(axis), int(side), f.real, f.imag))

    def surface_quantities(self, region, static=False):
        """Surface charge density and currents of the current state on `region`
        (x0, x1, y0, y1, z0, z1): array (nq, z, y, x); 3-D rho, Jex, Jey, Jez, Jhx,
        Jhy, Jhz; TM Jez, Jhx, Jhy; TE rho, Jex, Jey, Jhz.  static=True: the entry point that allocates
        nothing (regions of at most FDTD_SURFACE_STATIC_MAX_CELLS cells, a grid
        already set up on the device)."""
        r = [int(v) for v in region]
        nq = {"ThreeD": 7, "TM": 3, "TE": 4}.get(self.mode, 1)  # TEM: refused by the library
        shape = (r[5] - r[4], r[3] - r[2], r[1] - r[0])
        out = np.zeros(nq * int(np.prod(shape)))
        fn = load_library().fdtd_surface_quantities_static if static else load_library().fdtd_surface_quantities
        _check(fn(self.h, (ctypes.c_int * 6)(*r), _dptr(out)))
        return out.reshape((nq,) + shape)

    def set_graph(self, on=True):
        """Capture chunks of steps in hipGraphs and replay them (fewer launches for small grids)."""
        _check(load_library().fdtd_set_graph(self.h, 1 if on else 0))

    def get_tile(self):
        v = [ctypes.c_int() for _ in range(3)]
        _check(load_library().fdtd_get_tile(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_adjoint_tile(self, bx, by, kc):
        """Tile of the adjoint gradient's reverse sweep and recompute ((0, 0, 0): the grid's tile)."""
        _check(load_library().fdtd_set_adjoint_tile(self.h, bx, by, kc))

    def autotune(self, trial_steps=20):
        """Pick the fastest tile by timing; the simulation state is left unchanged."""
        v = [ctypes.c_int() for _ in range(3)]
        _check(load_library().fdtd_autotune(self.h, trial_steps, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_profiling(self, on=True):
        _check(load_library().fdtd_set_profiling(self.h, 1 if on else 0))

    def kernel_time(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_long()
        _check(load_library().fdtd_kernel_time(self.h, kernel, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # ---- design and adjoint
    def set_design(self, box, eps_r):
        """Per-cell eps_r over box [i0,i1,j0,j1,k0,k1); eps_r shaped (k, j, i) or flat, x fastest."""
        b = (ctypes.c_int * 6)(*box)
        e = _f64(eps_r).reshape(-1)
        _check(load_library().fdtd_set_design(self.h, b, _dptr(e)))
        self.design_box = list(box)
        self.design_shape = (box[5] - box[4], box[3] - box[2], box[1] - box[0])

    def adjoint_eps_gradient(self, nsteps, dft, comp_mask=4, checkpoint_every=100, history_mb=4096.0):
        """Step nsteps from step 0 and return (J, dJ/d eps_r) for J = sum |F|^2 of monitor `dft`."""
        grad = np.zeros(int(np.prod(self.design_shape)))
        obj = ctypes.c_double()
        _check(load_library().fdtd_adjoint_eps_gradient(self.h, nsteps, checkpoint_every, float(history_mb), dft,
                                                        comp_mask, ctypes.byref(obj), _dptr(grad)))
        return obj.value, grad.reshape(self.design_shape)


def adjoint_eps_gradient_batch(grids, nsteps, dft, comp_mask=4, checkpoint_every=100, history_mb=4096.0):
    """Gradients of several designs (grids at step 0 with a design box each) computed concurrently;
    returns a list of (J, dJ/d eps_r) in grid order."""
    n = len(grids)
    hs = (_p * n)(*[g.h for g in grids])
    grads = [np.zeros(int(np.prod(g.design_shape))) for g in grids]
    gp = (_dp * n)(*[_dptr(x) for x in grads])
    obj = np.zeros(n)
    _check(load_library().fdtd_adjoint_eps_gradient_batch(hs, n, nsteps, checkpoint_every, float(history_mb), dft,
                                                          comp_mask, _dptr(obj), gp))
    return [(float(obj[q]), grads[q].reshape(grids[q].design_shape)) for q in range(n)]


class Batch:
    """Independent 3-D grids stepped together (one launch per pass for all,
    hipGraph-captured chunks).  The grids must outlive the batch."""

    def __init__(self, grids, graph=True):
        L = load_library()
        self.grids = list(grids)
        arr = (_p * len(self.grids))(*[g.h for g in self.grids])
        h = _p()
        _check(L.fdtd_batch_create(arr, len(self.grids), ctypes.byref(h)))
        self.h = h
        if not graph:
            _check(L.fdtd_batch_set_graph(self.h, 0))

    def step(self, n=1):
        _check(load_library().fdtd_batch_step(self.h, n))

    def synchronize(self):
        _check(load_library().fdtd_batch_synchronize(self.h))

    def set_tile(self, bx, by, kc):
        _check(load_library().fdtd_batch_set_tile(self.h, bx, by, kc))

    def set_persistent(self, wg_per_cu):
        """Persistent passes: wg_per_cu workgroups per compute unit walk the tile table (0: one per tile)."""
        _check(load_library().fdtd_batch_set_persistent(self.h, int(wg_per_cu)))

    def autotune(self, trial_steps=10):
        """Pick the fastest tile by timing whole batch steps; the grids' states are left unchanged."""
        v = [ctypes.c_int() for _ in range(3)]
        _check(load_library().fdtd_batch_autotune(self.h, trial_steps, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def close(self):
        if getattr(self, "h", None):
            load_library().fdtd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
