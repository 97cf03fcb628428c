"""ctypes binding of the FP64 CPU oracle (oracle/refcpu.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` as the checker.  The product package
(``fdtd_amd``) never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MODES = {"TEM": 1, "TM": 2, "ThreeD": 3, "TE": 4}
# boundary kinds (names of BoundaryUpdates.hpp:17-26; numbering the build's, same as include/fdtd_hip.h)
BC = {"NONE": 0, "PEC": 1, "PMC": 2, "Symmetric": 3, "Antisymmetric": 4, "Periodic": 5}

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    global _LIB
    if _LIB is None:
        # REFCPU_LIB: another build of the same source (the sanitizer build of
        # oracle/Makefile's `asan` target, which the CPU tests run once)
        path = os.environ.get("REFCPU_LIB") or os.path.join(_HERE, "librefcpu.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle refcpu`")
        L = ctypes.CDLL(path)
        for name, res, args in [
            ("rc_eps0", _d, []), ("rc_mu0", _d, []), ("rc_c0", _d, []), ("rc_imp0", _d, []),
            ("rc_pml_coeffs", None, [_d, _d, _d, _d, _dp]),
            ("rc_pml_model", None, [_d, _d, _d, _d, _d, _d, _dp]),
            ("rc_pml_profile", None, [_i, _i, _d, _d, _d, _d, _d, _d, _dp, _dp]),
            ("rc_aux_create", _p, [_i, _d, _d]), ("rc_aux_destroy", None, [_p]),
            ("rc_aux_length", _i, [_p]), ("rc_aux_step", None, [_p, _d]),
            ("rc_aux_interp_E", _d, [_p, _d]), ("rc_aux_interp_H", _d, [_p, _d]),
            ("rc_aux_state", None, [_p, _dp, _dp]),
            ("rc_create", _p, [_i, _i, _i, _i, _d, _d]), ("rc_destroy", None, [_p]),
            ("rc_set_pml_axis", _i, [_p, _i, _i, _dp]),
            ("rc_set_material", _i, [_p, _i, _d, _d, _i, _dp]),
            ("rc_set_material_magnetic", _i, [_p, _i, _i, _dp]),
            ("rc_set_magnetized_drude", _i, [_p, _i, _d, _d, _d, _d, _dp]),
            ("rc_set_fluid_drude", _i, [_p, _i, _d, _i, _dp]),
            ("rc_set_bloch_boundary", _i, [_p, _i, _i, _d, _d]),
            ("rc_get_probes_imag", _i, [_p, _dp, _i]),
            ("rc_set_fluid_state", _i, [_p, _dp, _dp]),
            ("rc_set_matid", _i, [_p, ctypes.c_char_p]),
            ("rc_add_source", _i, [_p, _i, _i, _i, _i, _i, _dp, _i]),
            ("rc_set_tfsf", _i, [_p, _ip, _dp, _dp, _dp, _dp, _i, _dp, _i]),
            ("rc_add_tfsf_plane", _i, [_p, _i, _i, _ip, _dp, _dp, _dp, _dp, _i, _dp, _i]),
            ("rc_add_dft", _i, [_p, _ip, _i, _i, _dp, _dp, _i]),
            ("rc_get_dft", _i, [_p, _i, _dp]),
            ("rc_add_probe", _i, [_p, _i, _i, _i, _i, _i]),
            ("rc_get_probes", _i, [_p, _dp, _i]),
            ("rc_get_field", _i, [_p, _i, _i, _i, _dp]),
            ("rc_set_field", _i, [_p, _i, _i, _i, _dp]),
            ("rc_step", _i, [_p, _i]), ("rc_nthreads", _i, [_i]),
            ("rc_set_design", _i, [_p, _ip, _dp]), ("rc_sync_derived", None, [_p]),
            ("rc_set_boundary", _i, [_p, _i, _i, _i]),
            ("rc_bc_mask", ctypes.c_double, [_i, _i, _i, _i]),
            ("rc_apply_boundaries", None, [_p, _i]),
            ("rc_surface_quantities", _i, [_p, _ip, _dp]),
            ("rc_adjoint_eps_gradient", _i, [_p, _i, _i, _i, _dp, _dp]),
            ("rc_adjoint_reset", None, [_p]), ("rc_adjoint_step", None, [_p, ctypes.c_long]),
            ("rc_adjoint_finish", None, [_p]),
            ("rc_get_adjoint", _i, [_p, _i, _i, _dp]), ("rc_set_adjoint", _i, [_p, _i, _i, _dp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def constants():
    L = lib()
    return {"eps0": L.rc_eps0(), "mu0": L.rc_mu0(), "c0": L.rc_c0(), "imp0": L.rc_imp0()}


def pml_coeffs(K, S, A, dt):
    out = np.zeros(4)
    lib().rc_pml_coeffs(K, S, A, dt, _dptr(out))
    return out


def pml_model(x, m, ma, smax, kmax, amax):
    out = np.zeros(3)
    lib().rc_pml_model(x, m, ma, smax, kmax, amax, _dptr(out))
    return out


def pml_profile(L, n, m, ma, smax, kmax, amax, dt):
    """[L,7] arrays (K,S,A,B,C,F,G) for E and H of the placement idiom."""
    e = np.zeros((L, 7))
    h = np.zeros((L, 7))
    lib().rc_pml_profile(L, n, m, ma, smax, kmax, amax, dt, _dptr(e), _dptr(h))
    return e, h


class Aux:
    def __init__(self, ncells, dx, dt):
        self.h = lib().rc_aux_create(ncells, dx, dt)
        self.L = lib().rc_aux_length(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().rc_aux_destroy(self.h)
            self.h = None

    def step(self, s):
        lib().rc_aux_step(self.h, s)

    def interp_E(self, loc):
        return lib().rc_aux_interp_E(self.h, loc)

    def interp_H(self, loc):
        return lib().rc_aux_interp_H(self.h, loc)

    def state(self):
        e = np.zeros(self.L)
        h = np.zeros(self.L)
        lib().rc_aux_state(self.h, _dptr(e), _dptr(h))
        return e, h


class Sim:
    """FP64 oracle grid.  Field codes: 0 D, 1 E, 2 B, 3 H, 4 P, 5 J."""

    def __init__(self, mode, nx, ny=1, nz=1, dx=1.0, dt=1.0):
        self.mode = mode
        self.shape = (nz if mode == "ThreeD" else 1, ny if mode != "TEM" else 1, nx)
        self.h = lib().rc_create(MODES[mode], nx, ny, nz, dx, dt)
        self.dft_meta = []
        self._dt = dt
        self._dx = dx

    def __del__(self):
        if getattr(self, "h", None):
            lib().rc_destroy(self.h)
            self.h = None

    def set_pml_axis(self, axis, ft, coeffs):
        c = _f64(coeffs)
        assert lib().rc_set_pml_axis(self.h, axis, ft, _dptr(c)) == 0

    def set_pml_idiom(self, axis, n, m, ma, smax, kmax, amax):
        """PML placement idiom (tests/testTM.cpp:46-71) along one axis."""
        L = (self.shape[2], self.shape[1], self.shape[0])[axis]
        e, h = pml_profile(L, n, m, ma, smax, kmax, amax, self._dt)
        self.set_pml_axis(axis, 0, e[:, 3:7])
        self.set_pml_axis(axis, 1, h[:, 3:7])

    def set_material(self, mid, eps_r, mu_r=1.0, poles=()):
        arr = _f64(np.array(poles, dtype=np.float64).reshape(-1)) if len(poles) else np.zeros(1)
        assert lib().rc_set_material(self.h, mid, eps_r, mu_r, len(poles), _dptr(arr)) == 0

    def set_material_magnetic(self, mid, poles=()):
        """Magnetic side of material mid: poles (kind 1 Drude, 2 Lorentz, 3 conductive, freq, gamma, weight)
        on B, H, M, K (mu_r comes from set_material).  get(8, c, pole) is M, get(9, c, pole) K."""
        arr = _f64(np.array(poles, dtype=np.float64).reshape(-1)) if len(poles) else np.zeros(1)
        assert lib().rc_set_material_magnetic(self.h, mid, len(poles), _dptr(arr)) == 0

    def set_magnetized_drude(self, mid, eps_r, mu_r, freq, gamma, bias=None):
        """bias: the fixed field [T] (3 values), or None for the cell's own B."""
        b = None if bias is None else _f64(np.asarray(bias, dtype=np.float64).reshape(3))
        assert lib().rc_set_magnetized_drude(self.h, mid, eps_r, mu_r, freq, gamma, None if b is None else _dptr(b)) == 0

    def set_fluid_drude(self, mid, mu_r=1.0, magnetization=0, bias=None):
        b = None if bias is None else _f64(np.asarray(bias, dtype=np.float64).reshape(3))
        assert lib().rc_set_fluid_drude(self.h, mid, mu_r, magnetization, None if b is None else _dptr(b)) == 0

    def set_fluid_state(self, ne, nu_m):
        ne, nu_m = _f64(np.asarray(ne).reshape(-1)), _f64(np.asarray(nu_m).reshape(-1))
        assert ne.size == nu_m.size == int(np.prod(self.shape))
        assert lib().rc_set_fluid_state(self.h, _dptr(ne), _dptr(nu_m)) == 0

    def set_matid(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.uint8).reshape(-1)
        assert ids.size == int(np.prod(self.shape))
        assert lib().rc_set_matid(self.h, ids.tobytes()) == 0

    def add_source(self, field, comp, i, j, k, table):
        """field: 0 (D), 2 (B), 1 (E) or 3 (H), the field codes of this class;
        rc_add_source numbers them D 0, B 1, E 2, H 3."""
        code = {0: 0, 2: 1, 1: 2, 3: 3}[field]
        t = _f64(table)
        assert lib().rc_add_source(self.h, code, comp, i, j, k, _dptr(t), t.size) == 0

    def set_tfsf(self, box, kvec, pol, polH, origin, aux_ncells, signal):
        b = (ctypes.c_int * 6)(*box)
        k, p, ph, o, s = (_f64(x) for x in (kvec, pol, polH, origin, signal))
        assert lib().rc_set_tfsf(self.h, b, _dptr(k), _dptr(p), _dptr(ph), _dptr(o), aux_ncells,
                                 _dptr(s), s.size) == 0

    def add_tfsf_plane(self, axis, side, cells, kvec, pol, polH, origin, aux_ncells, signal):
        c = (ctypes.c_int * 6)(*cells)
        k, p, ph, o, s = (_f64(x) for x in (kvec, pol, polH, origin, signal))
        assert lib().rc_add_tfsf_plane(self.h, axis, side, c, _dptr(k), _dptr(p), _dptr(ph), _dptr(o),
                                       aux_ncells, _dptr(s), s.size) == 0

    def add_dft(self, region, comp_mask, phaseE, phaseH):
        pe, ph = _f64(phaseE), _f64(phaseH)
        nphase, nfreq = pe.shape[0], pe.shape[1]
        r = (ctypes.c_int * 6)(*region)
        idx = lib().rc_add_dft(self.h, r, comp_mask, nfreq, _dptr(pe), _dptr(ph), nphase)
        ncomp = bin(comp_mask).count("1")
        ncell = (region[1] - region[0]) * (region[3] - region[2]) * (region[5] - region[4])
        self.dft_meta.append((ncomp, nfreq, ncell, region))
        return idx

    def get_dft(self, idx):
        ncomp, nfreq, ncell, region = self.dft_meta[idx]
        out = np.zeros(ncomp * nfreq * 2 * ncell)
        assert lib().rc_get_dft(self.h, idx, _dptr(out)) == 0
        shp = (region[5] - region[4], region[3] - region[2], region[1] - region[0])
        return out.reshape(ncomp, nfreq, 2, *shp)

    def add_probe(self, field, comp, i, j, k=0):
        lib().rc_add_probe(self.h, field, comp, i, j, k)
        self._nprobes = getattr(self, "_nprobes", 0) + 1

    def probes(self, max_steps=1 << 20, imag=False):
        out = np.zeros(max_steps * self._nprobes) if max_steps < (1 << 20) else None
        if out is None:
            out = np.zeros(self._steps_done() * self._nprobes)
            max_steps = self._steps_done()
        n = (lib().rc_get_probes_imag if imag else lib().rc_get_probes)(self.h, _dptr(out), max_steps)
        assert n >= 0
        return out[: n * self._nprobes].reshape(n, self._nprobes)

    def set_bloch_boundary(self, axis, side, factor):
        f = complex(factor)
        assert lib().rc_set_bloch_boundary(self.h, axis, side, f.real, f.imag) == 0

    def _steps_done(self):
        return getattr(self, "_nstep", 0)

    def get(self, field, comp, pole=0, imag=False):
        out = np.zeros(int(np.prod(self.shape)))
        if lib().rc_get_field(self.h, field | (16 if imag else 0), comp, pole, _dptr(out)) != 0:
            return None
        return out.reshape(self.shape)

    def set(self, field, comp, arr, pole=0):
        a = _f64(arr).reshape(-1)
        assert lib().rc_set_field(self.h, field, comp, pole, _dptr(a)) == 0

    def step(self, n=1):
        assert lib().rc_step(self.h, n) == 0
        self._nstep = self._steps_done() + n

    # ---- boundary conditions (BoundaryUpdates.hpp; placement build-defined)
    def set_boundary(self, axis, side, kind):
        """kind: one of BC (name or number); axis 0..2, side 0 (min) / 1 (max)."""
        t = BC[kind] if isinstance(kind, str) else int(kind)
        assert lib().rc_set_boundary(self.h, axis, side, t) == 0, (axis, side, kind)

    def apply_boundaries(self, ft):
        lib().rc_apply_boundaries(self.h, ft)

    def surface_quantities(self, region):
        """(nq, z, y, x) array: 3-D rho, Jex, Jey, Jez, Jhx, Jhy, Jhz; TM Jez, Jhx, Jhy;
        TE rho, Jex, Jey, Jhz."""
        r = [int(v) for v in region]
        nq = {"ThreeD": 7, "TM": 3, "TE": 4}[self.mode]
        shape = (r[5] - r[4], r[3] - r[2], r[1] - r[0])
        out = np.zeros(nq * int(np.prod(shape)))
        assert lib().rc_surface_quantities(self.h, (ctypes.c_int * 6)(*r), _dptr(out)) == 0, region
        return out.reshape((nq,) + shape)

    # ---- adjoint (build-defined)
    def set_design(self, box, eps_r):
        b = (ctypes.c_int * 6)(*box)
        e = _f64(eps_r).reshape(-1)
        assert lib().rc_set_design(self.h, b, _dptr(e)) == 0
        self.design_shape = (box[5] - box[4], box[3] - box[2], box[1] - box[0])

    def sync_derived(self):
        lib().rc_sync_derived(self.h)

    def adjoint_eps_gradient(self, nsteps, dft, comp_mask=4):
        grad = np.zeros(int(np.prod(self.design_shape)))
        obj = ctypes.c_double()
        assert lib().rc_adjoint_eps_gradient(self.h, nsteps, dft, comp_mask, ctypes.byref(obj), _dptr(grad)) == 0
        self._nstep = self._steps_done() + nsteps
        return obj.value, grad.reshape(self.design_shape)

    def adjoint_reset(self):
        lib().rc_adjoint_reset(self.h)

    def adjoint_step(self, n):
        lib().rc_adjoint_step(self.h, n)

    def adjoint_finish(self):
        lib().rc_adjoint_finish(self.h)

    def get_adjoint(self, field, comp):
        out = np.zeros(int(np.prod(self.shape)))
        if lib().rc_get_adjoint(self.h, field, comp, _dptr(out)) != 0:
            return None
        return out.reshape(self.shape)

    def set_adjoint(self, field, comp, arr):
        a = _f64(arr).reshape(-1)
        assert lib().rc_set_adjoint(self.h, field, comp, _dptr(a)) == 0


def nthreads(n=0):
    return lib().rc_nthreads(n)

