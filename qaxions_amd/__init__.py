"""qaxions_amd -- MI355X-native kick-drift-kick Schroedinger-Poisson step.

Python access to ``libqx.so`` (HIP kernels for gfx950 behind the C ABI in
``include/qx.h``).  There is no CPU fallback: if the library is missing or no
GPU is visible, every entry point raises.

    >>> from qaxions_amd import Lattice
    >>> lat = Lattice(dim=3, N=64, L=20.0, precision="fp32")
    >>> lat.ic_grf(seed=0)
    >>> lat.update_potential()
    >>> lat.leapfrog(ds=5e-4, nsteps=10)
    >>> psi = lat.psi          # complex128 numpy array, shape (N,)*dim

``Lattice`` is the Python face of one reference ``Field`` (src/field.h:13-97);
the pybind11 module ``qaxions_amd.fieldtest`` keeps the reference's own
test entry points (src/examples/fieldtest.cxx:99-105).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

__all__ = ["Lattice", "Ensemble", "Snapshot", "Corrector", "QxError", "lib", "LIB_PATH", "physics_static"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqx.so")

QX_OK, QX_EINVAL, QX_ENOMEM, QX_EDEVICE, QX_ESTATE = range(5)


class QxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"qx status {status}: {msg}")
        self.status = status


class Physics(C.Structure):
    _fields_ = [("pref", C.c_double), ("rho_bar", C.c_double),
                ("k0_mode", C.c_int), ("rho_bar_mode", C.c_int)]


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i64p = C.POINTER(C.c_int64)
_lib = None

# every symbol include/qx.h declares, with its ctypes signature
SIGNATURES = {
    "qx_create": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "qx_destroy": (C.c_int, [C.c_void_p]),
    "qx_set_physics": (C.c_int, [C.c_void_p, C.POINTER(Physics)]),
    "qx_get_physics": (C.c_int, [C.c_void_p, C.POINTER(Physics)]),
    "qx_upload_psi": (C.c_int, [C.c_void_p, _dp]),
    "qx_download_psi": (C.c_int, [C.c_void_p, _dp]),
    "qx_upload_V": (C.c_int, [C.c_void_p, _dp]),
    "qx_download_V": (C.c_int, [C.c_void_p, _dp]),
    "qx_copy_state": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qx_device_state": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "qx_snap_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "qx_snap_take": (C.c_int, [C.c_void_p]),
    "qx_snap_wait": (C.c_int, [C.c_void_p, _dp, _dp]),
    "qx_snap_destroy": (C.c_int, [C.c_void_p]),
    "qx_kick": (C.c_int, [C.c_void_p, C.c_double]),
    "qx_drift": (C.c_int, [C.c_void_p, C.c_double]),
    "qx_update_potential": (C.c_int, [C.c_void_p, _dp]),
    "qx_poisson_half": (C.c_int, [C.c_void_p, _dp]),
    "qx_compute_energy": (C.c_int, [C.c_void_p, _dp]),
    "qx_density_stats": (C.c_int, [C.c_void_p, _dp, _dp]),
    "qx_density_stats_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                           C.c_void_p, C.c_void_p]),
    "qx_energy": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "qx_vmax": (C.c_double, [C.c_void_p]),
    "qx_spectrum": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _i64p, _dp]),
    "qx_leapfrog": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int]),
    "qx_leapfrog_schedule": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "qx_propagate": (C.c_int, [C.c_void_p, C.c_double]),
    "qx_fft_c2c": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "qx_fft_r2c": (C.c_int, [C.c_void_p, _dp, _dp]),
    "qx_fft_c2r": (C.c_int, [C.c_void_p, _dp, _dp]),
    "qx_fft_r2c_V": (C.c_int, [C.c_void_p, C.c_double, _dp]),
    "qx_sites": (C.c_int64, [C.c_void_p]),
    "qx_ksites": (C.c_int64, [C.c_void_p]),
    "qx_global_N": (C.c_int, [C.c_void_p]),
    "qx_ic_solitons": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_double]),
    "qx_ic_grf": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double]),
    "qx_ic_blobs": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_double]),
    "qx_step_pass_count": (C.c_int, [C.c_void_p]),
    "qx_step_pass_name": (C.c_char_p, [C.c_void_p, C.c_int]),
    "qx_time_step_passes": (C.c_int, [C.c_void_p, C.c_double, C.c_int, _dp, _dp, C.c_int]),
    "qx_ens_create": (C.c_int, [C.c_int, C.POINTER(C.c_int), _dp, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "qx_ens_destroy": (C.c_int, [C.c_void_p]),
    "qx_ens_set_physics": (C.c_int, [C.c_void_p, C.POINTER(Physics)]),
    "qx_ens_set_graph_steps": (C.c_int, [C.c_void_p, C.c_int]),
    "qx_ens_upload_psi": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "qx_ens_download_psi": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "qx_ens_ic_grf": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_double]),
    "qx_ens_update_potential": (C.c_int, [C.c_void_p]),
    "qx_ens_leapfrog": (C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int]),
    "qx_ens_tune": (C.c_int, [C.c_void_p, _dp, C.c_int, C.POINTER(C.c_int)]),
    "qx_ens_set_tile_variants": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "qx_ens_classes": (C.c_int, [C.c_void_p]),
    "qx_ens_class_size": (C.c_int, [C.c_void_p, C.c_int]),
    "qx_ens_mass": (C.c_int, [C.c_void_p, _dp]),
    "qx_ens_size": (C.c_int, [C.c_void_p]),
    "qx_ens_sites": (C.c_int64, [C.c_void_p]),
    "qx_ens_stream": (C.c_void_p, [C.c_void_p]),
    "qx_dim": (C.c_int, [C.c_void_p]),
    "qx_precision_of": (C.c_int, [C.c_void_p]),
    "qx_net_param_count": (C.c_int, []),
    "qx_net_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "qx_net_destroy": (C.c_int, [C.c_void_p]),
    "qx_net_set_params": (C.c_int, [C.c_void_p, _fp]),
    "qx_net_get_params": (C.c_int, [C.c_void_p, _fp]),
    "qx_net_set_output_scale": (C.c_int, [C.c_void_p, C.c_float]),
    "qx_net_output_scale": (C.c_float, [C.c_void_p]),
    "qx_net_set_input_scale": (C.c_int, [C.c_void_p, C.c_float]),
    "qx_net_input_scale": (C.c_float, [C.c_void_p]),
    "qx_net_train_step_lattices": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, _dp]),
    "qx_net_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "qx_net_forward_host": (C.c_int, [C.c_void_p, _fp, _fp]),
    "qx_net_loss_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp, _fp]),
    "qx_net_loss_grad_host": (C.c_int, [C.c_void_p, _fp, _fp, _dp, _fp]),
    "qx_net_adam_step": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float]),
    "qx_net_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, _dp]),
    "qx_net_correct": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qx_net_pair_mse": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _dp]),
    "qx_net_coarse_steps": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.POINTER(C.c_int)]),
    "qx_net_set_energy_tol": (C.c_int, [C.c_void_p, C.c_double]),
    "qx_net_stream": (C.c_void_p, [C.c_void_p]),
    "qx_net_last_error": (C.c_char_p, []),
    "qx_tune": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_int)]),
    "qx_set_tile_variants": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "qx_set_stream_io": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    "qx_stream": (C.c_void_p, [C.c_void_p]),
    "qx_synchronize": (C.c_int, [C.c_void_p]),
    "qx_last_error": (C.c_char_p, []),
    "qx_version": (C.c_char_p, []),
}


def lib():
    """Load libqx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m qaxions_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status):
    if status != QX_OK:
        raise QxError(status, lib().qx_last_error().decode())


def _p(a):
    return a.ctypes.data_as(_dp)


def physics_static(norm=1.0, k0_mode=1, rho_bar_mode=0):
    """STATIC cosmology constants (src/field.cxx:73-80): pref = 4 pi norm, rho_bar = 1."""
    return dict(pref=4.0 * math.pi * norm, rho_bar=1.0, k0_mode=k0_mode,
                rho_bar_mode=rho_bar_mode)


class Lattice:
    """One device-resident lattice (a reference ``Field``) behind the C ABI."""

    def __init__(self, dim, N, L, precision="fp32", device=0, pref=1.0, rho_bar=1.0,
                 k0_mode=1, rho_bar_mode=0):
        self.dim, self.N, self.L = int(dim), int(N), float(L)
        self.precision = precision
        prec = {"fp32": 0, "fp64": 1}[precision]
        h = C.c_void_p()
        _check(lib().qx_create(self.dim, self.N, self.L, prec, int(device), C.byref(h)))
        self._h = h
        self.set_physics(pref=pref, rho_bar=rho_bar, k0_mode=k0_mode, rho_bar_mode=rho_bar_mode)

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().qx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def shape(self):
        return (self.N,) * self.dim

    @property
    def half_shape(self):
        return (self.N,) * (self.dim - 1) + (self.N // 2 + 1,)

    @property
    def sites(self):
        return int(lib().qx_sites(self._h))

    @property
    def ksites(self):
        return int(lib().qx_ksites(self._h))

    # -- physics --------------------------------------------------------------
    def set_physics(self, pref, rho_bar=1.0, k0_mode=1, rho_bar_mode=0):
        ph = Physics(float(pref), float(rho_bar), int(k0_mode), int(rho_bar_mode))
        _check(lib().qx_set_physics(self._h, C.byref(ph)))

    def get_physics(self):
        ph = Physics()
        _check(lib().qx_get_physics(self._h, C.byref(ph)))
        return dict(pref=ph.pref, rho_bar=ph.rho_bar, k0_mode=ph.k0_mode,
                    rho_bar_mode=ph.rho_bar_mode)

    # -- state ------------------------------------------------------------------
    @property
    def psi(self):
        out = np.empty(self.shape, np.complex128)
        _check(lib().qx_download_psi(self._h, _p(out.view(np.float64))))
        return out

    @psi.setter
    def psi(self, value):
        a = np.ascontiguousarray(value, dtype=np.complex128)
        if a.shape != self.shape:
            raise ValueError(f"psi must have shape {self.shape}")
        _check(lib().qx_upload_psi(self._h, _p(a.view(np.float64))))

    @property
    def V(self):
        out = np.empty(self.shape, np.float64)
        _check(lib().qx_download_V(self._h, _p(out)))
        return out

    @V.setter
    def V(self, value):
        a = np.ascontiguousarray(value, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"V must have shape {self.shape}")
        _check(lib().qx_upload_V(self._h, _p(a)))

    def copy_from(self, other):
        """device-to-device copy of psi, V and the potential spectrum"""
        _check(lib().qx_copy_state(self._h, other.handle))

    # -- operators ---------------------------------------------------------------
    def kick(self, dt):
        _check(lib().qx_kick(self._h, float(dt)))

    def drift(self, dt):
        _check(lib().qx_drift(self._h, float(dt)))

    def update_potential(self):
        v = C.c_double()
        _check(lib().qx_update_potential(self._h, C.byref(v)))
        return v.value

    def poisson_half(self, vhat):
        a = np.ascontiguousarray(vhat, dtype=np.complex128).copy()
        if a.shape != self.half_shape:
            raise ValueError(f"half spectrum must have shape {self.half_shape}")
        _check(lib().qx_poisson_half(self._h, _p(a.view(np.float64))))
        return a

    def compute_energy(self):
        r = C.c_double()
        _check(lib().qx_compute_energy(self._h, C.byref(r)))
        return r.value

    def density_stats(self):
        r, m = C.c_double(), C.c_double()
        _check(lib().qx_density_stats(self._h, C.byref(r), C.byref(m)))
        return r.value, m.value

    def energy(self):
        m, k, p = C.c_double(), C.c_double(), C.c_double()
        _check(lib().qx_energy(self._h, C.byref(m), C.byref(k), C.byref(p)))
        return m.value, k.value, p.value

    @property
    def vmax(self):
        return float(lib().qx_vmax(self._h))

    def spectrum(self, source="V", nbins=None, spb=1):
        src = {"V": 0, "rho": 1}[source]
        if nbins is None:
            nbins = self.N // 2 + 1
        P = np.zeros(nbins)
        cnt = np.zeros(nbins, np.int64)
        k = np.zeros(nbins)
        _check(lib().qx_spectrum(self._h, src, int(nbins), int(spb), _p(P),
                                 cnt.ctypes.data_as(_i64p), _p(k)))
        return P, cnt, k

    def leapfrog(self, ds, nsteps, first_half=True, last_half=True):
        _check(lib().qx_leapfrog(self._h, float(ds), int(nsteps), int(bool(first_half)),
                                 int(bool(last_half))))

    def leapfrog_schedule(self, drift_dt, kick_dt, pref):
        """merged-kick loop with per-step constants: K(kick_dt[0]), then per
        step j D(drift_dt[j]), P with pref[j], K(kick_dt[j+1]) (0: none)"""
        dd = np.ascontiguousarray(drift_dt, np.float64)
        kk = np.ascontiguousarray(kick_dt, np.float64)
        pp = np.ascontiguousarray(pref, np.float64)
        if kk.size != dd.size + 1 or pp.size != dd.size:
            raise ValueError("need len(kick_dt) == len(drift_dt) + 1 == len(pref) + 1")
        _check(lib().qx_leapfrog_schedule(self._h, int(dd.size), _p(dd), _p(kk), _p(pp)))

    def propagate(self, ds):
        _check(lib().qx_propagate(self._h, float(ds)))

    # -- FFTBackend ------------------------------------------------------------------
    def fft_c2c(self, x, sign):
        a = np.ascontiguousarray(x, dtype=np.complex128).copy()
        _check(lib().qx_fft_c2c(self._h, _p(a.view(np.float64)), int(sign)))
        return a

    def fft_r2c(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(self.half_shape, np.complex128)
        _check(lib().qx_fft_r2c(self._h, _p(a), _p(out.view(np.float64))))
        return out

    def fft_r2c_V(self, scale=1.0):
        """forward r2c of the lattice's own V, times scale (reference layout)"""
        out = np.empty(self.half_shape, np.complex128)
        _check(lib().qx_fft_r2c_V(self._h, float(scale), _p(out.view(np.float64))))
        return out

    def fft_c2r(self, xh):
        a = np.ascontiguousarray(xh, dtype=np.complex128)
        out = np.empty(self.shape, np.float64)
        _check(lib().qx_fft_c2r(self._h, _p(a.view(np.float64)), _p(out)))
        return out

    # -- initial conditions -------------------------------------------------------------
    def ic_solitons(self, solitons, bkg=1.0):
        t = np.ascontiguousarray(np.asarray(solitons, np.float64).reshape(-1, 5))
        _check(lib().qx_ic_solitons(self._h, _p(t), t.shape[0], float(bkg)))

    def ic_grf(self, seed, amp=0.1):
        _check(lib().qx_ic_grf(self._h, int(seed), float(amp)))

    def ic_blobs(self, table, bkg=0.1):
        t = np.ascontiguousarray(np.asarray(table, np.float64).reshape(-1, 6))
        _check(lib().qx_ic_blobs(self._h, _p(t), t.shape[0], float(bkg)))

    # -- measurement ------------------------------------------------------------------------
    def time_step_passes(self, ds, reps=10):
        """Per-pass average ms (HIP events) and minimum bytes of the fused step."""
        n = int(lib().qx_step_pass_count(self._h))
        ms = np.zeros(n)
        by = np.zeros(n)
        _check(lib().qx_time_step_passes(self._h, float(ds), int(reps), _p(ms), _p(by), n))
        names = [lib().qx_step_pass_name(self._h, i).decode() for i in range(n)]
        return names, ms, by

    def tune(self, ds, reps=5, cache=None):
        """Pick the strided tile variant of every pass by measurement.  With a
        cache path (JSON), a previous result for this (dim, N, precision) is
        reused, else the new one is stored."""
        import json
        key = f"{self.dim}d-{self.N}-{self.precision}"
        if cache and os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
            if key in db:
                v = (C.c_int * 8)(*db[key])
                _check(lib().qx_set_tile_variants(self._h, v))
                return list(db[key])
        v = (C.c_int * 8)()
        _check(lib().qx_tune(self._h, float(ds), int(reps), v))
        out = list(v)
        if cache:
            db = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    db = json.load(f)
            db[key] = out
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(db, f, indent=1)
        return out

    def set_tile_variants(self, v):
        _check(lib().qx_set_tile_variants(self._h, (C.c_int * 8)(*v)))

    def set_stream_io(self, on):
        """Streaming (non-temporal) loads and stores in the step's passes;
        returns the setting in force (False where N < 512)."""
        out = C.c_int()
        _check(lib().qx_set_stream_io(self._h, 1 if on else 0, C.byref(out)))
        return bool(out.value)

    # -- plumbing ---------------------------------------------------------------------------
    @property
    def stream(self):
        return lib().qx_stream(self._h)

    def synchronize(self):
        _check(lib().qx_synchronize(self._h))


class Snapshot:
    """Asynchronous copy of a lattice's psi and/or V (qx_snap_*): take()
    queues device copies in stream order and returns; wait() returns the
    arrays as they were at take() while later steps run on."""

    def __init__(self, lattice, psi=True, V=True):
        what = (1 if psi else 0) | (2 if V else 0)
        h = C.c_void_p()
        _check(lib().qx_snap_create(lattice._h, what, C.byref(h)))
        self._h, self._lat, self._psi, self._V = h, lattice, bool(psi), bool(V)

    def take(self):
        _check(lib().qx_snap_take(self._h))

    def wait(self):
        psi = np.empty(self._lat.shape, np.complex128) if self._psi else None
        V = np.empty(self._lat.shape, np.float64) if self._V else None
        _check(lib().qx_snap_wait(self._h, None if psi is None else _p(psi.view(np.float64)),
                                  None if V is None else _p(V)))
        return psi, V

    def close(self):
        if getattr(self, "_h", None):
            lib().qx_snap_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def blob_table(nblobs, seed, L):
    """Blob parameters (rho_c, r_c, x, y, z, phi) drawn from the keyed RNG:
    rho_c ~ U[1,10], r_c ~ U[0.3,1], centre ~ U[0,L)^3, phi ~ U[0,2pi)."""
    from .rng import uniform
    t = np.zeros((nblobs, 6))
    for j in range(nblobs):
        st = 16 + j
        t[j] = [1.0 + 9.0 * uniform(seed, st, 0), 0.3 + 0.7 * uniform(seed, st, 1),
                L * uniform(seed, st, 2), L * uniform(seed, st, 3), L * uniform(seed, st, 4),
                2.0 * math.pi * uniform(seed, st, 5)]
    return t


class Ensemble:
    """Independent 2D lattices of uneven sizes advanced together (one launch
    per size class and pass, hipGraph-captured steps)."""

    def __init__(self, Ns, Ls, precision="fp32", device=0, pref=1.0, rho_bar=1.0, k0_mode=1,
                 rho_bar_mode=0, graph_steps=8):
        self.Ns = [int(n) for n in Ns]
        self.Ls = [float(x) for x in Ls]
        self.precision = precision
        n = len(self.Ns)
        h = C.c_void_p()
        _check(lib().qx_ens_create(n, (C.c_int * n)(*self.Ns), (C.c_double * n)(*self.Ls),
                                   {"fp32": 0, "fp64": 1}[precision], int(device), C.byref(h)))
        self._h = h
        ph = Physics(float(pref), float(rho_bar), int(k0_mode), int(rho_bar_mode))
        _check(lib().qx_ens_set_physics(self._h, C.byref(ph)))
        self.set_graph_steps(graph_steps)

    def close(self):
        if getattr(self, "_h", None):
            lib().qx_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.Ns)

    @property
    def sites(self):
        return int(lib().qx_ens_sites(self._h))

    @property
    def stream(self):
        return lib().qx_ens_stream(self._h)

    def set_graph_steps(self, k):
        _check(lib().qx_ens_set_graph_steps(self._h, int(k)))

    def ic_grf(self, seeds, amp=0.1):
        n = len(self.Ns)
        _check(lib().qx_ens_ic_grf(self._h, (C.c_uint64 * n)(*[int(x) for x in seeds]), float(amp)))

    def set_psi(self, m, psi):
        a = np.ascontiguousarray(psi, dtype=np.complex128)
        assert a.shape == (self.Ns[m],) * 2
        _check(lib().qx_ens_upload_psi(self._h, int(m), _p(a.view(np.float64))))

    def psi(self, m):
        out = np.empty((self.Ns[m],) * 2, np.complex128)
        _check(lib().qx_ens_download_psi(self._h, int(m), _p(out.view(np.float64))))
        return out

    def update_potential(self):
        _check(lib().qx_ens_update_potential(self._h))

    def _classes(self):
        n = int(lib().qx_ens_classes(self._h))
        return [int(lib().qx_ens_class_size(self._h, i)) for i in range(n)]

    def tune(self, ds, reps=3, cache=None):
        """Autotune each size class's drift tile by timing its steps (the
        state is kept); returns {N: variant} (0 full tile, 1 half tile).
        With a cache path (JSON), a result stored for the same class make-up
        (sizes, member counts, precision) is reinstalled instead of timed,
        else the new result is stored."""
        import json
        sizes = self._classes()
        key = "ens-" + self.precision + "-" + ",".join(f"{n}x{list(self.Ns).count(n)}" for n in sizes)
        db = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
            if key in db:
                got = {int(k): int(v) for k, v in db[key].items()}
                self.set_tile_variants(got)
                return got
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(ds, np.float64), (len(self.Ns),)))
        v = (C.c_int * max(len(sizes), 1))()
        _check(lib().qx_ens_tune(self._h, _p(d), int(reps), v))
        out = {n: int(v[i]) for i, n in enumerate(sizes)}
        if cache:
            db[key] = {str(k): x for k, x in out.items()}
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as f:
                json.dump(db, f, indent=1)
        return out

    def set_tile_variants(self, variants):
        """Reinstall {N: variant} (or one variant per class in increasing N)."""
        sizes = self._classes()
        vals = [variants[n] for n in sizes] if isinstance(variants, dict) else list(variants)
        if len(vals) != len(sizes):
            raise ValueError(f"{len(sizes)} size classes, {len(vals)} variants")
        _check(lib().qx_ens_set_tile_variants(self._h, (C.c_int * max(len(vals), 1))(*vals)))

    def leapfrog(self, ds, nsteps, first_half=True, last_half=True):
        n = len(self.Ns)
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(ds, np.float64), (n,)))
        _check(lib().qx_ens_leapfrog(self._h, _p(d), int(nsteps), int(bool(first_half)),
                                     int(bool(last_half))))

    def mass(self):
        out = np.zeros(len(self.Ns))
        _check(lib().qx_ens_mass(self._h, _p(out)))
        return out


def _check_net(status):
    if status != QX_OK:
        raise QxError(status, lib().qx_net_last_error().decode())


class Corrector:
    """Learned coarse-step corrector psi <- psi + s net(psi, a V) on H x W
    lattices (s: output scale, a: input scale of the potential channel)."""

    def __init__(self, H, W, device=0, params=None):
        self.H, self.W = int(H), int(W)
        h = C.c_void_p()
        _check_net(lib().qx_net_create(self.H, self.W, int(device), C.byref(h)))
        self._h = h
        self.n_params = int(lib().qx_net_param_count())
        if params is not None:
            self.params = params

    def close(self):
        if getattr(self, "_h", None):
            lib().qx_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.zeros(self.n_params, np.float32)
        _check_net(lib().qx_net_get_params(self._h, out.ctypes.data_as(_fp)))
        return out

    @params.setter
    def params(self, theta):
        t = np.ascontiguousarray(theta, np.float32)
        assert t.size == self.n_params
        _check_net(lib().qx_net_set_params(self._h, t.ctypes.data_as(_fp)))

    @property
    def output_scale(self):
        return float(lib().qx_net_output_scale(self._h))

    @output_scale.setter
    def output_scale(self, s):
        _check_net(lib().qx_net_set_output_scale(self._h, float(s)))

    @property
    def input_scale(self):
        return float(lib().qx_net_input_scale(self._h))

    @input_scale.setter
    def input_scale(self, a):
        _check_net(lib().qx_net_set_input_scale(self._h, float(a)))

    def forward(self, x):
        """x: [H, W, 3] float array (Re psi, Im psi, a V) -> x[..., :2] + s net(x)"""
        x = np.ascontiguousarray(x, np.float32)
        assert x.shape == (self.H, self.W, 3)
        y = np.empty((self.H, self.W, 2), np.float32)
        _check_net(lib().qx_net_forward_host(self._h, x.ctypes.data_as(_fp), y.ctypes.data_as(_fp)))
        return y

    def loss_grad(self, x, t):
        """x: [H, W, 3] input, t: [H, W, 2] target -> (loss, gradient)"""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        assert x.shape == (self.H, self.W, 3) and t.shape == (self.H, self.W, 2)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _check_net(lib().qx_net_loss_grad_host(self._h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp),
                                               C.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def train_step_lattices(self, coarse, fine, lr):
        """one Adam step on the pair ((coarse.psi, coarse.V) -> fine.psi), both
        FP32 Lattices"""
        loss = C.c_double()
        _check_net(lib().qx_net_train_step_lattices(self._h, coarse.handle, fine.handle, float(lr),
                                                    C.byref(loss)))
        return loss.value

    def pair_mse(self, coarse, fine):
        """mean((coarse.psi - fine.psi)^2): the identity corrector's loss"""
        pc, pf = C.c_void_p(), C.c_void_p()
        _check(lib().qx_device_state(coarse.handle, C.byref(pc), None))
        _check(lib().qx_device_state(fine.handle, C.byref(pf), None))
        m = C.c_double()
        _check_net(lib().qx_net_pair_mse(self._h, pc, pf, C.byref(m)))
        return m.value

    def correct(self, lat):
        _check_net(lib().qx_net_correct(self._h, lat.handle))

    def set_energy_tol(self, tol):
        """coarse_steps' energy gate: chunks whose relative energy change
        exceeds tol fall back to fine steps (0 turns it off)"""
        _check_net(lib().qx_net_set_energy_tol(self._h, float(tol)))

    def coarse_steps(self, lat, dt, k, nsteps, check_every=10, mass_tol=1e-6):
        """nsteps corrected coarse steps of k*dt; returns the number of chunks
        that fell back to fine stepping"""
        fb = C.c_int()
        _check_net(lib().qx_net_coarse_steps(self._h, lat.handle, float(dt), int(k), int(nsteps),
                                             int(check_every), float(mass_tol), C.byref(fb)))
        return int(fb.value)
