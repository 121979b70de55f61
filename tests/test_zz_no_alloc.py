"""qx_density_stats_noalloc (the entry point that allocates nothing) against
qx_density_stats on the same psi, through the C ABI: two calls in a row, the
second on a smaller lattice, in both precisions; results bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle as o

pytestmark = pytest.mark.gpu


def _hip():
    """the HIP runtime libqx.so is linked against: the caller's output
    buffer is allocated here, by the caller, as the entry point allows"""
    for line in open("/proc/self/maps"):
        if "libamdhip64.so" in line:
            path = line.split()[-1]
            break
    else:
        raise RuntimeError("libamdhip64 is not mapped")
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def _stats_noalloc(qx, lat, out_dev):
    psi_dev, v_dev = C.c_void_p(), C.c_void_p()
    assert qx.lib().qx_device_state(lat.handle, C.byref(psi_dev), C.byref(v_dev)) == 0
    prec = {"fp32": 0, "fp64": 1}[lat.precision]
    st = qx.lib().qx_density_stats_noalloc(psi_dev, lat.dim, lat.N, lat.L, prec,
                                           qx.lib().qx_stream(lat.handle), out_dev)
    assert st == 0, qx.lib().qx_last_error()
    assert qx.lib().qx_synchronize(lat.handle) == 0
    got = np.full(2, np.nan)
    assert _hip().hipMemcpy(got.ctypes.data, out_dev, 16, 2) == 0   # device to host
    return got


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_noalloc_density_stats_equals_density_stats(qx, prec):
    cases = [(3, 64, 20.0, 4), (2, 32, 10.0, 5)]   # the second call on a smaller lattice
    first = qx.Lattice(*cases[0][:3], prec)   # makes the library's device current
    hip = _hip()
    out = C.c_void_p()
    assert hip.hipMalloc(C.byref(out), 16) == 0
    assert hip.hipMemset(out, 0xFF, 16) == 0
    first.close()
    for dim, N, L, seed in cases:
        print(f"[case] noalloc {prec} {dim}D {N}", flush=True)
        lat = qx.Lattice(dim, N, L, prec)
        lat.psi = o.ic_grf(dim, N, L, seed)
        got = _stats_noalloc(qx, lat, out)
        want = np.array(lat.density_stats())
        print(f"noalloc {got} density_stats {want}")
        assert np.array_equal(got, want), (got, want)
        # and the oracle's statistics of the same state
        r, s = o.density_stats(lat.psi)
        assert got[0] == pytest.approx(r, rel=1e-6 if prec == "fp32" else 1e-13)
        assert got[1] == pytest.approx(s * (L / N) ** dim, rel=1e-6 if prec == "fp32" else 1e-13)
        lat.close()
    assert hip.hipFree(out) == 0
