"""The C ABI library loads and exports every symbol include/qx.h declares;
argument validation happens on the host before any device call (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "qx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    for s in ("qx_create", "qx_destroy", "qx_kick", "qx_drift", "qx_update_potential",
              "qx_leapfrog", "qx_spectrum", "qx_fft_c2c", "qx_fft_r2c", "qx_fft_c2r",
              "qx_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol(qx):
    lib = ctypes.CDLL(qx.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding covers the whole ABI
    assert set(declared_symbols()) <= set(qx.SIGNATURES)


def test_shim_and_pybind_module_load(qx):
    ctypes.CDLL(os.path.join(ROOT, "qaxions_amd", "libqxshim.so"))
    import qaxions_amd.fieldtest as ft
    for name in ("test_kick", "test_drift", "test_potential", "test_kick_drift_step"):
        assert callable(getattr(ft, name))
    import qaxions_amd.ictest as ic
    assert callable(ic.test_solitons) and callable(ic.test_solitons_3D)


@pytest.mark.parametrize("dim,N,L,prec,msg", [
    (4, 8, 1.0, "fp32", "Dimension"), (1, 8, 1.0, "fp64", "Dimension"),
    (2, 12, 1.0, "fp32", "power of two"), (3, 2048, 1.0, "fp32", "power of two"),
    (2, 8, -1.0, "fp64", "positive"), (2, 1, 1.0, "fp64", "power of two")])
def test_create_rejects_bad_arguments(qx, dim, N, L, prec, msg):
    with pytest.raises(qx.QxError) as e:
        qx.Lattice(dim, N, L, precision=prec)
    assert e.value.status == 1  # QX_EINVAL
    assert msg in str(e.value)


def test_null_context_is_an_error(qx):
    lib = qx.lib()
    assert lib.qx_kick(None, 0.1) == 1
    assert b"null" in lib.qx_last_error()
    assert lib.qx_destroy(None) == 0


@pytest.mark.parametrize("dim,N,L,msg", [(4, 8, 1.0, "Dimension"), (3, 2048, 1.0, "power of two"),
                                         (2, 12, 1.0, "power of two"), (2, 8, 0.0, "positive")])
def test_noalloc_density_stats_rejects_bad_sizes(qx, dim, N, L, msg):
    """qx_density_stats_noalloc checks its lattice on the host, before any
    device call (no GPU here)."""
    lib = qx.lib()
    buf = ctypes.create_string_buffer(16)
    assert lib.qx_density_stats_noalloc(ctypes.addressof(buf), dim, N, L, 0, None, ctypes.addressof(buf)) == 1
    assert msg.encode() in lib.qx_last_error()
    assert lib.qx_density_stats_noalloc(None, 2, 8, 1.0, 0, None, ctypes.addressof(buf)) == 1
