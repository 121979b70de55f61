"""Build the native parts of qaxions_amd in-tree.

  libqx.so          HIP kernels + C ABI (include/qx.h), hipcc --offload-arch=gfx950
  libqxshim.so      C++ Field/FFTBackend/Spectrum mirror of the reference, over the C ABI
  fieldtest*.so     pybind11 module with the reference's fieldtest entry points
  ictest*.so        pybind11 module with the reference's ictest entry points
  libqxio.so        HDF5 configuration files (IO, SnapshotWriter) and measure(),
                    built when an HDF5 C library is found (HDF5_DIR, default /opt/conda)
  h5test*.so        pybind11 module with the reference's h5test entry points
  qx_run            the reference's time-loop program over the device Field (C++)

Run ``python -m qaxions_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QX_OFFLOAD_ARCH", "gfx950")

# packed-FP32 instructions are disabled for all device code
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]

LIB = os.path.join(HERE, "libqx.so")
SHIM = os.path.join(HERE, "libqxshim.so")
IOLIB = os.path.join(HERE, "libqxio.so")
RUN = os.path.join(HERE, "qx_run")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "unknown target feature" not in l and "packed-fp32-ops" not in l
                    and l.strip() and not l.startswith("1 warning"))
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}")
    return out


def _hip_objects():
    srcs = ["qx_lib_f32.hip", "qx_lib_f64.hip", "qx_lib.hip", "qx_net.hip"]
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(HERE, "_build", s.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    return objs, cmds


def build_lib():
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    objs, cmds = _hip_objects()
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", LIB + ".tmp"])
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_shim():
    """C++ host mirror of the reference Field/FFTBackend/Spectrum + pybind11 module."""
    import pybind11
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(CSRC, "shim")]
    shim_srcs = [os.path.join(CSRC, "shim", f) for f in
                 ("field.cxx", "spectrum.cxx", "solitons.cxx", "utils.cxx")]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc, *shim_srcs,
          "-L", HERE, "-lqx", "-Wl,-rpath,$ORIGIN", "-o", SHIM + ".tmp"])
    os.replace(SHIM + ".tmp", SHIM)
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    mod = os.path.join(HERE, "fieldtest" + ext)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc,
          "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
          os.path.join(CSRC, "shim", "fieldtest.cxx"),
          "-L", HERE, "-lqxshim", "-lqx", "-Wl,-rpath,$ORIGIN", "-o", mod + ".tmp"])
    os.replace(mod + ".tmp", mod)
    icmod = os.path.join(HERE, "ictest" + ext)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc,
          "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
          os.path.join(CSRC, "shim", "ictest.cxx"),
          "-L", HERE, "-lqxshim", "-lqx", "-Wl,-rpath,$ORIGIN", "-o", icmod + ".tmp"])
    os.replace(icmod + ".tmp", icmod)
    return SHIM, mod


def hdf5_found() -> bool:
    return (os.path.isfile(os.path.join(HDF5_DIR, "include", "hdf5.h")) and
            os.path.isfile(os.path.join(HDF5_DIR, "lib", "libhdf5.so")))


def build_io():
    """libqxio.so and the h5test module.  libqxio's RUNPATH names the HDF5
    directory after $ORIGIN; the C++ runtime it needs is the one its loader
    (libqxshim, h5test or a program) has mapped already."""
    import pybind11
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(CSRC, "shim"),
           "-I", os.path.join(HDF5_DIR, "include")]
    h5lib = os.path.join(HDF5_DIR, "lib")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc,
          os.path.join(CSRC, "shim", "io.cxx"), os.path.join(CSRC, "shim", "meas.cxx"),
          "-L", HERE, "-lqxshim", "-lqx", "-L", h5lib, "-lhdf5", "-pthread",
          "-Wl,-rpath,$ORIGIN:" + h5lib, "-o", IOLIB + ".tmp"])
    os.replace(IOLIB + ".tmp", IOLIB)
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    mod = os.path.join(HERE, "h5test" + ext)
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", *inc,
          "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
          os.path.join(CSRC, "shim", "h5test.cxx"),
          "-L", HERE, "-lqxio", "-lqxshim", "-lqx", "-Wl,-rpath,$ORIGIN", "-o", mod + ".tmp"])
    os.replace(mod + ".tmp", mod)
    # the reference's program over the device Field, a C++ caller of the
    # libraries (links the C++ runtime ahead of libqxio)
    _run(["g++", "-O2", "-std=c++17", *inc, os.path.join(CSRC, "shim", "qx_run.cxx"),
          "-L", HERE, "-lqxio", "-lqxshim", "-lqx", "-Wl,-rpath,$ORIGIN", "-o", RUN + ".tmp"])
    os.replace(RUN + ".tmp", RUN)
    return IOLIB, mod


def build(verbose: bool = True):
    lib = build_lib()
    if verbose:
        print("built", lib)
    shim, mod = build_shim()
    if verbose:
        print("built", shim)
        print("built", mod)
    if hdf5_found():
        iolib, h5mod = build_io()
        if verbose:
            print("built", iolib)
            print("built", h5mod)
    elif verbose:
        print("HDF5 not found under", HDF5_DIR, "- libqxio.so (IO, measure) not built")


if __name__ == "__main__":
    build()
    sys.exit(0)
