"""Writes tests/golden/golden_field.npz: the expected values of the reference's
own known-answer tests (tests/test_field.py:12-129, 56 cases) plus whole
arrays for a few of them, computed with the numpy restatement in
tests/refmath.py.  The reference itself is C++ over FFTW and is not built
here; its Python tests are restated from their text, not imported.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import refmath as rm  # noqa: E402


def main():
    out = {}
    for n, N, dt, idx in rm.PARAMS:
        psi = rm.solitons(n, N)
        out[f"kick_{n}_{N}_{dt}_{idx}"] = rm.kick(psi, dt).flat[idx].real
        out[f"drift_{n}_{N}_{dt}_{idx}"] = rm.drift(psi, dt).flat[idx].real
        out[f"kds_{n}_{N}_{dt}_{idx}"] = rm.kick_drift_step(psi, dt).flat[idx].real
    for n, N, idx in rm.PARAMS_POT:
        psi = rm.solitons(n, N)
        out[f"potential_{n}_{N}_{idx}"] = rm.potential(psi).flat[idx].real
    for n, N in [(2, 8), (3, 8), (2, 16)]:
        psi = rm.solitons(n, N)
        out[f"arr_psi0_{n}_{N}"] = psi
        out[f"arr_potential_{n}_{N}"] = rm.potential(psi)
        out[f"arr_kick_{n}_{N}"] = rm.kick(psi, 0.1)
        out[f"arr_drift_{n}_{N}"] = rm.drift(psi, 0.1)
        out[f"arr_kds_{n}_{N}"] = rm.kick_drift_step(psi, 0.1)
    np.savez_compressed(os.path.join(HERE, "golden_field.npz"), **out)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
