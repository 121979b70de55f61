"""Keyed counter-based random numbers (SplitMix64 finaliser, Steele, Lea and
Flood 2014), bit-identical to the device generator used by the IC kernels:
u64(seed, stream, i) = mix(mix(seed ^ stream*C1) + (i+1)*C2)."""
from __future__ import annotations

M64 = (1 << 64) - 1


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def u64(seed, stream, idx):
    key = _mix64((seed ^ (stream * 0xD6E8FEB86659FD93)) & M64)
    return _mix64((key + (idx + 1) * 0x9E3779B97F4A7C15) & M64)


def uniform(seed, stream, idx):
    return (u64(seed, stream, idx) >> 11) * 2.0 ** -53
