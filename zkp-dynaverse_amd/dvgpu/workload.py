"""Seeded synthetic workloads of the benchmark configurations.

ChaCha20 seeds as used throughout: SRS trapdoor tau = uniform nonzero Fr, SRS
G1 points [tau^i] g1 (generated on the device by fixed-base setup), scalars
uniform Fr.  Everything here is host-side input preparation; the group and
transform work runs on the GPU.
"""
import numpy as np

from .codec import FR_MODULUS
from .rng import ChaCha20Stream


def tau_powers(tau: int, n: int) -> np.ndarray:
    """tau^0 .. tau^(n-1) as an (n, 32) canonical little-endian array."""
    out = bytearray(32 * n)
    p = 1
    r = FR_MODULUS
    for i in range(n):
        out[32 * i:32 * i + 32] = p.to_bytes(32, "little")
        p = p * tau % r
    return np.frombuffer(bytes(out), dtype=np.uint8).reshape(n, 32)


def srs_scalars(seed: int, n: int):
    """(tau, tau powers) for an SRS of n G1 points."""
    rng = ChaCha20Stream(seed)
    tau = rng.nonzero_fr()
    return tau, tau_powers(tau, n)


def poly_eval(coeff_bytes: np.ndarray, x: int) -> int:
    """sum_i c_i x^i mod r for an (n, 32) canonical coefficient array."""
    acc = 0
    r = FR_MODULUS
    arr = np.ascontiguousarray(coeff_bytes)
    for i in range(arr.shape[0] - 1, -1, -1):
        acc = (acc * x + int.from_bytes(arr[i].tobytes(), "little")) % r
    return acc


def small_sequence(seed: int = 1, n: int = 256, updates: int = 8, k: int = 4, workload: str = "hotspot",
                   hot: int = 16):
    """The plain n = 256 update sequence: (tau, coefficients (n, 32), updates).

    tau uniform nonzero Fr, coefficients uniform Fr; each of `updates`
    updates rewrites k indices (repeats allowed; a repeat rewrites twice, the
    last value wins) drawn from a fixed seed-chosen set of `hot` indices
    ("hotspot") or uniformly over n ("spreading"), each with a fresh uniform
    Fr value.  An update is a list of (index, new value)."""
    if workload not in ("hotspot", "spreading"):
        raise ValueError("workload is hotspot or spreading")
    rng = ChaCha20Stream(seed)
    tau = rng.nonzero_fr()
    coeffs = rng.uniform_fr_array(n)
    hot_set = []
    while len(hot_set) < hot:
        i = rng.below(n)
        if i not in hot_set:
            hot_set.append(i)
    ups = []
    for _ in range(updates):
        idx = [hot_set[rng.below(hot)] if workload == "hotspot" else rng.below(n) for _ in range(k)]
        ups.append(list(zip(idx, rng.uniform_fr(k))))
    return tau, coeffs, ups
