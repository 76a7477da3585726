"""Input builders shared by the tests (test-side only)."""
from oracle import bn254 as o


def progression_points(n, a, b):
    """P_i = [b + a*i] G for i < n, by successive affine additions.

    Returns (points, discrete_logs).  The logs make every MSM over these
    points checkable in closed form: sum s_i P_i = [sum s_i k_i mod r] G.
    """
    A = o.g1_mul(a)
    P = o.g1_mul(b)
    pts, logs = [], []
    k = b % o.R
    for _ in range(n):
        pts.append(P)
        logs.append(k)
        P = o.add(P, A)
        k = (k + a) % o.R
    return pts, logs


def closed_form_msm(logs, scalars, mask_identity=()):
    """[sum s_i k_i] G, skipping indices listed as identity points."""
    skip = set(mask_identity)
    e = sum(s * k for i, (s, k) in enumerate(zip(scalars, logs)) if i not in skip) % o.R
    return o.g1_mul(e)
