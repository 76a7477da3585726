"""CPU oracle for the BN254 proving hot path (Python big-int restatement).

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker.

It restates, from the textbook definitions, what the reference computes with
arkworks 0.4 (ark-bn254 0.4.0, ark-ec/ff/poly 0.4.2, a crates.io dependency
that is absent from the reference checkout):

* Fq, Fr, Fq2 arithmetic and the BN254 curves E: y^2 = x^3 + 3 over Fq and
  E': y^2 = x^3 + 3/(9+u) over Fq2 (``crates/gpu/src/codec.rs:1-18``).
* ``straight_msm``: left-to-right double-and-add MSM, the same algorithm as
  ``crates/gpu/src/lib.rs:120-139`` (``CpuBackend::msm``).
* ``ntt``: natural-order radix-2 DFT over <omega_n>, omega_n derived from the
  2^20-th root the wrapper passes (``crates/gpu/src/icicle.rs:429-456``,
  ``Radix2EvaluationDomain`` in ``crates/gpu/src/lib.rs:99-117``); inverse
  carries the 1/n factor.
* ``fixed_base``: [s]G followed by normalisation
  (``crates/protocol/src/setup.rs:13-31``).
* G2 wire parsing and subgroup predicates with masks 1/2/4
  (``crates/gpu/src/g2_validation.rs:214-299``,
  ``native/include/icicle_g2_wire.cuh:9-35``,
  ``native/src/icicle_g2_validation.cu:20-116``).

Every function here is pinned by a check in ``tests/test_oracle.py``
(closed forms from the reference's own tests, brute-force DFT, group-order
identities, cross-check against the C oracle).  Functions without such a
check would say "parity unpinned" here; none are left in that state.
"""

from __future__ import annotations

# --------------------------------------------------------------------------
# Moduli and curve constants
# --------------------------------------------------------------------------
Q = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47  # Fq
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001  # Fr
B1 = 3
G1_GEN = (1, 2)
# EIP-197 / ark-bn254 G2 generator, (c0, c1) pairs.
G2_GEN = (
    (10857046999023057135944570762232829481370756359578518086990519993285655852781,
     11559732032986387107991004021392285783925812861821192530917403151452391805634),
    (8495653923123431417604973247489272438418190587263600148770280649306958101930,
     4082367875863433681332203403145435568316851327593401208105741076214120093531),
)
BN_SEED = 4965661367192848881  # BN parameter x; 63 bits, popcount 28
LAMBDA = 6 * BN_SEED * BN_SEED  # psi acts as [6x^2] on the r-torsion of E'
FR_GENERATOR = 5  # multiplicative generator arkworks derives roots from
TWO_ADICITY = 28


def fr_root_of_unity(n: int) -> int:
    """omega_n = 5^((r-1)/n) for a power of two n <= 2^28."""
    assert n >= 1 and n & (n - 1) == 0 and n <= (1 << TWO_ADICITY)
    return pow(FR_GENERATOR, (R - 1) // n, R)


ROOT_2_20 = fr_root_of_unity(1 << 20)

# --------------------------------------------------------------------------
# Fq2 = Fq[u]/(u^2+1)
# --------------------------------------------------------------------------


def f2_add(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def f2_sub(a, b):
    return ((a[0] - b[0]) % Q, (a[1] - b[1]) % Q)


def f2_neg(a):
    return ((-a[0]) % Q, (-a[1]) % Q)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def f2_inv(a):
    n = (a[0] * a[0] + a[1] * a[1]) % Q
    ni = pow(n, Q - 2, Q)
    return ((a[0] * ni) % Q, (-a[1] * ni) % Q)


def f2_pow(a, e):
    out = (1, 0)
    while e:
        if e & 1:
            out = f2_mul(out, a)
        a = f2_mul(a, a)
        e >>= 1
    return out


def f2_conj(a):
    return (a[0], (-a[1]) % Q)


XI = (9, 1)
B2 = f2_mul((3, 0), f2_inv(XI))  # twist coefficient 3/(9+u)

# --------------------------------------------------------------------------
# Field abstraction so the curve code serves both G1 (Fq) and G2 (Fq2)
# --------------------------------------------------------------------------


class _Fq:
    zero = 0
    one = 1
    b = B1

    @staticmethod
    def add(a, b):
        return (a + b) % Q

    @staticmethod
    def sub(a, b):
        return (a - b) % Q

    @staticmethod
    def mul(a, b):
        return (a * b) % Q

    @staticmethod
    def neg(a):
        return (-a) % Q

    @staticmethod
    def inv(a):
        return pow(a, Q - 2, Q)

    @staticmethod
    def is_zero(a):
        return a == 0


class _Fq2:
    zero = (0, 0)
    one = (1, 0)
    b = B2
    add = staticmethod(f2_add)
    sub = staticmethod(f2_sub)
    mul = staticmethod(f2_mul)
    neg = staticmethod(f2_neg)
    inv = staticmethod(f2_inv)

    @staticmethod
    def is_zero(a):
        return a == (0, 0)


FQ = _Fq
FQ2 = _Fq2

# --------------------------------------------------------------------------
# Affine group law.  None is the point at infinity.  Affine arithmetic with an
# inversion per step is slow but obviously correct; it is the oracle.
# --------------------------------------------------------------------------


def on_curve(p, F=FQ):
    if p is None:
        return True
    x, y = p
    return F.sub(F.mul(y, y), F.add(F.mul(F.mul(x, x), x), F.b)) == F.zero


def neg(p, F=FQ):
    if p is None:
        return None
    return (p[0], F.neg(p[1]))


def add(p, q, F=FQ):
    if p is None:
        return q
    if q is None:
        return p
    x1, y1 = p
    x2, y2 = q
    if x1 == x2:
        if F.add(y1, y2) == F.zero:
            return None
        return double(p, F)
    lam = F.mul(F.sub(y2, y1), F.inv(F.sub(x2, x1)))
    x3 = F.sub(F.sub(F.mul(lam, lam), x1), x2)
    y3 = F.sub(F.mul(lam, F.sub(x1, x3)), y1)
    return (x3, y3)


def double(p, F=FQ):
    if p is None:
        return None
    x, y = p
    if F.is_zero(y):
        return None
    three_x2 = F.mul(F.add(F.add(x, x), x), x)
    lam = F.mul(three_x2, F.inv(F.add(y, y)))
    x3 = F.sub(F.mul(lam, lam), F.add(x, x))
    y3 = F.sub(F.mul(lam, F.sub(x, x3)), y)
    return (x3, y3)


# Jacobian arithmetic for the scalar-multiplication inner loops (same group
# law, fewer inversions); converted back to affine at the end.

def _jac_double(P, F):
    X, Y, Z = P
    if F.is_zero(Z) or F.is_zero(Y):
        return (F.one, F.one, F.zero)
    A = F.mul(X, X)
    Bv = F.mul(Y, Y)
    C = F.mul(Bv, Bv)
    t = F.add(X, Bv)
    D = F.sub(F.sub(F.mul(t, t), A), C)
    D = F.add(D, D)
    E = F.add(F.add(A, A), A)
    Fv = F.mul(E, E)
    X3 = F.sub(Fv, F.add(D, D))
    C8 = F.add(C, C)
    C8 = F.add(C8, C8)
    C8 = F.add(C8, C8)
    Y3 = F.sub(F.mul(E, F.sub(D, X3)), C8)
    Z3 = F.mul(F.add(Y, Y), Z)
    return (X3, Y3, Z3)


def _jac_add(P, Qp, F):
    X1, Y1, Z1 = P
    X2, Y2, Z2 = Qp
    if F.is_zero(Z1):
        return Qp
    if F.is_zero(Z2):
        return P
    Z1Z1 = F.mul(Z1, Z1)
    Z2Z2 = F.mul(Z2, Z2)
    U1 = F.mul(X1, Z2Z2)
    U2 = F.mul(X2, Z1Z1)
    S1 = F.mul(F.mul(Y1, Z2), Z2Z2)
    S2 = F.mul(F.mul(Y2, Z1), Z1Z1)
    if U1 == U2:
        if S1 == S2:
            return _jac_double(P, F)
        return (F.one, F.one, F.zero)
    H = F.sub(U2, U1)
    Rr = F.sub(S2, S1)
    H2 = F.mul(H, H)
    H3 = F.mul(H2, H)
    U1H2 = F.mul(U1, H2)
    X3 = F.sub(F.sub(F.mul(Rr, Rr), H3), F.add(U1H2, U1H2))
    Y3 = F.sub(F.mul(Rr, F.sub(U1H2, X3)), F.mul(S1, H3))
    Z3 = F.mul(F.mul(Z1, Z2), H)
    return (X3, Y3, Z3)


def _to_jac(p, F):
    if p is None:
        return (F.one, F.one, F.zero)
    return (p[0], p[1], F.one)


def _from_jac(P, F):
    X, Y, Z = P
    if F.is_zero(Z):
        return None
    zi = F.inv(Z)
    zi2 = F.mul(zi, zi)
    return (F.mul(X, zi2), F.mul(Y, F.mul(zi2, zi)))


def mul(p, k: int, F=FQ):
    """[k]p for an unreduced non-negative integer k (double-and-add)."""
    assert k >= 0
    acc = _to_jac(None, F)
    base = _to_jac(p, F)
    while k:
        if k & 1:
            acc = _jac_add(acc, base, F)
        base = _jac_double(base, F)
        k >>= 1
    return _from_jac(acc, F)


def straight_msm(points, scalars, F=FQ):
    """sum_i s_i * P_i by per-term double-and-add (``lib.rs:122-139``)."""
    assert len(points) == len(scalars)
    acc = _to_jac(None, F)
    for p, s in zip(points, scalars):
        assert 0 <= s < R
        cur = _to_jac(p, F)
        k = s
        while k:
            if k & 1:
                acc = _jac_add(acc, cur, F)
            cur = _jac_double(cur, F)
            k >>= 1
    return _from_jac(acc, F)


def g1_mul(k):
    return mul(G1_GEN, k % R, FQ)


def g2_mul(k):
    return mul(G2_GEN, k % R, FQ2)


# --------------------------------------------------------------------------
# Wire codec (crates/gpu/src/codec.rs:32-77; resident.rs encode_g2)
# --------------------------------------------------------------------------


def fr_to_bytes(v: int) -> bytes:
    return int(v).to_bytes(32, "little")


def bytes_to_int(b: bytes) -> int:
    return int.from_bytes(bytes(b), "little")


def g1_to_bytes(p) -> bytes:
    if p is None:
        return bytes(64)
    return p[0].to_bytes(32, "little") + p[1].to_bytes(32, "little")


def g1_from_bytes(b: bytes):
    b = bytes(b)
    assert len(b) == 64
    if b == bytes(64):
        return None
    x = int.from_bytes(b[:32], "little")
    y = int.from_bytes(b[32:], "little")
    assert x < Q and y < Q and on_curve((x, y)), "invalid G1 bytes"
    return (x, y)


def g2_to_bytes(p) -> bytes:
    if p is None:
        return bytes(128)
    (x0, x1), (y0, y1) = p
    return b"".join(v.to_bytes(32, "little") for v in (x0, x1, y0, y1))


def g2_from_bytes(b: bytes):
    b = bytes(b)
    assert len(b) == 128
    if b == bytes(128):
        return None
    v = [int.from_bytes(b[32 * i:32 * i + 32], "little") for i in range(4)]
    return ((v[0], v[1]), (v[2], v[3]))


# --------------------------------------------------------------------------
# NTT (natural order in and out; inverse scaled by 1/n)
# --------------------------------------------------------------------------


def dft_bruteforce(values, inverse=False):
    """O(n^2) evaluation sum_j v_j w^(jk): the definition itself."""
    n = len(values)
    w = fr_root_of_unity(n)
    if inverse:
        w = pow(w, R - 2, R)
    out = []
    for k in range(n):
        wk = pow(w, k, R)
        acc = 0
        t = 1
        for v in values:
            acc = (acc + v * t) % R
            t = (t * wk) % R
        out.append(acc)
    if inverse:
        ninv = pow(n, R - 2, R)
        out = [(o * ninv) % R for o in out]
    return out


def ntt(values, inverse=False):
    """Iterative radix-2 DIT transform, equal to ``dft_bruteforce``."""
    n = len(values)
    assert n >= 1 and n & (n - 1) == 0
    a = [v % R for v in values]
    if n == 1:
        return a
    logn = n.bit_length() - 1
    rev = [int(bin(i)[2:].zfill(logn)[::-1], 2) for i in range(n)]
    a = [a[rev[i]] for i in range(n)]
    w_n = fr_root_of_unity(n)
    if inverse:
        w_n = pow(w_n, R - 2, R)
    h = 1
    while h < n:
        w = pow(w_n, n // (2 * h), R)
        tw = [1] * h
        for j in range(1, h):
            tw[j] = (tw[j - 1] * w) % R
        for start in range(0, n, 2 * h):
            for j in range(h):
                u = a[start + j]
                v = (a[start + j + h] * tw[j]) % R
                a[start + j] = (u + v) % R
                a[start + j + h] = (u - v) % R
        h *= 2
    if inverse:
        ninv = pow(n, R - 2, R)
        a = [(x * ninv) % R for x in a]
    return a


# --------------------------------------------------------------------------
# Fixed-base setup (crates/protocol/src/setup.rs:13-31)
# --------------------------------------------------------------------------


def fixed_base_g1(scalars):
    return [g1_mul(s) for s in scalars]


def fixed_base_g2(scalars):
    return [g2_mul(s) for s in scalars]


# --------------------------------------------------------------------------
# G2 wire validation (g2_validation.rs:214-299, icicle_g2_wire.cuh:9-35)
# --------------------------------------------------------------------------
MASK_ENCODING, MASK_CURVE, MASK_SUBGROUP = 1, 2, 4


def g2_parse_wire(b: bytes):
    """Return (status, point).  status 1 = encoding error, else 0.

    All-zero is infinity.  Byte 127 carries arkworks' two flag bits: 0xC0 is
    invalid, 0x40 marks infinity (coordinates must still be canonical), and
    the flag bits are masked off the top limb before the canonical check.
    """
    b = bytes(b)
    assert len(b) == 128
    if b == bytes(128):
        return 0, None
    flags = b[127] & 0xC0
    if flags == 0xC0:
        return 1, None
    raw = bytearray(b)
    raw[127] &= 0x3F
    v = [int.from_bytes(raw[32 * i:32 * i + 32], "little") for i in range(4)]
    if any(x >= Q for x in v):
        return 1, None
    if flags == 0x40:
        return 0, None
    return 0, ((v[0], v[1]), (v[2], v[3]))


_PSI_CX = f2_pow(XI, (Q - 1) // 3)
_PSI_CY = f2_pow(XI, (Q - 1) // 2)


def psi(p):
    """Untwist-Frobenius-twist endomorphism on E'(Fq2)."""
    if p is None:
        return None
    x, y = p
    return (f2_mul(f2_conj(x), _PSI_CX), f2_mul(f2_conj(y), _PSI_CY))


def subgroup_full_order(p):
    return mul(p, R, FQ2) is None


def subgroup_frobenius_127(p):
    return psi(p) == mul(p, LAMBDA, FQ2)


def subgroup_frobenius_63(p):
    a = mul(p, BN_SEED, FQ2)
    b = psi(a)
    c = psi(b)
    lhs = add(add(add(a, p, FQ2), b, FQ2), c, FQ2)
    return lhs == double(psi(c), FQ2)


SUBGROUP_ALGORITHMS = {1: subgroup_full_order, 2: subgroup_frobenius_127, 3: subgroup_frobenius_63}


def g2_masks(wire: bytes, algorithm: int):
    """Per-point masks and the 32-byte report fields for a wire array."""
    assert len(wire) % 128 == 0
    pred = SUBGROUP_ALGORITHMS[algorithm]
    masks = []
    for i in range(len(wire) // 128):
        status, p = g2_parse_wire(wire[128 * i:128 * i + 128])
        if status:
            masks.append(MASK_ENCODING)
        elif not on_curve(p, FQ2):
            masks.append(MASK_CURVE)
        elif p is not None and not pred(p):
            masks.append(MASK_SUBGROUP)
        else:
            masks.append(0)
    invalid = [i for i, m in enumerate(masks) if m]
    agg = 0
    for m in masks:
        agg |= m
    report = dict(checked_points=len(masks), invalid_points=len(invalid),
                  first_invalid=invalid[0] if invalid else (1 << 64) - 1,
                  failure_mask=agg, algorithm=algorithm)
    return masks, report


def twist_point_from_x(x0: int):
    """Smallest twist point (x0 + k, 0 i + ...) found by incrementing x.c0.

    Twist points of E'(Fq2) outside the r-torsion exist because the twist
    has cofactor > 1; these feed the non-subgroup corpus.
    """
    k = x0
    while True:
        x = (k % Q, 1)
        rhs = f2_add(f2_mul(f2_mul(x, x), x), B2)
        y = f2_sqrt(rhs)
        if y is not None:
            return (x, y)
        k += 1


def _fq_sqrt(a):
    a %= Q
    if a == 0:
        return 0
    if pow(a, (Q - 1) // 2, Q) != 1:
        return None
    return pow(a, (Q + 1) // 4, Q)  # Q = 3 mod 4


def f2_sqrt(a):
    """Square root in Fq2 (complex method, Q = 3 mod 4)."""
    a0, a1 = a
    if a1 == 0:
        s = _fq_sqrt(a0)
        if s is not None:
            return (s, 0)
        s = _fq_sqrt((-a0) % Q)
        return (0, s) if s is not None else None
    norm = (a0 * a0 + a1 * a1) % Q
    alpha = _fq_sqrt(norm)
    if alpha is None:
        return None
    inv2 = pow(2, Q - 2, Q)
    for cand in ((a0 + alpha) * inv2 % Q, (a0 - alpha) * inv2 % Q):
        x0 = _fq_sqrt(cand)
        if x0 is not None and x0 != 0:
            x1 = a1 * pow(2 * x0, Q - 2, Q) % Q
            r = (x0, x1)
            if f2_mul(r, r) == (a0 % Q, a1 % Q):
                return r
    return None


# --------------------------------------------------------------------------
# Group NTT over G1 (resident ECNTT rows; operators.rs:413-440)
# --------------------------------------------------------------------------


def ecntt(points, inverse=False):
    """out_k = sum_j [w^(jk)] P_j over <w_n>, inverse scaled by n^-1: the DFT
    definition evaluated in the group, O(n^2) scalar multiplications."""
    n = len(points)
    w = fr_root_of_unity(n)
    if inverse:
        w = pow(w, R - 2, R)
    ninv = pow(n, R - 2, R) if inverse else 1
    out = []
    for k in range(n):
        acc = None
        for j, p in enumerate(points):
            acc = add(acc, mul(p, pow(w, j * k, R) * ninv % R))
        out.append(acc)
    return out


def ecntt_fft(points, inverse=False):
    """The same transform as ``ecntt`` by the iterative radix-2 DIT form of
    ``ntt`` with group additions and twiddle scalar multiplications."""
    n = len(points)
    assert n >= 1 and n & (n - 1) == 0
    a = list(points)
    if n == 1:
        return a
    logn = n.bit_length() - 1
    a = [a[int(bin(i)[2:].zfill(logn)[::-1], 2)] for i in range(n)]
    w_n = fr_root_of_unity(n)
    if inverse:
        w_n = pow(w_n, R - 2, R)
    h = 1
    while h < n:
        w = pow(w_n, n // (2 * h), R)
        for start in range(0, n, 2 * h):
            t = 1
            for j in range(h):
                u, v = a[start + j], mul(a[start + j + h], t)
                a[start + j], a[start + j + h] = add(u, v), add(u, neg(v))
                t = t * w % R
        h *= 2
    if inverse:
        ninv = pow(n, R - 2, R)
        a = [mul(p, ninv) for p in a]
    return a


# --------------------------------------------------------------------------
# KZG opening cache (crates/protocol/src/kzg.rs:103-127, 217-275)
# --------------------------------------------------------------------------


def kzg_srs(tau: int, max_degree: int):
    """KzgParams::from_tau's G1 powers [tau^i] G1, i <= max_degree (kzg.rs:107-127)."""
    return [g1_mul(pow(tau, i, R)) for i in range(max_degree + 1)]


def kzg_opening_closed_form(tau: int, coefficients, z: int):
    """The single-point opening witness [q(tau)] G1 of q = (f - f(z)) / (X - z)
    (KzgParams::open_single, kzg.rs:202-210) from the trapdoor, tau != z."""
    def ev(x):
        acc = 0
        for c in reversed(coefficients):
            acc = (acc * x + c) % R
        return acc
    return g1_mul((ev(tau) - ev(z)) * pow((tau - z) % R, R - 2, R) % R)


def kzg_openings(powers, coefficients, n: int, offset=None):
    """KzgParams::precompute_single_openings (kzg.rs:217-275): openings at
    offset * w_n^i (offset None: the subgroup itself), i < n, in order."""
    if len(coefficients) > n or len(coefficients) > len(powers):
        raise ValueError("polynomial exceeds the domain or the SRS")
    degree = max(len(coefficients) - 1, 0)
    if degree == 0:
        return [None] * n
    m = 1
    while m < 2 * degree - 1:
        m *= 2
    reversed_c = list(reversed(coefficients[1:])) + [0] * (m - degree)
    scalar_fft = ntt(reversed_c, False)
    prepared = ecntt_fft(list(powers[:degree]) + [None] * (m - degree), False)
    conv = ecntt_fft([mul(p, s) for p, s in zip(prepared, scalar_fft)], True)
    h = list(reversed(conv[:degree])) + [None] * (n - degree)
    if offset is not None:
        h = [mul(p, pow(offset, j, R)) for j, p in enumerate(h)]
    return ecntt_fft(h, False)


# --------------------------------------------------------------------------
# Optimal ate pairing (what arkworks' Bn254::multi_pairing computes; the
# reference's cache checks, crates/protocol/src/dynaverse/rebuilding/
# machine.rs:255-390, compare two such products)
# --------------------------------------------------------------------------
# Fq12 as Fq[w]/(w^12 - 18 w^6 + 82): w^6 = xi = 9 + u, so u = w^6 - 9 and
# u^2 = -1.  Elements are tuples of 12 coefficients (w^0 first).  This flat
# form is independent of the Fq2/Fq6/Fq12 tower the device uses.

F12_ONE = (1,) + (0,) * 11
ATE_LOOP = 6 * BN_SEED + 2


def f12_mul(a, b):
    t = [0] * 23
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                t[i + j] += x * y
    for k in range(22, 11, -1):       # w^k = 18 w^(k-6) - 82 w^(k-12)
        c = t[k]
        if c:
            t[k - 6] += 18 * c
            t[k - 12] -= 82 * c
    return tuple(v % Q for v in t[:12])


def f12_pow(a, e):
    out = F12_ONE
    while e:
        if e & 1:
            out = f12_mul(out, a)
        a = f12_mul(a, a)
        e >>= 1
    return out


def f12_from_f2(a, k):
    """The Fq2 element a times w^k (0 <= k < 6) in flat form."""
    out = [0] * 12
    out[k] = (a[0] - 9 * a[1]) % Q
    out[k + 6] = a[1] % Q
    return out


def _line(t, q, p):
    """Line through the twist points t, q (tangent when equal) evaluated at
    the untwisted P = (xp, yp): with x' = x w^2, y' = y w^3, slope' = m w,
    l = yp - (m xp) w + (m x_t - y_t) w^3."""
    (x1, y1), (x2, y2) = t, q
    if x1 != x2:
        m = f2_mul(f2_sub(y2, y1), f2_inv(f2_sub(x2, x1)))
    else:
        assert y1 == y2, "vertical line"
        m = f2_mul(f2_mul((3, 0), f2_mul(x1, x1)), f2_inv(f2_add(y1, y1)))
    xp, yp = p
    c1 = f2_neg(f2_mul(m, (xp, 0)))
    c3 = f2_sub(f2_mul(m, x1), y1)
    out = [0] * 12
    out[0] = yp
    for k, c in ((1, c1), (3, c3)):
        e = f12_from_f2(c, k)
        out = [(u + v) % Q for u, v in zip(out, e)]
    return tuple(out)


def miller_loop(p, q):
    """f_{6x+2,Q}(P) l_{T,pi(Q)}(P) l_{T',-pi^2(Q)}(P); 1 if either is infinity."""
    if p is None or q is None:
        return F12_ONE
    f, t = F12_ONE, q
    for i in range(ATE_LOOP.bit_length() - 2, -1, -1):
        f = f12_mul(f12_mul(f, f), _line(t, t, p))
        t = double(t, FQ2)
        if (ATE_LOOP >> i) & 1:
            f = f12_mul(f, _line(t, q, p))
            t = add(t, q, FQ2)
    q1 = psi(q)
    q2 = neg(psi(q1), FQ2)
    f = f12_mul(f, _line(t, q1, p))
    t = add(t, q1, FQ2)
    return f12_mul(f, _line(t, q2, p))


FINAL_EXP = (Q ** 12 - 1) // R


def final_exponentiation(f):
    return f12_pow(f, FINAL_EXP)


def pairing(p, q):
    return final_exponentiation(miller_loop(p, q))


def pairing_product(pairs):
    """prod e(P_i, Q_i) with one final exponentiation."""
    f = F12_ONE
    for p, q in pairs:
        f = f12_mul(f, miller_loop(p, q))
    return final_exponentiation(f)


# --------------------------------------------------------------------------
# Polynomial tools (crates/protocol/src/domain/fast.rs, domain.rs:303-337):
# the quadratic textbook forms the reference's own tests use as oracles
# (fast_tests.rs:5-22): products, long division, Horner, Lagrange.
# --------------------------------------------------------------------------


def poly_trim(p):
    p = [c % R for c in p] or [0]
    while len(p) > 1 and p[-1] == 0:
        p.pop()
    return p


def poly_mul(a, b):
    """multiply_polynomials (domain.rs:326-337): [0] for an empty operand."""
    if not a or not b:
        return [0]
    out = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] = (out[i + j] + x * y) % R
    return poly_trim(out)


def poly_divide_exact(num, den):
    """Long division; ValueError for a zero divisor or a nonzero remainder."""
    num, den = poly_trim(num), poly_trim(den)
    if den == [0]:
        raise ValueError("cannot divide by zero polynomial")
    if len(num) < len(den):
        if any(num):
            raise ValueError("nonzero remainder")
        return [0]
    rem = list(num)
    q = [0] * (len(num) - len(den) + 1)
    li = pow(den[-1], R - 2, R)
    for k in range(len(q) - 1, -1, -1):
        c = rem[k + len(den) - 1] * li % R
        q[k] = c
        for j, d in enumerate(den):
            rem[k + j] = (rem[k + j] - c * d) % R
    if any(rem):
        raise ValueError("nonzero remainder")
    return poly_trim(q)


def poly_eval(p, x):
    acc = 0
    for c in reversed(p):
        acc = (acc * x + c) % R
    return acc


def lagrange_interpolate(points, values):
    """naive_interpolate (fast_tests.rs:5-22)."""
    res = [0] * max(len(points), 1)
    for i, xi in enumerate(points):
        num, den = [1], 1
        for j, xj in enumerate(points):
            if i != j:
                num = poly_mul(num, [(-xj) % R, 1])
                den = den * (xi - xj) % R
        w = values[i] * pow(den, R - 2, R) % R
        for k, c in enumerate(num):
            res[k] = (res[k] + w * c) % R
    return poly_trim(res)


def cauchy_sums_direct(points, weights):
    """sum_{j != i} w_j / (x_i - x_j) (fast_tests.rs:45-54)."""
    return [sum(weights[j] * pow((xi - xj) % R, R - 2, R) for j, xj in enumerate(points) if j != i) % R
            for i, xi in enumerate(points)]
