"""Generate bn254_constants.hpp: BN254 moduli and derived constants.

Field elements live in registers as nine 29-bit limbs (radix 2^29, R = 2^261
Montgomery form) and in memory as eight little-endian 32-bit words.  Every
constant below is derived from the two moduli and the curve definitions:
Fq  q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
Fr  r = 21888242871839275222246405745257275088548364400416034343698204186575808495617
E : y^2 = x^3 + 3 over Fq;  E': y^2 = x^3 + 3/(9+u) over Fq2 = Fq[u]/(u^2+1).
Run: python gen_constants.py > bn254_constants.hpp
"""
Q = 0x30644E72E131A029B85045B68181585D97816A916871CA8D3C208C16D87CFD47
R = 0x30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001
LB, NL = 29, 9
RM = 1 << (LB * NL)  # Montgomery radix 2^261
M = (1 << LB) - 1


def l29(v):
    assert 0 <= v < RM
    return [(v >> (LB * i)) & M for i in range(NL)]


def w32(v):
    assert 0 <= v < 1 << 256
    return [(v >> (32 * i)) & 0xFFFFFFFF for i in range(8)]


def borrow_form(k_times_mod):
    """Limbs of K*mod rearranged so limbs 0..7 are >= 2^29 - 1 (>= 2^31 - 4
    for the 'big' form), letting a + K - b run limb-wise without borrows."""
    base = l29(k_times_mod)
    out = list(base)
    out[0] += 1 << LB
    for i in range(1, 8):
        out[i] += (1 << LB) - 1
    out[8] -= 1
    assert sum(v << (LB * i) for i, v in enumerate(out)) == k_times_mod
    assert all(0 <= v < 1 << 32 for v in out)
    return out


def borrow_form_big(k_times_mod):
    base = l29(k_times_mod)
    out = list(base)
    out[0] += 4 << LB
    for i in range(1, 8):
        out[i] += (4 << LB) - 4
    out[8] -= 4
    assert sum(v << (LB * i) for i, v in enumerate(out)) == k_times_mod
    assert all(0 <= v < 1 << 32 for v in out)
    return out


def arr(name, vals, typ="uint32_t"):
    return "    static constexpr %s %s[%d] = {%s};\n" % (
        typ, name, len(vals), ", ".join("0x%08xu" % v for v in vals))


def field_struct(name, mod):
    inv = (-pow(mod, -1, 1 << LB)) % (1 << LB)
    top = mod >> (LB * 8)
    s = "struct %s {\n" % name
    s += arr("MOD", l29(mod))
    s += arr("MOD32", w32(mod))
    s += arr("NEG", l29(RM - mod))            # 2^261 - mod
    s += arr("ONE", l29(RM % mod))            # R mod m
    s += arr("R2", l29(RM * RM % mod))        # R^2 mod m
    s += arr("SUB4", borrow_form(4 * mod))    # 4m, borrow-adjusted
    s += arr("SUB8", borrow_form_big(8 * mod))  # 8m, big borrow-adjusted
    s += "    static constexpr uint32_t INV = 0x%08xu;\n" % inv
    s += "    static constexpr uint32_t TOPDIV = %du;  // (m >> 232) + 1\n" % (top + 1)
    s += "};\n"
    return s


def mont(v, mod):
    return l29(v * RM % mod)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def f2_pow(a, e):
    out = (1, 0)
    while e:
        if e & 1:
            out = f2_mul(out, a)
        a = f2_mul(a, a)
        e >>= 1
    return out


def _g1_mul(k):
    """[k]G1 in affine coordinates (plain double-and-add; None = infinity)."""
    def add(p, q):
        if p is None:
            return q
        if q is None:
            return p
        if p[0] == q[0] and (p[1] + q[1]) % Q == 0:
            return None
        if p == q:
            m = 3 * p[0] * p[0] * pow(2 * p[1], Q - 2, Q) % Q
        else:
            m = (q[1] - p[1]) * pow(q[0] - p[0], Q - 2, Q) % Q
        x = (m * m - p[0] - q[0]) % Q
        return x, (m * (p[0] - x) - p[1]) % Q
    acc, base = None, (1, 2)
    while k:
        if k & 1:
            acc = add(acc, base)
        base = add(base, base)
        k >>= 1
    return acc


def glv_constants():
    """GLV endomorphism phi(x, y) = (beta x, y) = [lambda] on G1, with the short
    lattice basis of {(a, b) : a + b lambda = 0 mod r} from the extended
    Euclidean algorithm on (r, lambda) (Gallant-Lambert-Vanstone 2001, sec. 4)
    and the reciprocals that round b2 k / r and -b1 k / r for a split
    k = k1 + k2 lambda with |k1|, |k2| < 2^126."""
    beta = next(c for c in (pow(g, (Q - 1) // 3, Q) for g in range(2, 50)) if c != 1)
    lam = next(c for c in (pow(g, (R - 1) // 3, R) for g in range(2, 50)) if c != 1)
    g1 = _g1_mul(1)
    if _g1_mul(lam) != (beta * g1[0] % Q, g1[1]):
        lam = lam * lam % R
    assert _g1_mul(lam) == (beta * g1[0] % Q, g1[1])
    seq = [(R, 0), (lam, 1)]
    while seq[-1][0]:
        (ra, ta), (rb, tb) = seq[-2], seq[-1]
        seq.append((ra - ra // rb * rb, ta - ra // rb * tb))
    m = max(i for i, (r_, _) in enumerate(seq) if r_ * r_ >= R)
    a1, b1 = seq[m + 1][0], -seq[m + 1][1]
    a2, b2 = min(((seq[m][0], -seq[m][1]), (seq[m + 2][0], -seq[m + 2][1])), key=lambda v: v[0] ** 2 + v[1] ** 2)
    assert (a1 + b1 * lam) % R == 0 and (a2 + b2 * lam) % R == 0 and a1 * b2 - a2 * b1 == R
    assert b1 == -a2 and a1 > 0 and a2 > 0 and b2 > 0
    rc1, rc2 = (b2 << 384) // R, (a2 << 384) // R

    def words(name, v, n):
        assert 0 <= v < 1 << (32 * n)
        return "constexpr uint32_t %s[%d] = {%s};" % (
            name, n, ", ".join("0x%08xu" % ((v >> (32 * i)) & 0xFFFFFFFF) for i in range(n)))
    return ["// GLV: phi(x, y) = (GLV_BETA x, y) = [GLV_LAMBDA](x, y); basis (a1, -a2), (a2, b2).",
            "constexpr uint32_t GLV_BETA[%d] = {%s};" % (NL, ", ".join("0x%08xu" % x for x in mont(beta, Q))),
            words("GLV_LAMBDA", lam, 8), words("GLV_A1", a1, 4), words("GLV_A2", a2, 2),
            words("GLV_B2", b2, 4), words("GLV_RC1", rc1, 9), words("GLV_RC2", rc2, 7)]


def pairing_constants(xi):
    """Optimal ate pairing: loop count 6x + 2, Frobenius coefficients
    gamma_k = xi^(k (q - 1) / 6) of w^k (k = 1..5) for Fq12 = Fq2[w]/(w^6 - xi),
    and the hard part (q^4 - q^2 + 1) / r of the final exponent."""
    x = 4965661367192848881
    ate = 6 * x + 2
    hard = (Q ** 4 - Q ** 2 + 1) // R
    assert hard * R == Q ** 4 - Q ** 2 + 1
    out = ["// Optimal ate pairing: PAIRING_ATE = 6x + 2, FROB_Ck = xi^(k(q-1)/6), hard exponent words.",
           "constexpr uint64_t PAIRING_ATE_LO = 0x%016xull;" % (ate & (2 ** 64 - 1)),
           "constexpr uint32_t PAIRING_ATE_HI = %du;" % (ate >> 64),
           "constexpr int PAIRING_ATE_BITS = %d;" % ate.bit_length()]
    for k in range(1, 6):
        g = f2_pow(xi, k * (Q - 1) // 6)
        out.append("constexpr uint32_t FROB_C%d_0[%d] = {%s};" % (k, NL, ", ".join("0x%08xu" % v for v in mont(g[0], Q))))
        out.append("constexpr uint32_t FROB_C%d_1[%d] = {%s};" % (k, NL, ", ".join("0x%08xu" % v for v in mont(g[1], Q))))
    nw = (hard.bit_length() + 31) // 32
    out.append("constexpr int HARD_EXP_BITS = %d;" % hard.bit_length())
    out.append("constexpr uint32_t HARD_EXP[%d] = {%s};" % (
        nw, ", ".join("0x%08xu" % ((hard >> (32 * i)) & 0xFFFFFFFF) for i in range(nw))))
    return out


def main():
    xi = (9, 1)
    n = (81 + 1) % Q
    ni = pow(n, Q - 2, Q)
    xi_inv = (9 * ni % Q, (-ni) % Q)
    b2 = f2_mul((3, 0), xi_inv)
    psi_x = f2_pow(xi, (Q - 1) // 3)
    psi_y = f2_pow(xi, (Q - 1) // 2)
    g2 = ((10857046999023057135944570762232829481370756359578518086990519993285655852781,
           11559732032986387107991004021392285783925812861821192530917403151452391805634),
          (8495653923123431417604973247489272438418190587263600148770280649306958101930,
           4082367875863433681332203403145435568316851327593401208105741076214120093531))
    out = ["// Generated by gen_constants.py; do not edit.", "#pragma once", "#include <stdint.h>",
           "namespace bn254 {",
           "constexpr int LIMB_BITS = %d;" % LB, "constexpr int LIMBS = %d;" % NL,
           "constexpr uint32_t LIMB_MASK = 0x%08xu;" % M,
           field_struct("FqParams", Q), field_struct("FrParams", R)]
    consts = {
        "FQ_B3": mont(3, Q),                   # curve b of E
        "FQ2_B_C0": mont(b2[0], Q), "FQ2_B_C1": mont(b2[1], Q),
        "PSI_X_C0": mont(psi_x[0], Q), "PSI_X_C1": mont(psi_x[1], Q),
        "PSI_Y_C0": mont(psi_y[0], Q), "PSI_Y_C1": mont(psi_y[1], Q),
        "G1_X": mont(1, Q), "G1_Y": mont(2, Q),
        "G2_X_C0": mont(g2[0][0], Q), "G2_X_C1": mont(g2[0][1], Q),
        "G2_Y_C0": mont(g2[1][0], Q), "G2_Y_C1": mont(g2[1][1], Q),
    }
    for k, v in consts.items():
        out.append("constexpr uint32_t %s[%d] = {%s};" % (k, NL, ", ".join("0x%08xu" % x for x in v)))
    out.append("constexpr uint32_t FR_MOD32[8] = {%s};" % ", ".join("0x%08xu" % x for x in w32(R)))
    out.append("constexpr uint32_t FQ_MOD32[8] = {%s};" % ", ".join("0x%08xu" % x for x in w32(Q)))
    out.append("constexpr uint64_t BN_SEED = 4965661367192848881ull;")
    lam = 6 * 4965661367192848881 ** 2
    out.append("constexpr uint32_t G2_LAMBDA[4] = {%s};" % ", ".join(
        "0x%08xu" % ((lam >> (32 * i)) & 0xFFFFFFFF) for i in range(4)))
    out.extend(glv_constants())
    out.extend(pairing_constants(xi))
    out.append("}  // namespace bn254")
    print("\n".join(out))


if __name__ == "__main__":
    main()
