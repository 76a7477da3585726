"""Seeded G2 wire corpus (test-side): subgroup points, non-subgroup coset
points, infinity forms, encoding errors and off-curve points, each with the
mask the reference's validator assigns (g2_validation/tests.rs:32-44,79-103)."""
from oracle import bn254 as o

Q = o.Q


def twist_cofactor_point():
    """A point of E'(Fq2) with a nonzero component outside the r-torsion:
    T = [r] T0 for a non-subgroup twist point T0, so T has cofactor order."""
    t0 = o.twist_point_from_x(1)
    t = o.mul(t0, o.R, o.FQ2)
    assert t is not None
    return t


def corpus(scalars, coset_every=2):
    """Points [s_i] g2; every `coset_every`-th (odd i) gets T added, which
    leaves the subgroup.  Returns (wire bytes, expected masks)."""
    T = twist_cofactor_point()
    wire, masks = [], []
    for i, s in enumerate(scalars):
        p = o.g2_mul(s)
        if coset_every and i % coset_every == 1:
            p = o.add(p, T, o.FQ2)
            masks.append(4)
        else:
            masks.append(0)
        wire.append(o.g2_to_bytes(p))
    return b"".join(wire), masks


def encoding_cases():
    """(wire, mask) pairs for the parser: infinity forms, flags, q and q+1
    coordinates, bit flips."""
    g = o.G2_GEN
    good = bytearray(o.g2_to_bytes(g))
    cases = [(bytes(128), 0)]
    inf_flag = bytearray(128)
    inf_flag[127] = 0x40
    cases.append((bytes(inf_flag), 0))                       # flagged infinity
    both = bytearray(good)
    both[127] |= 0xC0
    cases.append((bytes(both), 1))                           # both flags
    for j in range(4):                                       # coordinate = q
        bad = bytearray(good)
        bad[32 * j:32 * j + 32] = Q.to_bytes(32, "little")
        cases.append((bytes(bad), 1))
    for j in range(4):                                       # coordinate = q + 1
        bad = bytearray(good)
        bad[32 * j:32 * j + 32] = (Q + 1).to_bytes(32, "little")
        cases.append((bytes(bad), 1))
    flagged_noncanon = bytearray(128)
    flagged_noncanon[0:32] = Q.to_bytes(32, "little")
    flagged_noncanon[127] = 0x40
    cases.append((bytes(flagged_noncanon), 1))               # infinity flag, bad coords
    for byte in (0, 33, 64, 100):                            # bit flips -> off curve
        bad = bytearray(good)
        bad[byte] ^= 0x01
        cases.append((bytes(bad), 2))
    cases.append((bytes(good), 0))
    return cases
