"""The MSM's signed-digit recoding, restated on the host: digits in
[-(2^(c-1) - 1), 2^(c-1)] from one addition of K = sum_k (2^(c-1) - 1) 2^(kc)
equal the carry-propagating recoding (both represent s with the same digit
set, which makes the representation unique), for every window width the
kernels accept, including W c > 256 (nine-word sums)."""
import random

from dvgpu.codec import FR_MODULUS as R


def carry_recoding(s, c):
    W = -(-255 // c)
    B = 1 << (c - 1)
    out, carry = [], 0
    for w in range(W):
        v = ((s >> (w * c)) & ((1 << c) - 1)) + carry
        if v > B:
            out.append(v - (1 << c))
            carry = 1
        else:
            out.append(v)
            carry = 0
    assert carry == 0
    return out


def offset_recoding(s, c):
    W = -(-255 // c)
    B = 1 << (c - 1)
    K = sum((B - 1) << (k * c) for k in range(W))
    t = s + K
    assert t < 1 << (9 * 32)
    return [((t >> (w * c)) & ((1 << c) - 1)) - (B - 1) for w in range(W)]


def test_offset_recoding_equals_carry_recoding():
    rng = random.Random(11)
    specials = [0, 1, 2, R - 1, R - 2, (1 << 253), (1 << 254) - 1 if (1 << 254) - 1 < R else R - 3]
    for c in range(4, 21):
        W = -(-255 // c)
        for s in specials + [rng.randrange(R) for _ in range(300)]:
            a, b = carry_recoding(s, c), offset_recoding(s, c)
            assert a == b, (c, s)
            assert sum(d << (w * c) for w, d in enumerate(a)) == s
            assert all(-(1 << (c - 1)) < d <= (1 << (c - 1)) for d in a)
        assert W * c >= 255
