"""Seeded synthetic workload generator: ChaCha20 (RFC 8439) keystream.

Key = the u64 seed little-endian, zero padded to 32 bytes; nonce = 0; block
counter from 0.  Uniform Fr: successive 32-byte keystream chunks, masked to
254 bits, rejected when >= r (BN254 scalar field).  Vectorised with numpy so
2^20 scalars take about a second.
"""
import numpy as np

from .codec import FR_MODULUS

_CONST = np.array([0x61707865, 0x3320646E, 0x79622D32, 0x6B206574], dtype=np.uint32)


def _rotl(x, n):
    return (x << np.uint32(n)) | (x >> np.uint32(32 - n))


def chacha20_blocks(seed: int, first_block: int, count: int) -> np.ndarray:
    """Keystream blocks [first_block, first_block + count) as (count, 64) uint8."""
    key = np.frombuffer(int(seed).to_bytes(8, "little") + bytes(24), dtype="<u4").astype(np.uint32)
    ctr = np.arange(first_block, first_block + count, dtype=np.uint64).astype(np.uint32)
    state = [np.full(count, c, dtype=np.uint32) for c in _CONST]
    state += [np.full(count, k, dtype=np.uint32) for k in key]
    state += [ctr, np.zeros(count, np.uint32), np.zeros(count, np.uint32), np.zeros(count, np.uint32)]
    x = [s.copy() for s in state]

    def qr(a, b, c, d):
        x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 16)
        x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 12)
        x[a] += x[b]; x[d] ^= x[a]; x[d] = _rotl(x[d], 8)
        x[c] += x[d]; x[b] ^= x[c]; x[b] = _rotl(x[b], 7)

    with np.errstate(over="ignore"):
        for _ in range(10):
            qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
            qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
        out = np.stack([x[i] + state[i] for i in range(16)], axis=1)
    return out.astype("<u4").view(np.uint8).reshape(count, 64)


class ChaCha20Stream:
    """Sequential keystream reader."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.block = 0
        self.buf = b""

    def read(self, nbytes: int) -> bytes:
        while len(self.buf) < nbytes:
            need = max(1, (nbytes - len(self.buf) + 63) // 64)
            blocks = chacha20_blocks(self.seed, self.block, need)
            self.block += need
            self.buf += blocks.tobytes()
        out, self.buf = self.buf[:nbytes], self.buf[nbytes:]
        return out

    def uniform_fr_array(self, n: int) -> np.ndarray:
        """n uniform Fr scalars as an (n, 32) uint8 canonical LE array."""
        out = np.zeros((n, 32), dtype=np.uint8)
        have = 0
        r_words = np.array([(FR_MODULUS >> (64 * i)) & (2**64 - 1) for i in range(4)], dtype=np.uint64)
        while have < n:
            want = int((n - have) * 1.34) + 16
            self._last = self.read(32 * want)
            cand = np.frombuffer(self._last, dtype=np.uint8).reshape(want, 32).copy()
            cand[:, 31] &= 0x3F
            words = cand.view("<u8").reshape(want, 4)
            lt = np.zeros(want, dtype=bool)
            eq = np.ones(want, dtype=bool)
            for i in (3, 2, 1, 0):
                lt |= eq & (words[:, i] < r_words[i])
                eq &= words[:, i] == r_words[i]
            idx = np.nonzero(lt)[0]
            take = min(n - have, idx.shape[0])
            out[have:have + take] = cand[idx[:take]]
            have += take
            if take:
                # Return keystream past the last accepted candidate, so draws
                # are exactly sequential whatever the batch size.
                used = int(idx[take - 1]) + 1
                self.buf = self._last[32 * used:] + self.buf
        return out

    def uniform_fr(self, n: int):
        arr = self.uniform_fr_array(n)
        return [int.from_bytes(arr[i].tobytes(), "little") for i in range(n)]

    def nonzero_fr(self) -> int:
        while True:
            v = self.uniform_fr(1)[0]
            if v:
                return v

    def below(self, bound: int) -> int:
        """Uniform integer in [0, bound) by rejection on 64-bit draws."""
        limit = (1 << 64) - ((1 << 64) % bound)
        while True:
            v = int.from_bytes(self.read(8), "little")
            if v < limit:
                return v % bound
