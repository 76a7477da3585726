"""The four-step transform cursor (dvgpu.resident.FourStepPlan) on the CPU,
over an in-memory backend with the semantics of the reference's
CpuRebuildBackend transpose and row transforms
(crates/protocol/src/incremental_rebuild/operators.rs:387-445), checked
against the oracle's natural-order NTT.  The shapes follow the reference's
cursor tests (operators_tests.rs:27-50 and 280-303: tile 3, row batch 2,
lengths 1..128, retry of a committed cursor)."""
import copy

import pytest

from dvgpu.resident import FR, G1, FourStepPlan
from oracle import bn254 as o


class ListBackend:
    """Buffers are Python lists; every call is checked like the resident
    engine checks it, before anything is written."""

    def __init__(self):
        self.bufs, self.kinds, self.calls = [], [], 0

    def alloc(self, kind, n):
        self.bufs.append([0] * n)
        self.kinds.append(kind)
        return len(self.bufs) - 1

    def buffer_kind(self, h):
        return self.kinds[h]

    def buffer_len(self, h):
        return len(self.bufs[h])

    def transpose(self, kind, src, dst, rows, cols, start, length, twiddle=None):
        assert src != dst and start + length <= rows * cols <= len(self.bufs[dst])
        s, d = self.bufs[src], self.bufs[dst]
        for i in range(start, start + length):
            row, col = divmod(i, cols)
            f = pow(twiddle, row * col, o.R) if twiddle is not None else 1
            d[col * rows + row] = s[i] * f % o.R
        self.calls += 1

    def transform_rows(self, kind, src, dst, row_len, first_row, row_count, inverse):
        assert src != dst and (first_row + row_count) * row_len <= len(self.bufs[dst])
        s, d = self.bufs[src], self.bufs[dst]
        for r in range(first_row, first_row + row_count):
            d[r * row_len:(r + 1) * row_len] = o.ntt(s[r * row_len:(r + 1) * row_len], inverse)
        self.calls += 1

    def transform_four_step(self, kind, src, scratch, dst, n, inverse):
        self.bufs[dst][:n] = o.ntt(self.bufs[src][:n], inverse)
        self.calls += 1


def values(n):
    return [(i * i + 3 * i + 11) % o.R for i in range(n)]


def drain(plan, backend):
    steps = 0
    while True:
        steps += 1
        if plan.next_step(backend):
            return steps
        assert steps < 10_000


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128])
@pytest.mark.parametrize("inverse", [False, True])
def test_cursor_matches_natural_ntt(n, inverse):
    be = ListBackend()
    i, s, c = be.alloc(FR, n), be.alloc(FR, n), be.alloc(FR, n)
    be.bufs[i][:] = values(n)
    plan = FourStepPlan(FR, n, i, s, c, 3, 2, inverse)
    a, b = plan.dimensions()
    assert a * b == n and a <= b <= 2 * a
    assert drain(plan, be) == plan.total_steps == be.calls
    assert plan.is_complete() and plan.next_step(be)
    assert be.bufs[c] == o.ntt(values(n), inverse)


def test_session_65536_rebuild_shape_step_count():
    """An n = 65536 session's rebuild transforms have length 2n = 2^17 with the minimum
    session tile 2^ceil(17/2) = 512 and row batch 16 (program.rs:84-104,
    144-150): 256 x 512, 3 * 256 tiles + 512/16 + 256/16 row batches."""
    plan = FourStepPlan(FR, 1 << 17, 0, 1, 2, 512, 16, False)
    assert plan.dimensions() == (256, 512)
    assert plan.total_steps == 3 * 256 + 32 + 16
    init = plan.initialization_plan()
    assert (init.tile_len, init.row_batch, init.total_steps) == (1 << 16, 128, 3 * 2 + 4 + 2)


def test_invalid_plans_and_buffers():
    for args in [(FR, 12, 0, 1, 2, 3, 2), (FR, 16, 0, 1, 2, 0, 2), (FR, 16, 0, 1, 2, 3, 0),
                 (FR, 16, 0, 0, 2, 3, 2), (FR, 16, 0, 1, 1, 3, 2), (FR, 1 << 21, 0, 1, 2, 3, 2)]:
        with pytest.raises(ValueError):
            FourStepPlan(*args, False)
    be = ListBackend()
    i, s, c, short, pts = be.alloc(FR, 16), be.alloc(FR, 16), be.alloc(FR, 16), be.alloc(FR, 8), be.alloc(G1, 16)
    for bad in [FourStepPlan(FR, 16, i, s, short, 3, 2, False), FourStepPlan(FR, 16, i, pts, c, 3, 2, False)]:
        with pytest.raises(ValueError):
            bad.next_step(be)
        assert (bad.phase, bad.cursor, be.calls) == (0, 0, 0)     # nothing ran, cursor unchanged
    plan = FourStepPlan(FR, 16, i, s, c, 3, 2, False)
    plan.next_step(be)
    with pytest.raises(ValueError):
        plan.execute_packet(be)                                   # a packet needs an untouched plan


def test_committed_cursor_retried_from_its_copy():
    """A caller that abandons a partly run cursor resumes from the copy it
    committed (operators_tests.rs:280-303); the abandoned copy's progress
    does not leak into the committed one."""
    be = ListBackend()
    i, s, c = be.alloc(FR, 32), be.alloc(FR, 32), be.alloc(FR, 32)
    be.bufs[i][:] = values(32)
    committed = FourStepPlan(FR, 32, i, s, c, 3, 2, False)
    tentative = copy.copy(committed)
    for _ in range(15):
        assert not tentative.next_step(be)
    assert (committed.phase, committed.cursor) == (0, 0)
    be.bufs[s][:], be.bufs[c][:] = [0] * 32, [0] * 32             # what a rollback restores
    drain(committed, be)
    assert be.bufs[c] == o.ntt(values(32))
