"""Host logic of ibpm_amd.advance_concurrently (no GPU): one thread per batch,
every batch advanced by the same step count, the first error re-raised after
all threads have ended, and batches sharing a context refused."""
import threading
import types

import pytest

import ibpm_amd


class FakeBatch:
    """Stands in for ibpm_amd.Batch: a context handle and an advance()."""

    def __init__(self, handle, fail=False, delay=None):
        self.ctx = types.SimpleNamespace(h=types.SimpleNamespace(value=handle))
        self.fail, self.delay = fail, delay
        self.steps, self.threads = 0, set()

    def advance(self, n):
        self.threads.add(threading.get_ident())
        if self.delay is not None:   # an Event or a Barrier
            self.delay.wait(5.0)
        if self.fail:
            raise ibpm_amd.IbpmError("advance failed")
        self.steps += n


def test_every_batch_advances_on_its_own_thread():
    meet = threading.Barrier(3)   # all three advances run at once (thread ids stay distinct)
    bs = [FakeBatch(h, delay=meet) for h in (11, 12, 13)]
    ibpm_amd.advance_concurrently(bs, 7)
    assert [b.steps for b in bs] == [7, 7, 7]
    tids = set().union(*(b.threads for b in bs))
    assert len(tids) == 3 and threading.get_ident() not in tids


def test_first_error_is_raised_after_every_thread_ends():
    gate = threading.Event()
    slow = FakeBatch(21, delay=gate)
    bad = FakeBatch(22, fail=True)
    timer = threading.Timer(0.2, gate.set)   # the slow batch finishes after the failing one
    timer.start()
    with pytest.raises(ibpm_amd.IbpmError):
        ibpm_amd.advance_concurrently([slow, bad], 3)
    assert slow.steps == 3   # joined before the error came back
    timer.cancel()


def test_batches_sharing_a_context_are_refused():
    bs = [FakeBatch(31), FakeBatch(31)]
    with pytest.raises(ValueError):
        ibpm_amd.advance_concurrently(bs, 1)
    assert [b.steps for b in bs] == [0, 0]
