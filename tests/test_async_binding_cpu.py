"""Host logic of the Python binding's run_async/sync (no GPU): the arrays an
in-flight run writes into stay referenced until sync(), a refused call
replaces nothing, and sync() with nothing pending returns (None, None)."""
import numpy as np
import pytest


class FakeLib:
    """Stands in for the C library: pic_run_async succeeds once, then refuses
    with PIC_ESTATE until pic_sync."""

    def __init__(self):
        self.busy = False
        self.target = None

    def pic_run_async(self, h, nsteps, D, S):
        if self.busy:
            return -5
        self.busy = True
        self.target = (D, S)
        return 0

    def pic_sync(self, h):
        self.busy = False
        return 0

    def pic_last_error(self, h):
        return b"a pic_run_async run is in flight; call pic_sync first"

    def pic_destroy(self, h):
        return 0


def fake_context():
    from picgolf.core import Config, Context
    c = Context.__new__(Context)
    c._lib = FakeLib()
    c._h = object()
    c.configs = [Config(N=64, P=2048, W=400.0, dt=1 / 384)]
    c.N = 64
    return c


def test_refused_run_async_keeps_the_in_flight_arrays():
    from picgolf._lib import PicError
    c = fake_context()
    c.run_async(16)
    D, S = c._pending
    assert D.shape == (1, 16, 4) and S.shape == (1, 16)
    with pytest.raises(PicError) as e:
        c.run_async(1)
    assert e.value.code == -5
    assert c._pending[0] is D and c._pending[1] is S
    got = c.sync()
    assert got[0] is D and got[1] is S
    assert c.sync() == (None, None)
    c._h = None


def test_sync_without_a_run_returns_nothing():
    c = fake_context()
    assert c.sync() == (None, None)
    c._h = None
