"""Handle lifetimes in the Python layer (host logic, no GPU): an engine is
destroyed before the context it was made on, even when a reference cycle
keeps the engine object alive past the point where the context is closed.

Every dspx_*_destroy entry point synchronises on its context's stream
(engine->ctx->stream), so destroying an engine after dspx_ctx_destroy would
read freed memory.  The library is replaced here by a stand-in that records
the order of create and destroy calls."""
import ctypes as C
import gc

import pytest


class _FakeLib:
    """Records create/destroy calls; handles are distinct integers."""

    def __init__(self):
        self.calls = []
        self.live_ctx = set()
        self.next = 100

    def _new(self, out):
        self.next += 1
        out._obj.value = self.next
        return self.next

    def dspx_ctx_create(self, device, stream, out):
        self.live_ctx.add(self._new(out))
        return 0

    def dspx_ctx_destroy(self, h):
        h = getattr(h, "value", h)
        self.calls.append(("ctx_destroy", h))
        self.live_ctx.discard(h)

    def dspx_sk_create(self, ctx, desc, out):
        self._new(out)
        return 0

    def dspx_sk_destroy(self, h):
        # the real entry point dereferences the engine's context here
        self.calls.append(("sk_destroy", getattr(h, "value", h), bool(self.live_ctx)))


@pytest.fixture
def fake(dspx, monkeypatch):
    f = _FakeLib()
    monkeypatch.setattr(dspx, "_lib", f)
    return f


def _raises_with_cycle(dspx, ctx):
    """Like `with pytest.raises(...) as e:` in a test: the exception's
    traceback holds this frame, whose local e holds the exception."""
    sk = dspx.SpectralKurtosis(ctx, 4, 2, 32)
    try:
        raise dspx.DspxError(dspx.E_STATE, "no limits")
    except dspx.DspxError as err:
        e = err
    assert e.code == dspx.E_STATE and sk.h


def test_cycle_keeps_engine_alive(dspx, fake):
    ctx = dspx.Context(0, None)
    gc.disable()
    try:
        n0 = len(dspx._live_engines)
        _raises_with_cycle(dspx, ctx)
        # the engine outlives the call that made it (until the cyclic GC runs)
        assert len(dspx._live_engines) == n0 + 1
        ctx.close()
    finally:
        gc.enable()
        gc.collect()
    kinds = [c[0] for c in fake.calls]
    assert kinds == ["sk_destroy", "ctx_destroy"]
    # the engine was destroyed while its context was still alive
    assert fake.calls[0][2] is True


def test_close_order_and_idempotence(dspx, fake):
    ctx = dspx.Context(0, None)
    sk = dspx.SpectralKurtosis(ctx, 4, 2, 32)
    ctx.close()
    assert sk.h is None and ctx.h is None
    sk.close()
    ctx.close()
    dspx._close_all()
    assert [c[0] for c in fake.calls] == ["sk_destroy", "ctx_destroy"]
