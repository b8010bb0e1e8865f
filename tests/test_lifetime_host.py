"""Handle lifetime of the Python binding, on the CPU (no device).

A Context and a Stepper that refer to each other through a frame kept by a traceback form a
garbage cycle (pytest.raises(...) as e keeps e -> traceback -> frame -> {st, ctx}). CPython's
cycle collector clears weak references before it runs finalisers and finalises the objects of
a cycle in no fixed order, so the binding may not rely on closing dependents first. The
library's contract (include/of_hip.h: of_context_destroy) is that a stepper or adjoint run
holds a reference of its own on its context, and the native context is freed with the last
reference. These tests drive the binding against a stand-in library that enforces that
contract and fails on any use of a freed context, and check that both collection orders free
everything exactly once.
"""
from __future__ import annotations

import gc
import types

import pytest

import solver_amd as sa


class FakeLib:
    """Stand-in for libsolver_amd.so implementing the handle contract of of_hip.h."""

    def __init__(self):
        self.log = []
        self.contexts = {}
        self.objects = {}
        self.next = 1

    def _new(self):
        h, self.next = self.next, self.next + 1
        return h

    @staticmethod
    def _out(byref_arg, value):
        byref_arg._obj.value = value

    def _context(self, h):
        c = self.contexts[h]
        if c["freed"]:
            raise AssertionError("use of a freed native context")
        return c

    def _release(self, h):
        c = self._context(h)
        c["refs"] -= 1
        if c["refs"] == 0:
            c["freed"] = True
            self.log.append(("free_context", h))

    def of_context_create(self, device, stream, out):
        h = self._new()
        self.contexts[h] = {"refs": 1, "freed": False}
        self._out(out, h)
        return sa.OF_OK

    def of_context_destroy(self, h):
        self.log.append(("of_context_destroy", h.value))
        self._release(h.value)
        return sa.OF_OK

    def _create_dependent(self, kind, ctx, out):
        self._context(ctx.value)["refs"] += 1
        h = self._new()
        self.objects[h] = (kind, ctx.value)
        self._out(out, h)
        return sa.OF_OK

    def _destroy_dependent(self, kind, h):
        k, ctx = self.objects.pop(h.value)
        assert k == kind
        self.log.append((f"of_{kind}_destroy", h.value))
        self._context(ctx)  # the native destroy reads its context's device and stream
        self._release(ctx)
        return sa.OF_OK

    def of_stepper_create(self, ctx, cfg, out):
        return self._create_dependent("stepper", ctx, out)

    def of_stepper_destroy(self, h):
        return self._destroy_dependent("stepper", h)

    def of_adjoint_create(self, ctx, cfg, every, out):
        return self._create_dependent("adjoint", ctx, out)

    def of_adjoint_destroy(self, h):
        return self._destroy_dependent("adjoint", h)

    def of_last_error(self):
        return b""


@pytest.fixture
def fake(monkeypatch):
    lib = FakeLib()
    monkeypatch.setattr(sa, "_lib", lib)
    fake_torch = types.SimpleNamespace(cuda=types.SimpleNamespace(is_available=lambda: True,
                                                                 synchronize=lambda: None))
    monkeypatch.setattr(sa, "_torch", lambda: fake_torch)
    gc.collect()
    yield lib
    gc.collect()


def make_cycle(kind):
    ctx = sa.Context(0)
    if kind == "stepper":
        obj = sa.Stepper(ctx, sa.TAYLOR_GREEN, 8, 8, 8, 0.01, 0.01)
    else:
        obj = sa.Adjoint(ctx, 8, 8, 8, 0.01, 0.01, 2)
    try:
        raise sa.SolverError("keep this frame alive through the traceback")
    except sa.SolverError as e:
        err = e
    # err -> traceback -> this frame -> {ctx, obj, err}: a cycle only the collector frees
    return err


@pytest.mark.parametrize("kind", ["stepper", "adjoint"])
@pytest.mark.parametrize("context_first", [True, False])
def test_garbage_cycle_frees_context_after_its_dependent(fake, kind, context_first):
    gc.disable()
    try:
        make_cycle(kind)
        ctx = next(o for o in gc.get_objects() if isinstance(o, sa.Context))
        dep = next(o for o in gc.get_objects() if isinstance(o, (sa.Stepper, sa.Adjoint)))
        # either finaliser may run first inside one collection; force each order explicitly
        first, second = (ctx, dep) if context_first else (dep, ctx)
        first.close()
        second.close()
        del ctx, dep, first, second
        gc.collect()
    finally:
        gc.enable()
    names = [n for n, _ in fake.log]
    assert names.count("of_context_destroy") == 1
    assert names.count(f"of_{kind}_destroy") == 1
    # the native context is freed once, and only after the dependent that uses it
    assert names.count("free_context") == 1
    assert names.index("free_context") > names.index(f"of_{kind}_destroy")


@pytest.mark.parametrize("kind", ["stepper", "adjoint"])
def test_collector_alone_frees_everything(fake, kind):
    make_cycle(kind)
    gc.collect()
    names = [n for n, _ in fake.log]
    assert names.count("of_context_destroy") == 1
    assert names.count(f"of_{kind}_destroy") == 1
    assert names.count("free_context") == 1
    assert all(c["freed"] for c in fake.contexts.values())
    assert not fake.objects


def test_stand_in_flags_the_old_order(fake):
    """The stand-in catches the failure this guards against: a destroy of a context that frees
    it while a stepper still refers to it makes the stepper's destroy fail."""
    ctx = sa.Context(0)
    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, 8, 8, 8, 0.01, 0.01)
    fake.contexts[ctx.h.value]["refs"] = 1  # a library without the dependent's reference
    ctx.close()
    with pytest.raises(AssertionError, match="freed native context"):
        st.close()


def test_close_is_idempotent(fake):
    ctx = sa.Context(0)
    st = sa.Stepper(ctx, sa.TAYLOR_GREEN, 8, 8, 8, 0.01, 0.01)
    st.close()
    st.close()
    ctx.close()
    ctx.close()
    names = [n for n, _ in fake.log]
    assert names == ["of_stepper_destroy", "of_context_destroy", "free_context"]
