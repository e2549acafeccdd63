"""roctx ranges (no GPU): a no-op unless PYGFMC_ROCTX=1; with it, ranges
nest through the ROCm roctx library when that is installed."""
import importlib

import pytest


def fresh(monkeypatch, value):
    if value is None:
        monkeypatch.delenv("PYGFMC_ROCTX", raising=False)
    else:
        monkeypatch.setenv("PYGFMC_ROCTX", value)
    from pygfmc_amd import trace
    return importlib.reload(trace)


def test_span_is_a_noop_by_default(monkeypatch):
    trace = fresh(monkeypatch, None)
    assert trace.library() is None
    with trace.span("x") as level:
        assert level is None


def test_spans_nest_when_enabled(monkeypatch):
    trace = fresh(monkeypatch, "1")
    if trace.library() is None:
        pytest.skip("no roctx library")
    with trace.span("outer") as a:
        with trace.span("inner") as b:
            assert b == a + 1
    fresh(monkeypatch, None)
