"""Bounded grid-barrier waits and launched widths of cooperative kernels.

Every grid barrier wait is bounded (CDS_BARRIER_TIMEOUT_US, default 60 s):
a workgroup whose wait runs out marks the barrier aborted and ends, every
other workgroup ends at its next barrier, and the call returns
CDS_ERR_HIP instead of hanging.  Here the limit is set to one microsecond
from the host, so the first barrier where any workgroup waits gives up:
every kernel still ends normally (no fault, no hang), the call fails with
the barrier message, and the next call on the same engine, with the
default limit, returns the oracle's answer.  The result also reports the
widths the cooperative kernels ran at.
"""
import os

import pytest

import cds_amd
from oracle_ref import oracle_solve
from test_gpu_parity import _announce, compare

pytestmark = pytest.mark.gpu

TUNING = {"list_threads": 0, "list_chunk": 0, "peel_blocks": 4, "peel_coarse_shift": -1,
          "replay_blocks": 0, "flow_blocks": 0}


def _graph():
    return cds_amd.Graph.chung_lu(3000, 2.3, 10.0, 30, 0.85, seed=21)


def _solve_with_limit(engine, dg, k, us):
    os.environ["CDS_BARRIER_TIMEOUT_US"] = str(us)
    try:
        return engine.solve(dg, k, vertex_arrays=True)
    finally:
        del os.environ["CDS_BARRIER_TIMEOUT_US"]


def _engine(env=None):
    saved = {k: os.environ.get(k) for k in (env or {})}
    os.environ.update(env or {})
    try:
        return cds_amd.Engine(device=0, tune_cache="")
    finally:
        for key, v in saved.items():
            if v is None:
                del os.environ[key]
            else:
                os.environ[key] = v


def test_peel_barrier_wait_gives_up_then_engine_recovers():
    g = _graph()
    want = oracle_solve(g, 3)
    with _engine() as eng:
        eng.set_tuning(TUNING)  # the persistent peel on 4 workgroups
        dg = eng.upload(g)
        _announce("peel barrier limit 1 us")
        with pytest.raises(cds_amd.CdsError) as ei:
            _solve_with_limit(eng, dg, 3, 1)
        assert ei.value.status == 3 and "barrier" in str(ei.value), str(ei.value)
        _announce("peel default limit")
        got = eng.solve(dg, 3, vertex_arrays=True)
        compare(got, want, "after a given-up peel barrier")
        assert got["peel_grid"] == 4 and got["coop_shrinks"] == 0
        dg.free()


def test_flow_barrier_wait_gives_up_then_engine_recovers():
    g = _graph()
    want = oracle_solve(g, 3)
    with _engine({"CDS_FLOW_BLOCKS": "3"}) as eng:
        eng.set_tuning(dict(TUNING, peel_blocks=1))  # peel as a plain launch
        dg = eng.upload(g)
        got = eng.solve(dg, 3, vertex_arrays=True)
        assert got["flow_launches"] > 0 and got["flow_grid"] == 3
        assert got["peel_grid"] == 1 and got["coop_shrinks"] == 0
        _announce("flow barrier limit 1 us")
        with pytest.raises(cds_amd.CdsError) as ei:
            _solve_with_limit(eng, dg, 3, 1)
        assert ei.value.status == 3 and "max flow" in str(ei.value), str(ei.value)
        _announce("flow default limit")
        compare(eng.solve(dg, 3, vertex_arrays=True), want, "after a given-up flow barrier")
        dg.free()
