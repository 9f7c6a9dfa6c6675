"""Serving loop on the GPU: each result is the same alone and in the mix.

Requests of the batched workload shape (uneven qubit count, nonzero count and
depth) arrive as a Poisson stream; every exported state must equal, bit for
bit, the state of the same request served alone through the same gate step
(FP64 and FP32), and FP64 states must match the oracle within 1e-12.
"""
import sys

import numpy as np
import pytest

from sparqc_amd import gen
from sparqc_amd.serve import ServeLoop, pin_requests, run_solo
from tests.helpers import assert_same_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def requests(seed, n):
    return gen.batch_circuits(seed, n_circuits=n, nq_range=(16, 30), log2_nnz=(6.0, 12.0), depth_range=(40, 220))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_mix_equals_solo(oracle, prec):
    reqs = requests(21, 14)
    arr = gen.poisson_arrivals(5, 3000.0, len(reqs))
    say("serve mix", prec, len(reqs), "requests")
    with ServeLoop(prec=prec, step_gates=32, hbm_budget_bytes=1 << 30) as s:
        rep = s.run(reqs, arr, keep_results=True)
    say("report", {k: rep[k] for k in ("wall_s", "steps", "peak_live", "p50_latency_s", "p99_latency_s")})
    assert rep["peak_live"] >= 2, "requests should overlap on the device"
    for i, r in enumerate(reqs):
        solo = run_solo(r, prec=prec, step_gates=32)
        mix = rep["results"][i]
        assert_same_state(mix["state"], solo["state"], atol=0.0, label="request %d" % i)
        assert mix["updates"] == solo["updates"]
        if prec == "f64":
            o = oracle.Sparse(r["nq"])
            o.load(r["keys"], r["re"], r["im"])
            o.apply(r["gates"])
            assert_same_state(mix["state"], o.export(), atol=1e-12, label="oracle %d" % i)
            assert mix["updates"] == o.updates()
    assert rep["p50_latency_s"] <= rep["p99_latency_s"]


def test_budget_limited_serving_completes(oracle):
    """A budget that holds only a few requests at a time still serves all."""
    reqs = requests(22, 10)
    arr = np.zeros(len(reqs))
    with ServeLoop(prec="f64", step_gates=64, hbm_budget_bytes=1 << 20) as s:
        rep = s.run(reqs, arr, keep_results=True)
    say("budget-limited", rep["peak_live"], rep["peak_pool_bytes"])
    assert rep["peak_pool_bytes"] <= 1 << 20
    assert rep["peak_live"] < len(reqs)
    for i, r in enumerate(reqs):
        o = oracle.Sparse(r["nq"])
        o.load(r["keys"], r["re"], r["im"])
        o.apply(r["gates"])
        assert_same_state(rep["results"][i]["state"], o.export(), atol=1e-12, label="request %d" % i)


def test_cross_bucket_requests_served_beside_the_batch(oracle):
    """Requests branching on 6 qubits (more than a bucket holds) go to the
    single-circuit engine; with the batched ones they all match the oracle."""
    reqs = requests(23, 6)
    for k in (1, 4):
        w = gen.circuit_large(40 + k, nq=22, n_init=64, n_gates=150, n_branch=6)
        reqs[k] = dict(nq=22, keys=w["keys"], re=w["re"], im=w["im"], gates=w["gates"])
    arr = gen.poisson_arrivals(6, 2000.0, len(reqs))
    with ServeLoop(prec="f64", step_gates=32, hbm_budget_bytes=1 << 30) as s:
        rep = s.run(reqs, arr, keep_results=True)
    say("single-engine requests:", rep["single_engine_requests"])
    assert rep["single_engine_requests"] == 2
    for i, r in enumerate(reqs):
        o = oracle.Sparse(r["nq"])
        o.load(r["keys"], r["re"], r["im"])
        o.apply(r["gates"])
        assert_same_state(rep["results"][i]["state"], o.export(), atol=1e-12, label="request %d" % i)
        assert rep["results"][i]["updates"] == o.updates()


def test_pinned_requests_mix_equals_solo(oracle):
    """Requests held in page-locked buffers (admitted without a host copy)
    give the same results as the same requests admitted by copy."""
    reqs = requests(24, 10)
    pinned = pin_requests(reqs)
    arr = gen.poisson_arrivals(7, 3000.0, len(reqs))
    out = []
    for rs in (pinned, reqs):
        with ServeLoop(prec="f64", step_gates=32, hbm_budget_bytes=1 << 30) as s:
            out.append(s.run(rs, arr, keep_results=True))
    for i, r in enumerate(reqs):
        assert_same_state(out[0]["results"][i]["state"], out[1]["results"][i]["state"], atol=0.0,
                          label="request %d pinned vs copied" % i)
        assert out[0]["results"][i]["updates"] == out[1]["results"][i]["updates"]
