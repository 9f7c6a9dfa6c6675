"""Parity at the sizes the benchmarks run.

* 48 qubits, 2^20 distinct keys, the seeded 2000-gate 60/35/5 circuit (the
  headline workload; nonzeros grow to 2^24): FP64 engine against the oracle --
  bit-exact key set, amplitudes within 1e-12, equal update counts.  The FP32
  form of the same run (renormalisation every 100 gates, gate by gate and with
  fused blocks) against the FP64 oracle: infidelity <= 1e-6, max |a32 - a64|
  printed and bounded.
* 256 independent circuits of 16-40 qubits, 2^8-2^18 initial keys, depth
  50-500 (the batched workload, generator defaults): every circuit of the FP64
  batch run against the oracle, and the FP32 batch run against the same
  oracle states.
* Serving at 1000 requests/s with requests drawn like the batched workload:
  every result equals the same request served alone, to the bit, and the
  oracle within 1e-12.
* 2^18 keys branching on 6 qubits (cross-bucket launches that grow the
  stores) against the oracle.

The oracle needs about a minute on 16 host cores for the 48-qubit circuit;
these tests carry their own time limits.
"""
import sys
import time

import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gen
from sparqc_amd.serve import ServeLoop, run_solo
from tests.helpers import assert_same_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_chunked(O, w, label, chunk=200):
    """Run the oracle in chunks, printing progress (a long silent call looks hung)."""
    s = O.Sparse(w["nq"])
    s.load(w["keys"], w["re"], w["im"])
    gl = w["gates"]
    t0 = time.perf_counter()
    for a in range(0, len(gl), chunk):
        s.apply(gl[a:a + chunk])
        say("  oracle %s: %d/%d gates, %d nonzeros, %.1f s" % (label, min(a + chunk, len(gl)), len(gl), s.nnz(),
                                                                time.perf_counter() - t0))
    return s


def fidelity(got, want):
    """1 - |<want|got>|^2 / (|want|^2 |got|^2) over the key union, and max |got - want|."""
    gk, gr, gi = got
    wk, wr, wi = want
    keys = np.union1d(gk, wk)
    a = np.zeros(keys.size, dtype=np.complex128)
    b = np.zeros(keys.size, dtype=np.complex128)
    a[np.searchsorted(keys, gk)] = gr + 1j * gi
    b[np.searchsorted(keys, wk)] = wr + 1j * wi
    ov = np.vdot(b, a)
    infid = 1.0 - abs(ov) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
    return float(infid), float(np.max(np.abs(a - b)))


@pytest.fixture(scope="module")
def headline(oracle):
    """The benchmarked 48-qubit workload and its oracle result."""
    w = gen.circuit_large(1, nq=48, n_init=1 << 20, n_gates=2000)
    say("48-qubit 2^20-key 2000-gate workload: oracle on %d threads" % oracle.get_threads())
    s = oracle_chunked(oracle, w, "48q")
    return w, s.export(), s.updates()


@pytest.mark.timeout(900)
def test_headline_fp64_full_size(headline):
    w, want, upd = headline
    say("GPU FP64 48q 2^20 keys 2000 gates")
    with E.Engine(48, prec="f64") as e:
        e.load(w["keys"], w["re"], w["im"])
        e.prepare(w["gates"])
        e.apply(w["gates"])
        got = e.export()
        st = e.stats()
    say("  %d nonzeros (oracle %d), updates %d (oracle %d)" % (got[0].size, want[0].size, st["updates"], upd))
    assert_same_state(got, want, atol=1e-12, label="48q FP64")
    assert st["updates"] == upd
    assert got[0].size > (1 << 23)               # the state did grow toward 2^24


@pytest.mark.timeout(900)
@pytest.mark.parametrize("blocks", [0, 2])
def test_headline_fp32_full_size(headline, blocks):
    w, want, _ = headline
    say("GPU FP32 48q renorm every 100, fused blocks %d" % blocks)
    with E.Engine(48, prec="f32", renorm_every=100, fuse_blocks=blocks) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.prepare(w["gates"])
        e.apply(w["gates"])
        got = e.export()
    infid, dmax = fidelity(got, want)
    n2 = float(np.sum(got[1] ** 2 + got[2] ** 2))
    say("  FP32 (blocks %d): 1-F = %.3e, max |a32 - a64| = %.3e, norm2 %.9f, keys %d vs %d"
        % (blocks, infid, dmax, n2, got[0].size, want[0].size))
    assert infid <= 1e-6
    assert dmax <= 1e-5
    assert abs(n2 - 1.0) <= 1e-5


@pytest.fixture(scope="module")
def batch_workload(oracle):
    cs = gen.batch_circuits(1, n_circuits=256)
    say("256-circuit batch: oracle per circuit")
    t0 = time.perf_counter()
    want = []
    for i, c in enumerate(cs):
        s = oracle.Sparse(c["nq"])
        s.load(c["keys"], c["re"], c["im"])
        s.apply(c["gates"])
        want.append((s.export(), s.updates()))
        if i % 32 == 31:
            say("  oracle: %d/256 circuits, %.1f s" % (i + 1, time.perf_counter() - t0))
    return cs, want


@pytest.mark.timeout(900)
def test_batch_256_circuits_full_size(batch_workload):
    cs, want = batch_workload
    say("GPU FP64 batch of 256 circuits (captured run, two tiles per work item as the autotuner picks)")
    with E.Batch(prec="f64") as b:
        b.set_tuning(tiles_per_item=2)
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.run()
        b.sync()
        for i, cid in enumerate(ids):
            got = b.export(cid)
            assert_same_state(got, want[i][0], atol=1e-12, label="circuit %d (nq %d)" % (i, cs[i]["nq"]))
            assert b.info(cid)["updates"] == want[i][1], "circuit %d update count" % i
    say("  all 256 circuits match the oracle")


@pytest.mark.timeout(900)
def test_batch_256_circuits_fp32_full_size(batch_workload):
    cs, want = batch_workload
    say("GPU FP32 batch of 256 circuits (renorm every 100, two tiles per work item)")
    worst_f, worst_d = 0.0, 0.0
    with E.Batch(prec="f32", renorm_every=100) as b:
        b.set_tuning(tiles_per_item=2)
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.run()
        b.sync()
        for i, cid in enumerate(ids):
            f, d = fidelity(b.export(cid), want[i][0])
            worst_f, worst_d = max(worst_f, f), max(worst_d, d)
    say("  worst 1-F %.3e, worst max |a32 - a64| %.3e" % (worst_f, worst_d))
    assert worst_f <= 1e-6
    assert worst_d <= 1e-5


@pytest.mark.timeout(900)
def test_serving_1000_per_s_full_size(oracle):
    reqs = gen.batch_circuits(31, n_circuits=48)
    arr = gen.poisson_arrivals(32, 1000.0, len(reqs))
    say("serve 48 requests of the batched workload at 1000/s")
    with ServeLoop(prec="f64", step_gates=64, hbm_budget_bytes=200 * 10**9) as s:
        rep = s.run(reqs, arr, keep_results=True)
    say("  report", {k: rep[k] for k in ("wall_s", "steps", "peak_live", "p50_latency_s", "p99_latency_s")})
    assert rep["peak_live"] >= 2
    for i, r in enumerate(reqs):
        mix = rep["results"][i]
        solo = run_solo(r, prec="f64", step_gates=64)
        assert_same_state(mix["state"], solo["state"], atol=0.0, label="request %d mix vs solo" % i)
        assert mix["updates"] == solo["updates"]
        o = oracle.Sparse(r["nq"])
        o.load(r["keys"], r["re"], r["im"])
        o.apply(r["gates"])
        assert_same_state(mix["state"], o.export(), atol=1e-12, label="request %d vs oracle" % i)
        assert mix["updates"] == o.updates()
        if i % 8 == 7:
            say("  %d/%d requests checked" % (i + 1, len(reqs)))


@pytest.mark.timeout(900)
def test_wide_branching_with_store_growth(oracle):
    """2^18 keys branching on 6 qubits (2 of them outside the buckets): the
    cross-bucket launches grow the stores while the state reaches 2^24
    nonzeros, with grids larger than the device holds at once, so an output
    store that aliased the input would corrupt the state."""
    w = gen.circuit_large(1, nq=48, n_init=1 << 18, n_gates=2000, n_branch=6)
    w = dict(w, gates=w["gates"][:500])
    say("48 qubits, 2^18 keys, 6 branching qubits, 500 gates")
    with E.Engine(48) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.apply(w["gates"])
        got = e.export()
        st = e.stats()
    s = oracle_chunked(oracle, w, "wide")
    say("  %d nonzeros, %d cross-bucket launches" % (got[0].size, st["pairs"]))
    assert st["pairs"] > 0
    assert_same_state(got, s.export(), atol=1e-12, label="wide branching")
    assert st["updates"] == s.updates()
