"""Host logic of the serving loop (sparqc_amd/serve.py) on a model batch engine.

The model admits by bytes against a budget, advances each circuit by the gate
step per call and reports finished circuits like engine.Batch; the clock is
virtual (each step takes a fixed time), so admission order, the head-of-line
wait on the budget and the latency bookkeeping are checked exactly.
"""
import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd.serve import ServeLoop, percentile


class ModelBatch:
    prec = "f64"

    def __init__(self, budget, clock):
        self.budget, self.used, self.clock = budget, 0, clock
        self.c = {}
        self.next_id = 0
        self.log = []

    def try_add(self, nq, keys, re, im, gates):
        if nq > 40:                              # the model's stand-in for a cross-bucket circuit
            raise E.SqeError(E.SQE_ENOTLOCAL, "gate 0 crosses buckets")
        need = int(keys.size) * 100
        if self.used + need > self.budget:
            return None
        cid = self.next_id
        self.next_id += 1
        self.c[cid] = dict(left=len(gates), bytes=need, n=len(gates), nnz=int(keys.size))
        self.used += need
        self.log.append(("add", cid, self.clock.t))
        return cid

    def usage(self):
        return {"used": self.used, "budget": self.budget,
                "active": sum(1 for v in self.c.values() if v["left"] > 0)}

    def step(self, k):
        for v in self.c.values():
            v["left"] = max(0, v["left"] - k)
        self.clock.t += 0.01                     # a step takes 10 ms
        self.log.append(("step", self.clock.t))

    def finished(self):
        return [cid for cid, v in self.c.items() if v["left"] == 0]

    def export(self, cid, sort=True):
        v = self.c[cid]
        return (np.arange(v["nnz"], dtype=np.uint64), np.zeros(v["nnz"]), np.zeros(v["nnz"]))

    def export_buffer(self, cap=0):
        return Buffer(cap)

    def export_bound(self, cid):
        return self.c[cid]["nnz"]

    def export_start(self, cid, out):
        """The export completes at the next poll (no device in the model)."""
        v = self.c[cid]
        keys, re, im = self.export(cid)
        out.keys, out.re, out.im = keys, re, im
        self.next_ticket = getattr(self, "next_ticket", 0) + 1
        self.tickets = getattr(self, "tickets", {})
        self.tickets[self.next_ticket] = (v["nnz"], {"updates": v["n"] * v["nnz"], "nnz": v["nnz"]})
        self.log.append(("export", cid, self.clock.t))
        return self.next_ticket

    def export_poll(self, ticket, wait=False):
        return self.tickets.pop(ticket)

    def info(self, cid):
        v = self.c[cid]
        return {"done": int(v["left"] == 0), "updates": v["n"] * v["nnz"], "nnz": v["nnz"]}

    def remove(self, cid):
        self.used -= self.c.pop(cid)["bytes"]


class Buffer:
    keys = re = im = None

    def __init__(self, cap=0):
        self.cap = cap

    def free(self):
        pass


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


def req(nkeys, ngates):
    return dict(nq=20, keys=np.arange(nkeys, dtype=np.uint64), re=np.ones(nkeys), im=np.zeros(nkeys),
                gates=[None] * ngates)


def test_fifo_admission_and_latency():
    clk = Clock()
    mb = ModelBatch(budget=10**9, clock=clk)
    s = ServeLoop(batch=mb, step_gates=10)
    reqs = [req(10, 25), req(10, 5), req(10, 40)]
    rep = s.run(reqs, [0.0, 0.0, 0.015], clock=clk, sleep=clk.sleep)
    # r0: 3 steps (done at 0.03); r1: 1 step (0.01); r2 admitted after step 2 (t=0.02) -> 4 steps -> 0.06
    assert rep["latency_s"][1] == pytest.approx(0.01)
    assert rep["latency_s"][0] == pytest.approx(0.03)
    assert rep["latency_s"][2] == pytest.approx(0.06 - 0.015)
    assert rep["steps"] == 6
    assert rep["updates"] == 25 * 10 + 5 * 10 + 40 * 10
    assert rep["p50_latency_s"] == pytest.approx(percentile(rep["latency_s"], 50))


def test_budget_head_of_line_wait():
    """A request that does not fit waits (FIFO) until finished circuits free room."""
    clk = Clock()
    mb = ModelBatch(budget=2000, clock=clk)
    s = ServeLoop(batch=mb, step_gates=10)
    reqs = [req(15, 30), req(10, 10), req(2, 10)]
    rep = s.run(reqs, [0.0, 0.0, 0.0], clock=clk, sleep=clk.sleep)
    adds = [e for e in mb.log if e[0] == "add"]
    assert [a[1] for a in adds] == [0, 1, 2]
    # request 1 (1000 bytes) cannot join request 0 (1500 of 2000), nor can the
    # later small request 2 jump the queue: both wait for request 0 to finish
    assert adds[1][2] == pytest.approx(0.03) and adds[2][2] == pytest.approx(0.03)
    assert rep["queue_wait_s"][1] == pytest.approx(0.03)
    assert all(x is not None for x in rep["latency_s"])


def test_idle_gap_sleeps_until_next_arrival():
    clk = Clock()
    mb = ModelBatch(budget=10**9, clock=clk)
    s = ServeLoop(batch=mb, step_gates=100)
    rep = s.run([req(4, 10), req(4, 10)], [0.0, 1.0], clock=clk, sleep=clk.sleep)
    assert rep["latency_s"][0] == pytest.approx(0.01)
    assert rep["latency_s"][1] == pytest.approx(0.01, abs=1e-9)
    assert rep["steps"] == 2


def test_request_larger_than_budget_is_an_error():
    clk = Clock()
    mb = ModelBatch(budget=100, clock=clk)
    s = ServeLoop(batch=mb, step_gates=10)
    with pytest.raises(E.SqeError):
        s.run([req(5, 10)], [0.0], clock=clk, sleep=clk.sleep)


def test_failed_export_is_reported_and_the_rest_complete():
    """An export the engine rejects (a duplicate key in that request's state)
    is reported in the report's errors; the other requests are served."""
    clk = Clock()
    mb = ModelBatch(budget=10**9, clock=clk)
    poll = mb.export_poll
    bad = {}

    def failing_poll(ticket, wait=False):
        if ticket in bad:
            mb.tickets.pop(ticket)
            raise E.SqeError(E.SQE_EINVAL, "duplicate key in the circuit's load")
        return poll(ticket, wait)

    start = mb.export_start

    def marking_start(cid, out):
        t = start(cid, out)
        if cid == 1:
            bad[t] = True
        return t

    mb.export_poll, mb.export_start = failing_poll, marking_start
    s = ServeLoop(batch=mb, step_gates=10)
    rep = s.run([req(4, 10), req(4, 10), req(4, 20)], [0.0, 0.0, 0.0], clock=clk, sleep=clk.sleep)
    assert list(rep["errors"]) == [1]
    assert all(x is not None for x in rep["latency_s"])
    assert rep["updates"] == 10 * 4 + 20 * 4


def test_cross_bucket_request_served_on_single_engine():
    """A request the batch refuses as not bucket-local (SQE_ENOTLOCAL) is
    served on the single-circuit engine at its turn; the others are batched."""
    clk = Clock()
    mb = ModelBatch(budget=10**9, clock=clk)
    s = ServeLoop(batch=mb, step_gates=10)
    ran = []

    def single(r):
        ran.append(r["nq"])
        clk.t += 0.005                          # the single engine takes 5 ms
        n = int(r["keys"].size)
        return (np.arange(n, dtype=np.uint64), np.zeros(n), np.zeros(n)), 7, n

    s.run_single = single
    wide = dict(req(6, 30), nq=48)
    rep = s.run([req(10, 10), wide, req(10, 10)], [0.0, 0.0, 0.0], clock=clk, sleep=clk.sleep, keep_results=True)
    assert ran == [48]
    assert rep["single_engine_requests"] == 1
    assert rep["latency_s"][1] == pytest.approx(0.005)
    assert rep["results"][1]["updates"] == 7 and rep["results"][1]["nnz"] == 6
    assert [e[1] for e in mb.log if e[0] == "add"] == [0, 1]     # requests 0 and 2 were batched
    assert rep["updates"] == 10 * 10 + 7 + 10 * 10
