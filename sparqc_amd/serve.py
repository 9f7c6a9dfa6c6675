"""Serving loop over the batch engine: open arrivals, HBM-budget admission,
continuous batching at gate-step granularity (no reference counterpart; the
reference runs one circuit per process, emulator/src/main.cu:87-148).

Requests arrive at given times (seconds from the start).  The loop
  1. admits every request that has arrived, in arrival order, while its
     pool bytes fit the engine's HBM budget (a request that does not fit waits
     at the head of the queue until finished circuits free room);
  2. advances every admitted, unfinished circuit by `step_gates` gates with
     one batched launch per bucket size (circuits that arrived while others
     were running join at the next step, so their work overlaps on the GPU);
  3. starts the export of each circuit that finished -- asynchronous, on the
     engine's export stream, into a reused page-locked buffer of the
     circuit's size class (buffers are kept for reuse, never freed inside the
     loop) -- and releases its pool space at once;
  4. collects the exports whose copies have completed and records each
     request's arrival -> result latency.
A request may carry its state in a page-locked buffer (`pinned`, see
pin_requests): it is then admitted without a host-side copy.
Only an idle loop waits for the device (for the oldest export, or sleeps until
the next arrival), so admissions, steps and result copies overlap.

A request the batch engine cannot hold -- one whose branching gates target
more qubits than a bucket holds, so some of its gates cross buckets
(SQE_ENOTLOCAL) -- is served on the single-circuit engine instead (one engine
per qubit count and precision, kept for reuse), inside the loop at its turn
in the arrival order; its latency counts like any other.

Each circuit advances by the same gate step from its own first gate whatever
else is in the batch, so its result does not depend on the mix: the solo and
mixed runs of a circuit agree to the bit (checked by tests/test_gpu_serve.py).
"""
import time
from collections import deque

import numpy as np

from . import engine as E


def percentile(values, q):
    if not values:
        return None
    return float(np.percentile(np.asarray(values, dtype=np.float64), q))


class ServeLoop:
    def __init__(self, prec="f64", hbm_budget_bytes=200 * 10**9, step_gates=64, renorm_every=None,
                 timing=False, device=0, batch=None):
        """batch: an existing engine.Batch (or an object with its methods) to
        serve on instead of a new one."""
        if renorm_every is None:
            renorm_every = 100 if prec in ("f32", E.F32) else 0
        if batch is None:
            batch = E.Batch(prec=prec, renorm_every=renorm_every, hbm_budget_bytes=hbm_budget_bytes,
                            timing=timing, device=device)
        self.batch = batch
        self._pool = {}                         # capacity -> page-locked export buffers, reused
        self._solo = {}                         # qubit count -> single-circuit engine for wide requests
        self._renorm_every = renorm_every
        self._device = device
        self.step_gates = int(step_gates)
        self.prec = getattr(batch, "prec", prec)

    @staticmethod
    def _size_class(n):
        """Buffer capacities are powers of two (at least 4096 entries), so a
        freed buffer serves every later export of its class."""
        return 1 << max(12, (int(n) - 1).bit_length())

    def _acquire(self, need):
        bufs = self._pool.get(self._size_class(need))
        if bufs:
            return bufs.pop()
        return self.batch.export_buffer(self._size_class(need))

    def _release(self, buf):
        self._pool.setdefault(buf.cap, []).append(buf)

    def reserve(self, count, cap):
        """Allocate `count` page-locked export buffers for exports of up to
        `cap` entries ahead of time (allocating inside the loop stalls it)."""
        c = self._size_class(cap)
        have = len(self._pool.get(c, []))
        for _ in range(count - have):
            self._release(self.batch.export_buffer(c))

    def close(self):
        for bufs in self._pool.values():
            for buf in bufs:
                buf.free()
        self._pool = {}
        for e in self._solo.values():
            e.close()
        self._solo = {}
        self.batch.close()

    def run_single(self, r):
        """A request with cross-bucket gates on the single-circuit engine:
        (state sorted by key, updates, nonzeros)."""
        e = self._solo.get(r["nq"])
        if e is None:
            e = E.Engine(r["nq"], prec=self.prec, renorm_every=self._renorm_every or 0, device=self._device)
            self._solo[r["nq"]] = e
        u0 = e.stats()["updates"]
        e.load(r["keys"], r["re"], r["im"])
        e.apply(r["gates"])
        state = e.export()
        st = e.stats()
        return state, st["updates"] - u0, st["nnz"]

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, requests, arrivals, keep_results=False, clock=time.perf_counter, sleep=time.sleep):
        """Serve `requests` (circuit dicts: nq, keys, re, im, gates) arriving at
        `arrivals` (seconds, nondecreasing).  Returns a report dict; with
        keep_results the exported states (sorted by key) are in
        report["results"][i].

        Nothing in the loop waits for the device except when it has nothing
        else to do: admissions enqueue their uploads, steps their launches,
        and a finished circuit's export runs on the engine's export stream
        into a page-locked buffer while later steps run; the loop collects
        exports whose copies have completed.  A request whose state the engine
        rejects (a duplicate key) is reported in report["errors"]."""
        n = len(requests)
        arrivals = np.asarray(arrivals, dtype=np.float64)
        assert arrivals.size == n and np.all(np.diff(arrivals) >= 0)
        b = self.batch
        nxt = 0                                 # next request not yet admitted
        live = {}                               # circuit id -> request index
        pending = deque()                       # (ticket, request index, buffer) of exports in flight
        latency = [None] * n
        admit_t = [None] * n
        results = [None] * n if keep_results else None
        errors = {}
        updates = 0
        steps = 0
        peak_live = 0
        peak_bytes = 0
        peak_pending = 0
        single = 0                              # requests served on the single-circuit engine
        t0 = clock()

        def collect(wait):
            nonlocal updates
            while pending:
                tk, i, buf = pending[0]
                try:
                    r = b.export_poll(tk, wait=wait)
                except E.SqeError as ex:
                    pending.popleft()
                    self._release(buf)
                    errors[i] = str(ex)
                    latency[i] = clock() - t0 - arrivals[i]
                    continue
                if r is None:
                    return
                pending.popleft()
                cnt, info = r
                updates += info["updates"]
                latency[i] = clock() - t0 - arrivals[i]
                if keep_results:
                    k = buf.keys[:cnt].copy()
                    o = np.argsort(k, kind="stable")
                    results[i] = dict(state=(k[o], buf.re[:cnt][o].copy(), buf.im[:cnt][o].copy()),
                                      updates=info["updates"], nnz=info["nnz"])
                self._release(buf)
                wait = False                    # waited for the oldest; take the rest only if done

        while nxt < n or live or pending:
            now = clock() - t0
            # 1. admission (FIFO, by HBM budget)
            while nxt < n and arrivals[nxt] <= now:
                r = requests[nxt]
                try:
                    if "pinned" in r:
                        cid = b.try_add(r["nq"], r["keys"], r["re"], r["im"], r["gates"], pinned=r["pinned"])
                    else:
                        cid = b.try_add(r["nq"], r["keys"], r["re"], r["im"], r["gates"])
                except E.SqeError as ex:
                    if ex.status != E.SQE_ENOTLOCAL:
                        raise
                    state, upd, nnz = self.run_single(r)
                    updates += upd
                    single += 1
                    latency[nxt] = clock() - t0 - arrivals[nxt]
                    admit_t[nxt] = clock() - t0
                    if keep_results:
                        results[nxt] = dict(state=state, updates=upd, nnz=nnz)
                    nxt += 1
                    continue
                if cid is None:
                    if not live:
                        raise E.SqeError(E.SQE_ETABLE_FULL,
                                         "request %d does not fit the HBM budget on an empty device" % nxt)
                    break
                live[cid] = nxt
                admit_t[nxt] = clock() - t0
                nxt += 1
            peak_live = max(peak_live, len(live))
            if live:
                u = b.usage()
                peak_bytes = max(peak_bytes, u["used"])
                # 2. one step of every admitted circuit
                b.step(self.step_gates)
                steps += 1
                # 3. finished circuits: start their exports, free their pool space
                for cid in b.finished():
                    i = live.pop(cid)
                    buf = self._acquire(b.export_bound(cid))
                    pending.append((b.export_start(cid, buf), i, buf))
                    b.remove(cid)
                peak_pending = max(peak_pending, len(pending))
            # 4. exports whose copies are done
            if pending:
                idle = not live and (nxt >= n or arrivals[nxt] > clock() - t0)
                collect(wait=idle)
            elif not live and nxt < n:
                wait = arrivals[nxt] - (clock() - t0)
                if wait > 0:
                    sleep(min(wait, 0.05))
        wall = clock() - t0
        lat = [x for x in latency if x is not None]
        rep = dict(requests=n, wall_s=wall, steps=steps, step_gates=self.step_gates, prec=self.prec,
                   updates=updates, updates_per_s=updates / wall if wall > 0 else 0.0,
                   p50_latency_s=percentile(lat, 50), p99_latency_s=percentile(lat, 99),
                   mean_latency_s=float(np.mean(lat)) if lat else None,
                   peak_live=peak_live, peak_pool_bytes=peak_bytes, peak_exports_in_flight=peak_pending,
                   single_engine_requests=single,
                   latency_s=latency, errors=errors,
                   queue_wait_s=[(a - arrivals[i]) if a is not None else None for i, a in enumerate(admit_t)])
        if keep_results:
            rep["results"] = results
        return rep


def pin_requests(requests):
    """Copies of the requests whose states sit in page-locked buffers
    (engine.PinnedArrays), as a front end that receives requests into such
    buffers would hold them: admission then uploads them without a host-side
    copy (sqe_batch_add_pinned).  The buffers stay allocated with the
    returned requests."""
    out = []
    for r in requests:
        n = int(r["keys"].size)
        buf = E.PinnedArrays(n)
        buf.keys[:n] = r["keys"]
        buf.re[:n] = r["re"]
        buf.im[:n] = r["im"]
        out.append(dict(r, pinned=(buf, n)))
    return out


def run_solo(request, prec="f64", step_gates=64, renorm_every=None):
    """One request alone through the same batch engine and gate step."""
    with ServeLoop(prec=prec, step_gates=step_gates, renorm_every=renorm_every,
                   hbm_budget_bytes=0) as s:
        rep = s.run([request], [0.0], keep_results=True)
    return rep["results"][0]
