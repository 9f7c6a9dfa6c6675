"""GPU parity of the batch engine (sqe_batch_* through the C ABI).

Every circuit run inside a batch must give what the oracle gives for it alone:
bit-exact keys and FP64 amplitudes within 1e-12 (same gate arithmetic as the
single-circuit engine), the same update count, and -- for FP32 with FP64
renormalisation -- agreement with the FP32 oracle within 1e-6.
"""
import sys

import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gates as G
from sparqc_amd import gen
from tests.helpers import assert_same_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_batch(seed, n, nq=(14, 22), log2_nnz=(4.0, 11.0), depth=(40, 160)):
    return gen.batch_circuits(seed, n_circuits=n, nq_range=nq, log2_nnz=log2_nnz, depth_range=depth)


def oracle_of(O, c, f32=False, renorm_every=0):
    s = O.Sparse(c["nq"], f32=f32, thr=1e-7 if f32 else 1e-14, renorm_every=renorm_every)
    s.load(c["keys"], c["re"], c["im"])
    s.apply(c["gates"])
    return s.export(), s.updates()


def test_batch_matches_oracle_f64(oracle):
    cs = small_batch(5, 24)
    say("batch f64: %d circuits" % len(cs))
    with E.Batch() as b:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.run()
        b.sync()
        for cid, c in zip(ids, cs):
            info = b.info(cid)
            assert info["done"] == 1 and info["gates_done"] == len(c["gates"])
            want, upd = oracle_of(oracle, c)
            assert_same_state(b.export(cid), want, atol=1e-12, label="circuit %d" % cid)
            assert info["updates"] == upd, "circuit %d: updates %d vs %d" % (cid, info["updates"], upd)
            assert info["nnz"] == want[0].size


def test_batch_steps_equal_full_run(oracle):
    """Advancing in steps of a few gates gives the same states as one run."""
    cs = small_batch(6, 10)
    with E.Batch() as b1, E.Batch() as b2:
        i1 = [b1.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        i2 = [b2.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b1.run()
        steps = 0
        while b2.usage()["active"]:
            b2.step(7)
            steps += 1
        say("stepped batch: %d steps" % steps)
        assert steps >= max(len(c["gates"]) for c in cs) // 7
        for a, c in zip(zip(i1, i2), cs):
            assert_same_state(b2.export(a[1]), b1.export(a[0]), atol=1e-12, label="circuit %d" % a[0])
            assert b1.info(a[0])["updates"] == b2.info(a[1])["updates"]


def test_batch_f32_renorm(oracle):
    cs = small_batch(7, 12, depth=(150, 320))
    say("batch f32 renorm: %d circuits" % len(cs))
    with E.Batch(prec="f32", renorm_every=100) as b:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.run()
        for cid, c in zip(ids, cs):
            want, upd = oracle_of(oracle, c, f32=True, renorm_every=100)
            got = b.export(cid)
            # keys equal except entries within a few ulps of the FP32 prune threshold
            gk, wk = got[0], want[0]
            diff = np.setxor1d(gk, wk)
            assert diff.size <= max(2, gk.size // 1000), "circuit %d: %d keys differ" % (cid, diff.size)
            common, ig, iw = np.intersect1d(gk, wk, return_indices=True)
            d = max(np.max(np.abs(got[1][ig] - want[1][iw])), np.max(np.abs(got[2][ig] - want[2][iw])))
            assert d <= 1e-6, "circuit %d: amplitude difference %.3e" % (cid, d)
            n2 = float(np.sum(got[1] ** 2 + got[2] ** 2))
            assert abs(n2 - 1.0) < 1e-5


def test_batch_f32_one_pass_against_step_mode(oracle):
    """One-pass renormalisation (norm checkpoints, one scale at the end): a run
    is two sweep launches per bucket size instead of one step per period, and
    the states agree with the step-by-step batch and the FP32 oracle."""
    cs = small_batch(9, 10, depth=(250, 480))
    out = {}
    for mode in (1, 0):
        with E.Batch(prec="f32", renorm_every=100, renorm_mode=mode) as b:
            ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
            sizes = {b.info(i)["bucket_bits"] for i in ids}
            l0 = b.tuning()["launches"]
            b.run()
            b.sync()
            launches = b.tuning()["launches"] - l0
            out[mode] = ([b.export(i) for i in ids], [b.info(i)["updates"] for i in ids], launches, len(sizes))
        say("renorm_mode %d: %d sweep launches, %d bucket sizes" % (mode, launches, len(sizes)))
    one, step = out[1], out[0]
    assert one[2] <= 2 * one[3] < step[2]
    for i, c in enumerate(cs):
        a, b = one[0][i], step[0][i]
        common, ia, ib = np.intersect1d(a[0], b[0], return_indices=True)
        assert common.size >= b[0].size - 2
        d = max(np.max(np.abs(a[1][ia] - b[1][ib])), np.max(np.abs(a[2][ia] - b[2][ib])))
        assert d <= 1e-6, "circuit %d: %.3e" % (i, d)
        assert one[1][i] == step[1][i]
        want, _ = oracle_of(oracle, c, f32=True, renorm_every=100)
        common, ia, iw = np.intersect1d(a[0], want[0], return_indices=True)
        d = max(np.max(np.abs(a[1][ia] - want[1][iw])), np.max(np.abs(a[2][ia] - want[2][iw])))
        assert d <= 1e-6, "circuit %d vs oracle: %.3e" % (i, d)
        assert abs(float(np.sum(a[1] ** 2 + a[2] ** 2)) - 1.0) < 1e-5


def test_batch_f32_one_pass_reports_a_drifting_norm():
    """A non-unitary gate (2 x identity) moves the checkpoint norm far beyond
    2^-11: the one-pass circuit is reported failed, its neighbour is not."""
    cs = small_batch(10, 2, depth=(120, 150))
    bad = list(cs[0]["gates"])
    u = G.make_gate(G.U, int(cs[0]["branch"][0]), matrix=[[2.0, 0.0], [0.0, 2.0]])
    bad.insert(10, u)
    with E.Batch(prec="f32", renorm_every=100) as b:
        i0 = b.add(cs[0]["nq"], cs[0]["keys"], cs[0]["re"], cs[0]["im"], G.gate_array(bad))
        i1 = b.add(cs[1]["nq"], cs[1]["keys"], cs[1]["re"], cs[1]["im"], cs[1]["gates"])
        b.run()
        b.sync()
        with pytest.raises(E.SqeError) as ex:
            b.info(i0)
        assert ex.value.status == E.SQE_EDEVICE and "one-pass" in str(ex.value)
        assert b.info(i1)["done"] == 1
        b.export(i1)


def test_batch_f32_matches_solo_engine():
    """FP32 batch and the single-circuit FP32 engine agree (both renormalise in FP64)."""
    cs = small_batch(8, 6, depth=(120, 220))
    with E.Batch(prec="f32", renorm_every=100) as b:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.run()
        for cid, c in zip(ids, cs):
            with E.Engine(c["nq"], prec="f32", renorm_every=100) as e:
                e.load(c["keys"], c["re"], c["im"])
                e.apply(c["gates"])
                solo = e.export()
            got = b.export(cid)
            common, ig, iw = np.intersect1d(got[0], solo[0], return_indices=True)
            assert common.size >= solo[0].size - 2
            d = max(np.max(np.abs(got[1][ig] - solo[1][iw])), np.max(np.abs(got[2][ig] - solo[2][iw])))
            assert d <= 1e-6


def test_admission_budget_and_reuse(oracle):
    """Over-budget circuits are refused; removing finished ones frees room."""
    cs = small_batch(9, 8, log2_nnz=(9.0, 10.0))
    with E.Batch(hbm_budget_bytes=400_000) as b:
        admitted, refused = [], []
        for c in cs:
            cid = b.try_add(c["nq"], c["keys"], c["re"], c["im"], c["gates"])
            (admitted if cid is not None else refused).append((cid, c))
        u = b.usage()
        say("admitted", len(admitted), "refused", len(refused), u)
        assert admitted and refused, "budget should split this batch"
        assert u["used"] <= u["budget"]
        b.run()
        for cid, c in admitted:
            want, _ = oracle_of(oracle, c)
            assert_same_state(b.export(cid), want, atol=1e-12)
            b.remove(cid)
        assert b.usage()["used"] == 0
        again = []
        for _, c in refused:
            cid = b.try_add(c["nq"], c["keys"], c["re"], c["im"], c["gates"])
            if cid is not None:
                again.append((cid, c))
        assert again
        b.run()
        for cid, c in again:
            want, _ = oracle_of(oracle, c)
            assert_same_state(b.export(cid), want, atol=1e-12, label="readmitted %d" % cid)


def test_autotune_keeps_results(oracle):
    cs = small_batch(10, 16)
    with E.Batch() as b:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        t = b.autotune()
        say("autotune best", t)
        assert t["block"] in (64, 128, 256) and t["tiles_per_item"] >= 1
        # the load factor is chosen among the timed candidates and kept for
        # later admissions
        assert t["load_factor"] in (0.25, 0.375, 0.5, 0.625, 0.75) and t["load_ms"] > 0
        assert b.tuning()["load_factor"] == t["load_factor"]
        b.run()
        for cid, c in zip(ids, cs):
            want, upd = oracle_of(oracle, c)
            assert_same_state(b.export(cid), want, atol=1e-12)
            assert b.info(cid)["updates"] == upd
        # a circuit admitted after tuning (its index at the tuned load factor)
        extra = small_batch(12, 2)
        eid = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in extra]
        b.run()
        for cid, c in zip(eid, extra):
            want, upd = oracle_of(oracle, c)
            assert_same_state(b.export(cid), want, atol=1e-12)
            assert b.info(cid)["updates"] == upd


@pytest.mark.parametrize("block,gran", [(64, 1), (128, 4), (256, 2)])
def test_tuning_grid(oracle, block, gran):
    cs = small_batch(11, 6)
    with E.Batch() as b:
        b.set_tuning(block=block, tiles_per_item=gran)
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.run()
        for cid, c in zip(ids, cs):
            want, _ = oracle_of(oracle, c)
            assert_same_state(b.export(cid), want, atol=1e-12)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_two_tiles_per_pass_equal_one(prec):
    """Work items of two or more tiles run 8-slot buckets two tiles per pass
    (an odd item ends with a half-empty pass): the states, update counts and,
    in FP32, the one-pass norm checkpoints (one partial per tile) are those of
    one tile per pass, to the bit."""
    cs = small_batch(13, 8, depth=(150, 300))
    runs = {}
    for gran in (1, 2, 3):
        with E.Batch(prec=prec, renorm_every=100 if prec == "f32" else 0) as b:
            b.set_tuning(tiles_per_item=gran)
            ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
            sizes = [b.info(i)["bucket_bits"] for i in ids]
            b.run()
            b.sync()
            runs[gran] = ([b.export(i) for i in ids], [b.info(i)["updates"] for i in ids])
    say("bucket bits", sizes)
    assert 3 in sizes, "the workload should have 8-slot buckets"
    for gran in (2, 3):
        for i in range(len(cs)):
            assert_same_state(runs[gran][0][i], runs[1][0][i], atol=0.0, label="gran %d circuit %d" % (gran, i))
            assert runs[gran][1][i] == runs[1][1][i]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_one_launch_equals_split(oracle, prec):
    """Circuits of 0-3 bucket qubits in one sweep launch per step (each
    wavefront branches on its circuit's bucket size) give the states and
    update counts of one launch per bucket size, to the bit, for both the
    captured run and direct steps; the FP64 run is also the oracle's."""
    cs = small_batch(14, 12)
    runs = {}
    for one in (0, 1):
        with E.Batch(prec=prec, renorm_every=100 if prec == "f32" else 0) as b:
            b.set_tuning(tiles_per_item=2, one_launch=one)
            assert b.tuning()["one_launch"] == one
            ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
            sizes = sorted({b.info(i)["bucket_bits"] for i in ids})
            b.run()
            b.sync()
            launches = b.tuning()["launches"]
            runs[one] = ([b.export(i) for i in ids], [b.info(i)["updates"] for i in ids], launches)
    say("bucket bits", sizes, "launches split / one", runs[0][2], runs[1][2])
    assert len([s for s in sizes if s <= 3]) >= 2, "the workload should mix bucket sizes"
    assert runs[1][2] < runs[0][2]
    for i in range(len(cs)):
        assert_same_state(runs[1][0][i], runs[0][0][i], atol=0.0, label="circuit %d" % i)
        assert runs[1][1][i] == runs[0][1][i]
    if prec == "f64":
        for i, c in enumerate(cs):
            want, upd = oracle_of(oracle, c)
            assert_same_state(runs[1][0][i], want, atol=1e-12, label="oracle circuit %d" % i)
            assert runs[1][1][i] == upd
    # direct steps of 64 gates (the serving path)
    stepped = {}
    for one in (0, 1):
        with E.Batch(prec=prec, renorm_every=100 if prec == "f32" else 0) as b:
            b.set_tuning(one_launch=one)
            ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
            while b.usage()["active"]:
                b.step(64)
            stepped[one] = [b.export(i) for i in ids]
    for i in range(len(cs)):
        assert_same_state(stepped[1][i], stepped[0][i], atol=0.0, label="stepped circuit %d" % i)


def test_batch_tuning_rejects_bad_one_launch():
    with E.Batch() as b:
        with pytest.raises(E.SqeError) as ex:
            b.set_tuning(one_launch=2)
        assert ex.value.status == E.SQE_EINVAL


def test_batch_rejects_cross_bucket_circuit():
    """Branching on more qubits than a bucket holds is refused with ENOTLOCAL."""
    nq = 12
    gl = G.gate_array([G.make_gate(G.H, q) for q in range(6)])
    with E.Batch() as b:
        with pytest.raises(E.SqeError) as ex:
            b.add(nq, np.zeros(1, dtype=np.uint64), np.ones(1), np.zeros(1), gl)
        assert ex.value.status == E.SQE_ENOTLOCAL
        assert b.usage()["used"] == 0


def test_batch_snapshot_restore_repeats(oracle):
    cs = small_batch(12, 8)
    with E.Batch() as b:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.snapshot_save()
        b.run()
        first = [b.export(i) for i in ids]
        upd = [b.info(i)["updates"] for i in ids]
        b.snapshot_restore()
        assert b.usage()["active"] == len(ids)
        b.run()
        for k, i in enumerate(ids):
            assert_same_state(b.export(i), first[k], atol=0.0, label="circuit %d" % i)
            assert b.info(i)["updates"] == upd[k]
        b.remove(ids[0])
        with pytest.raises(E.SqeError):
            b.snapshot_restore()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_graph_replays_equal_eager(oracle, prec):
    """run() replays one captured hipGraph of the whole gate sequence; every
    replay (after a snapshot restore, with eager work and synchronisation in
    between) must equal the directly launched run to the bit."""
    cs = small_batch(31, 10, depth=(120, 260))
    ren = 100 if prec == "f32" else 0
    with E.Batch(prec=prec, renorm_every=ren, eager=True) as be, E.Batch(prec=prec, renorm_every=ren) as bg:
        ie = [be.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        ig = [bg.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        be.snapshot_save()
        be.run()
        want = [be.export(i) for i in ie]
        upd = [be.info(i)["updates"] for i in ie]
        bg.snapshot_save()
        for rep in range(4):
            say("graph replay", prec, rep)
            if rep:
                bg.snapshot_restore()
                bg.export(ig[rep % len(ig)])          # eager work between replays
                bg.sync()
            bg.run()
            bg.sync()
            for k, i in enumerate(ig):
                assert_same_state(bg.export(i), want[k], atol=0.0, label="replay %d circuit %d" % (rep, i))
                assert bg.info(i)["updates"] == upd[k]
                assert bg.info(i)["done"] == 1
        # a run from mid-program positions, after the set changed, is planned
        # anew: the same sequence on both batches gives the same bits
        c = cs[0]
        got = []
        for b, ids in ((be, ie), (bg, ig)):
            b.snapshot_restore()
            b.step(17)
            b.remove(ids[0])
            nid = b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"])
            b.run()
            got.append([b.export(i) for i in ids[1:]] + [b.export(nid)])
        for k, (x, y) in enumerate(zip(*got)):
            assert_same_state(y, x, atol=0.0, label="mid-run circuit %d" % k)
        if prec == "f64":
            o, u = oracle_of(oracle, c)
            assert_same_state(got[1][-1], o, atol=1e-12, label="oracle")


def test_export_into_pinned_arrays():
    """Export into reused page-locked arrays (grown on demand) equals the
    plain export."""
    cs = small_batch(41, 6)
    with E.Batch() as b:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.run()
        pin = E.PinnedArrays(16)
        for i in ids:
            want = b.export(i)
            got = b.export(i, out=pin)
            assert_same_state(tuple(np.array(x) for x in got), want, atol=0.0, label="circuit %d" % i)
        assert pin.cap >= max(b.info(i)["nnz"] for i in ids)
        pin.free()


def sorted_state(keys, re, im):
    o = np.argsort(keys, kind="stable")
    return keys[o], re[o], im[o]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_async_export_equals_export(oracle, prec):
    """sqe_batch_export_start / _poll give what sqe_batch_export gives, with
    the circuit removed right after the start and new circuits admitted into
    its pool space (and stepped) before the copy is collected."""
    cs = small_batch(41, 12)
    extra = small_batch(42, 6)
    kw = dict(prec=prec, renorm_every=100) if prec == "f32" else {}
    with E.Batch(**kw) as b, E.Batch(**kw) as ref:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        rids = [ref.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        b.run()
        ref.run()
        bufs, tickets = [], []
        for cid in ids:
            buf = b.export_buffer()
            tickets.append(b.export_start(cid, buf))
            bufs.append(buf)
            b.remove(cid)
        more = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in extra]
        b.step(16)
        for k, (tk, buf) in enumerate(zip(tickets, bufs)):
            n, info = b.export_poll(tk, wait=True)
            got = sorted_state(buf.keys[:n].copy(), buf.re[:n].copy(), buf.im[:n].copy())
            want = ref.export(rids[k])
            assert_same_state(got, want, atol=0.0, label="circuit %d" % k)
            rinfo = ref.info(rids[k])
            assert info["updates"] == rinfo["updates"] and info["nnz"] == rinfo["nnz"]
            assert info["done"] == 1 and info["gates_done"] == len(cs[k]["gates"])
        b.run()
        if prec == "f64":
            for cid, c in zip(more, extra):
                want, upd = oracle_of(oracle, c)
                assert_same_state(b.export(cid), want, atol=1e-12, label="later circuit %d" % cid)
                assert b.info(cid)["updates"] == upd
        for buf in bufs:
            buf.free()


def test_async_export_argument_checks():
    c = small_batch(43, 1)[0]
    with E.Batch() as b:
        cid = b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"])
        b.run()
        L = E.lib()
        t = E.ctypes.c_int64()
        bound = E.ctypes.c_uint64()
        small = E.PinnedArrays(1024)
        rc = L.sqe_batch_export_start(b.h, cid, E._ptr(small.keys, E._u64p), E._ptr(small.re, E._f64p),
                                      E._ptr(small.im, E._f64p), 1, E.ctypes.byref(t), E.ctypes.byref(bound))
        assert rc == E.SQE_EINVAL and bound.value >= c["keys"].size
        # pageable numpy arrays are refused (the copy runs as a kernel)
        cap = bound.value
        k, r, i = np.zeros(cap, np.uint64), np.zeros(cap), np.zeros(cap)
        rc = L.sqe_batch_export_start(b.h, cid, E._ptr(k, E._u64p), E._ptr(r, E._f64p), E._ptr(i, E._f64p), cap,
                                      E.ctypes.byref(t), E.ctypes.byref(bound))
        assert rc == E.SQE_EINVAL
        assert b"page-locked" in L.sqe_batch_last_error()
        with pytest.raises(E.SqeError):
            b.export_poll(12345)
        small.free()
        assert_same_state(b.export(cid), b.export(cid), atol=0.0)


def test_duplicate_key_reported_for_its_circuit(oracle):
    """Admission does not wait for the device: a duplicate key in one
    circuit's state is reported by info / export / export_poll for that
    circuit, and the circuits around it are unaffected."""
    cs = small_batch(44, 3)
    bad = dict(cs[1])
    bad["keys"] = bad["keys"].copy()
    bad["keys"][-1] = bad["keys"][0]
    with E.Batch() as b:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in (cs[0], bad, cs[2])]
        b.run()
        for which in (0, 2):
            want, upd = oracle_of(oracle, cs[which])
            assert_same_state(b.export(ids[which]), want, atol=1e-12)
        with pytest.raises(E.SqeError) as ex:
            b.info(ids[1])
        assert ex.value.status == E.SQE_EINVAL
        with pytest.raises(E.SqeError) as ex:
            b.export(ids[1])
        assert ex.value.status == E.SQE_EINVAL
        buf = b.export_buffer()
        tk = b.export_start(ids[1], buf)
        with pytest.raises(E.SqeError) as ex:
            b.export_poll(tk, wait=True)
        assert ex.value.status == E.SQE_EINVAL
        with pytest.raises(E.SqeError):
            b.export_poll(tk)                    # the ticket was retired
        buf.free()
        b.sync()


def test_add_pinned_equals_add(oracle):
    """Admission straight from page-locked arrays gives the same states and
    counts as the copying admission, zero amplitudes included (they load as
    absent entries); a key beyond nq is reported for its circuit only."""
    cs = small_batch(14, 8)
    # give one circuit zero amplitudes
    z = dict(cs[2])
    z["re"] = z["re"].copy()
    z["im"] = z["im"].copy()
    z["re"][::3] = 0.0
    z["im"][::3] = 0.0
    cs[2] = z
    bufs = []
    with E.Batch() as a, E.Batch() as b:
        ia = [a.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        ib = []
        for c in cs:
            n = c["keys"].size
            buf = E.PinnedArrays(n)
            buf.keys[:n] = c["keys"]
            buf.re[:n] = c["re"]
            buf.im[:n] = c["im"]
            bufs.append(buf)
            ib.append(b.add_pinned(c["nq"], buf, n, c["gates"]))
        # a key with a bit beyond nq, admitted pinned
        bad = cs[0]
        n = bad["keys"].size
        bb = E.PinnedArrays(n)
        bb.keys[:n] = bad["keys"]
        bb.keys[0] = np.uint64(1) << np.uint64(bad["nq"])
        bb.re[:n] = bad["re"]
        bb.im[:n] = bad["im"]
        bufs.append(bb)
        ibad = b.add_pinned(bad["nq"], bb, n, bad["gates"])
        a.run()
        b.run()
        b.sync()
        for k, c in enumerate(cs):
            ga, gb = a.export(ia[k]), b.export(ib[k])
            assert_same_state(gb, ga, atol=0.0, label="circuit %d pinned vs copied" % k)
            assert a.info(ia[k])["updates"] == b.info(ib[k])["updates"]
            want, upd = oracle_of(oracle, c)
            assert_same_state(gb, want, atol=1e-12, label="circuit %d vs oracle" % k)
            assert b.info(ib[k])["updates"] == upd
        with pytest.raises(E.SqeError) as ex:
            b.info(ibad)
        assert ex.value.status == E.SQE_EINVAL and "beyond nq" in str(ex.value)
    for buf in bufs:
        buf.free()
