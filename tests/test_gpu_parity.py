"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bar: bit-exact key sets; FP64 amplitudes within 1e-12 absolute (gates on bucket
qubits use the oracle's operation order; consecutive factors of gates on base
qubits are multiplied together before they reach the amplitudes, a few ulps
apart from the oracle's one-gate-at-a-time order; a single gate agrees to the
bit); FP32 against the FP32 oracle within 1e-6 with key sets equal away from
the prune threshold.
"""
import math
import sys

import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gates as G
from sparqc_amd import gen
from tests.helpers import assert_same_state, random_gates, random_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_run(O, nq, keys, re, im, gates, f32=False, thr=None, renorm_every=0):
    s = O.Sparse(nq, f32=f32, thr=(1e-7 if f32 else 1e-14) if thr is None else thr,
                 renorm_every=renorm_every)
    s.load(keys, re, im)
    s.apply(gates)
    return s.export(), s


def gpu_run(nq, keys, re, im, gates, **kw):
    with E.Engine(nq, **kw) as e:
        e.load(keys, re, im)
        e.apply(gates)
        out = e.export()
        st = e.stats()
    return out, st


@pytest.mark.parametrize("bits", [0, 1, 2, 3, 4])
def test_single_gates_every_kind(oracle, bits):
    """Each gate kind, on bucket and base targets, with and without controls."""
    nq = 9
    keys, re, im = random_state(11 + bits, nq, 180)
    for kind in range(11):
        for t in (0, 4, 8):
            for ctrls in ((), (2,), (1, 6)):
                if t in ctrls:
                    continue
                g = G.gate_array([G.make_gate(kind, t, ctrls, theta=1.234,
                                              matrix=([[0.6, 0.8j], [0.8j, 0.6]] if kind == G.U else
                                                      [np.exp(0.3j), np.exp(-1.1j)]))])
                say("kind", G.KIND_NAMES[kind], "t", t, "ctrls", ctrls, "bits", bits)
                want, _ = oracle_run(oracle, nq, keys, re, im, g)
                got, _ = gpu_run(nq, keys, re, im, g, bucket_bits=bits)
                assert_same_state(got, want, atol=0.0, label="%s t=%d c=%s" % (G.KIND_NAMES[kind], t, ctrls))


def _class_cases():
    """sqe_apply's three classes: (class, 2x2 matrix, equivalent gate kind and its matrix argument)."""
    d0, d1 = np.exp(0.3j), np.exp(-1.1j)
    u01, u10 = np.exp(0.7j), -np.exp(-0.2j)
    c, s = math.cos(0.45), math.sin(0.45)
    ub = np.array([[c * np.exp(0.4j), -s * np.exp(1.3j)], [s * np.exp(-1.3j), c * np.exp(-0.4j)]])
    return [(G.CLS_DIAG, np.array([[d0, 0], [0, d1]]), G.DIAG, [d0, d1]),
            (G.CLS_PERM, np.array([[0, u01], [u10, 0]]), G.PERM, [u01, u10]),
            (G.CLS_BRANCH, ub, G.U, ub)]


@pytest.mark.parametrize("bits", [0, 2, 4])
def test_sqe_apply_each_class(oracle, bits):
    """The boundary call sqe_apply (a host matrix and a class, the reference's
    Mapping__ gate seam) against the oracle's DIAG / PERM / U records: one call
    at a time agrees to the bit, on bucket and base targets, with controls of
    either value."""
    nq = 9
    keys, re, im = random_state(41 + bits, nq, 200)
    for cls, m, kind, arg in _class_cases():
        for t in (0, 5, 8):
            for ctrls, cv in (((), None), ((2,), None), ((1, 6), 1 << 6), ((3,), 0)):
                if t in ctrls:
                    continue
                say("sqe_apply class", cls, "t", t, "ctrls", ctrls, "cv", cv, "bits", bits)
                g = G.gate_array([G.make_gate(kind, t, ctrls, matrix=arg, ctrl_val=cv)])
                want, _ = oracle_run(oracle, nq, keys, re, im, g)
                with E.Engine(nq, bucket_bits=bits) as e:
                    e.load(keys, re, im)
                    e.apply_matrix(t, m, cls, controls=ctrls, ctrl_val=cv)
                    got = e.export()
                assert_same_state(got, want, atol=0.0, label="class %d t=%d c=%s" % (cls, t, ctrls))


def test_sqe_apply_sequence(oracle):
    """A run of sqe_apply calls (one planned program) against the oracle."""
    nq = 12
    keys, re, im = random_state(43, nq, 300)
    rng = gen.SplitMix64(44)
    cases = _class_cases()
    recs = []
    with E.Engine(nq) as e:
        e.load(keys, re, im)
        for _ in range(150):
            cls, m, kind, arg = cases[rng.randint(3)]
            t = rng.randint(nq)
            c = (t + 1 + rng.randint(nq - 1)) % nq
            ctrls = (c,) if rng.uniform() < 0.5 else ()
            e.apply_matrix(t, m, cls, controls=ctrls)
            recs.append(G.make_gate(kind, t, ctrls, matrix=arg))
        got = e.export()
        st = e.stats()
    want, orc = oracle_run(oracle, nq, keys, re, im, G.gate_array(recs))
    assert_same_state(got, want, atol=1e-12, label="sqe_apply sequence")
    assert st["updates"] == orc.updates()


@pytest.mark.parametrize("bits,window", [(-1, 0), (0, 0), (1, 3), (2, 1), (3, 0), (4, 7)])
def test_random_circuits(oracle, bits, window):
    """Random mixed circuits with random controls (cross-bucket and local paths)."""
    nq = 12
    for seed in (1, 2):
        keys, re, im = random_state(100 + seed, nq, 64)
        gl = random_gates(200 + seed, nq, 300)
        say("random circuit seed", seed, "bits", bits, "window", window)
        want, orc = oracle_run(oracle, nq, keys, re, im, gl)
        got, st = gpu_run(nq, keys, re, im, gl, bucket_bits=bits, fuse_window=window)
        assert_same_state(got, want, atol=1e-12, label="seed %d" % seed)
        assert st["updates"] == orc.updates()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_small_workload_vs_oracle_and_dense(oracle, seed):
    """20-qubit, 200-gate workload: oracle (bit-exact) and dense FP64 statevector."""
    w = gen.circuit_small(seed)
    say("circuit_small seed", seed)
    want, _ = oracle_run(oracle, w["nq"], w["keys"], w["re"], w["im"], w["gates"])
    got, st = gpu_run(w["nq"], w["keys"], w["re"], w["im"], w["gates"])
    assert_same_state(got, want, atol=1e-12, label="circuit_small %d" % seed)
    dense = oracle.run_dense(w["nq"], w["keys"], w["re"], w["im"], w["gates"])
    sp = np.zeros(1 << w["nq"], dtype=complex)
    sp[got[0].astype(np.int64)] = got[1] + 1j * got[2]
    assert np.max(np.abs(dense - sp)) <= 1e-12


def test_large_scaled_24q_vs_dense(oracle):
    """The 48-qubit workload's gate mix scaled to 24 qubits (dense check possible)."""
    w = gen.circuit_large(7, nq=24, n_init=1 << 12, n_gates=2000)
    say("circuit_large nq=24")
    want, _ = oracle_run(oracle, 24, w["keys"], w["re"], w["im"], w["gates"])
    got, st = gpu_run(24, w["keys"], w["re"], w["im"], w["gates"])
    assert_same_state(got, want, atol=1e-12, label="large-24q")
    dense = oracle.run_dense(24, w["keys"], w["re"], w["im"], w["gates"])
    sp = np.zeros(1 << 24, dtype=complex)
    sp[got[0].astype(np.int64)] = got[1] + 1j * got[2]
    assert np.max(np.abs(dense - sp)) <= 1e-12
    assert st["pairs"] == 0          # every gate ran inside fused sweeps


def test_large_48q_medium(oracle):
    w = gen.circuit_large(3, nq=48, n_init=1 << 14, n_gates=600)
    say("circuit_large nq=48 n_init=2^14")
    want, orc = oracle_run(oracle, 48, w["keys"], w["re"], w["im"], w["gates"])
    got, st = gpu_run(48, w["keys"], w["re"], w["im"], w["gates"])
    assert_same_state(got, want, atol=1e-12, label="large-48q")


def test_fp32_against_fp32_oracle_and_fp64(oracle):
    w = gen.circuit_large(5, nq=48, n_init=1 << 12, n_gates=1000)
    say("fp32 nq=48")
    (wk, wr, wi), _ = oracle_run(oracle, 48, w["keys"], w["re"], w["im"], w["gates"], f32=True,
                                 renorm_every=100)
    gk, gr, gi = gpu_run(48, w["keys"], w["re"], w["im"], w["gates"], prec="f32", renorm_every=100)[0]
    common, ia, ib = np.intersect1d(gk, wk, return_indices=True)
    only = np.setxor1d(gk, wk)
    assert only.size <= max(2, gk.size // 10000)
    assert np.max(np.abs(gr[ia] - wr[ib])) <= 1e-6
    assert np.max(np.abs(gi[ia] - wi[ib])) <= 1e-6
    (dk, dr, di), _ = oracle_run(oracle, 48, w["keys"], w["re"], w["im"], w["gates"])
    common, ia, ib = np.intersect1d(gk, dk, return_indices=True)
    ov = np.sum((dr[ib] - 1j * di[ib]) * (gr[ia] + 1j * gi[ia]))
    infid = 1.0 - abs(ov) ** 2 / (np.sum(dr ** 2 + di ** 2) * np.sum(gr ** 2 + gi ** 2))
    say("fp32 infidelity", infid)
    assert infid <= 1e-6


def test_all_h_closed_form():
    """H on every qubit of |0...0> gives 2^(-n/2) on every key (emulator/src/main.cu:142-143)."""
    nq = 16
    w = gen.circuit_all_h(nq)
    k, r, i = gpu_run(nq, w["keys"], w["re"], w["im"], w["gates"])[0]
    assert k.size == 1 << nq
    assert np.array_equal(k, np.arange(1 << nq, dtype=np.uint64))
    assert np.max(np.abs(r - 2.0 ** (-nq / 2))) <= 1e-15 and np.max(np.abs(i)) == 0.0


def test_inverse_circuit_returns_initial_state():
    w = gen.circuit_large(9, nq=48, n_init=1 << 13, n_gates=800)
    inv = G.gate_array([G.inverse_gate(g) for g in w["gates"][::-1]])
    with E.Engine(48) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.apply(w["gates"])
        e.apply(inv)
        k, r, i = e.export()
    o = np.argsort(w["keys"])
    assert np.array_equal(k, w["keys"][o])
    assert np.max(np.abs(r - w["re"][o])) <= 1e-12 and np.max(np.abs(i - w["im"][o])) <= 1e-12


def test_snapshot_restore_repeats_bit_exactly():
    w = gen.circuit_large(4, nq=48, n_init=1 << 12, n_gates=400)
    with E.Engine(48) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.prepare(w["gates"])
        e.snapshot_save()
        e.apply(w["gates"])
        a = e.export()
        e.snapshot_restore()
        e.apply(w["gates"])
        b = e.export()
    assert_same_state(b, a, atol=0.0)


def test_errors_are_reported():
    with E.Engine(10) as e:
        with pytest.raises(E.SqeError) as ex:
            e.apply(G.gate_array([G.make_gate(G.H, 12)]))
        assert ex.value.status == E.SQE_EINVAL
        with pytest.raises(E.SqeError) as ex:
            e.apply(G.gate_array([G.make_gate(G.H, 3)]))
            e.sync()
        assert ex.value.status == E.SQE_EINVAL          # no state loaded
        e.load(np.array([3, 5, 3], dtype=np.uint64), np.ones(3) / 3 ** 0.5, np.zeros(3))
        e.apply(G.gate_array([G.make_gate(G.H, 1)]))
        with pytest.raises(E.SqeError) as ex:
            e.sync()
        assert ex.value.status == E.SQE_EINVAL          # duplicate key
    e = E.Engine(20, hbm_budget_bytes=1 << 16)
    e.load(np.arange(1 << 14, dtype=np.uint64), np.ones(1 << 14) / 128.0, np.zeros(1 << 14))
    assert e.nnz() == 1 << 14                           # staged on the host, nothing allocated yet
    with pytest.raises(E.SqeError) as ex:
        e.apply(G.gate_array([G.make_gate(G.H, 1)]))
        e.sync()
    assert ex.value.status == E.SQE_ETABLE_FULL
    e.close()


def test_export_where_and_checkpoint(oracle, tmp_path):
    """The device-side pattern filter equals filtering the full export; a
    restart file restores the state exactly and the engine continues from it."""
    nq = 14
    keys, re, im = random_state(77, nq, 3000)
    gl = random_gates(78, nq, 60, p_ctrl=0.3)
    with E.Engine(nq) as e:
        e.load(keys, re, im)
        e.apply(gl)
        full = e.export()
        for mask, value in ((0, 0), (1 << 3, 1 << 3), ((1 << 13) | 0b101, 0b100), ((1 << nq) - 1, int(full[0][0]))):
            sel = (full[0] & np.uint64(mask)) == np.uint64(value)
            got = e.export(where=(mask, value))
            assert_same_state(got, tuple(x[sel] for x in full), atol=0.0, label="pattern %x/%x" % (mask, value))
        with pytest.raises(E.SqeError):
            e.export(where=(1, 2))
        path = str(tmp_path / "state.bin")
        e.checkpoint_save(path)
        more = random_gates(79, nq, 20)
        e.apply(more)
        after = e.export()
    with E.Engine(nq) as r:
        r.checkpoint_load(path)
        assert_same_state(r.export(), full, atol=0.0, label="restored")
        assert r.stats()["gates"] == len(gl)
        r.apply(more)
        assert_same_state(r.export(), after, atol=1e-12, label="continued")
    with E.Engine(nq + 1) as bad:
        with pytest.raises(E.SqeError):
            bad.checkpoint_load(path)


def _sample_restated(keys, re, im, seed, shots):
    """The device sampler's arithmetic in numpy (same order of operations):
    tiles of 4096 entries, 16 per thread summed in order, thread offsets and
    tile offsets by sequential scans, a counter-based splitmix64 uniform per
    shot, then the first entry whose cumulative |a|^2 exceeds u * total."""
    n = keys.size
    p = re * re + im * im
    per_thread, tile = 16, 4096
    nt = (n + tile - 1) // tile
    local = np.empty(n)
    tsum = np.empty(nt)
    for t in range(nt):
        seg = p[t * tile:(t + 1) * tile]
        m = seg.size
        pad = np.zeros(tile)
        pad[:m] = seg
        run = np.cumsum(pad.reshape(256, per_thread), axis=1)
        sums = run[:, -1]
        off = np.concatenate([[0.0], np.cumsum(sums)[:-1]])
        loc = (off[:, None] + run).reshape(-1)[:m]
        local[t * tile:t * tile + m] = loc
        tsum[t] = loc[-1]
    offs = np.concatenate([[0.0], np.cumsum(tsum)[:-1]])
    total = np.cumsum(tsum)[-1]
    u = (gen.SplitMix64(seed).u64_array(shots) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    target = u * total
    out = np.empty(shots, dtype=np.uint64)
    for k in range(shots):
        t = int(np.searchsorted(offs, target[k], side="right")) - 1
        a, b = t * tile, min(n, (t + 1) * tile)
        j = a + int(np.searchsorted(offs[t] + local[a:b], target[k], side="right"))
        out[k] = keys[min(j, b - 1)]
    return out


def test_sample_measurement():
    """Sampling: equal to the restated arithmetic draw for draw, frequencies
    within 5 sigma of |a|^2, deterministic per seed."""
    # three entries with probabilities 0.5 / 0.3 / 0.2
    keys = np.array([3, 9, 12], dtype=np.uint64)
    amp = np.sqrt(np.array([0.5, 0.3, 0.2]))
    with E.Engine(6) as e:
        e.load(keys, amp * np.cos([0.3, 1.0, 2.0]), amp * np.sin([0.3, 1.0, 2.0]))
        shots = 200000
        s = e.sample(shots, seed=5)
        for key, p in zip(keys, [0.5, 0.3, 0.2]):
            f = np.mean(s == key)
            assert abs(f - p) < 5 * math.sqrt(p * (1 - p) / shots), (key, f, p)
        assert np.array_equal(s, e.sample(shots, seed=5))
        assert not np.array_equal(s[:1000], e.sample(1000, seed=6))
        k, r, i = e.export(sort=False)
        assert np.array_equal(s[:5000], _sample_restated(k, r, i, 5, 5000))
    # a 14-qubit state of ~12000 entries (three tiles) after a circuit
    nq = 14
    k0, r0, i0 = random_state(91, nq, 3000)
    gl = random_gates(92, nq, 40, kinds=[G.H, G.RY, G.X, G.RZ], p_ctrl=0.2)
    with E.Engine(nq) as e:
        e.load(k0, r0, i0)
        e.apply(gl)
        k, r, i = e.export(sort=False)
        assert k.size > 4096
        shots = 20000
        s = e.sample(shots, seed=123)
        assert np.array_equal(s, _sample_restated(k, r, i, 123, shots))
        p = (r * r + i * i) / np.sum(r * r + i * i)
        for q in (0, 5, 13):
            want = float(np.sum(p[(k >> np.uint64(q)) & np.uint64(1) == 1]))
            got = float(np.mean((s >> np.uint64(q)) & np.uint64(1)))
            assert abs(got - want) < 5 * math.sqrt(max(want * (1 - want), 1e-6) / shots), (q, got, want)


def test_edge_cases(oracle):
    """Empty and all-negligible states, zero amplitudes, the top qubit of a
    63-qubit key, keys outside the register."""
    # empty load: gates run on nothing, export is empty
    with E.Engine(12) as e:
        e.load(np.zeros(0, dtype=np.uint64), np.zeros(0), np.zeros(0))
        e.apply(G.gate_array([G.make_gate(G.H, 3), G.make_gate(G.X, 5, controls=(3,))]))
        k, r, i = e.export()
        assert k.size == 0 and e.nnz() == 0
    # zero amplitudes are not entries; entries below the threshold are pruned
    # by the first branching gate that touches them and never exported
    with E.Engine(12) as e:
        keys = np.array([1, 2, 4, 8], dtype=np.uint64)
        e.load(keys, np.array([0.0, 1e-15, 0.6, 0.8]), np.array([0.0, 0.0, 0.0, 0.0]))
        k, _, _ = e.export()
        assert list(k) == [4, 8]
        e.apply(G.gate_array([G.make_gate(G.H, 0), G.make_gate(G.H, 0)]))
        got = e.export()
        o = oracle.Sparse(12)
        o.load(keys, np.array([0.0, 1e-15, 0.6, 0.8]), np.zeros(4))
        o.apply(G.gate_array([G.make_gate(G.H, 0), G.make_gate(G.H, 0)]))
        assert_same_state(got, o.export(), atol=1e-12, label="negligible entries")
    # 63 qubits: gates on and controlled by the top qubit, against the oracle
    nq = 63
    top = 1 << 62
    keys = np.array([0, 5, top, top | 5, top | (1 << 40), (1 << 61) | 3], dtype=np.uint64)
    amp = np.linspace(0.2, 0.7, keys.size)
    amp = amp / np.linalg.norm(amp)
    gl = G.gate_array([G.make_gate(G.H, 62), G.make_gate(G.X, 61, controls=(62,)), G.make_gate(G.RY, 0, theta=0.4),
                       G.make_gate(G.Z, 62, controls=(0,)), G.make_gate(G.X, 62), G.make_gate(G.H, 40)])
    with E.Engine(nq) as e:
        e.load(keys, amp, np.zeros(keys.size))
        e.apply(gl)
        got = e.export()
    o = oracle.Sparse(nq)
    o.load(keys, amp, np.zeros(keys.size))
    o.apply(gl)
    assert_same_state(got, o.export(), atol=1e-12, label="63 qubits")
    # a key outside the register is rejected
    with E.Engine(10) as e:
        with pytest.raises(E.SqeError) as ex:
            e.load(np.array([1 << 10], dtype=np.uint64), np.ones(1), np.zeros(1))
        assert ex.value.status == E.SQE_EINVAL


@pytest.mark.parametrize("bits", [2, 4])
def test_frames_on_and_off_agree(oracle, bits):
    """Pauli gates on bucket qubits through the per-bucket frame against the
    same gates through the amplitude registers: both are exact, so the
    states differ only where the pending factor, which then carries the
    frame's signs, is multiplied in with a fused multiply-add (ulps)."""
    nq = 14
    keys, re, im = random_state(61 + bits, nq, 400)
    gl = random_gates(62 + bits, nq, 500, kinds=[G.X, G.Y, G.Z, G.H, G.RY, G.RZ, G.PHASE, G.U], p_ctrl=0.6)
    for prec in ("f64", "f32"):
        outs = []
        for frames in (True, False):
            with E.Engine(nq, prec=prec, bucket_bits=bits, frames=frames) as e:
                e.load(keys, re, im)
                e.apply(gl)
                outs.append((e.export(), e.stats()["updates"]))
        (a, ua), (b, ub) = outs
        say("frames on/off", prec, "bits", bits, "entries", a[0].size)
        assert_same_state(a, b, atol=1e-12 if prec == "f64" else 1e-6, label="frames on/off %s" % prec)
        assert ua == ub
        if prec == "f64":
            want, _ = oracle_run(oracle, nq, keys, re, im, gl)
            assert_same_state(a, want, atol=1e-12, label="frames bits %d" % bits)
