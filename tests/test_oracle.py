"""Pinning of the CPU oracle (no GPU).

The reference ships no tests or golden vectors, so the oracle is pinned by what
the mathematics fixes: the reference's own gate matrices
(emulator/include/QLib.h:61-154) applied densely with numpy, textbook gate
identities, an independent dense statevector, exact integer closed forms for
the key transforms, the all-H scenario of emulator/src/main.cu:142-143,
inverse circuits, invariants, and the componentwise prune predicate of
emulator/include/data_t.h:50-60 at the threshold.
"""
import math

import numpy as np
import pytest

from sparqc_amd import gates as G
from sparqc_amd import gen
from tests.helpers import assert_same_state, random_gates, random_state

THR = 1e-14

# the reference's matrices, transcribed as numbers (QLib.h:61-154)
INV_SQRT2 = 1.0 / math.sqrt(2)
M_X = np.array([[0, 1], [1, 0]], dtype=complex)                          # QLib.h:64-71
M_Z = np.array([[1, 0], [0, -1]], dtype=complex)                         # QLib.h:74-81
M_H = np.array([[INV_SQRT2, INV_SQRT2], [INV_SQRT2, -INV_SQRT2]], dtype=complex)  # QLib.h:84-91


def m_ry(a):                                                             # QLib.h:93-104
    c, s = math.cos(a / 2), math.sin(a / 2)
    return np.array([[c, -s], [s, c]], dtype=complex)


def m_rz(a):                                                             # QLib.h:106-120
    c, s = math.cos(a / 2), math.sin(a / 2)
    return np.array([[c - 1j * s, 0], [0, c + 1j * s]])


def m_phase(a):                                                          # QLib.h:141-154
    return np.array([[1, 0], [0, math.cos(a) + 1j * math.sin(a)]])


def m_rx(a):   # textbook (the reference has no mRx; its rx_core, device_data.cuh:266-287)
    c, s = math.cos(a / 2), math.sin(a / 2)
    return np.array([[c, -1j * s], [-1j * s, c]])


M_Y = np.array([[0, -1j], [1j, 0]])


def run(oracle, nq, keys, re, im, gates, **kw):
    s = oracle.Sparse(nq, **kw)
    s.load(keys, re, im)
    s.apply(gates)
    return s


def as_dense(nq, k, r, i):
    v = np.zeros(1 << nq, dtype=complex)
    v[np.asarray(k, dtype=np.int64)] = r + 1j * i
    return v


def apply_dense_numpy(v, nq, g):
    """Independent numpy application of a (controlled) 2x2 gate to a dense vector."""
    m = G.matrix_of(g)
    t = int(g["target"])
    cm, cv = int(g["ctrl_mask"]), int(g["ctrl_val"])
    out = v.copy()
    idx = np.arange(1 << nq, dtype=np.int64)
    lo = idx[((idx >> t) & 1) == 0]
    sel = lo[(lo & cm) == cv]
    hi = sel | (1 << t)
    a0, a1 = v[sel], v[hi]
    out[sel] = m[0, 0] * a0 + m[0, 1] * a1
    out[hi] = m[1, 0] * a0 + m[1, 1] * a1
    return out


ANGLES = [0.0, 0.3, 1.234, math.pi / 2, math.pi, 4.0, 2 * math.pi - 0.01]


@pytest.mark.parametrize("theta", ANGLES)
def test_gate_records_match_reference_matrices(theta):
    assert np.allclose(G.matrix_of(G.make_gate(G.X, 0)), M_X, atol=0)
    assert np.allclose(G.matrix_of(G.make_gate(G.Z, 0)), M_Z, atol=0)
    assert np.array_equal(G.matrix_of(G.make_gate(G.H, 0)), M_H)
    assert np.allclose(G.matrix_of(G.make_gate(G.Y, 0)), M_Y, atol=0)
    assert np.allclose(G.matrix_of(G.make_gate(G.RY, 0, theta=theta)), m_ry(theta), atol=1e-16)
    assert np.allclose(G.matrix_of(G.make_gate(G.RZ, 0, theta=theta)), m_rz(theta), atol=1e-16)
    assert np.allclose(G.matrix_of(G.make_gate(G.PHASE, 0, theta=theta)), m_phase(theta), atol=1e-16)
    assert np.allclose(G.matrix_of(G.make_gate(G.RX, 0, theta=theta)), m_rx(theta), atol=1e-16)


@pytest.mark.parametrize("kind", range(11))
def test_each_gate_against_dense_numpy(oracle, kind):
    """Oracle (sparse, reference formulas) vs numpy matrix products, with controls."""
    nq = 5
    k = np.arange(1 << nq, dtype=np.uint64)
    rng = gen.SplitMix64(kind + 1)
    re, im = rng.normal_complex_array(1 << nq)
    for t in range(nq):
        for ctrls, cv in [((), None), (((t + 1) % nq,), None),
                          (((t + 2) % nq, (t + 3) % nq), 1 << ((t + 3) % nq))]:
            g = G.make_gate(kind, t, ctrls, theta=1.1 + t,
                            matrix=([[0.6, 0.8j], [0.8j, 0.6]] if kind == G.U else
                                    [np.exp(0.7j), np.exp(-0.2j)]), ctrl_val=cv)
            s = run(oracle, nq, k, re, im, G.gate_array([g]))
            kk, rr, ii = s.export()
            got = as_dense(nq, kk, rr, ii)
            want = apply_dense_numpy(re + 1j * im, nq, g)
            assert np.max(np.abs(got - want)) <= 1e-15


def test_identities(oracle):
    nq = 3
    keys, re, im = random_state(5, nq, 8)
    v = re + 1j * im
    o = np.argsort(keys)

    def dense_of(gl):
        s = run(oracle, nq, keys, re, im, G.gate_array(gl))
        return as_dense(nq, *s.export())

    v0 = as_dense(nq, keys, re, im)
    h, x, z = G.make_gate(G.H, 1), G.make_gate(G.X, 1), G.make_gate(G.Z, 1)
    assert np.max(np.abs(dense_of([h, h]) - v0)) < 1e-15                       # HH = I
    assert np.max(np.abs(dense_of([x, x]) - v0)) == 0.0                        # XX = I
    assert np.max(np.abs(dense_of([h, z, h]) - dense_of([x]))) < 1e-15         # HZH = X
    th = 0.77
    rz, ph = G.make_gate(G.RZ, 2, theta=th), G.make_gate(G.PHASE, 2, theta=th)
    assert np.max(np.abs(dense_of([rz]) - np.exp(-0.5j * th) * dense_of([ph]))) < 1e-15
    rx = G.make_gate(G.RX, 0, theta=th)
    h0, rz0 = G.make_gate(G.H, 0), G.make_gate(G.RZ, 0, theta=th)
    assert np.max(np.abs(dense_of([rx]) - dense_of([h0, rz0, h0]))) < 1e-15    # Rx = H Rz H
    s = run(oracle, 1, np.zeros(1, np.uint64), np.ones(1), np.zeros(1),
            G.gate_array([G.make_gate(G.RY, 0, theta=math.pi)]))
    k, r, i = s.export()                                                        # Ry(pi)|0> = |1>
    assert list(k) == [1] and abs(r[0] - 1) < 1e-15
    del v, o


@pytest.mark.parametrize("nq", [4, 6])
def test_key_transforms_brute_force(oracle, nq):
    """X, CX, CCX on every basis key: exact integer closed forms."""
    keys = np.arange(1 << nq, dtype=np.uint64)
    amps = np.arange(1, (1 << nq) + 1, dtype=np.float64)
    for t in range(nq):
        for c in range(nq):
            for c2 in range(nq):
                if len({t, c, c2}) < 3 and not (c == c2 and c != t):
                    continue
                ctrls = (c,) if c == c2 else (c, c2)
                g = G.make_gate(G.X, t, ctrls)
                s = run(oracle, nq, keys, amps, np.zeros_like(amps), G.gate_array([g]))
                k, r, _ = s.export()
                cm = sum(1 << q for q in ctrls)
                for key in range(1 << nq):
                    dst = key ^ ((1 << t) if (key & cm) == cm else 0)
                    j = np.searchsorted(k, dst)
                    assert r[j] == key + 1


def test_all_h_scenario(oracle):
    """emulator/src/main.cu:142-143: H on all qubits of |0...0>."""
    nq = 12
    w = gen.circuit_all_h(nq)
    s = run(oracle, nq, w["keys"], w["re"], w["im"], w["gates"])
    k, r, i = s.export()
    assert np.array_equal(k, np.arange(1 << nq, dtype=np.uint64))
    assert np.max(np.abs(r - 2.0 ** (-nq / 2))) < 1e-15 and np.all(i == 0)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_small_workload_sparse_equals_dense(oracle, seed):
    w = gen.circuit_small(seed)
    k, r, i = run(oracle, w["nq"], w["keys"], w["re"], w["im"], w["gates"]).export()
    d = oracle.run_dense(w["nq"], w["keys"], w["re"], w["im"], w["gates"])
    assert np.max(np.abs(as_dense(w["nq"], k, r, i) - d)) <= 1e-12
    keep = (np.abs(d.real) >= THR) | (np.abs(d.imag) >= THR)
    assert np.array_equal(k, np.nonzero(keep)[0].astype(np.uint64))


def test_dense_against_numpy_reference(oracle):
    """The C dense statevector against per-gate numpy products (independent code)."""
    nq = 8
    keys, re, im = random_state(3, nq, 40)
    gl = random_gates(4, nq, 120)
    v = as_dense(nq, keys, re, im)
    for g in gl:
        v = apply_dense_numpy(v, nq, g)
    d = oracle.run_dense(nq, keys, re, im, gl)
    assert np.max(np.abs(d - v)) <= 1e-13


def test_large_mix_scaled_sparse_equals_dense(oracle):
    w = gen.circuit_large(2, nq=16, n_init=1 << 8, n_gates=600)
    k, r, i = run(oracle, 16, w["keys"], w["re"], w["im"], w["gates"]).export()
    d = oracle.run_dense(16, w["keys"], w["re"], w["im"], w["gates"])
    assert np.max(np.abs(as_dense(16, k, r, i) - d)) <= 1e-12


def test_random_circuits_sparse_equals_dense(oracle):
    for seed in range(3):
        nq = 10
        keys, re, im = random_state(50 + seed, nq, 30)
        gl = random_gates(60 + seed, nq, 250)
        k, r, i = run(oracle, nq, keys, re, im, gl).export()
        d = oracle.run_dense(nq, keys, re, im, gl)
        assert np.max(np.abs(as_dense(nq, k, r, i) - d)) <= 1e-12


def test_inverse_circuit(oracle):
    w = gen.circuit_large(6, nq=40, n_init=1 << 10, n_gates=400)
    inv = G.gate_array([G.inverse_gate(g) for g in w["gates"][::-1]])
    s = run(oracle, 40, w["keys"], w["re"], w["im"], w["gates"])
    s.apply(inv)
    k, r, i = s.export()
    o = np.argsort(w["keys"])
    assert np.array_equal(k, w["keys"][o])
    assert np.max(np.abs(r - w["re"][o])) < 1e-12 and np.max(np.abs(i - w["im"][o])) < 1e-12


def test_invariants(oracle):
    nq = 30
    keys, re, im = random_state(8, nq, 500)
    diag = random_gates(9, nq, 200, kinds=[G.Z, G.PHASE, G.RZ, G.DIAG])
    s = run(oracle, nq, keys, re, im, diag)
    k, r, i = s.export()
    o = np.argsort(keys)
    assert np.array_equal(k, keys[o])                                           # key set kept
    assert np.max(np.abs(np.hypot(r, i) - np.hypot(re, im)[o])) < 1e-15        # |a| kept
    perm = random_gates(10, nq, 200, kinds=[G.X])
    s = run(oracle, nq, keys, re, im, perm)
    k, r, i = s.export()
    assert np.array_equal(np.sort(r), np.sort(re))                              # amplitudes permuted
    assert s.nnz() == keys.size
    k12, r12, i12 = random_state(12, 12, 64)                                   # <= 2^12 entries
    br = random_gates(11, 12, 100, kinds=[G.H, G.RX, G.RY, G.U])
    s = run(oracle, 12, k12, r12, i12, br)
    assert abs(s.norm2() - 1.0) < 1e-12                                         # unitary


def test_prune_predicate_is_componentwise_and_on_branching_only(oracle):
    f = 1.0 / math.sqrt(2.0)
    # H on a lone element gives f*v on both keys of the pair
    cases = [((0.9e-14 / f, 0.9e-14 / f), False),     # both components below: pruned
             ((0.9e-14 / f, 1.1e-14 / f), True),      # one component above: kept
             ((1.1e-14 / f, 0.0), True),
             ((0.75e-14 / f, 0.75e-14 / f), False)]   # |a| ~ 1.06e-14 > thr, still pruned
    for (a, b), keep in cases:
        s = run(oracle, 2, np.array([0], np.uint64), np.array([a]), np.array([b]),
                G.gate_array([G.make_gate(G.H, 0)]))
        assert (s.nnz() == 2) == keep
    # diagonal / permutation gates never prune (update_ah_diag/swap, device_data.cuh:108-131)
    keys = np.array([0, 2], np.uint64)
    s = run(oracle, 3, keys, np.array([1e-16, 1.0]), np.array([0.0, 0.0]),
            G.gate_array([G.make_gate(G.RZ, 0, theta=0.4), G.make_gate(G.X, 2)]))
    assert s.nnz() == 2 and s.export()[0].size == 1        # present, dropped only at readout
    s.apply(G.gate_array([G.make_gate(G.H, 1)]))           # a branching gate prunes it
    assert s.nnz() == 2                                    # 1.0 splits into two, 1e-16 is gone
    # exact cancellation: H H on a fresh pair returns to one key
    s = run(oracle, 2, np.array([0], np.uint64), np.array([1.0]), np.array([0.0]),
            G.gate_array([G.make_gate(G.H, 1), G.make_gate(G.H, 1)]))
    assert s.nnz() == 1


def test_updates_count_is_sum_of_nonzeros_entering_each_gate(oracle):
    s = run(oracle, 3, np.array([0], np.uint64), np.array([1.0]), np.array([0.0]),
            G.gate_array([G.make_gate(G.H, 0), G.make_gate(G.H, 1), G.make_gate(G.Z, 2),
                          G.make_gate(G.H, 1)]))
    assert s.updates() == 1 + 2 + 4 + 4
    assert s.nnz() == 2


def test_fp32_form_close_to_fp64(oracle):
    w = gen.circuit_large(3, nq=40, n_init=1 << 10, n_gates=600)
    d = run(oracle, 40, w["keys"], w["re"], w["im"], w["gates"])
    f = run(oracle, 40, w["keys"], w["re"], w["im"], w["gates"], f32=True, thr=1e-7, renorm_every=100)
    dk, dr, di = d.export()
    fk, fr, fi = f.export()
    c, ia, ib = np.intersect1d(dk, fk, return_indices=True)
    big = np.hypot(dr, di) > 1e-6
    assert np.all(np.isin(dk[big], fk))
    ov = np.sum((dr[ia] - 1j * di[ia]) * (fr[ib] + 1j * fi[ib]))
    infid = 1 - abs(ov) ** 2 / (d.norm2() * f.norm2())
    assert infid < 1e-10
    assert abs(f.norm2() - 1.0) < 1e-6


def test_oracle_threads_do_not_change_results(oracle):
    w = gen.circuit_small(2)
    oracle.set_threads(1)
    a = run(oracle, w["nq"], w["keys"], w["re"], w["im"], w["gates"]).export()
    oracle.set_threads(4)
    b = run(oracle, w["nq"], w["keys"], w["re"], w["im"], w["gates"]).export()
    assert_same_state(a, b, atol=0.0)


def test_fc_engine_matches_map_engine(oracle):
    """The FC restatement (dense slots, address book, two buffers) used as the
    20-qubit CPU baseline gives the map engine's state and update count on the
    20-qubit workload and on random controlled circuits."""
    from sparqc_amd import gen
    from tests.helpers import random_gates, random_state
    for seed in (1, 2, 3):
        w = gen.circuit_small(seed)
        f = oracle.FC(20)
        f.load(w["keys"], w["re"], w["im"])
        f.apply(w["gates"])
        s = oracle.Sparse(20)
        s.load(w["keys"], w["re"], w["im"])
        s.apply(w["gates"])
        a, b = f.export(), s.export()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
        assert f.updates() == s.updates()
    keys, re, im = random_state(7, 10, 60)
    gl = random_gates(8, 10, 400)
    f = oracle.FC(10)
    f.load(keys, re, im)
    f.apply(gl)
    s = oracle.Sparse(10)
    s.load(keys, re, im)
    s.apply(gl)
    a, b = f.export(), s.export()
    assert np.array_equal(a[0], b[0]) and np.max(np.abs(a[1] - b[1])) == 0.0
    assert f.updates() == s.updates()
