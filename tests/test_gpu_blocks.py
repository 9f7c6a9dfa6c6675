"""FP32 fused-block variant (windows of bucket-qubit gates applied as dense
16x16 blocks for 4 bucket qubits, 32x32 for 5, with MFMA) against the FP32
oracle and the FP64 engine.

A block prunes once, on its outputs, where gate-by-gate execution prunes after
every branching gate; both prune at 1e-7 componentwise with FP64 decisions near
the threshold, so key sets agree away from the threshold and amplitudes within
1e-5 (FP32 accumulation over the block).
"""
import sys

import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gates as G
from sparqc_amd import gen
from tests.helpers import random_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dense_block_circuit(seed, nq=16, layers=30, bq=(1, 4, 7, 10)):
    """Layers of H / RY / RZ / PHASE / X on the bucket qubits bq, each layer
    followed by diagonal, X and controlled gates elsewhere (long fusable
    windows)."""
    rng = gen.SplitMix64(seed)
    bq = list(bq)
    others = [q for q in range(nq) if q not in bq]
    gl = []
    for _ in range(layers):
        for q in bq:
            k = [G.H, G.RY, G.RZ, G.PHASE, G.X, G.RX][rng.randint(6)]
            gl.append(G.make_gate(k, q, theta=rng.uniform() * 6.28))
            if rng.uniform() < 0.5:
                t = others[rng.randint(len(others))]
                gl.append(G.make_gate(G.RZ, t, theta=rng.uniform() * 6.28))
        c, t = others[rng.randint(len(others))], others[rng.randint(len(others))]
        if c != t:
            gl.append(G.make_gate(G.X, t, controls=(c,)))
        gl.append(G.make_gate(G.PHASE, others[rng.randint(len(others))], theta=rng.uniform() * 6.28))
    return G.gate_array(gl)


def compare(got, want, atol, max_key_diff):
    gk, gr, gi = got
    wk, wr, wi = want
    diff = np.setxor1d(gk, wk)
    assert diff.size <= max_key_diff, "%d keys differ" % diff.size
    common, ia, ib = np.intersect1d(gk, wk, return_indices=True)
    d = max(np.max(np.abs(gr[ia] - wr[ib])), np.max(np.abs(gi[ia] - wi[ib])))
    assert d <= atol, "amplitude difference %.3e" % d
    return d


@pytest.mark.parametrize("seed", [1, 2])
def test_blocks_match_fp32_oracle_and_fp64(oracle, seed):
    nq = 16
    keys, re, im = random_state(300 + seed, nq, 300)
    gl = dense_block_circuit(seed, nq)
    say("fused blocks seed", seed, "gates", len(gl))
    with E.Engine(nq, prec="f32", fuse_blocks=2, renorm_every=100, bucket_bits=4) as e:
        e.load(keys, re, im)
        e.apply(gl)
        got = e.export()
        st = e.stats()
    assert st["blocks"] > 0
    o32 = oracle.Sparse(nq, f32=True, thr=1e-7, renorm_every=100)
    o32.load(keys, re, im)
    o32.apply(gl)
    d32 = compare(got, o32.export(), 1e-5, max(2, got[0].size // 1000))
    o64 = oracle.Sparse(nq)
    o64.load(keys, re, im)
    o64.apply(gl)
    d64 = compare(got, o64.export(), 1e-5, max(2, got[0].size // 1000))
    say("max diff vs fp32 oracle %.2e, vs fp64 %.2e, sweeps %d" % (d32, d64, st["sweeps"]))
    n2 = float(np.sum(got[1] ** 2 + got[2] ** 2))
    assert abs(n2 - 1.0) < 1e-5


@pytest.mark.parametrize("seed", [1, 2])
def test_blocks_32x32_match_fp32_oracle_and_fp64(oracle, seed):
    """Five bucket qubits: windows become 32x32 complex blocks (two 32-row
    MFMA tiles per half-wavefront chain)."""
    nq = 18
    bq = (1, 4, 7, 10, 13)
    keys, re, im = random_state(500 + seed, nq, 400)
    gl = dense_block_circuit(seed, nq, layers=24, bq=bq)
    say("32x32 fused blocks seed", seed, "gates", len(gl))
    with E.Engine(nq, prec="f32", fuse_blocks=2, renorm_every=100, bucket_bits=5) as e:
        e.load(keys, re, im)
        e.apply(gl)
        got = e.export()
        st = e.stats()
    assert st["bucket_bits"] == 5 and st["bucket_mask"] == sum(1 << q for q in bq)
    assert st["blocks"] > 0 and st["block_gates"] >= 2 * st["blocks"]
    o32 = oracle.Sparse(nq, f32=True, thr=1e-7, renorm_every=100)
    o32.load(keys, re, im)
    o32.apply(gl)
    d32 = compare(got, o32.export(), 1e-5, max(2, got[0].size // 1000))
    o64 = oracle.Sparse(nq)
    o64.load(keys, re, im)
    o64.apply(gl)
    d64 = compare(got, o64.export(), 1e-5, max(2, got[0].size // 1000))
    say("blocks %d holding %d gates; max diff vs fp32 oracle %.2e, vs fp64 %.2e"
        % (st["blocks"], st["block_gates"], d32, d64))
    n2 = float(np.sum(got[1] ** 2 + got[2] ** 2))
    assert abs(n2 - 1.0) < 1e-5


def test_fp32_32_slot_buckets_gate_by_gate(oracle):
    """Five bucket qubits without fusion: the 32-slot FP32 sweep applies the
    window gates one by one (per-pair band test, no fused blocks) and agrees
    with the FP32 oracle."""
    nq = 18
    bq = (1, 4, 7, 10, 13)
    keys, re, im = random_state(520, nq, 400)
    gl = dense_block_circuit(3, nq, layers=24, bq=bq)
    with E.Engine(nq, prec="f32", renorm_every=100, bucket_bits=5) as e:
        e.load(keys, re, im)
        e.apply(gl)
        got = e.export()
        st = e.stats()
    assert st["bucket_bits"] == 5 and st["blocks"] == 0
    o32 = oracle.Sparse(nq, f32=True, thr=1e-7, renorm_every=100)
    o32.load(keys, re, im)
    o32.apply(gl)
    d32 = compare(got, o32.export(), 1e-5, max(2, got[0].size // 1000))
    say("32-slot FP32 gate by gate: %d entries, max diff vs fp32 oracle %.2e" % (got[0].size, d32))
    assert abs(float(np.sum(got[1] ** 2 + got[2] ** 2)) - 1.0) < 1e-5


def test_blocks_on_large_mix_close_to_fp64(oracle):
    """The 48-qubit mix (2^12 keys): windows are short, results as the FP32 form."""
    w = gen.circuit_large(5, nq=48, n_init=1 << 12, n_gates=600)
    with E.Engine(48, prec="f32", fuse_blocks=2, renorm_every=100) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.apply(w["gates"])
        got = e.export()
    o64 = oracle.Sparse(48)
    o64.load(w["keys"], w["re"], w["im"])
    o64.apply(w["gates"])
    wk, wr, wi = o64.export()
    common, ia, ib = np.intersect1d(got[0], wk, return_indices=True)
    ov = np.sum((wr[ib] - 1j * wi[ib]) * (got[1][ia] + 1j * got[2][ia]))
    infid = 1.0 - abs(ov) ** 2 / (np.sum(wr ** 2 + wi ** 2) * np.sum(got[1] ** 2 + got[2] ** 2))
    say("fused-block infidelity %.3e" % infid)
    assert infid <= 1e-6


def test_blocks_ignored_for_fp64():
    """The flag only applies to FP32 (16- or 32-slot buckets); FP64 runs gate by gate."""
    nq = 12
    keys, re, im = random_state(5, nq, 64)
    gl = dense_block_circuit(3, nq=12, layers=5)
    with E.Engine(nq, prec="f64", fuse_blocks=True, bucket_bits=4) as e, E.Engine(nq, prec="f64", bucket_bits=4) as f:
        e.load(keys, re, im)
        f.load(keys, re, im)
        e.apply(gl)
        f.apply(gl)
        a, b = e.export(), f.export()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_more_than_4096_blocks_in_one_flush(oracle):
    """4200 fused blocks in one program: each block's side-table entry (128
    units of 8 doubles) lies beyond 2^19 units after block 4096, so a short
    index field would hand later blocks another block's matrix.  The state
    must follow the FP64 oracle through all 29400 gates."""
    nq = 16
    bq = [1, 4, 7, 10]
    others = [q for q in range(nq) if q not in bq]
    rng = gen.SplitMix64(77)
    gl = []
    n_windows = 4200
    for _ in range(n_windows):
        for _k in range(6):
            q = bq[rng.randint(4)]
            k = [G.H, G.RY, G.RZ, G.PHASE, G.X, G.RX][rng.randint(6)]
            gl.append(G.make_gate(k, q, theta=rng.uniform() * 6.28))
        # CZ between a bucket qubit and another qubit closes the window
        gl.append(G.make_gate(G.Z, bq[rng.randint(4)], controls=(others[rng.randint(len(others))],)))
    gl = G.gate_array(gl)
    keys, re, im = random_state(91, nq, 300)
    say("blocks regression: %d gates, %d windows of 6" % (len(gl), n_windows))
    with E.Engine(nq, prec="f32", fuse_blocks=6, bucket_bits=4) as e:
        e.load(keys, re, im)
        e.apply(gl)
        got = e.export()
        st = e.stats()
    assert st["sweeps"] == 1, st["sweeps"]          # one flush, one program
    o64 = oracle.Sparse(nq)
    o64.load(keys, re, im)
    o64.apply(gl)
    want = o64.export()
    o32 = oracle.Sparse(nq, f32=True, thr=1e-7)
    o32.load(keys, re, im)
    o32.apply(gl)
    w32 = o32.export()
    # keys agree away from the FP32 threshold
    big = want[0][np.maximum(np.abs(want[1]), np.abs(want[2])) > 1e-5]
    assert np.all(np.isin(big, got[0])), "keys with |a| > 1e-5 missing"
    d64 = compare(got, want, 1e-4, max(2, got[0].size // 100))
    d32 = compare(w32, want, 1e-4, max(2, got[0].size // 100))
    say("max |a - a64|: fused blocks %.2e, FP32 gate by gate %.2e" % (d64, d32))
    assert d64 <= 4 * d32 + 1e-6
