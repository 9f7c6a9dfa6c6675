"""FP32 prune decisions near the threshold, against a closed form.

The prune rule (emulator/include/data_t.h:50-60) drops a branching output
whose |re| and |im| are both below the threshold.  The FP32 engine decides
outputs with max(|re|,|im|) in [thr/4, 4 thr) on an FP64 recomputation; that
recomputation runs only for wavefronts that hold such an output.  Here
H(q0) H(q0) is the identity, and the state pairs a small amplitude t on key
2j with a large one L on key 2j+1, so the intermediate outputs (t +- L)/sqrt2
are far above the threshold and the final ones are t and L: key 2j survives
exactly when |t| >= thr.  The values of t sit on both sides of the threshold,
inside and outside the band, and whole runs of buckets hold no in-band value,
so both sides of the wavefront-uniform branch are taken.
"""
import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gates as G

pytestmark = pytest.mark.gpu

THR = 1e-7
L_AMP = 1e-5
# multiples of the threshold: outside the band [thr/4, 4 thr) ...
FAR = [0.05, 0.1, 0.2, 4.5, 10.0, -0.2, -6.0]
# ... and inside it, both sides of the threshold, 1e-3 away from it at least
NEAR = [0.26, 0.5, 0.99, 0.999, 1.001, 1.01, 2.0, 3.9, -0.999, -1.001]


def band_state(n_pairs):
    j = np.arange(n_pairs, dtype=np.uint64)
    # runs of 512 pairs alternate between far-only values and a mix of both
    run = (np.arange(n_pairs) // 512) % 2
    far = np.array(FAR)[np.arange(n_pairs) % len(FAR)]
    mixed_set = np.array(FAR + NEAR)
    mixed = mixed_set[(np.arange(n_pairs) * 7) % len(mixed_set)]
    t = np.where(run == 0, far, mixed) * THR
    keys = np.empty(2 * n_pairs, dtype=np.uint64)
    keys[0::2] = 2 * j
    keys[1::2] = 2 * j + 1
    re = np.empty(2 * n_pairs)
    re[0::2] = t
    re[1::2] = L_AMP
    im = np.zeros(2 * n_pairs)
    return keys, re, im, t


@pytest.mark.parametrize("fuse_blocks", [0, 2])
def test_fp32_prune_band_decisions(fuse_blocks):
    n_pairs = 1 << 12
    nq = 13
    keys, re, im, t = band_state(n_pairs)
    gl = [G.make_gate(G.H, 0), G.make_gate(G.H, 0)]
    with E.Engine(nq, prec="f32", prune_thr=THR, bucket_bits=4, fuse_blocks=fuse_blocks) as e:
        e.load(keys, re, im)
        e.apply(gl)
        gk, gr, gi = e.export()
    keep = np.abs(t) >= THR
    j = np.arange(n_pairs, dtype=np.uint64)
    want_k = np.sort(np.concatenate([2 * j[keep], 2 * j + 1]))
    print("fuse_blocks %d: kept %d of %d small amplitudes" % (fuse_blocks, int(keep.sum()), n_pairs), flush=True)
    assert np.array_equal(gk, want_k)
    want = dict(zip((2 * j[keep]).tolist(), t[keep].tolist()))
    want.update({int(k): L_AMP for k in (2 * j + 1)})
    w = np.array([want[int(k)] for k in gk])
    # FP32 arithmetic on values of size L: a few FP32 ulps of L
    assert np.max(np.abs(gr - w)) <= 1e-11
    assert np.max(np.abs(gi)) <= 1e-11


def test_fp64_prune_threshold_decisions():
    thr = 1e-14
    n_pairs = 1 << 12
    keys, re, im, t = band_state(n_pairs)
    re[0::2] *= thr / THR
    t = t * (thr / THR)
    re[1::2] = 1e-12
    gl = [G.make_gate(G.H, 0), G.make_gate(G.H, 0)]
    with E.Engine(13, prec="f64", prune_thr=thr, bucket_bits=4) as e:
        e.load(keys, re, im)
        e.apply(gl)
        gk, gr, gi = e.export()
    keep = np.abs(t) >= thr
    j = np.arange(n_pairs, dtype=np.uint64)
    assert np.array_equal(gk, np.sort(np.concatenate([2 * j[keep], 2 * j + 1])))
