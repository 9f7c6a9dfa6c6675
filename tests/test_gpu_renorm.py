"""FP32 renormalisation in one pass against the step-by-step form.

With renorm_every = R and renorm_mode = 1 the FP32 engine applies a queued
run of bucket-local gates in one pass: it records the FP64 norm at every R-th gate without
rescaling, then applies 1/sqrt(norm) of the last point; its prune decisions
are taken on amplitudes that lack the pending scale (so a decision can differ
from the step-by-step run's only for a value within that scale's distance
from 1 of the threshold), and a norm more than 2^-11 from 1 makes the engine
rerun the gates step by step.  Both forms must give the same state within
FP32 rounding (or agree to the bit, when the check made it rerun); the
forced rerun (renorm_mode 2) must give the step-by-step state to the bit.
"""
import sys

import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gen

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(w, **kw):
    with E.Engine(w["nq"], prec="f32", renorm_every=100, **kw) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.apply(w["gates"])
        out = e.export()
        st = e.stats()
    return out, st


@pytest.mark.parametrize("blocks", [0, 2])
def test_one_pass_matches_step_by_step(blocks):
    w = gen.circuit_large(11, nq=48, n_init=1 << 14, n_gates=1050)
    say("one-pass renormalisation, blocks", blocks)
    a, sa = run(w, renorm_mode=1, fuse_blocks=blocks)
    b, sb = run(w, renorm_mode=0, fuse_blocks=blocks)
    c, sc = run(w, renorm_mode=2, fuse_blocks=blocks)
    say("  sweeps: one pass %d (reruns %d), step by step %d, forced rerun %d (reruns %d)"
        % (sa["sweeps"], sa["reruns"], sb["sweeps"], sc["sweeps"], sc["reruns"]))
    assert sc["reruns"] == 1 and sb["reruns"] == 0
    assert sa["reruns"] == 1 or sa["sweeps"] == 2     # the pass and the final scale
    # the forced rerun is the step-by-step run
    assert np.array_equal(c[0], b[0]) and np.array_equal(c[1], b[1]) and np.array_equal(c[2], b[2])
    assert sc["updates"] == sb["updates"] and sa["updates"] == sb["updates"]
    assert np.array_equal(a[0], b[0])
    d = max(np.max(np.abs(a[1] - b[1])), np.max(np.abs(a[2] - b[2])))
    say("  max |one pass - step by step| = %.2e" % d)
    assert d <= 1e-6
    n2 = float(np.sum(a[1] ** 2 + a[2] ** 2))
    assert abs(n2 - 1.0) < 1e-5


def test_one_pass_across_flushes_and_sync_points():
    """Several runs in a row (a sync between them) and a run that ends
    between renormalisation points keep the step-by-step gate positions."""
    w = gen.circuit_large(12, nq=48, n_init=1 << 12, n_gates=900)
    gl = w["gates"]
    outs = []
    for mode in (1, 0):
        with E.Engine(48, prec="f32", renorm_every=100, renorm_mode=mode) as e:
            e.load(w["keys"], w["re"], w["im"])
            for a, b in ((0, 130), (130, 420), (420, 900)):
                e.apply(gl[a:b])
                e.sync()
            outs.append(e.export())
    a, b = outs
    assert np.array_equal(a[0], b[0])
    assert max(np.max(np.abs(a[1] - b[1])), np.max(np.abs(a[2] - b[2]))) <= 1e-6
