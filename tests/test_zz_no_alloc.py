"""sqe_norm2_noalloc (the allocation-free variant of sqe_norm2, fixed device
scratch) against sqe_norm2 through the C ABI: identical results on two calls in
a row, the second on a smaller state; queued gates and oversize states are
refused with a status, not handled by allocating."""
import ctypes
import sys

import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gen
from tests.helpers import random_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def norms(e):
    """(sqe_norm2, sqe_norm2_noalloc) of the engine's current state, through the C ABI."""
    L = E.lib()
    a, b = ctypes.c_double(), ctypes.c_double()
    assert L.sqe_norm2(e.h, ctypes.byref(a)) == E.SQE_OK, L.sqe_last_error()
    rc = L.sqe_norm2_noalloc(e.h, ctypes.byref(b))
    assert rc == E.SQE_OK, L.sqe_last_error()
    return a.value, b.value


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_norm2_noalloc_matches_norm2(prec):
    w = gen.circuit_large(3, nq=48, n_init=1 << 16, n_gates=400)
    small = gen.circuit_large(4, nq=48, n_init=1 << 9, n_gates=100)
    say("norm2_noalloc", prec, "2^16 keys then 2^9 keys")
    with E.Engine(48, prec=prec) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.apply(w["gates"])
        e.sync()
        n_ref, n_na = norms(e)
        say("  large: %.17g %.17g" % (n_ref, n_na))
        assert n_na == n_ref and abs(n_na - 1.0) < 1e-4
        # the second call, on a smaller state, must not see the first call's partials
        e.load(small["keys"], small["re"] * 0.5, small["im"] * 0.5)
        e.apply(small["gates"])
        e.sync()
        n_ref2, n_na2 = norms(e)
        say("  small: %.17g %.17g" % (n_ref2, n_na2))
        assert n_na2 == n_ref2 and abs(n_na2 - 0.25) < 1e-4


def test_norm2_noalloc_refuses_queued_gates_and_oversize_state():
    L = E.lib()
    out = ctypes.c_double()
    keys, re, im = random_state(5, 20, 1000)
    with E.Engine(20) as e:
        e.load(keys, re, im)
        e.apply(gen.circuit_small(1)["gates"][:10])
        assert L.sqe_norm2_noalloc(e.h, ctypes.byref(out)) == E.SQE_EINVAL
        assert b"queued" in L.sqe_last_error()
        e.sync()
        assert L.sqe_norm2_noalloc(e.h, ctypes.byref(out)) == E.SQE_OK
        assert abs(out.value - e.norm2()) == 0.0
    # more buckets than the fixed scratch covers: one key per bucket
    n = (16384 * 256) + 4096
    rng = gen.SplitMix64(9)
    k = gen.distinct_keys(rng, n, 40)
    r = np.full(n, 1.0 / np.sqrt(n))
    with E.Engine(40, bucket_bits=0) as e:
        e.load(k, r, np.zeros(n))
        e.sync()
        e.prepare([])
        assert L.sqe_norm2_noalloc(e.h, ctypes.byref(out)) == E.SQE_EINVAL
        assert b"fixed scratch" in L.sqe_last_error()
        assert abs(e.norm2() - 1.0) < 1e-9
