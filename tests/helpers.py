"""Shared helpers of the test suite (comparison against the CPU oracle)."""
import numpy as np

from sparqc_amd import gates as G
from sparqc_amd import gen


def assert_same_state(got, want, atol=1e-12, label=""):
    """Bit-exact key sets, amplitudes within atol (both exports sorted by key)."""
    gk, gr, gi = got
    wk, wr, wi = want
    assert gk.size == wk.size, "%s: %d entries vs %d expected" % (label, gk.size, wk.size)
    assert np.array_equal(gk, wk), "%s: key sets differ (%d keys only on one side)" % (
        label, np.setxor1d(gk, wk).size)
    if gk.size:
        d = max(np.max(np.abs(gr - wr)), np.max(np.abs(gi - wi)))
        assert d <= atol, "%s: max amplitude difference %.3e > %.1e" % (label, d, atol)


def random_state(seed, nq, n):
    rng = gen.SplitMix64(seed)
    return gen.random_state(rng, n, nq)


def random_gates(seed, nq, n, kinds=None, p_ctrl=0.3, max_ctrl=2):
    """Any gate kind, random targets and random controls (with random values)."""
    rng = gen.SplitMix64(seed)
    kinds = list(range(11)) if kinds is None else list(kinds)
    out = []
    for _ in range(n):
        k = kinds[rng.randint(len(kinds))]
        t = rng.randint(nq)
        ctrls = []
        if rng.uniform() < p_ctrl:
            nc = 1 + rng.randint(max_ctrl)
            pool = [q for q in range(nq) if q != t]
            for _c in range(nc):
                c = pool[rng.randint(len(pool))]
                if c not in ctrls:
                    ctrls.append(c)
        cm = 0
        for c in ctrls:
            cm |= 1 << c
        cv = cm & rng.next_u64()
        th = rng.uniform() * 6.283185307179586
        if k == G.U:
            a, b = rng.uniform() * 6.28, rng.uniform() * 6.28
            c, s = np.cos(th / 2), np.sin(th / 2)
            m = np.array([[c * np.exp(1j * a), -s * np.exp(1j * b)],
                          [s * np.exp(-1j * b), c * np.exp(-1j * a)]])
            g = G.make_gate(k, t, ctrls, matrix=m, ctrl_val=cv)
        elif k in (G.DIAG, G.PERM):
            a, b = rng.uniform() * 6.28, rng.uniform() * 6.28
            g = G.make_gate(k, t, ctrls, matrix=[np.exp(1j * a), np.exp(1j * b)], ctrl_val=cv)
        else:
            g = G.make_gate(k, t, ctrls, theta=th, ctrl_val=cv)
        out.append(g)
    return G.gate_array(out)
