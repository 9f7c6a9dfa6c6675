"""Seeded synthetic workloads for the sparse statevector engine.

Randomness comes from a counter-based splitmix64 stream written here (not numpy's
or the C++ standard library's distributions, whose outputs are implementation
defined), so the same seed gives the same circuit and state on every machine.

Workloads:

* ``circuit_small``  -- 20 qubits from |0...0>, 200 gates drawn uniformly from
  {H, RY, X, CX, RZ, CZ}; H/RY only on a fixed random set of 10 "active" qubits.
* ``circuit_large``  -- 48 qubits, 2^20 distinct uniform keys in [0, 2^48) with
  normalised complex-normal amplitudes; 2000 gates, 60% diagonal {RZ, PHASE, CZ},
  35% permutation {X, CX, CCX}, 5% branching {H, RY} on 4 fixed qubits.
  Branching qubits are never CX/CCX controls, so the number of nonzeros stays
  below n_init * 2^4.
* ``batch_circuits`` -- independent circuits of 16..40 qubits, 2^8..2^18 initial
  keys, depth 50..500 and 3 branching qubits each.
* ``circuit_layers`` -- the large state under layers of single-qubit gates on the
  4 branching qubits separated by diagonal / CX gates elsewhere (long windows of
  bucket-qubit gates, the shape the fused dense blocks are for).

The reference's own driver applies H to every qubit of |0...0>
(emulator/src/main.cu:142-143); ``circuit_all_h`` reproduces that scenario.
"""
import math

import numpy as np

from . import gates as G

_GOLD = 0x9E3779B97F4A7C15
_M1 = 0xBF58476D1CE4E5B9
_M2 = 0x94D049BB133111EB
_MASK = (1 << 64) - 1


def _mix_scalar(z):
    z = (z ^ (z >> 30)) * _M1 & _MASK
    z = (z ^ (z >> 27)) * _M2 & _MASK
    return z ^ (z >> 31)


def _mix_array(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(_M1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(_M2)
    return z ^ (z >> np.uint64(31))


class SplitMix64:
    """Counter-based splitmix64: output n is mix(seed + (n+1)*golden)."""

    def __init__(self, seed):
        self.state = int(seed) & _MASK

    def next_u64(self):
        self.state = (self.state + _GOLD) & _MASK
        return _mix_scalar(self.state)

    def u64_array(self, n):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + idx * np.uint64(_GOLD)
        self.state = (self.state + n * _GOLD) & _MASK
        return _mix_array(z)

    def uniform(self):
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def uniform_array(self, n):
        return (self.u64_array(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def randint(self, n):
        """Integer in [0, n) by multiply-shift of a 64-bit draw."""
        return (self.next_u64() * n) >> 64

    def sample(self, n, k):
        """k distinct integers from range(n), partial Fisher-Yates."""
        a = list(range(n))
        for i in range(k):
            j = i + self.randint(n - i)
            a[i], a[j] = a[j], a[i]
        return a[:k]

    def normal_complex_array(self, n):
        """n complex normals by Box-Muller (radius from u1 in (0,1], angle from u2)."""
        u = self.uniform_array(2 * n)
        u1 = 1.0 - u[0::2]
        u2 = u[1::2]
        r = np.sqrt(-2.0 * np.log(u1))
        return r * np.cos(2.0 * math.pi * u2), r * np.sin(2.0 * math.pi * u2)


def distinct_keys(rng, n, nq):
    """n distinct keys uniform in [0, 2^nq), in order of first draw."""
    if n > (1 << nq):
        raise ValueError("more keys than basis states")
    out = np.zeros(0, dtype=np.uint64)
    shift = np.uint64(64 - nq)
    while out.size < n:
        need = n - out.size
        draw = rng.u64_array(need + need // 8 + 16) >> shift
        cand = np.concatenate([out, draw])
        _, first = np.unique(cand, return_index=True)
        keep = np.sort(first)
        out = cand[keep][:n]
    return out


def random_state(rng, n, nq):
    keys = distinct_keys(rng, n, nq)
    re, im = rng.normal_complex_array(n)
    s = math.sqrt(float(np.sum(re * re + im * im)))
    return keys, re / s, im / s


def _two_distinct(rng, nq, exclude_ctrl=()):
    pool = [q for q in range(nq) if q not in exclude_ctrl]
    c = pool[rng.randint(len(pool))]
    t = rng.randint(nq - 1)
    if t >= c:
        t += 1
    return c, t


def circuit_small(seed, nq=20, n_gates=200, n_active=10):
    """20-qubit, 200-gate circuit from |0...0> (the CPU-runnable workload)."""
    rng = SplitMix64(seed)
    active = rng.sample(nq, n_active)
    out = []
    for _ in range(n_gates):
        kind = rng.randint(6)
        if kind == 0:
            out.append(G.make_gate(G.H, active[rng.randint(n_active)]))
        elif kind == 1:
            t = active[rng.randint(n_active)]
            out.append(G.make_gate(G.RY, t, theta=rng.uniform() * 2 * math.pi))
        elif kind == 2:
            out.append(G.make_gate(G.X, rng.randint(nq)))
        elif kind == 3:
            c, t = _two_distinct(rng, nq)
            out.append(G.make_gate(G.X, t, controls=(c,)))
        elif kind == 4:
            t = rng.randint(nq)
            out.append(G.make_gate(G.RZ, t, theta=rng.uniform() * 2 * math.pi))
        else:
            c, t = _two_distinct(rng, nq)
            out.append(G.make_gate(G.Z, t, controls=(c,)))
    keys = np.zeros(1, dtype=np.uint64)
    return dict(nq=nq, keys=keys, re=np.ones(1), im=np.zeros(1),
                gates=G.gate_array(out), active=active, thr=1e-14)


def mixed_gates(rng, nq, n_gates, branch_qubits, p_diag=0.6, p_perm=0.35):
    """Gate list with the diagonal / permutation / branching mix of the large workload."""
    out = []
    bq = list(branch_qubits)
    for _ in range(n_gates):
        u = rng.uniform()
        if u < p_diag:
            kind = rng.randint(3)
            if kind == 0:
                out.append(G.make_gate(G.RZ, rng.randint(nq), theta=rng.uniform() * 2 * math.pi))
            elif kind == 1:
                out.append(G.make_gate(G.PHASE, rng.randint(nq), theta=rng.uniform() * 2 * math.pi))
            else:
                c, t = _two_distinct(rng, nq)
                out.append(G.make_gate(G.Z, t, controls=(c,)))
        elif u < p_diag + p_perm:
            kind = rng.randint(3)
            if kind == 0:
                out.append(G.make_gate(G.X, rng.randint(nq)))
            elif kind == 1:
                c, t = _two_distinct(rng, nq, exclude_ctrl=bq)
                out.append(G.make_gate(G.X, t, controls=(c,)))
            else:
                pool = [q for q in range(nq) if q not in bq]
                c1 = pool[rng.randint(len(pool))]
                pool2 = [q for q in pool if q != c1]
                c2 = pool2[rng.randint(len(pool2))]
                rest = [q for q in range(nq) if q not in (c1, c2)]
                t = rest[rng.randint(len(rest))]
                out.append(G.make_gate(G.X, t, controls=(c1, c2)))
        else:
            t = bq[rng.randint(len(bq))]
            if rng.randint(2) == 0:
                out.append(G.make_gate(G.H, t))
            else:
                out.append(G.make_gate(G.RY, t, theta=rng.uniform() * 2 * math.pi))
    return G.gate_array(out)


def circuit_large(seed, nq=48, n_init=1 << 20, n_gates=2000, n_branch=4):
    """48-qubit state of n_init random keys and a 2000-gate mixed circuit."""
    rng = SplitMix64(seed)
    branch = rng.sample(nq, n_branch)
    keys, re, im = random_state(rng, n_init, nq)
    gl = mixed_gates(rng, nq, n_gates, branch)
    return dict(nq=nq, keys=keys, re=re, im=im, gates=gl, branch=branch, thr=1e-14)


def batch_circuits(seed, n_circuits=256, nq_range=(16, 40), log2_nnz=(8.0, 18.0),
                   depth_range=(50, 500), n_branch=3):
    """Independent circuits of uneven qubit count, nonzero count and depth.

    The initial nonzero count is floor(2^U[a,b]) clamped to 2^(nq-1)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n_circuits):
        nq = nq_range[0] + rng.randint(nq_range[1] - nq_range[0] + 1)
        n0 = int(math.floor(2.0 ** (log2_nnz[0] + rng.uniform() * (log2_nnz[1] - log2_nnz[0]))))
        n0 = max(1, min(n0, 1 << (nq - 1)))
        depth = depth_range[0] + rng.randint(depth_range[1] - depth_range[0] + 1)
        sub = SplitMix64(rng.next_u64())
        branch = sub.sample(nq, n_branch)
        keys, re, im = random_state(sub, n0, nq)
        gl = mixed_gates(sub, nq, depth, branch)
        out.append(dict(nq=nq, keys=keys, re=re, im=im, gates=gl, branch=branch, thr=1e-14))
    return out


def circuit_layers(seed, nq=48, n_init=1 << 20, n_gates=2000, n_branch=4):
    """The large state with a layered circuit: each layer applies a random
    single-qubit gate (H, RY, RX, RZ, PHASE, X) to every one of the n_branch
    branching qubits, then diagonal and CX gates on other qubits (one RZ or
    PHASE per layer qubit and one CX per layer, controls off the branching
    qubits).  Gates on the branching qubits come in long windows that commute
    with everything between them."""
    rng = SplitMix64(seed)
    branch = rng.sample(nq, n_branch)
    keys, re, im = random_state(rng, n_init, nq)
    others = [q for q in range(nq) if q not in branch]
    out = []
    while len(out) < n_gates:
        for q in branch:
            k = (G.H, G.RY, G.RX, G.RZ, G.PHASE, G.X)[rng.randint(6)]
            out.append(G.make_gate(k, q, theta=rng.uniform() * 2 * math.pi))
            t = others[rng.randint(len(others))]
            out.append(G.make_gate((G.RZ, G.PHASE)[rng.randint(2)], t, theta=rng.uniform() * 2 * math.pi))
        c, t = _two_distinct(rng, nq, exclude_ctrl=branch)
        if t not in branch:
            out.append(G.make_gate(G.X, t, controls=(c,)))
    return dict(nq=nq, keys=keys, re=re, im=im, gates=G.gate_array(out[:n_gates]), branch=branch, thr=1e-14)


def circuit_all_h(nq):
    """H on every qubit of |0...0> (emulator/src/main.cu:142-143)."""
    gl = G.gate_array([G.make_gate(G.H, q) for q in range(nq)])
    return dict(nq=nq, keys=np.zeros(1, dtype=np.uint64), re=np.ones(1), im=np.zeros(1),
                gates=gl, thr=1e-14)


def poisson_arrivals(seed, rate, n):
    """Arrival times (seconds) of n requests at the given rate (exponential gaps)."""
    rng = SplitMix64(seed)
    u = 1.0 - rng.uniform_array(n)
    return np.cumsum(-np.log(u) / rate)
