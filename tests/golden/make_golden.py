"""Writes the golden fixtures of tests/test_golden.py (run from the repo root:
``python tests/golden/make_golden.py``).

The reference ships no golden vectors or tests for this path, so the fixtures
are the CPU oracle's outputs on small seeded inputs (the oracle itself is pinned
by tests/test_oracle.py) plus one closed form.  Each .npz holds the input state,
the gate records (sqe_gate bytes) and the expected sorted output state.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O          # noqa: E402
from sparqc_amd import gates as G       # noqa: E402
from sparqc_amd import gen              # noqa: E402
from tests.helpers import random_gates, random_state   # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def cases():
    w = gen.circuit_small(1)
    yield "c1_seed1_first80", dict(nq=20, keys=w["keys"], re=w["re"], im=w["im"], gates=w["gates"][:80],
                                   prec="f64", renorm_every=0)
    k, r, i = random_state(77, 12, 64)
    yield "mixed_12q_150", dict(nq=12, keys=k, re=r, im=i, gates=random_gates(78, 12, 150), prec="f64",
                                renorm_every=0)
    w = gen.circuit_large(3, nq=48, n_init=1 << 8, n_gates=300)
    yield "large48_2e8keys_300", dict(nq=48, keys=w["keys"], re=w["re"], im=w["im"], gates=w["gates"],
                                      prec="f64", renorm_every=0)
    yield "large48_2e8keys_300_fp32", dict(nq=48, keys=w["keys"], re=w["re"], im=w["im"], gates=w["gates"],
                                           prec="f32", renorm_every=100)


def main():
    for name, c in cases():
        f32 = c["prec"] == "f32"
        s = O.Sparse(c["nq"], f32=f32, thr=1e-7 if f32 else 1e-14, renorm_every=c["renorm_every"])
        s.load(c["keys"], c["re"], c["im"])
        s.apply(c["gates"])
        ok, orr, oi = s.export()
        gb = np.ascontiguousarray(G.gate_array(c["gates"])).view(np.uint8)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), nq=np.int64(c["nq"]), keys=c["keys"], re=c["re"],
                            im=c["im"], gates=gb, prec=np.array(c["prec"]),
                            renorm_every=np.int64(c["renorm_every"]), out_keys=ok, out_re=orr, out_im=oi,
                            updates=np.uint64(s.updates()))
        print(name, c["nq"], len(c["gates"]), ok.size)
    # closed form: H on each of 10 qubits of |0...0> -> 2^-5 on all 1024 keys
    # (the scenario of the reference's driver, emulator/src/main.cu:142-143)
    gl = G.gate_array([G.make_gate(G.H, q) for q in range(10)])
    np.savez_compressed(os.path.join(HERE, "all_h_10q.npz"), nq=np.int64(10), keys=np.zeros(1, np.uint64),
                        re=np.ones(1), im=np.zeros(1), gates=np.ascontiguousarray(gl).view(np.uint8),
                        prec=np.array("f64"), renorm_every=np.int64(0),
                        out_keys=np.arange(1024, dtype=np.uint64), out_re=np.full(1024, 2.0 ** -5),
                        out_im=np.zeros(1024), updates=np.uint64(sum(2 ** q for q in range(10))))


if __name__ == "__main__":
    main()
