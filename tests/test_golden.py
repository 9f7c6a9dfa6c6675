"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

The reference holds no golden vectors for this path; the fixtures are the
oracle's outputs on small seeded inputs plus the closed form of H on every
qubit of |0...0> (emulator/src/main.cu:142-143).  CPU: the oracle reproduces
them bit for bit (the closed form within 1e-15).  GPU: the engine matches them
(keys exact, FP64 amplitudes within 1e-12, FP32 within 1e-6) with the same
update count.
"""
import glob
import os

import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gates as G
from tests.helpers import assert_same_state

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def load(path):
    d = np.load(path, allow_pickle=False)
    gates = np.frombuffer(d["gates"].tobytes(), dtype=G.GATE_DTYPE)
    return dict(nq=int(d["nq"]), keys=d["keys"], re=d["re"], im=d["im"], gates=gates, prec=str(d["prec"]),
                renorm_every=int(d["renorm_every"]), out=(d["out_keys"], d["out_re"], d["out_im"]),
                updates=int(d["updates"]))


def ids(paths):
    return [os.path.basename(p)[:-4] for p in paths]


def test_fixtures_present():
    assert len(FIXTURES) >= 5


@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_oracle_reproduces_fixture(oracle, path):
    c = load(path)
    f32 = c["prec"] == "f32"
    s = oracle.Sparse(c["nq"], f32=f32, thr=1e-7 if f32 else 1e-14, renorm_every=c["renorm_every"])
    s.load(c["keys"], c["re"], c["im"])
    s.apply(c["gates"])
    closed = os.path.basename(path).startswith("all_h")
    assert_same_state(s.export(), c["out"], atol=1e-15 if closed else 0.0, label=os.path.basename(path))
    assert s.updates() == c["updates"]


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=ids(FIXTURES))
def test_engine_matches_fixture(path):
    c = load(path)
    print("golden", os.path.basename(path), flush=True)
    with E.Engine(c["nq"], prec=c["prec"], renorm_every=c["renorm_every"]) as e:
        e.load(c["keys"], c["re"], c["im"])
        e.apply(c["gates"])
        got = e.export()
        upd = e.stats()["updates"]
    if c["prec"] == "f32":
        common, ia, ib = np.intersect1d(got[0], c["out"][0], return_indices=True)
        assert np.setxor1d(got[0], c["out"][0]).size <= 2
        assert np.max(np.abs(got[1][ia] - c["out"][1][ib])) <= 1e-6
        assert np.max(np.abs(got[2][ia] - c["out"][2][ib])) <= 1e-6
    else:
        assert_same_state(got, c["out"], atol=1e-12, label=os.path.basename(path))
    assert upd == c["updates"]
