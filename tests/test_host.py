"""Host-side checks that need no GPU: generators, gate records, the C ABI library."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from sparqc_amd import engine as E
from sparqc_amd import gates as G
from sparqc_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_splitmix_known_values():
    # splitmix64 reference outputs for seed 0 (Vigna's published generator)
    r = gen.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                               0x06C45D188009454F]
    r1, r2 = gen.SplitMix64(42), gen.SplitMix64(42)
    a = r1.u64_array(1000)
    assert np.array_equal(a, np.array([r2.next_u64() for _ in range(1000)], dtype=np.uint64))


def test_generators_are_deterministic_and_well_formed():
    a = gen.circuit_large(3, nq=48, n_init=1 << 12, n_gates=2000)
    b = gen.circuit_large(3, nq=48, n_init=1 << 12, n_gates=2000)
    assert np.array_equal(a["keys"], b["keys"]) and np.array_equal(a["gates"], b["gates"])
    k = a["keys"]
    assert np.unique(k).size == k.size and int(k.max()) < (1 << 48)
    assert abs(float(np.sum(a["re"] ** 2 + a["im"] ** 2)) - 1.0) < 1e-12
    gl = a["gates"]
    cls = np.array([G.gate_class(x) for x in gl["kind"]])
    assert abs(np.mean(cls == G.CLS_DIAG) - 0.60) < 0.04
    assert abs(np.mean(cls == G.CLS_PERM) - 0.35) < 0.04
    assert abs(np.mean(cls == G.CLS_BRANCH) - 0.05) < 0.02
    bq = set(a["branch"])
    bmask = sum(1 << q for q in bq)
    for g in gl:
        if G.gate_class(g["kind"]) == G.CLS_BRANCH:
            assert int(g["target"]) in bq and int(g["ctrl_mask"]) == 0
        if int(g["kind"]) == G.X:
            assert int(g["ctrl_mask"]) & bmask == 0       # branching qubits never control X
    s = gen.circuit_small(1)
    assert s["nq"] == 20 and len(s["gates"]) == 200
    act = set(s["active"])
    for g in s["gates"]:
        if int(g["kind"]) in (G.H, G.RY):
            assert int(g["target"]) in act


def test_batch_generator_ranges():
    cs = gen.batch_circuits(5, n_circuits=40)
    for c in cs:
        assert 16 <= c["nq"] <= 40
        assert 1 <= c["keys"].size <= min(1 << 18, 1 << (c["nq"] - 1))
        assert 50 <= len(c["gates"]) <= 500
        assert len(c["branch"]) == 3


def test_poisson_arrivals():
    t = gen.poisson_arrivals(1, 1000.0, 20000)
    gaps = np.diff(t)
    assert np.all(gaps > 0) and abs(np.mean(gaps) - 1e-3) < 5e-5


def test_gate_record_layout_matches_header():
    hdr = open(os.path.join(ROOT, "include", "sqe.h")).read()
    assert "SQE_GATE_PERM = 10" in hdr
    assert G.GATE_DTYPE.itemsize == 88
    assert G.GATE_DTYPE.fields["p"][1] == 24
    g = G.make_gate(G.RY, 3, theta=1.0)
    assert g["p"][0] == math.cos(0.5) and g["p"][1] == math.sin(0.5)
    with pytest.raises(ValueError):
        G.make_gate(G.X, 2, controls=(2,))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "sqe.h")).read()
    declared = set(re.findall(r"\b(sqe_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(E.EXPORTS)
    L = E.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_struct_layouts():
    assert ctypes.sizeof(E.Config) == 80
    cfg = E.Config()
    E.lib().sqe_config_default(ctypes.byref(cfg), 30)
    assert cfg.nq == 30 and cfg.bucket_bits == -1 and cfg.max_load == 0.5


def test_struct_mirrors_match_the_header(tmp_path):
    """Every public struct of include/sqe.h, compiled by the host C compiler,
    has the size and field offsets of its Python mirror (ctypes structures
    and the gate record dtype)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    mirrors = {"sqe_config": E.Config, "sqe_stats": E.Stats, "sqe_batch_circuit": E.BatchCircuit,
               "sqe_batch_tuning": E.BatchTuning}
    want = {}
    for cname, py in mirrors.items():
        want[cname] = (ctypes.sizeof(py), {f: getattr(py, f).offset for f, _ in py._fields_})
    want["sqe_gate"] = (G.GATE_DTYPE.itemsize, {f: G.GATE_DTYPE.fields[f][1] for f in G.GATE_DTYPE.names})
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "sqe.h"', "int main(void) {"]
    for cname, (_, fields) in want.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in fields:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            s, f, v = ln.split()
            got[(s, f)] = int(v)
    for cname, (size, fields) in want.items():
        assert got[(cname, "size")] == size, cname
        for f, off in fields.items():
            assert got[(cname, f)] == off, (cname, f)


def test_create_without_gpu_reports_no_device():
    try:
        e = E.Engine(10)
    except E.SqeError as ex:
        assert ex.status == E.SQE_EDEVICE
        assert "no HIP device" in str(ex)
    else:                                   # a GPU is present: nothing to check here
        e.close()


def test_bad_config_rejected_before_device():
    with pytest.raises(E.SqeError) as ex:
        E.Engine(64)
    assert ex.value.status == E.SQE_EINVAL
    with pytest.raises(E.SqeError) as ex:
        E.Engine(20, bucket_bits=5)          # FP64 buckets hold at most 2^4 slots
    assert ex.value.status == E.SQE_EINVAL


def test_program_encoder_side_table_index():
    """Fused blocks and general-U records address their side-table entry by a
    64-bit index of their own: more than 4096 blocks (each takes 128 units)
    and more than 2^19 U records in one program keep their own coefficients."""
    import subprocess
    from sparqc_amd import build as B
    srcs = [B.ENC_SRC, os.path.join(B.CSRC, "sqe_plan.hpp"), os.path.join(B.CSRC, "sqe_device.hpp")]
    if not os.path.exists(B.ENC_BIN) or os.path.getmtime(B.ENC_BIN) < max(os.path.getmtime(p) for p in srcs):
        B.build_encode_check()
    r = subprocess.run([B.ENC_BIN, "4200", str((1 << 19) + 300)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "encode_check ok" in r.stdout
    assert ">= 2^19" in r.stdout
