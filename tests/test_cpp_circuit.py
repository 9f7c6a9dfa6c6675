"""The reference's C++ circuit API (include/sparqc_circuit.hpp) over libsqe.so.

tests/cpp/circuit_check.cpp drives Circuit__ the way the reference's driver
does (emulator/src/main.cu:87-148) and checks closed forms: H on every qubit,
a Bell pair, the Toffoli truth table, Rx = H Rz H and Rz = e^{-i t/2} Phase on a
random state, controlled rotations, and errors thrown as std::string.
"""
import os
import subprocess

import pytest

from sparqc_amd import build as B


def _binary():
    if not os.path.exists(B.CHECK_BIN) or os.path.getmtime(B.CHECK_BIN) < os.path.getmtime(B.CHECK_SRC):
        B.build_circuit_check()
    return B.CHECK_BIN


def test_circuit_api_without_gpu_reports_std_string():
    """Off the GPU the constructor throws the engine's message as std::string."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "NO_GPU" in r.stdout


@pytest.mark.gpu
def test_circuit_api_on_gpu():
    print("circuit_check on GPU 0", flush=True)
    r = subprocess.run([_binary()], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], flush=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert r.stdout.strip().endswith("OK")
