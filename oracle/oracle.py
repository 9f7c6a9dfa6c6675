"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use this
module, and only as the checker.  The engine (``sparqc_amd``) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_load.restype = ctypes.c_int
        L.orc_load.argtypes = [ctypes.c_void_p, _u64p, _f64p, _f64p, ctypes.c_uint64]
        L.orc_apply.restype = ctypes.c_int
        L.orc_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        L.orc_nnz.restype = ctypes.c_uint64
        L.orc_nnz.argtypes = [ctypes.c_void_p]
        L.orc_updates.restype = ctypes.c_uint64
        L.orc_updates.argtypes = [ctypes.c_void_p]
        L.orc_norm2.restype = ctypes.c_double
        L.orc_norm2.argtypes = [ctypes.c_void_p]
        L.orc_export.restype = ctypes.c_uint64
        L.orc_export.argtypes = [ctypes.c_void_p, _u64p, _f64p, _f64p, ctypes.c_uint64]
        L.fc_create.restype = ctypes.c_void_p
        L.fc_create.argtypes = [ctypes.c_int, ctypes.c_double]
        L.fc_destroy.argtypes = [ctypes.c_void_p]
        L.fc_load.restype = ctypes.c_int
        L.fc_load.argtypes = [ctypes.c_void_p, _u64p, _f64p, _f64p, ctypes.c_uint64]
        L.fc_apply.restype = ctypes.c_int
        L.fc_apply.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        L.fc_updates.restype = ctypes.c_uint64
        L.fc_updates.argtypes = [ctypes.c_void_p]
        L.fc_nnz.restype = ctypes.c_uint64
        L.fc_nnz.argtypes = [ctypes.c_void_p]
        L.fc_export.restype = ctypes.c_uint64
        L.fc_export.argtypes = [ctypes.c_void_p, _u64p, _f64p, _f64p, ctypes.c_uint64]
        L.orc_dense_run.restype = ctypes.c_int
        L.orc_dense_run.argtypes = [ctypes.c_int, _u64p, _f64p, _f64p, ctypes.c_uint64,
                                    ctypes.c_void_p, ctypes.c_uint64, _f64p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


class Sparse:
    """Map-semantics sparse engine (FP64, or FP32 with FP64 near-threshold decisions)."""

    def __init__(self, nq, f32=False, thr=1e-14, renorm_every=0):
        self.h = lib().orc_create(nq, 1 if f32 else 0, thr, renorm_every)
        if not self.h:
            raise ValueError("bad qubit count")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def load(self, keys, re, im):
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        r = np.ascontiguousarray(re, dtype=np.float64)
        i = np.ascontiguousarray(im, dtype=np.float64)
        rc = lib().orc_load(self.h, _p(k, _u64p), _p(r, _f64p), _p(i, _f64p), k.size)
        if rc != 0:
            raise ValueError("oracle load failed (%d)" % rc)

    def apply(self, gates):
        from sparqc_amd.gates import gate_array
        g = gate_array(gates)
        rc = lib().orc_apply(self.h, g.ctypes.data, g.size)
        if rc != 0:
            raise ValueError("oracle apply failed")

    def nnz(self):
        return int(lib().orc_nnz(self.h))

    def norm2(self):
        return float(lib().orc_norm2(self.h))

    def updates(self):
        """Sum over applied gates of the entries entering each gate."""
        return int(lib().orc_updates(self.h))

    def export(self):
        n = self.nnz()
        k = np.zeros(n, dtype=np.uint64)
        r = np.zeros(n)
        i = np.zeros(n)
        m = lib().orc_export(self.h, _p(k, _u64p), _p(r, _f64p), _p(i, _f64p), n)
        return k[:m], r[:m], i[:m]


class FC:
    """The reference's full-capacity table on the CPU (dense slots addressed by
    key, an address book, two buffers): the CPU baseline of small-qubit
    workloads, nq <= 26."""

    def __init__(self, nq, thr=1e-14):
        self.h = lib().fc_create(nq, thr)
        if not self.h:
            raise ValueError("FC engine holds at most 26 qubits")

    def __del__(self):
        if getattr(self, "h", None):
            lib().fc_destroy(self.h)
            self.h = None

    def load(self, keys, re, im):
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        r = np.ascontiguousarray(re, dtype=np.float64)
        i = np.ascontiguousarray(im, dtype=np.float64)
        rc = lib().fc_load(self.h, _p(k, _u64p), _p(r, _f64p), _p(i, _f64p), k.size)
        if rc != 0:
            raise ValueError("FC load failed (%d)" % rc)

    def apply(self, gates):
        from sparqc_amd.gates import gate_array
        g = gate_array(gates)
        if lib().fc_apply(self.h, g.ctypes.data, g.size) != 0:
            raise ValueError("FC apply failed")

    def updates(self):
        return int(lib().fc_updates(self.h))

    def nnz(self):
        return int(lib().fc_nnz(self.h))

    def export(self):
        n = self.nnz()
        k = np.zeros(n, dtype=np.uint64)
        r = np.zeros(n)
        i = np.zeros(n)
        m = lib().fc_export(self.h, _p(k, _u64p), _p(r, _f64p), _p(i, _f64p), n)
        return k[:m], r[:m], i[:m]


def run_sparse(w, gates=None, f32=False, thr=None, renorm_every=0):
    """Run a workload dict (keys/re/im/gates/nq/thr) through the oracle; return export."""
    s = Sparse(w["nq"], f32=f32, thr=w["thr"] if thr is None else thr, renorm_every=renorm_every)
    s.load(w["keys"], w["re"], w["im"])
    s.apply(w["gates"] if gates is None else gates)
    return s.export()


def run_dense(nq, keys, re, im, gates):
    from sparqc_amd.gates import gate_array
    g = gate_array(gates)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    r = np.ascontiguousarray(re, dtype=np.float64)
    i = np.ascontiguousarray(im, dtype=np.float64)
    N = 1 << nq
    outr = np.zeros(N)
    outi = np.zeros(N)
    rc = lib().orc_dense_run(nq, _p(k, _u64p), _p(r, _f64p), _p(i, _f64p), k.size,
                             g.ctypes.data, g.size, _p(outr, _f64p), _p(outi, _f64p))
    if rc != 0:
        raise ValueError("dense oracle failed")
    return outr + 1j * outi
