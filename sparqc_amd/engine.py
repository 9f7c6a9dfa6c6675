"""Python binding of the C ABI in include/sqe.h (libsqe.so, built in-tree).

The engine never falls back to a CPU path: if libsqe.so is missing this module
raises, and every call that needs a GPU reports the engine's own error.
"""
import ctypes
import os

import numpy as np

from . import gates as G

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsqe.so")

SQE_OK, SQE_EINVAL, SQE_ENOMEM, SQE_ETABLE_FULL, SQE_EHIP, SQE_EDEVICE, SQE_ENOTLOCAL = 0, -1, -2, -3, -4, -5, -6
STATUS_NAMES = {0: "SQE_OK", -1: "SQE_EINVAL", -2: "SQE_ENOMEM", -3: "SQE_ETABLE_FULL", -6: "SQE_ENOTLOCAL",
                -4: "SQE_EHIP", -5: "SQE_EDEVICE"}
F64, F32 = 0, 1

# every entry point declared in include/sqe.h
EXPORTS = ["sqe_config_default", "sqe_create", "sqe_destroy", "sqe_set_basis", "sqe_load",
           "sqe_apply", "sqe_apply_gates", "sqe_nnz", "sqe_norm2", "sqe_norm2_noalloc", "sqe_export", "sqe_export_where",
           "sqe_checkpoint_save", "sqe_checkpoint_load", "sqe_sample", "sqe_sync",
           "sqe_get_stats", "sqe_prepare", "sqe_snapshot_save", "sqe_snapshot_restore",
           "sqe_last_error",
           "sqe_batch_create", "sqe_batch_destroy", "sqe_batch_add", "sqe_batch_add_pinned", "sqe_batch_step",
           "sqe_batch_run",
           "sqe_batch_sync", "sqe_batch_info", "sqe_batch_export", "sqe_batch_remove", "sqe_batch_finished",
           "sqe_batch_export_start", "sqe_batch_export_poll",
           "sqe_batch_usage",
           "sqe_batch_get_tuning", "sqe_batch_set_tuning", "sqe_batch_autotune", "sqe_batch_snapshot_save",
           "sqe_batch_snapshot_restore", "sqe_batch_last_error", "sqe_host_alloc", "sqe_host_free"]


class SqeError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


class Config(ctypes.Structure):
    _fields_ = [("nq", ctypes.c_uint32), ("prec", ctypes.c_uint32), ("prune_thr", ctypes.c_double),
                ("max_load", ctypes.c_double), ("hbm_budget_bytes", ctypes.c_uint64),
                ("device", ctypes.c_int32), ("bucket_bits", ctypes.c_int32),
                ("renorm_every", ctypes.c_uint32), ("fuse_window", ctypes.c_uint32),
                ("timing", ctypes.c_uint32), ("mem_reserve_ppm", ctypes.c_uint32),
                ("fuse_blocks", ctypes.c_uint32), ("batch_eager", ctypes.c_uint32),
                ("renorm_mode", ctypes.c_uint32), ("frame_off", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 2)]


class Stats(ctypes.Structure):
    _fields_ = [("gates", ctypes.c_uint64), ("updates", ctypes.c_uint64), ("nnz", ctypes.c_uint64),
                ("buckets", ctypes.c_uint64), ("capacity", ctypes.c_uint64),
                ("bytes_device", ctypes.c_uint64), ("bucket_bits", ctypes.c_int32),
                ("reruns", ctypes.c_uint32), ("bucket_mask", ctypes.c_uint64), ("sweeps", ctypes.c_uint64),
                ("sweep_ops", ctypes.c_uint64), ("pairs", ctypes.c_uint64), ("remaps", ctypes.c_uint64),
                ("moved", ctypes.c_uint64), ("sweep_ms", ctypes.c_double), ("pair_ms", ctypes.c_double), ("other_ms", ctypes.c_double),
                ("moved_pair", ctypes.c_uint64), ("blocks", ctypes.c_uint64), ("block_gates", ctypes.c_uint64)]


class BatchCircuit(ctypes.Structure):
    _fields_ = [("done", ctypes.c_uint32), ("bucket_bits", ctypes.c_int32), ("gates_total", ctypes.c_uint64),
                ("gates_done", ctypes.c_uint64), ("updates", ctypes.c_uint64), ("nnz", ctypes.c_uint64),
                ("buckets", ctypes.c_uint64), ("bytes", ctypes.c_uint64)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_uint32), ("tiles_per_item", ctypes.c_uint32),
                ("load_factor", ctypes.c_double), ("sweep_ms", ctypes.c_double), ("launches", ctypes.c_uint64),
                ("load_ms", ctypes.c_double), ("one_launch", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


assert ctypes.sizeof(Config) == 80
assert ctypes.sizeof(Stats) == 152
assert ctypes.sizeof(BatchCircuit) == 56
assert ctypes.sizeof(BatchTuning) == 48

_lib = None
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)


def lib():
    """Load libsqe.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libsqe.so not built: run `python -m sparqc_amd.build` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        st = ctypes.c_int
        L.sqe_config_default.argtypes = [ctypes.POINTER(Config), ctypes.c_uint32]
        L.sqe_config_default.restype = None
        L.sqe_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
        L.sqe_create.restype = st
        L.sqe_destroy.argtypes = [vp]
        L.sqe_destroy.restype = st
        L.sqe_set_basis.argtypes = [vp, ctypes.c_uint64]
        L.sqe_set_basis.restype = st
        L.sqe_load.argtypes = [vp, _u64p, _f64p, _f64p, ctypes.c_uint64]
        L.sqe_load.restype = st
        L.sqe_apply.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, _f64p, ctypes.c_int]
        L.sqe_apply.restype = st
        L.sqe_apply_gates.argtypes = [vp, vp, ctypes.c_uint64]
        L.sqe_apply_gates.restype = st
        L.sqe_prepare.argtypes = [vp, vp, ctypes.c_uint64]
        L.sqe_prepare.restype = st
        L.sqe_nnz.argtypes = [vp, _u64p]
        L.sqe_nnz.restype = st
        L.sqe_norm2.argtypes = [vp, _f64p]
        L.sqe_norm2.restype = st
        L.sqe_norm2_noalloc.argtypes = [vp, _f64p]
        L.sqe_norm2_noalloc.restype = st
        L.sqe_export.argtypes = [vp, _u64p, _f64p, _f64p, ctypes.c_uint64, _u64p]
        L.sqe_export.restype = st
        L.sqe_export_where.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, _u64p, _f64p, _f64p, ctypes.c_uint64,
                                       _u64p]
        L.sqe_export_where.restype = st
        L.sqe_sample.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, _u64p]
        L.sqe_sample.restype = st
        L.sqe_checkpoint_save.argtypes = [vp, ctypes.c_char_p]
        L.sqe_checkpoint_save.restype = st
        L.sqe_checkpoint_load.argtypes = [vp, ctypes.c_char_p]
        L.sqe_checkpoint_load.restype = st
        L.sqe_sync.argtypes = [vp]
        L.sqe_sync.restype = st
        L.sqe_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.sqe_get_stats.restype = st
        L.sqe_snapshot_save.argtypes = [vp]
        L.sqe_snapshot_save.restype = st
        L.sqe_snapshot_restore.argtypes = [vp]
        L.sqe_snapshot_restore.restype = st
        L.sqe_last_error.argtypes = []
        L.sqe_last_error.restype = ctypes.c_char_p
        L.sqe_batch_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
        L.sqe_batch_add.argtypes = [vp, ctypes.c_uint32, _u64p, _f64p, _f64p, ctypes.c_uint64, vp,
                                    ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64)]
        L.sqe_batch_add_pinned.argtypes = [vp, ctypes.c_uint32, _u64p, _f64p, _f64p, ctypes.c_uint64, vp,
                                           ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64)]
        L.sqe_batch_step.argtypes = [vp, ctypes.c_uint32]
        L.sqe_batch_info.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(BatchCircuit)]
        L.sqe_batch_export.argtypes = [vp, ctypes.c_int64, _u64p, _f64p, _f64p, ctypes.c_uint64, _u64p]
        L.sqe_batch_remove.argtypes = [vp, ctypes.c_int64]
        L.sqe_batch_export_start.argtypes = [vp, ctypes.c_int64, _u64p, _f64p, _f64p, ctypes.c_uint64,
                                             ctypes.POINTER(ctypes.c_int64), _u64p]
        L.sqe_batch_export_poll.argtypes = [vp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _u64p,
                                            ctypes.POINTER(BatchCircuit)]
        L.sqe_batch_usage.argtypes = [vp, _u64p, _u64p, _u64p]
        L.sqe_batch_finished.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.c_uint64, _u64p]
        L.sqe_batch_get_tuning.argtypes = [vp, ctypes.POINTER(BatchTuning)]
        L.sqe_batch_set_tuning.argtypes = [vp, ctypes.POINTER(BatchTuning)]
        L.sqe_batch_autotune.argtypes = [vp, ctypes.POINTER(BatchTuning)]
        for f in ("sqe_batch_destroy", "sqe_batch_run", "sqe_batch_sync", "sqe_batch_snapshot_save",
                  "sqe_batch_snapshot_restore"):
            getattr(L, f).argtypes = [vp]
        for f in EXPORTS:
            if f.startswith("sqe_batch_") and f != "sqe_batch_last_error":
                getattr(L, f).restype = st
        L.sqe_batch_last_error.argtypes = []
        L.sqe_batch_last_error.restype = ctypes.c_char_p
        L.sqe_host_alloc.argtypes = [ctypes.c_uint64, ctypes.POINTER(vp)]
        L.sqe_host_alloc.restype = st
        L.sqe_host_free.argtypes = [vp]
        L.sqe_host_free.restype = st
        _lib = L
    return _lib


def _check(rc):
    if rc != SQE_OK:
        raise SqeError(rc, lib().sqe_last_error().decode())


def _bcheck(rc):
    if rc != SQE_OK:
        raise SqeError(rc, lib().sqe_batch_last_error().decode())


class PinnedArrays:
    """Page-locked host arrays (sqe_host_alloc) for keys / re / im of up to
    `cap` entries, reused across exports; grow() replaces them when a larger
    export comes."""

    def __init__(self, cap=0):
        self._p = None
        self.cap = 0
        self.keys = self.re = self.im = None
        self.grow(cap)

    def grow(self, cap):
        cap = int(cap)
        if cap <= self.cap and self._p is not None:
            return
        self.free()
        cap = (max(cap, 1024) + 63) & ~63          # keeps re / im 16-byte aligned
        p = ctypes.c_void_p()
        _bcheck(lib().sqe_host_alloc(3 * 8 * cap, ctypes.byref(p)))
        self._p = p
        self.cap = cap
        buf = (ctypes.c_uint8 * (3 * 8 * cap)).from_address(p.value)
        self.keys = np.frombuffer(buf, dtype=np.uint64, count=cap, offset=0)
        self.re = np.frombuffer(buf, dtype=np.float64, count=cap, offset=8 * cap)
        self.im = np.frombuffer(buf, dtype=np.float64, count=cap, offset=16 * cap)
        # the arrays' addresses as call arguments, made once (the buffer
        # does not move until grow or free)
        self.kp = ctypes.cast(p.value, _u64p)
        self.rp = ctypes.cast(p.value + 8 * cap, _f64p)
        self.ip = ctypes.cast(p.value + 16 * cap, _f64p)

    def free(self):
        if self._p is not None:
            self.keys = self.re = self.im = None
            self.kp = self.rp = self.ip = None
            lib().sqe_host_free(self._p)
            self._p = None
            self.cap = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a, t):
    return a.ctypes.data_as(t)


class Engine:
    """One sparse statevector on one GPU (owns its device memory and stream)."""

    def __init__(self, nq, prec="f64", prune_thr=0.0, bucket_bits=-1, renorm_every=0,
                 fuse_window=0, timing=False, device=0, hbm_budget_bytes=0, max_load=0.5,
                 fuse_blocks=False, renorm_mode=1, frames=True):
        L = lib()
        cfg = Config()
        L.sqe_config_default(ctypes.byref(cfg), nq)
        cfg.fuse_blocks = int(fuse_blocks) if fuse_blocks is not True else 1
        cfg.renorm_mode = int(renorm_mode)
        cfg.frame_off = 0 if frames else 1
        cfg.prec = F32 if prec in ("f32", F32) else F64
        cfg.prune_thr = prune_thr
        cfg.bucket_bits = bucket_bits
        cfg.renorm_every = renorm_every
        cfg.fuse_window = fuse_window
        cfg.timing = 1 if timing else 0
        cfg.device = device
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.max_load = max_load
        self.cfg = cfg
        self.nq = nq
        self.prec = "f32" if cfg.prec == F32 else "f64"
        h = ctypes.c_void_p()
        _check(L.sqe_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sqe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_basis(self, key=0):
        _check(lib().sqe_set_basis(self.h, int(key)))

    def load(self, keys, re, im):
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        r = np.ascontiguousarray(re, dtype=np.float64)
        i = np.ascontiguousarray(im, dtype=np.float64)
        if not (k.size == r.size == i.size):
            raise ValueError("keys/re/im sizes differ")
        _check(lib().sqe_load(self.h, _ptr(k, _u64p), _ptr(r, _f64p), _ptr(i, _f64p), k.size))

    def apply(self, gates):
        g = G.gate_array(gates)
        _check(lib().sqe_apply_gates(self.h, g.ctypes.data, g.size))

    def apply_matrix(self, target, u, cls, controls=(), ctrl_val=None):
        cm = 0
        for c in controls:
            cm |= 1 << int(c)
        cv = cm if ctrl_val is None else int(ctrl_val)
        m = np.asarray(u, dtype=np.complex128).reshape(2, 2)
        flat = np.array([m[0, 0].real, m[0, 0].imag, m[0, 1].real, m[0, 1].imag,
                         m[1, 0].real, m[1, 0].imag, m[1, 1].real, m[1, 1].imag])
        _check(lib().sqe_apply(self.h, int(target), cm, cv, _ptr(flat, _f64p), int(cls)))

    def prepare(self, gates):
        g = G.gate_array(gates)
        _check(lib().sqe_prepare(self.h, g.ctypes.data, g.size))

    def sync(self):
        _check(lib().sqe_sync(self.h))

    def nnz(self):
        n = ctypes.c_uint64()
        _check(lib().sqe_nnz(self.h, ctypes.byref(n)))
        return n.value

    def norm2(self):
        x = ctypes.c_double()
        _check(lib().sqe_norm2(self.h, ctypes.byref(x)))
        return x.value

    def export(self, sort=True, where=None):
        """Non-negligible entries as (keys, re, im); sorted by key by default.
        where=(mask, value): only keys with key & mask == value."""
        mask, value = (0, 0) if where is None else (int(where[0]), int(where[1]))
        n = ctypes.c_uint64()
        rc = lib().sqe_export_where(self.h, mask, value, None, None, None, 0, ctypes.byref(n))
        if rc not in (SQE_OK, SQE_EINVAL):
            _check(rc)
        cap = n.value
        k = np.zeros(cap, dtype=np.uint64)
        r = np.zeros(cap)
        i = np.zeros(cap)
        _check(lib().sqe_export_where(self.h, mask, value, _ptr(k, _u64p), _ptr(r, _f64p), _ptr(i, _f64p), cap,
                                      ctypes.byref(n)))
        m = n.value
        k, r, i = k[:m], r[:m], i[:m]
        if sort:
            o = np.argsort(k, kind="stable")
            k, r, i = k[o], r[o], i[o]
        return k, r, i

    def sample(self, shots, seed=0):
        """`shots` basis keys drawn with probability |a|^2 (uint64 array)."""
        out = np.zeros(int(shots), dtype=np.uint64)
        _check(lib().sqe_sample(self.h, int(shots), int(seed) & ((1 << 64) - 1), _ptr(out, _u64p)))
        return out

    def checkpoint_save(self, path):
        _check(lib().sqe_checkpoint_save(self.h, os.fsencode(path)))

    def checkpoint_load(self, path):
        _check(lib().sqe_checkpoint_load(self.h, os.fsencode(path)))

    def stats(self):
        s = Stats()
        _check(lib().sqe_get_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def snapshot_save(self):
        _check(lib().sqe_snapshot_save(self.h))

    def snapshot_restore(self):
        _check(lib().sqe_snapshot_restore(self.h))


class Batch:
    """Many independent circuits on one GPU, advanced together (sqe_batch_*)."""

    def __init__(self, prec="f64", prune_thr=0.0, bucket_bits=-1, renorm_every=0, timing=False,
                 device=0, hbm_budget_bytes=0, max_load=0.5, eager=False, frames=True, renorm_mode=1):
        """eager: run() launches every step directly instead of replaying one
        captured hipGraph of the whole run.  renorm_mode (FP32 with
        renorm_every): 1 renormalises in one pass (norm checkpoints, one
        scale at the end), 0 stops every circuit at each renormalisation
        point."""
        L = lib()
        cfg = Config()
        L.sqe_config_default(ctypes.byref(cfg), 1)
        cfg.batch_eager = 1 if eager else 0
        cfg.renorm_mode = 1 if int(renorm_mode) == 1 else 0
        cfg.frame_off = 0 if frames else 1
        cfg.prec = F32 if prec in ("f32", F32) else F64
        cfg.prune_thr = prune_thr
        cfg.bucket_bits = bucket_bits
        cfg.renorm_every = renorm_every
        cfg.timing = 1 if timing else 0
        cfg.device = device
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.max_load = max_load
        self.prec = "f32" if cfg.prec == F32 else "f64"
        h = ctypes.c_void_p()
        _bcheck(L.sqe_batch_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sqe_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, nq, keys, re, im, gates):
        """Admit one circuit; returns its id (raises SqeError SQE_ETABLE_FULL when over budget)."""
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        r = np.ascontiguousarray(re, dtype=np.float64)
        i = np.ascontiguousarray(im, dtype=np.float64)
        if not (k.size == r.size == i.size):
            raise ValueError("keys/re/im sizes differ")
        g = G.gate_array(gates)
        cid = ctypes.c_int64()
        _bcheck(lib().sqe_batch_add(self.h, int(nq), _ptr(k, _u64p), _ptr(r, _f64p), _ptr(i, _f64p), k.size,
                                    g.ctypes.data, g.size, ctypes.byref(cid)))
        return cid.value

    def add_pinned(self, nq, state, n, gates):
        """Admit a circuit whose state is the first n entries of `state`, a
        PinnedArrays (page-locked): uploaded without a host-side copy; the
        arrays must stay unchanged until the upload has completed (after the
        next sync, or once the circuit's export has completed)."""
        if n > state.cap:
            raise ValueError("n exceeds the buffer's capacity")
        g = G.gate_array(gates)
        cid = ctypes.c_int64()
        _bcheck(lib().sqe_batch_add_pinned(self.h, int(nq), state.kp, state.rp, state.ip, int(n), g.ctypes.data,
                                           g.size, ctypes.byref(cid)))
        return cid.value

    def try_add(self, nq, keys, re, im, gates, pinned=None):
        """Like add (add_pinned with `pinned` = (PinnedArrays, n)), but returns
        None when the circuit does not fit the budget now."""
        try:
            if pinned is not None:
                return self.add_pinned(nq, pinned[0], pinned[1], gates)
            return self.add(nq, keys, re, im, gates)
        except SqeError as ex:
            if ex.status == SQE_ETABLE_FULL:
                return None
            raise

    def step(self, max_gates=0):
        _bcheck(lib().sqe_batch_step(self.h, int(max_gates)))

    def run(self):
        _bcheck(lib().sqe_batch_run(self.h))

    def sync(self):
        _bcheck(lib().sqe_batch_sync(self.h))

    def info(self, cid):
        c = BatchCircuit()
        _bcheck(lib().sqe_batch_info(self.h, int(cid), ctypes.byref(c)))
        return {f: getattr(c, f) for f, _ in BatchCircuit._fields_}

    def export(self, cid, sort=True, out=None):
        """(keys, re, im) of circuit cid.  out: a PinnedArrays to export into
        (grown when too small); the result is then views of its arrays, valid
        until its next use."""
        n = ctypes.c_uint64()
        if out is not None:
            rc = lib().sqe_batch_export(self.h, int(cid), _ptr(out.keys, _u64p), _ptr(out.re, _f64p),
                                        _ptr(out.im, _f64p), out.cap, ctypes.byref(n))
            if rc == SQE_EINVAL and n.value > out.cap:
                out.grow(n.value)
                rc = lib().sqe_batch_export(self.h, int(cid), _ptr(out.keys, _u64p), _ptr(out.re, _f64p),
                                            _ptr(out.im, _f64p), out.cap, ctypes.byref(n))
            _bcheck(rc)
            k, r, i = out.keys, out.re, out.im
        else:
            rc = lib().sqe_batch_export(self.h, int(cid), None, None, None, 0, ctypes.byref(n))
            if rc not in (SQE_OK, SQE_EINVAL):
                _bcheck(rc)
            cap = n.value
            k = np.zeros(cap, dtype=np.uint64)
            r = np.zeros(cap)
            i = np.zeros(cap)
            _bcheck(lib().sqe_batch_export(self.h, int(cid), _ptr(k, _u64p), _ptr(r, _f64p), _ptr(i, _f64p), cap,
                                           ctypes.byref(n)))
        m = n.value
        k, r, i = k[:m], r[:m], i[:m]
        if sort:
            o = np.argsort(k, kind="stable")
            k, r, i = k[o], r[o], i[o]
        return k, r, i

    def remove(self, cid):
        _bcheck(lib().sqe_batch_remove(self.h, int(cid)))

    @staticmethod
    def export_buffer(cap=0):
        """A page-locked buffer for export_start."""
        return PinnedArrays(cap)

    def export_bound(self, cid):
        """Entries an export of circuit cid may need (its slot count): the
        capacity export_start asks of its buffer."""
        t = ctypes.c_int64()
        bound = ctypes.c_uint64()
        rc = lib().sqe_batch_export_start(self.h, int(cid), None, None, None, 0, ctypes.byref(t),
                                          ctypes.byref(bound))
        if rc != SQE_EINVAL or bound.value == 0:
            _bcheck(rc)
        return bound.value

    def export_start(self, cid, out):
        """Start an asynchronous export of circuit cid into `out` (a
        PinnedArrays, grown to the circuit's slot count when smaller); returns
        a ticket for export_poll.  `out` belongs to the export until then."""
        t = ctypes.c_int64()
        bound = ctypes.c_uint64()
        rc = lib().sqe_batch_export_start(self.h, int(cid), out.kp, out.rp, out.ip, out.cap, ctypes.byref(t),
                                          ctypes.byref(bound))
        if rc == SQE_EINVAL and bound.value > out.cap:
            out.grow(bound.value)
            rc = lib().sqe_batch_export_start(self.h, int(cid), out.kp, out.rp, out.ip, out.cap, ctypes.byref(t),
                                              ctypes.byref(bound))
        _bcheck(rc)
        return t.value

    def export_poll(self, ticket, wait=False):
        """None while the export runs; then (n, info): the entry count (the
        entries are out.keys[:n] etc. of its PinnedArrays) and the circuit's
        counters at export."""
        pa = self.__dict__.get("_poll_args")
        if pa is None:                              # reused by every poll (polls run often)
            pa = self._poll_args = (ctypes.c_int32(), ctypes.c_uint64(), BatchCircuit())
        done, n, c = pa
        _bcheck(lib().sqe_batch_export_poll(self.h, int(ticket), 1 if wait else 0, ctypes.byref(done),
                                            ctypes.byref(n), ctypes.byref(c)))
        if not done.value:
            return None
        return n.value, {f: getattr(c, f) for f, _ in BatchCircuit._fields_}

    def finished(self):
        """Ids of admitted circuits that have all their gates enqueued (no wait)."""
        n = ctypes.c_uint64()
        cap = 256
        while True:
            ids = (ctypes.c_int64 * cap)()
            rc = lib().sqe_batch_finished(self.h, ids, cap, ctypes.byref(n))
            if rc == SQE_OK:
                return list(ids[:n.value])
            if n.value <= cap:
                _bcheck(rc)
            cap = int(n.value)

    def usage(self):
        u, b, a = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _bcheck(lib().sqe_batch_usage(self.h, ctypes.byref(u), ctypes.byref(b), ctypes.byref(a)))
        return {"used": u.value, "budget": b.value, "active": a.value}

    def tuning(self):
        t = BatchTuning()
        _bcheck(lib().sqe_batch_get_tuning(self.h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in BatchTuning._fields_}

    def set_tuning(self, block=None, tiles_per_item=None, load_factor=None, one_launch=None):
        t = BatchTuning()
        _bcheck(lib().sqe_batch_get_tuning(self.h, ctypes.byref(t)))
        if one_launch is not None:
            t.one_launch = int(one_launch)
        if block is not None:
            t.block = block
        if tiles_per_item is not None:
            t.tiles_per_item = tiles_per_item
        if load_factor is not None:
            t.load_factor = load_factor
        _bcheck(lib().sqe_batch_set_tuning(self.h, ctypes.byref(t)))

    def snapshot_save(self):
        _bcheck(lib().sqe_batch_snapshot_save(self.h))

    def snapshot_restore(self):
        _bcheck(lib().sqe_batch_snapshot_restore(self.h))

    def autotune(self):
        t = BatchTuning()
        _bcheck(lib().sqe_batch_autotune(self.h, ctypes.byref(t)))
        return {f: getattr(t, f) for f, _ in BatchTuning._fields_}


def run(w, prec="f64", **kw):
    """Load a workload dict, apply its gates, return the sorted export."""
    with Engine(w["nq"], prec=prec, **kw) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.apply(w["gates"])
        return e.export()
