"""Gate records shared by the engine, the generators and the tests.

A gate is one 88-byte record (``GATE_DTYPE``): kind, target qubit, control mask,
required control values and eight host-computed FP64 coefficients.  The layout is
the ``sqe_gate`` struct of ``include/sqe.h``.

Kinds and coefficient conventions follow the reference's gate set and its angle
conventions:

* ``X, Y, Z, H`` -- ``Circuit__::x/y/z/h`` (emulator/include/circuit.h:45-63);
  H uses ``f1 = 1./sqrt(2.)`` (emulator/device/device_data.cuh:205).
* ``PHASE(theta)`` -- ``diag(1, e^{i theta})``; the reference passes theta
  unhalved (emulator/device/table_manager.cuh:333).
* ``RZ/RX/RY(theta)`` -- the reference halves theta before the kernel
  (emulator/device/table_manager.cuh:334-336); p = (cos(theta/2), sin(theta/2)).
* ``U`` -- a general 2x2 complex matrix (row-major u00,u01,u10,u11),
  ``DIAG`` -- a general diagonal (d0, d1) and ``PERM`` -- an anti-diagonal
  [[0, u01], [u10, 0]]; these are additions of this engine (controlled gates
  and general unitaries are absent from the reference).

Controls: a gate acts only on basis keys k with ``(k & ctrl_mask) == ctrl_val``.
CX, CZ and CCX are X and Z with one or two controls.
"""
import math

import numpy as np

X, Y, Z, H, PHASE, RZ, RX, RY, U, DIAG, PERM = range(11)
KIND_NAMES = ["X", "Y", "Z", "H", "PHASE", "RZ", "RX", "RY", "U", "DIAG", "PERM"]

# gate classes (sqe_class in include/sqe.h)
CLS_DIAG, CLS_PERM, CLS_BRANCH = 0, 1, 2
KIND_CLASS = {X: CLS_PERM, Y: CLS_PERM, Z: CLS_DIAG, H: CLS_BRANCH,
              PHASE: CLS_DIAG, RZ: CLS_DIAG, RX: CLS_BRANCH, RY: CLS_BRANCH,
              U: CLS_BRANCH, DIAG: CLS_DIAG, PERM: CLS_PERM}

GATE_DTYPE = np.dtype([("kind", "<u4"), ("target", "<u4"), ("ctrl_mask", "<u8"),
                       ("ctrl_val", "<u8"), ("p", "<f8", (8,))])
assert GATE_DTYPE.itemsize == 88

INV_SQRT2 = 1.0 / math.sqrt(2.0)


def _ctrl(controls):
    m = 0
    for c in controls:
        m |= 1 << int(c)
    return m


def make_gate(kind, target, controls=(), theta=None, matrix=None, ctrl_val=None):
    """Return one gate record (a numpy void of GATE_DTYPE).

    ``matrix``: the 2x2 matrix for U, (d0, d1) for DIAG, (u01, u10) for PERM.

    ``theta`` is the full rotation angle, as in the reference API; halving for
    RZ/RX/RY is done here on the host (emulator/device/table_manager.cuh:334-336).
    """
    g = np.zeros((), dtype=GATE_DTYPE)
    g["kind"] = kind
    g["target"] = target
    cm = _ctrl(controls)
    if cm >> target & 1:
        raise ValueError("target qubit is also a control")
    g["ctrl_mask"] = cm
    g["ctrl_val"] = cm if ctrl_val is None else int(ctrl_val)
    p = np.zeros(8)
    if kind == H:
        p[0] = INV_SQRT2
    elif kind == PHASE:
        p[0], p[1] = math.cos(theta), math.sin(theta)
    elif kind in (RZ, RX, RY):
        p[0], p[1] = math.cos(theta / 2.0), math.sin(theta / 2.0)
    elif kind == U:
        m = np.asarray(matrix, dtype=np.complex128).reshape(2, 2)
        p[:] = [m[0, 0].real, m[0, 0].imag, m[0, 1].real, m[0, 1].imag,
                m[1, 0].real, m[1, 0].imag, m[1, 1].real, m[1, 1].imag]
    elif kind in (DIAG, PERM):
        d = np.asarray(matrix, dtype=np.complex128).reshape(2)
        p[:4] = [d[0].real, d[0].imag, d[1].real, d[1].imag]
    g["p"] = p
    return g


def gate_array(gates):
    """Stack gate records into a contiguous GATE_DTYPE array."""
    if isinstance(gates, np.ndarray) and gates.dtype == GATE_DTYPE:
        return np.ascontiguousarray(gates.reshape(-1))
    out = np.zeros(len(gates), dtype=GATE_DTYPE)
    for i, g in enumerate(gates):
        out[i] = g
    return out


def gate_class(kind):
    return KIND_CLASS[int(kind)]


def matrix_of(g):
    """2x2 complex matrix a gate record applies (for dense references in tests)."""
    k = int(g["kind"])
    p = g["p"]
    if k == X:
        return np.array([[0, 1], [1, 0]], dtype=complex)
    if k == Y:
        return np.array([[0, -1j], [1j, 0]], dtype=complex)
    if k == Z:
        return np.array([[1, 0], [0, -1]], dtype=complex)
    if k == H:
        return p[0] * np.array([[1, 1], [1, -1]], dtype=complex)
    if k == PHASE:
        return np.array([[1, 0], [0, p[0] + 1j * p[1]]])
    if k == RZ:
        return np.array([[p[0] - 1j * p[1], 0], [0, p[0] + 1j * p[1]]])
    if k == RX:
        return np.array([[p[0], -1j * p[1]], [-1j * p[1], p[0]]])
    if k == RY:
        return np.array([[p[0], -p[1]], [p[1], p[0]]], dtype=complex)
    if k == U:
        return np.array([[p[0] + 1j * p[1], p[2] + 1j * p[3]],
                         [p[4] + 1j * p[5], p[6] + 1j * p[7]]])
    if k == DIAG:
        return np.array([[p[0] + 1j * p[1], 0], [0, p[2] + 1j * p[3]]])
    if k == PERM:
        return np.array([[0, p[0] + 1j * p[1]], [p[2] + 1j * p[3], 0]])
    raise ValueError(k)


def inverse_gate(g):
    """The inverse of one gate record (used by circuit-then-inverse tests)."""
    k = int(g["kind"])
    h = g.copy()
    p = g["p"]
    if k in (X, Y, Z, H):
        return h
    if k in (PHASE, RZ, RX, RY):
        q = p.copy()
        q[1] = -p[1]
        h["p"] = q
        return h
    m = matrix_of(g).conj().T
    if k == U:
        h["p"] = [m[0, 0].real, m[0, 0].imag, m[0, 1].real, m[0, 1].imag,
                  m[1, 0].real, m[1, 0].imag, m[1, 1].real, m[1, 1].imag]
    elif k == DIAG:
        h["p"] = [m[0, 0].real, m[0, 0].imag, m[1, 1].real, m[1, 1].imag, 0, 0, 0, 0]
    else:
        h["p"] = [m[0, 1].real, m[0, 1].imag, m[1, 0].real, m[1, 0].imag, 0, 0, 0, 0]
    return h
