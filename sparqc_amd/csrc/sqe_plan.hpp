// Planner shared by the single-circuit engine and the batch engine: choice of
// the bucket qubits, classification of gates as bucket-local or cross-bucket,
// and encoding of bucket-local gates into sweep-program records.
#pragma once
#include <algorithm>
#include <array>
#include <cstring>
#include <complex>
#include <cstdint>
#include <vector>

#include "../../include/sqe.h"
#include "sqe_device.hpp"

namespace sqe {

inline bool kind_is_branch(uint32_t k) { return k == SQE_GATE_H || k == SQE_GATE_RX || k == SQE_GATE_RY || k == SQE_GATE_U; }
inline bool kind_is_perm(uint32_t k) { return k == SQE_GATE_X || k == SQE_GATE_Y || k == SQE_GATE_PERM; }

inline uint32_t pext_host(uint64_t v, const QMap& qm) {
    uint32_t s = 0;
    for (int j = 0; j < qm.mb; ++j) s |= (uint32_t)((v >> qm.q[j]) & 1ull) << j;
    return s;
}

inline QMap make_qmap(const std::vector<int>& qs) {
    QMap m{};
    std::vector<int> q(qs);
    std::sort(q.begin(), q.end());
    m.mb = (int)q.size();
    m.mask = 0;
    for (size_t j = 0; j < q.size(); ++j) {
        m.q[j] = (uint8_t)q[j];
        m.mask |= 1ull << q[j];
    }
    return m;
}

inline bool gate_local(const sqe_gate& g, const QMap& qm) {
    if ((qm.mask >> g.target) & 1ull) return true;
    if (kind_is_branch(g.kind)) return false;
    if (kind_is_perm(g.kind)) return (g.ctrl_mask & qm.mask) == 0;
    return true;
}

inline uint64_t count_cross(const std::vector<sqe_gate>& gl, const QMap& qm) {
    uint64_t c = 0;
    for (auto& g : gl) c += gate_local(g, qm) ? 0 : 1;
    return c;
}

// Bucket qubits: the most frequent branching targets (ties: lower qubit).
inline QMap choose_map(int nq, int want, const sqe_gate* gl, size_t ng) {
    const int mmax = 4;
    std::vector<uint64_t> cnt(nq, 0);
    for (size_t k = 0; k < ng; ++k)
        if (kind_is_branch(gl[k].kind)) cnt[gl[k].target]++;
    std::vector<int> order(nq);
    for (int i = 0; i < nq; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
    std::vector<int> pick;
    if (want < 0) {
        for (int i = 0; i < nq && (int)pick.size() < mmax; ++i)
            if (cnt[order[i]] > 0) pick.push_back(order[i]);
    } else {
        want = std::min(want, std::min(5, nq));
        for (int i = 0; i < nq && (int)pick.size() < want; ++i) pick.push_back(order[i]);
    }
    return make_qmap(pick);
}

inline DOp plan_op(const sqe_gate& g, const QMap& qm) {
    DOp d{};
    d.kind = g.kind;
    const bool tin = (qm.mask >> g.target) & 1ull;
    d.tslot = -1;
    d.tbase = 0;
    if (tin) {
        for (int j = 0; j < qm.mb; ++j)
            if (qm.q[j] == g.target) d.tslot = j;
    } else {
        d.tbase = 1ull << g.target;
    }
    d.cmB = g.ctrl_mask & ~qm.mask;
    d.cvB = g.ctrl_val & ~qm.mask;
    d.cmS = pext_host(g.ctrl_mask, qm);
    d.cvS = pext_host(g.ctrl_val, qm);
    for (int k = 0; k < 8; ++k) d.p[k] = g.p[k];
    return d;
}

// A Pauli gate on a bucket qubit that the sweep applies to the per-bucket
// frame (sqe_device.hpp) as a base record: X / Y / Z on bucket qubit j with
// controls on base qubits only (or none), or Z on a base qubit b with exactly
// one bucket-qubit control q_j -- a CZ, which acts on buckets whose base bit b
// is 1 as Z_j (control value 1) or -Z_j (control value 0).  With the frame
// logical = g Z^fz X^fx (registers):
//   X_j: fx ^= e_j and g *= (-1)^fz_j    (factor selected by the fz_j bit)
//   Z_j: fz ^= e_j
//   Y_j = i X_j Z_j: fz ^= e_j, fx ^= e_j, g *= (fz_j ? i : -i) (fz_j before)
// Returns false when the gate is not of this form.
inline bool frame_record(const DOp& d, const QMap& qm, bool f32, MOp* out) {
    int j = -1;
    uint64_t cm = d.cmB, cv = d.cvB;
    double f0r = 1, f0i = 0, f1r = 1, f1i = 0;
    uint64_t flip = 0;
    if (d.tslot >= 0 && d.cmS == 0 && (d.kind == KX || d.kind == KY || d.kind == KZ)) {
        j = d.tslot;
        const uint64_t xb = 1ull << qm.q[j], zb = 1ull << (FRAME_Z_BIT0 + j);
        if (d.kind == KX) { flip = xb; f1r = -1; }
        else if (d.kind == KZ) { flip = zb; }
        else { flip = xb | zb; f0r = 0; f0i = -1; f1r = 0; f1i = 1; }
    } else if (d.tslot < 0 && d.kind == KZ && d.cmS != 0 && (d.cmS & (d.cmS - 1)) == 0) {
        j = __builtin_ctz(d.cmS);
        cm |= d.tbase;
        cv |= d.tbase;
        flip = 1ull << (FRAME_Z_BIT0 + j);
        if (!(d.cvS & d.cmS)) { f0r = -1; f1r = -1; }
    } else {
        return false;
    }
    MOp m{};
    m.hdr = H_BASE_PERM | (uint32_t)(FRAME_Z_BIT0 + j) << 7;
    m.ctl = 0;
    m.cmB = cm;
    m.cvB = cv;
    if (f32) {
        const float f[4] = {(float)f0r, (float)f0i, (float)f1r, (float)f1i};
        std::memcpy(&m.c[0], f, 16);
    } else {
        m.c[0] = f0r; m.c[1] = f0i; m.c[2] = f1r; m.c[3] = f1i;
    }
    m.c[4] = coef_index_bits(flip);
    *out = m;
    return true;
}

// Pack a planned op into a 64-byte sweep record: pick the handler and turn the
// gate's coefficients into the handler's (see the handler table in
// sqe_device.hpp).  General U keeps its eight coefficients in the side table.
// f32: the program runs on FP32 amplitudes; a base-qubit op (whose factor
// the sweep folds in FP32 there) then carries c0 / c1 as FP32 pairs in the
// first two coefficient slots: c[0] = (f0r, f0i), c[1] = (f1r, f1i).
// qm non-null: Pauli gates of the frame_record form become frame records
// (*framed is set when one does).
inline MOp encode_op(const DOp& d, std::vector<double>& coef, bool f32 = false, const QMap* qm = nullptr,
                     bool* framed = nullptr) {
    if (qm) {
        MOp fr;
        if (frame_record(d, *qm, f32, &fr)) {
            if (framed) *framed = true;
            return fr;
        }
    }
    MOp m{};
    const bool slot = d.tslot >= 0;
    const uint32_t j = slot ? (uint32_t)d.tslot : 0u;
    const uint32_t target = (!slot && d.tbase) ? (uint32_t)__builtin_ctzll(d.tbase) : 0u;
    uint32_t h = H_RENORM;
    uint64_t ci = 0;
    bool side = false;
    const double* p = d.p;
    auto set = [&](double a, double b, double c, double e, double f) {
        m.c[0] = a; m.c[1] = b; m.c[2] = c; m.c[3] = e; m.c[4] = f;
    };
    switch (d.kind) {
    case SQE_GATE_Z: set(1, 0, -1, 0, 0); h = slot ? H_SLOT_DIAG : H_BASE_DIAG; break;
    case SQE_GATE_PHASE: set(1, 0, p[0], p[1], 0); h = slot ? H_SLOT_DIAG : H_BASE_DIAG; break;
    case SQE_GATE_RZ: set(p[0], -p[1], p[0], p[1], 0); h = slot ? H_SLOT_DIAG : H_BASE_DIAG; break;
    case SQE_GATE_DIAG: set(p[0], p[1], p[2], p[3], 0); h = slot ? H_SLOT_DIAG : H_BASE_DIAG; break;
    case SQE_GATE_X: set(1, 0, 1, 0, 0); h = slot ? H_SLOT_X : H_BASE_X; break;           // base: factor 1
    case SQE_GATE_Y: set(0, 1, 0, -1, 0); h = slot ? H_SLOT_PERM : H_BASE_PERM; break;      // u10 = i, u01 = -i
    case SQE_GATE_PERM: set(p[2], p[3], p[0], p[1], 0); h = slot ? H_SLOT_PERM : H_BASE_PERM; break;
    case SQE_GATE_H: set(p[0], 0, 0, 0, 0); h = H_SLOT_H; break;
    case SQE_GATE_RX: {
        // the general handler: [[c, -is], [-is, c]] as 8 coefficients
        const double u[8] = {p[0], 0.0, 0.0, -p[1], 0.0, -p[1], p[0], 0.0};
        h = H_SLOT_U;
        ci = coef.size() / 8;
        side = true;
        coef.insert(coef.end(), u, u + 8);
    } break;
    case SQE_GATE_RY: set(p[0], p[1], 0, 0, 0); h = H_SLOT_RY; break;
    case SQE_GATE_U:
        h = H_SLOT_U;
        ci = coef.size() / 8;
        side = true;
        coef.insert(coef.end(), p, p + 8);
        break;
    default: h = H_RENORM; break;                                                           // K_RENORM
    }
    if (f32 && h <= H_BASE_PERM) {
        const float f[4] = {(float)m.c[0], (float)m.c[1], (float)m.c[2], (float)m.c[3]};
        m.c[0] = m.c[1] = m.c[2] = m.c[3] = 0.0;
        std::memcpy(&m.c[0], f, 16);
    }
    m.hdr = h | j << 4 | target << 7;
    if (side) m.c[4] = coef_index_bits(ci);
    if (h <= H_BASE_PERM) m.c[4] = coef_index_bits((h == H_BASE_X || h == H_BASE_PERM) ? d.tbase : 0ull);
    m.ctl = (d.cmS & 0xffu) | (d.cvS & 0xffu) << 8;
    m.cmB = d.cmB;
    m.cvB = d.cvB;
    return m;
}

// ---------------------------------------------------------------------------
// Fused blocks (FP32, 16- or 32-slot buckets): a window of gates on bucket
// qubits without controls is multiplied into one dense D x D complex matrix
// (D = 2^MB: 16x16 for 4 bucket qubits, 32x32 for 5) that the sweep applies
// with MFMA.  Gates on base qubits without bucket controls commute with such
// gates (they only scale or relabel whole buckets) and are emitted ahead of
// the window; any other gate closes it.  A window of fewer than min_gates
// gates runs gate by gate: a dense block costs 4 D^2 MFMA FMAs per bucket
// whatever it holds, so it pays only for long windows.  Prune decisions are
// made once, on the block's outputs.
// ---------------------------------------------------------------------------
using cx = std::complex<double>;
using BlockMat = std::vector<cx>;                // D x D, row-major, out x in

// 2x2 matrix of a gate record, row-major (the conventions of sqe.h).
inline void gate_matrix(const sqe_gate& g, cx m[4]) {
    const double* p = g.p;
    const cx i1(0.0, 1.0);
    switch (g.kind) {
    case SQE_GATE_X: m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; break;
    case SQE_GATE_Y: m[0] = 0; m[1] = -i1; m[2] = i1; m[3] = 0; break;
    case SQE_GATE_Z: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = -1; break;
    case SQE_GATE_H: m[0] = p[0]; m[1] = p[0]; m[2] = p[0]; m[3] = -p[0]; break;
    case SQE_GATE_PHASE: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = cx(p[0], p[1]); break;
    case SQE_GATE_RZ: m[0] = cx(p[0], -p[1]); m[1] = 0; m[2] = 0; m[3] = cx(p[0], p[1]); break;
    case SQE_GATE_RX: m[0] = p[0]; m[1] = cx(0, -p[1]); m[2] = cx(0, -p[1]); m[3] = p[0]; break;
    case SQE_GATE_RY: m[0] = p[0]; m[1] = -p[1]; m[2] = p[1]; m[3] = p[0]; break;
    case SQE_GATE_U: for (int k = 0; k < 4; ++k) m[k] = cx(p[2 * k], p[2 * k + 1]); break;
    case SQE_GATE_DIAG: m[0] = cx(p[0], p[1]); m[1] = 0; m[2] = 0; m[3] = cx(p[2], p[3]); break;
    case SQE_GATE_PERM: m[0] = 0; m[1] = cx(p[0], p[1]); m[2] = cx(p[2], p[3]); m[3] = 0; break;
    default: m[0] = 1; m[1] = 0; m[2] = 0; m[3] = 1; break;
    }
}

constexpr uint32_t K_BLOCK = 17;                  // engine-internal op: fused dense block
constexpr uint32_t K_NORM_POINT = 18;             // engine-internal op: norm checkpoint p[0]

// The planned op becomes a frame record (see frame_record).
inline bool is_frame_op(const DOp& d) {
    if (d.tslot >= 0) return d.cmS == 0 && (d.kind == KX || d.kind == KY || d.kind == KZ);
    return d.kind == KZ && d.cmS != 0 && (d.cmS & (d.cmS - 1)) == 0;
}

// The op reads or writes the amplitude registers (so a pending frame must be
// applied first): gates on bucket qubits, base-qubit gates restricted by
// bucket controls and fused blocks.  Base-class gates, renormalisations and
// norm points only touch the base key and the pending factor.
inline bool touches_amplitudes(const DOp& d) {
    if (d.kind == K_BLOCK) return true;
    if (d.kind == K_RENORM || d.kind == K_NORM_POINT) return false;
    return d.tslot >= 0 || d.cmS != 0;
}

// Fused dense blocks (FP32, 16- or 32-slot buckets) built from a program's
// gates in order.  Plain gates on bucket qubits (no controls) go into the
// open window, whose product is one D x D matrix (D = 2^MB).  Base-class gates commute with the
// window (they scale or relabel whole buckets) and are emitted ahead of it.
// Frame gates (Pauli gates on a bucket qubit whose controls are base
// qubits, see frame_record) act on one bucket qubit with a per-bucket
// condition, so they cannot enter a shared matrix; they are held back and
// emitted after the block, which is exact as long as the window's later gates
// commute with them: a held gate on qubit j blocks only later window gates on
// j that it does not commute with (a Z-type gate commutes with diagonal
// gates, an X-type gate with X), and later base gates that flip one of its
// control bits.  Any other gate closes the window.  A window of fewer than
// min_gates gates runs gate by gate; prune decisions are made once, on the
// block's outputs.
struct BlockBuilder {
    bool on = false;                              // fusion enabled for this program
    bool frames = false;                          // frame gates may be held back past the window
    uint32_t min_gates = 6;                       // smaller windows run gate by gate
    QMap qm{};
    BlockMat m;                                   // D x D, row-major, out x in
    uint32_t count = 0;
    std::vector<DOp> window;                      // the window's gates, in order
    std::vector<DOp> held;                        // frame gates held back behind the window
    uint32_t held_x = 0, held_z = 0;              // bucket slots with held X-type / Z-type gates
    uint64_t held_ctl = 0;                        // base bits the held gates' conditions read
    std::vector<BlockMat>* blocks = nullptr;

    static bool plain_slot(const sqe_gate& g, const QMap& qm) {
        return ((qm.mask >> g.target) & 1ull) && g.ctrl_mask == 0;
    }
    static bool commutes(const sqe_gate& g, const QMap& qm) {   // base-class gate
        return !((qm.mask >> g.target) & 1ull) && (g.ctrl_mask & qm.mask) == 0;
    }
    bool open() const { return count > 0 || !held.empty(); }
    // the window can take plain slot gate g past the held gates
    bool passes_held(const sqe_gate& g, int j) const {
        const bool diag = g.kind == KZ || g.kind == KRZ || g.kind == KPHASE || g.kind == KDIAG;
        if ((held_z >> j) & 1u) { if (!diag) return false; }
        if ((held_x >> j) & 1u) { if (g.kind != KX) return false; }
        return true;
    }
    // base-class gate g may be emitted ahead of the held gates
    bool base_passes_held(const sqe_gate& g) const {
        const bool flips = g.kind == KX || g.kind == KY || g.kind == KPERM;
        return !(flips && ((held_ctl >> g.target) & 1ull));
    }
    void hold(const DOp& d) {
        int j = d.tslot >= 0 ? d.tslot : __builtin_ctz(d.cmS);
        if (d.kind == KX || d.kind == KY) held_x |= 1u << j;
        if (d.kind == KZ || d.kind == KY) held_z |= 1u << j;
        held_ctl |= d.cmB | d.tbase;
        held.push_back(d);
    }
    int slot_of(const sqe_gate& g) const {
        for (int k = 0; k < qm.mb; ++k)
            if (qm.q[k] == g.target) return k;
        return 0;
    }
    void add(const sqe_gate& g) {
        const int D = 1 << qm.mb;
        if (count == 0) {
            m.assign((size_t)D * D, cx(0.0, 0.0));
            for (int r = 0; r < D; ++r) m[(size_t)r * D + r] = 1.0;
            window.clear();
        }
        window.push_back(plan_op(g, qm));
        const int j = slot_of(g);
        cx u[4];
        gate_matrix(g, u);
        BlockMat out((size_t)D * D);
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) {
                // (G M)[r][c] = sum over s of G[r][s] M[s][c]; G mixes r with r ^ 2^j
                const int r0 = r & ~(1 << j), r1 = r | (1 << j), br = (r >> j) & 1;
                out[(size_t)r * D + c] = u[br * 2] * m[(size_t)r0 * D + c] + u[br * 2 + 1] * m[(size_t)r1 * D + c];
            }
        m.swap(out);
        ++count;
    }
    void flush(std::vector<DOp>& prog) {
        if (count > 0 && count < min_gates) prog.insert(prog.end(), window.begin(), window.end());
        if (count >= min_gates) {
            DOp d{};
            d.kind = K_BLOCK;
            d.tslot = -1;
            d.p[0] = (double)blocks->size();
            d.p[1] = (double)count;
            blocks->push_back(m);
            prog.push_back(d);
        }
        count = 0;
        prog.insert(prog.end(), held.begin(), held.end());
        held.clear();
        held_x = held_z = 0;
        held_ctl = 0;
    }
    // Route one bucket-local gate: into the window, held back, ahead of the
    // window, or (after closing it) straight into the program.
    void route(const sqe_gate& g, std::vector<DOp>& prog) {
        if (plain_slot(g, qm)) {
            if (!passes_held(g, slot_of(g))) flush(prog);
            add(g);
            return;
        }
        const DOp d = plan_op(g, qm);
        if (commutes(g, qm)) {
            if (!base_passes_held(g)) flush(prog);
            prog.push_back(d);                     // ahead of the open window
            return;
        }
        if (frames && is_frame_op(d)) {
            hold(d);
            return;
        }
        flush(prog);
        prog.push_back(d);
    }
};

// A program encoded into sweep records.  start[i] is the first record of
// planned op i (start[n] = the record count), so a run of planned ops maps to
// a run of records.  With frames, Pauli gates on bucket qubits become frame
// records and an H_FRAME record is put before the next op that touches the
// amplitude registers while frame records are outstanding.
struct Encoded {
    std::vector<MOp> ops;
    std::vector<double> coef;
    std::vector<size_t> start;
    bool framed = false;
};

// Record of a block (D x D complex, D = 16 or 32): the side table gets the
// 2D x 2D real form R = [[Mr, -Mi], [Mi, Mr]] as the A operands of the
// v_mfma_f32_32x32x2f32 chains -- D / 16 tiles of 32 output rows, D steps of
// two input rows each (tile q, step t, lane l: R[32 q + l % 32][2t + l / 32],
// FP32) -- then M itself in FP64 (row-major re, im) for the FP64
// recomputation of near-threshold outputs.  The side table takes 64 D^2 / 8
// units of 8 doubles: 128 for 16x16, 512 for 32x32.
inline MOp encode_block(const DOp& d, const BlockMat& m, std::vector<double>& coef) {
    MOp r{};
    const int D = m.size() == 1024 ? 32 : 16;
    const uint64_t ci = coef.size() / 8;
    const int tiles = D / 16, steps = D;
    std::vector<float> tab((size_t)tiles * steps * 64);
    auto R = [&](int i, int k) -> double {
        const cx v = m[(size_t)(i % D) * D + (k % D)];
        if (i < D) return k < D ? v.real() : -v.imag();
        return k < D ? v.imag() : v.real();
    };
    for (int q = 0; q < tiles; ++q)
        for (int t = 0; t < steps; ++t)
            for (int l = 0; l < 64; ++l)
                tab[((size_t)q * steps + t) * 64 + l] = (float)R(32 * q + l % 32, 2 * t + l / 32);
    const size_t at = coef.size();
    coef.resize(at + tab.size() / 2);
    std::memcpy(&coef[at], tab.data(), tab.size() * sizeof(float));
    for (size_t k = 0; k < m.size(); ++k) {
        coef.push_back(m[k].real());
        coef.push_back(m[k].imag());
    }
    r.hdr = H_SLOT_BLOCK;
    r.ctl = 0;
    r.c[0] = d.p[1];                               // gates in the block
    r.c[4] = coef_index_bits(ci);
    return r;
}

inline DOp renorm_op() {
    DOp d{};
    d.kind = K_RENORM;
    d.tslot = -1;
    return d;
}

inline DOp norm_point_op(uint32_t k) {
    DOp d{};
    d.kind = K_NORM_POINT;
    d.tslot = -1;
    d.p[0] = (double)k;
    return d;
}

inline MOp frame_apply_record() {
    MOp m{};
    m.hdr = H_FRAME;
    return m;
}

// Encode planned ops (blocks index into `blocks`); qm non-null enables frames.
inline void encode_program(const std::vector<DOp>& prog, const std::vector<BlockMat>& blocks, bool f32,
                           const QMap* qm, Encoded& out) {
    out.ops.clear();
    out.coef.clear();
    out.start.assign(prog.size() + 1, 0);
    out.framed = false;
    out.ops.reserve(prog.size() + 8);
    bool outstanding = false;
    for (size_t i = 0; i < prog.size(); ++i) {
        const DOp& d = prog[i];
        out.start[i] = out.ops.size();
        if (qm && outstanding && touches_amplitudes(d) && !is_frame_op(d)) {
            out.ops.push_back(frame_apply_record());
            outstanding = false;
        }
        if (d.kind == K_BLOCK) {
            out.ops.push_back(encode_block(d, blocks[(size_t)d.p[0]], out.coef));
        } else if (d.kind == K_NORM_POINT) {
            MOp m{};
            m.hdr = H_NORM_POINT | (uint32_t)d.p[0] << 13;
            out.ops.push_back(m);
        } else {
            bool fr = false;
            out.ops.push_back(encode_op(d, out.coef, f32, qm, &fr));
            if (fr) outstanding = out.framed = true;
        }
    }
    out.start[prog.size()] = out.ops.size();
}

}  // namespace sqe
