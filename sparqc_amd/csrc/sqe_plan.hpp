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

// Pack a planned op into a 64-byte sweep record: pick the handler and turn the
// gate's coefficients into the handler's (see the handler table in
// sqe_device.hpp).  General U keeps its eight coefficients in the side table.
// f32: the program runs on FP32 amplitudes; a base-qubit op (whose factor
// the sweep folds in FP32 there) then carries c0 / c1 as FP32 pairs in the
// first two coefficient slots: c[0] = (f0r, f0i), c[1] = (f1r, f1i).
inline MOp encode_op(const DOp& d, std::vector<double>& coef, bool f32 = false) {
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
    m.ctl = (d.cmS & 0xffu) | (d.cvS & 0xffu) << 8;
    m.cmB = d.cmB;
    m.cvB = d.cvB;
    return m;
}

// ---------------------------------------------------------------------------
// Fused blocks (FP32, 16-slot buckets): a window of gates on bucket qubits
// without controls is multiplied into one dense 16x16 complex matrix that the
// sweep applies with MFMA.  Gates on base qubits without bucket controls
// commute with such gates (they only scale or relabel whole buckets) and are
// emitted ahead of the window; any other gate closes it.  A window
// of fewer than min_gates gates runs gate by gate: a dense block costs 2048
// MFMA FMAs per bucket whatever it holds, so it pays only for long windows.
// Prune decisions are made once, on the block's outputs.
// ---------------------------------------------------------------------------
using cx = std::complex<double>;

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

struct BlockBuilder {
    bool on = false;                              // fusion enabled for this program
    uint32_t min_gates = 6;                       // smaller windows run gate by gate
    QMap qm{};
    std::array<cx, 256> m{};                      // 16x16, row-major, out x in
    uint32_t count = 0;
    std::vector<DOp> window;                      // the window's gates, in order
    std::vector<std::array<cx, 256>>* blocks = nullptr;

    static bool plain_slot(const sqe_gate& g, const QMap& qm) {
        return ((qm.mask >> g.target) & 1ull) && g.ctrl_mask == 0;
    }
    static bool commutes(const sqe_gate& g, const QMap& qm) {   // base-class gate
        return !((qm.mask >> g.target) & 1ull) && (g.ctrl_mask & qm.mask) == 0;
    }
    void add(const sqe_gate& g) {
        if (count == 0) {
            for (int r = 0; r < 16; ++r)
                for (int c = 0; c < 16; ++c) m[r * 16 + c] = r == c ? 1.0 : 0.0;
            window.clear();
        }
        window.push_back(plan_op(g, qm));
        int j = 0;
        for (int k = 0; k < qm.mb; ++k)
            if (qm.q[k] == g.target) j = k;
        cx u[4];
        gate_matrix(g, u);
        std::array<cx, 256> out{};
        for (int r = 0; r < 16; ++r)
            for (int c = 0; c < 16; ++c) {
                // (G M)[r][c] = sum over s of G[r][s] M[s][c]; G mixes r with r ^ 2^j
                const int r0 = r & ~(1 << j), r1 = r | (1 << j), br = (r >> j) & 1;
                out[r * 16 + c] = u[br * 2] * m[r0 * 16 + c] + u[br * 2 + 1] * m[r1 * 16 + c];
            }
        m = out;
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
    }
};

// Record of a block: the side table gets the 32x32 real form R = [[Mr, -Mi],
// [Mi, Mr]] as the A operands of 16 v_mfma_f32_32x32x2f32 steps (step t, lane
// l: R[l % 32][2t + l / 32], FP32), then M itself in FP64 (row-major re, im)
// for the FP64 recomputation of near-threshold outputs.
inline MOp encode_block(const DOp& d, const std::array<cx, 256>& m, std::vector<double>& coef) {
    MOp r{};
    const uint64_t ci = coef.size() / 8;
    std::vector<float> tab(1024);
    auto R = [&](int i, int k) -> double {
        const cx v = m[(i & 15) * 16 + (k & 15)];
        if (i < 16) return k < 16 ? v.real() : -v.imag();
        return k < 16 ? v.imag() : v.real();
    };
    for (int t = 0; t < 16; ++t)
        for (int l = 0; l < 64; ++l) tab[t * 64 + l] = (float)R(l % 32, 2 * t + l / 32);
    const size_t at = coef.size();
    coef.resize(at + 512);
    std::memcpy(&coef[at], tab.data(), 1024 * sizeof(float));
    for (int k = 0; k < 256; ++k) {
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

}  // namespace sqe
