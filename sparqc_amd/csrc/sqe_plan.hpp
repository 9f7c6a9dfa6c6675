// Planner shared by the single-circuit engine and the batch engine: choice of
// the bucket qubits, classification of gates as bucket-local or cross-bucket,
// and encoding of bucket-local gates into sweep-program records.
#pragma once
#include <algorithm>
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
inline MOp encode_op(const DOp& d, std::vector<double>& coef) {
    MOp m{};
    const bool slot = d.tslot >= 0;
    const uint32_t j = slot ? (uint32_t)d.tslot : 0u;
    const uint32_t target = (!slot && d.tbase) ? (uint32_t)__builtin_ctzll(d.tbase) : 0u;
    uint32_t h = H_RENORM, ci = NO_COEF;
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
    case SQE_GATE_RX: set(p[0], 0, -p[1], 0, -p[1]); h = H_SLOT_U2; break;
    case SQE_GATE_RY: set(p[0], -p[1], 0, p[1], 0); h = H_SLOT_U2; break;
    case SQE_GATE_U:
        h = H_SLOT_U;
        ci = (uint32_t)(coef.size() / 8);
        coef.insert(coef.end(), p, p + 8);
        break;
    default: h = H_RENORM; break;                                                           // K_RENORM
    }
    m.hdr = h | j << 4 | target << 7 | ci << 13;
    m.ctl = (d.cmS & 0xffu) | (d.cvS & 0xffu) << 8;
    m.cmB = d.cmB;
    m.cvB = d.cvB;
    return m;
}

inline DOp renorm_op() {
    DOp d{};
    d.kind = K_RENORM;
    d.tslot = -1;
    return d;
}

}  // namespace sqe
