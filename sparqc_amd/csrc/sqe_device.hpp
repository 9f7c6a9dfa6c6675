// Device-side building blocks of the sparse statevector engine (gfx950).
//
// Data model.  The state is a set of buckets.  A bucket holds the amplitudes of
// the 2^MB basis keys that share every key bit outside the bucket-qubit set M
// (|M| = MB <= 5); slot s of a bucket is the key whose M-bits, read in
// increasing qubit order, spell s.  A key absent from the state has amplitude 0.
// Buckets are located by an open-addressing hash index (base key -> bucket id).
// With MB = 0 a bucket is a single key and the table is a plain key -> amplitude
// map, the reference's model (emulator/device/table.cuh:13-26).
//
// Layout in HBM (structure of arrays, tiles of 64 buckets = one wavefront):
//   base[b]                                  uint64 base key, ~0 = padding
//   amp[((b >> 6) * S + s) * 64 + (b & 63)]  complex (double2 / float2)
// so that thread-per-bucket sweeps issue fully coalesced 64-lane loads for
// every slot.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace sqe {

constexpr uint64_t EMPTY_KEY = ~0ull;
constexpr uint32_t NONE = 0xffffffffu;
constexpr int TILE = 64;          // buckets per tile (one wavefront)
constexpr int BLOCK = 256;        // threads per workgroup for bucket kernels
// Batched sweeps of circuits of 0..SQE_MIXED_MAX_MB bucket qubits go out as
// one launch (k_sweep_batch_mixed, launch id MB_MIXED).
#ifndef SQE_MIXED_MAX_MB
#define SQE_MIXED_MAX_MB 3
#endif
constexpr int MB_MIXED = 8;

// gate kinds (sqe_gate.kind); K_RENORM is an engine-internal op
enum : uint32_t { KX = 0, KY, KZ, KH, KPHASE, KRZ, KRX, KRY, KU, KDIAG, KPERM, K_RENORM = 16 };

__host__ __device__ constexpr bool is_branch(uint32_t k) {
    return k == KH || k == KRX || k == KRY || k == KU;
}
__host__ __device__ constexpr bool is_perm(uint32_t k) { return k == KX || k == KY || k == KPERM; }

// One gate after planning against the bucket-qubit set.
struct DOp {
    uint32_t kind;
    int32_t tslot;      // slot bit of the target when it is a bucket qubit, else -1
    uint64_t tbase;     // 1 << target when the target is not a bucket qubit, else 0
    uint64_t cmB, cvB;  // controls on base bits: (base & cmB) == cvB
    uint32_t cmS, cvS;  // controls on slot bits: (s & cmS) == cvS
    double p[8];        // FP64 coefficients (host computed)
};

// Sweep programs: each gate is pre-classified on the host into one of a few
// handlers whose coefficients already encode the gate kind, so the kernel
// dispatches on a handful of cases instead of kind x target-bit x controls.
//   H_BASE_DIAG  diagonal gate on a base qubit: factor f0 / f1 by the base bit
//   H_BASE_X     X on a base qubit: flip the bucket's base key (factors 1, 1)
//   H_BASE_PERM  Y / PERM on a base qubit: factor by the base bit, then flip
//                (without bucket controls these three only multiply the
//                bucket's pending FP64 factor, which is applied to the slots
//                before the next branching gate and before the store)
//   H_SLOT_DIAG  diagonal gate on bucket qubit J: factor f0 / f1 by slot bit J
//   H_SLOT_X     X on bucket qubit J: swap slot pairs
//   H_SLOT_PERM  Y / PERM on bucket qubit J: swap slot pairs with factors
//   H_SLOT_H     H on bucket qubit J (prunes)
//   H_SLOT_RY    RY on bucket qubit J: [[c, -s], [s, c]] (prunes)
//   H_SLOT_U     any 2x2 on bucket qubit J, 8 coefficients in a side table
//                (prunes)
//   H_RENORM     scale by the device-computed 1/sqrt(norm2)
//   H_SLOT_BLOCK dense 2^MB x 2^MB complex block of fused bucket-qubit gates
//                (FP32, MB = 4 or 5, MFMA; side table: MFMA operand table,
//                then the FP64 matrix; c[0] = number of gates) (prunes)
// (RX on a bucket qubit runs in H_SLOT_U: a separate handler would add a
// path to the slot switch and push the FP64 sweep past its register budget.)
// Coefficients c[]: DIAG f0r, f0i, f1r, f1i (Z: 1,0,-1,0; PHASE: 1,0,c,s;
// RZ: c,-s,c,s); PERM u10r, u10i, u01r, u01i (the factor for the element
// leaving bit 0, then for the element leaving bit 1); H f; RY c, s (cos and
// sin of theta/2).  With these the arithmetic of
// a slot gate is operation for operation the reference formula of its kind
// (products with exact 0 or +-1 only change the sign of zero results); base
// gates differ from it only by the folding of consecutive factors (a few ulps).
//   H_FRAME      apply the per-bucket Pauli frame (below) to the amplitude
//                registers; the planner puts one before every op that reads or
//                writes them after frame records
//   H_NORM_POINT FP64 sum of |a|^2 of the bucket (pending factor included),
//                written per wavefront for checkpoint hdr >> 13 (the one-pass
//                renormalisation of the FP32 form); no amplitude changes
enum : uint32_t {
    H_BASE_DIAG = 0, H_BASE_X, H_BASE_PERM, H_SLOT_DIAG, H_SLOT_X, H_SLOT_PERM, H_SLOT_H, H_SLOT_RY,
    H_SLOT_U, H_RENORM, H_SLOT_BLOCK, H_NORM_POINT, H_FRAME
};
// handlers whose outputs are pruned (they mix slot pairs)
__host__ __device__ constexpr bool branching_handler(uint32_t h) {
    return h == H_SLOT_H || h == H_SLOT_RY || h == H_SLOT_U || h == H_SLOT_BLOCK;
}

// Base-qubit records (H_BASE_*) select their factor c0 / c1 by the base-key
// bit at `target` (hdr bits 7-12) and then XOR the key with the 64-bit mask
// in the bits of c[4] (zero for diagonal gates, 1 << target for X / Y / PERM).
//
// Per-bucket Pauli frame.  A Pauli gate on a bucket qubit whose controls are
// base qubits only (CX / CCX from base qubits onto a bucket qubit, CZ between
// a base and a bucket qubit, X / Y / Z on a bucket qubit) is applied as such a
// base record: it toggles frame bits kept in the bucket's base-key register,
// which are zero in every stored key -- the X part at the bucket qubits'
// own positions and the Z part at bits FRAME_Z_BIT0 + j -- and multiplies the
// pending factor by the sign or phase the reordering produces.  The logical
// bucket is g * Z^fz X^fx (registers): before any op that reads or writes the
// amplitude registers, and before the store, the sweep applies the frame (slot
// swaps and sign flips, exact) and clears the bits.  Keys of at most 59 qubits
// leave room for the Z part.
constexpr int FRAME_Z_BIT0 = 59;
__host__ __device__ constexpr uint64_t frame_z_mask(int mb) {
    return mb <= 0 ? 0ull : (((1ull << mb) - 1) << FRAME_Z_BIT0);
}

// One sweep op: 16 dwords, staged 64 at a time in the lanes of a wavefront.
// hdr = handler | J << 4 | target << 7; ctl = cmS | cvS << 8.  Ops that read
// the coefficient side table (H_SLOT_U, H_SLOT_BLOCK) carry its index, in
// units of 8 doubles, as a full 64-bit integer in the bits of c[4]: a field of
// its own, so no program is too long for it.
struct MOp {
    uint32_t hdr;
    uint32_t ctl;
    uint64_t cmB, cvB;
    double c[5];
};
static_assert(sizeof(MOp) == 64, "MOp is sixteen dwords");
__host__ __device__ inline double coef_index_bits(uint64_t ci) { return __builtin_bit_cast(double, ci); }
__host__ __device__ inline uint64_t coef_index_of(double c4) { return __builtin_bit_cast(uint64_t, c4); }

// The same pointer in the constant address space: a load through it at a
// wavefront-uniform address becomes a scalar load into SGPRs.  Only for memory
// that no kernel writes while the reading kernel runs (the program and the
// coefficient side table).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_const_space(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}

struct Slot {           // hash index slot: 16 B, 8 slots = one 128 B line
    uint64_t key;
    uint64_t val;
};

struct QMap {           // bucket qubits, increasing
    uint64_t mask;
    int32_t mb;
    uint8_t q[8];
};

// Engine counters, folded from the per-block slots by k_stats_reduce.
struct Stats {
    unsigned long long updates;      // sum over gates of nonzeros entering the gate (monotonic)
    unsigned long long moved;        // sum over launches of nonzeros read + nonzeros written (monotonic)
    unsigned long long nnz;          // nonzero slots after the last counting launch
    unsigned long long empty;        // all-zero live buckets after the last counting launch
    unsigned long long err;          // sticky error bits
    unsigned long long moved_pair;   // the part of `moved` from cross-bucket launches (monotonic)
    unsigned long long pad[2];
};

// One slot per workgroup: counting kernels write their block's totals here
// with plain stores (no same-address atomics across thousands of blocks);
// upd / moved accumulate across launches until k_stats_reduce folds them.
struct BlockStats {
    unsigned long long upd, moved, nnz, empty, moved_pair;
    unsigned long long pad[3];
};

enum : unsigned long long { ERR_PROBE = 1, ERR_DUP = 2, ERR_MISSING = 4, ERR_CAP = 8, ERR_SCAN = 16, ERR_RENORM = 32,
                           ERR_KEY = 64 };

// Plain complex value (trivially copyable, so register arrays of it are fully
// scalarised); 8/16-byte aligned for single 64/128-bit memory accesses.
template <class R>
struct alignas(2 * sizeof(R)) Cplx {
    R x, y;
};
using cd = Cplx<double>;
using cf = Cplx<float>;
__device__ __forceinline__ cd mkd(double a, double b) { return cd{a, b}; }
__device__ __forceinline__ cf mkf(float a, float b) { return cf{a, b}; }

template <class R> struct CT { using T = Cplx<R>; };

__device__ __forceinline__ uint64_t hmix(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33; return k;
}

__device__ __forceinline__ uint32_t pext_q(uint64_t key, const QMap& qm) {
    uint32_t s = 0;
    for (int j = 0; j < qm.mb; ++j) s |= (uint32_t)((key >> qm.q[j]) & 1ull) << j;
    return s;
}
__device__ __forceinline__ uint64_t pdep_q(uint32_t s, const QMap& qm) {
    uint64_t k = 0;
    for (int j = 0; j < qm.mb; ++j) k |= (uint64_t)((s >> j) & 1u) << qm.q[j];
    return k;
}

__device__ __forceinline__ size_t amp_index(uint64_t b, int s, int S) {
    return ((size_t)(b >> 6) * (size_t)S + (size_t)s) * TILE + (size_t)(b & 63);
}

template <int N, class F, int... I>
__device__ __forceinline__ void static_for_impl(F& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl<N>(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------
// Gate arithmetic.  FP64 follows the reference's per-gate formulas and their
// operation order (emulator/device/device_data.cuh:159-319,
// emulator/include/data_t.h:63-80); the build uses -ffp-contract=off so no
// multiply-add is fused.  FP32 uses explicit fused multiply-adds.
// v0 / v1 are the amplitudes of the keys with target bit 0 / 1.
// ---------------------------------------------------------------------------
template <uint32_t K>
__device__ __forceinline__ void pair64(const double* p, cd v0, cd v1, cd& n0, cd& n1) {
    if constexpr (K == KX) { n0 = v1; n1 = v0; }
    else if constexpr (K == KY) { n0 = mkd(v1.y, -v1.x); n1 = mkd(-v0.y, v0.x); }
    else if constexpr (K == KZ) { n0 = v0; n1 = mkd(-v1.x, -v1.y); }
    else if constexpr (K == KH) {
        const double f = p[0];
        n0 = mkd(f * (v0.x + v1.x), f * (v0.y + v1.y));
        n1 = mkd(f * (v0.x - v1.x), f * (v0.y - v1.y));
    } else if constexpr (K == KPHASE) {
        const double c = p[0], s = p[1];
        n0 = v0;
        n1 = mkd(c * v1.x - s * v1.y, c * v1.y + s * v1.x);
    } else if constexpr (K == KRZ) {
        const double c = p[0], s = p[1];
        n0 = mkd(c * v0.x + s * v0.y, c * v0.y - s * v0.x);
        n1 = mkd(c * v1.x - s * v1.y, c * v1.y + s * v1.x);
    } else if constexpr (K == KRX) {
        const double c = p[0], s = p[1];
        n0 = mkd(c * v0.x + s * v1.y, c * v0.y - s * v1.x);
        n1 = mkd(c * v1.x + s * v0.y, c * v1.y - s * v0.x);
    } else if constexpr (K == KRY) {
        const double c = p[0], s = p[1];
        n0 = mkd(c * v0.x - s * v1.x, c * v0.y - s * v1.y);
        n1 = mkd(c * v1.x + s * v0.x, c * v1.y + s * v0.y);
    } else if constexpr (K == KU) {
        n0 = mkd((p[0] * v0.x - p[1] * v0.y) + (p[2] * v1.x - p[3] * v1.y),
                          (p[0] * v0.y + p[1] * v0.x) + (p[2] * v1.y + p[3] * v1.x));
        n1 = mkd((p[4] * v0.x - p[5] * v0.y) + (p[6] * v1.x - p[7] * v1.y),
                          (p[4] * v0.y + p[5] * v0.x) + (p[6] * v1.y + p[7] * v1.x));
    } else if constexpr (K == KDIAG) {
        n0 = mkd(p[0] * v0.x - p[1] * v0.y, p[0] * v0.y + p[1] * v0.x);
        n1 = mkd(p[2] * v1.x - p[3] * v1.y, p[2] * v1.y + p[3] * v1.x);
    } else if constexpr (K == KPERM) {
        n0 = mkd(p[0] * v1.x - p[1] * v1.y, p[0] * v1.y + p[1] * v1.x);
        n1 = mkd(p[2] * v0.x - p[3] * v0.y, p[2] * v0.y + p[3] * v0.x);
    } else { n0 = v0; n1 = v1; }
}

template <uint32_t K>
__device__ __forceinline__ void pair32(const double* pd, cf v0, cf v1, cf& n0, cf& n1) {
    float p[8];
    constexpr int NP = (K == KU) ? 8 : (K == KDIAG || K == KPERM) ? 4 : 2;
    for (int k = 0; k < 8; ++k) p[k] = k < NP ? (float)pd[k] : 0.0f;
    if constexpr (K == KX) { n0 = v1; n1 = v0; }
    else if constexpr (K == KY) { n0 = mkf(v1.y, -v1.x); n1 = mkf(-v0.y, v0.x); }
    else if constexpr (K == KZ) { n0 = v0; n1 = mkf(-v1.x, -v1.y); }
    else if constexpr (K == KH) {
        const float f = p[0];
        n0 = mkf(f * (v0.x + v1.x), f * (v0.y + v1.y));
        n1 = mkf(f * (v0.x - v1.x), f * (v0.y - v1.y));
    } else if constexpr (K == KPHASE) {
        const float c = p[0], s = p[1];
        n0 = v0;
        n1 = mkf(__fmaf_rn(c, v1.x, -(s * v1.y)), __fmaf_rn(c, v1.y, s * v1.x));
    } else if constexpr (K == KRZ) {
        const float c = p[0], s = p[1];
        n0 = mkf(__fmaf_rn(c, v0.x, s * v0.y), __fmaf_rn(c, v0.y, -(s * v0.x)));
        n1 = mkf(__fmaf_rn(c, v1.x, -(s * v1.y)), __fmaf_rn(c, v1.y, s * v1.x));
    } else if constexpr (K == KRX) {
        const float c = p[0], s = p[1];
        n0 = mkf(__fmaf_rn(c, v0.x, s * v1.y), __fmaf_rn(c, v0.y, -(s * v1.x)));
        n1 = mkf(__fmaf_rn(c, v1.x, s * v0.y), __fmaf_rn(c, v1.y, -(s * v0.x)));
    } else if constexpr (K == KRY) {
        const float c = p[0], s = p[1];
        n0 = mkf(__fmaf_rn(c, v0.x, -(s * v1.x)), __fmaf_rn(c, v0.y, -(s * v1.y)));
        n1 = mkf(__fmaf_rn(c, v1.x, s * v0.x), __fmaf_rn(c, v1.y, s * v0.y));
    } else if constexpr (K == KU) {
        float t;
        t = p[0] * v0.x; t = __fmaf_rn(-p[1], v0.y, t); t = __fmaf_rn(p[2], v1.x, t); n0.x = __fmaf_rn(-p[3], v1.y, t);
        t = p[0] * v0.y; t = __fmaf_rn(p[1], v0.x, t); t = __fmaf_rn(p[2], v1.y, t); n0.y = __fmaf_rn(p[3], v1.x, t);
        t = p[4] * v0.x; t = __fmaf_rn(-p[5], v0.y, t); t = __fmaf_rn(p[6], v1.x, t); n1.x = __fmaf_rn(-p[7], v1.y, t);
        t = p[4] * v0.y; t = __fmaf_rn(p[5], v0.x, t); t = __fmaf_rn(p[6], v1.y, t); n1.y = __fmaf_rn(p[7], v1.x, t);
    } else if constexpr (K == KDIAG) {
        n0 = mkf(__fmaf_rn(p[0], v0.x, -(p[1] * v0.y)), __fmaf_rn(p[0], v0.y, p[1] * v0.x));
        n1 = mkf(__fmaf_rn(p[2], v1.x, -(p[3] * v1.y)), __fmaf_rn(p[2], v1.y, p[3] * v1.x));
    } else if constexpr (K == KPERM) {
        n0 = mkf(__fmaf_rn(p[0], v1.x, -(p[1] * v1.y)), __fmaf_rn(p[0], v1.y, p[1] * v1.x));
        n1 = mkf(__fmaf_rn(p[2], v0.x, -(p[3] * v0.y)), __fmaf_rn(p[2], v0.y, p[3] * v0.x));
    } else { n0 = v0; n1 = v1; }
}

__device__ __forceinline__ bool negl64(cd v, double thr) {
    return fabs(v.x) < thr && fabs(v.y) < thr;
}
__device__ __forceinline__ bool nonzero(cd v) { return v.x != 0.0 || v.y != 0.0; }
__device__ __forceinline__ bool nonzero(cf v) { return v.x != 0.0f || v.y != 0.0f; }

// Apply a branching pair and prune its outputs (componentwise |re|,|im| < thr,
// emulator/include/data_t.h:50-60).  FP32: decisions for outputs with
// max(|re|,|im|) in [thr/4, 4 thr) are redone in FP64 from the FP32 inputs.
// Band edges [thr/4, 4 thr) as FP32 values.  The round trip through a
// scalar register keeps the compiler from widening the FP32 compares back to
// FP64 against the double edges (FP64 compares, and two uniform doubles held
// in VGPRs, that the FP32 sweep spilled).
struct Band32 {
    float lo, hi;
};
__device__ __forceinline__ float opaque_uniform(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, v)));
}
__device__ __forceinline__ Band32 band32(double thr) {
    return Band32{opaque_uniform((float)(0.25 * thr)), opaque_uniform((float)(4.0 * thr))};
}
template <uint32_t K>
__device__ __forceinline__ void branch_pair(const DOp& op, double thr, cd v0, cd v1, cd& n0, cd& n1) {
    pair64<K>(op.p, v0, v1, n0, n1);
    if (negl64(n0, thr)) n0 = mkd(0.0, 0.0);
    if (negl64(n1, thr)) n1 = mkd(0.0, 0.0);
}
template <uint32_t K>
__device__ __forceinline__ void branch_pair(const DOp& op, double thr, cf v0, cf v1, cf& n0, cf& n1) {
    pair32<K>(op.p, v0, v1, n0, n1);
    // band [thr/4, 4 thr) tested in FP32: an output within an ulp of an edge
    // is far from thr, where the FP32 and FP64 decisions agree
    const Band32 band = band32(thr);
    const float m0 = fmaxf(fabsf(n0.x), fabsf(n0.y));
    const float m1 = fmaxf(fabsf(n1.x), fabsf(n1.y));
    bool k0 = m0 >= band.hi, k1 = m1 >= band.hi;
    const bool near0 = !k0 && m0 >= band.lo, near1 = !k1 && m1 >= band.lo;
    // the FP64 recomputation runs when any lane has an output in the band (a
    // uniform branch, no per-pair exec-mask save and restore)
    if (__builtin_amdgcn_ballot_w64(near0 || near1) != 0) {
        cd w0 = mkd(v0.x, v0.y), w1 = mkd(v1.x, v1.y), d0, d1;
        pair64<K>(op.p, w0, w1, d0, d1);
        k0 = k0 || (near0 && !negl64(d0, thr));
        k1 = k1 || (near1 && !negl64(d1, thr));
    }
    if (!k0) n0 = mkf(0.0f, 0.0f);
    if (!k1) n1 = mkf(0.0f, 0.0f);
}

template <uint32_t K, class C>
__device__ __forceinline__ void any_pair(const DOp& op, double thr, C v0, C v1, C& n0, C& n1) {
    if constexpr (is_branch(K)) branch_pair<K>(op, thr, v0, v1, n0, n1);
    else if constexpr (sizeof(C) == 16) pair64<K>(op.p, v0, v1, n0, n1);
    else pair32<K>(op.p, v0, v1, n0, n1);
}

template <class C>
__device__ __forceinline__ C csel(bool c, C a, C b) { return c ? a : b; }

// ---------------------------------------------------------------------------
// Wavefront-cooperative hash probing: 8 lanes probe one key over an aligned
// window of 8 slots (one 128-byte line) and decide with a 64-lane ballot; the
// 8 groups of a wavefront serve their 8 lanes' keys in turn.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t group_bits(unsigned long long ball, int g) {
    return (uint32_t)((ball >> (g * 8)) & 0xffull);
}

// Returns the value stored for qkey (NONE when absent).  All 64 lanes of the
// wavefront must call it; `active` marks lanes with a query.
__device__ __forceinline__ uint64_t coop_find(const Slot* __restrict__ tab, uint64_t mask, uint64_t qkey,
                                              bool active, unsigned long long* err) {
    const int lane = threadIdx.x & 63, g = lane >> 3, gl = lane & 7;
    uint64_t result = NONE;
    for (int r = 0; r < 8; ++r) {
        const int src = (g << 3) | r;
        const uint64_t key = __shfl(qkey, src);
        const int act = __shfl((int)active, src);
        if (act) {
            uint64_t pos = hmix(key) & mask & ~7ull;
            for (uint64_t it = 0;; ++it) {
                const Slot sl = tab[(pos + gl) & mask];
                const uint32_t m = group_bits(__ballot(sl.key == key), g);
                const uint32_t e = group_bits(__ballot(sl.key == EMPTY_KEY), g);
                if (m) {
                    const uint64_t v = __shfl(sl.val, (g << 3) + __ffs(m) - 1);
                    if (gl == r) result = v;
                    break;
                }
                if (e) break;
                pos = (pos + 8) & mask;
                if (it > (mask >> 3)) { if (gl == 0) atomicOr(err, ERR_PROBE); break; }
            }
        }
    }
    return result;
}

// Insert qkey if absent.  With `val` != NONE the inserting lane stores it.
// Returns true when this lane's key was newly inserted.
__device__ __forceinline__ bool coop_insert(Slot* __restrict__ tab, uint64_t mask, uint64_t qkey, uint64_t val,
                                            bool active, unsigned long long* err) {
    const int lane = threadIdx.x & 63, g = lane >> 3, gl = lane & 7;
    bool inserted = false;
    for (int r = 0; r < 8; ++r) {
        const int src = (g << 3) | r;
        const uint64_t key = __shfl(qkey, src);
        const uint64_t v = __shfl(val, src);
        const int act = __shfl((int)active, src);
        if (act) {
            uint64_t pos = hmix(key) & mask & ~7ull;
            for (uint64_t it = 0;; ++it) {
                const Slot sl = tab[(pos + gl) & mask];
                const uint32_t m = group_bits(__ballot(sl.key == key), g);
                if (m) break;                                   // already present
                const uint32_t e = group_bits(__ballot(sl.key == EMPTY_KEY), g);
                if (e) {
                    const int leader = __ffs(e) - 1;
                    unsigned long long old = 0;
                    if (gl == leader) {
                        Slot* s = &tab[(pos + leader) & mask];
                        old = atomicCAS((unsigned long long*)&s->key, (unsigned long long)EMPTY_KEY,
                                        (unsigned long long)key);
                        if (old == EMPTY_KEY && v != NONE) s->val = v;
                    }
                    old = __shfl(old, (g << 3) + leader);
                    if (old == EMPTY_KEY) { if (gl == r) inserted = true; break; }
                    if (old == key) break;
                    continue;                                   // lost the slot: re-read the window
                }
                pos = (pos + 8) & mask;
                if (it > (mask >> 3)) { if (gl == 0) atomicOr(err, ERR_PROBE); break; }
            }
        }
    }
    return inserted;
}

// ---------------------------------------------------------------------------
// Single-pass device-wide exclusive scan (decoupled look-back, Merrill and
// Garland, "Single-pass Parallel Prefix Scan with Decoupled Look-back",
// NVIDIA TR 2016-002), with the look-back done by a whole wavefront.  Tiles are
// handed out in start order by a ticket counter that the last tile resets, so
// a tile only ever waits on tiles that are already running.  Status words
// carry flag (2 bits) | epoch (22 bits) | value (40 bits); a word from an older
// launch (other epoch) reads as "not ready".
// ---------------------------------------------------------------------------
struct ScanCtl {
    unsigned long long* status;
    unsigned int* ticket;
    unsigned long long* err;   // sticky error word (ERR_SCAN on a look-back that never completes)
    uint32_t epoch;
};

constexpr unsigned long long ST_AGG = 1ull << 62, ST_PRE = 2ull << 62;
constexpr unsigned long long VAL_MASK = (1ull << 40) - 1;

__device__ __forceinline__ unsigned long long st_pack(unsigned long long flag, uint32_t epoch, unsigned long long v) {
    return flag | ((unsigned long long)(epoch & 0x3fffff) << 40) | (v & VAL_MASK);
}

// Block-level exclusive scan of `count`; returns this thread's global output
// offset.  `tile` is the ticket of this block, `ntiles` the grid size.  The
// last tile writes the total over all tiles to *gtotal; every thread gets the
// inclusive prefix of its own tile in *ltotal (the total, in the last tile).
__device__ __forceinline__ unsigned long long block_scan(uint32_t count, ScanCtl sc, uint32_t tile,
                                                         uint32_t ntiles, unsigned long long* gtotal,
                                                         unsigned long long* ltotal) {
    __shared__ unsigned long long s_wave[BLOCK / 64];
    __shared__ unsigned long long s_prefix;
    __shared__ unsigned long long s_total;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // inclusive wave scan
    unsigned long long x = count;
    for (int d = 1; d < 64; d <<= 1) {
        unsigned long long y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    unsigned long long wave_off = 0, agg = 0;
    for (int i = 0; i < BLOCK / 64; ++i) {
        if (i < w) wave_off += s_wave[i];
        agg += s_wave[i];
    }
    if (w == 0) {
        // Wavefront 0 looks back 64 predecessors at a time (lane k reads tile
        // hi - k): it folds in every aggregate up to the nearest inclusive
        // prefix, or all 64 when none is there yet, and waits only while a
        // nearer predecessor has not published its aggregate.  Every
        // predecessor holds an earlier ticket and is running, so the wait
        // ends; the bound only turns a broken invariant into a reported error.
        // A status word carries its value with its flag and epoch, so the
        // device-scope atomics can be relaxed: release / acquire would write
        // back / invalidate the XCD's whole L2 at every publication and poll.
        unsigned long long prefix = 0;
        if (tile == 0) {
            if (lane == 0)
                __hip_atomic_store(&sc.status[0], st_pack(ST_PRE, sc.epoch, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&sc.status[tile], st_pack(ST_AGG, sc.epoch, agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            int hi = (int)tile - 1;
            for (unsigned long long spins = 0;; ++spins) {
                if (spins > (1ull << 24)) {
                    if (lane == 0) atomicOr(sc.err, ERR_SCAN);
                    break;
                }
                const int j = hi - lane;
                unsigned long long st = st_pack(ST_PRE, sc.epoch, 0);     // before tile 0: empty prefix
                if (j >= 0) st = __hip_atomic_load(&sc.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t ep = (uint32_t)((st >> 40) & 0x3fffff);
                const unsigned long long fl = st & (3ull << 62);
                const bool ready = ep == (sc.epoch & 0x3fffff) && fl != 0;
                const unsigned long long okm = __ballot(ready);
                const unsigned long long prem = __ballot(ready && fl == ST_PRE);
                const int first = prem ? __ffsll((long long)prem) - 1 : 63;
                const unsigned long long need = (2ull << first) - 1;       // lanes 0..first
                if ((okm & need) != need) continue;                         // a nearer one is not ready
                unsigned long long v = lane <= first ? (st & VAL_MASK) : 0ull;
                for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
                prefix += v;
                if (prem) break;
                hi -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&sc.status[tile], st_pack(ST_PRE, sc.epoch, prefix + agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_prefix = prefix;
            s_total = prefix + agg;
            if (tile == ntiles - 1) {
                *gtotal = prefix + agg;
                __hip_atomic_store(sc.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    __syncthreads();
    *ltotal = s_total;
    return s_prefix + wave_off + (x - count);
}

__device__ __forceinline__ uint32_t take_ticket(ScanCtl sc) {
    __shared__ uint32_t s_tile;
    if (threadIdx.x == 0) s_tile = atomicAdd(sc.ticket, 1u);
    __syncthreads();
    return s_tile;
}

// Block reduction of the counters into this workgroup's BlockStats slot.
__device__ __forceinline__ void block_stats(BlockStats* bs, unsigned long long upd, unsigned long long nnz,
                                            unsigned long long empty, unsigned long long moved = 0,
                                            bool pair = false) {
    __shared__ unsigned long long s_red[4][BLOCK / 64];
    for (int d = 32; d >= 1; d >>= 1) {
        upd += __shfl_xor(upd, d);
        nnz += __shfl_xor(nnz, d);
        empty += __shfl_xor(empty, d);
        moved += __shfl_xor(moved, d);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_red[0][w] = upd;
        s_red[1][w] = nnz;
        s_red[2][w] = empty;
        s_red[3][w] = moved;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k)
            for (int i = 0; i < BLOCK / 64; ++i) t[k] += s_red[k][i];
        BlockStats& b = bs[blockIdx.x];
        b.upd += t[0];
        b.moved += t[3];
        if (pair) b.moved_pair += t[3];
        b.nnz = t[1];
        b.empty = t[2];
    }
}

}  // namespace sqe
