// Kernel templates of the sparse statevector engine (gfx950); instantiated
// per precision in sqe_kernels_f64.hip / sqe_kernels_f32.hip.
//
// Reference path replaced (ivanNovikau/SparQC, emulator/device/):
//   FC__::gate_sq / gate_sq_par   device_table_FC.cuh:113-205  -> k_sweep, k_pair
//   *_core gate arithmetic         device_data.cuh:159-319      -> pair64 / pair32
//   update_ah_sup prune+append     device_data.cuh:134-156      -> branch_pair + block_scan
//   device_zero_arrays rebuild     device_data.cuh:62-105       -> out-of-place stores, k_compact
//   Table__::get_statevector       table.cuh:47-75              -> k_export
// Norm reduction, controlled gates, hashing of 64-bit keys and the FP32 form
// have no reference counterpart.
#pragma once
#include <algorithm>
#include <type_traits>
#include "sqe_device.hpp"
#include "sqe_launch.hpp"

namespace sqe {

// ---------------------------------------------------------------------------
// Fused local sweep: one thread owns one bucket, keeps its 2^MB amplitudes in
// registers and applies a whole run of bucket-local gates before writing back.
// ---------------------------------------------------------------------------

// Complex rotation by (c, s): (c x - s y, c y + s x), the reference's phase
// formula (emulator/device/device_data.cuh:221-224); FP32 with fused multiply-add.
__device__ __forceinline__ cd rot(double c, double s, cd v) {
    return mkd(c * v.x - s * v.y, c * v.y + s * v.x);
}
__device__ __forceinline__ cf rot(float c, float s, cf v) {
    return mkf(__fmaf_rn(c, v.x, -(s * v.y)), __fmaf_rn(c, v.y, s * v.x));
}
// The same rotation with fused multiply-adds in FP64 too: for the folded
// per-bucket factor, whose rounding already differs from gate-by-gate
// evaluation by a few ulps.
__device__ __forceinline__ cd rot_fma(double c, double s, cd v) {
    return mkd(__fma_rn(c, v.x, -(s * v.y)), __fma_rn(c, v.y, s * v.x));
}
__device__ __forceinline__ cf rot_fma(float c, float s, cf v) { return rot(c, s, v); }

// RY [[c, -s], [s, c]] with c, s of theta/2, the reference formula
// (emulator/device/device_data.cuh:290-319) operation for operation; FP32
// with fused multiply-add.
__device__ __forceinline__ void ry_pair(double c, double s, cd v0, cd v1, cd& n0, cd& n1) {
    n0 = mkd(c * v0.x - s * v1.x, c * v0.y - s * v1.y);
    n1 = mkd(s * v0.x + c * v1.x, s * v0.y + c * v1.y);
}
__device__ __forceinline__ void ry_pair(float c, float s, cf v0, cf v1, cf& n0, cf& n1) {
    n0 = mkf(__fmaf_rn(c, v0.x, -(s * v1.x)), __fmaf_rn(c, v0.y, -(s * v1.y)));
    n1 = mkf(__fmaf_rn(c, v1.x, s * v0.x), __fmaf_rn(c, v1.y, s * v0.y));
}
__device__ __forceinline__ void hpair(double f, cd v0, cd v1, cd& n0, cd& n1) {
    n0 = mkd(f * (v0.x + v1.x), f * (v0.y + v1.y));
    n1 = mkd(f * (v0.x - v1.x), f * (v0.y - v1.y));
}
__device__ __forceinline__ void hpair(float f, cf v0, cf v1, cf& n0, cf& n1) {
    n0 = mkf(f * (v0.x + v1.x), f * (v0.y + v1.y));
    n1 = mkf(f * (v0.x - v1.x), f * (v0.y - v1.y));
}

// Prune a branching output (componentwise |re|,|im| < thr,
// emulator/include/data_t.h:50-60).  FP32: outputs with max(|re|,|im|) in
// [thr/4, 4 thr) are decided on an FP64 recomputation.  The band edges are
// compared in FP32: an output within an ulp of an edge is far from thr, where
// the FP32 and FP64 decisions agree, so the edge rounding cannot change a
// decision.
__device__ __forceinline__ cd prune(cd n, double thr) { return negl64(n, thr) ? mkd(0.0, 0.0) : n; }

template <int MB, class R>
__device__ __forceinline__ uint32_t count_nz(const typename CT<R>::T (&a)[1 << MB]) {
    uint32_t c = 0;
    static_for<(1 << MB)>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        c += nonzero(a[s]) ? 1u : 0u;
    });
    return c;
}

// base ^ (pred ? flip : 0): one v_bitop3 per word, S0 ^ (S1 & S2) (truth
// table 0x78 with S0 = 0xF0, S1 = 0xCC, S2 = 0xAA), on a 0 / ~0 lane mask.
// Written as a select, the compiler moved the uniform flip into VGPRs and
// selected each word (six instructions instead of three).
__device__ __forceinline__ uint64_t flip_if(uint64_t base, uint64_t flip, bool pred) {
    const uint32_t m = pred ? ~0u : 0u;
    const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)base, (uint32_t)flip, m, 0x78);
    const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(base >> 32), (uint32_t)(flip >> 32), m, 0x78);
    return ((uint64_t)hi << 32) | lo;
}

// -v where c holds (a sign flip of both components: one XOR per FP64 high
// word or FP32 word with a lane mask).
__device__ __forceinline__ Cplx<double> negate_if(Cplx<double> v, bool c) {
    const uint64_t m = c ? 0x8000000000000000ull : 0ull;
    return Cplx<double>{__builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v.x) ^ m),
                        __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v.y) ^ m)};
}
__device__ __forceinline__ Cplx<float> negate_if(Cplx<float> v, bool c) {
    const uint32_t m = c ? 0x80000000u : 0u;
    return Cplx<float>{__builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v.x) ^ m),
                       __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v.y) ^ m)};
}

// Decoded sweep op (uniform across the wavefront).
struct SOp {
    uint32_t h, j, cmS, cvS;
    uint64_t ci;                      // coefficient side-table index (H_SLOT_U, H_SLOT_BLOCK)
    uint64_t tbase, cmB, cvB;
    double c[5];
};

template <int S, int J, int Q>
struct PairIdx {
    static constexpr int s0 = ((Q >> J) << (J + 1)) | (Q & ((1 << J) - 1));
    static constexpr int s1 = s0 | (1 << J);
};

// Slot handlers with compile-time target bit J, applied to the EPT buckets a
// thread owns (a[e], predicate pred[e]).  MASK = false is the common case of a
// gate without controls: every pair of every bucket is updated, straight-line.
// MASK = true covers bucket controls (a scalar condition per slot pair) and
// base controls (the lane predicate): the pair's new values are computed and
// selected, still without branches, so no pair leaves a merge of old and new
// register copies behind.
// Returns true when it counted the bucket's nonzero outputs into nzc (FP64
// H / RY handlers without controls, from their prune decisions; in FP32 the
// extra live counter cost more than the recount it saves).
template <int MB, class R, int EPT, int J, bool MASK>
__device__ __forceinline__ bool slot_handler(const SOp& op, typename CT<R>::T (&a)[EPT][1 << MB],
                                             const bool (&pred)[EPT], const R (&pr)[EPT],
                                             const R (&pi)[EPT], double thr, const double* __restrict__ coef,
                                             uint32_t (&nzc)[EPT]) {
    using C = typename CT<R>::T;
    constexpr int HALF = (1 << MB) / 2;
    constexpr bool F32 = sizeof(R) == 4;
    const bool pall = op.cmB == 0;               // every bucket takes the gate
    // apply f(x, y, n0, n1) to every pair (s0 = slot with bit J clear, s1 = set);
    // with FLUSH (branching gates) the bucket's pending factor is multiplied
    // into both inputs first (in the amplitude precision), so every slot
    // leaves with its true value
    auto pairs = [&](auto&& f, auto flush_c) __attribute__((always_inline)) {
        constexpr bool FLUSH = decltype(flush_c)::value;
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
            constexpr int e = decltype(ec)::value;
            static_for<HALF>([&](auto qc) __attribute__((always_inline)) {
                constexpr int s0 = PairIdx<1 << MB, J, decltype(qc)::value>::s0;
                constexpr int s1 = PairIdx<1 << MB, J, decltype(qc)::value>::s1;
                C x = a[e][s0], y = a[e][s1];
                if constexpr (FLUSH) {
                    const R fr = (R)pr[e], fi = (R)pi[e];
                    x = rot_fma(fr, fi, x);
                    y = rot_fma(fr, fi, y);
                }
                C n0, n1;
                f(x, y, n0, n1, nzc[e]);
                if constexpr (MASK) {
                    const bool m = (((uint32_t)s0 & op.cmS) == op.cvS) && (pall || pred[e]);
                    a[e][s0] = csel(m, n0, x);
                    a[e][s1] = csel(m, n1, y);
                } else {
                    a[e][s0] = n0;
                    a[e][s1] = n1;
                }
            });
        });
    };
    // prune both outputs and count the ones kept (an exact zero is never kept)
    auto prune_pair = [&](C x, C y, C& n0, C& n1, uint32_t& kept, auto&& exact) __attribute__((always_inline)) {
        if constexpr (F32) {
            // keep above the band, prune below it, and decide the band on
            // the FP64 recomputation; the recomputation runs when any lane of
            // the wavefront has an output in the band (a uniform branch, so
            // no per-pair exec-mask save and restore)
            const Band32 band = band32(thr);
            const float m0 = fmaxf(fabsf(n0.x), fabsf(n0.y)), m1 = fmaxf(fabsf(n1.x), fabsf(n1.y));
            bool k0 = m0 >= band.hi, k1 = m1 >= band.hi;
            const bool near0 = !k0 && m0 >= band.lo, near1 = !k1 && m1 >= band.lo;
            if (__builtin_amdgcn_ballot_w64(near0 || near1) != 0) {
                cd e0, e1;
                exact(mkd(x.x, x.y), mkd(y.x, y.y), e0, e1);
                k0 = k0 || (near0 && !negl64(e0, thr));
                k1 = k1 || (near1 && !negl64(e1, thr));
            }
            n0 = k0 ? n0 : mkf(0.0f, 0.0f);
            n1 = k1 ? n1 : mkf(0.0f, 0.0f);
            kept += (k0 ? 1u : 0u) + (k1 ? 1u : 0u);
        } else {
            const bool k0 = !negl64(n0, thr), k1 = !negl64(n1, thr);
            n0 = k0 ? n0 : mkd(0.0, 0.0);
            n1 = k1 ? n1 : mkd(0.0, 0.0);
            kept += (k0 ? 1u : 0u) + (k1 ? 1u : 0u);
        }
    };
    // FP64 branching gate without controls: apply it to every pair, keeping
    // per lane the smallest max(|re|,|im|) of its outputs; when no lane of
    // the wavefront has an output below the threshold (full buckets, the
    // usual case once the branching qubits have mixed) nothing is pruned and
    // every slot counts as nonzero, so the componentwise prune and count run
    // only behind that wavefront-uniform test.  The outputs are the same
    // either way.
    auto unprune_then_prune = [&](auto&& gate) __attribute__((always_inline)) {
        R mm[EPT];
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) { mm[decltype(ec)::value] = (R)__builtin_inf(); });
        pairs([&](C x, C y, C& n0, C& n1, uint32_t&) __attribute__((always_inline)) {
            gate(x, y, n0, n1);
        }, std::true_type{});
        bool small = false;
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
            constexpr int e = decltype(ec)::value;
            static_for<(1 << MB)>([&](auto sc) __attribute__((always_inline)) {
                constexpr int sl = decltype(sc)::value;
                mm[e] = fmin(mm[e], fmax(fabs(a[e][sl].x), fabs(a[e][sl].y)));
            });
            small |= !(mm[e] >= (R)thr);
        });
        if (__builtin_amdgcn_ballot_w64(small) != 0) {
            static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
                constexpr int e = decltype(ec)::value;
                uint32_t kept = 0;
                static_for<(1 << MB)>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int sl = decltype(sc)::value;
                    // componentwise |re|, |im| < thr (the prune predicate of negl64)
                    const bool k = !(fabs((double)a[e][sl].x) < thr && fabs((double)a[e][sl].y) < thr);
                    a[e][sl] = k ? a[e][sl] : C{0, 0};
                    kept += k ? 1u : 0u;
                });
                nzc[e] = kept;
            });
        } else {
            static_for<EPT>([&](auto ec) __attribute__((always_inline)) { nzc[decltype(ec)::value] = 1u << MB; });
        }
    };
    // FP32 branching gate without controls (up to 16 slots; 32-slot buckets,
    // compiled for one wavefront per SIMD, run slower with the extra pass):
    // the pending factor goes into the registers first, then a dry run of
    // the gate keeps per lane only the smallest max(|re|,|im|) of its outputs.  When every output of the
    // wavefront is at or above the band's upper edge, the per-pair band test
    // could not drop or re-decide any of them: the gate runs again, storing,
    // and every slot counts as nonzero.  Otherwise the caller runs the
    // per-pair band test on the flushed inputs (the same outputs and
    // decisions as testing during the first pass).
    auto fp32_clear_pass = [&](auto&& gate) __attribute__((always_inline)) -> bool {
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
            constexpr int e = decltype(ec)::value;
            static_for<(1 << MB)>([&](auto sc) __attribute__((always_inline)) {
                constexpr int sl = decltype(sc)::value;
                a[e][sl] = rot_fma((R)pr[e], (R)pi[e], a[e][sl]);
            });
        });
        const R hi = (R)band32(thr).hi;
        bool low = false;
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
            constexpr int e = decltype(ec)::value;
            R mm = (R)__builtin_inf();
            static_for<HALF>([&](auto qc) __attribute__((always_inline)) {
                constexpr int s0 = PairIdx<1 << MB, J, decltype(qc)::value>::s0;
                constexpr int s1 = PairIdx<1 << MB, J, decltype(qc)::value>::s1;
                C n0, n1;
                gate(a[e][s0], a[e][s1], n0, n1);
                mm = fmin(mm, fmin(fmax(fabs(n0.x), fabs(n0.y)), fmax(fabs(n1.x), fabs(n1.y))));
            });
            low |= !(mm >= hi);
        });
        if (__builtin_amdgcn_ballot_w64(low) != 0) return false;
        pairs([&](C x, C y, C& n0, C& n1, uint32_t&) __attribute__((always_inline)) { gate(x, y, n0, n1); },
              std::false_type{});
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) { nzc[decltype(ec)::value] = 1u << MB; });
        return true;
    };
    switch (op.h) {
    case H_SLOT_DIAG: {
        const R f0r = (R)op.c[0], f0i = (R)op.c[1], f1r = (R)op.c[2], f1i = (R)op.c[3];
        pairs([&](C x, C y, C& n0, C& n1, uint32_t&) __attribute__((always_inline)) {
            n0 = rot(f0r, f0i, x);
            n1 = rot(f1r, f1i, y);
        }, std::false_type{});
    } break;
    case H_SLOT_X:
        pairs([&](C x, C y, C& n0, C& n1, uint32_t&) __attribute__((always_inline)) {
            n0 = y;
            n1 = x;
        }, std::false_type{});
        break;
    case H_SLOT_PERM: {
        const R u10r = (R)op.c[0], u10i = (R)op.c[1], u01r = (R)op.c[2], u01i = (R)op.c[3];
        pairs([&](C x, C y, C& n0, C& n1, uint32_t&) __attribute__((always_inline)) {
            n0 = rot(u01r, u01i, y);
            n1 = rot(u10r, u10i, x);
        }, std::false_type{});
    } break;
    case H_SLOT_H: {
        const R f = (R)op.c[0];
        if constexpr (!F32 && !MASK) {
            unprune_then_prune([&](C x, C y, C& n0, C& n1) __attribute__((always_inline)) { hpair(f, x, y, n0, n1); });
            return true;
        }
        if constexpr (F32 && !MASK && MB <= 4) {
            if (fp32_clear_pass([&](C x, C y, C& n0, C& n1) __attribute__((always_inline)) { hpair(f, x, y, n0, n1); }))
                return true;
        }
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) { nzc[decltype(ec)::value] = 0; });
        pairs([&](C x, C y, C& n0, C& n1, uint32_t& k) __attribute__((always_inline)) {
            hpair(f, x, y, n0, n1);
            prune_pair(x, y, n0, n1, k, [&](cd ex, cd ey, cd& e0, cd& e1) __attribute__((always_inline)) {
                hpair(op.c[0], ex, ey, e0, e1);
            });
        }, std::integral_constant<bool, !(F32 && !MASK && MB <= 4)>{});
        return !MASK && !F32;
    }
    case H_SLOT_RY: {
        const R c = (R)op.c[0], sn = (R)op.c[1];
        if constexpr (!F32 && !MASK) {
            unprune_then_prune([&](C x, C y, C& n0, C& n1) __attribute__((always_inline)) {
                ry_pair(c, sn, x, y, n0, n1);
            });
            return true;
        }
        if constexpr (F32 && !MASK && MB <= 4) {
            if (fp32_clear_pass([&](C x, C y, C& n0, C& n1) __attribute__((always_inline)) {
                    ry_pair(c, sn, x, y, n0, n1);
                }))
                return true;
        }
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) { nzc[decltype(ec)::value] = 0; });
        pairs([&](C x, C y, C& n0, C& n1, uint32_t& k) __attribute__((always_inline)) {
            ry_pair(c, sn, x, y, n0, n1);
            prune_pair(x, y, n0, n1, k, [&](cd ex, cd ey, cd& e0, cd& e1) __attribute__((always_inline)) {
                ry_pair(op.c[0], op.c[1], ex, ey, e0, e1);
            });
        }, std::integral_constant<bool, !(F32 && !MASK && MB <= 4)>{});
        return !MASK && !F32;
    }

    case H_SLOT_U: {
        DOp d{};
        const auto* __restrict__ cq = as_const_space(coef + op.ci * 8);
        for (int k = 0; k < 8; ++k) d.p[k] = cq[k];
        pairs([&](C x, C y, C& n0, C& n1, uint32_t&) __attribute__((always_inline)) {
            branch_pair<KU>(d, thr, x, y, n0, n1);
        }, std::true_type{});
    } break;
    default:
        break;
    }
    return false;
}

template <int MB, class R, int EPT, bool MASK>
__device__ __forceinline__ bool slot_dispatch_j(const SOp& op, typename CT<R>::T (&a)[EPT][1 << MB],
                                                const bool (&pred)[EPT], const R (&pr)[EPT],
                                                const R (&pi)[EPT], double thr, const double* __restrict__ coef,
                                                uint32_t (&nzc)[EPT]) {
    switch (op.j) {
    case 0: if constexpr (MB > 0) return slot_handler<MB, R, EPT, 0, MASK>(op, a, pred, pr, pi, thr, coef, nzc); break;
    case 1: if constexpr (MB > 1) return slot_handler<MB, R, EPT, 1, MASK>(op, a, pred, pr, pi, thr, coef, nzc); break;
    case 2: if constexpr (MB > 2) return slot_handler<MB, R, EPT, 2, MASK>(op, a, pred, pr, pi, thr, coef, nzc); break;
    case 3: if constexpr (MB > 3) return slot_handler<MB, R, EPT, 3, MASK>(op, a, pred, pr, pi, thr, coef, nzc); break;
    case 4: if constexpr (MB > 4) return slot_handler<MB, R, EPT, 4, MASK>(op, a, pred, pr, pi, thr, coef, nzc); break;
    default: break;
    }
    return false;
}

template <int MB, class R, int EPT>
__device__ __forceinline__ bool slot_dispatch(const SOp& op, typename CT<R>::T (&a)[EPT][1 << MB],
                                              const bool (&pred)[EPT], const R (&pr)[EPT],
                                              const R (&pi)[EPT], double thr, const double* __restrict__ coef,
                                              uint32_t (&nzc)[EPT]) {
    if (op.cmS == 0 && op.cmB == 0) return slot_dispatch_j<MB, R, EPT, false>(op, a, pred, pr, pi, thr, coef, nzc);
    return slot_dispatch_j<MB, R, EPT, true>(op, a, pred, pr, pi, thr, coef, nzc);
}

// Per-lane factor of a base-qubit DIAG / PERM op (identity where the base
// controls fail).
// (DIAG: f0 / f1 by the bit; PERM: u10 for the element leaving bit 0, u01 for
// the one leaving bit 1 -- both stored as c[0..1] / c[2..3]).
__device__ __forceinline__ void base_factor(const SOp& op, bool b, bool pred, double& fr, double& fi) {
    fr = pred ? (b ? op.c[2] : op.c[0]) : 1.0;
    fi = pred ? (b ? op.c[3] : op.c[1]) : 0.0;
}

// Multiply every slot by the pending per-bucket factor (pr, pi), computed in
// FP64: the reference's phase formula (emulator/device/device_data.cuh:221-224)
// with the folded factor.
template <int MB>
__device__ __forceinline__ void apply_phase(Cplx<float> (&a)[1 << MB], float pr, float pi) {
    static_for<(1 << MB)>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        a[s] = rot(pr, pi, a[s]);
    });
}
template <int MB>
__device__ __forceinline__ void apply_phase(Cplx<double> (&a)[1 << MB], double pr, double pi) {
    static_for<(1 << MB)>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        a[s] = rot(pr, pi, a[s]);
    });
}

// Fused dense block on a D-slot FP32 bucket (D = 16 or 32; one bucket per
// lane, 64 per wavefront): Y = M X with M the D x D complex matrix of the
// block, computed as the real product R X' (R = [[Mr, -Mi], [Mi, Mr]],
// 2D x 2D; X' = D real parts over D imaginary parts, 2D x 64 buckets) by
// v_mfma_f32_32x32x2f32 chains: one per 32-row tile of R (D / 16 of them) and
// per half of the wavefront's buckets (0-31, 32-63), over D steps of two
// input rows.  The lane-per-bucket registers become B operands with one
// v_permlane32_swap per row pair, and the outputs go back the same way.  The
// A operands come from the side table in the lane order the instruction reads
// (tile q, step t, lane l: R[32 q + l % 32][2t + l / 32]).  Outputs are pruned
// componentwise; those within 4x of the threshold are decided on an FP64
// recomputation from the same inputs.  Needs all 64 lanes active
// (wavefront-uniform control flow).
typedef float v16f __attribute__((ext_vector_type(16)));
template <int D>
__device__ __forceinline__ void block_apply(Cplx<float> (&a)[D], float pr, float pi, const double* __restrict__ side,
                                            double thr) {
    static_assert(D == 16 || D == 32, "blocks of 4 or 5 bucket qubits");
    constexpr int Q = D / 16;                                     // 32-row tiles of R
    const uint32_t lane = threadIdx.x & 63;
    const Band32 band = band32(thr);
    const float* __restrict__ tab = (const float*)side;          // Q tiles x D steps x 64 lanes
    const double* __restrict__ m64 = side + Q * D * 32;           // D x D complex, FP64
    float x[2 * D];
    {
        const float fr = pr, fi = pi;
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const Cplx<float> v = rot(fr, fi, a[s]);
            x[s] = v.x;
            x[D + s] = v.y;
        }
    }
    float y[2 * D];
    // register r of lane l holds output row 8 (r / 4) + 4 (l / 32) + r % 4 of
    // its tile for bucket l % 32 of its half; one swap per register pair
    // brings rows back to the lane of their bucket
    auto unpack = [&](const v16f& c0, const v16f& c1, int row0) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(c0[r]), __float_as_uint(c1[r]), false,
                                                             false);
            const int i0 = row0 + 8 * (r / 4) + (r % 4);
            y[i0] = __uint_as_float(sw[0]);
            y[i0 + 4] = __uint_as_float(sw[1]);
        }
    };
    if constexpr (D == 16) {
        float av[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) av[t] = tab[t * 64 + lane];
        v16f acc0 = {}, acc1 = {};
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[2 * t]), __float_as_uint(x[2 * t + 1]),
                                                             false, false);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], __uint_as_float(sw[0]), acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], __uint_as_float(sw[1]), acc1, 0, 0, 0);
        }
        unpack(acc0, acc1, 0);
    } else {
        v16f acc00 = {}, acc01 = {}, acc10 = {}, acc11 = {};
#pragma unroll
        for (int t = 0; t < D; ++t) {
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[2 * t]), __float_as_uint(x[2 * t + 1]),
                                                             false, false);
            const float av0 = tab[t * 64 + lane], av1 = tab[(D + t) * 64 + lane];
            acc00 = __builtin_amdgcn_mfma_f32_32x32x2f32(av0, __uint_as_float(sw[0]), acc00, 0, 0, 0);
            acc01 = __builtin_amdgcn_mfma_f32_32x32x2f32(av0, __uint_as_float(sw[1]), acc01, 0, 0, 0);
            acc10 = __builtin_amdgcn_mfma_f32_32x32x2f32(av1, __uint_as_float(sw[0]), acc10, 0, 0, 0);
            acc11 = __builtin_amdgcn_mfma_f32_32x32x2f32(av1, __uint_as_float(sw[1]), acc11, 0, 0, 0);
        }
        unpack(acc00, acc01, 0);
        unpack(acc10, acc11, 32);
    }
#pragma unroll
    for (int s = 0; s < D; ++s) {
        Cplx<float> n = mkf(y[s], y[D + s]);
        // keep above the band, prune below it; the FP64 row recomputation
        // runs when any lane has this output in the band (a uniform branch)
        const float m = fmaxf(fabsf(n.x), fabsf(n.y));
        bool keep = m >= band.hi;
        const bool near = !keep && m >= band.lo;
        if (__builtin_amdgcn_ballot_w64(near) != 0) {
            double er = 0.0, ei = 0.0;
            for (int k = 0; k < D; ++k) {
                const double mr = m64[(s * D + k) * 2], mi = m64[(s * D + k) * 2 + 1];
                const double xr = x[k], xi = x[D + k];
                er += mr * xr - mi * xi;
                ei += mr * xi + mi * xr;
            }
            keep = keep || (near && !negl64(mkd(er, ei), thr));
        }
        a[s] = keep ? n : mkf(0.0f, 0.0f);
    }
}

// Apply the per-bucket Pauli frame held in the base-key registers (see
// sqe_device.hpp) to the amplitude registers and clear it: for each bucket
// qubit j, swap the slot pairs along bit j in buckets whose frame has X_j, then
// negate the slots with bit j set in buckets whose frame has Z_j.  Both are
// exact.  Wavefront-uniform branches skip the bits no bucket of the
// wavefront has set.
template <int MB, class R, int EPT>
__device__ __forceinline__ void frame_apply(typename CT<R>::T (&a)[EPT][1 << MB], uint64_t (&base)[EPT],
                                            const QMap& qm) {
    using C = typename CT<R>::T;
    constexpr int HALF = (1 << MB) / 2;
    const uint64_t fmask = qm.mask | frame_z_mask(MB);
    bool any = false;
    static_for<EPT>([&](auto ec) __attribute__((always_inline)) { any |= (base[ec] & fmask) != 0; });
    if (__builtin_amdgcn_ballot_w64(any) == 0) return;
    static_for<MB>([&](auto jc) __attribute__((always_inline)) {
        constexpr int J = decltype(jc)::value;
        const uint32_t qj = qm.q[J];
        bool bx[EPT], bz[EPT];
        bool anyx = false, anyz = false;
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
            constexpr int e = decltype(ec)::value;
            bx[e] = ((base[e] >> qj) & 1ull) != 0;
            bz[e] = ((base[e] >> (FRAME_Z_BIT0 + J)) & 1ull) != 0;
            anyx |= bx[e];
            anyz |= bz[e];
        });
        if (__builtin_amdgcn_ballot_w64(anyx) != 0) {
            static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
                constexpr int e = decltype(ec)::value;
                static_for<HALF>([&](auto qc) __attribute__((always_inline)) {
                    constexpr int s0 = PairIdx<1 << MB, J, decltype(qc)::value>::s0;
                    constexpr int s1 = PairIdx<1 << MB, J, decltype(qc)::value>::s1;
                    const C x = a[e][s0], y = a[e][s1];
                    a[e][s0] = csel(bx[e], y, x);
                    a[e][s1] = csel(bx[e], x, y);
                });
            });
        }
        if (__builtin_amdgcn_ballot_w64(anyz) != 0) {
            static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
                constexpr int e = decltype(ec)::value;
                static_for<HALF>([&](auto qc) __attribute__((always_inline)) {
                    constexpr int s1 = PairIdx<1 << MB, J, decltype(qc)::value>::s1;
                    a[e][s1] = negate_if(a[e][s1], bz[e]);
                });
            });
        }
    });
    static_for<EPT>([&](auto ec) __attribute__((always_inline)) { base[ec] &= ~fmask; });
}

// Per-thread counters of one sweep group.
struct SweepCount {
    unsigned long long upd, nnz_in, nnz_out, empty;
    double norm;                      // FP64 sum of |a|^2 of the stored amplitudes
};

// Ops that touch no amplitude register: gates on a base qubit without bucket
// controls (X flips the base key; diagonal and Y/PERM factors are folded into
// the bucket's pending FP64 factor).
__device__ __forceinline__ bool base_class(uint32_t h, uint32_t cmS) { return h <= H_BASE_PERM && cmS == 0; }

// A wavefront's program staging area and the FP64 identity factor (1, 0) of
// failed controls: one copy per workgroup, which every sweep a kernel inlines
// shares (a kernel that dispatches on the bucket size, or that runs one- and
// two-tile passes, still stages through 4 KB per wavefront).
struct SweepLds {
    uint4 prog[BLOCK / 64][64 * 4];
    double one[BLOCK / 64][2];
};
__device__ __forceinline__ SweepLds& sweep_lds() {
    __shared__ SweepLds s;
    return s;
}

// Apply one program to EPT tiles (64 buckets each) with one wavefront: every op
// is decoded once and applied to all of the thread's buckets.  tile[e] < 0
// marks a missing tile (the wavefront computes on zeros and stores nothing).
// Tiles live at basep / ampv in the tile-interleaved layout of sqe_device.hpp.
//
// The program runs as alternating runs of base-class ops, which only update
// the base keys and the pending factors (the amplitude registers stay
// untouched, so the loop carries them without copies), and slot ops, which
// work on the amplitude registers.  The pending factor is a scalar per bucket,
// so it commutes with every bucket-local gate; it is multiplied in before a
// branching op (whose prune test needs true values) and before the store.
// SREC: slot-op records are read with scalar loads from the program in
// memory (k_sweep); otherwise from the LDS copy (k_sweep_batch, whose
// program pointer comes from a per-wave task record).
template <int MB, class R, int EPT, bool BLOCKS = false, bool SREC = false>
__device__ __forceinline__ void sweep_group(uint64_t* __restrict__ basep, void* ampv, const int64_t (&tile)[EPT],
                                            const MOp* __restrict__ ops, uint32_t nops,
                                            const double* __restrict__ coef, double thr,
                                            const double* __restrict__ scale, int store, SweepCount& out,
                                            const SweepCtx& ctx) {
    using C = typename CT<R>::T;
    constexpr int S = 1 << MB;
    const uint32_t lane = threadIdx.x & 63;
    C* __restrict__ amp = (C*)ampv;
    C a[EPT][S];
    uint64_t base[EPT], base0[EPT];
    bool valid[EPT];
    uint32_t cnt[EPT];
    R pr[EPT], pi[EPT];               // pending per-bucket factor (in the amplitude precision)
    out.upd = 0;
    out.nnz_in = 0;
    static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
        constexpr int e = decltype(ec)::value;
        base[e] = EMPTY_KEY;
        if (tile[e] >= 0) {
            const uint64_t b = (uint64_t)tile[e] * TILE + lane;
            base[e] = basep[b];
            static_for<S>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                a[e][s] = amp[amp_index(b, s, S)];
            });
        } else {
            static_for<S>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                a[e][s].x = 0; a[e][s].y = 0;
            });
        }
        base0[e] = base[e];
        valid[e] = base[e] != EMPTY_KEY;
        cnt[e] = valid[e] ? count_nz<MB, R>(a[e]) : 0u;
        out.nnz_in += cnt[e];
        pr[e] = (R)1;
        pi[e] = (R)0;
    });
    // The program is staged 64 ops at a time through LDS (4 KB per wavefront,
    // one coalesced 1 KB load per 16 dwords), with ballot masks of the
    // base-class, base-controlled and renormalisation ops: a gate on a base
    // qubit then reads its record with broadcast LDS loads straight into
    // registers that the VALU uses as operands; a bucket-qubit gate takes its
    // record into scalar registers (scalar loads with SREC, else LDS and
    // v_readfirstlane).  The nonzero count that every gate adds to the update
    // counter changes only at branching gates, so gates are counted in a
    // scalar and multiplied in when the count changes.
    SweepLds& lds = sweep_lds();
    uint4* __restrict__ sp = lds.prog[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
    double* __restrict__ one = lds.one[__builtin_amdgcn_readfirstlane(threadIdx.x >> 6)];
    if constexpr (sizeof(R) == 8) {
        one[0] = 1.0;                         // every lane stores the same pair, so each
        one[1] = 0.0;                         // lane's later read sees its own store
    }
    const uint4* __restrict__ recs = (const uint4*)ops;
    uint32_t c0 = 0;
    uint64_t bmask = 0;               // bit k: op c0 + k is a base-class op
    uint64_t cmask = 0;               // bit k: op c0 + k has base controls
    uint64_t rmask = 0;               // bit k: op c0 + k is a renormalisation
    uint64_t nmask = 0;               // bit k: op c0 + k is a norm point
    auto stage = [&](uint32_t from) __attribute__((always_inline)) {
        c0 = from;
        const uint32_t mine = from + lane < nops ? from + lane : nops - 1;
        uint4 q0 = recs[(size_t)mine * 4], q1 = recs[(size_t)mine * 4 + 1];
        uint4 q2 = recs[(size_t)mine * 4 + 2], q3 = recs[(size_t)mine * 4 + 3];
        const bool ren = (q0.x & 15u) == H_RENORM;
        if (ren) {
            // a renormalisation is the diagonal factor (f, f) with the device
            // computed f = 1/sqrt(norm2): give it that record in LDS (FP32
            // programs: c[0] = (f, 0), c[1] = (f, 0) as FP32 pairs)
            q0.z = 0; q0.w = 0; q1.x = 0; q1.y = 0;
            if constexpr (sizeof(R) == 4) {
                const uint32_t fb = __builtin_bit_cast(uint32_t, (float)*scale);
                q1.z = fb; q1.w = 0; q2.x = fb; q2.y = 0;
                q2.z = 0; q2.w = 0; q3.x = 0; q3.y = 0;
            } else {
                const uint64_t fb = __builtin_bit_cast(uint64_t, *scale);
                const uint32_t lo = (uint32_t)fb, hi = (uint32_t)(fb >> 32);
                q1.z = lo; q1.w = hi; q2.x = 0; q2.y = 0;
                q2.z = lo; q2.w = hi; q3.x = 0; q3.y = 0;
            }
        }
        bmask = __ballot(from + lane < nops && (ren || base_class(q0.x & 15u, q0.y & 0xffu)));
        cmask = __ballot(from + lane < nops && (q0.z | q0.w) != 0);
        rmask = __ballot(from + lane < nops && ren);
        nmask = __ballot(from + lane < nops && (q0.x & 15u) == H_NORM_POINT);
        sp[lane * 4] = q0;
        sp[lane * 4 + 1] = q1;
        sp[lane * 4 + 2] = q2;
        sp[lane * 4 + 3] = q3;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    };
    if (nops) stage(0);
    uint32_t pend = 0;                // gates since cnt[] last changed
    auto fold_upd = [&]() __attribute__((always_inline)) {
        static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
            constexpr int e = decltype(ec)::value;
            out.upd += (unsigned long long)cnt[e] * pend;
        });
        pend = 0;
    };
    auto u64of = [](uint32_t lo, uint32_t hi) __attribute__((always_inline)) {
        return (uint64_t)lo | (uint64_t)hi << 32;
    };
    auto f64of = [](uint32_t lo, uint32_t hi) __attribute__((always_inline)) {
        return __builtin_bit_cast(double, (uint64_t)lo | (uint64_t)hi << 32);
    };
    uint32_t i = 0;
    while (i < nops) {
        // ---- run of base ops: base keys and pending factors only ----
        // one branch-free form for X / diagonal / Y-PERM: factor c0 for a
        // bucket whose target bit is 0, c1 for bit 1 (X: 1 and 1), then the
        // base key flips for X / Y / PERM; lanes whose base controls fail take
        // factor 1 and keep their key.
        // CTRL: the run holds gates with base controls (one more compare per
        // gate); both forms are straight-line, so a trip of four gates is one
        // block and its LDS reads can all be in flight together.
        auto base_op = [&](uint32_t jj, auto ctrl_c) __attribute__((always_inline)) {
            constexpr bool CTRL = decltype(ctrl_c)::value;
            const uint4 q0 = sp[jj * 4], q1 = sp[jj * 4 + 1], q2 = sp[jj * 4 + 2], q3 = sp[jj * 4 + 3];
            const uint32_t hdr = __builtin_amdgcn_readfirstlane(q0.x);
            const uint64_t tbase = 1ull << ((hdr >> 7) & 63u);
            const uint64_t flip = u64of(q3.z, q3.w);          // key bits to toggle (c[4])
            const uint64_t cmB = u64of(q0.z, q0.w), cvB = u64of(q1.x, q1.y);
            const uint32_t tpos = (hdr >> 7) & 63u;
            const bool thi = tpos >= 32;                    // the target bit's word (uniform)
            // FP64: each lane reads its own factor, (c0r, c0i) at byte 24 of the
            // staged record or (c1r, c1i) at byte 40, at an address picked by its
            // base bit (a controlled gate's failing lanes read the identity
            // instead): three address instructions replace eight 32-bit selects
            // of the uniform doubles.  FP32 pairs (bytes 24 and 32) are selected
            // as values, one instruction per float.
            const char* __restrict__ cbase = reinterpret_cast<const char*>(&sp[jj * 4]) + 24;
            static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
                constexpr int e = decltype(ec)::value;
                R fr, fi;
                if constexpr (sizeof(R) == 8) {
                    const uint32_t w = thi ? (uint32_t)(base[e] >> 32) : (uint32_t)base[e];
                    const uint32_t bitv = __builtin_amdgcn_ubfe(w, tpos & 31u, 1u);
                    const R* cp = reinterpret_cast<const R*>(cbase + bitv * 16u);
                    if constexpr (CTRL) {
                        const bool pred = valid[e] && ((base[e] & cmB) == cvB);
                        cp = pred ? cp : one;
                        base[e] = flip_if(base[e], flip, pred);
                    }
                    fr = cp[0];
                    fi = cp[1];
                } else {
                    const bool bit = (base[e] & tbase) != 0;
                    fr = bit ? __builtin_bit_cast(float, q2.x) : __builtin_bit_cast(float, q1.z);
                    fi = bit ? __builtin_bit_cast(float, q2.y) : __builtin_bit_cast(float, q1.w);
                    if constexpr (CTRL) {
                        const bool pred = valid[e] && ((base[e] & cmB) == cvB);
                        fr = pred ? fr : (R)1;
                        fi = pred ? fi : (R)0;
                        base[e] = flip_if(base[e], flip, pred);
                    }
                }
                if constexpr (!CTRL) {
                    // a padding bucket (valid false) holds zeros, so its factor
                    // is irrelevant; its key flips too but is never stored
                    base[e] ^= flip;
                }
                const Cplx<R> f = rot_fma(pr[e], pi[e], Cplx<R>{fr, fi});
                pr[e] = f.x;
                pi[e] = f.y;
            });
        };
        auto base_run = [&](uint32_t jj, uint32_t n, auto ctrl_c) __attribute__((always_inline)) {
            uint32_t k = 0;
            for (; k + 4 <= n; k += 4) {               // four ops per trip: independent work to overlap
                base_op(jj + k, ctrl_c);
                base_op(jj + k + 1, ctrl_c);
                base_op(jj + k + 2, ctrl_c);
                base_op(jj + k + 3, ctrl_c);
            }
            for (; k < n; ++k) base_op(jj + k, ctrl_c);
        };
        // each trip of the loop is either a run of base ops (within the
        // staged chunk) or one slot op, so the amplitude registers pass
        // through one merge per trip
        if (i - c0 >= 64) stage(i);
        const uint32_t jj = i - c0;
        const uint64_t rest = bmask >> jj;              // base-class ops from i on, in this chunk
        const uint32_t n = rest == ~0ull ? 64u : (uint32_t)__builtin_ctzll(~rest);
        if (n != 0) {
            const uint64_t run = n == 64 ? ~0ull : ((1ull << n) - 1);
            if ((cmask >> jj) & run) base_run(jj, n, std::true_type{});
            else base_run(jj, n, std::false_type{});
            i += n;
            pend += n - (uint32_t)__builtin_popcountll((rmask >> jj) & run);   // renormalisations are not gates
        } else if ((nmask >> jj) & 1ull) {
            // ---- norm point: FP64 norm of the state at this point of the
            // program (the frame and the pending factor's phase leave |a|
            // unchanged); reads the amplitude registers only ----
            // (one partial per tile: slot ctx.wave + e; a missing tile writes none)
            const uint32_t k = __builtin_amdgcn_readfirstlane(sp[jj * 4].x) >> 13;
            static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
                constexpr int e = decltype(ec)::value;
                double t = 0.0;
                static_for<S>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    t += (double)a[e][s].x * (double)a[e][s].x + (double)a[e][s].y * (double)a[e][s].y;
                });
                double v = t * ((double)pr[e] * (double)pr[e] + (double)pi[e] * (double)pi[e]);
                for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
                if (lane == 0 && ctx.ckpt && tile[e] >= 0) ctx.ckpt[(size_t)k * ctx.ckpt_stride + ctx.wave + e] = v;
            });
            ++i;
        } else {
            // ---- one slot op: amplitude registers ----
            ++i;
            SOp op;
            uint32_t hdr, ctl;
            if constexpr (SREC) {
                // the record goes straight from memory into scalar registers
                // (uniform address, program read-only during the launch)
                const auto* __restrict__ rec = as_const_space(ops + (c0 + jj));
                hdr = rec->hdr;
                ctl = rec->ctl;
                op.cmB = rec->cmB;
                op.cvB = rec->cvB;
                for (int k = 0; k < 5; ++k) op.c[k] = rec->c[k];
            } else {
                // from the staged copy in LDS, moved to scalar registers
                const uint4 q0 = sp[jj * 4], q1 = sp[jj * 4 + 1], q2 = sp[jj * 4 + 2], q3 = sp[jj * 4 + 3];
                auto sg = [](uint32_t v) __attribute__((always_inline)) { return __builtin_amdgcn_readfirstlane(v); };
                hdr = sg(q0.x);
                ctl = sg(q0.y);
                op.cmB = u64of(sg(q0.z), sg(q0.w));
                op.cvB = u64of(sg(q1.x), sg(q1.y));
                op.c[0] = f64of(sg(q1.z), sg(q1.w));
                op.c[1] = f64of(sg(q2.x), sg(q2.y));
                op.c[2] = f64of(sg(q2.z), sg(q2.w));
                op.c[3] = f64of(sg(q3.x), sg(q3.y));
                op.c[4] = f64of(sg(q3.z), sg(q3.w));
            }
            op.h = hdr & 15u;
            op.cmS = ctl & 0xffu;
            if (op.h == H_FRAME) {
                // apply the per-bucket frame before the next amplitude op
                frame_apply<MB, R, EPT>(a, base, ctx.qm);
            } else {
                ++pend;
                op.cvS = (ctl >> 8) & 0xffu;
                op.j = (hdr >> 4) & 7u;
                op.tbase = 1ull << ((hdr >> 7) & 63u);
                op.ci = coef_index_of(op.c[4]);
                bool pred[EPT];
                static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
                    constexpr int e = decltype(ec)::value;
                    pred[e] = valid[e] && ((base[e] & op.cmB) == op.cvB);
                });
                if (op.h <= H_BASE_PERM) {
                    // diagonal / Y / PERM on a base qubit, restricted by bucket controls
                    static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
                        constexpr int e = decltype(ec)::value;
                        const bool bit = (base[e] & op.tbase) != 0;
                        R cr, ci;
                        if constexpr (sizeof(R) == 4) {
                            // FP32 programs carry c0 / c1 as FP32 pairs in c[0], c[1]
                            const uint64_t w = __builtin_bit_cast(uint64_t, bit ? op.c[1] : op.c[0]);
                            cr = pred[e] ? __builtin_bit_cast(float, (uint32_t)w) : 1.0f;
                            ci = pred[e] ? __builtin_bit_cast(float, (uint32_t)(w >> 32)) : 0.0f;
                        } else {
                            double fr, fi;
                            base_factor(op, bit, pred[e], fr, fi);
                            cr = fr;
                            ci = fi;
                        }
                        static_for<S>([&](auto sc) __attribute__((always_inline)) {
                            constexpr int s = decltype(sc)::value;
                            const bool on = ((uint32_t)s & op.cmS) == op.cvS;
                            a[e][s] = csel(on, rot(cr, ci, a[e][s]), a[e][s]);
                        });
                    });
                } else if constexpr (MB > 0) {
                    const bool branch = branching_handler(op.h);
                    uint32_t nzc[EPT];
                    bool counted = false;
                    if constexpr (BLOCKS && (MB == 4 || MB == 5) && sizeof(R) == 4 && EPT == 1) {
                        if (op.h == H_SLOT_BLOCK) {
                            block_apply<1 << MB>(a[0], pr[0], pi[0], coef + (size_t)op.ci * 8, thr);
                            pend += (uint32_t)op.c[0] - 1;       // the block's gates (lower bound of their nonzeros)
                        } else {
                            counted = slot_dispatch<MB, R, EPT>(op, a, pred, pr, pi, thr, coef, nzc);
                        }
                    } else {
                        counted = slot_dispatch<MB, R, EPT>(op, a, pred, pr, pi, thr, coef, nzc);
                    }
                    if (branch) {
                        // the branching handlers multiplied the pending factor in
                        fold_upd();
                        static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
                            constexpr int e = decltype(ec)::value;
                            pr[e] = (R)1;
                            pi[e] = (R)0;
                            // (a padding bucket holds zeros: its kept count is 0)
                            cnt[e] = counted ? nzc[e] : (valid[e] ? count_nz<MB, R>(a[e]) : 0u);
                        });
                    }
                }
            }
        }
    }
    fold_upd();
    if (ctx.frames) frame_apply<MB, R, EPT>(a, base, ctx.qm);
    out.nnz_out = 0;
    out.empty = 0;
    out.norm = 0.0;
    static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
        constexpr int e = decltype(ec)::value;
        apply_phase<MB>(a[e], pr[e], pi[e]);
        static_for<S>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            out.norm += (double)a[e][s].x * (double)a[e][s].x + (double)a[e][s].y * (double)a[e][s].y;
        });
        if (tile[e] >= 0 && store) {
            // (padding buckets may carry flipped keys; only valid keys are stored)
            const uint64_t b = (uint64_t)tile[e] * TILE + lane;
            C* __restrict__ dst = ctx.out_amp ? (C*)ctx.out_amp : amp;
            static_for<S>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                dst[amp_index(b, s, S)] = a[e][s];
            });
            if (ctx.out_base) ctx.out_base[b] = valid[e] ? base[e] : EMPTY_KEY;
            else if (valid[e] && base[e] != base0[e]) basep[b] = base[e];
        }
        if (ctx.itab) {
            // the index of the stored state, built here instead of by a
            // separate pass (all 64 lanes take part in the cooperative insert)
            const uint64_t b = (uint64_t)(tile[e] >= 0 ? tile[e] : 0) * TILE + lane;
            coop_insert(ctx.itab, ctx.itmask, base[e], b, tile[e] >= 0 && store && valid[e], ctx.err);
        }
        out.nnz_out += cnt[e];
        out.empty += (valid[e] && cnt[e] == 0) ? 1ull : 0ull;
    });
}

// Wait until this wave's vector memory operations have completed
// (s_waitcnt vmcnt(0); the immediate is the gfx9-family encoding, whose
// vmcnt field is bits 3:0 and 15:14 with expcnt / lgkmcnt left at their
// maxima).  The last-workgroup norm fold relies on it: a partial written with
// a device-scope relaxed atomic store is globally visible once the store has
// completed, which is what vmcnt(0) waits for on gfx950, so the relaxed
// counter increment that follows it cannot overtake it.  This is the ordering
// the LLVM AMDGPU memory model gives GFX940+ for agent-scope monotonic
// atomics (stores with sc1 write through to memory; loads of the partials
// by the last workgroup are agent-scope monotonic loads that miss the
// non-coherent caches).  The compiler-level fences keep the store and the
// count on their sides of the wait.  Built for gfx950 only.
#if defined(__HIP_DEVICE_COMPILE__) && !defined(__gfx950__)
#error "the sweep's last-workgroup norm fold is written for gfx950's memory model"
#endif
__device__ __forceinline__ void wait_vector_memory() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Minimum wavefronts per SIMD the sweep kernels are compiled for (caps the
// VGPRs at 512 / value): the amplitude registers of a bucket take 4 * 2^MB
// (FP64) or 2 * 2^MB (FP32) VGPRs; with 16 slots FP64 fits 3 waves (about
// 160 VGPRs) and FP32 runs 4 (a few spilled dwords, measured faster than 3).
// SHORT: the FP32 16-slot sweep of a short program (a few gates, bound by
// HBM rather than instruction issue) is compiled for 3 waves instead, where
// it does not spill: at 4 waves every lane writes and rereads its spill slots
// once per launch, extra traffic comparable to a tenth of the state's bytes.
// BLOCKS: the FP32 16-slot sweep with fused dense blocks is compiled for 3
// waves too (the block's 32 inputs, 32 outputs and 32 accumulators spill at
// 4 waves; measured 1.64x faster on the layered circuit at 3).  32 slots
// (FP32 only): 2 waves (236 VGPRs, no spills; 4 % faster than 1), and 1 with
// fused 32x32 blocks (337 VGPRs).
#ifndef SQE_OCC_F64_MB4
#define SQE_OCC_F64_MB4 3
#endif
#ifndef SQE_OCC_F32_MB4
#define SQE_OCC_F32_MB4 4
#endif
#ifndef SQE_OCC_F32_MB5
#define SQE_OCC_F32_MB5 2
#endif
template <int MB, class R, bool SHORT = false, bool BLOCKS = false>
struct SweepOcc {
    static constexpr int value =
        MB >= 5 ? (BLOCKS ? 1 : SQE_OCC_F32_MB5) : (MB == 4 ? (sizeof(R) == 8 || SHORT || BLOCKS ? SQE_OCC_F64_MB4 : SQE_OCC_F32_MB4) : 3);
};
constexpr uint32_t SHORT_PROGRAM_OPS = 8;     // programs up to this length take the SHORT sweep

// The batched sweep: as the single-circuit sweep, except FP32 16-slot buckets
// at 3 waves per SIMD (at 4 the batch kernel's per-item state pushed about
// 190 bytes per lane of VGPRs to scratch; at 3, 28 bytes) and 32-slot buckets
// at 1.
template <int MB, class R>
struct BatchOcc {
    static constexpr int value = MB >= 5 ? 1 : (MB == 4 && sizeof(R) == 4 ? 3 : SweepOcc<MB, R>::value);
};

// One program over the engine's store: wavefront g of the grid takes tiles
// EPT*g .. EPT*g + EPT-1.
template <int MB, class R, int EPT, bool BLOCKS = false, bool SHORT = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(SweepOcc<MB, R, SHORT, BLOCKS>::value))) void k_sweep(
    SweepArgs A) {
    const uint64_t nb = *A.nb;
    const int64_t ntiles = (int64_t)((nb + 63) >> 6);
    // tile ownership is decided per wavefront from a scalar wave index, so the
    // gate loop is uniform control flow for the compiler
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t g = (int64_t)blockIdx.x * (BLOCK / 64) + wave;
    int64_t tile[EPT];
    bool any = false;
    static_for<EPT>([&](auto ec) __attribute__((always_inline)) {
        constexpr int e = decltype(ec)::value;
        const int64_t t = g * EPT + e;
        tile[e] = t < ntiles ? t : -1;
        any |= t < ntiles;
    });
    SweepCount cnt{0, 0, 0, 0, 0.0};
    const SweepCtx ctx{A.qm, A.frames, A.ckpt_stride, (uint32_t)g, A.out_base, A.out_amp, A.ckpt, A.itab, A.itmask,
                       &A.stats->err};
    if (any)
        sweep_group<MB, R, EPT, BLOCKS, true>(A.base, A.amp, tile, A.ops, A.nops, A.coef, A.thr, A.scale, A.store, cnt,
                                              ctx);
    block_stats(A.bstats, cnt.upd, cnt.nnz_out, cnt.empty, A.store ? cnt.nnz_in + cnt.nnz_out : 0ull);
    if (A.norm_out) {
        // the FP64 norm of the stored state, one partial per workgroup (for a
        // renormalisation right after this launch)
        __shared__ double s_norm[BLOCK / 64];
        double v = cnt.norm;
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
        if ((threadIdx.x & 63) == 0) s_norm[threadIdx.x >> 6] = v;
        __syncthreads();
        __shared__ unsigned int s_last;
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < BLOCK / 64; ++w) t += s_norm[w];
            // The partial goes out with a device-coherent store and is
            // complete (vmcnt 0) before this workgroup counts itself done;
            // the last workgroup reads the partials with device-coherent loads.
            // (A release fence would also write back the XCD's whole L2.)
            __hip_atomic_store(&A.norm_out[blockIdx.x], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wait_vector_memory();
            s_last = __hip_atomic_fetch_add(A.norm_done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                     gridDim.x - 1;
        }
        __syncthreads();
        if (s_last) {
            // every other workgroup's partial is in: fold them here (same
            // fixed order as a separate reduction kernel)
            __shared__ double s_w[BLOCK / 64];
            double acc = 0.0;
            for (uint32_t i = threadIdx.x; i < gridDim.x; i += BLOCK)
                acc += __hip_atomic_load(&A.norm_out[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
            if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = 0.0;
                for (int i = 0; i < BLOCK / 64; ++i) t += s_w[i];
                *A.norm2 = t;
                *A.scale_out = t > 0.0 ? 1.0 / sqrt(t) : 1.0;
                __hip_atomic_store(A.norm_done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Count-only pass (no program, no store): nonzeros and empty buckets.
template <int MB, class R>
__global__ __launch_bounds__(BLOCK) void k_count(SweepArgs A) {
    const uint64_t nb = *A.nb;
    const int64_t ntiles = (int64_t)((nb + 63) >> 6);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = (int64_t)blockIdx.x * (BLOCK / 64) + wave;
    const int64_t tile[1] = {t < ntiles ? t : -1};
    SweepCount cnt{0, 0, 0, 0, 0.0};
    const SweepCtx ctx{A.qm, 0, 0, 0, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
    if (t < ntiles) sweep_group<MB, R, 1>(A.base, A.amp, tile, nullptr, 0, nullptr, A.thr, nullptr, 0, cnt, ctx);
    block_stats(A.bstats, 0ull, cnt.nnz_out, cnt.empty, 0ull);
}

// ---------------------------------------------------------------------------
// Batched sweep: many independent circuits, each with its own bucket region
// and program, advanced by one launch.  One wavefront per work item (a circuit
// and a run of `gran` of its tiles); the host orders the items longest program
// first, so the in-order dispatch of workgroups balances uneven depths and
// sizes.  Per-item counts go to counts[2*item .. 2*item+1] (updates, nonzeros)
// and are folded per circuit by k_batch_fold in a fixed order.
// ---------------------------------------------------------------------------
// One work item: its tiles of its circuit's program (upd / nnz accumulate the
// wavefront's counts).
template <int MB, class R>
__device__ __forceinline__ void batch_item(const BatchArgs& A, const BItem& item, const BTask& t,
                                           unsigned long long& upd, unsigned long long& nnz) {
    auto run = [&](auto ept_c) __attribute__((always_inline)) {
        constexpr int BEPT = decltype(ept_c)::value;
        for (uint32_t g = 0; g < A.gran; g += BEPT) {
            const uint32_t tile = item.tile + g;
            if (tile >= t.ntiles) break;
            int64_t tl[BEPT];
            static_for<BEPT>([&](auto ec) __attribute__((always_inline)) {
                constexpr uint32_t e = decltype(ec)::value;
                tl[e] = (g + e < A.gran && tile + e < t.ntiles) ? (int64_t)(tile + e) : -1;
            });
            SweepCount c{0, 0, 0, 0};
            // (FP32 one-pass renormalisation: norm points write one partial per tile)
            const SweepCtx ctx{t.qm, t.frames, t.ntiles, tile, nullptr, nullptr, t.ckpt, nullptr, 0, nullptr};
            sweep_group<MB, R, BEPT>(t.base, t.amp, tl, t.ops, t.nops, t.coef, A.thr, t.scale, 1, c, ctx);
            upd += c.upd;
            nnz += c.nnz_out;
        }
    };
    // 8-slot buckets with two or more tiles per work item: two tiles per pass
    // (each op decoded once for both; the registers of two 8-slot buckets are
    // those of one 16-slot bucket)
    if constexpr (MB == 3) {
        if (A.gran >= 2) run(std::integral_constant<int, 2>{});
        else run(std::integral_constant<int, 1>{});
    } else {
        run(std::integral_constant<int, 1>{});
    }
}

__device__ __forceinline__ void batch_counts(const BatchArgs& A, uint32_t it, unsigned long long upd,
                                             unsigned long long nnz) {
    for (int d = 32; d >= 1; d >>= 1) {
        upd += __shfl_xor(upd, d);
        nnz += __shfl_xor(nnz, d);
    }
    if ((threadIdx.x & 63) == 0) {
        A.counts[2 * (uint64_t)it] = upd;
        A.counts[2 * (uint64_t)it + 1] = nnz;
    }
}

template <int MB, class R>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(BatchOcc<MB, R>::value))) void k_sweep_batch(
    BatchArgs A) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t it = blockIdx.x * (blockDim.x >> 6) + wave;
    if (it >= A.nitems) return;
    const BItem item = A.items[it];
    const BTask t = A.tasks[item.task];
    unsigned long long upd = 0, nnz = 0;
    batch_item<MB, R>(A, item, t, upd, nnz);
    batch_counts(A, it, upd, nnz);
}

// The whole run's circuits of up to SQE_MIXED_MAX_MB bucket qubits in one
// launch: each wavefront takes the item the grid order gives it (longest
// programs first, whatever their bucket size) and applies the sweep of its
// circuit's bucket size (a wavefront-uniform branch).  Compiled for 3 waves
// per SIMD, the budget of the two-tile 8-slot sweep.
template <class R>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void k_sweep_batch_mixed(BatchArgs A) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t it = blockIdx.x * (blockDim.x >> 6) + wave;
    if (it >= A.nitems) return;
    const BItem item = A.items[it];
    const BTask t = A.tasks[item.task];
    unsigned long long upd = 0, nnz = 0;
    switch (__builtin_amdgcn_readfirstlane(t.qm.mb)) {
    case 0: if constexpr (0 <= SQE_MIXED_MAX_MB) batch_item<0, R>(A, item, t, upd, nnz); break;
    case 1: if constexpr (1 <= SQE_MIXED_MAX_MB) batch_item<1, R>(A, item, t, upd, nnz); break;
    case 2: if constexpr (2 <= SQE_MIXED_MAX_MB) batch_item<2, R>(A, item, t, upd, nnz); break;
    case 3: if constexpr (3 <= SQE_MIXED_MAX_MB) batch_item<3, R>(A, item, t, upd, nnz); break;
    case 4: if constexpr (4 <= SQE_MIXED_MAX_MB) batch_item<4, R>(A, item, t, upd, nnz); break;
    default: break;
    }
    batch_counts(A, it, upd, nnz);
}

// FP64 sum of |a|^2 of each work item's tiles -> partials[item] (FP32 renorm);
// k_norm_batch_final folds them per circuit.
template <int MB, class R>
__global__ __launch_bounds__(BLOCK) void k_norm_batch(BatchArgs A) {
    using C = typename CT<R>::T;
    constexpr int S = 1 << MB;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (uint32_t it = blockIdx.x * (BLOCK / 64) + wave; it < A.nitems; it += gridDim.x * (BLOCK / 64)) {
        const BItem item = A.items[it];
        const BTask t = A.tasks[item.task];
        double acc = 0.0;
        for (uint32_t g = 0; g < A.gran; ++g) {
            const uint32_t tile = item.tile + g;
            if (tile >= t.ntiles) break;
            const uint64_t b = (uint64_t)tile * TILE + lane;
            const C* __restrict__ amp = (const C*)t.amp;
            static_for<S>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                const C v = amp[amp_index(b, s, S)];
                acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
            });
        }
        for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
        if (lane == 0) A.partials[it] = acc;
    }
}

// ---------------------------------------------------------------------------
// Cross-bucket gate (target is a base qubit): pairs bucket `base` with bucket
// `base ^ (1 << target)` found through the hash index.  The bucket with target
// bit 0 owns the pair when it exists, otherwise the bit-1 bucket does
// (the reference's ownership rule, emulator/device/device_table_FC.cuh:95-108).
// Outputs are compacted into the other store with a single-pass scan;
// all-zero buckets are dropped.
// ---------------------------------------------------------------------------
template <int MB, class R, uint32_t K>
__global__ __launch_bounds__(BLOCK) void k_pair(PairArgs A) {
    using C = typename CT<R>::T;
    constexpr int S = 1 << MB;
    const uint32_t tile = take_ticket(A.scan);
    const uint64_t b = (uint64_t)tile * BLOCK + threadIdx.x;
    const uint64_t nb = *A.in_nb;
    const uint64_t nb_pad = (nb + 63) & ~63ull;
    const bool live = b < nb_pad;
    const C* __restrict__ in_amp = (const C*)A.in_amp;
    C* __restrict__ out_amp = (C*)A.out_amp;
    const DOp& op = A.op;

    const uint64_t base = live ? A.in_base[b] : EMPTY_KEY;
    const bool valid = base != EMPTY_KEY;
    const bool ctrl_ok = valid && ((base & op.cmB) == op.cvB);
    const uint64_t pbase = base ^ op.tbase;
    const uint64_t j = coop_find(A.tab, A.tmask, pbase, ctrl_ok, &A.stats->err);
    const bool bit = (base & op.tbase) != 0;
    const bool owner = ctrl_ok && !(bit && j != NONE);

    C own[S], par[S];
    static_for<S>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        own[s].x = 0; own[s].y = 0; par[s].x = 0; par[s].y = 0;
    });
    // a bucket whose pair its partner owns is read once, by the owner: it
    // loads nothing itself, and the owner counts the partner's nonzeros
    const bool reads = owner || (valid && !ctrl_ok);
    if (reads) {
        static_for<S>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            own[s] = in_amp[amp_index(b, s, S)];
        });
    }
    uint32_t nz_in = reads ? count_nz<MB, R>(own) : 0u;
    if (owner && j != NONE) {
        static_for<S>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            par[s] = in_amp[amp_index(j, s, S)];
        });
        nz_in += count_nz<MB, R>(par);
    }
    uint32_t count = 0;
    bool lo_nz = false, hi_nz = false;
    if (owner) {
        // own/par -> lo (target bit 0) / hi (target bit 1), in place
        static_for<S>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            const C x = own[s], y = par[s];
            C v0 = bit ? y : x, v1 = bit ? x : y;
            if (((uint32_t)s & op.cmS) == op.cvS) {
                C n0, n1;
                any_pair<K>(op, A.thr, v0, v1, n0, n1);
                v0 = n0; v1 = n1;
            }
            own[s] = v0; par[s] = v1;
            lo_nz |= nonzero(v0);
            hi_nz |= nonzero(v1);
        });
        count = (lo_nz ? 1u : 0u) + (hi_nz ? 1u : 0u);
    } else if (valid && !ctrl_ok) {
        count = nz_in ? 1u : 0u;      // untouched bucket is carried over (dropped if all zero)
    }
    unsigned long long total = 0;
    unsigned long long off = block_scan(count, A.scan, tile, gridDim.x, A.out_nb, &total);
    uint32_t nz_out = 0;
    if (count) {
        if (owner) {
            if (lo_nz) {
                if (off < A.out_cap) {
                    A.out_base[off] = base & ~op.tbase;
                    static_for<S>([&](auto sc) __attribute__((always_inline)) {
                        constexpr int s = decltype(sc)::value;
                        out_amp[amp_index(off, s, S)] = own[s];
                    });
                }
                ++off;
            }
            if (hi_nz && off < A.out_cap) {
                A.out_base[off] = base | op.tbase;
                static_for<S>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    out_amp[amp_index(off, s, S)] = par[s];
                });
            }
            nz_out = (lo_nz ? count_nz<MB, R>(own) : 0u) + (hi_nz ? count_nz<MB, R>(par) : 0u);
        } else if (off < A.out_cap) {
            A.out_base[off] = base;
            static_for<S>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                out_amp[amp_index(off, s, S)] = own[s];
            });
            nz_out = nz_in;
        }
    }
    // the last tile pads the output up to a whole tile of buckets
    if (tile == gridDim.x - 1) {
        if (total > A.out_cap) {
            if (threadIdx.x == 0) atomicOr(&A.stats->err, ERR_CAP);
        } else {
            const uint64_t pad_end = (total + 63) & ~63ull;
            for (uint64_t p = total + threadIdx.x; p < pad_end && p < A.out_cap; p += BLOCK) {
                A.out_base[p] = EMPTY_KEY;
                static_for<S>([&](auto sc) __attribute__((always_inline)) {
                    constexpr int s = decltype(sc)::value;
                    C z; z.x = 0; z.y = 0;
                    out_amp[amp_index(p, s, S)] = z;
                });
            }
        }
    }
    block_stats(A.bstats, nz_in, nz_out, 0ull, (unsigned long long)nz_in + nz_out, true);
    (void)nb;
}

// Drop all-zero buckets (threshold pruning leaves them behind) with a scan.
template <int MB, class R>
__global__ __launch_bounds__(BLOCK) void k_compact(CompactArgs A) {
    using C = typename CT<R>::T;
    constexpr int S = 1 << MB;
    const uint32_t tile = take_ticket(A.scan);
    const uint64_t b = (uint64_t)tile * BLOCK + threadIdx.x;
    const uint64_t nb_pad = (*A.in_nb + 63) & ~63ull;
    const bool live = b < nb_pad;
    const C* __restrict__ in_amp = (const C*)A.in_amp;
    C* __restrict__ out_amp = (C*)A.out_amp;
    const uint64_t base = live ? A.in_base[b] : EMPTY_KEY;
    const bool valid = base != EMPTY_KEY;
    C a[S];
    static_for<S>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        if (valid) a[s] = in_amp[amp_index(b, s, S)]; else { a[s].x = 0; a[s].y = 0; }
    });
    const uint32_t nz = count_nz<MB, R>(a);
    const uint32_t count = (valid && nz) ? 1u : 0u;
    unsigned long long total = 0;
    const unsigned long long off = block_scan(count, A.scan, tile, gridDim.x, A.out_nb, &total);
    if (count && off < A.out_cap) {
        A.out_base[off] = base;
        static_for<S>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            out_amp[amp_index(off, s, S)] = a[s];
        });
    }
    if (tile == gridDim.x - 1) {
        const uint64_t pad_end = (total + 63) & ~63ull;
        for (uint64_t p = total + threadIdx.x; p < pad_end && p < A.out_cap; p += BLOCK) {
            A.out_base[p] = EMPTY_KEY;
            static_for<S>([&](auto sc) __attribute__((always_inline)) {
                constexpr int s = decltype(sc)::value;
                C z; z.x = 0; z.y = 0;
                out_amp[amp_index(p, s, S)] = z;
            });
        }
    }
    block_stats(A.bstats, 0ull, count ? nz : 0u, 0ull);
}

// Export non-negligible (or all nonzero) entries as FP64 (key, re, im) arrays
// in bucket order (the reference's readout filter, emulator/device/table.cuh:58-73).
template <int MB, class R>
__global__ __launch_bounds__(BLOCK) void k_export(ExportArgs A) {
    using C = typename CT<R>::T;
    constexpr int S = 1 << MB;
    const uint32_t tile = take_ticket(A.scan);
    static_assert(SNAP_WORDS == 4, "snapshot source selection below");
    if (A.snap && blockIdx.x == 0 && threadIdx.x < SNAP_WORDS) {
        const unsigned long long* src = A.snap_src[0];
        if (threadIdx.x == 1) src = A.snap_src[1];
        if (threadIdx.x == 2) src = A.snap_src[2];
        if (threadIdx.x == 3) src = A.snap_src[3];
        A.snap[threadIdx.x] = *src;
    }
    const uint64_t b = (uint64_t)tile * BLOCK + threadIdx.x;
    const uint64_t nb_pad = (*A.nb + 63) & ~63ull;
    const bool live = b < nb_pad;
    const C* __restrict__ amp = (const C*)A.amp;
    const uint64_t base = live ? A.base[b] : EMPTY_KEY;
    const bool valid = base != EMPTY_KEY;
    C a[S];
    uint32_t keep = 0, count = 0;
    static_for<S>([&](auto sc) __attribute__((always_inline)) {
        constexpr int s = decltype(sc)::value;
        if (valid) a[s] = amp[amp_index(b, s, S)]; else { a[s].x = 0; a[s].y = 0; }
        const cd v = mkd((double)a[s].x, (double)a[s].y);
        const bool k = (A.filter ? !negl64(v, A.thr) : nonzero(v)) &&
                       (((base | pdep_q((uint32_t)s, A.qm)) & A.kmask) == A.kval);
        if (valid && k) { keep |= 1u << s; ++count; }
    });
    unsigned long long total = 0;
    unsigned long long off = block_scan(count, A.scan, tile, gridDim.x, A.total, &total);
    if (count) {
        static_for<S>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            if ((keep >> s) & 1u) {
                if (off < A.cap) {
                    A.keys[off] = base | pdep_q((uint32_t)s, A.qm);
                    A.re[off] = (double)a[s].x;
                    A.im[off] = (double)a[s].y;
                }
                ++off;
            }
        });
    }
}

// Norm: per-block FP64 partial sums in a fixed order, then one block.
template <int MB, class R>
__global__ __launch_bounds__(BLOCK) void k_norm_partial(NormArgs A) {
    using C = typename CT<R>::T;
    constexpr int S = 1 << MB;
    __shared__ double s_w[BLOCK / 64];
    const uint64_t b = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint64_t nb_pad = (*A.nb + 63) & ~63ull;
    const C* __restrict__ amp = (const C*)A.amp;
    double acc = 0.0;
    if (b < nb_pad) {
        static_for<S>([&](auto sc) __attribute__((always_inline)) {
            constexpr int s = decltype(sc)::value;
            const C v = amp[amp_index(b, s, S)];
            acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        });
    }
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < BLOCK / 64; ++i) t += s_w[i];
        A.partials[blockIdx.x] = t;
    }
}

// Scatter each input amplitude into its bucket slot; a second write to the
// same (bucket, slot) flags a duplicate key.
template <class R>
__global__ __launch_bounds__(BLOCK) void k_load_scatter(LoadArgs A) {
    using C = typename CT<R>::T;
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const bool act = i < A.n && !(A.keys[i] & A.keymask);     // (a key beyond nq was reported at insert)
    const uint64_t key = act ? A.keys[i] : EMPTY_KEY;
    const uint64_t j = coop_find(A.tab, A.tmask, key & ~A.qm.mask, act, &A.stats->err);
    if (!act) return;
    if (j == NONE) { atomicOr(&A.stats->err, ERR_MISSING); return; }
    const uint32_t s = pext_q(key, A.qm);
    const uint32_t old = atomicOr(&A.occ[j], 1u << s);
    if (old & (1u << s)) { atomicOr(&A.stats->err, ERR_DUP); return; }
    C v;
    v.x = (R)A.re[i];
    v.y = (R)A.im[i];
    ((C*)A.amp)[amp_index(j, (int)s, 1 << A.qm.mb)] = v;
}


// ---------------------------------------------------------------------------
// per-precision launchers (MB = 5 with FP64 is not instantiated: register budget)
// ---------------------------------------------------------------------------
#define SQE_MB_SWITCH(mb, ...)                                                     \
    switch (mb) {                                                                  \
    case 0: { constexpr int MB_ = 0; __VA_ARGS__; } break;                         \
    case 1: { constexpr int MB_ = 1; __VA_ARGS__; } break;                         \
    case 2: { constexpr int MB_ = 2; __VA_ARGS__; } break;                         \
    case 3: { constexpr int MB_ = 3; __VA_ARGS__; } break;                         \
    case 4: { constexpr int MB_ = 4; __VA_ARGS__; } break;                         \
    case 5: { constexpr int MB_ = (sizeof(R) == 8) ? 4 : 5;                        \
              if (sizeof(R) == 8) return hipErrorInvalidValue; __VA_ARGS__; } break;  \
    default: return hipErrorInvalidValue;                                          \
    }

template <class R>
struct Launch {
    // Buckets (tiles) per thread for the single-program sweep: one keeps the
    // kernels within 3-4 wavefronts per SIMD.
    static constexpr int EPT = 1;
    static hipError_t sweep(int mb, hipStream_t st, uint64_t ntiles, const SweepArgs& a) {
        const uint32_t grid = sweep_grid(mb, ntiles);
        if (a.nops == 0 && !a.store) {
            SQE_MB_SWITCH(mb, hipLaunchKernelGGL((k_count<MB_, R>), dim3(grid), dim3(BLOCK), 0, st, a));
        } else if ((mb == 4 || mb == 5) && a.blocks) {
            if constexpr (sizeof(R) == 4) {
                if (mb == 4) hipLaunchKernelGGL((k_sweep<4, R, EPT, true>), dim3(grid), dim3(BLOCK), 0, st, a);
                else hipLaunchKernelGGL((k_sweep<5, R, EPT, true>), dim3(grid), dim3(BLOCK), 0, st, a);
            } else {
                return hipErrorInvalidValue;
            }
        } else if (sizeof(R) == 4 && mb == 4 && a.nops <= SHORT_PROGRAM_OPS) {
            if constexpr (sizeof(R) == 4)
                hipLaunchKernelGGL((k_sweep<4, R, EPT, false, true>), dim3(grid), dim3(BLOCK), 0, st, a);
        } else {
            SQE_MB_SWITCH(mb, hipLaunchKernelGGL((k_sweep<MB_, R, EPT>), dim3(grid), dim3(BLOCK), 0, st, a));
        }
        return hipGetLastError();
    }
    static uint32_t sweep_grid(int, uint64_t ntiles) {
        const uint64_t waves = (ntiles + EPT - 1) / EPT;
        return (uint32_t)std::max<uint64_t>(1, (waves + BLOCK / 64 - 1) / (BLOCK / 64));
    }
    template <int MB>
    static void pair_kind(uint32_t kind, hipStream_t st, uint32_t grid, const PairArgs& a) {
        switch (kind) {
        case KX: hipLaunchKernelGGL((k_pair<MB, R, KX>), dim3(grid), dim3(BLOCK), 0, st, a); break;
        case KY: hipLaunchKernelGGL((k_pair<MB, R, KY>), dim3(grid), dim3(BLOCK), 0, st, a); break;
        case KPERM: hipLaunchKernelGGL((k_pair<MB, R, KPERM>), dim3(grid), dim3(BLOCK), 0, st, a); break;
        case KH: hipLaunchKernelGGL((k_pair<MB, R, KH>), dim3(grid), dim3(BLOCK), 0, st, a); break;
        case KRX: hipLaunchKernelGGL((k_pair<MB, R, KRX>), dim3(grid), dim3(BLOCK), 0, st, a); break;
        case KRY: hipLaunchKernelGGL((k_pair<MB, R, KRY>), dim3(grid), dim3(BLOCK), 0, st, a); break;
        case KU: hipLaunchKernelGGL((k_pair<MB, R, KU>), dim3(grid), dim3(BLOCK), 0, st, a); break;
        default: break;
        }
    }
    static hipError_t pair(int mb, uint32_t kind, hipStream_t st, uint32_t grid, const PairArgs& a) {
        SQE_MB_SWITCH(mb, pair_kind<MB_>(kind, st, grid, a));
        return hipGetLastError();
    }
    static hipError_t compact(int mb, hipStream_t st, uint32_t grid, const CompactArgs& a) {
        SQE_MB_SWITCH(mb, hipLaunchKernelGGL((k_compact<MB_, R>), dim3(grid), dim3(BLOCK), 0, st, a));
        return hipGetLastError();
    }
    static hipError_t exportk(int mb, hipStream_t st, uint32_t grid, const ExportArgs& a) {
        SQE_MB_SWITCH(mb, hipLaunchKernelGGL((k_export<MB_, R>), dim3(grid), dim3(BLOCK), 0, st, a));
        return hipGetLastError();
    }
    static hipError_t norm_partial(int mb, hipStream_t st, uint32_t grid, const NormArgs& a) {
        SQE_MB_SWITCH(mb, hipLaunchKernelGGL((k_norm_partial<MB_, R>), dim3(grid), dim3(BLOCK), 0, st, a));
        return hipGetLastError();
    }
    static hipError_t sweep_batch(int mb, hipStream_t st, uint32_t grid, uint32_t block, const BatchArgs& a) {
        if (mb == MB_MIXED) {
            hipLaunchKernelGGL(k_sweep_batch_mixed<R>, dim3(grid), dim3(block), 0, st, a);
            return hipGetLastError();
        }
        SQE_MB_SWITCH(mb, hipLaunchKernelGGL((k_sweep_batch<MB_, R>), dim3(grid), dim3(block), 0, st, a));
        return hipGetLastError();
    }
    static hipError_t norm_batch_partial(int mb, hipStream_t st, uint32_t grid, const BatchArgs& a) {
        SQE_MB_SWITCH(mb, hipLaunchKernelGGL((k_norm_batch<MB_, R>), dim3(grid), dim3(BLOCK), 0, st, a));
        return hipGetLastError();
    }
    static hipError_t scatter(hipStream_t st, uint32_t grid, const LoadArgs& a) {
        hipLaunchKernelGGL(k_load_scatter<R>, dim3(grid), dim3(BLOCK), 0, st, a);
        return hipGetLastError();
    }
};

}  // namespace sqe
