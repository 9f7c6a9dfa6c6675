// Non-templated kernels (fills, hash index, load path, norm finish) and the
// public launchers of the sparse statevector engine (gfx950).
//
// Reference path replaced (ivanNovikau/SparQC, emulator/device/):
//   device_init_arrays / set zero  device_data.cuh:48-59,
//                                  device_table_FC.cuh:39-49    -> k_fill_*, load path
#include "sqe_kernels.hpp"

namespace sqe {

extern template struct Launch<double>;
extern template struct Launch<float>;

// ---------------------------------------------------------------------------
// fills (used instead of memset so that captured graphs never hold memset nodes)
// ---------------------------------------------------------------------------
__global__ void k_fill_u64(uint64_t* __restrict__ p, size_t n, uint64_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_slots(Slot* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        Slot s;
        s.key = EMPTY_KEY;
        s.val = NONE;
        p[i] = s;
    }
}

__global__ void k_fill_set(FillSet S) {
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
    for (uint32_t j = 0; j < S.k && j < FILL_RANGES; ++j) {
        uint64_t* __restrict__ p = S.r[j].p;
        const uint64_t v0 = S.r[j].v0, v1 = S.r[j].v1;
        for (size_t i = t0; i < S.r[j].n; i += stride) p[i] = (i & 1) ? v1 : v0;
    }
}

// Device staging -> host-mapped arrays.  Each lane moves 16 bytes of each
// array per pass, so a wavefront writes 1 KiB contiguous runs across the host
// link.
__global__ __launch_bounds__(BLOCK) void k_copy_out(CopyOutArgs A) {
    const unsigned long long total = A.hdr[0];
    const uint64_t n = total < A.cap ? total : A.cap;
    const uint64_t t0 = (uint64_t)blockIdx.x * BLOCK + threadIdx.x, stride = (uint64_t)gridDim.x * BLOCK;
    if (A.vec) {
        const uint64_t nc = n >> 1;
        for (uint64_t i = t0; i < nc; i += stride) {
            ((uint4*)A.okeys)[i] = ((const uint4*)A.keys)[i];
            ((uint4*)A.ore)[i] = ((const uint4*)A.re)[i];
            ((uint4*)A.oim)[i] = ((const uint4*)A.im)[i];
        }
        if ((n & 1) && t0 == 0) {
            A.okeys[n - 1] = A.keys[n - 1];
            A.ore[n - 1] = A.re[n - 1];
            A.oim[n - 1] = A.im[n - 1];
        }
    } else {
        for (uint64_t i = t0; i < n; i += stride) {
            A.okeys[i] = A.keys[i];
            A.ore[i] = A.re[i];
            A.oim[i] = A.im[i];
        }
    }
    if (t0 <= (uint64_t)SNAP_WORDS) A.rec[t0] = A.hdr[t0];
}

__global__ __launch_bounds__(BLOCK) void k_norm_final(const double* __restrict__ partials, uint32_t n,
                                                      double* __restrict__ out_norm2, double* __restrict__ out_scale) {
    __shared__ double s_w[BLOCK / 64];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) acc += partials[i];
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < BLOCK / 64; ++i) t += s_w[i];
        *out_norm2 = t;
        *out_scale = t > 0.0 ? 1.0 / sqrt(t) : 1.0;
    }
}

// One-pass renormalisation of a run of gates: checkpoint j's norm N_j is the
// fixed-order sum of its per-wavefront partials ckpt[j * nw + w], one
// workgroup per checkpoint (k_ckpt_rows).  The run took every prune decision
// after checkpoint j on amplitudes that still lack the scale 1/sqrt(N_j) the
// step-by-step run applies there (the run's amplitudes carry the product of
// the earlier scales, which N_j includes).  The run stands when each such
// scale is within tol of 1 (k_ckpt_fold): then a decision can differ from the
// step-by-step run's only for a value within that relative distance of the
// threshold (its contribution to the state is below the threshold), while a
// larger scale would move decisions materially and the run is redone step by
// step.  Writes the norm and the scale 1/sqrt(N) of the last checkpoint,
// which the run's result still needs, and the verdict (*flag = 0: stands,
// 1: rerun step by step).
__global__ __launch_bounds__(BLOCK) void k_ckpt_rows(const double* __restrict__ ckpt, uint32_t nw,
                                                     double* __restrict__ norms) {
    __shared__ double s_w[BLOCK / 64];
    const uint32_t j = blockIdx.x;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < nw; i += BLOCK) acc += ckpt[(size_t)j * nw + i];
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < BLOCK / 64; ++i) t += s_w[i];
        norms[j] = t;
    }
}

__global__ void k_ckpt_fold(const double* __restrict__ norms, uint32_t nk, double tol, int force_fail,
                            double* __restrict__ norm2, double* __restrict__ scale,
                            unsigned long long* __restrict__ flag) {
    if (threadIdx.x != 0) return;
    double last = 1.0;
    bool bad = false;
    for (uint32_t j = 0; j < nk; ++j) {
        const double t = norms[j];
        last = t;
        bad = bad || !(t > 0.0) || !(fabs(1.0 / sqrt(t) - 1.0) < tol);
    }
    *norm2 = last;
    *scale = last > 0.0 ? 1.0 / sqrt(last) : 1.0;
    *flag = (bad || force_fail) ? 1ull : 0ull;
}

hipError_t launch_ckpt_fold(hipStream_t st, double* ckpt, uint32_t nw, uint32_t nk, double tol, int force_fail,
                            double* norm2, double* scale, unsigned long long* flag) {
    double* norms = ckpt + (size_t)nk * nw;
    if (nk) hipLaunchKernelGGL(k_ckpt_rows, dim3(nk), dim3(BLOCK), 0, st, ckpt, nw, norms);
    hipLaunchKernelGGL(k_ckpt_fold, dim3(1), dim3(64), 0, st, norms, nk, tol, force_fail, norm2, scale, flag);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Hash index and load path
// ---------------------------------------------------------------------------
// (base of bucket b) -> b for every live bucket of a store.
__global__ __launch_bounds__(BLOCK) void k_index_build(IndexArgs A) {
    const uint64_t b = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    const uint64_t nb_pad = (*A.nb + 63) & ~63ull;
    const uint64_t base = b < nb_pad ? A.base[b] : EMPTY_KEY;
    coop_insert(A.tab, A.tmask, base, b, base != EMPTY_KEY, &A.stats->err);
}

// Insert the base (key with bucket bits cleared) of every input key.
// A key with bits in keymask (beyond the circuit's qubits) is reported and
// left out.
__global__ __launch_bounds__(BLOCK) void k_load_insert(LoadArgs A) {
    const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    bool act = i < A.n;
    const uint64_t key = act ? A.keys[i] : EMPTY_KEY;
    if (act && (key & A.keymask)) {
        atomicOr(&A.stats->err, ERR_KEY);
        act = false;
    }
    coop_insert(A.tab, A.tmask, key & ~A.qm.mask, NONE, act, &A.stats->err);
}

// Number the occupied index slots in slot order (scan) -> bucket ids.
__global__ __launch_bounds__(BLOCK) void k_index_enumerate(EnumArgs A) {
    const uint32_t tile = take_ticket(A.scan);
    const uint64_t i = (uint64_t)tile * BLOCK + threadIdx.x;
    const bool in = i < A.tcap;
    const uint64_t key = in ? A.tab[i].key : EMPTY_KEY;
    const uint32_t count = key != EMPTY_KEY ? 1u : 0u;
    unsigned long long total = 0;
    const unsigned long long id = block_scan(count, A.scan, tile, gridDim.x, A.out_nb, &total);
    if (count) {
        A.tab[i].val = id;
        if (id < A.out_cap) A.out_base[id] = key; else atomicOr(&A.stats->err, ERR_CAP);
    }
    if (tile == gridDim.x - 1) {
        const uint64_t pad_end = (total + 63) & ~63ull;
        for (uint64_t p = total + threadIdx.x; p < pad_end && p < A.out_cap; p += BLOCK) A.out_base[p] = EMPTY_KEY;
    }
}

// Fold the per-block slots: upd / moved over every slot used since the last
// fold (then cleared), nnz / empty over the slots of the last counting launch.
__global__ __launch_bounds__(BLOCK) void k_stats_reduce(BlockStats* __restrict__ bs, uint32_t n_acc,
                                                        uint32_t n_last, Stats* __restrict__ st) {
    __shared__ unsigned long long s_red[5][BLOCK / 64];
    unsigned long long u = 0, m = 0, z = 0, e = 0, mp = 0;
    for (uint32_t i = threadIdx.x; i < n_acc; i += BLOCK) {
        u += bs[i].upd;
        m += bs[i].moved;
        mp += bs[i].moved_pair;
        bs[i].upd = 0;
        bs[i].moved = 0;
        bs[i].moved_pair = 0;
    }
    for (uint32_t i = threadIdx.x; i < n_last; i += BLOCK) {
        z += bs[i].nnz;
        e += bs[i].empty;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        u += __shfl_xor(u, d);
        m += __shfl_xor(m, d);
        z += __shfl_xor(z, d);
        e += __shfl_xor(e, d);
        mp += __shfl_xor(mp, d);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_red[0][w] = u;
        s_red[1][w] = m;
        s_red[2][w] = z;
        s_red[3][w] = e;
        s_red[4][w] = mp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < 5; ++k)
            for (int i = 0; i < BLOCK / 64; ++i) t[k] += s_red[k][i];
        st->updates += t[0];
        st->moved += t[1];
        st->moved_pair += t[4];
        if (n_last) {
            st->nnz = t[2];
            st->empty = t[3];
        }
    }
}

// ---------------------------------------------------------------------------
// public launchers
// ---------------------------------------------------------------------------
static inline uint32_t fill_grid(size_t n) {
    size_t g = (n + BLOCK - 1) / BLOCK;
    if (g > 4096) g = 4096;
    if (g == 0) g = 1;
    return (uint32_t)g;
}

hipError_t launch_fill_u64(hipStream_t st, uint64_t* p, size_t n, uint64_t v) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(n)), dim3(BLOCK), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_set(hipStream_t st, const FillSet& s) {
    size_t most = 0;
    for (uint32_t j = 0; j < s.k && j < FILL_RANGES; ++j) most = s.r[j].n > most ? s.r[j].n : most;
    if (!most) return hipSuccess;
    hipLaunchKernelGGL(k_fill_set, dim3(fill_grid(most)), dim3(BLOCK), 0, st, s);
    return hipGetLastError();
}

hipError_t launch_copy_out(hipStream_t st, uint32_t grid, const CopyOutArgs& a) {
    hipLaunchKernelGGL(k_copy_out, dim3(grid ? grid : 1), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_fill_slots(hipStream_t st, Slot* p, size_t n) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_slots, dim3(fill_grid(n)), dim3(BLOCK), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_sweep(int mb, bool f32, hipStream_t st, uint64_t ntiles, const SweepArgs& a) {
    return f32 ? Launch<float>::sweep(mb, st, ntiles, a) : Launch<double>::sweep(mb, st, ntiles, a);
}

uint32_t sweep_grid(int mb, bool f32, uint64_t ntiles) {
    return f32 ? Launch<float>::sweep_grid(mb, ntiles) : Launch<double>::sweep_grid(mb, ntiles);
}

hipError_t launch_pair(int mb, bool f32, uint32_t kind, hipStream_t st, uint32_t grid, const PairArgs& a) {
    if (!grid) return hipSuccess;
    if (!(is_perm(kind) || is_branch(kind))) return hipErrorInvalidValue;
    return f32 ? Launch<float>::pair(mb, kind, st, grid, a) : Launch<double>::pair(mb, kind, st, grid, a);
}

hipError_t launch_compact(int mb, bool f32, hipStream_t st, uint32_t grid, const CompactArgs& a) {
    if (!grid) return hipSuccess;
    return f32 ? Launch<float>::compact(mb, st, grid, a) : Launch<double>::compact(mb, st, grid, a);
}

hipError_t launch_export(int mb, bool f32, hipStream_t st, uint32_t grid, const ExportArgs& a) {
    if (!grid) return hipSuccess;
    return f32 ? Launch<float>::exportk(mb, st, grid, a) : Launch<double>::exportk(mb, st, grid, a);
}

hipError_t launch_norm(int mb, bool f32, hipStream_t st, uint32_t grid, const NormArgs& a, double* norm2, double* scale) {
    if (!grid) return hipSuccess;
    hipError_t r = f32 ? Launch<float>::norm_partial(mb, st, grid, a) : Launch<double>::norm_partial(mb, st, grid, a);
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(BLOCK), 0, st, (const double*)a.partials, grid, norm2, scale);
    return hipGetLastError();
}

hipError_t launch_norm_final(hipStream_t st, const double* partials, uint32_t n, double* norm2, double* scale) {
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(BLOCK), 0, st, partials, n, norm2, scale);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Measurement sampling over the staged export (keys, re, im; n entries):
// per-tile inclusive prefix sums of |a|^2, a fixed-order scan of the tile
// sums, then one thread per shot draws a counter-based uniform and
// binary-searches the cumulative distribution.
// ---------------------------------------------------------------------------
constexpr uint32_t SAMPLE_PER_THREAD = 16;
constexpr uint32_t SAMPLE_TILE = BLOCK * SAMPLE_PER_THREAD;

__global__ __launch_bounds__(BLOCK) void k_prob_tiles(const double* __restrict__ re, const double* __restrict__ im,
                                                      uint64_t n, double* __restrict__ local,
                                                      double* __restrict__ tile_sum) {
    __shared__ double s_off[BLOCK];
    const uint64_t t0 = (uint64_t)blockIdx.x * SAMPLE_TILE + (uint64_t)threadIdx.x * SAMPLE_PER_THREAD;
    double run[SAMPLE_PER_THREAD];
    double acc = 0.0;
    for (uint32_t k = 0; k < SAMPLE_PER_THREAD; ++k) {
        const uint64_t i = t0 + k;
        if (i < n) acc += re[i] * re[i] + im[i] * im[i];
        run[k] = acc;
    }
    s_off[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {                       // exclusive scan of the thread sums, in order
        double o = 0.0;
        for (uint32_t t = 0; t < BLOCK; ++t) {
            const double v = s_off[t];
            s_off[t] = o;
            o += v;
        }
    }
    __syncthreads();
    const double off = s_off[threadIdx.x];
    for (uint32_t k = 0; k < SAMPLE_PER_THREAD; ++k) {
        const uint64_t i = t0 + k;
        if (i < n) {
            const double v = off + run[k];
            local[i] = v;
            if (i == n - 1 || k == SAMPLE_PER_THREAD - 1 && threadIdx.x == BLOCK - 1) tile_sum[blockIdx.x] = v;
        }
    }
}

// offsets[t] = sum of tile_sum[0..t) in order; total = offsets[nt-1] + tile_sum[nt-1].
__global__ void k_prob_offsets(const double* __restrict__ tile_sum, uint32_t nt, double* __restrict__ offsets,
                               double* __restrict__ total) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double o = 0.0;
    for (uint32_t t = 0; t < nt; ++t) {
        offsets[t] = o;
        o = o + tile_sum[t];
    }
    *total = o;
}

__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(BLOCK) void k_sample(const uint64_t* __restrict__ keys, const double* __restrict__ local,
                                                  const double* __restrict__ offsets, uint32_t nt,
                                                  const double* __restrict__ total, uint64_t n, uint64_t seed,
                                                  uint64_t shots, uint64_t* __restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= shots) return;
    const double u = (double)(splitmix_at(seed, k) >> 11) * (1.0 / 9007199254740992.0);
    const double target = u * *total;
    // last tile whose offset is <= target
    uint32_t lo = 0, hi = nt;                    // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (offsets[mid] <= target) lo = mid; else hi = mid;
    }
    const double base = offsets[lo];
    const uint64_t first = (uint64_t)lo * SAMPLE_TILE;
    const uint64_t last = first + SAMPLE_TILE < n ? first + SAMPLE_TILE : n;   // [first, last)
    // first entry of the tile whose cumulative probability exceeds target
    uint64_t a = first, b = last;
    while (a < b) {
        const uint64_t mid = (a + b) / 2;
        if (base + local[mid] > target) b = mid; else a = mid + 1;
    }
    out[k] = keys[a < last ? a : last - 1];
}

hipError_t launch_sample(hipStream_t st, const uint64_t* keys, const double* re, const double* im, uint64_t n,
                         double* local, double* tile_sum, double* offsets, double* total, uint64_t seed,
                         uint64_t shots, uint64_t* out) {
    if (!n || !shots) return hipSuccess;
    const uint32_t nt = (uint32_t)((n + SAMPLE_TILE - 1) / SAMPLE_TILE);
    hipLaunchKernelGGL(k_prob_tiles, dim3(nt), dim3(BLOCK), 0, st, re, im, n, local, tile_sum);
    hipLaunchKernelGGL(k_prob_offsets, dim3(1), dim3(64), 0, st, (const double*)tile_sum, nt, offsets, total);
    hipLaunchKernelGGL(k_sample, dim3((uint32_t)((shots + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, keys,
                       (const double*)local, (const double*)offsets, nt, (const double*)total, n, seed, shots, out);
    return hipGetLastError();
}

uint64_t sample_tiles(uint64_t n) { return (n + SAMPLE_TILE - 1) / SAMPLE_TILE; }

// One wavefront per circuit: fold its items' partial sums in a fixed order.
__global__ __launch_bounds__(BLOCK) void k_norm_batch_final(const BTask* __restrict__ tasks, uint32_t ntasks,
                                                            const double* __restrict__ partials) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t ti = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    if (ti >= ntasks) return;
    const BTask t = tasks[ti];
    double acc = 0.0;
    for (uint32_t i = lane; i < t.nitem; i += 64) acc += partials[t.item0 + i];
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
    if (lane == 0) {
        *t.norm2 = acc;
        *t.scale = acc > 0.0 ? 1.0 / sqrt(acc) : 1.0;
    }
}

// FP32 one-pass renormalisation of batched circuits, one wavefront per
// circuit: checkpoint k's norm N_k is the fixed-order sum of the circuit's
// per-tile partials ckpt[k * ntiles + tile] (the same for the circuit alone
// or in any mix); the scale of the last checkpoint, 1/sqrt(N), goes to the
// circuit's scale slot for its final renormalisation record.  A checkpoint
// whose scale lies more than tol from 1 (the prune decisions of the pass
// would then lack a material scale) marks the circuit with ERR_RENORM.
__global__ __launch_bounds__(BLOCK) void k_ckpt_batch(const BTask* __restrict__ tasks, uint32_t ntasks, double tol) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t ti = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    if (ti >= ntasks) return;
    const BTask t = tasks[ti];
    double last = 1.0;
    bool bad = false;
    for (uint32_t k = 0; k < t.nck; ++k) {
        double acc = 0.0;
        for (uint32_t i = lane; i < t.ntiles; i += 64) acc += t.ckpt[(size_t)k * t.ntiles + i];
        for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d);
        last = acc;
        bad = bad || !(acc > 0.0) || !(fabs(1.0 / sqrt(acc) - 1.0) < tol);
    }
    if (lane == 0) {
        *t.norm2 = last;
        *t.scale = last > 0.0 ? 1.0 / sqrt(last) : 1.0;
        if (bad) atomicOr(t.err, ERR_RENORM);
    }
}

hipError_t launch_ckpt_batch(hipStream_t st, const BTask* tasks, uint32_t ntasks, double tol) {
    if (!ntasks) return hipSuccess;
    hipLaunchKernelGGL(k_ckpt_batch, dim3((ntasks + BLOCK / 64 - 1) / (BLOCK / 64)), dim3(BLOCK), 0, st, tasks, ntasks,
                       tol);
    return hipGetLastError();
}

// One wavefront per circuit: fold its items' counts in a fixed order (each
// circuit appears in one task of a launch, so the plain update is race-free).
__global__ __launch_bounds__(BLOCK) void k_batch_fold(const BTask* __restrict__ tasks, uint32_t ntasks,
                                                      const unsigned long long* __restrict__ counts) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t ti = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    if (ti >= ntasks) return;
    const BTask t = tasks[ti];
    unsigned long long u = 0, z = 0;
    for (uint32_t i = lane; i < t.nitem; i += 64) {
        u += counts[2 * (uint64_t)(t.item0 + i)];
        z += counts[2 * (uint64_t)(t.item0 + i) + 1];
    }
    for (int d = 32; d >= 1; d >>= 1) {
        u += __shfl_xor(u, d);
        z += __shfl_xor(z, d);
    }
    if (lane == 0) {
        *t.upd += u;
        *t.nnz = z;
    }
}

hipError_t launch_batch_fold(hipStream_t st, const BTask* tasks, uint32_t ntasks, const unsigned long long* counts) {
    if (!ntasks) return hipSuccess;
    hipLaunchKernelGGL(k_batch_fold, dim3((ntasks + BLOCK / 64 - 1) / (BLOCK / 64)), dim3(BLOCK), 0, st, tasks,
                       ntasks, counts);
    return hipGetLastError();
}

hipError_t launch_sweep_batch(int mb, bool f32, hipStream_t st, uint32_t block, const BatchArgs& a) {
    if (!a.nitems) return hipSuccess;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    const uint32_t grid = (a.nitems + block / 64 - 1) / (block / 64);
    return f32 ? Launch<float>::sweep_batch(mb, st, grid, block, a) : Launch<double>::sweep_batch(mb, st, grid, block, a);
}

hipError_t launch_norm_batch(int mb, bool f32, hipStream_t st, uint32_t grid, uint32_t ntasks, const BatchArgs& a) {
    const hipError_t r = f32 ? Launch<float>::norm_batch_partial(mb, st, grid, a)
                             : Launch<double>::norm_batch_partial(mb, st, grid, a);
    if (r != hipSuccess || !ntasks) return r;
    hipLaunchKernelGGL(k_norm_batch_final, dim3((ntasks + BLOCK / 64 - 1) / (BLOCK / 64)), dim3(BLOCK), 0, st,
                       a.tasks, ntasks, (const double*)a.partials);
    return hipGetLastError();
}

hipError_t launch_stats_reduce(hipStream_t st, BlockStats* bs, uint32_t n_acc, uint32_t n_last, Stats* stats) {
    hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(BLOCK), 0, st, bs, n_acc, n_last, stats);
    return hipGetLastError();
}

hipError_t launch_index_build(hipStream_t st, uint32_t grid, const IndexArgs& a) {
    if (!grid) return hipSuccess;
    hipLaunchKernelGGL(k_index_build, dim3(grid), dim3(BLOCK), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_index_enumerate(hipStream_t st, uint32_t grid, const EnumArgs& e) {
    hipLaunchKernelGGL(k_index_enumerate, dim3(grid), dim3(BLOCK), 0, st, e);
    return hipGetLastError();
}

hipError_t launch_load(bool f32, hipStream_t st, const LoadArgs& a, const EnumArgs& e, uint32_t enum_grid) {
    const uint32_t grid = (uint32_t)((a.n + BLOCK - 1) / BLOCK);
    if (grid) hipLaunchKernelGGL(k_load_insert, dim3(grid), dim3(BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_index_enumerate, dim3(enum_grid), dim3(BLOCK), 0, st, e);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess || !grid) return r;
    return f32 ? Launch<float>::scatter(st, grid, a) : Launch<double>::scatter(st, grid, a);
}

}  // namespace sqe
