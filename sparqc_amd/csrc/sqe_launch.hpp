// Kernel argument blocks and launchers shared by the engine host code and the
// kernels (the only interface between sqe_engine.cpp and sqe_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sqe_device.hpp"

namespace sqe {

struct SweepArgs {
    uint64_t* base;
    void* amp;
    const unsigned long long* nb;
    const MOp* ops;
    const double* coef;     // side table: 8 coefficients per general-U gate, fused blocks
    uint32_t nops;
    int32_t store;
    int32_t blocks;         // the program holds fused blocks (FP32, 16 slots)
    double* norm_out;       // non-null: per-workgroup FP64 partial norms of the stored state;
                            // the last workgroup to finish folds them into *norm2 / *scale_out
    unsigned int* norm_done;// workgroups finished (zero between launches; the last one resets it)
    double* norm2;
    double* scale_out;
    double thr;
    const double* scale;
    Stats* stats;
    BlockStats* bstats;
    QMap qm;                // bucket qubits (positions of the frame's X part)
    int32_t frames;         // the program holds frame records: apply the frame before amplitude ops
    uint32_t ckpt_stride;   // wavefronts of the grid (row length of ckpt)
    uint64_t* out_base;     // non-null: store the result here (out of place) instead of in place
    void* out_amp;
    double* ckpt;           // [checkpoint * ckpt_stride + wavefront]: FP64 norms of H_NORM_POINT ops
    Slot* itab;             // non-null: insert (base key -> bucket) of every stored bucket into this cleared index
    uint64_t itmask;
};

// Per-launch context of sweep_group beyond the program.
struct SweepCtx {
    QMap qm;
    int32_t frames;
    uint32_t ckpt_stride;
    uint32_t wave;          // this wavefront's index in the grid
    uint64_t* out_base;
    void* out_amp;
    double* ckpt;
    Slot* itab;             // non-null: index the stored buckets (base key -> bucket) here
    uint64_t itmask;
    unsigned long long* err;
};

// One circuit's share of a batched launch.
struct BTask {
    uint64_t* base;              // bucket base keys of the circuit's region
    void* amp;                   // its amplitudes (tile-interleaved)
    const MOp* ops;              // program segment of this launch
    const double* coef;          // side table of general-U coefficients
    double* scale;               // renormalisation factor slot
    unsigned long long* upd;     // per-circuit counters (folded after each launch)
    unsigned long long* nnz;
    double* norm2;               // per-circuit FP64 norm
    uint32_t nops;
    uint32_t ntiles;
    uint32_t item0;              // first work item of this circuit (items are grouped by circuit)
    uint32_t nitem;
    QMap qm;                     // the circuit's bucket qubits
    int32_t frames;              // its program holds frame records
    int32_t pad;
    double* ckpt;                // FP32 one-pass renormalisation: norm partials, ntiles per checkpoint
    uint32_t nck;                // its checkpoints
    uint32_t pad2;
    unsigned long long* err;     // the circuit's error word
};

struct BItem {                   // a circuit and the first of `gran` consecutive tiles
    uint32_t task;
    uint32_t tile;
};

struct BatchArgs {
    const BTask* tasks;
    const BItem* items;
    uint32_t nitems;
    uint32_t gran;
    double thr;
    double* partials;            // per-item FP64 partial sums (norm)
    unsigned long long* counts;  // per-item (updates, nonzeros) of a sweep
};

struct PairArgs {
    const uint64_t* in_base;
    const void* in_amp;
    const unsigned long long* in_nb;
    const Slot* tab;
    uint64_t tmask;
    DOp op;
    double thr;
    uint64_t* out_base;
    void* out_amp;
    unsigned long long* out_nb;
    uint64_t out_cap;
    ScanCtl scan;
    Stats* stats;
    BlockStats* bstats;
};

struct CompactArgs {
    const uint64_t* in_base;
    const void* in_amp;
    const unsigned long long* in_nb;
    uint64_t* out_base;
    void* out_amp;
    unsigned long long* out_nb;
    uint64_t out_cap;
    ScanCtl scan;
    Stats* stats;
    BlockStats* bstats;
};

constexpr int SNAP_WORDS = 4;

struct ExportArgs {
    const uint64_t* base;
    const void* amp;
    const unsigned long long* nb;
    QMap qm;
    double thr;
    int32_t filter;
    uint64_t kmask, kval;        // keep only keys with (key & kmask) == kval
    uint64_t* keys;
    double* re;
    double* im;
    uint64_t cap;
    unsigned long long* total;
    ScanCtl scan;
    // optional: the first workgroup copies *snap_src[k] to snap[k] (a
    // circuit's counters, taken in stream order with its export)
    const unsigned long long* snap_src[SNAP_WORDS];
    unsigned long long* snap;
};

// An export's entries from device staging to host-mapped (or device)
// arrays: hdr[0] entries (at most cap) of keys / re / im, then the header's
// 1 + SNAP_WORDS words (entry count, counter snapshot) to rec.  vec: every
// array 16-byte aligned (16-byte copies).
struct CopyOutArgs {
    const unsigned long long* hdr;
    const uint64_t* keys;
    const double* re;
    const double* im;
    uint64_t* okeys;
    double* ore;
    double* oim;
    uint64_t cap;
    unsigned long long* rec;
    int32_t vec;
};

// Up to FILL_RANGES clears in one launch: p[i] = (i odd ? v1 : v0).
struct FillRange {
    uint64_t* p;
    uint64_t n;
    uint64_t v0, v1;
};
constexpr int FILL_RANGES = 8;
struct FillSet {
    FillRange r[FILL_RANGES];
    uint32_t k;
};

struct NormArgs {
    const void* amp;
    const unsigned long long* nb;
    double* partials;
};

struct IndexArgs {
    const uint64_t* base;
    const unsigned long long* nb;
    Slot* tab;
    uint64_t tmask;
    Stats* stats;
};

struct LoadArgs {
    const uint64_t* keys;
    const double* re;
    const double* im;
    uint64_t n;
    uint64_t keymask;            // key bits that must be clear (0: checked on the host)
    QMap qm;
    Slot* tab;
    uint64_t tmask;
    void* amp;
    unsigned int* occ;
    Stats* stats;
};

struct EnumArgs {
    Slot* tab;
    uint64_t tcap;
    ScanCtl scan;
    uint64_t* out_base;
    unsigned long long* out_nb;
    uint64_t out_cap;
    Stats* stats;
};

hipError_t launch_fill_u64(hipStream_t st, uint64_t* p, size_t n, uint64_t v);
hipError_t launch_fill_slots(hipStream_t st, Slot* p, size_t n);
hipError_t launch_fill_set(hipStream_t st, const FillSet& s);
hipError_t launch_copy_out(hipStream_t st, uint32_t grid, const CopyOutArgs& a);
hipError_t launch_sweep(int mb, bool f32, hipStream_t st, uint64_t ntiles, const SweepArgs& a);
uint32_t sweep_grid(int mb, bool f32, uint64_t ntiles);   // workgroups launch_sweep uses
hipError_t launch_pair(int mb, bool f32, uint32_t kind, hipStream_t st, uint32_t grid, const PairArgs& a);
hipError_t launch_compact(int mb, bool f32, hipStream_t st, uint32_t grid, const CompactArgs& a);
hipError_t launch_export(int mb, bool f32, hipStream_t st, uint32_t grid, const ExportArgs& a);
// measurement sampling over n staged entries: `shots` keys into out; scratch
// local[n], tile_sum / offsets[sample_tiles(n)], total[1]
hipError_t launch_sample(hipStream_t st, const uint64_t* keys, const double* re, const double* im, uint64_t n,
                         double* local, double* tile_sum, double* offsets, double* total, uint64_t seed,
                         uint64_t shots, uint64_t* out);
uint64_t sample_tiles(uint64_t n);
hipError_t launch_norm(int mb, bool f32, hipStream_t st, uint32_t grid, const NormArgs& a, double* norm2,
                       double* scale);
hipError_t launch_index_build(hipStream_t st, uint32_t grid, const IndexArgs& a);
hipError_t launch_index_enumerate(hipStream_t st, uint32_t grid, const EnumArgs& e);
hipError_t launch_ckpt_batch(hipStream_t st, const BTask* tasks, uint32_t ntasks, double tol);
// norm2 and scale from n per-workgroup partials (of a sweep with norm_out)
hipError_t launch_norm_final(hipStream_t st, const double* partials, uint32_t n, double* norm2, double* scale);
// One wavefront per work item: grid = ceil(nitems / (block / 64)).
hipError_t launch_sweep_batch(int mb, bool f32, hipStream_t st, uint32_t block, const BatchArgs& a);
// Per-circuit FP64 norms of the tasks' regions -> norm2 and scale = 1/sqrt(norm2),
// summed in a fixed order (independent of the other circuits in the launch).
hipError_t launch_norm_batch(int mb, bool f32, hipStream_t st, uint32_t grid, uint32_t ntasks, const BatchArgs& a);
// Fold the sweep's per-item counts per circuit: *upd += updates, *nnz = nonzeros.
hipError_t launch_batch_fold(hipStream_t st, const BTask* tasks, uint32_t ntasks, const unsigned long long* counts);
hipError_t launch_stats_reduce(hipStream_t st, BlockStats* bs, uint32_t n_acc, uint32_t n_last, Stats* stats);
hipError_t launch_load(bool f32, hipStream_t st, const LoadArgs& a, const EnumArgs& e, uint32_t enum_grid);
// Checkpoint norms of a one-pass renormalised run -> norm2 / scale of the last
// checkpoint and the verdict flag (0: the run stands); see k_ckpt_fold.
// ckpt holds nk rows of nw partials followed by room for nk row sums
hipError_t launch_ckpt_fold(hipStream_t st, double* ckpt, uint32_t nw, uint32_t nk, double tol, int force_fail,
                            double* norm2, double* scale, unsigned long long* flag);

// Fixed per-device scratch of the allocation-free entry points: room for
// NOALLOC_PARTIALS FP64 per-workgroup partials (128 KB).  noalloc_scratch_init
// looks up its device address on `device` once (at engine creation);
// noalloc_scratch returns it (nullptr before the lookup).
constexpr uint32_t NOALLOC_PARTIALS = 16384;
hipError_t noalloc_scratch_init(int device);
double* noalloc_scratch(int device);

}  // namespace sqe
