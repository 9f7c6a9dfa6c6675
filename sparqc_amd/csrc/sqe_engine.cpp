// Host side of the sparse statevector engine: C ABI, planner, memory and
// launch sequencing.  Replaces the reference's TableManager__ / Mapping__ /
// Table__ host layer (ivanNovikau/SparQC emulator/device/table_manager.cuh,
// mapping_GHT.cuh, table.cuh).
//
// Execution model.  Gates are queued on the host.  When the state is observed
// (nnz / norm / export / sync) or the queue is long, the planner walks the
// queue against the bucket-qubit set M: gates that act inside a bucket are
// batched into fused sweeps (one launch applies the whole run in registers);
// a branching gate on a non-bucket qubit, or a permutation whose controls
// include bucket qubits, becomes one cross-bucket launch through the hash
// index.  M is chosen from the branching targets of the first queued batch and
// re-chosen (re-bucketing through the insert path) when later batches would
// otherwise be dominated by cross-bucket gates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sqe.h"
#include "sqe_launch.hpp"
#include "sqe_plan.hpp"

using namespace sqe;

#ifndef SQE_NORM_IN_SWEEP
#define SQE_NORM_IN_SWEEP 1
#endif

namespace {

thread_local std::string g_err;

sqe_status fail(sqe_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

// No C++ exception crosses the C ABI: host allocation failures (and anything
// else a standard container throws) become status codes.
template <class F>
sqe_status guard(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(SQE_ENOMEM, "host allocation failed");
    } catch (const std::exception& ex) {
        return fail(SQE_EINVAL, std::string("host error: ") + ex.what());
    }
}

#define HIPCK(x)                                                                         \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(SQE_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define SQCK(x)                              \
    do {                                     \
        sqe_status s_ = (x);                 \
        if (s_ != SQE_OK) return s_;         \
    } while (0)

inline uint64_t round64(uint64_t x) { return (x + 63) & ~63ull; }
inline uint64_t pow2_at_least(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}
inline uint32_t tiles_of(uint64_t n) { return (uint32_t)((n + BLOCK - 1) / BLOCK); }

struct Store {
    uint64_t* base = nullptr;
    void* amp = nullptr;
    uint64_t cap = 0;   // buckets (multiple of 64)
    size_t abytes = 0;  // bytes of amp
    int mb = -1;        // slot bits the amp buffer was sized for
};

struct Timed {
    hipEvent_t a, b;
    int cls;            // 0 sweep, 1 pair, 2 other
};

enum StepKind { ST_SWEEP, ST_PAIR, ST_NORM };

struct LazyRun {
    bool pending = false;               // a verdict is outstanding
    int in = 0;                         // the store holding the run's input
    uint64_t gates0 = 0, sweeps0 = 0, sweep_ops0 = 0, blocks0 = 0, block_gates0 = 0;
    std::vector<sqe_gate> gl;           // the run's gates
};
struct Step {
    StepKind kind;
    uint64_t off, n;    // sweep: ops [off, off+n) of the uploaded program
    DOp op;             // pair
};

}  // namespace

struct sqe_engine {
    sqe_config cfg;
    bool f32 = false;
    double thr = 1e-14;
    hipStream_t st = nullptr;

    bool have_state = false;
    bool on_device = false;
    QMap qm{};
    Store s[2];
    int cur = 0;
    Store snap;
    bool snap_valid = false;
    QMap snap_qm{};
    uint64_t snap_gates = 0;                // gate count of the snapshot (renormalisation points follow it)

    unsigned long long* meta = nullptr;     // [0],[1] bucket counts of the stores, [2] export total, [3] snapshot count,
                                            // [4] finished-workgroup count of a norm-folding sweep
    Stats* stats = nullptr;
    BlockStats* bstats = nullptr;           // per-workgroup counter slots
    uint64_t bstats_cap = 0;
    uint32_t acc_hi = 0;                    // slots holding unfolded upd / moved
    uint32_t last_grid = 0;                 // grid of the last counting launch
    double* dscal = nullptr;                // [0] norm2, [1] 1/sqrt(norm2)
    double* partials = nullptr;
    uint64_t partials_cap = 0;
    Slot* tab = nullptr;
    uint64_t tcap = 0;
    bool tab_valid = false;
    unsigned long long* scan_status = nullptr;
    uint64_t scan_cap = 0;
    unsigned int* ticket = nullptr;
    uint32_t epoch = 0;
    char* dprog = nullptr;                  // sweep program: MOp[] then coefficient table
    uint64_t dprog_cap = 0;                 // bytes
    char* hprog = nullptr;                  // pinned host staging of the program
    uint64_t hprog_cap = 0;
    hipEvent_t upload_done = nullptr;
    bool upload_pending = false;
    unsigned int* occ = nullptr;
    uint64_t occ_cap = 0;
    uint64_t* xkeys = nullptr;
    double* xre = nullptr;
    double* xim = nullptr;
    uint64_t xcap = 0;

    std::vector<uint64_t> hkeys;
    std::vector<double> hre, him;
    std::vector<sqe_gate> pending;
    uint64_t nb_upper = 0;
    uint64_t gates = 0;
    uint64_t bytes = 0;
    bool renorm_pending = false;
    bool prog_frames = false;               // the uploaded program holds frame records
    LazyRun lz;                             // last one-pass renormalised run, until its verdict is read
    unsigned long long* lz_dev = nullptr;   // its verdict flag (word 0) and Stats backup (words 2-18)
    double* ckpt = nullptr;                 // norm-point partials
    uint64_t ckpt_cap = 0;
    uint64_t n_reruns = 0;                  // one-pass runs rerun step by step

    uint64_t n_sweeps = 0, n_sweep_ops = 0, n_pairs = 0, n_remaps = 0;
    uint64_t n_blocks = 0, n_block_gates = 0;   // fused dense blocks launched, gates they hold
    std::vector<Timed> timed;
    std::vector<hipEvent_t> ev_pool;
    double t_ms[3] = {0, 0, 0};

    int S() const { return 1 << qm.mb; }
    size_t csize() const { return f32 ? 8 : 16; }
    size_t amp_bytes(uint64_t cap) const { return (size_t)cap * (size_t)S() * csize(); }
};

namespace {

sqe_status dalloc(sqe_engine* e, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return SQE_OK;
    if (e->bytes + bytes > e->cfg.hbm_budget_bytes)
        return fail(SQE_ETABLE_FULL, "state needs " + std::to_string(e->bytes + bytes) +
                                         " device bytes, budget is " + std::to_string(e->cfg.hbm_budget_bytes));
    hipError_t r = hipMalloc(p, bytes);
    if (r != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(SQE_ENOMEM, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(r));
    }
    e->bytes += bytes;
    return SQE_OK;
}

void dfree(sqe_engine* e, void* p, size_t bytes) {
    if (p) {
        (void)hipFree(p);
        e->bytes -= bytes;
    }
}

void free_store(sqe_engine* e, Store& s) {
    dfree(e, s.base, s.cap * 8);
    dfree(e, s.amp, s.abytes);
    s = Store{};
}

sqe_status alloc_store(sqe_engine* e, Store& s, uint64_t cap) {
    cap = round64(std::max<uint64_t>(cap, 64));
    Store n;
    n.cap = cap;
    n.abytes = e->amp_bytes(cap);
    n.mb = e->qm.mb;
    SQCK(dalloc(e, (void**)&n.base, cap * 8));
    sqe_status r = dalloc(e, &n.amp, n.abytes);
    if (r != SQE_OK) {
        dfree(e, n.base, cap * 8);
        return r;
    }
    s = n;
    return SQE_OK;
}

// ---- timing -----------------------------------------------------------------
hipEvent_t get_event(sqe_engine* e) {
    if (!e->ev_pool.empty()) {
        hipEvent_t ev = e->ev_pool.back();
        e->ev_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

struct TimeScope {
    sqe_engine* e;
    Timed t{};
    bool on;
    TimeScope(sqe_engine* e_, int cls) : e(e_), on(e_->cfg.timing != 0) {
        if (!on) return;
        t.cls = cls;
        t.a = get_event(e);
        t.b = get_event(e);
        if (!t.a || !t.b) { on = false; return; }
        (void)hipEventRecord(t.a, e->st);
    }
    ~TimeScope() {
        if (!on) return;
        (void)hipEventRecord(t.b, e->st);
        e->timed.push_back(t);
    }
};

void collect_times(sqe_engine* e) {
    for (auto& t : e->timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) e->t_ms[t.cls] += ms;
        e->ev_pool.push_back(t.a);
        e->ev_pool.push_back(t.b);
    }
    e->timed.clear();
}

// ---- scan control -------------------------------------------------------------
sqe_status next_scan(sqe_engine* e, uint32_t ntiles, ScanCtl* out) {
    if (ntiles > e->scan_cap) {
        HIPCK(hipStreamSynchronize(e->st));
        dfree(e, e->scan_status, e->scan_cap * 8);
        e->scan_status = nullptr;
        const uint64_t cap = std::max<uint64_t>(1024, pow2_at_least(ntiles));
        SQCK(dalloc(e, (void**)&e->scan_status, cap * 8));
        e->scan_cap = cap;
        HIPCK(launch_fill_u64(e->st, (uint64_t*)e->scan_status, cap, 0));
        e->epoch = 0;
    }
    if (++e->epoch >= (1u << 22) - 1) {
        HIPCK(launch_fill_u64(e->st, (uint64_t*)e->scan_status, e->scan_cap, 0));
        e->epoch = 1;
    }
    out->status = e->scan_status;
    out->ticket = e->ticket;
    out->err = &e->stats->err;
    out->epoch = e->epoch;
    return SQE_OK;
}

// Make room for `grid` per-workgroup counter slots (folding what they hold first).
sqe_status prep_bstats(sqe_engine* e, uint32_t grid) {
    if (grid <= e->bstats_cap) return SQE_OK;
    if (e->acc_hi) {
        HIPCK(launch_stats_reduce(e->st, e->bstats, e->acc_hi, 0, e->stats));
        e->acc_hi = 0;
    }
    HIPCK(hipStreamSynchronize(e->st));
    dfree(e, e->bstats, e->bstats_cap * sizeof(BlockStats));
    e->bstats = nullptr;
    e->bstats_cap = 0;
    e->last_grid = 0;
    const uint64_t cap = std::max<uint64_t>(4096, pow2_at_least(grid));
    SQCK(dalloc(e, (void**)&e->bstats, cap * sizeof(BlockStats)));
    e->bstats_cap = cap;
    HIPCK(launch_fill_u64(e->st, (uint64_t*)e->bstats, cap * (sizeof(BlockStats) / 8), 0));
    return SQE_OK;
}

void note_counting_launch(sqe_engine* e, uint32_t grid) {
    e->acc_hi = std::max(e->acc_hi, grid);
    e->last_grid = grid;
}

sqe_status reduce_stats(sqe_engine* e) {
    if (!e->acc_hi && !e->last_grid) return SQE_OK;
    HIPCK(launch_stats_reduce(e->st, e->bstats, e->acc_hi, e->last_grid, e->stats));
    e->acc_hi = 0;
    return SQE_OK;
}

sqe_status read_stats(sqe_engine* e, Stats* h) {
    SQCK(reduce_stats(e));
    HIPCK(hipMemcpyAsync(h, e->stats, sizeof(Stats), hipMemcpyDeviceToHost, e->st));
    HIPCK(hipStreamSynchronize(e->st));
    return SQE_OK;
}

sqe_status check_device_errors(sqe_engine* e) {
    Stats h;
    SQCK(read_stats(e, &h));
    if (h.err) {
        const unsigned long long zero = 0;
        HIPCK(hipMemcpy(&e->stats->err, &zero, sizeof(zero), hipMemcpyHostToDevice));
        std::string m = "device reported:";
        if (h.err & ERR_DUP) m += " duplicate key in load;";
        if (h.err & ERR_MISSING) m += " key missing from index;";
        if (h.err & ERR_PROBE) m += " hash probe exhausted;";
        if (h.err & ERR_CAP) m += " output capacity exceeded;";
        if (h.err & ERR_SCAN) m += " scan look-back did not complete;";
        return fail((h.err & ERR_DUP) ? SQE_EINVAL : SQE_EHIP, m);
    }
    return SQE_OK;
}

sqe_status read_nb(sqe_engine* e, uint64_t* nb) {
    unsigned long long v = 0;
    HIPCK(hipMemcpyAsync(&v, &e->meta[e->cur], sizeof(v), hipMemcpyDeviceToHost, e->st));
    HIPCK(hipStreamSynchronize(e->st));
    *nb = v;
    return SQE_OK;
}

// ---- state transfer ---------------------------------------------------------------
sqe_status ensure_staging(sqe_engine* e, uint64_t n) {
    if (n <= e->xcap) return SQE_OK;
    HIPCK(hipStreamSynchronize(e->st));
    dfree(e, e->xkeys, e->xcap * 8);
    dfree(e, e->xre, e->xcap * 8);
    dfree(e, e->xim, e->xcap * 8);
    e->xkeys = nullptr; e->xre = nullptr; e->xim = nullptr; e->xcap = 0;
    const uint64_t cap = std::max<uint64_t>(n, 1024);
    SQCK(dalloc(e, (void**)&e->xkeys, cap * 8));
    SQCK(dalloc(e, (void**)&e->xre, cap * 8));
    SQCK(dalloc(e, (void**)&e->xim, cap * 8));
    e->xcap = cap;
    return SQE_OK;
}

sqe_status ensure_index(sqe_engine* e, uint64_t nkeys) {
    const double ml = e->cfg.max_load > 0.0 && e->cfg.max_load < 0.95 ? e->cfg.max_load : 0.5;
    // (a power of two holding nkeys at load factor ml; an exact power of two of
    // keys must not round up to twice the table)
    const uint64_t want = std::max<uint64_t>(1024, pow2_at_least((uint64_t)std::ceil((double)nkeys / ml)));
    if (want > e->tcap || want * 8 < e->tcap) {
        HIPCK(hipStreamSynchronize(e->st));
        dfree(e, e->tab, e->tcap * sizeof(Slot));
        e->tab = nullptr;
        e->tcap = 0;
        SQCK(dalloc(e, (void**)&e->tab, want * sizeof(Slot)));
        e->tcap = want;
    }
    return SQE_OK;
}

sqe_status ensure_stores(sqe_engine* e, uint64_t cap) {
    cap = round64(std::max<uint64_t>(cap, 256));
    for (int i = 0; i < 2; ++i) {
        if (e->s[i].cap >= cap && e->s[i].mb == e->qm.mb) continue;
        HIPCK(hipStreamSynchronize(e->st));
        free_store(e, e->s[i]);
        SQCK(alloc_store(e, e->s[i], cap));
    }
    return SQE_OK;
}

// Build the bucket store from n FP64 entries staged in xkeys/xre/xim.
sqe_status load_from_staging(sqe_engine* e, uint64_t n) {
    const uint64_t S = (uint64_t)e->S();
    // buckets <= n; leave headroom for growth by cross-bucket gates
    const uint64_t cap = round64(std::max<uint64_t>(n + n / 2, 256));
    if (e->s[0].cap < cap || e->s[1].cap < cap || e->s[0].cap > 4 * cap || e->s[0].mb != e->qm.mb ||
        e->s[1].mb != e->qm.mb) {
        HIPCK(hipStreamSynchronize(e->st));
        free_store(e, e->s[0]);
        free_store(e, e->s[1]);
        SQCK(alloc_store(e, e->s[0], cap));
        SQCK(alloc_store(e, e->s[1], cap));
    }
    e->cur = 0;
    Store& o = e->s[0];
    if (o.cap > e->occ_cap) {
        HIPCK(hipStreamSynchronize(e->st));
        dfree(e, e->occ, e->occ_cap * 4);
        e->occ = nullptr;
        e->occ_cap = 0;
        SQCK(dalloc(e, (void**)&e->occ, o.cap * 4));
        e->occ_cap = o.cap;
    }
    SQCK(ensure_index(e, n));
    HIPCK(launch_fill_slots(e->st, e->tab, e->tcap));
    HIPCK(launch_fill_u64(e->st, (uint64_t*)o.amp, e->amp_bytes(o.cap) / 8, 0));
    HIPCK(launch_fill_u64(e->st, (uint64_t*)e->occ, o.cap / 2, 0));
    LoadArgs la{};
    la.keys = e->xkeys; la.re = e->xre; la.im = e->xim; la.n = n; la.qm = e->qm;
    la.tab = e->tab; la.tmask = e->tcap - 1; la.amp = o.amp; la.occ = e->occ; la.stats = e->stats;
    EnumArgs ea{};
    ea.tab = e->tab; ea.tcap = e->tcap; ea.out_base = o.base; ea.out_nb = &e->meta[0];
    ea.out_cap = o.cap; ea.stats = e->stats;
    const uint32_t eg = tiles_of(e->tcap);
    SQCK(next_scan(e, eg, &ea.scan));
    {
        TimeScope ts(e, 2);
        HIPCK(launch_load(e->f32, e->st, la, ea, eg));
    }
    SQCK(check_device_errors(e));
    uint64_t nb = 0;
    SQCK(read_nb(e, &nb));
    e->nb_upper = nb;
    e->tab_valid = true;
    e->on_device = true;
    (void)S;
    return SQE_OK;
}

sqe_status materialize(sqe_engine* e, const std::vector<sqe_gate>& gl) {
    e->qm = choose_map((int)e->cfg.nq, e->cfg.bucket_bits, gl.data(), gl.size());
    const uint64_t n = e->hkeys.size();
    SQCK(ensure_staging(e, std::max<uint64_t>(n, 1)));
    if (n) {
        HIPCK(hipMemcpyAsync(e->xkeys, e->hkeys.data(), n * 8, hipMemcpyHostToDevice, e->st));
        HIPCK(hipMemcpyAsync(e->xre, e->hre.data(), n * 8, hipMemcpyHostToDevice, e->st));
        HIPCK(hipMemcpyAsync(e->xim, e->him.data(), n * 8, hipMemcpyHostToDevice, e->st));
    }
    SQCK(load_from_staging(e, n));
    std::vector<uint64_t>().swap(e->hkeys);
    std::vector<double>().swap(e->hre);
    std::vector<double>().swap(e->him);
    return SQE_OK;
}

// Export the current store into the FP64 staging arrays; returns the count.
sqe_status export_to_staging(sqe_engine* e, int filter, uint64_t* count, uint64_t kmask = 0, uint64_t kval = 0) {
    uint64_t nb = 0;
    SQCK(read_nb(e, &nb));
    const uint64_t bound = std::max<uint64_t>(nb * (uint64_t)e->S(), 1);
    SQCK(ensure_staging(e, bound));
    ExportArgs xa{};
    const Store& o = e->s[e->cur];
    xa.base = o.base; xa.amp = o.amp; xa.nb = &e->meta[e->cur]; xa.qm = e->qm; xa.thr = e->thr;
    xa.filter = filter; xa.keys = e->xkeys; xa.re = e->xre; xa.im = e->xim; xa.cap = e->xcap;
    xa.kmask = kmask; xa.kval = kval;
    xa.total = &e->meta[2];
    const uint32_t g = std::max<uint32_t>(1, tiles_of(round64(nb)));
    SQCK(next_scan(e, g, &xa.scan));
    {
        TimeScope ts(e, 2);
        HIPCK(launch_export(e->qm.mb, e->f32, e->st, g, xa));
    }
    unsigned long long tot = 0;
    HIPCK(hipMemcpyAsync(&tot, &e->meta[2], sizeof(tot), hipMemcpyDeviceToHost, e->st));
    HIPCK(hipStreamSynchronize(e->st));
    *count = tot;
    return SQE_OK;
}

sqe_status remap(sqe_engine* e, const QMap& nm) {
    uint64_t n = 0;
    SQCK(export_to_staging(e, 0, &n));
    e->qm = nm;
    ++e->n_remaps;
    return load_from_staging(e, n);
}

// Size and clear the index for the current state (before a rebuild, or
// before a sweep that inserts its buckets as it stores them).
sqe_status clear_index(sqe_engine* e) {
    SQCK(ensure_index(e, e->nb_upper));
    HIPCK(launch_fill_slots(e->st, e->tab, e->tcap));
    return SQE_OK;
}

sqe_status rebuild_index(sqe_engine* e) {
    SQCK(clear_index(e));
    IndexArgs ia{};
    const Store& o = e->s[e->cur];
    ia.base = o.base; ia.nb = &e->meta[e->cur]; ia.tab = e->tab; ia.tmask = e->tcap - 1; ia.stats = e->stats;
    {
        TimeScope ts(e, 2);
        HIPCK(launch_index_build(e->st, tiles_of(round64(e->nb_upper)), ia));
    }
    e->tab_valid = true;
    return SQE_OK;
}

// Buckets the key space allows: one per base key, 2^(nq - bucket bits).
uint64_t max_buckets(const sqe_engine* e) {
    const int bits = (int)e->cfg.nq - e->qm.mb;
    return bits >= 62 ? (1ull << 62) : (1ull << std::max(bits, 0));
}

// Stores of at most this many bytes each are sized for the whole key space
// at their first growth: a cross-bucket gate then never needs the exact
// bucket count from the device (a host round trip per gate).
constexpr uint64_t FULL_SPACE_STORE_BYTES = 256ull << 20;

// Grow both stores so that `need` buckets fit, keeping the current state.
// `need` = 0 asks for room for one cross-bucket gate (2 * buckets + 64).
sqe_status grow(sqe_engine* e, uint64_t need) {
    uint64_t nb = 0;
    SQCK(read_nb(e, &nb));
    e->nb_upper = nb;
    if (!need) need = 2 * nb + 64;
    if (need <= e->s[0].cap && need <= e->s[1].cap) return SQE_OK;
    uint64_t cap = round64(std::max<uint64_t>(need, e->s[0].cap * 2));
    if (max_buckets(e) <= (1ull << 32)) {
        const uint64_t full = round64(2 * max_buckets(e) + 64);
        const uint64_t full_bytes = full * 8 + e->amp_bytes(full);
        if (full > cap && full_bytes <= FULL_SPACE_STORE_BYTES &&
            e->bytes + 2 * full_bytes <= e->cfg.hbm_budget_bytes)
            cap = full;
    }
    const int other = e->cur ^ 1;
    free_store(e, e->s[other]);
    SQCK(alloc_store(e, e->s[other], cap));
    const uint64_t nbp = round64(nb);
    HIPCK(hipMemcpyAsync(e->s[other].base, e->s[e->cur].base, nbp * 8, hipMemcpyDeviceToDevice, e->st));
    HIPCK(hipMemcpyAsync(e->s[other].amp, e->s[e->cur].amp, e->amp_bytes(nbp), hipMemcpyDeviceToDevice, e->st));
    HIPCK(hipMemcpyAsync(&e->meta[other], &e->meta[e->cur], 8, hipMemcpyDeviceToDevice, e->st));
    HIPCK(hipStreamSynchronize(e->st));
    free_store(e, e->s[e->cur]);
    SQCK(alloc_store(e, e->s[e->cur], cap));
    e->cur = other;
    return SQE_OK;
}

sqe_status ensure_partials(sqe_engine* e, uint32_t g) {
    if (g <= e->partials_cap) return SQE_OK;
    HIPCK(hipStreamSynchronize(e->st));
    dfree(e, e->partials, e->partials_cap * 8);
    e->partials = nullptr;
    e->partials_cap = 0;
    SQCK(dalloc(e, (void**)&e->partials, (uint64_t)g * 8));
    e->partials_cap = g;
    return SQE_OK;
}

// A sweep with norm_grid != nullptr also writes per-workgroup norm partials
// (into e->partials) and returns their count.
// Options of one sweep launch.
struct SweepOpt {
    bool blocks = false;                // the program holds fused blocks
    uint32_t* norm_grid = nullptr;      // non-null: per-workgroup norm partials into e->partials (count returned)
    int out_store = -1;                 // >= 0: write the result to this store (out of place)
    double* ckpt = nullptr;             // norm-point partials, ckpt_stride per checkpoint
    uint32_t* ckpt_stride = nullptr;    // receives the wavefront count of the grid
    bool index = false;                 // insert every stored bucket into the (cleared) hash index
};

sqe_status run_sweep(sqe_engine* e, const MOp* mops, const double* coef, uint64_t nops, int store,
                     const SweepOpt& opt = SweepOpt{}) {
    SweepArgs a{};
    a.blocks = opt.blocks ? 1 : 0;
    Store& o = e->s[e->cur];
    a.base = o.base; a.amp = o.amp; a.nb = &e->meta[e->cur]; a.ops = mops; a.coef = coef; a.nops = (uint32_t)nops;
    a.store = store; a.thr = e->thr; a.scale = &e->dscal[1]; a.stats = e->stats;
    a.qm = e->qm;
    a.frames = (mops && e->prog_frames) ? 1 : 0;
    if (opt.out_store >= 0) {
        a.out_base = e->s[opt.out_store].base;
        a.out_amp = e->s[opt.out_store].amp;
    }
    const uint64_t ntiles = std::max<uint64_t>(1, round64(e->nb_upper) / 64);
    const uint32_t g = sweep_grid(e->qm.mb, e->f32, ntiles);
    a.ckpt = opt.ckpt;
    a.ckpt_stride = g * (BLOCK / 64);
    if (opt.ckpt_stride) *opt.ckpt_stride = a.ckpt_stride;
    if (opt.index) {
        a.itab = e->tab;
        a.itmask = e->tcap - 1;
    }
    SQCK(prep_bstats(e, g));
    a.bstats = e->bstats;
    if (opt.norm_grid) {
        SQCK(ensure_partials(e, g));
        a.norm_out = e->partials;
        a.norm_done = (unsigned int*)&e->meta[4];
        a.norm2 = &e->dscal[0];
        a.scale_out = &e->dscal[1];
        *opt.norm_grid = g;
    }
    {
        TimeScope ts(e, 0);
        HIPCK(launch_sweep(e->qm.mb, e->f32, e->st, ntiles, a));
    }
    note_counting_launch(e, g);
    return SQE_OK;
}

sqe_status run_pair(sqe_engine* e, const DOp& op) {
    if (!e->tab_valid) SQCK(rebuild_index(e));
    // (grow() moves the state to the other store: pick the output store after it)
    if (2 * e->nb_upper + 64 > e->s[e->cur ^ 1].cap) SQCK(grow(e, 0));
    const int other = e->cur ^ 1;
    PairArgs a{};
    const Store& i = e->s[e->cur];
    const Store& o = e->s[other];
    a.in_base = i.base; a.in_amp = i.amp; a.in_nb = &e->meta[e->cur]; a.tab = e->tab; a.tmask = e->tcap - 1;
    a.op = op; a.thr = e->thr; a.out_base = o.base; a.out_amp = o.amp; a.out_nb = &e->meta[other];
    a.out_cap = o.cap; a.stats = e->stats;
    const uint32_t g = std::max<uint32_t>(1, tiles_of(round64(e->nb_upper)));
    SQCK(next_scan(e, g, &a.scan));
    SQCK(prep_bstats(e, g));
    a.bstats = e->bstats;
    {
        TimeScope ts(e, 1);
        HIPCK(launch_pair(e->qm.mb, e->f32, op.kind, e->st, g, a));
    }
    note_counting_launch(e, g);
    e->cur = other;
    e->nb_upper = std::min<uint64_t>(std::min<uint64_t>(2 * e->nb_upper, o.cap), max_buckets(e));
    e->tab_valid = false;
    ++e->n_pairs;
    return SQE_OK;
}

sqe_status run_norm(sqe_engine* e) {
    const uint32_t g = std::max<uint32_t>(1, tiles_of(round64(e->nb_upper)));
    SQCK(ensure_partials(e, g));
    NormArgs a{};
    a.amp = e->s[e->cur].amp; a.nb = &e->meta[e->cur]; a.partials = e->partials;
    TimeScope ts(e, 2);
    HIPCK(launch_norm(e->qm.mb, e->f32, e->st, g, a, &e->dscal[0], &e->dscal[1]));
    return SQE_OK;
}

sqe_status compact(sqe_engine* e) {
    const int other = e->cur ^ 1;
    CompactArgs a{};
    const Store& i = e->s[e->cur];
    const Store& o = e->s[other];
    a.in_base = i.base; a.in_amp = i.amp; a.in_nb = &e->meta[e->cur]; a.out_base = o.base; a.out_amp = o.amp;
    a.out_nb = &e->meta[other]; a.out_cap = o.cap; a.stats = e->stats;
    const uint32_t g = std::max<uint32_t>(1, tiles_of(round64(e->nb_upper)));
    SQCK(next_scan(e, g, &a.scan));
    SQCK(prep_bstats(e, g));
    a.bstats = e->bstats;
    {
        TimeScope ts(e, 2);
        HIPCK(launch_compact(e->qm.mb, e->f32, e->st, g, a));
    }
    note_counting_launch(e, g);
    e->cur = other;
    e->tab_valid = false;
    return SQE_OK;
}

// Upload an encoded program (records, then the coefficient table) through
// the pinned staging buffer; *dm / *dc receive its device addresses.
sqe_status upload_program(sqe_engine* e, const Encoded& enc, const MOp** dm, const double** dc) {
    const size_t mbytes = enc.ops.size() * sizeof(MOp);
    const size_t cbytes = enc.coef.size() * sizeof(double);
    const size_t coff = (mbytes + 255) & ~(size_t)255;
    const size_t total = coff + cbytes;
    *dm = nullptr;
    *dc = nullptr;
    if (enc.ops.empty()) return SQE_OK;
    if (e->upload_pending) {
        HIPCK(hipEventSynchronize(e->upload_done));
        e->upload_pending = false;
    }
    if (total > e->hprog_cap) {
        if (e->hprog) (void)hipHostFree(e->hprog);
        e->hprog = nullptr;
        e->hprog_cap = 0;
        const uint64_t cap = pow2_at_least(total);
        if (hipHostMalloc((void**)&e->hprog, cap, hipHostMallocDefault) != hipSuccess)
            return fail(SQE_ENOMEM, "pinned program buffer");
        e->hprog_cap = cap;
    }
    if (total > e->dprog_cap) {
        HIPCK(hipStreamSynchronize(e->st));
        dfree(e, e->dprog, e->dprog_cap);
        e->dprog = nullptr;
        e->dprog_cap = 0;
        const uint64_t cap = pow2_at_least(total);
        SQCK(dalloc(e, (void**)&e->dprog, cap));
        e->dprog_cap = cap;
    }
    std::memcpy(e->hprog, enc.ops.data(), mbytes);
    if (cbytes) std::memcpy(e->hprog + coff, enc.coef.data(), cbytes);
    HIPCK(hipMemcpyAsync(e->dprog, e->hprog, total, hipMemcpyHostToDevice, e->st));
    HIPCK(hipEventRecord(e->upload_done, e->st));
    e->upload_pending = true;
    e->prog_frames = enc.framed;
    *dm = (const MOp*)e->dprog;
    *dc = (const double*)(e->dprog + coff);
    return SQE_OK;
}

bool frames_on(const sqe_engine* e) {
    // keys of at most 59 qubits leave the frame's Z bits free
    return !e->cfg.frame_off && e->cfg.nq <= (uint32_t)FRAME_Z_BIT0 && e->qm.mb > 0;
}

void setup_blocks(sqe_engine* e, BlockBuilder& bb, std::vector<BlockMat>* blocks) {
    bb.on = e->cfg.fuse_blocks && e->f32 && (e->qm.mb == 4 || e->qm.mb == 5);
    if (e->cfg.fuse_blocks >= 2) bb.min_gates = e->cfg.fuse_blocks;
    bb.frames = frames_on(e);
    bb.qm = e->qm;
    bb.blocks = blocks;
}

// Plan and launch gates step by step: fused sweeps, cross-bucket launches and,
// with renorm_every, one launch per renormalisation period (norm, then the
// scale enters the next sweep).
sqe_status flush_exact(sqe_engine* e, const std::vector<sqe_gate>& gl) {
    std::vector<DOp> prog;
    std::vector<Step> steps;
    std::vector<BlockMat> blocks;
    BlockBuilder bb;
    setup_blocks(e, bb, &blocks);
    const uint64_t W = e->cfg.fuse_window ? e->cfg.fuse_window : ~0ull;
    uint64_t run_start = 0;
    auto close_run = [&]() {
        bb.flush(prog);
        uint64_t n = prog.size() - run_start;
        uint64_t off = run_start;
        while (n) {
            const uint64_t k = std::min(n, W);
            steps.push_back(Step{ST_SWEEP, off, k, DOp{}});
            off += k;
            n -= k;
        }
        run_start = prog.size();
    };
    auto open_run = [&]() {
        if (e->renorm_pending) {
            prog.push_back(renorm_op());
            e->renorm_pending = false;
        }
    };
    for (const auto& g : gl) {
        if (gate_local(g, e->qm)) {
            if (prog.size() == run_start && !bb.open()) open_run();
            if (!bb.on) prog.push_back(plan_op(g, e->qm));
            else bb.route(g, prog);
        } else {
            if (e->renorm_pending) open_run();
            close_run();
            steps.push_back(Step{ST_PAIR, 0, 0, plan_op(g, e->qm)});
        }
        ++e->gates;
        if (e->cfg.renorm_every && e->gates % e->cfg.renorm_every == 0) {
            close_run();
            steps.push_back(Step{ST_NORM, 0, 0, DOp{}});
            e->renorm_pending = true;
        }
    }
    if (e->renorm_pending) open_run();
    close_run();
    Encoded enc;
    encode_program(prog, blocks, e->f32, frames_on(e) ? &e->qm : nullptr, enc);
    const MOp* dm = nullptr;
    const double* dc = nullptr;
    SQCK(upload_program(e, enc, &dm, &dc));
    uint32_t fused_norm = 0;              // partials of the last sweep, when it computed them
    for (size_t si = 0; si < steps.size(); ++si) {
        const Step& s = steps[si];
        switch (s.kind) {
        case ST_SWEEP: {
            bool has_block = false;
            for (uint64_t i = s.off; i < s.off + s.n; ++i) has_block |= prog[i].kind == K_BLOCK;
            // a renormalisation right after this sweep takes the norm from its
            // epilogue (SQE_NORM_IN_SWEEP=0 at build time: a separate
            // reduction instead, for debugging)
            const bool norm_next = SQE_NORM_IN_SWEEP && si + 1 < steps.size() && steps[si + 1].kind == ST_NORM;
            SweepOpt so;
            so.blocks = has_block;
            so.norm_grid = norm_next ? &fused_norm : nullptr;
            bool flips = false;
            for (uint64_t i = s.off; i < s.off + s.n; ++i) flips |= prog[i].tbase && kind_is_perm(prog[i].kind);
            // a sweep that changes base keys right before a cross-bucket gate
            // rebuilds the index as it stores its buckets (no separate pass)
            so.index = flips && si + 1 < steps.size() && steps[si + 1].kind == ST_PAIR;
            if (so.index) SQCK(clear_index(e));
            SQCK(run_sweep(e, dm + enc.start[s.off], dc, enc.start[s.off + s.n] - enc.start[s.off], 1, so));
            ++e->n_sweeps;
            for (uint64_t i = s.off; i < s.off + s.n; ++i) {
                const DOp& d = prog[i];
                if (d.kind == K_BLOCK) {
                    e->n_sweep_ops += (uint64_t)d.p[1];
                    ++e->n_blocks;
                    e->n_block_gates += (uint64_t)d.p[1];
                } else if (d.kind != K_RENORM && d.kind != K_NORM_POINT) {
                    ++e->n_sweep_ops;
                }
            }
            if (flips) e->tab_valid = so.index;
        } break;
        case ST_PAIR: SQCK(run_pair(e, s.op)); fused_norm = 0; break;
        case ST_NORM:
            if (!fused_norm) SQCK(run_norm(e));   // else the sweep's last workgroup folded the norm
            fused_norm = 0;
            break;
        }
    }
    return reduce_stats(e);
}

// The one-pass form applies when every gate is bucket-local (FP32 with
// renormalisation) and the other store can take the result.
bool lazy_applies(const sqe_engine* e, const std::vector<sqe_gate>& gl) {
    if (!e->f32 || !e->cfg.renorm_every || e->cfg.renorm_mode == 0 || e->renorm_pending) return false;
    const Store& in = e->s[e->cur];
    const Store& out = e->s[e->cur ^ 1];
    if (!out.base || out.cap < round64(e->nb_upper) || out.mb != in.mb || out.mb != e->qm.mb) return false;
    for (const auto& g : gl)
        if (!gate_local(g, e->qm)) return false;
    return true;
}

// One pass over a run of bucket-local gates with renormalisation every R
// gates: the sweep reads the state once (out of place, so the input stays
// for a rerun) and records the FP64 norm at each renormalisation point
// without rescaling; a fold kernel takes the norms, checks that no prune
// decision could have changed under the scales the step-by-step run applies,
// and leaves 1/sqrt(norm) of the last point, which one more in-place pass
// applies.  lazy_settle() reads the verdict and reruns step by step if needed.
sqe_status flush_lazy(sqe_engine* e, const std::vector<sqe_gate>& gl) {
    LazyRun& L = e->lz;
    L.gates0 = e->gates;
    L.in = e->cur;
    L.sweeps0 = e->n_sweeps;
    L.sweep_ops0 = e->n_sweep_ops;
    L.blocks0 = e->n_blocks;
    L.block_gates0 = e->n_block_gates;
    std::vector<DOp> prog;
    std::vector<BlockMat> blocks;
    BlockBuilder bb;
    setup_blocks(e, bb, &blocks);
    uint32_t k = 0;
    for (const auto& g : gl) {
        if (!bb.on) prog.push_back(plan_op(g, e->qm));
        else bb.route(g, prog);
        ++e->gates;
        if (e->gates % e->cfg.renorm_every == 0) {
            bb.flush(prog);
            prog.push_back(norm_point_op(k++));
        }
    }
    bb.flush(prog);
    const size_t n_main = prog.size();
    if (k) prog.push_back(renorm_op());
    Encoded enc;
    encode_program(prog, blocks, e->f32, frames_on(e) ? &e->qm : nullptr, enc);
    const MOp* dm = nullptr;
    const double* dc = nullptr;
    SQCK(upload_program(e, enc, &dm, &dc));
    // device words: [0] verdict flag, [2..] the Stats backup
    if (!e->lz_dev) SQCK(dalloc(e, (void**)&e->lz_dev, 256));
    SQCK(reduce_stats(e));
    HIPCK(hipMemcpyAsync(e->lz_dev + 2, e->stats, sizeof(Stats), hipMemcpyDeviceToDevice, e->st));
    HIPCK(launch_fill_u64(e->st, (uint64_t*)e->lz_dev, 1, 0));
    const int out = e->cur ^ 1;
    HIPCK(hipMemcpyAsync(&e->meta[out], &e->meta[e->cur], 8, hipMemcpyDeviceToDevice, e->st));
    // checkpoint partials: one row of wavefronts per norm point
    const uint64_t ntiles = std::max<uint64_t>(1, round64(e->nb_upper) / 64);
    const uint64_t waves = (uint64_t)sweep_grid(e->qm.mb, e->f32, ntiles) * (BLOCK / 64);
    // (rows of partials, then the rows' sums)
    if ((uint64_t)k * (waves + 1) > e->ckpt_cap) {
        HIPCK(hipStreamSynchronize(e->st));
        dfree(e, e->ckpt, e->ckpt_cap * 8);
        e->ckpt = nullptr;
        e->ckpt_cap = 0;
        const uint64_t cap = pow2_at_least((uint64_t)k * (waves + 1));
        SQCK(dalloc(e, (void**)&e->ckpt, cap * 8));
        e->ckpt_cap = cap;
    }
    // wavefronts of the grid without a tile write no partials: start from zeros
    if (k) HIPCK(launch_fill_u64(e->st, (uint64_t*)e->ckpt, (uint64_t)k * waves, 0));
    const uint64_t W = e->cfg.fuse_window ? e->cfg.fuse_window : ~0ull;
    bool first = true;
    uint32_t stride = 0;
    for (uint64_t off = 0; off < n_main || first;) {
        const uint64_t n = std::min<uint64_t>(n_main - off, W);
        bool has_block = false;
        for (uint64_t i = off; i < off + n; ++i) has_block |= prog[i].kind == K_BLOCK;
        SweepOpt so;
        so.blocks = has_block;
        so.out_store = first ? out : -1;
        so.ckpt = e->ckpt;
        so.ckpt_stride = &stride;
        SQCK(run_sweep(e, dm + enc.start[off], dc, enc.start[off + n] - enc.start[off], 1, so));
        if (first) e->cur = out;
        first = false;
        ++e->n_sweeps;
        for (uint64_t i = off; i < off + n; ++i) {
            const DOp& d = prog[i];
            if (d.kind == K_BLOCK) {
                e->n_sweep_ops += (uint64_t)d.p[1];
                ++e->n_blocks;
                e->n_block_gates += (uint64_t)d.p[1];
            } else if (d.kind != K_RENORM && d.kind != K_NORM_POINT) {
                ++e->n_sweep_ops;
            }
        }
        off += n;
        if (n == 0) break;
    }
    e->tab_valid = false;
    if (k) {
        HIPCK(launch_ckpt_fold(e->st, e->ckpt, stride, k, 0x1p-11, e->cfg.renorm_mode == 2 ? 1 : 0,
                               &e->dscal[0], &e->dscal[1], e->lz_dev));
        SQCK(run_sweep(e, dm + enc.start[n_main], dc, enc.start[n_main + 1] - enc.start[n_main], 1));
        ++e->n_sweeps;
    } else if (e->cfg.renorm_mode == 2) {
        HIPCK(launch_fill_u64(e->st, (uint64_t*)e->lz_dev, 1, 1));
    }
    L.gl = gl;
    L.pending = true;
    return reduce_stats(e);
}

// Read the verdict of the last one-pass run; rerun its gates step by step
// from its input (counters restored) when it does not stand.
sqe_status lazy_settle(sqe_engine* e) {
    LazyRun& L = e->lz;
    if (!L.pending) return SQE_OK;
    L.pending = false;
    unsigned long long flag = 0;
    HIPCK(hipMemcpyAsync(&flag, e->lz_dev, 8, hipMemcpyDeviceToHost, e->st));
    HIPCK(hipStreamSynchronize(e->st));
    if (!flag) return SQE_OK;
    ++e->n_reruns;
    e->cur = L.in;
    e->gates = L.gates0;
    e->n_sweeps = L.sweeps0;
    e->n_sweep_ops = L.sweep_ops0;
    e->n_blocks = L.blocks0;
    e->n_block_gates = L.block_gates0;
    e->renorm_pending = false;
    e->tab_valid = false;
    HIPCK(hipMemcpyAsync(e->stats, e->lz_dev + 2, sizeof(Stats), hipMemcpyDeviceToDevice, e->st));
    if (e->bstats_cap) HIPCK(launch_fill_u64(e->st, (uint64_t*)e->bstats, e->bstats_cap * (sizeof(BlockStats) / 8), 0));
    e->acc_hi = 0;
    e->last_grid = 0;
    std::vector<sqe_gate> gl;
    gl.swap(L.gl);
    return flush_exact(e, gl);
}

// Plan and launch the queued gates.
sqe_status flush(sqe_engine* e) {
    SQCK(lazy_settle(e));
    if (e->pending.empty()) return SQE_OK;
    if (!e->have_state) return fail(SQE_EINVAL, "no state: call sqe_set_basis or sqe_load first");
    std::vector<sqe_gate> gl;
    gl.swap(e->pending);
    if (!e->on_device) {
        SQCK(materialize(e, gl));
    } else if (e->cfg.bucket_bits < 0) {
        const QMap best = choose_map((int)e->cfg.nq, e->cfg.bucket_bits, gl.data(), gl.size());
        const uint64_t now = count_cross(gl, e->qm), then = count_cross(gl, best);
        if (best.mask != e->qm.mask && now >= 8 && then * 2 < now) SQCK(remap(e, best));
    }
    if (lazy_applies(e, gl)) return flush_lazy(e, gl);
    return flush_exact(e, gl);
}

sqe_status enter(sqe_engine* e) {
    if (!e) return fail(SQE_EINVAL, "null engine");
    HIPCK(hipSetDevice(e->cfg.device));
    return SQE_OK;
}

sqe_status full_sync(sqe_engine* e) {
    SQCK(flush(e));
    SQCK(lazy_settle(e));
    HIPCK(hipStreamSynchronize(e->st));
    SQCK(check_device_errors(e));
    collect_times(e);
    return SQE_OK;
}

sqe_status count_now(sqe_engine* e, Stats* h) {
    SQCK(full_sync(e));
    if (!e->on_device) {
        std::memset(h, 0, sizeof(*h));
        h->nnz = e->hkeys.size();
        return SQE_OK;
    }
    SQCK(run_sweep(e, nullptr, nullptr, 0, 0));
    SQCK(read_stats(e, h));
    collect_times(e);
    return SQE_OK;
}

void destroy_all(sqe_engine* e) {
    if (e->st) (void)hipStreamSynchronize(e->st);
    free_store(e, e->s[0]);
    free_store(e, e->s[1]);
    free_store(e, e->snap);
    dfree(e, e->meta, 64);
    dfree(e, e->stats, sizeof(Stats));
    dfree(e, e->bstats, e->bstats_cap * sizeof(BlockStats));
    dfree(e, e->dscal, 16);
    dfree(e, e->partials, e->partials_cap * 8);
    dfree(e, e->tab, e->tcap * sizeof(Slot));
    dfree(e, e->scan_status, e->scan_cap * 8);
    dfree(e, e->ticket, 64);
    dfree(e, e->dprog, e->dprog_cap);
    dfree(e, e->occ, e->occ_cap * 4);
    dfree(e, e->xkeys, e->xcap * 8);
    dfree(e, e->xre, e->xcap * 8);
    dfree(e, e->xim, e->xcap * 8);
    dfree(e, e->lz_dev, 256);
    dfree(e, e->ckpt, e->ckpt_cap * 8);
    if (e->hprog) (void)hipHostFree(e->hprog);
    for (auto& t : e->timed) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto ev : e->ev_pool) (void)hipEventDestroy(ev);
    if (e->upload_done) (void)hipEventDestroy(e->upload_done);
    if (e->st) (void)hipStreamDestroy(e->st);
}

}  // namespace

extern "C" {

void sqe_config_default(sqe_config* c, uint32_t nq) {
    std::memset(c, 0, sizeof(*c));
    c->nq = nq;
    c->prec = SQE_F64;
    c->prune_thr = 0.0;
    c->max_load = 0.5;
    c->hbm_budget_bytes = 0;
    c->device = 0;
    c->bucket_bits = -1;
    c->renorm_every = 0;
    c->renorm_mode = 1;
    c->fuse_window = 0;
    c->timing = 0;
}

sqe_status sqe_create(const sqe_config* cfg, sqe_engine** out) {
    return guard([&]() -> sqe_status {
        if (!cfg || !out) return fail(SQE_EINVAL, "null argument");
        *out = nullptr;
        if (cfg->nq < 1 || cfg->nq > 63) return fail(SQE_EINVAL, "nq must be in 1..63");
        if (cfg->prec != SQE_F64 && cfg->prec != SQE_F32) return fail(SQE_EINVAL, "bad precision");
        if (cfg->bucket_bits > (cfg->prec == SQE_F32 ? 5 : 4)) return fail(SQE_EINVAL, "bucket_bits too large");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
            (void)hipGetLastError();
            return fail(SQE_EDEVICE, "no HIP device");
        }
        if (cfg->device < 0 || cfg->device >= ndev) return fail(SQE_EINVAL, "bad device ordinal");
        HIPCK(hipSetDevice(cfg->device));
        sqe_engine* e = new (std::nothrow) sqe_engine();
        if (!e) return fail(SQE_ENOMEM, "host allocation");
        e->cfg = *cfg;
        e->f32 = cfg->prec == SQE_F32;
        e->thr = cfg->prune_thr > 0.0 ? cfg->prune_thr : (e->f32 ? 1e-7 : 1e-14);
        if (e->cfg.hbm_budget_bytes == 0) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete e; return fail(SQE_EHIP, "hipMemGetInfo"); }
            const double keep = e->cfg.mem_reserve_ppm ? std::min(1.0, e->cfg.mem_reserve_ppm * 1e-6) : 0.1;
            e->cfg.hbm_budget_bytes = (uint64_t)((1.0 - keep) * (double)fr);
        }
        sqe_status r = SQE_OK;
        if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess) r = fail(SQE_EHIP, "stream");
        if (r == SQE_OK && hipEventCreateWithFlags(&e->upload_done, hipEventDisableTiming) != hipSuccess)
            r = fail(SQE_EHIP, "event");
        if (r == SQE_OK) r = dalloc(e, (void**)&e->meta, 64);
        if (r == SQE_OK) r = dalloc(e, (void**)&e->stats, sizeof(Stats));
        if (r == SQE_OK) r = dalloc(e, (void**)&e->dscal, 16);
        if (r == SQE_OK) r = dalloc(e, (void**)&e->ticket, 64);
        if (r == SQE_OK && hipMemsetAsync(e->meta, 0, 64, e->st) != hipSuccess) r = fail(SQE_EHIP, "memset");
        if (r == SQE_OK && hipMemsetAsync(e->stats, 0, sizeof(Stats), e->st) != hipSuccess) r = fail(SQE_EHIP, "memset");
        if (r == SQE_OK && hipMemsetAsync(e->dscal, 0, 16, e->st) != hipSuccess) r = fail(SQE_EHIP, "memset");
        if (r == SQE_OK && hipMemsetAsync(e->ticket, 0, 64, e->st) != hipSuccess) r = fail(SQE_EHIP, "memset");
        if (r == SQE_OK && hipStreamSynchronize(e->st) != hipSuccess) r = fail(SQE_EHIP, "sync");
    // the allocation-free entry points' scratch on this device
    if (r == SQE_OK && noalloc_scratch_init(cfg->device) != hipSuccess) r = fail(SQE_EHIP, "device scratch lookup");
        if (r != SQE_OK) {
            destroy_all(e);
            delete e;
            return r;
        }
        *out = e;
        return SQE_OK;
    });
}

sqe_status sqe_destroy(sqe_engine* e) {
    return guard([&]() -> sqe_status {
        if (!e) return SQE_OK;
        (void)hipSetDevice(e->cfg.device);
        destroy_all(e);
        delete e;
        return SQE_OK;
    });
}

sqe_status sqe_load(sqe_engine* e, const uint64_t* keys, const double* re, const double* im, uint64_t n) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (n && (!keys || !re || !im)) return fail(SQE_EINVAL, "null array");
        const uint64_t lim = (e->cfg.nq >= 64) ? ~0ull : ((1ull << e->cfg.nq) - 1);
        std::vector<uint64_t> k;
        std::vector<double> r, i;
        k.reserve(n); r.reserve(n); i.reserve(n);
        for (uint64_t j = 0; j < n; ++j) {
            if (keys[j] & ~lim) return fail(SQE_EINVAL, "key has bits beyond nq");
            if (re[j] == 0.0 && im[j] == 0.0) continue;
            k.push_back(keys[j]); r.push_back(re[j]); i.push_back(im[j]);
        }
        // queued gates, and the verdict of a one-pass run, belonged to the
        // state being replaced
        e->pending.clear();
        e->lz.pending = false;
        e->lz.gl.clear();
        HIPCK(hipStreamSynchronize(e->st));
        e->hkeys.swap(k);
        e->hre.swap(r);
        e->him.swap(i);
        e->have_state = true;
        e->on_device = false;
        e->tab_valid = false;
        e->renorm_pending = false;
        // a new state starts a new gate count (renormalisation points follow it)
        e->gates = 0;
        return SQE_OK;
    });
}

sqe_status sqe_set_basis(sqe_engine* e, uint64_t key) {
    return guard([&]() -> sqe_status {
        const double one = 1.0, zero = 0.0;
        return sqe_load(e, &key, &one, &zero, 1);
    });
}

sqe_status sqe_apply_gates(sqe_engine* e, const sqe_gate* g, uint64_t n) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (n && !g) return fail(SQE_EINVAL, "null gates");
        const uint64_t lim = (1ull << e->cfg.nq) - 1;
        for (uint64_t j = 0; j < n; ++j) {
            const sqe_gate& x = g[j];
            if (x.kind > SQE_GATE_PERM) return fail(SQE_EINVAL, "gate " + std::to_string(j) + ": bad kind");
            if (x.target >= e->cfg.nq) return fail(SQE_EINVAL, "gate " + std::to_string(j) + ": target out of range");
            if (x.ctrl_mask & ~lim) return fail(SQE_EINVAL, "gate " + std::to_string(j) + ": control out of range");
            if ((x.ctrl_mask >> x.target) & 1ull) return fail(SQE_EINVAL, "gate " + std::to_string(j) + ": target is a control");
            if (x.ctrl_val & ~x.ctrl_mask) return fail(SQE_EINVAL, "gate " + std::to_string(j) + ": ctrl_val outside ctrl_mask");
        }
        e->pending.insert(e->pending.end(), g, g + n);
        // bounded batches bound the host planning memory and the program upload
        if (e->pending.size() >= (1u << 17)) return flush(e);
        return SQE_OK;
    });
}

sqe_status sqe_apply(sqe_engine* e, uint32_t target, uint64_t ctrl_mask, uint64_t ctrl_val, const double u[8],
                     sqe_class cls) {
    return guard([&]() -> sqe_status {
        if (!u) return fail(SQE_EINVAL, "null matrix");
        sqe_gate g{};
        g.target = target;
        g.ctrl_mask = ctrl_mask;
        g.ctrl_val = ctrl_val;
        switch (cls) {
        case SQE_DIAG:
            g.kind = SQE_GATE_DIAG;
            g.p[0] = u[0]; g.p[1] = u[1]; g.p[2] = u[6]; g.p[3] = u[7];
            break;
        case SQE_PERM:
            g.kind = SQE_GATE_PERM;
            g.p[0] = u[2]; g.p[1] = u[3]; g.p[2] = u[4]; g.p[3] = u[5];
            break;
        case SQE_BRANCH:
            g.kind = SQE_GATE_U;
            for (int k = 0; k < 8; ++k) g.p[k] = u[k];
            break;
        default:
            return fail(SQE_EINVAL, "bad gate class");
        }
        return sqe_apply_gates(e, &g, 1);
    });
}

sqe_status sqe_sync(sqe_engine* e) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        SQCK(full_sync(e));
        if (e->on_device) {
            // drop buckets left all-zero by pruning once they are a quarter of the store
            Stats h;
            SQCK(read_stats(e, &h));
            uint64_t nb = 0;
            SQCK(read_nb(e, &nb));
            e->nb_upper = nb;                 // exact again (cross-bucket gates only bound it)
            if (h.empty * 4 > nb && nb > 1024) {
                SQCK(compact(e));
                HIPCK(hipStreamSynchronize(e->st));
                SQCK(read_nb(e, &e->nb_upper));
                collect_times(e);
            }
        }
        return SQE_OK;
    });
}

sqe_status sqe_nnz(sqe_engine* e, uint64_t* n) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (!n) return fail(SQE_EINVAL, "null output");
        Stats h;
        SQCK(count_now(e, &h));
        *n = h.nnz;
        return SQE_OK;
    });
}

sqe_status sqe_norm2(sqe_engine* e, double* out) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (!out) return fail(SQE_EINVAL, "null output");
        SQCK(full_sync(e));
        if (!e->on_device) {
            double s = 0.0;
            for (size_t i = 0; i < e->hkeys.size(); ++i) s += e->hre[i] * e->hre[i] + e->him[i] * e->him[i];
            *out = s;
            return SQE_OK;
        }
        SQCK(run_norm(e));
        HIPCK(hipMemcpyAsync(out, &e->dscal[0], 8, hipMemcpyDeviceToHost, e->st));
        HIPCK(hipStreamSynchronize(e->st));
        collect_times(e);
        return SQE_OK;
    });
}

sqe_status sqe_norm2_noalloc(sqe_engine* e, double* out) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (!out) return fail(SQE_EINVAL, "null output");
        // queued gates would have to be planned and uploaded, which can allocate
        if (!e->pending.empty() || e->lz.pending)
            return fail(SQE_EINVAL, "sqe_norm2_noalloc: gates are queued; call sqe_sync first");
        if (!e->have_state || !e->on_device) {
            double s = 0.0;
            for (size_t i = 0; i < e->hkeys.size(); ++i) s += e->hre[i] * e->hre[i] + e->him[i] * e->him[i];
            *out = s;
            return SQE_OK;
        }
        const uint32_t g = std::max<uint32_t>(1, tiles_of(round64(e->nb_upper)));
        if (g > NOALLOC_PARTIALS)
            return fail(SQE_EINVAL, "sqe_norm2_noalloc: the state spans " + std::to_string(e->nb_upper) +
                                        " buckets; at most " + std::to_string((uint64_t)NOALLOC_PARTIALS * BLOCK) +
                                        " fit its fixed scratch (use sqe_norm2)");
        double* scratch = noalloc_scratch(e->cfg.device);
        if (!scratch) return fail(SQE_EHIP, "sqe_norm2_noalloc: device scratch not set up");
        // every partial the final fold reads is written first by this call
        NormArgs a{};
        a.amp = e->s[e->cur].amp;
        a.nb = &e->meta[e->cur];
        a.partials = scratch;
        HIPCK(launch_norm(e->qm.mb, e->f32, e->st, g, a, &e->dscal[0], &e->dscal[1]));
        HIPCK(hipMemcpyAsync(out, &e->dscal[0], 8, hipMemcpyDeviceToHost, e->st));
        HIPCK(hipStreamSynchronize(e->st));
        return SQE_OK;
    });
}

sqe_status sqe_export_where(sqe_engine* e, uint64_t mask, uint64_t value, uint64_t* keys, double* re, double* im,
                            uint64_t cap, uint64_t* n) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (!n) return fail(SQE_EINVAL, "null output");
        if (value & ~mask) return fail(SQE_EINVAL, "pattern value has bits outside its mask");
        SQCK(full_sync(e));
        if (!e->have_state) { *n = 0; return SQE_OK; }
        if (!e->on_device) {
            // nothing applied yet: filter the staged host copy
            auto keep = [&](size_t i) {
                return !(std::fabs(e->hre[i]) < e->thr && std::fabs(e->him[i]) < e->thr) &&
                       (e->hkeys[i] & mask) == value;
            };
            uint64_t c = 0;
            for (size_t i = 0; i < e->hkeys.size(); ++i) c += keep(i) ? 1 : 0;
            *n = c;
            if (c > cap) return fail(SQE_EINVAL, "export capacity too small");
            uint64_t j = 0;
            for (size_t i = 0; i < e->hkeys.size(); ++i) {
                if (!keep(i)) continue;
                double r = e->hre[i], m = e->him[i];
                if (e->f32) { r = (double)(float)r; m = (double)(float)m; }
                keys[j] = e->hkeys[i]; re[j] = r; im[j] = m; ++j;
            }
            return SQE_OK;
        }
        uint64_t tot = 0;
        SQCK(export_to_staging(e, 1, &tot, mask, value));
        *n = tot;
        if (tot > cap) return fail(SQE_EINVAL, "export capacity too small (" + std::to_string(tot) + " needed)");
        if (tot) {
            if (!keys || !re || !im) return fail(SQE_EINVAL, "null array");
            HIPCK(hipMemcpyAsync(keys, e->xkeys, tot * 8, hipMemcpyDeviceToHost, e->st));
            HIPCK(hipMemcpyAsync(re, e->xre, tot * 8, hipMemcpyDeviceToHost, e->st));
            HIPCK(hipMemcpyAsync(im, e->xim, tot * 8, hipMemcpyDeviceToHost, e->st));
            HIPCK(hipStreamSynchronize(e->st));
        }
        collect_times(e);
        return SQE_OK;
    });
}

sqe_status sqe_export(sqe_engine* e, uint64_t* keys, double* re, double* im, uint64_t cap, uint64_t* n) {
    return guard([&]() -> sqe_status {
        return sqe_export_where(e, 0, 0, keys, re, im, cap, n);
    });
}

sqe_status sqe_sample(sqe_engine* e, uint64_t shots, uint64_t seed, uint64_t* keys) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (shots && !keys) return fail(SQE_EINVAL, "null output");
        if (shots > (1ull << 30)) return fail(SQE_EINVAL, "at most 2^30 shots per call");
        SQCK(full_sync(e));
        if (!shots) return SQE_OK;
        if (!e->have_state) return fail(SQE_EINVAL, "no state");
        if (!e->on_device) SQCK(materialize(e, std::vector<sqe_gate>()));   // a loaded state: place it first
        uint64_t n = 0;
        SQCK(export_to_staging(e, 1, &n));
        if (!n) return fail(SQE_EINVAL, "the state has no non-negligible entry");
        const uint64_t nt = sample_tiles(n);
        const size_t bytes = (size_t)(n + 2 * nt + 1 + shots) * 8;
        char* scratch = nullptr;
        SQCK(dalloc(e, (void**)&scratch, bytes));
        double* local = (double*)scratch;
        double* tsum = local + n;
        double* offs = tsum + nt;
        double* total = offs + nt;
        uint64_t* out = (uint64_t*)(total + 1);
        hipError_t r = launch_sample(e->st, e->xkeys, e->xre, e->xim, n, local, tsum, offs, total, seed, shots, out);
        if (r == hipSuccess) r = hipMemcpyAsync(keys, out, shots * 8, hipMemcpyDeviceToHost, e->st);
        if (r == hipSuccess) r = hipStreamSynchronize(e->st);
        dfree(e, scratch, bytes);
        if (r != hipSuccess) {
            (void)hipGetLastError();
            return fail(SQE_EHIP, std::string("sampling: ") + hipGetErrorString(r));
        }
        return SQE_OK;
    });
}

// Flat restart file: a 64-byte header (magic, version, qubits, precision,
// entry count, gates applied), then the keys, the real parts and the
// imaginary parts of the exported state, little-endian.
namespace {
struct CkptHeader {
    char magic[8];
    uint32_t version, nq, prec, pad;
    uint64_t count, gates;
    uint64_t reserved[3];
};
static_assert(sizeof(CkptHeader) == 64, "checkpoint header");
const char CKPT_MAGIC[8] = {'S', 'Q', 'E', 'S', 'T', 'A', 'T', 'E'};
}  // namespace

sqe_status sqe_checkpoint_save(sqe_engine* e, const char* path) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (!path) return fail(SQE_EINVAL, "null path");
        uint64_t n = 0;
        const sqe_status q = sqe_export(e, nullptr, nullptr, nullptr, 0, &n);   // size query
        if (q != SQE_OK && q != SQE_EINVAL) return q;
        std::vector<uint64_t> k(n);
        std::vector<double> r(n), m(n);
        uint64_t got = 0;
        SQCK(sqe_export(e, k.data(), r.data(), m.data(), n, &got));
        CkptHeader h{};
        std::memcpy(h.magic, CKPT_MAGIC, 8);
        h.version = 1;
        h.nq = e->cfg.nq;
        h.prec = e->cfg.prec;
        h.count = got;
        h.gates = e->gates;
        FILE* f = std::fopen(path, "wb");
        if (!f) return fail(SQE_EINVAL, std::string("cannot open ") + path + " for writing");
        bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
        ok = ok && (got == 0 || (std::fwrite(k.data(), 8, got, f) == got && std::fwrite(r.data(), 8, got, f) == got &&
                                 std::fwrite(m.data(), 8, got, f) == got));
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return fail(SQE_EINVAL, std::string("write failed: ") + path);
        return SQE_OK;
    });
}

sqe_status sqe_checkpoint_load(sqe_engine* e, const char* path) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (!path) return fail(SQE_EINVAL, "null path");
        FILE* f = std::fopen(path, "rb");
        if (!f) return fail(SQE_EINVAL, std::string("cannot open ") + path);
        CkptHeader h{};
        bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, CKPT_MAGIC, 8) == 0 && h.version == 1 &&
                  (h.prec == SQE_F64 || h.prec == SQE_F32);
        // the header's count must match the file's size before anything is sized by it
        long fsize = -1;
        if (ok && std::fseek(f, 0, SEEK_END) == 0) fsize = std::ftell(f);
        ok = ok && fsize >= (long)sizeof(h) && h.count <= ((uint64_t)fsize - sizeof(h)) / 24 &&
             (uint64_t)fsize == sizeof(h) + 24 * h.count && std::fseek(f, (long)sizeof(h), SEEK_SET) == 0;
        if (ok && h.nq != e->cfg.nq) {
            std::fclose(f);
            return fail(SQE_EINVAL, "checkpoint has " + std::to_string(h.nq) + " qubits, engine " +
                                        std::to_string(e->cfg.nq));
        }
        std::vector<uint64_t> k;
        std::vector<double> r, m;
        if (ok) {
            try {
                k.resize(h.count);
                r.resize(h.count);
                m.resize(h.count);
            } catch (const std::exception&) {
                std::fclose(f);
                return fail(SQE_ENOMEM, "host memory for " + std::to_string(h.count) + " checkpoint entries");
            }
            ok = h.count == 0 || (std::fread(k.data(), 8, h.count, f) == h.count &&
                                  std::fread(r.data(), 8, h.count, f) == h.count &&
                                  std::fread(m.data(), 8, h.count, f) == h.count);
        }
        std::fclose(f);
        if (!ok) return fail(SQE_EINVAL, std::string("not a readable state checkpoint: ") + path);
        SQCK(sqe_load(e, k.data(), r.data(), m.data(), h.count));
        e->gates = h.gates;
        return SQE_OK;
    });
}

sqe_status sqe_get_stats(sqe_engine* e, sqe_stats* out) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (!out) return fail(SQE_EINVAL, "null output");
        Stats h;
        SQCK(count_now(e, &h));
        std::memset(out, 0, sizeof(*out));
        out->gates = e->gates;
        out->updates = h.updates;
        out->nnz = h.nnz;
        uint64_t nb = 0;
        if (e->on_device) SQCK(read_nb(e, &nb));
        out->buckets = nb;
        out->capacity = e->s[e->cur].cap;
        out->bytes_device = e->bytes;
        out->bucket_bits = e->qm.mb;
        out->bucket_mask = e->qm.mask;
        out->sweeps = e->n_sweeps;
        out->sweep_ops = e->n_sweep_ops;
        out->pairs = e->n_pairs;
        out->remaps = e->n_remaps;
        out->reruns = (uint32_t)e->n_reruns;
        out->moved = h.moved;
        out->moved_pair = h.moved_pair;
        out->blocks = e->n_blocks;
        out->block_gates = e->n_block_gates;
        out->sweep_ms = e->t_ms[0];
        out->pair_ms = e->t_ms[1];
        out->other_ms = e->t_ms[2];
        return SQE_OK;
    });
}

sqe_status sqe_prepare(sqe_engine* e, const sqe_gate* g, uint64_t n) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (n && !g) return fail(SQE_EINVAL, "null gates");
        SQCK(full_sync(e));
        if (!e->have_state) return fail(SQE_EINVAL, "no state");
        std::vector<sqe_gate> gl(g, g + n);
        if (!e->on_device) return materialize(e, gl);
        const QMap best = choose_map((int)e->cfg.nq, e->cfg.bucket_bits, gl.data(), gl.size());
        if (best.mask != e->qm.mask) SQCK(remap(e, best));
        return SQE_OK;
    });
}

sqe_status sqe_snapshot_save(sqe_engine* e) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        SQCK(full_sync(e));
        if (!e->have_state) return fail(SQE_EINVAL, "no state");
        if (!e->on_device) SQCK(materialize(e, std::vector<sqe_gate>()));
        uint64_t nb = 0;
        SQCK(read_nb(e, &nb));
        const uint64_t nbp = round64(std::max<uint64_t>(nb, 1));
        if (e->snap.cap < nbp || e->snap.mb != e->qm.mb) {
            free_store(e, e->snap);
            SQCK(alloc_store(e, e->snap, nbp));
        }
        HIPCK(hipMemcpyAsync(e->snap.base, e->s[e->cur].base, nbp * 8, hipMemcpyDeviceToDevice, e->st));
        HIPCK(hipMemcpyAsync(e->snap.amp, e->s[e->cur].amp, e->amp_bytes(nbp), hipMemcpyDeviceToDevice, e->st));
        HIPCK(hipMemcpyAsync(&e->meta[3], &e->meta[e->cur], 8, hipMemcpyDeviceToDevice, e->st));
        HIPCK(hipStreamSynchronize(e->st));
        e->snap_qm = e->qm;
        e->snap_gates = e->gates;
        e->snap_valid = true;
        return SQE_OK;
    });
}

sqe_status sqe_snapshot_restore(sqe_engine* e) {
    return guard([&]() -> sqe_status {
        SQCK(enter(e));
        if (!e->snap_valid) return fail(SQE_EINVAL, "no snapshot");
        e->pending.clear();
        e->lz.pending = false;
        e->lz.gl.clear();
        HIPCK(hipStreamSynchronize(e->st));
        unsigned long long nb = 0;
        HIPCK(hipMemcpy(&nb, &e->meta[3], 8, hipMemcpyDeviceToHost));
        const uint64_t nbp = round64(std::max<uint64_t>(nb, 1));
        e->qm = e->snap_qm;
        SQCK(ensure_stores(e, nbp));
        HIPCK(hipMemcpyAsync(e->s[e->cur].base, e->snap.base, nbp * 8, hipMemcpyDeviceToDevice, e->st));
        HIPCK(hipMemcpyAsync(e->s[e->cur].amp, e->snap.amp, e->amp_bytes(nbp), hipMemcpyDeviceToDevice, e->st));
        HIPCK(hipMemcpyAsync(&e->meta[e->cur], &e->meta[3], 8, hipMemcpyDeviceToDevice, e->st));
        e->nb_upper = nb;
        e->tab_valid = false;
        e->renorm_pending = false;
        e->gates = e->snap_gates;
        e->on_device = true;
        e->have_state = true;
        return SQE_OK;
    });
}

const char* sqe_last_error(void) { return g_err.c_str(); }

}  // extern "C"
