// Batch engine: many independent circuits resident in one device pool and
// advanced together by persistent launches (no reference counterpart; the
// reference runs one circuit per process, emulator/src/main.cu:87-148).
//
// Each admitted circuit gets a region of the pool (bucket bases + tile-
// interleaved amplitudes, the layout of sqe_device.hpp), its own bucket-qubit
// set and its own pre-encoded sweep program.  sqe_batch_step advances every
// active circuit by up to N gates with one k_sweep_batch launch per bucket
// size; FP32 circuits stop at their renormalisation points, where a batched
// FP64 norm runs before the next step.  Admission is by bytes of the pool
// against the configured HBM budget, so a serving loop can admit arrivals as
// finished circuits are removed (continuous batching at step granularity).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <new>
#include <thread>
#include <stdexcept>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sqe.h"
#include "sqe_launch.hpp"
#include "sqe_plan.hpp"

using namespace sqe;

namespace {

thread_local std::string b_err;

sqe_status bfail(sqe_status s, const std::string& m) {
    b_err = m;
    return s;
}

// No C++ exception crosses the C ABI: host allocation failures (and anything
// else a standard container throws) become status codes.
template <class F>
sqe_status guard(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return bfail(SQE_ENOMEM, "host allocation failed");
    } catch (const std::exception& ex) {
        return bfail(SQE_EINVAL, std::string("host error: ") + ex.what());
    }
}

#define BHIP(x)                                                                          \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return bfail(SQE_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define BCK(x)                       \
    do {                             \
        sqe_status s_ = (x);         \
        if (s_ != SQE_OK) return s_; \
    } while (0)

inline uint64_t p2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// First-fit allocator over a range of units (tiles or amplitude rows).
struct RangeAlloc {
    std::map<uint64_t, uint64_t> free_;   // start -> length
    uint64_t cap = 0;
    void grow(uint64_t new_cap) {
        if (new_cap <= cap) return;
        // merge with a trailing free range
        if (!free_.empty()) {
            auto it = std::prev(free_.end());
            if (it->first + it->second == cap) {
                it->second += new_cap - cap;
                cap = new_cap;
                return;
            }
        }
        free_[cap] = new_cap - cap;
        cap = new_cap;
    }
    bool take(uint64_t n, uint64_t* at) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *at = it->first;
                const uint64_t rest = it->second - n, st = it->first + n;
                free_.erase(it);
                if (rest) free_[st] = rest;
                return true;
            }
        }
        return false;
    }
    void give(uint64_t at, uint64_t n) {
        auto it = free_.emplace(at, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used_extent() const {        // end of the last used unit
        if (free_.empty()) return cap;
        auto it = std::prev(free_.end());
        return it->first + it->second == cap ? it->first : cap;
    }
};

// Staging that asynchronous copies read or write, used as a ring: each
// reservation stays busy until the event recorded after the work that uses it
// has completed, so the host never waits for the device unless the ring has
// come round to a reservation still in use.  Page-locked host memory for
// uploads, device memory for export output.
struct Ring {
    char* p = nullptr;
    uint64_t cap = 0, head = 0;
    bool host = false;
    struct Span {
        uint64_t begin, end;
        hipEvent_t ev;
    };
    std::deque<Span> busy;                // in reservation order
    std::vector<hipEvent_t> idle;         // recorded events whose work is done
};

constexpr uint64_t RING_ALIGN = 256;
inline uint64_t ring_round(uint64_t n) { return (std::max<uint64_t>(n, 1) + RING_ALIGN - 1) & ~(RING_ALIGN - 1); }

// One exported circuit on its way to the caller's arrays.
struct XTicket {
    uint32_t rec = 0;                     // host record slot
    hipEvent_t ev = nullptr;              // after the copy to the caller
    sqe_batch_circuit info{};             // host-side fields at export time
    uint64_t cap = 0;
};

struct Circ {
    bool live = false;
    uint32_t nq = 0;
    QMap qm{};
    uint64_t tile0 = 0, tiles = 0;        // base pool tiles
    uint64_t row0 = 0, rows = 0;          // amp pool rows (64 amplitudes)
    uint64_t prog0 = 0, nops = 0;         // program pool ops
    uint64_t coef0 = 0, ncoef = 0;        // coefficient pool doubles
    uint64_t pos = 0;                     // next op
    std::vector<uint32_t> renorm_at;      // op indices of RENORM ops
    std::vector<uint32_t> gate_at;        // op index of each gate's record (increasing)
    uint64_t gates = 0;
    uint64_t bytes = 0;
    bool frames = false;                  // the program holds frame records
    uint64_t ck0 = 0, nck = 0;            // FP32 one-pass renormalisation: checkpoint partials (nck x tiles)
};

struct Pool {
    char* p = nullptr;
    uint64_t bytes = 0;
};

}  // namespace

// Helper threads for the host copies of large admissions (copying a
// request's keys and amplitudes into the page-locked upload ring is the bulk of
// an admission's host time).  run(parts, f) calls f(0) on the calling thread
// and f(1) .. f(parts - 1) on the helpers, and returns when all are done.
class CopyPool {
public:
    static constexpr int HELPERS = 3;
    CopyPool() {
        for (int i = 0; i < HELPERS; ++i) th_.emplace_back([this, i] { work(i + 1); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void run(int parts, const std::function<void(int)>& f) {
        parts = std::max(1, std::min(parts, HELPERS + 1));
        {
            std::lock_guard<std::mutex> g(mu_);
            job_ = &f;
            parts_ = parts;
            left_.store(parts - 1, std::memory_order_relaxed);
            ++gen_;
        }
        if (parts > 1) cv_.notify_all();
        f(0);
        while (left_.load(std::memory_order_acquire) > 0) std::this_thread::yield();
    }

private:
    void work(int part) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)>* f = nullptr;
            int parts = 0;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                f = job_;
                parts = parts_;
            }
            if (part < parts) {
                (*f)(part);
                left_.fetch_sub(1, std::memory_order_release);
            }
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_;
    const std::function<void(int)>* job_ = nullptr;
    int parts_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::atomic<int> left_{0};
};

struct sqe_batch {
    sqe_config cfg;
    bool f32 = false;
    double thr = 1e-14;
    size_t csize = 16;
    hipStream_t st = nullptr;
    std::vector<Circ> c;                   // indexed by slot (= id)
    std::vector<uint32_t> free_slots;
    RangeAlloc ta, ra, pa, ca, ka;         // tiles, rows, program ops, coefficients, checkpoint partials
    Pool base, amp, prog, coef, ck;
    uint64_t used = 0;                     // admitted bytes
    // per-slot device arrays
    uint64_t nslot = 0;
    unsigned long long *upd = nullptr, *nnz = nullptr, *meta = nullptr;
    double *norm2 = nullptr, *scale = nullptr;
    // load path scratch
    Slot* tab = nullptr;
    uint64_t tcap = 0;
    uint64_t *xkeys = nullptr;
    double *xre = nullptr, *xim = nullptr;
    uint64_t xcap = 0;
    unsigned int* occ = nullptr;
    uint64_t occ_cap = 0;
    Stats* stats = nullptr;
    unsigned long long* scan_status = nullptr;
    uint64_t scan_cap = 0;
    unsigned int* ticket = nullptr;
    uint32_t epoch = 0;
    unsigned long long* total = nullptr;
    // page-locked staging of admissions' and steps' uploads
    Ring up;
    // asynchronous exports: device staging of their output, the stream that
    // copies it to the caller, page-locked records of entry counts and counters
    Ring xr;
    hipStream_t xst = nullptr;
    std::map<int64_t, XTicket> xt;
    int64_t next_ticket = 1;
    unsigned long long *hrec = nullptr, *drec = nullptr;
    std::vector<uint32_t> rec_free;
    hipEvent_t xev = nullptr;                // export kernel done -> copy stream may start
    std::vector<hipEvent_t> evpool;          // idle ticket events
    unsigned long long* hsmall = nullptr;    // page-locked words for small readbacks
    Stats* cst = nullptr;                    // per-slot load error words (err field)
    double* partials = nullptr;
    uint64_t pcap = 0;
    BTask* dtask = nullptr;
    BItem* ditem = nullptr;
    uint64_t dcap_t = 0, dcap_i = 0;
    unsigned long long* counts = nullptr;     // per-item sweep counts
    uint64_t ccap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device-side snapshot of every admitted circuit (pools, positions, counters)
    char *snap_base = nullptr, *snap_amp = nullptr;
    unsigned long long* snap_cnt = nullptr;
    uint64_t snap_bcap = 0, snap_acap = 0, snap_ccap = 0, snap_bext = 0, snap_aext = 0;
    std::vector<uint64_t> snap_pos;
    uint64_t generation = 0, snap_generation = ~0ull;
    uint64_t layout = 0;                   // bumped whenever a device pointer a launch uses may change
    // captured gate sequence of a whole run (sqe_batch_run): valid for the
    // layout, tuning and start positions it was planned from
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t g_layout = ~0ull;
    uint32_t g_block = 0, g_gran = 0;
    bool g_one = false;
    std::vector<uint64_t> g_pos0, g_pos1;  // positions before / after the run
    uint64_t g_steps = 0, g_launches = 0;
    BTask* gtask = nullptr;
    BItem* gitem = nullptr;
    unsigned long long* gcounts = nullptr;
    double* gpartials = nullptr;
    uint64_t gcap_t = 0, gcap_i = 0, gcap_c = 0, gcap_p = 0;
    std::vector<hipStream_t> side;           // branches of the captured graph (default priority)
    std::vector<hipEvent_t> gev;             // fork / join events of the capture
    // tuning and counters
    uint32_t block = BLOCK, gran = 1;
    bool one_launch = true;                  // circuits of 0..SQE_MIXED_MAX_MB bucket qubits share a launch
    double sweep_ms = 0.0;
    double load_ms = 0.0;                   // index part of the load path, last autotune
    std::unique_ptr<CopyPool> copy_pool;     // helper threads for large admission copies (made on first use)
    uint64_t launches = 0, steps = 0, bytes_dev = 0;
};

namespace {

sqe_status dal(sqe_batch* b, void** p, size_t n) {
    *p = nullptr;
    if (!n) return SQE_OK;
    hipError_t r = hipMalloc(p, n);
    if (r != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return bfail(SQE_ENOMEM, "hipMalloc(" + std::to_string(n) + ") failed");
    }
    b->bytes_dev += n;
    return SQE_OK;
}
void dfr(sqe_batch* b, void* p, size_t n) {
    if (p) {
        (void)hipFree(p);
        b->bytes_dev -= n;
    }
}

constexpr uint32_t REC_WORDS = 8;          // per export record: entry count, counter snapshot, spare
constexpr uint32_t REC_SLOTS = 4096;       // exports in flight at once
constexpr uint64_t UP_RING_MIN = 64ull << 20;
constexpr uint64_t XR_RING_MIN = 256ull << 20;

sqe_status ring_event(Ring& r, hipEvent_t* ev) {
    if (!r.idle.empty()) {
        *ev = r.idle.back();
        r.idle.pop_back();
        return SQE_OK;
    }
    BHIP(hipEventCreateWithFlags(ev, hipEventDisableTiming));
    return SQE_OK;
}

// Retire the oldest reservation if its work has completed (or after waiting
// for it, with `wait`); *popped tells whether one was retired.
sqe_status ring_pop(Ring& r, bool wait, bool* popped) {
    *popped = false;
    if (r.busy.empty()) return SQE_OK;
    const Ring::Span sp = r.busy.front();
    if (wait) {
        BHIP(hipEventSynchronize(sp.ev));
    } else {
        const hipError_t q = hipEventQuery(sp.ev);
        if (q == hipErrorNotReady) return SQE_OK;
        BHIP(q);
    }
    r.idle.push_back(sp.ev);
    r.busy.pop_front();
    *popped = true;
    return SQE_OK;
}

sqe_status ring_release(sqe_batch* b, Ring& r) {
    for (bool pp = true; pp;) BCK(ring_pop(r, true, &pp));
    if (r.p) {
        if (r.host) (void)hipHostFree(r.p);
        else dfr(b, r.p, r.cap);
    }
    r.p = nullptr;
    r.cap = 0;
    r.head = 0;
    return SQE_OK;
}

// Reserve n bytes of the ring at *off (RING_ALIGN-aligned), growing it to at
// least min_cap bytes the first time.
sqe_status ring_take(sqe_batch* b, Ring& r, uint64_t n, uint64_t min_cap, uint64_t* off) {
    n = ring_round(n);
    if (n > r.cap) {
        BCK(ring_release(b, r));
        const uint64_t cap = std::max<uint64_t>(p2(n), min_cap);
        if (r.host) {
            if (hipHostMalloc((void**)&r.p, cap, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                r.p = nullptr;
                return bfail(SQE_ENOMEM, "hipHostMalloc(" + std::to_string(cap) + ") failed");
            }
        } else {
            BCK(dal(b, (void**)&r.p, cap));
        }
        r.cap = cap;
    }
    for (bool pp = true; pp;) BCK(ring_pop(r, false, &pp));
    for (;;) {
        const uint64_t s0 = r.head + n <= r.cap ? r.head : 0;
        bool clash = false;
        for (const Ring::Span& sp : r.busy) clash |= s0 < sp.end && sp.begin < s0 + n;
        if (!clash) {
            *off = s0;
            r.head = s0 + n;
            return SQE_OK;
        }
        bool pp = false;
        BCK(ring_pop(r, true, &pp));       // the oldest reservation is the next in ring order
    }
}

// The reservation of n bytes at off stays busy until the work enqueued on st
// so far has completed.
sqe_status ring_mark(Ring& r, uint64_t off, uint64_t n, hipStream_t st) {
    hipEvent_t ev = nullptr;
    BCK(ring_event(r, &ev));
    const hipError_t e = hipEventRecord(ev, st);
    if (e != hipSuccess) {
        r.idle.push_back(ev);
        BHIP(e);
    }
    r.busy.push_back(Ring::Span{off, off + ring_round(n), ev});
    return SQE_OK;
}

void ring_destroy(sqe_batch* b, Ring& r) {
    (void)ring_release(b, r);
    for (hipEvent_t e : r.idle) (void)hipEventDestroy(e);
    r.idle.clear();
}

// Grow a pool so that `need` bytes fit, keeping its used contents.
sqe_status pool_fit(sqe_batch* b, Pool& pl, uint64_t need, uint64_t keep) {
    if (need <= pl.bytes) return SQE_OK;
    const uint64_t nb = std::max<uint64_t>(p2(need), 1ull << 24);
    char* np = nullptr;
    BCK(dal(b, (void**)&np, nb));
    if (pl.p && keep) BHIP(hipMemcpyAsync(np, pl.p, std::min(keep, pl.bytes), hipMemcpyDeviceToDevice, b->st));
    BHIP(hipStreamSynchronize(b->st));
    dfr(b, pl.p, pl.bytes);
    pl.p = np;
    pl.bytes = nb;
    ++b->layout;
    return SQE_OK;
}

sqe_status slots_fit(sqe_batch* b, uint64_t need) {
    if (need <= b->nslot) return SQE_OK;
    const uint64_t n = std::max<uint64_t>(p2(need), 256);
    BHIP(hipStreamSynchronize(b->st));
    auto regrow = [&](void** ptr, size_t elem) -> sqe_status {
        void* np = nullptr;
        BCK(dal(b, &np, n * elem));
        BHIP(hipMemset(np, 0, n * elem));
        if (*ptr) BHIP(hipMemcpy(np, *ptr, b->nslot * elem, hipMemcpyDeviceToDevice));
        dfr(b, *ptr, b->nslot * elem);
        *ptr = np;
        return SQE_OK;
    };
    BCK(regrow((void**)&b->upd, 8));
    BCK(regrow((void**)&b->nnz, 8));
    BCK(regrow((void**)&b->meta, 8));
    BCK(regrow((void**)&b->norm2, 8));
    BCK(regrow((void**)&b->scale, 8));
    BCK(regrow((void**)&b->cst, sizeof(Stats)));
    b->nslot = n;
    ++b->layout;
    return SQE_OK;
}

template <class T>
sqe_status buf_fit(sqe_batch* b, T** p, uint64_t* cap, uint64_t need) {
    if (need <= *cap) return SQE_OK;
    BHIP(hipStreamSynchronize(b->st));
    dfr(b, *p, *cap * sizeof(T));
    *p = nullptr;
    const uint64_t n = std::max<uint64_t>(p2(need), 1024);
    BCK(dal(b, (void**)p, n * sizeof(T)));
    *cap = n;
    return SQE_OK;
}

sqe_status next_scan(sqe_batch* b, uint32_t ntiles, ScanCtl* out) {
    if (ntiles > b->scan_cap) {
        BHIP(hipStreamSynchronize(b->st));
        dfr(b, b->scan_status, b->scan_cap * 8);
        b->scan_status = nullptr;
        const uint64_t cap = std::max<uint64_t>(1024, p2(ntiles));
        BCK(dal(b, (void**)&b->scan_status, cap * 8));
        b->scan_cap = cap;
        BHIP(launch_fill_u64(b->st, (uint64_t*)b->scan_status, cap, 0));
        b->epoch = 0;
    }
    if (++b->epoch >= (1u << 22) - 1) {
        BHIP(launch_fill_u64(b->st, (uint64_t*)b->scan_status, b->scan_cap, 0));
        b->epoch = 1;
    }
    *out = ScanCtl{b->scan_status, b->ticket, &b->stats->err, b->epoch};
    return SQE_OK;
}

sqe_status check_err(sqe_batch* b) {
    unsigned long long err = 0;
    BHIP(hipMemcpyAsync(&err, &b->stats->err, 8, hipMemcpyDeviceToHost, b->st));
    BHIP(hipStreamSynchronize(b->st));
    if (err) {
        const unsigned long long z = 0;
        BHIP(hipMemcpy(&b->stats->err, &z, 8, hipMemcpyHostToDevice));
        return bfail((err & ERR_DUP) ? SQE_EINVAL : SQE_EHIP,
                     (err & ERR_DUP) ? "duplicate key in load" : "device-side error in batch");
    }
    return SQE_OK;
}

// A circuit's counters (updates, nonzeros, buckets) and its load error word,
// after the work enqueued so far.
sqe_status read_counters(sqe_batch* b, uint32_t id, unsigned long long v[4]) {
    const unsigned long long* src[4] = {b->upd + id, b->nnz + id, b->meta + id, &b->cst[id].err};
    for (int k = 0; k < 4; ++k) BHIP(hipMemcpyAsync(b->hsmall + k, src[k], 8, hipMemcpyDeviceToHost, b->st));
    BHIP(hipStreamSynchronize(b->st));
    for (int k = 0; k < 4; ++k) v[k] = b->hsmall[k];
    return SQE_OK;
}

sqe_status load_fail(unsigned long long err) {
    if (err & ERR_KEY) return bfail(SQE_EINVAL, "a key of the circuit's load has bits beyond nq");
    if (err & ERR_RENORM)
        return bfail(SQE_EDEVICE, "FP32 one-pass renormalisation: a checkpoint norm moved by more than 2^-11 "
                                  "(admit the circuit with renorm_mode 0 to renormalise step by step)");
    return bfail((err & ERR_DUP) ? SQE_EINVAL : SQE_EHIP,
                 (err & ERR_DUP) ? "duplicate key in the circuit's load" : "device-side error in the circuit's load");
}

// The device address of a caller's export array: page-locked host memory
// (sqe_host_alloc) or device memory; pageable memory is refused, since the
// copy runs as a kernel.
sqe_status device_view(void* p, void** out) {
    hipPointerAttribute_t at{};
    const hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) (void)hipGetLastError();
    if (e == hipSuccess && at.type == hipMemoryTypeDevice) {
        *out = p;
        return SQE_OK;
    }
    if (e == hipSuccess && at.type == hipMemoryTypeHost && hipHostGetDevicePointer(out, p, 0) == hipSuccess)
        return SQE_OK;
    (void)hipGetLastError();
    return bfail(SQE_EINVAL, "export arrays must be page-locked (sqe_host_alloc) or device memory");
}

// FP64 key / re / im staging arrays of at least `need` entries: one
// allocation of three arrays of xcap entries (exports use them at stride
// xcap; an admission packs its m entries at stride m, so one copy uploads them).
sqe_status stage_fit(sqe_batch* b, uint64_t need) {
    if (need <= b->xcap) return SQE_OK;
    BHIP(hipStreamSynchronize(b->st));
    dfr(b, b->xkeys, b->xcap * 24);
    b->xkeys = nullptr;
    b->xre = nullptr;
    b->xim = nullptr;
    b->xcap = 0;
    const uint64_t nc = std::max<uint64_t>(1024, p2(need));
    BCK(dal(b, (void**)&b->xkeys, nc * 24));
    b->xre = (double*)(b->xkeys + nc);
    b->xim = b->xre + nc;
    b->xcap = nc;
    return SQE_OK;
}

// Copy the n input entries to k / r / i on `parts` threads when every
// amplitude is nonzero (each part copies its range in place and checks it);
// otherwise, or for small inputs, compact on the calling thread.  Returns
// the count of entries kept and ORs every input key into *keybits.
uint64_t compact_entries(const uint64_t*, const double*, const double*, uint64_t, uint64_t*, double*, double*,
                         uint64_t*);
uint64_t copy_entries(sqe_batch* b, const uint64_t* keys, const double* re, const double* im, uint64_t n,
                      uint64_t* k, double* r, double* i, uint64_t* keybits) {
    constexpr uint64_t PAR_MIN = 1 << 15;          // entries: smaller copies stay on one thread
    if (n < PAR_MIN) return compact_entries(keys, re, im, n, k, r, i, keybits);
    if (!b->copy_pool) b->copy_pool.reset(new CopyPool());
    const int parts = CopyPool::HELPERS + 1;
    uint64_t bits[parts] = {}, zeros[parts] = {};
    b->copy_pool->run(parts, [&](int p) {
        const uint64_t j0 = n * (uint64_t)p / parts, j1 = n * (uint64_t)(p + 1) / parts;
        std::memcpy(k + j0, keys + j0, (j1 - j0) * 8);
        std::memcpy(r + j0, re + j0, (j1 - j0) * 8);
        std::memcpy(i + j0, im + j0, (j1 - j0) * 8);
        uint64_t acc = 0, z = 0;
        for (uint64_t j = j0; j < j1; ++j) {
            acc |= keys[j];
            z += (re[j] == 0.0 && im[j] == 0.0) ? 1 : 0;
        }
        bits[p] = acc;
        zeros[p] = z;
    });
    uint64_t acc = 0, z = 0;
    for (int p = 0; p < parts; ++p) { acc |= bits[p]; z += zeros[p]; }
    if (z) return compact_entries(keys, re, im, n, k, r, i, keybits);
    *keybits = acc;
    return n;
}

// Copy the entries with a nonzero amplitude (re != 0 or im != 0) of the n
// input entries, in input order, to k / r / i; returns their count and ORs
// every input key into *keybits.  Blocks of 64 entries whose amplitudes are
// all nonzero (the usual case) are copied with memcpy.
uint64_t compact_entries(const uint64_t* keys, const double* re, const double* im, uint64_t n, uint64_t* k,
                         double* r, double* i, uint64_t* keybits) {
    uint64_t q = 0, acc = 0;
    for (uint64_t j0 = 0; j0 < n; j0 += 64) {
        const uint64_t e = std::min<uint64_t>(n - j0, 64);
        uint64_t nz = 0;
        for (uint64_t j = 0; j < e; ++j) {
            acc |= keys[j0 + j];
            nz += (re[j0 + j] != 0.0 || im[j0 + j] != 0.0) ? 1 : 0;
        }
        if (nz == e) {
            std::memcpy(k + q, keys + j0, e * 8);
            std::memcpy(r + q, re + j0, e * 8);
            std::memcpy(i + q, im + j0, e * 8);
            q += e;
        } else {
            for (uint64_t j = j0; j < j0 + e; ++j) {
                if (re[j] == 0.0 && im[j] == 0.0) continue;
                k[q] = keys[j]; r[q] = re[j]; i[q] = im[j]; ++q;
            }
        }
    }
    *keybits = acc;
    return q;
}

char* amp_ptr(sqe_batch* b, const Circ& c) { return b->amp.p + c.row0 * 64 * b->csize; }
uint64_t* base_ptr(sqe_batch* b, const Circ& c) { return (uint64_t*)b->base.p + c.tile0 * 64; }

void release(sqe_batch* b, uint32_t id) {
    Circ& c = b->c[id];
    if (!c.live) return;
    b->ta.give(c.tile0, c.tiles);
    b->ra.give(c.row0, c.rows);
    if (c.nops) b->pa.give(c.prog0, c.nops);
    if (c.ncoef) b->ca.give(c.coef0, c.ncoef);
    if (c.nck) b->ka.give(c.ck0, c.nck * c.tiles);
    b->used -= c.bytes;
    c = Circ{};
    b->free_slots.push_back(id);
    ++b->generation;
    ++b->layout;
}

// One batched launch: a sweep (then the fold of its per-item counters) or an
// FP64 norm, over a task list (one task per circuit) and its work items.
struct Group {
    bool norm = false;
    bool fold = false;                      // norm group of one-pass circuits: fold their checkpoints
    int mb = 0;
    std::vector<BTask> tasks;
    std::vector<BItem> items;
    size_t task_off = 0, item_off = 0;     // offsets in the packed upload
};

constexpr uint32_t NORM_TILES = 4;         // tiles per norm work item

// Whether the run's sweeps go out as one launch for every circuit of
// 0..SQE_MIXED_MAX_MB bucket qubits (k_sweep_batch_mixed): with one_launch
// set, when circuits of two or more such bucket sizes have ops left.
bool use_mixed(const sqe_batch* b) {
    if (!b->one_launch) return false;
    uint32_t seen = 0;
    for (const Circ& c : b->c)
        if (c.live && c.pos < c.nops && c.qm.mb <= SQE_MIXED_MAX_MB) seen |= 1u << c.qm.mb;
    return __builtin_popcount(seen) >= 2;
}

// Plan one step from positions pos[] (advanced in place): every live circuit
// with ops left runs to its next RENORM op (at most max_ops ops when
// max_ops > 0), one sweep group per bucket size (with `mixed`, one group for
// bucket sizes 0..SQE_MIXED_MAX_MB) with the longest programs first; then one
// norm group per bucket size for the circuits that stopped before a RENORM
// op.  Returns false when no circuit has ops left.
bool plan_step(sqe_batch* b, uint64_t max_ops, bool mixed, std::vector<uint64_t>& pos, std::vector<Group>& out) {
    struct Seg { uint32_t id; uint64_t from, to; };
    std::vector<Seg> segs;
    for (uint32_t id = 0; id < b->c.size(); ++id) {
        const Circ& c = b->c[id];
        if (!c.live || pos[id] >= c.nops) continue;
        uint64_t end = max_ops ? std::min<uint64_t>(c.nops, pos[id] + max_ops) : c.nops;
        for (uint32_t r : c.renorm_at)
            if (r > pos[id] && r < end) { end = r; break; }
        segs.push_back(Seg{id, pos[id], end});
    }
    if (segs.empty()) return false;
    std::sort(segs.begin(), segs.end(), [](const Seg& x, const Seg& y) { return x.to - x.from > y.to - y.from; });
    std::vector<uint32_t> need_norm;
    for (int mb = 0; mb <= 5; ++mb) {
        if (mixed && mb < SQE_MIXED_MAX_MB) continue;
        const bool mix = mixed && mb == SQE_MIXED_MAX_MB;
        Group g;
        g.mb = mix ? MB_MIXED : mb;
        for (const Seg& s : segs) {
            const Circ& c = b->c[s.id];
            if (mix ? c.qm.mb > SQE_MIXED_MAX_MB : c.qm.mb != mb) continue;
            BTask t{};
            t.base = base_ptr(b, c);
            t.amp = amp_ptr(b, c);
            t.ops = (const MOp*)b->prog.p + c.prog0 + s.from;
            t.coef = (const double*)b->coef.p + c.coef0;
            t.scale = b->scale + s.id;
            t.upd = b->upd + s.id;
            t.nnz = b->nnz + s.id;
            t.norm2 = b->norm2 + s.id;
            t.nops = (uint32_t)(s.to - s.from);
            t.ntiles = (uint32_t)c.tiles;
            t.qm = c.qm;
            t.frames = c.frames ? 1 : 0;
            t.ckpt = c.nck ? (double*)b->ck.p + c.ck0 : nullptr;
            t.nck = (uint32_t)c.nck;
            t.item0 = (uint32_t)g.items.size();
            const uint32_t ti = (uint32_t)g.tasks.size();
            for (uint64_t tl = 0; tl < c.tiles; tl += b->gran) g.items.push_back(BItem{ti, (uint32_t)tl});
            t.nitem = (uint32_t)g.items.size() - t.item0;
            g.tasks.push_back(t);
            bool at_renorm = false;
            for (uint32_t r : c.renorm_at) at_renorm |= (r == s.to);
            if (at_renorm) need_norm.push_back(s.id);
        }
        if (!g.tasks.empty()) out.push_back(std::move(g));
    }
    for (const Seg& s : segs) pos[s.id] = s.to;
    for (int mb = 0; mb <= 5; ++mb) {
        Group g;
        g.norm = true;
        g.mb = mb;
        for (uint32_t id : need_norm) {
            const Circ& c = b->c[id];
            if (c.qm.mb != mb || c.nck) continue;
            BTask t{};
            t.amp = amp_ptr(b, c);
            t.norm2 = b->norm2 + id;
            t.scale = b->scale + id;
            t.ntiles = (uint32_t)c.tiles;
            t.item0 = (uint32_t)g.items.size();
            const uint32_t ti = (uint32_t)g.tasks.size();
            for (uint64_t tl = 0; tl < c.tiles; tl += NORM_TILES) g.items.push_back(BItem{ti, (uint32_t)tl});
            t.nitem = (uint32_t)g.items.size() - t.item0;
            g.tasks.push_back(t);
        }
        if (!g.tasks.empty()) out.push_back(std::move(g));
    }
    // one-pass circuits at their final renormalisation: fold the checkpoints
    // (one group per bucket size, so a captured run keeps each fold on the
    // branch of the sweep that wrote its partials)
    for (int mb = 0; mb <= 5; ++mb) {
        Group f;
        f.norm = true;
        f.fold = true;
        f.mb = mb;
        for (uint32_t id : need_norm) {
            const Circ& c = b->c[id];
            if (!c.nck || c.qm.mb != mb) continue;
            BTask t{};
            t.norm2 = b->norm2 + id;
            t.scale = b->scale + id;
            t.ntiles = (uint32_t)c.tiles;
            t.ckpt = (double*)b->ck.p + c.ck0;
            t.nck = (uint32_t)c.nck;
            t.err = &b->cst[id].err;
            f.tasks.push_back(t);
        }
        if (!f.tasks.empty()) out.push_back(std::move(f));
    }
    return true;
}

// Pack the groups' tasks and items back to back (recording each group's
// offsets) and return the largest sweep and norm item counts.
void pack(std::vector<Group>& gs, std::vector<BTask>& T, std::vector<BItem>& I, uint64_t* max_sweep,
          uint64_t* max_norm) {
    *max_sweep = 0;
    *max_norm = 0;
    for (Group& g : gs) {
        g.task_off = T.size();
        g.item_off = I.size();
        T.insert(T.end(), g.tasks.begin(), g.tasks.end());
        I.insert(I.end(), g.items.begin(), g.items.end());
        uint64_t& m = g.norm ? *max_norm : *max_sweep;
        m = std::max<uint64_t>(m, g.items.size());
    }
}

// Enqueue one group on stream st (no host synchronisation unless `timed`).
sqe_status launch_group(sqe_batch* b, hipStream_t st, const Group& g, const BTask* dt, const BItem* di,
                        unsigned long long* counts, double* partials, bool timed) {
    const BTask* t = dt + g.task_off;
    const BItem* it = di + g.item_off;
    if (!g.norm) {
        BatchArgs a{t, it, (uint32_t)g.items.size(), b->gran, b->thr, nullptr, counts};
        if (timed) BHIP(hipEventRecord(b->ev0, st));
        BHIP(launch_sweep_batch(g.mb, b->f32, st, b->block, a));
        if (timed) {
            BHIP(hipEventRecord(b->ev1, st));
            BHIP(hipEventSynchronize(b->ev1));
            float ms = 0.f;
            BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            b->sweep_ms += ms;
        }
        BHIP(launch_batch_fold(st, t, (uint32_t)g.tasks.size(), counts));
    } else if (g.fold) {
        BHIP(launch_ckpt_batch(st, t, (uint32_t)g.tasks.size(), 0x1p-11));
    } else {
        BatchArgs a{t, it, (uint32_t)g.items.size(), NORM_TILES, b->thr, partials, nullptr};
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(4096, ((uint32_t)g.items.size() + 3) / 4));
        BHIP(launch_norm_batch(g.mb, b->f32, st, grid, (uint32_t)g.tasks.size(), a));
    }
    return SQE_OK;
}

// One step launched directly: its groups are uploaded together (one host
// synchronisation per step, for the reused pinned staging), then enqueued.
sqe_status step_all(sqe_batch* b, uint64_t max_gates) {
    std::vector<uint64_t> pos(b->c.size());
    for (size_t i = 0; i < b->c.size(); ++i) pos[i] = b->c[i].pos;
    std::vector<Group> gs;
    if (!plan_step(b, max_gates, use_mixed(b), pos, gs)) return SQE_OK;
    std::vector<BTask> T;
    std::vector<BItem> I;
    uint64_t ms = 0, mn = 0;
    pack(gs, T, I, &ms, &mn);
    BCK(buf_fit(b, &b->dtask, &b->dcap_t, T.size()));
    BCK(buf_fit(b, &b->ditem, &b->dcap_i, I.size()));
    BCK(buf_fit(b, &b->counts, &b->ccap, std::max<uint64_t>(1, 2 * ms)));
    BCK(buf_fit(b, &b->partials, &b->pcap, std::max<uint64_t>(1, mn)));
    // task and item lists through the upload ring (the device lists are
    // rewritten in stream order, after the previous step has read them)
    const uint64_t tb = T.size() * sizeof(BTask), ib = I.size() * sizeof(BItem);
    uint64_t off = 0;
    BCK(ring_take(b, b->up, ring_round(tb) + ib, UP_RING_MIN, &off));
    char* hp = b->up.p + off;
    std::memcpy(hp, T.data(), tb);
    std::memcpy(hp + ring_round(tb), I.data(), ib);
    BHIP(hipMemcpyAsync(b->dtask, hp, tb, hipMemcpyHostToDevice, b->st));
    BHIP(hipMemcpyAsync(b->ditem, hp + ring_round(tb), ib, hipMemcpyHostToDevice, b->st));
    BCK(ring_mark(b->up, off, ring_round(tb) + ib, b->st));
    for (const Group& g : gs) {
        BCK(launch_group(b, b->st, g, b->dtask, b->ditem, b->counts, b->partials, b->cfg.timing != 0));
        if (!g.norm) ++b->launches;
    }
    for (size_t i = 0; i < b->c.size(); ++i) b->c[i].pos = pos[i];
    ++b->steps;
    return SQE_OK;
}

void drop_graph(sqe_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->g_layout = ~0ull;
}

// The rest of every circuit's program as one captured graph: all steps are
// planned on the host (the schedule depends only on the programs and the
// start positions), their task and item lists uploaded once, and the kernel
// sequence (sweep, fold and norm per step and bucket size) captured from the
// stream.  A run with the same layout, tuning and start positions (a run
// after a snapshot restore) replays the instantiated graph.
sqe_status run_graph(sqe_batch* b) {
    std::vector<uint64_t> pos0(b->c.size());
    for (size_t i = 0; i < b->c.size(); ++i) pos0[i] = b->c[i].pos;
    const bool hit = b->gexec && b->g_layout == b->layout && b->g_block == b->block && b->g_gran == b->gran &&
                     b->g_one == b->one_launch &&
                     b->g_pos0 == pos0;
    if (!hit) {
        drop_graph(b);
        std::vector<Group> gs;
        std::vector<uint64_t> pos = pos0;
        uint64_t steps = 0, launches = 0;
        const bool mixed = use_mixed(b);
        while (plan_step(b, 0, mixed, pos, gs)) ++steps;
        if (gs.empty()) return SQE_OK;
        for (const Group& g : gs) launches += g.norm ? 0 : 1;
        std::vector<BTask> T;
        std::vector<BItem> I;
        uint64_t ms = 0, mn = 0;
        pack(gs, T, I, &ms, &mn);
        BCK(buf_fit(b, &b->gtask, &b->gcap_t, T.size()));
        BCK(buf_fit(b, &b->gitem, &b->gcap_i, I.size()));
        // every group has its own counter / partial slots (groups of
        // different bucket sizes run concurrently)
        BCK(buf_fit(b, &b->gcounts, &b->gcap_c, std::max<uint64_t>(1, 2 * I.size())));
        BCK(buf_fit(b, &b->gpartials, &b->gcap_p, std::max<uint64_t>(1, I.size())));
        (void)ms;
        (void)mn;
        BHIP(hipStreamSynchronize(b->st));
        BHIP(hipMemcpy(b->gtask, T.data(), T.size() * sizeof(BTask), hipMemcpyHostToDevice));
        BHIP(hipMemcpy(b->gitem, I.data(), I.size() * sizeof(BItem), hipMemcpyHostToDevice));
        // Circuits of different bucket sizes never share work, so each bucket
        // size is its own chain of launches (steps in order) on its own
        // branch of the graph: a small chain overlaps the large one instead
        // of waiting for its tail.  A mixed sweep and the norms and folds of
        // its circuits form one chain.
        std::vector<std::vector<const Group*>> chains(6);
        for (const Group& g : gs) chains[g.mb == MB_MIXED || (mixed && g.mb <= SQE_MIXED_MAX_MB) ? 0 : g.mb].push_back(&g);
        int nch = 0;
        for (const auto& c : chains) nch += c.empty() ? 0 : 1;
        while ((int)b->side.size() < nch - 1) {
            hipStream_t s2 = nullptr;
            BHIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
            b->side.push_back(s2);
        }
        while ((int)b->gev.size() < nch) {
            hipEvent_t ev = nullptr;
            BHIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            b->gev.push_back(ev);
        }
        BHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        sqe_status st = SQE_OK;
        hipError_t he = hipEventRecord(b->gev[0], b->st);          // fork
        int ci = 0;
        for (const auto& c : chains) {
            if (c.empty()) continue;
            hipStream_t s2 = ci == 0 ? b->st : b->side[ci - 1];
            if (ci > 0 && he == hipSuccess) he = hipStreamWaitEvent(s2, b->gev[0], 0);
            for (const Group* g : c) {
                if (st != SQE_OK || he != hipSuccess) break;
                st = launch_group(b, s2, *g, b->gtask, b->gitem, b->gcounts + 2 * g->item_off,
                                  b->gpartials + g->item_off, false);
            }
            if (ci > 0 && he == hipSuccess) he = hipEventRecord(b->gev[ci], s2);     // join
            ++ci;
        }
        for (int k = 1; k < ci && he == hipSuccess; ++k) he = hipStreamWaitEvent(b->st, b->gev[k], 0);
        if (st == SQE_OK && he != hipSuccess) st = bfail(SQE_EHIP, std::string("graph fork/join: ") + hipGetErrorString(he));
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(b->st, &graph);
        if (st != SQE_OK || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            return st != SQE_OK ? st : bfail(SQE_EHIP, std::string("graph capture: ") + hipGetErrorString(ec));
        }
        b->graph = graph;
        if (hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            drop_graph(b);
            return bfail(SQE_EHIP, "graph instantiate");
        }
        b->g_layout = b->layout;
        b->g_block = b->block;
        b->g_gran = b->gran;
        b->g_one = b->one_launch;
        b->g_pos0 = pos0;
        b->g_pos1 = pos;
        b->g_steps = steps;
        b->g_launches = launches;
    }
    const bool timed = b->cfg.timing != 0;
    if (timed) BHIP(hipEventRecord(b->ev0, b->st));
    BHIP(hipGraphLaunch(b->gexec, b->st));
    if (timed) {
        BHIP(hipEventRecord(b->ev1, b->st));
        BHIP(hipEventSynchronize(b->ev1));
        float ms = 0.f;
        BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->sweep_ms += ms;
    }
    for (size_t i = 0; i < b->c.size(); ++i) b->c[i].pos = b->g_pos1[i];
    b->steps += b->g_steps;
    b->launches += b->g_launches;
    return SQE_OK;
}

sqe_status snap_save(sqe_batch* b) {
    BHIP(hipStreamSynchronize(b->st));
    const uint64_t bext = b->ta.used_extent() * 64 * 8, aext = b->ra.used_extent() * 64 * b->csize;
    auto fit = [&](char** p, uint64_t* cap, uint64_t need) -> sqe_status {
        if (need <= *cap) return SQE_OK;
        dfr(b, *p, *cap);
        *p = nullptr;
        *cap = 0;
        BCK(dal(b, (void**)p, need));
        *cap = need;
        return SQE_OK;
    };
    BCK(fit(&b->snap_base, &b->snap_bcap, std::max<uint64_t>(bext, 8)));
    BCK(fit(&b->snap_amp, &b->snap_acap, std::max<uint64_t>(aext, 8)));
    BCK(fit((char**)&b->snap_cnt, &b->snap_ccap, 2 * b->nslot * 8));
    if (bext) BHIP(hipMemcpyAsync(b->snap_base, b->base.p, bext, hipMemcpyDeviceToDevice, b->st));
    if (aext) BHIP(hipMemcpyAsync(b->snap_amp, b->amp.p, aext, hipMemcpyDeviceToDevice, b->st));
    BHIP(hipMemcpyAsync(b->snap_cnt, b->upd, b->nslot * 8, hipMemcpyDeviceToDevice, b->st));
    BHIP(hipMemcpyAsync(b->snap_cnt + b->nslot, b->nnz, b->nslot * 8, hipMemcpyDeviceToDevice, b->st));
    BHIP(hipStreamSynchronize(b->st));
    b->snap_bext = bext;
    b->snap_aext = aext;
    b->snap_pos.clear();
    for (const Circ& c : b->c) b->snap_pos.push_back(c.pos);
    b->snap_generation = b->generation;
    return SQE_OK;
}

sqe_status snap_restore(sqe_batch* b) {
    if (b->snap_generation != b->generation)
        return bfail(SQE_EINVAL, "no snapshot of the current set of circuits (add / remove since the last save)");
    if (b->snap_bext) BHIP(hipMemcpyAsync(b->base.p, b->snap_base, b->snap_bext, hipMemcpyDeviceToDevice, b->st));
    if (b->snap_aext) BHIP(hipMemcpyAsync(b->amp.p, b->snap_amp, b->snap_aext, hipMemcpyDeviceToDevice, b->st));
    BHIP(hipMemcpyAsync(b->upd, b->snap_cnt, b->nslot * 8, hipMemcpyDeviceToDevice, b->st));
    BHIP(hipMemcpyAsync(b->nnz, b->snap_cnt + b->nslot, b->nslot * 8, hipMemcpyDeviceToDevice, b->st));
    for (size_t j = 0; j < b->c.size() && j < b->snap_pos.size(); ++j) b->c[j].pos = b->snap_pos[j];
    return SQE_OK;
}

}  // namespace

extern "C" {

sqe_status sqe_batch_create(const sqe_config* cfg, sqe_batch** out) {
    return guard([&]() -> sqe_status {
        if (!cfg || !out) return bfail(SQE_EINVAL, "null argument");
        *out = nullptr;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
            (void)hipGetLastError();
            return bfail(SQE_EDEVICE, "no HIP device");
        }
        if (cfg->prec != SQE_F64 && cfg->prec != SQE_F32) return bfail(SQE_EINVAL, "bad precision");
        if (cfg->device < 0 || cfg->device >= ndev) return bfail(SQE_EINVAL, "bad device ordinal");
        BHIP(hipSetDevice(cfg->device));
        sqe_batch* b = new (std::nothrow) sqe_batch();
        if (!b) return bfail(SQE_ENOMEM, "host allocation");
        b->cfg = *cfg;
        b->f32 = cfg->prec == SQE_F32;
        b->csize = b->f32 ? 8 : 16;
        b->thr = cfg->prune_thr > 0 ? cfg->prune_thr : (b->f32 ? 1e-7 : 1e-14);
        if (!b->cfg.hbm_budget_bytes) {
            size_t fr = 0, tot = 0;
            (void)hipMemGetInfo(&fr, &tot);
            const double keep = b->cfg.mem_reserve_ppm ? std::min(1.0, b->cfg.mem_reserve_ppm * 1e-6) : 0.1;
            b->cfg.hbm_budget_bytes = (uint64_t)((1.0 - keep) * (double)fr);
        }
        sqe_status r = SQE_OK;
        if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) r = bfail(SQE_EHIP, "stream");
        if (r == SQE_OK) r = dal(b, (void**)&b->stats, sizeof(Stats));
        if (r == SQE_OK) r = dal(b, (void**)&b->ticket, 64);
        if (r == SQE_OK) r = dal(b, (void**)&b->total, 64);
        if (r == SQE_OK && hipMemset(b->stats, 0, sizeof(Stats)) != hipSuccess) r = bfail(SQE_EHIP, "memset");
        if (r == SQE_OK && hipMemset(b->ticket, 0, 64) != hipSuccess) r = bfail(SQE_EHIP, "memset");
        if (r == SQE_OK && (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
                            hipEventCreateWithFlags(&b->xev, hipEventDisableTiming) != hipSuccess))
            r = bfail(SQE_EHIP, "event");
        if (r == SQE_OK && hipStreamCreateWithFlags(&b->xst, hipStreamNonBlocking) != hipSuccess)
            r = bfail(SQE_EHIP, "stream");
        if (r == SQE_OK && hipHostMalloc((void**)&b->hsmall, 64, hipHostMallocDefault) != hipSuccess)
            r = bfail(SQE_ENOMEM, "hipHostMalloc");
        b->up.host = true;
        if (r == SQE_OK) r = slots_fit(b, 256);
        if (r != SQE_OK) {
            sqe_batch_destroy(b);
            return r;
        }
        *out = b;
        return SQE_OK;
    });
}

sqe_status sqe_batch_destroy(sqe_batch* b) {
    return guard([&]() -> sqe_status {
        if (!b) return SQE_OK;
        (void)hipSetDevice(b->cfg.device);
        if (b->st) (void)hipStreamSynchronize(b->st);
        if (b->xst) (void)hipStreamSynchronize(b->xst);
        ring_destroy(b, b->up);
        ring_destroy(b, b->xr);
        for (auto& kv : b->xt) (void)hipEventDestroy(kv.second.ev);
        for (hipEvent_t ev : b->evpool) (void)hipEventDestroy(ev);
        if (b->xev) (void)hipEventDestroy(b->xev);
        if (b->hrec) (void)hipHostFree(b->hrec);
        if (b->hsmall) (void)hipHostFree(b->hsmall);
        if (b->xst) (void)hipStreamDestroy(b->xst);
        for (void* p : {(void*)b->cst, (void*)b->base.p, (void*)b->amp.p, (void*)b->prog.p, (void*)b->coef.p, (void*)b->ck.p, (void*)b->upd,
                        (void*)b->nnz, (void*)b->meta, (void*)b->norm2, (void*)b->scale, (void*)b->tab,
                        (void*)b->xkeys, (void*)b->occ, (void*)b->stats,
                        (void*)b->scan_status, (void*)b->ticket, (void*)b->total, (void*)b->dtask, (void*)b->ditem, (void*)b->partials,
                        (void*)b->counts, (void*)b->snap_base, (void*)b->snap_amp, (void*)b->snap_cnt, (void*)b->gtask,
                        (void*)b->gitem, (void*)b->gcounts, (void*)b->gpartials})
            if (p) (void)hipFree(p);
        drop_graph(b);
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        if (b->ev1) (void)hipEventDestroy(b->ev1);
        for (hipStream_t s2 : b->side) (void)hipStreamDestroy(s2);
        for (hipEvent_t ev : b->gev) (void)hipEventDestroy(ev);
        if (b->st) (void)hipStreamDestroy(b->st);
        delete b;
        return SQE_OK;
    });
}

// Admit a circuit (sqe_batch_add, sqe_batch_add_pinned).  Copied: the
// entries are checked, filtered and packed into the page-locked upload ring
// on the host.  Pinned: the caller's arrays are uploaded as they are (zero
// amplitudes load as absent entries of live buckets, keys beyond nq are
// reported by the load kernels in the circuit's error word).
sqe_status admit(sqe_batch* b, uint32_t nq, const uint64_t* keys, const double* re, const double* im, uint64_t n,
                 const sqe_gate* gates, uint64_t ng, int64_t* id_out, bool pinned) {
    return guard([&]() -> sqe_status {
        if (!b || !id_out) return bfail(SQE_EINVAL, "null argument");
        if (nq < 1 || nq > 63) return bfail(SQE_EINVAL, "nq must be in 1..63");
        if ((n && (!keys || !re || !im)) || (ng && !gates)) return bfail(SQE_EINVAL, "null array");
        BHIP(hipSetDevice(b->cfg.device));
        // program positions and per-launch op counts are 32-bit
        if (ng > (1ull << 30)) return bfail(SQE_EINVAL, "at most 2^30 gates per circuit");
        const uint64_t lim = (1ull << nq) - 1;
        for (uint64_t j = 0; j < ng; ++j) {
            const sqe_gate& x = gates[j];
            if (x.kind > SQE_GATE_PERM || x.target >= nq || (x.ctrl_mask & ~lim) || ((x.ctrl_mask >> x.target) & 1ull) ||
                (x.ctrl_val & ~x.ctrl_mask))
                return bfail(SQE_EINVAL, "gate " + std::to_string(j) + " is invalid");
        }
        const QMap qm = choose_map((int)nq, b->cfg.bucket_bits, gates, ng);
        for (uint64_t j = 0; j < ng; ++j)
            if (!gate_local(gates[j], qm))
                return bfail(SQE_ENOTLOCAL, "gate " + std::to_string(j) +
                                                " crosses buckets: run this circuit with the single-circuit engine");
        // program: gates, with a RENORM op after every renorm_every-th gate
        // (Pauli gates on bucket qubits as frame records when the keys leave
        // room for the frame).  FP32 with renorm_mode 1 renormalises in one
        // pass: an FP64 norm checkpoint after every renorm_every-th gate and
        // one RENORM op at the end, which applies the last checkpoint's scale
        // (a global scalar, so applying it late changes nothing but the prune
        // decisions' missing factor, checked within 2^-11 of 1).
        const bool onepass = b->f32 && b->cfg.renorm_every && b->cfg.renorm_mode == 1;
        std::vector<DOp> dprog;
        std::vector<size_t> gate_idx, ren_idx;
        dprog.reserve(ng + 8);
        uint32_t nck = 0;
        for (uint64_t j = 0; j < ng; ++j) {
            gate_idx.push_back(dprog.size());
            dprog.push_back(plan_op(gates[j], qm));
            if (b->cfg.renorm_every && (j + 1) % b->cfg.renorm_every == 0) {
                if (onepass) {
                    dprog.push_back(norm_point_op(nck++));
                } else {
                    ren_idx.push_back(dprog.size());
                    dprog.push_back(renorm_op());
                }
            }
        }
        if (nck) {
            ren_idx.push_back(dprog.size());
            dprog.push_back(renorm_op());
        }
        const bool frames = !b->cfg.frame_off && nq <= (uint32_t)FRAME_Z_BIT0 && qm.mb > 0;
        Encoded enc;
        encode_program(dprog, {}, b->f32, frames ? &qm : nullptr, enc);
        const std::vector<MOp>& prog = enc.ops;
        const std::vector<double>& coef = enc.coef;
        std::vector<uint32_t> renorm_at, gate_at;
        for (size_t r : ren_idx) renorm_at.push_back((uint32_t)enc.start[r]);
        for (size_t g : gate_idx) gate_at.push_back((uint32_t)(enc.start[g + 1] - 1));   // the gate's own record
        const bool framed = enc.framed;
        // the uploads go through the page-locked upload ring: program,
        // coefficients, then (copied admissions) the keys, real and imaginary
        // parts of the entries with a nonzero amplitude, packed (one pass over
        // the input: key range check, zero filter and copy); the device
        // staging they land in is rewritten in stream order
        const uint64_t pc = prog.size() * sizeof(MOp) + coef.size() * 8;
        uint64_t up_off = 0;
        BCK(ring_take(b, b->up, pc + (pinned ? 0 : 24 * n), UP_RING_MIN, &up_off));
        char* const hp0 = b->up.p + up_off;
        uint64_t* hk = (uint64_t*)(hp0 + pc);
        uint64_t m = n;
        auto unreserve = [&]() { b->up.head = up_off; };
        if (!pinned) {
            uint64_t keybits = 0;
            m = copy_entries(b, keys, re, im, n, hk, (double*)(hk + n), (double*)(hk + 2 * n), &keybits);
            if (m < n) {                             // pack at stride m
                std::memmove(hk + m, hk + n, m * 8);
                std::memmove(hk + 2 * m, hk + 2 * n, m * 8);
            }
            if (keybits & ~lim) {
                unreserve();
                return bfail(SQE_EINVAL, "key has bits beyond nq");
            }
        }
        // region sizes and admission
        const uint64_t S = 1ull << qm.mb;
        const uint64_t tiles = std::max<uint64_t>(1, (m + 63) / 64);
        const uint64_t rows = tiles * S;
        const uint64_t bytes = tiles * 64 * 8 + rows * 64 * b->csize + prog.size() * sizeof(MOp) + coef.size() * 8 +
                               (uint64_t)nck * tiles * 8;
        if (b->used + bytes > b->cfg.hbm_budget_bytes) {
            unreserve();
            return bfail(SQE_ETABLE_FULL, "admission: circuit needs " + std::to_string(bytes) + " bytes, " +
                                              std::to_string(b->cfg.hbm_budget_bytes - b->used) + " left in the budget");
        }
        uint32_t id;
        if (!b->free_slots.empty()) { id = b->free_slots.back(); b->free_slots.pop_back(); }
        else { id = (uint32_t)b->c.size(); b->c.emplace_back(); }
        BCK(slots_fit(b, id + 1));
        Circ c;
        c.nq = nq;
        c.qm = qm;
        c.tiles = tiles;
        c.rows = rows;
        c.nops = prog.size();
        c.ncoef = coef.size();
        c.renorm_at = renorm_at;
        c.gate_at = gate_at;
        c.gates = ng;
        c.bytes = bytes;
        c.frames = framed;
        c.nck = nck;
        auto place = [&](RangeAlloc& al, Pool& pl, uint64_t units, uint64_t unit_bytes, uint64_t* at) -> sqe_status {
            if (!units) { *at = 0; return SQE_OK; }
            if (!al.take(units, at)) {
                const uint64_t want = std::max<uint64_t>(al.cap * 2, al.cap + units);
                BCK(pool_fit(b, pl, want * unit_bytes, al.used_extent() * unit_bytes));
                al.grow(pl.bytes / unit_bytes);
                if (!al.take(units, at)) return bfail(SQE_ENOMEM, "pool allocation");
            }
            return SQE_OK;
        };
        int placed = 0;
        sqe_status st = place(b->ta, b->base, tiles, 64 * 8, &c.tile0);
        if (st == SQE_OK && ++placed) st = place(b->ra, b->amp, rows, 64 * b->csize, &c.row0);
        if (st == SQE_OK && ++placed) st = place(b->pa, b->prog, c.nops, sizeof(MOp), &c.prog0);
        if (st == SQE_OK && ++placed) st = place(b->ca, b->coef, c.ncoef, 8, &c.coef0);
        if (st == SQE_OK && ++placed) st = place(b->ka, b->ck, c.nck * c.tiles, 8, &c.ck0);
        if (st != SQE_OK) {
            if (placed > 0) b->ta.give(c.tile0, c.tiles);
            if (placed > 1) b->ra.give(c.row0, c.rows);
            if (placed > 2 && c.nops) b->pa.give(c.prog0, c.nops);
            if (placed > 3 && c.ncoef) b->ca.give(c.coef0, c.ncoef);
            b->free_slots.push_back(id);
            unreserve();
            return st;
        }
        c.live = true;
        b->c[id] = c;
        b->used += bytes;
        ++b->generation;
        ++b->layout;
        Circ& cc = b->c[id];
        BCK(stage_fit(b, std::max<uint64_t>(m, 1)));
        const uint64_t up = pc + (pinned ? 0 : 24 * m);
        b->up.head = up_off + ring_round(up);
        char* hp = hp0;
        if (cc.nops) {
            std::memcpy(hp, prog.data(), cc.nops * sizeof(MOp));
            BHIP(hipMemcpyAsync(b->prog.p + cc.prog0 * sizeof(MOp), hp, cc.nops * sizeof(MOp), hipMemcpyHostToDevice,
                                b->st));
            hp += cc.nops * sizeof(MOp);
        }
        if (cc.ncoef) {
            std::memcpy(hp, coef.data(), cc.ncoef * 8);
            BHIP(hipMemcpyAsync(b->coef.p + cc.coef0 * 8, hp, cc.ncoef * 8, hipMemcpyHostToDevice, b->st));
        }
        if (m && !pinned) BHIP(hipMemcpyAsync(b->xkeys, hk, m * 24, hipMemcpyHostToDevice, b->st));
        if (m && pinned) {
            // straight from the caller's arrays to the staging (packed at
            // stride m): one pitched copy when they are equally spaced, as the
            // three arrays of one sqe_host_alloc buffer are
            const char* k0 = (const char*)keys;
            const ptrdiff_t pitch = (const char*)re - k0;
            if (pitch > 0 && (const char*)im - (const char*)re == pitch && (uint64_t)pitch >= m * 8) {
                BHIP(hipMemcpy2DAsync(b->xkeys, m * 8, keys, (size_t)pitch, m * 8, 3, hipMemcpyHostToDevice, b->st));
            } else {
                BHIP(hipMemcpyAsync(b->xkeys, keys, m * 8, hipMemcpyHostToDevice, b->st));
                BHIP(hipMemcpyAsync(b->xkeys + m, re, m * 8, hipMemcpyHostToDevice, b->st));
                BHIP(hipMemcpyAsync(b->xkeys + 2 * m, im, m * 8, hipMemcpyHostToDevice, b->st));
            }
        }
        if (up) BCK(ring_mark(b->up, up_off, up, b->st));
        // load the state into the region (insert bases, number buckets, scatter)
        const double ml = b->cfg.max_load > 0 && b->cfg.max_load < 0.95 ? b->cfg.max_load : 0.5;
        const uint64_t tneed = std::max<uint64_t>(1024, p2((uint64_t)std::ceil((double)m / ml) + 8));
        if (tneed > b->tcap) {
            BHIP(hipStreamSynchronize(b->st));
            dfr(b, b->tab, b->tcap * sizeof(Slot));
            b->tab = nullptr;
            BCK(dal(b, (void**)&b->tab, tneed * sizeof(Slot)));
            b->tcap = tneed;
        }
        BCK(buf_fit(b, &b->occ, &b->occ_cap, tiles * 64));
        // one launch clears the index (its first tneed slots: the table is
        // kept at its largest size, so the cost follows the circuit's size),
        // the region, the occupancy bits and the circuit's counters
        FillSet fs{};
        fs.r[0] = FillRange{(uint64_t*)b->tab, 2 * tneed, EMPTY_KEY, NONE};
        fs.r[1] = FillRange{base_ptr(b, cc), tiles * 64, EMPTY_KEY, EMPTY_KEY};
        fs.r[2] = FillRange{(uint64_t*)amp_ptr(b, cc), rows * 64 * b->csize / 8, 0, 0};
        fs.r[3] = FillRange{(uint64_t*)b->occ, (tiles * 64 + 1) / 2, 0, 0};
        fs.r[4] = FillRange{(uint64_t*)(b->upd + id), 1, 0, 0};
        fs.r[5] = FillRange{(uint64_t*)(b->nnz + id), 1, 0, 0};
        fs.r[6] = FillRange{(uint64_t*)(b->cst + id), sizeof(Stats) / 8, 0, 0};
        fs.k = 7;
        BHIP(launch_fill_set(b->st, fs));
        // a load error (a duplicate key) lands in the circuit's own error word
        // and is reported by sqe_batch_info / sqe_batch_export for it, so
        // admission never waits for the device
        LoadArgs la{};
        la.keys = b->xkeys; la.re = (const double*)(b->xkeys + m); la.im = la.re + m; la.n = m; la.qm = qm;
        la.keymask = pinned ? ~lim : 0ull;
        la.tab = b->tab; la.tmask = tneed - 1; la.amp = amp_ptr(b, cc); la.occ = b->occ; la.stats = b->cst + id;
        EnumArgs ea{};
        ea.tab = b->tab; ea.tcap = tneed; ea.out_base = base_ptr(b, cc); ea.out_nb = b->meta + id;
        ea.out_cap = tiles * 64; ea.stats = b->cst + id;
        const uint32_t eg = (uint32_t)((tneed + BLOCK - 1) / BLOCK);
        BCK(next_scan(b, eg, &ea.scan));
        BHIP(launch_load(b->f32, b->st, la, ea, eg));
        *id_out = id;
        return SQE_OK;
    });
}

sqe_status sqe_batch_add(sqe_batch* b, uint32_t nq, const uint64_t* keys, const double* re, const double* im,
                         uint64_t n, const sqe_gate* gates, uint64_t ng, int64_t* id_out) {
    return admit(b, nq, keys, re, im, n, gates, ng, id_out, false);
}

sqe_status sqe_batch_add_pinned(sqe_batch* b, uint32_t nq, const uint64_t* keys, const double* re, const double* im,
                                uint64_t n, const sqe_gate* gates, uint64_t ng, int64_t* id_out) {
    return admit(b, nq, keys, re, im, n, gates, ng, id_out, true);
}

sqe_status sqe_batch_step(sqe_batch* b, uint32_t max_gates) {
    return guard([&]() -> sqe_status {
        if (!b) return bfail(SQE_EINVAL, "null batch");
        BHIP(hipSetDevice(b->cfg.device));
        return step_all(b, max_gates);
    });
}

sqe_status sqe_batch_run(sqe_batch* b) {
    return guard([&]() -> sqe_status {
        if (!b) return bfail(SQE_EINVAL, "null batch");
        BHIP(hipSetDevice(b->cfg.device));
        if (!b->cfg.batch_eager) return run_graph(b);
        for (;;) {
            bool more = false;
            for (const Circ& c : b->c) more |= c.live && c.pos < c.nops;
            if (!more) break;
            BCK(step_all(b, 0));
        }
        return SQE_OK;
    });
}

sqe_status sqe_batch_sync(sqe_batch* b) {
    return guard([&]() -> sqe_status {
        if (!b) return bfail(SQE_EINVAL, "null batch");
        BHIP(hipSetDevice(b->cfg.device));
        BHIP(hipStreamSynchronize(b->st));
        BHIP(hipStreamSynchronize(b->xst));
        return check_err(b);
    });
}

sqe_status sqe_batch_info(sqe_batch* b, int64_t id, sqe_batch_circuit* out) {
    return guard([&]() -> sqe_status {
        if (!b || !out || id < 0 || (uint64_t)id >= b->c.size() || !b->c[id].live)
            return bfail(SQE_EINVAL, "unknown circuit");
        BHIP(hipSetDevice(b->cfg.device));
        const Circ& c = b->c[id];
        unsigned long long v[4] = {0, 0, 0, 0};
        BCK(read_counters(b, (uint32_t)id, v));
        if (v[3]) return load_fail(v[3]);
        out->done = c.pos >= c.nops ? 1 : 0;
        out->gates_total = c.gates;
        out->gates_done = (uint64_t)(std::lower_bound(c.gate_at.begin(), c.gate_at.end(), (uint32_t)c.pos) -
                                     c.gate_at.begin());
        out->updates = v[0];
        out->nnz = v[1];
        out->buckets = v[2];
        out->bucket_bits = c.qm.mb;
        out->bytes = c.bytes;
        return SQE_OK;
    });
}

sqe_status sqe_batch_export(sqe_batch* b, int64_t id, uint64_t* keys, double* re, double* im, uint64_t cap,
                            uint64_t* n) {
    return guard([&]() -> sqe_status {
        if (!b || !n || id < 0 || (uint64_t)id >= b->c.size() || !b->c[id].live)
            return bfail(SQE_EINVAL, "unknown circuit");
        BHIP(hipSetDevice(b->cfg.device));
        const Circ& c = b->c[id];
        // staging for every amplitude slot of the circuit's region, so the bucket
        // count need not be read back first; the total and the error word come
        // back together
        const uint64_t nbk = c.tiles * 64;
        BCK(stage_fit(b, std::max<uint64_t>(nbk << c.qm.mb, 1)));
        ExportArgs xa{};
        xa.base = base_ptr(b, c); xa.amp = amp_ptr(b, c); xa.nb = b->meta + id; xa.qm = c.qm; xa.thr = b->thr;
        xa.filter = 1; xa.keys = b->xkeys; xa.re = b->xre; xa.im = b->xim; xa.cap = b->xcap; xa.total = b->total;
        const uint32_t g = std::max<uint32_t>(1, (uint32_t)((nbk + BLOCK - 1) / BLOCK));
        BCK(next_scan(b, g, &xa.scan));
        BHIP(launch_export(c.qm.mb, b->f32, b->st, g, xa));
        unsigned long long* hv = b->hsmall;
        BHIP(hipMemcpyAsync(&hv[0], b->total, 8, hipMemcpyDeviceToHost, b->st));
        BHIP(hipMemcpyAsync(&hv[1], &b->stats->err, 8, hipMemcpyDeviceToHost, b->st));
        BHIP(hipMemcpyAsync(&hv[2], &b->cst[id].err, 8, hipMemcpyDeviceToHost, b->st));
        BHIP(hipStreamSynchronize(b->st));
        if (hv[2]) return load_fail(hv[2]);
        if (hv[1]) BCK(check_err(b));
        const uint64_t tot = hv[0];
        *n = tot;
        if (tot > cap) return bfail(SQE_EINVAL, "export capacity too small (" + std::to_string(tot) + " needed)");
        if (tot) {
            if (!keys || !re || !im) return bfail(SQE_EINVAL, "null array");
            BHIP(hipMemcpyAsync(keys, b->xkeys, tot * 8, hipMemcpyDeviceToHost, b->st));
            BHIP(hipMemcpyAsync(re, b->xre, tot * 8, hipMemcpyDeviceToHost, b->st));
            BHIP(hipMemcpyAsync(im, b->xim, tot * 8, hipMemcpyDeviceToHost, b->st));
            BHIP(hipStreamSynchronize(b->st));
        }
        return SQE_OK;
    });
}

sqe_status sqe_batch_remove(sqe_batch* b, int64_t id) {
    return guard([&]() -> sqe_status {
        if (!b || id < 0 || (uint64_t)id >= b->c.size() || !b->c[id].live) return bfail(SQE_EINVAL, "unknown circuit");
        BHIP(hipSetDevice(b->cfg.device));
        // no wait: work already enqueued on the region (an export) runs before
        // anything a later admission enqueues there, and pool reallocation
        // synchronises before it frees
        release(b, (uint32_t)id);
        return SQE_OK;
    });
}

sqe_status sqe_batch_export_start(sqe_batch* b, int64_t id, uint64_t* keys, double* re, double* im, uint64_t cap,
                                  int64_t* ticket, uint64_t* bound) {
    return guard([&]() -> sqe_status {
        if (!b || !ticket) return bfail(SQE_EINVAL, "null argument");
        *ticket = 0;
        if (id < 0 || (uint64_t)id >= b->c.size() || !b->c[id].live) return bfail(SQE_EINVAL, "unknown circuit");
        BHIP(hipSetDevice(b->cfg.device));
        const Circ& c = b->c[id];
        const uint64_t nbk = c.tiles * 64, slots = nbk << c.qm.mb;
        if (bound) *bound = slots;
        if (cap < slots)
            return bfail(SQE_EINVAL, "export capacity below the circuit's slot count (" + std::to_string(slots) + ")");
        if (!keys || !re || !im) return bfail(SQE_EINVAL, "null array");
        void* dv[3] = {nullptr, nullptr, nullptr};
        void* hv[3] = {keys, re, im};
        for (int k = 0; k < 3; ++k) BCK(device_view(hv[k], &dv[k]));
        if (!b->hrec) {
            if (hipHostMalloc((void**)&b->hrec, REC_SLOTS * REC_WORDS * 8, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                b->hrec = nullptr;
                return bfail(SQE_ENOMEM, "hipHostMalloc");
            }
            BHIP(hipHostGetDevicePointer((void**)&b->drec, b->hrec, 0));
            for (uint32_t k = REC_SLOTS; k-- > 0;) b->rec_free.push_back(k);
        }
        if (b->rec_free.empty())
            return bfail(SQE_EINVAL, std::to_string(REC_SLOTS) + " exports in flight: poll some before starting more");
        // device staging: keys, real and imaginary parts of up to `slots`
        // entries (after one aligned spare unit)
        const uint64_t arr = ring_round(slots * 8), bytes = RING_ALIGN + 3 * arr;
        uint64_t off = 0;
        BCK(ring_take(b, b->xr, bytes, XR_RING_MIN, &off));
        char* sp = b->xr.p + off;
        uint64_t* sk = (uint64_t*)(sp + RING_ALIGN);
        double* sr = (double*)(sp + RING_ALIGN + arr);
        double* si = (double*)(sp + RING_ALIGN + 2 * arr);
        // the export kernel writes the header (entry count, counter
        // snapshot) straight into the ticket's record in page-locked host
        // memory; the caller reads it after the ticket's event, which follows
        // the kernel
        const uint32_t rec = b->rec_free.back();
        unsigned long long* rec_dev = b->drec + (uint64_t)rec * REC_WORDS;
        ExportArgs xa{};
        xa.base = base_ptr(b, c); xa.amp = amp_ptr(b, c); xa.nb = b->meta + id; xa.qm = c.qm; xa.thr = b->thr;
        xa.filter = 1; xa.keys = sk; xa.re = sr; xa.im = si; xa.cap = slots; xa.total = rec_dev;
        xa.snap_src[0] = b->upd + id; xa.snap_src[1] = b->nnz + id; xa.snap_src[2] = b->meta + id;
        xa.snap_src[3] = &b->cst[id].err;
        xa.snap = rec_dev + 1;
        const uint32_t g = std::max<uint32_t>(1, (uint32_t)((nbk + BLOCK - 1) / BLOCK));
        BCK(next_scan(b, g, &xa.scan));
        BHIP(launch_export(c.qm.mb, b->f32, b->st, g, xa));
        // the copy to the caller runs on the export stream, off the step
        // stream: the entries by DMA (one pitched copy of the three arrays when
        // the caller's arrays are equally spaced, as sqe_host_alloc buffers
        // split in three are; all `slots` staging entries, of which the
        // header's count are valid)
        BHIP(hipEventRecord(b->xev, b->st));
        BHIP(hipStreamWaitEvent(b->xst, b->xev, 0));
        const uint64_t wbytes = slots * 8;
        const char* k0 = (const char*)keys;
        const ptrdiff_t pitch = (const char*)re - k0;
        if (pitch > 0 && (const char*)im - (const char*)re == pitch && (uint64_t)pitch >= wbytes) {
            BHIP(hipMemcpy2DAsync(keys, (size_t)pitch, sk, (size_t)arr, wbytes, 3, hipMemcpyDeviceToHost, b->xst));
        } else {
            BHIP(hipMemcpyAsync(keys, sk, wbytes, hipMemcpyDeviceToHost, b->xst));
            BHIP(hipMemcpyAsync(re, sr, wbytes, hipMemcpyDeviceToHost, b->xst));
            BHIP(hipMemcpyAsync(im, si, wbytes, hipMemcpyDeviceToHost, b->xst));
        }
        BCK(ring_mark(b->xr, off, bytes, b->xst));
        hipEvent_t ev = nullptr;
        if (!b->evpool.empty()) { ev = b->evpool.back(); b->evpool.pop_back(); }
        else BHIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        const hipError_t er = hipEventRecord(ev, b->xst);
        if (er != hipSuccess) {
            b->evpool.push_back(ev);
            BHIP(er);
        }
        b->rec_free.pop_back();
        XTicket t;
        t.rec = rec;
        t.ev = ev;
        t.cap = cap;
        t.info.done = c.pos >= c.nops ? 1 : 0;
        t.info.bucket_bits = c.qm.mb;
        t.info.gates_total = c.gates;
        t.info.gates_done = (uint64_t)(std::lower_bound(c.gate_at.begin(), c.gate_at.end(), (uint32_t)c.pos) -
                                       c.gate_at.begin());
        t.info.bytes = c.bytes;
        const int64_t tk = b->next_ticket++;
        b->xt[tk] = t;
        *ticket = tk;
        return SQE_OK;
    });
}

sqe_status sqe_batch_export_poll(sqe_batch* b, int64_t ticket, int32_t wait, int32_t* done, uint64_t* n,
                                 sqe_batch_circuit* info) {
    return guard([&]() -> sqe_status {
        if (!b || !done) return bfail(SQE_EINVAL, "null argument");
        *done = 0;
        auto it = b->xt.find(ticket);
        if (it == b->xt.end()) return bfail(SQE_EINVAL, "unknown export ticket");
        BHIP(hipSetDevice(b->cfg.device));
        XTicket& t = it->second;
        if (wait) {
            BHIP(hipEventSynchronize(t.ev));
        } else {
            const hipError_t q = hipEventQuery(t.ev);
            if (q == hipErrorNotReady) return SQE_OK;
            BHIP(q);
        }
        const unsigned long long* r = b->hrec + (uint64_t)t.rec * REC_WORDS;
        sqe_batch_circuit ci = t.info;
        const uint64_t tot = r[0];
        ci.updates = r[1];
        ci.nnz = r[2];
        ci.buckets = r[3];
        const unsigned long long lerr = r[4];
        b->evpool.push_back(t.ev);
        b->rec_free.push_back(t.rec);
        b->xt.erase(it);
        *done = 1;
        if (n) *n = tot;
        if (info) *info = ci;
        if (lerr) return load_fail(lerr);
        return SQE_OK;
    });
}

sqe_status sqe_batch_finished(sqe_batch* b, int64_t* ids, uint64_t cap, uint64_t* n) {
    return guard([&]() -> sqe_status {
        if (!b || !n) return bfail(SQE_EINVAL, "null argument");
        uint64_t k = 0;
        for (uint32_t id = 0; id < b->c.size(); ++id) {
            const Circ& c = b->c[id];
            if (!c.live || c.pos < c.nops) continue;
            if (ids && k < cap) ids[k] = id;
            ++k;
        }
        *n = k;
        return k > cap ? bfail(SQE_EINVAL, "id capacity too small") : SQE_OK;
    });
}

sqe_status sqe_batch_usage(sqe_batch* b, uint64_t* used, uint64_t* budget, uint64_t* active) {
    return guard([&]() -> sqe_status {
        if (!b) return bfail(SQE_EINVAL, "null batch");
        if (used) *used = b->used;
        if (budget) *budget = b->cfg.hbm_budget_bytes;
        if (active) {
            uint64_t a = 0;
            for (const Circ& c : b->c) a += c.live && c.pos < c.nops ? 1 : 0;
            *active = a;
        }
        return SQE_OK;
    });
}

// Admission builds a hash index of the circuit's bucket bases sized by the
// load factor: a lower factor shortens probe sequences but makes the table
// (which the scan numbering reads whole) larger.  Time the index part of the
// load path -- table clear, cooperative insert of the bases, scan numbering --
// over every admitted circuit at each candidate factor, on the current
// bucket bases (one per bucket, as admission inserts them), with the
// numbering written to scratch so the circuits are untouched; keep the
// fastest for later admissions.
sqe_status tune_load_factor(sqe_batch* b) {
    std::vector<uint32_t> ids;
    uint64_t maxn = 0;
    for (uint32_t id = 0; id < b->c.size(); ++id)
        if (b->c[id].live) {
            ids.push_back(id);
            maxn = std::max<uint64_t>(maxn, b->c[id].tiles * 64);
        }
    if (ids.empty()) return SQE_OK;
    const double cands[] = {0.25, 0.375, 0.5, 0.625, 0.75};
    auto tsize = [](uint64_t n, double lf) {
        return std::max<uint64_t>(1024, p2((uint64_t)std::ceil((double)n / lf) + 8));
    };
    const uint64_t tmax = tsize(maxn, cands[0]);
    if (tmax > b->tcap) {
        BHIP(hipStreamSynchronize(b->st));
        dfr(b, b->tab, b->tcap * sizeof(Slot));
        b->tab = nullptr;
        BCK(dal(b, (void**)&b->tab, tmax * sizeof(Slot)));
        b->tcap = tmax;
    }
    // scratch: numbered bases, their count and an error word
    char* scratch = nullptr;
    const uint64_t sbytes = (maxn + 64) * 8 + 8 + sizeof(Stats);
    BCK(dal(b, (void**)&scratch, sbytes));
    uint64_t* out_base = (uint64_t*)scratch;
    unsigned long long* out_nb = (unsigned long long*)(scratch + (maxn + 64) * 8);
    Stats* st = (Stats*)(out_nb + 1);
    sqe_status rc = SQE_OK;
    double best = 1e300, best_lf = b->cfg.max_load > 0 && b->cfg.max_load < 0.95 ? b->cfg.max_load : 0.5;
    for (int rep = 0; rep < 2 && rc == SQE_OK; ++rep) {      // the first pass warms up
        for (double lf : cands) {
            hipError_t he = hipEventRecord(b->ev0, b->st);
            for (uint32_t id : ids) {
                if (he != hipSuccess || rc != SQE_OK) break;
                const Circ& c = b->c[id];
                const uint64_t tn = tsize(c.tiles * 64, lf);
                he = launch_fill_slots(b->st, b->tab, tn);
                IndexArgs ia{base_ptr(b, c), b->meta + id, b->tab, tn - 1, st};
                if (he == hipSuccess) he = launch_index_build(b->st, (uint32_t)((c.tiles * 64 + BLOCK - 1) / BLOCK), ia);
                EnumArgs ea{};
                ea.tab = b->tab; ea.tcap = tn; ea.out_base = out_base; ea.out_nb = out_nb;
                ea.out_cap = maxn + 64; ea.stats = st;
                const uint32_t eg = (uint32_t)((tn + BLOCK - 1) / BLOCK);
                rc = next_scan(b, eg, &ea.scan);
                if (rc == SQE_OK && he == hipSuccess) he = launch_index_enumerate(b->st, eg, ea);
            }
            if (he == hipSuccess) he = hipEventRecord(b->ev1, b->st);
            if (he == hipSuccess) he = hipEventSynchronize(b->ev1);
            float ms = 0.f;
            if (he == hipSuccess) he = hipEventElapsedTime(&ms, b->ev0, b->ev1);
            if (he != hipSuccess) rc = bfail(SQE_EHIP, std::string("load factor tuning: ") + hipGetErrorString(he));
            if (rc != SQE_OK) break;
            if (rep == 1 && ms < best) { best = ms; best_lf = lf; }
        }
    }
    (void)hipStreamSynchronize(b->st);
    dfr(b, scratch, sbytes);
    if (rc != SQE_OK) return rc;
    b->cfg.max_load = best_lf;
    b->load_ms = best;
    return SQE_OK;
}

sqe_status sqe_batch_get_tuning(sqe_batch* b, sqe_batch_tuning* t) {
    return guard([&]() -> sqe_status {
        if (!b || !t) return bfail(SQE_EINVAL, "null argument");
        t->block = b->block;
        t->tiles_per_item = b->gran;
        t->load_factor = b->cfg.max_load > 0 ? b->cfg.max_load : 0.5;
        t->sweep_ms = b->sweep_ms;
        t->launches = b->launches;
        t->load_ms = b->load_ms;
        t->one_launch = b->one_launch ? 1 : 0;
        t->reserved = 0;
        return SQE_OK;
    });
}

sqe_status sqe_batch_set_tuning(sqe_batch* b, const sqe_batch_tuning* t) {
    return guard([&]() -> sqe_status {
        if (!b || !t) return bfail(SQE_EINVAL, "null argument");
        if (t->block != 64 && t->block != 128 && t->block != 256) return bfail(SQE_EINVAL, "block must be 64/128/256");
        if (t->tiles_per_item < 1 || t->tiles_per_item > 64) return bfail(SQE_EINVAL, "tiles_per_item in 1..64");
        if (t->one_launch > 1) return bfail(SQE_EINVAL, "one_launch must be 0 or 1");
        b->block = t->block;
        b->gran = t->tiles_per_item;
        b->one_launch = t->one_launch != 0;
        if (t->load_factor > 0 && t->load_factor < 0.95) b->cfg.max_load = t->load_factor;
        return SQE_OK;
    });
}

constexpr int TUNE_REPS = 3;               // timed runs per autotuner candidate

// Time every (workgroup size, tiles per item, one_launch) candidate on the
// batch as it stands -- state and positions are saved and restored around each
// trial -- and keep the fastest.  With captured runs, each candidate is timed
// on a replay of its graph.
sqe_status sqe_batch_autotune(sqe_batch* b, sqe_batch_tuning* best_out) {
    return guard([&]() -> sqe_status {
        if (!b) return bfail(SQE_EINVAL, "null batch");
        BHIP(hipSetDevice(b->cfg.device));
        BCK(snap_save(b));
        struct Cand { uint32_t block, gran; bool one; };
        std::vector<Cand> cands;
        for (bool one : {false, true})
            for (uint32_t bl : {64u, 128u, 256u})
                for (uint32_t g : {1u, 2u, 4u, 8u}) cands.push_back(Cand{bl, g, one});
        double best = 1e300;
        Cand bc{b->block, b->gran, b->one_launch};
        const uint32_t timing = b->cfg.timing;
        b->cfg.timing = 0;
        for (const Cand& cd : cands) {
            b->block = cd.block;
            b->gran = cd.gran;
            b->one_launch = cd.one;
            if (!b->cfg.batch_eager) {
                // capture and instantiate this candidate's graph outside the timing
                BCK(sqe_batch_run(b));
                BCK(snap_restore(b));
                BHIP(hipStreamSynchronize(b->st));
            }
            // the fastest of TUNE_REPS replays: one replay's time varies by a
            // few percent, about as much as neighbouring candidates differ
            for (int rep = 0; rep < TUNE_REPS; ++rep) {
                BHIP(hipEventRecord(b->ev0, b->st));
                BCK(sqe_batch_run(b));
                BHIP(hipEventRecord(b->ev1, b->st));
                BHIP(hipEventSynchronize(b->ev1));
                float ms = 0.f;
                BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                if (ms < best) { best = ms; bc = cd; }
                BCK(snap_restore(b));
                BHIP(hipStreamSynchronize(b->st));
            }
        }
        b->cfg.timing = timing;
        b->block = bc.block;
        b->gran = bc.gran;
        b->one_launch = bc.one;
        BCK(tune_load_factor(b));
        if (best_out) {
            BCK(sqe_batch_get_tuning(b, best_out));
            best_out->sweep_ms = best;
        }
        return SQE_OK;
    });
}

sqe_status sqe_batch_snapshot_save(sqe_batch* b) {
    return guard([&]() -> sqe_status {
        if (!b) return bfail(SQE_EINVAL, "null batch");
        BHIP(hipSetDevice(b->cfg.device));
        return snap_save(b);
    });
}

sqe_status sqe_batch_snapshot_restore(sqe_batch* b) {
    return guard([&]() -> sqe_status {
        if (!b) return bfail(SQE_EINVAL, "null batch");
        BHIP(hipSetDevice(b->cfg.device));
        return snap_restore(b);
    });
}

const char* sqe_batch_last_error(void) { return b_err.c_str(); }

sqe_status sqe_host_alloc(uint64_t bytes, void** out) {
    return guard([&]() -> sqe_status {
        if (!out) return bfail(SQE_EINVAL, "null argument");
        *out = nullptr;
        if (!bytes) return SQE_OK;
        if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            *out = nullptr;
            return bfail(SQE_ENOMEM, "hipHostMalloc(" + std::to_string(bytes) + ") failed");
        }
        return SQE_OK;
    });
}

sqe_status sqe_host_free(void* p) {
    return guard([&]() -> sqe_status {
        if (p && hipHostFree(p) != hipSuccess) {
            (void)hipGetLastError();
            return bfail(SQE_EHIP, "hipHostFree failed");
        }
        return SQE_OK;
    });
}

}  // extern "C"
