// CPU oracle for the sparse statevector engine -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU, the gate semantics of the reference's GPU path
// (ivanNovikau/SparQC, emulator/) so that the HIP engine can be checked against
// it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it; the engine never does.
//
// What it restates (reference file:line):
//   * gate arithmetic, operation order included:
//       x_core   emulator/device/device_data.cuh:159-172
//       y_core   emulator/device/device_data.cuh:175-184
//       z_core   emulator/device/device_data.cuh:187-196
//       h_core   emulator/device/device_data.cuh:199-209 with C_PA/C_PS
//                emulator/include/data_t.h:63-80 (f1 = 1./sqrt(2.))
//       phase    emulator/device/device_data.cuh:212-231
//       rz       emulator/device/device_data.cuh:234-263
//       rx       emulator/device/device_data.cuh:266-287
//       ry       emulator/device/device_data.cuh:290-319
//   * angle conventions: theta/2 for rz/rx/ry, theta for phase
//       emulator/device/table_manager.cuh:333-336 (applied by the host caller)
//   * pair discovery and ownership (the bit-0 element owns the pair when present)
//       emulator/device/device_table_FC.cuh:62-110
//   * pruning only on outputs of branching gates, componentwise predicate
//       |re| < thr && |im| < thr
//       emulator/device/device_data.cuh:134-156, emulator/include/data_t.h:50-60,
//       ZERO_THRESHOLD emulator/include/sim_parameters_types.h:47
//   * readout drops negligible values   emulator/device/table.cuh:58-73
// The state is a map key -> amplitude (the reference's FC address book of
// present entries, emulator/device/table.cuh:15, without its duplicate-entry and
// race defects).  Controls (absent from the reference) restrict a gate to keys
// with (k & ctrl_mask) == ctrl_val.
//
// Added forms restated here as the checker of the engine's own additions:
//   * FP32 amplitudes with FMA gate arithmetic; near-threshold prune decisions
//     (max(|re|,|im|) in [thr/4, 4 thr)) redone in FP64 from the FP32 inputs;
//     renormalisation in FP64 every R gates.
//   * a dense FP64 statevector (nq <= 26) with the same per-gate arithmetic and
//     no pruning.
//
// Parity pinning: tests/test_oracle.py checks every gate against closed-form
// matrices, the sparse engine against the dense statevector on seeded circuits,
// the all-H scenario of emulator/src/main.cu:142-143, inverse circuits, norm and
// key-set invariants, and brute-force key transforms on small qubit counts.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Gate {               // layout of sqe_gate (88 bytes)
    uint32_t kind, target;
    uint64_t cmask, cval;
    double p[8];
};
static_assert(sizeof(Gate) == 88, "gate record layout");

enum { KX = 0, KY, KZ, KH, KPHASE, KRZ, KRX, KRY, KU, KDIAG, KPERM };

inline int gate_class(uint32_t k) {   // 0 diagonal, 1 permutation, 2 branching
    switch (k) {
    case KZ: case KPHASE: case KRZ: case KDIAG: return 0;
    case KX: case KY: case KPERM: return 1;
    default: return 2;
    }
}

template <class R> struct Cx { R r, i; };

// ---- FP64 arithmetic: the reference's formulas, no contraction -------------
inline void op64(const Gate& g, Cx<double> v0, Cx<double> v1, Cx<double>& n0, Cx<double>& n1) {
    const double* p = g.p;
    switch (g.kind) {
    case KX: n0 = v1; n1 = v0; break;
    case KY: n0 = {v1.i, -v1.r}; n1 = {-v0.i, v0.r}; break;
    case KZ: n0 = v0; n1 = {-v1.r, -v1.i}; break;
    case KH: {
        double f = p[0];
        n0 = {f * (v0.r + v1.r), f * (v0.i + v1.i)};
        n1 = {f * (v0.r - v1.r), f * (v0.i - v1.i)};
    } break;
    case KPHASE: {
        double c = p[0], s = p[1];
        n0 = v0;
        n1 = {c * v1.r - s * v1.i, c * v1.i + s * v1.r};
    } break;
    case KRZ: {
        double c = p[0], s = p[1];
        n0 = {c * v0.r + s * v0.i, c * v0.i - s * v0.r};
        n1 = {c * v1.r - s * v1.i, c * v1.i + s * v1.r};
    } break;
    case KRX: {
        double c = p[0], s = p[1];
        n0 = {c * v0.r + s * v1.i, c * v0.i - s * v1.r};
        n1 = {c * v1.r + s * v0.i, c * v1.i - s * v0.r};
    } break;
    case KRY: {
        double c = p[0], s = p[1];
        n0 = {c * v0.r - s * v1.r, c * v0.i - s * v1.i};
        n1 = {c * v1.r + s * v0.r, c * v1.i + s * v0.i};
    } break;
    case KU: {
        n0 = {(p[0] * v0.r - p[1] * v0.i) + (p[2] * v1.r - p[3] * v1.i),
              (p[0] * v0.i + p[1] * v0.r) + (p[2] * v1.i + p[3] * v1.r)};
        n1 = {(p[4] * v0.r - p[5] * v0.i) + (p[6] * v1.r - p[7] * v1.i),
              (p[4] * v0.i + p[5] * v0.r) + (p[6] * v1.i + p[7] * v1.r)};
    } break;
    case KDIAG: {
        n0 = {p[0] * v0.r - p[1] * v0.i, p[0] * v0.i + p[1] * v0.r};
        n1 = {p[2] * v1.r - p[3] * v1.i, p[2] * v1.i + p[3] * v1.r};
    } break;
    case KPERM: {
        n0 = {p[0] * v1.r - p[1] * v1.i, p[0] * v1.i + p[1] * v1.r};
        n1 = {p[2] * v0.r - p[3] * v0.i, p[2] * v0.i + p[3] * v0.r};
    } break;
    default: n0 = v0; n1 = v1;
    }
}

// ---- FP32 arithmetic: the same formulas with fused multiply-add ------------
inline void op32(const Gate& g, Cx<float> v0, Cx<float> v1, Cx<float>& n0, Cx<float>& n1) {
    float p[8];
    for (int k = 0; k < 8; ++k) p[k] = (float)g.p[k];
    switch (g.kind) {
    case KX: n0 = v1; n1 = v0; break;
    case KY: n0 = {v1.i, -v1.r}; n1 = {-v0.i, v0.r}; break;
    case KZ: n0 = v0; n1 = {-v1.r, -v1.i}; break;
    case KH: {
        float f = p[0];
        n0 = {f * (v0.r + v1.r), f * (v0.i + v1.i)};
        n1 = {f * (v0.r - v1.r), f * (v0.i - v1.i)};
    } break;
    case KPHASE: {
        float c = p[0], s = p[1];
        n0 = v0;
        n1 = {std::fma(c, v1.r, -(s * v1.i)), std::fma(c, v1.i, s * v1.r)};
    } break;
    case KRZ: {
        float c = p[0], s = p[1];
        n0 = {std::fma(c, v0.r, s * v0.i), std::fma(c, v0.i, -(s * v0.r))};
        n1 = {std::fma(c, v1.r, -(s * v1.i)), std::fma(c, v1.i, s * v1.r)};
    } break;
    case KRX: {
        float c = p[0], s = p[1];
        n0 = {std::fma(c, v0.r, s * v1.i), std::fma(c, v0.i, -(s * v1.r))};
        n1 = {std::fma(c, v1.r, s * v0.i), std::fma(c, v1.i, -(s * v0.r))};
    } break;
    case KRY: {
        float c = p[0], s = p[1];
        n0 = {std::fma(c, v0.r, -(s * v1.r)), std::fma(c, v0.i, -(s * v1.i))};
        n1 = {std::fma(c, v1.r, s * v0.r), std::fma(c, v1.i, s * v0.i)};
    } break;
    case KU: {
        float t;
        t = p[0] * v0.r; t = std::fma(-p[1], v0.i, t); t = std::fma(p[2], v1.r, t); n0.r = std::fma(-p[3], v1.i, t);
        t = p[0] * v0.i; t = std::fma(p[1], v0.r, t); t = std::fma(p[2], v1.i, t); n0.i = std::fma(p[3], v1.r, t);
        t = p[4] * v0.r; t = std::fma(-p[5], v0.i, t); t = std::fma(p[6], v1.r, t); n1.r = std::fma(-p[7], v1.i, t);
        t = p[4] * v0.i; t = std::fma(p[5], v0.r, t); t = std::fma(p[6], v1.i, t); n1.i = std::fma(p[7], v1.r, t);
    } break;
    case KDIAG: {
        n0 = {std::fma(p[0], v0.r, -(p[1] * v0.i)), std::fma(p[0], v0.i, p[1] * v0.r)};
        n1 = {std::fma(p[2], v1.r, -(p[3] * v1.i)), std::fma(p[2], v1.i, p[3] * v1.r)};
    } break;
    case KPERM: {
        n0 = {std::fma(p[0], v1.r, -(p[1] * v1.i)), std::fma(p[0], v1.i, p[1] * v1.r)};
        n1 = {std::fma(p[2], v0.r, -(p[3] * v0.i)), std::fma(p[2], v0.i, p[3] * v0.r)};
    } break;
    default: n0 = v0; n1 = v1;
    }
}

// Diagonal / permutation gates act on one element at a time: element with key
// bit b gets the b-th output of the pair formula with its partner taken as 0.
template <class R>
inline Cx<R> single(const Gate& g, Cx<R> v, int b);

template <>
inline Cx<double> single<double>(const Gate& g, Cx<double> v, int b) {
    Cx<double> z{0.0, 0.0}, n0, n1;
    if (b == 0) { op64(g, v, z, n0, n1); return gate_class(g.kind) == 1 ? n1 : n0; }
    op64(g, z, v, n0, n1);
    return gate_class(g.kind) == 1 ? n0 : n1;
}
template <>
inline Cx<float> single<float>(const Gate& g, Cx<float> v, int b) {
    Cx<float> z{0.0f, 0.0f}, n0, n1;
    if (b == 0) { op32(g, v, z, n0, n1); return gate_class(g.kind) == 1 ? n1 : n0; }
    op32(g, z, v, n0, n1);
    return gate_class(g.kind) == 1 ? n0 : n1;
}

inline bool negl(double r, double i, double thr) { return std::fabs(r) < thr && std::fabs(i) < thr; }

inline uint64_t hmix(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33; return k;
}

// open-addressing key -> index map used for partner lookups
struct Index {
    std::vector<uint64_t> keys;
    std::vector<uint64_t> vals;
    uint64_t mask = 0;
    static constexpr uint64_t EMPTY = ~0ULL;
    void build(const std::vector<uint64_t>& k) {
        uint64_t cap = 16;
        while (cap < 2 * k.size() + 16) cap <<= 1;
        mask = cap - 1;
        keys.assign(cap, EMPTY);
        vals.assign(cap, 0);
        const int64_t n = (int64_t)k.size();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint64_t h = hmix(k[i]) & mask;
            for (;;) {
                uint64_t expect = EMPTY;
                if (__atomic_compare_exchange_n(&keys[h], &expect, k[i], false,
                                                __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                    vals[h] = (uint64_t)i;
                    break;
                }
                h = (h + 1) & mask;
            }
        }
    }
    int64_t find(uint64_t key) const {
        uint64_t h = hmix(key) & mask;
        for (;;) {
            uint64_t k = keys[h];
            if (k == key) return (int64_t)vals[h];
            if (k == EMPTY) return -1;
            h = (h + 1) & mask;
        }
    }
};

int nthreads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
int tid() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

struct Engine {
    int nq;
    int f32;
    double thr;
    int renorm_every;
    uint64_t n_applied = 0;
    uint64_t updates = 0;          // sum over gates of the entries entering the gate
    std::vector<uint64_t> keys;
    std::vector<Cx<double>> a64;
    std::vector<Cx<float>> a32;

    template <class R> std::vector<Cx<R>>& amps();

    uint64_t size() const { return keys.size(); }

    template <class R> void apply(const Gate& g);
    template <class R> void branch(const Gate& g);
    template <class R> void renorm();
};
template <> std::vector<Cx<double>>& Engine::amps<double>() { return a64; }
template <> std::vector<Cx<float>>& Engine::amps<float>() { return a32; }

template <class R>
void Engine::apply(const Gate& g) {
    const uint64_t tb = 1ULL << g.target;
    auto& A = amps<R>();
    const int64_t n = (int64_t)keys.size();
    const int cls = gate_class(g.kind);
    updates += (uint64_t)n;
    if (cls == 0) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint64_t k = keys[i];
            if ((k & g.cmask) != g.cval) continue;
            A[i] = single<R>(g, A[i], (int)((k >> g.target) & 1));
        }
    } else if (cls == 1) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            uint64_t k = keys[i];
            if ((k & g.cmask) != g.cval) continue;
            A[i] = single<R>(g, A[i], (int)((k >> g.target) & 1));
            keys[i] = k ^ tb;
        }
    } else {
        branch<R>(g);
    }
    ++n_applied;
    if (renorm_every > 0 && n_applied % (uint64_t)renorm_every == 0) renorm<R>();
}

// FP64 re-decision for an FP32 output near the threshold.
inline bool keep32(const Gate& g, Cx<float> v0, Cx<float> v1, Cx<float> out, int which, double thr) {
    double m = std::max(std::fabs((double)out.r), std::fabs((double)out.i));
    if (m >= 4.0 * thr) return true;
    if (m < 0.25 * thr) return false;
    Cx<double> w0{(double)v0.r, (double)v0.i}, w1{(double)v1.r, (double)v1.i}, n0, n1;
    op64(g, w0, w1, n0, n1);
    const Cx<double>& o = which == 0 ? n0 : n1;
    return !negl(o.r, o.i, thr);
}

template <class R>
void Engine::branch(const Gate& g) {
    const uint64_t tb = 1ULL << g.target;
    auto& A = amps<R>();
    Index idx;
    idx.build(keys);
    const int64_t n = (int64_t)keys.size();
    const int T = nthreads();
    std::vector<std::vector<uint64_t>> ok(T);
    std::vector<std::vector<Cx<R>>> oa(T);
#pragma omp parallel
    {
        const int me = tid();
        auto& K = ok[me];
        auto& V = oa[me];
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const uint64_t k = keys[i];
            if ((k & g.cmask) != g.cval) { K.push_back(k); V.push_back(A[i]); continue; }
            const int b = (int)((k >> g.target) & 1);
            const int64_t j = idx.find(k ^ tb);
            if (b == 1 && j >= 0) continue;            // the bit-0 partner owns this pair
            Cx<R> zero{R(0), R(0)};
            Cx<R> own = A[i], par = j >= 0 ? A[j] : zero;
            Cx<R> v0 = b ? par : own, v1 = b ? own : par, n0, n1;
            bool k0, k1;
            if constexpr (sizeof(R) == 8) {
                op64(g, v0, v1, n0, n1);
                k0 = !negl(n0.r, n0.i, thr);
                k1 = !negl(n1.r, n1.i, thr);
            } else {
                op32(g, v0, v1, n0, n1);
                k0 = keep32(g, v0, v1, n0, 0, thr);
                k1 = keep32(g, v0, v1, n1, 1, thr);
            }
            if (k0) { K.push_back(k & ~tb); V.push_back(n0); }
            if (k1) { K.push_back(k | tb); V.push_back(n1); }
        }
    }
    std::vector<uint64_t> nk;
    std::vector<Cx<R>> na;
    size_t tot = 0;
    for (int t = 0; t < T; ++t) tot += ok[t].size();
    nk.reserve(tot);
    na.reserve(tot);
    for (int t = 0; t < T; ++t) {
        nk.insert(nk.end(), ok[t].begin(), ok[t].end());
        na.insert(na.end(), oa[t].begin(), oa[t].end());
    }
    keys.swap(nk);
    A.swap(na);
}

template <class R>
void Engine::renorm() {
    auto& A = amps<R>();
    double s = 0.0;
    for (auto& v : A) s += (double)v.r * (double)v.r + (double)v.i * (double)v.i;
    if (s <= 0.0) return;
    const double f = 1.0 / std::sqrt(s);
    for (auto& v : A) { v.r = (R)((double)v.r * f); v.i = (R)((double)v.i * f); }
}

}  // namespace

extern "C" {

void* orc_create(int nq, int f32, double thr, int renorm_every) {
    if (nq < 1 || nq > 63) return nullptr;
    Engine* e = new Engine();
    e->nq = nq;
    e->f32 = f32;
    e->thr = thr;
    e->renorm_every = renorm_every;
    return e;
}

void orc_destroy(void* h) { delete (Engine*)h; }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void) { return nthreads(); }

// Load distinct keys (zero amplitudes are dropped: an absent key has amplitude 0).
int orc_load(void* h, const uint64_t* keys, const double* re, const double* im, uint64_t n) {
    Engine* e = (Engine*)h;
    e->keys.clear(); e->a64.clear(); e->a32.clear();
    const uint64_t lim = e->nq == 64 ? ~0ULL : ((1ULL << e->nq) - 1);
    for (uint64_t i = 0; i < n; ++i) {
        if (keys[i] & ~lim) return -1;
        if (re[i] == 0.0 && im[i] == 0.0) continue;
        e->keys.push_back(keys[i]);
        if (e->f32) e->a32.push_back({(float)re[i], (float)im[i]});
        else e->a64.push_back({re[i], im[i]});
    }
    std::vector<uint64_t> s(e->keys);
    std::sort(s.begin(), s.end());
    if (std::adjacent_find(s.begin(), s.end()) != s.end()) return -2;
    e->n_applied = 0;
    return 0;
}

int orc_apply(void* h, const void* gates, uint64_t ng) {
    Engine* e = (Engine*)h;
    const Gate* g = (const Gate*)gates;
    for (uint64_t i = 0; i < ng; ++i) {
        if (g[i].target >= (uint32_t)e->nq) return -1;
        if (g[i].cmask >> g[i].target & 1) return -1;
        if (g[i].kind > KPERM) return -1;
        if (e->f32) e->apply<float>(g[i]);
        else e->apply<double>(g[i]);
    }
    return 0;
}

uint64_t orc_nnz(void* h) { return ((Engine*)h)->size(); }

uint64_t orc_updates(void* h) { return ((Engine*)h)->updates; }

double orc_norm2(void* h) {
    Engine* e = (Engine*)h;
    double s = 0.0;
    if (e->f32) for (auto& v : e->a32) s += (double)v.r * v.r + (double)v.i * v.i;
    else for (auto& v : e->a64) s += v.r * v.r + v.i * v.i;
    return s;
}

// Export non-negligible entries sorted by key; returns the count (or the count
// needed when cap is too small, writing nothing).
uint64_t orc_export(void* h, uint64_t* keys, double* re, double* im, uint64_t cap) {
    Engine* e = (Engine*)h;
    std::vector<uint64_t> order;
    const uint64_t n = e->size();
    order.reserve(n);
    for (uint64_t i = 0; i < n; ++i) {
        double r = e->f32 ? (double)e->a32[i].r : e->a64[i].r;
        double m = e->f32 ? (double)e->a32[i].i : e->a64[i].i;
        if (!negl(r, m, e->thr)) order.push_back(i);
    }
    if (order.size() > cap) return order.size();
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return e->keys[a] < e->keys[b]; });
    for (size_t j = 0; j < order.size(); ++j) {
        const uint64_t i = order[j];
        keys[j] = e->keys[i];
        re[j] = e->f32 ? (double)e->a32[i].r : e->a64[i].r;
        im[j] = e->f32 ? (double)e->a32[i].i : e->a64[i].i;
    }
    return order.size();
}

// Dense FP64 statevector: same per-pair arithmetic, no pruning.
int orc_dense_run(int nq, const uint64_t* keys, const double* re, const double* im, uint64_t n,
                  const void* gates, uint64_t ng, double* out_re, double* out_im) {
    if (nq < 1 || nq > 28) return -1;
    const uint64_t N = 1ULL << nq;
    std::vector<Cx<double>> v(N, Cx<double>{0.0, 0.0});
    for (uint64_t i = 0; i < n; ++i) {
        if (keys[i] >= N) return -1;
        v[keys[i]] = {re[i], im[i]};
    }
    const Gate* g = (const Gate*)gates;
    for (uint64_t q = 0; q < ng; ++q) {
        const Gate& G = g[q];
        if (G.target >= (uint32_t)nq) return -1;
        const uint64_t tb = 1ULL << G.target;
        const int64_t half = (int64_t)(N >> 1);
#pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < half; ++j) {
            // insert a 0 at bit position `target` of j
            const uint64_t lo = (uint64_t)j & (tb - 1);
            const uint64_t k0 = (((uint64_t)j - lo) << 1) | lo;
            if ((k0 & G.cmask) != G.cval) continue;
            const uint64_t k1 = k0 | tb;
            Cx<double> n0, n1;
            op64(G, v[k0], v[k1], n0, n1);
            v[k0] = n0;
            v[k1] = n1;
        }
    }
    for (uint64_t k = 0; k < N; ++k) { out_re[k] = v[k].r; out_im[k] = v[k].i; }
    return 0;
}

// ---------------------------------------------------------------------------
// FC engine: the reference's full-capacity table restated on the CPU, as the
// CPU baseline of the 20-qubit workload (timing only; checked against the
// map engine above).  emulator/device/device_table_FC.cuh and
// device_data.cuh:
//   * two dense buffers of 2^nq amplitudes addressed by the key itself
//     (identity hash, table.cuh:13-26, device_table_FC.cuh:31-36) and an
//     address book of the occupied keys (ah_);
//   * every gate walks the address book (gate_sq, device_table_FC.cuh:113-168)
//     and writes the other buffer; diagonal and permutation gates keep their
//     book position (update_ah_diag / update_ah_swap, device_data.cuh:108-131),
//     branching gates append their non-negligible outputs (update_ah_sup,
//     device_data.cuh:134-156); the old buffer's entries are then cleared
//     (device_zero_arrays, device_data.cuh:62-105) and the buffers swap;
//   * pairing (find_elements, device_table_FC.cuh:62-110): the key with the
//     target bit 0 owns the pair; the bit-1 key skips when its companion is
//     in the table.  (The reference tests the companion's value for
//     negligibility; a present companion with a negligible value would make
//     both keys process the pair and append it twice, a defect the map
//     engine and the GPU engine do not reproduce.  Presence is a nonzero
//     slot here: pruned outputs are never stored.)
// Appends use per-thread lists concatenated in thread order instead of the
// reference's atomic counter.
// ---------------------------------------------------------------------------
namespace {
struct FC {
    int nq = 0;
    double thr = 1e-14;
    uint64_t updates = 0;
    std::vector<Cx<double>> buf[2];
    int cur = 0;
    std::vector<uint64_t> book;
};

inline bool is_zero(const Cx<double>& v) { return v.r == 0.0 && v.i == 0.0; }

void fc_gate(FC& f, const Gate& g) {
    auto& A = f.buf[f.cur];
    auto& B = f.buf[f.cur ^ 1];
    const uint64_t tb = 1ULL << g.target;
    const int64_t n = (int64_t)f.book.size();
    const int cls = gate_class(g.kind);
    f.updates += (uint64_t)n;
    std::vector<uint64_t> nb;
    if (cls != 2) {
        nb.resize(f.book.size());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            const uint64_t k = f.book[i];
            const bool on = (k & g.cmask) == g.cval;
            const uint64_t k2 = (on && cls == 1) ? (k ^ tb) : k;
            B[k2] = on ? single<double>(g, A[k], (int)((k >> g.target) & 1)) : A[k];
            nb[i] = k2;
        }
    } else {
        const int T = nthreads();
        std::vector<std::vector<uint64_t>> part(T);
#pragma omp parallel
        {
            auto& P = part[tid()];
#pragma omp for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                const uint64_t k = f.book[i];
                if ((k & g.cmask) != g.cval) {
                    B[k] = A[k];
                    P.push_back(k);
                    continue;
                }
                const int b = (int)((k >> g.target) & 1);
                const uint64_t c = k ^ tb;
                if (b == 1 && !is_zero(A[c])) continue;         // the bit-0 key owns the pair
                const Cx<double> v0 = b ? A[c] : A[k], v1 = b ? A[k] : A[c];
                Cx<double> n0, n1;
                op64(g, v0, v1, n0, n1);
                const uint64_t k0 = k & ~tb, k1 = k | tb;
                if (!negl(n0.r, n0.i, f.thr)) { B[k0] = n0; P.push_back(k0); }
                if (!negl(n1.r, n1.i, f.thr)) { B[k1] = n1; P.push_back(k1); }
            }
        }
        size_t tot = 0;
        for (auto& P : part) tot += P.size();
        nb.reserve(tot);
        for (auto& P : part) nb.insert(nb.end(), P.begin(), P.end());
    }
    // clear the old buffer's entries, then swap
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) A[f.book[i]] = Cx<double>{0.0, 0.0};
    f.book.swap(nb);
    f.cur ^= 1;
}
}  // namespace

void* fc_create(int nq, double thr) {
    if (nq < 1 || nq > 26) return nullptr;
    FC* f = new FC();
    f->nq = nq;
    f->thr = thr;
    f->buf[0].assign(1ULL << nq, Cx<double>{0.0, 0.0});
    f->buf[1].assign(1ULL << nq, Cx<double>{0.0, 0.0});
    return f;
}

void fc_destroy(void* h) { delete (FC*)h; }

int fc_load(void* h, const uint64_t* keys, const double* re, const double* im, uint64_t n) {
    FC* f = (FC*)h;
    for (uint64_t k : f->book) f->buf[f->cur][k] = Cx<double>{0.0, 0.0};
    f->book.clear();
    const uint64_t N = 1ULL << f->nq;
    for (uint64_t i = 0; i < n; ++i) {
        if (keys[i] >= N) return -1;
        if (re[i] == 0.0 && im[i] == 0.0) continue;
        if (!is_zero(f->buf[f->cur][keys[i]])) return -2;
        f->buf[f->cur][keys[i]] = Cx<double>{re[i], im[i]};
        f->book.push_back(keys[i]);
    }
    f->updates = 0;
    return 0;
}

int fc_apply(void* h, const void* gates, uint64_t ng) {
    FC* f = (FC*)h;
    const Gate* g = (const Gate*)gates;
    for (uint64_t i = 0; i < ng; ++i) {
        if (g[i].target >= (uint32_t)f->nq || (g[i].cmask >> g[i].target & 1) || g[i].kind > KPERM) return -1;
        fc_gate(*f, g[i]);
    }
    return 0;
}

uint64_t fc_updates(void* h) { return ((FC*)h)->updates; }
uint64_t fc_nnz(void* h) { return ((FC*)h)->book.size(); }

// Non-negligible entries sorted by key (the reference's readout filter,
// table.cuh:58-73); returns the count, writing nothing when cap is too small.
uint64_t fc_export(void* h, uint64_t* keys, double* re, double* im, uint64_t cap) {
    FC* f = (FC*)h;
    std::vector<uint64_t> ks;
    for (uint64_t k : f->book) {
        const Cx<double>& v = f->buf[f->cur][k];
        if (!negl(v.r, v.i, f->thr)) ks.push_back(k);
    }
    if (ks.size() > cap) return ks.size();
    std::sort(ks.begin(), ks.end());
    for (size_t j = 0; j < ks.size(); ++j) {
        keys[j] = ks[j];
        re[j] = f->buf[f->cur][ks[j]].r;
        im[j] = f->buf[f->cur][ks[j]].i;
    }
    return ks.size();
}

}  // extern "C"
