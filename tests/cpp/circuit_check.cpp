// Self-checking driver of the Circuit__ API (include/sparqc_circuit.hpp) on
// libsqe.so, written the way the reference's driver uses its Circuit__
// (emulator/src/main.cu:87-148).  Exit 0: every check passed; 3: no GPU (the
// constructor reported it as a std::string, the reference's error
// convention); 1: a check failed.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <unistd.h>
#include <vector>

#include "sparqc_circuit.hpp"

using cd = std::complex<double>;
static int failures = 0;

#define CHECK(cond, ...)                              \
    do {                                              \
        if (!(cond)) {                                \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                 \
            std::printf("\n");                        \
            ++failures;                               \
        }                                             \
    } while (0)

static double max_diff(Circuit__& a, Circuit__& b, bool* same_keys) {
    std::vector<uint64_t> ka, kb;
    std::vector<cd> va, vb;
    a.export_state(ka, va);
    b.export_state(kb, vb);
    *same_keys = ka == kb;
    if (!*same_keys) return 1e300;
    double d = 0;
    for (size_t j = 0; j < va.size(); ++j) d = std::max(d, std::abs(va[j] - vb[j]));
    return d;
}

// splitmix64: seeded inputs without <random>'s implementation-defined streams
static uint64_t sm_state = 0x1234abcdull;
static uint64_t sm_next() {
    uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main() {
    hash_table_parameters hp;   // the reference driver's defaults (main.cu:30-34)
    hp.N_devices = 1;
    hp.N_subtables = 32;
    hp.coef_gpu_memory_reserve = 0.01;
    hp.log2_N_bucket_spots = 5;
    try {
        Circuit__ probe(4, hp);
    } catch (const std::string& e) {
        if (e.find("no HIP device") != std::string::npos) {
            std::printf("NO_GPU: %s\n", e.c_str());
            return 3;
        }
        std::printf("FAIL constructor: %s\n", e.c_str());
        return 1;
    }
    try {
        // 1. the reference driver's circuit: H on every qubit of |0...0>
        {
            const uint32_t nq = 16;
            Circuit__ oc(nq, hp);
            for (uint32_t q = 0; q < nq; ++q) oc.h(q);
            std::vector<uint64_t> k;
            std::vector<cd> v;
            oc.export_state(k, v);
            CHECK(k.size() == (1u << nq), "all-H: %zu entries", k.size());
            double d = 0;
            for (size_t j = 0; j < v.size(); ++j) d = std::max(d, std::abs(v[j] - cd(std::ldexp(1.0, -8), 0.0)));
            CHECK(d < 1e-12, "all-H amplitude error %.3e", d);
            CHECK(std::fabs(oc.norm2() - 1.0) < 1e-12, "all-H norm");
        }
        // 2. Bell pair and the MSB-first ket printout
        {
            Circuit__ oc(3, hp);
            oc.h(0);
            oc.cx(0, 1);
            std::vector<uint64_t> k;
            std::vector<cd> v;
            oc.export_state(k, v);
            CHECK(k.size() == 2 && k[0] == 0 && k[1] == 3, "bell keys");
            CHECK(std::abs(v[1] - cd(1 / std::sqrt(2.0), 0)) < 1e-15, "bell amplitude");
            CHECK(oc.ket(3) == "011", "ket string %s", oc.ket(3).c_str());
            oc.print_full_statevector();
            oc.print_raw_data();
        }
        // 3. Toffoli truth table
        for (uint64_t in = 0; in < 8; ++in) {
            Circuit__ oc(3, hp);
            oc.set_basis(in);
            oc.ccx(0, 1, 2);
            std::vector<uint64_t> k;
            std::vector<cd> v;
            oc.export_state(k, v);
            const uint64_t want = ((in & 3) == 3) ? in ^ 4 : in;
            CHECK(k.size() == 1 && k[0] == want, "ccx |%llu> -> %llu", (unsigned long long)in,
                  (unsigned long long)(k.empty() ? ~0ull : k[0]));
        }
        // 4. identities on a random 10-qubit state: Rx = H Rz H, Rz = e^{-i t/2} Phase
        {
            const uint32_t nq = 10;
            std::vector<uint64_t> keys;
            std::vector<cd> amps;
            double n2 = 0;
            for (uint64_t key = 0; key < (1u << nq); key += 1 + sm_next() % 3) {
                const cd a((double)(sm_next() >> 11) / 9007199254740992.0 - 0.5,
                           (double)(sm_next() >> 11) / 9007199254740992.0 - 0.5);
                keys.push_back(key);
                amps.push_back(a);
                n2 += std::norm(a);
            }
            for (auto& a : amps) a /= std::sqrt(n2);
            Circuit__ a(nq, hp), b(nq, hp), c(nq, hp);
            a.load_state(keys, amps);
            b.load_state(keys, amps);
            c.load_state(keys, amps);
            const double th = 0.7312;
            a.rx(3, th);
            b.h(3);
            b.rz(3, th);
            b.h(3);
            bool same = false;
            const double d = max_diff(a, b, &same);
            CHECK(same && d < 1e-12, "Rx = H Rz H: keys %d diff %.3e", (int)same, d);
            a.rz(5, th);
            c.rx(3, th);
            c.phase(5, th);
            std::vector<uint64_t> ka, kc;
            std::vector<cd> va, vc;
            a.export_state(ka, va);
            c.export_state(kc, vc);
            double e = ka == kc ? 0.0 : 1e300;
            const cd g = std::exp(cd(0, -th / 2));
            for (size_t j = 0; ka == kc && j < va.size(); ++j) e = std::max(e, std::abs(va[j] - g * vc[j]));
            CHECK(e < 1e-12, "Rz = e^{-it/2} Phase: %.3e", e);
            CHECK(std::fabs(a.norm2() - 1.0) < 1e-12, "norm after rotations");
        }
        // 5. a controlled rotation on a state that has its control off is a no-op
        {
            Circuit__ oc(4, hp);
            oc.cry(2, 0, 1.1);
            CHECK(oc.nnz() == 1, "cry with control |0>");
            oc.x(2);
            oc.cry(2, 0, 1.1);
            CHECK(oc.nnz() == 2, "cry with control |1>");
        }
        // 6. errors come back as std::string
        {
            Circuit__ oc(4, hp);
            bool thrown = false;
            try {
                oc.h(4);
            } catch (const std::string&) {
                thrown = true;
            }
            CHECK(thrown, "target out of range must throw");
        }
        // 7. the pattern view of get_state_full: H on 3 qubits, qubit 2 = 1,
        //    qubit 1 free, qubit 0 = 0 -> |100> and |110>, amplitude 2^-3/2
        {
            Circuit__ oc(3, hp);
            oc.h(0);
            oc.h(1);
            oc.h(2);
            std::string str;
            std::map<std::vector<short>, cd> data;
            oc.get_state_full({1, 2}, str, data, {1, -1, 0}, 3);
            CHECK(data.size() == 2, "pattern view: %zu entries", data.size());
            const std::string want = "   3.536e-01   0.000e+00j   |1>|00>\n   3.536e-01   0.000e+00j   |1>|10>\n";
            CHECK(str == want, "pattern view string:\n%s", str.c_str());
            if (data.size() == 2) {
                auto it = data.begin();
                CHECK((it->first == std::vector<short>{1, 0, 0}), "first pattern entry");
                CHECK(std::abs(it->second - cd(std::pow(2.0, -1.5), 0.0)) < 1e-15, "pattern amplitude");
            }
            std::string all;
            oc.get_state_full({}, all, data, {}, 2);
            CHECK(data.size() == 8, "no pattern: %zu entries", data.size());
        }
        // 8. restart files: the state and the gate count come back exactly
        {
            const std::string path = "/tmp/circuit_check_restart_" + std::to_string((long)getpid()) + ".bin";
            Circuit__ a(6, hp);
            a.h(0);
            a.ry(3, 0.7);
            a.cx(0, 5);
            a.rz(5, 1.3);
            a.save_restart(path);
            Circuit__ b(6, hp);
            b.load_restart(path);
            bool same = false;
            double d = max_diff(a, b, &same);
            CHECK(same && d == 0.0, "restart: same keys %d, max diff %.3e", (int)same, d);
            a.h(3);
            b.h(3);
            d = max_diff(a, b, &same);
            CHECK(same && d == 0.0, "after restart and one more gate: max diff %.3e", d);
            Circuit__ c(5, hp);
            bool thrown = false;
            try {
                c.load_restart(path);
            } catch (const std::string&) {
                thrown = true;
            }
            CHECK(thrown, "a restart file of 6 qubits must not load into 5");
            std::remove(path.c_str());
        }
        // 9. unitary(): the matrices of ry / rx / h / cry given as 2x2 unitaries
        //    produce the named gates' states (random 8-qubit state)
        {
            const uint32_t nq = 8;
            std::vector<uint64_t> keys;
            std::vector<cd> amps;
            double n2 = 0;
            for (uint64_t k = 0; k < (1u << nq); k += 1 + (sm_next() % 5)) {
                const cd a((double)(sm_next() >> 11) / 9007199254740992.0 - 0.5,
                           (double)(sm_next() >> 11) / 9007199254740992.0 - 0.5);
                keys.push_back(k);
                amps.push_back(a);
                n2 += std::norm(a);
            }
            for (auto& a : amps) a /= std::sqrt(n2);
            Circuit__ a(nq, hp), b(nq, hp);
            a.load_state(keys, amps);
            b.load_state(keys, amps);
            const double th = 0.9, c = std::cos(th / 2), sn = std::sin(th / 2), f = 1.0 / std::sqrt(2.0);
            const cd ry_m[4] = {c, -sn, sn, c};
            const cd rx_m[4] = {c, cd(0, -sn), cd(0, -sn), c};
            const cd h_m[4] = {f, f, f, -f};
            a.ry(2, th);
            b.unitary(2, ry_m);
            a.rx(5, th);
            b.unitary(5, rx_m);
            a.h(7);
            b.unitary(7, h_m);
            a.cry(1, 6, th);
            b.unitary(6, ry_m, {1});
            b.unitary(3, ry_m, {0}, 0);      // control value 0: the gate acts where qubit 0 is 0
            a.x(0);                          // the same through X conjugation of the control
            a.cry(0, 3, th);
            a.x(0);
            bool same = false;
            const double d = max_diff(a, b, &same);
            CHECK(same && d < 1e-12, "unitary(): same keys %d, max diff %.3e", (int)same, d);
        }
    } catch (const std::string& e) {
        std::printf("FAIL exception: %s\n", e.c_str());
        return 1;
    }
    if (failures) {
        std::printf("%d checks failed\n", failures);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
