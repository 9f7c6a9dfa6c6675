// sparqc_circuit.hpp -- the reference's circuit API on top of libsqe.so.
//
// Same class name, constructor and gate methods as the reference's Circuit__
// (ivanNovikau/SparQC emulator/include/circuit.h:10-83), so a driver written
// against it (emulator/src/main.cu:87-148) compiles unchanged against this
// header; the state lives in the MI355X engine behind the C ABI of sqe.h.
//
// Conventions kept from the reference:
//   - qubit t is bit t of the basis key, |0...0> is key 0, the constructor
//     prepares |0...0> (circuit.h:34-43, device_table_FC.cuh:39-49);
//   - phase(t, theta) is diag(1, e^{i theta}); rz / rx / ry take the full angle
//     and the gate uses theta / 2 (table_manager.cuh:333-336; QLib.h:57-157);
//   - errors are thrown as std::string (main.cu:44-76, table_manager.cuh:175).
// Added: controlled gates (cx, cz, ccx and controlled forms of every gate),
// load_state, norm2, nnz, sorted export, the unordered_map readout of
// form_unordered_statevector_DeviceToHost (circuit.h:111-126), the pattern
// view get_state_full (circuit.h:205-257) and restart files.
//
// Gates are queued by the engine and run fused on the GPU; readouts wait.
#ifndef SPARQC_CIRCUIT_HPP
#define SPARQC_CIRCUIT_HPP

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <initializer_list>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "sqe.h"

// The reference's table parameters (emulator/include/sim_parameters_types.h:133-145).
// N_devices must be 1 (one engine per GPU); coef_gpu_memory_reserve is the
// fraction of free device memory left unused (the rest is the engine's HBM
// budget); N_subtables and log2_N_bucket_spots describe the reference's
// cuckoo-table design and are accepted but not used (the engine chooses its
// bucket qubits from the gate list).
struct hash_table_parameters {
    uint32_t N_devices = 1;
    double coef_gpu_memory_reserve = 0.01;
    uint32_t N_subtables = 32;
    uint32_t log2_N_bucket_spots = 5;
};

class Circuit__ {
public:
    using amp_t = std::complex<double>;

    Circuit__(uint32_t nq, const hash_table_parameters& hp) : Circuit__(nq, hp, nullptr) {}

    // cfg (optional) overrides the engine configuration (precision, prune
    // threshold, renormalisation period, fuse window, ...); cfg->nq is ignored.
    Circuit__(uint32_t nq, const hash_table_parameters& hp, const sqe_config* cfg) : nq_(nq) {
        if (hp.N_devices != 1) throw std::string("Circuit__: this engine runs one circuit on one GPU (N_devices must be 1)");
        if (hp.coef_gpu_memory_reserve < 0 || hp.coef_gpu_memory_reserve >= 1)
            throw std::string("Circuit__: coef_gpu_memory_reserve must be in [0, 1)");
        sqe_config c;
        if (cfg) c = *cfg;
        else sqe_config_default(&c, nq);
        c.nq = nq;
        if (c.hbm_budget_bytes == 0) c.mem_reserve_ppm = (uint32_t)std::lround(hp.coef_gpu_memory_reserve * 1e6);
        if (c.hbm_budget_bytes == 0 && c.mem_reserve_ppm == 0) c.mem_reserve_ppm = 1;
        check(sqe_create(&c, &e_));
        check(sqe_set_basis(e_, 0));
    }
    Circuit__(const Circuit__&) = delete;
    Circuit__& operator=(const Circuit__&) = delete;
    ~Circuit__() {
        if (e_) sqe_destroy(e_);
    }

    uint32_t nq() const { return nq_; }
    sqe_engine* engine() { return e_; }

    // ---- single-qubit gates (circuit.h:45-83) ----
    void x(uint32_t t) { gate(SQE_GATE_X, t, {}); }
    void y(uint32_t t) { gate(SQE_GATE_Y, t, {}); }
    void z(uint32_t t) { gate(SQE_GATE_Z, t, {}); }
    void h(uint32_t t) { gate(SQE_GATE_H, t, {}); }
    void phase(uint32_t t, double theta) { gate(SQE_GATE_PHASE, t, {}, theta); }
    void rz(uint32_t t, double theta) { gate(SQE_GATE_RZ, t, {}, theta); }
    void rx(uint32_t t, double theta) { gate(SQE_GATE_RX, t, {}, theta); }
    void ry(uint32_t t, double theta) { gate(SQE_GATE_RY, t, {}, theta); }

    // ---- controlled gates (every control must be |1>) ----
    void cx(uint32_t c, uint32_t t) { gate(SQE_GATE_X, t, {c}); }
    void cy(uint32_t c, uint32_t t) { gate(SQE_GATE_Y, t, {c}); }
    void cz(uint32_t c, uint32_t t) { gate(SQE_GATE_Z, t, {c}); }
    void ch(uint32_t c, uint32_t t) { gate(SQE_GATE_H, t, {c}); }
    void ccx(uint32_t c1, uint32_t c2, uint32_t t) { gate(SQE_GATE_X, t, {c1, c2}); }
    void cphase(uint32_t c, uint32_t t, double theta) { gate(SQE_GATE_PHASE, t, {c}, theta); }
    void crz(uint32_t c, uint32_t t, double theta) { gate(SQE_GATE_RZ, t, {c}, theta); }
    void crx(uint32_t c, uint32_t t, double theta) { gate(SQE_GATE_RX, t, {c}, theta); }
    void cry(uint32_t c, uint32_t t, double theta) { gate(SQE_GATE_RY, t, {c}, theta); }

    // Any 2x2 unitary u (row-major) on t under controls: ctrl_val gives the
    // required value of each control bit (default: all ones).
    void unitary(uint32_t t, const amp_t u[4], const std::vector<uint32_t>& controls = {}, int64_t ctrl_val = -1) {
        sqe_gate g = make(SQE_GATE_U, t, controls, ctrl_val);
        for (int k = 0; k < 4; ++k) {
            g.p[2 * k] = u[k].real();
            g.p[2 * k + 1] = u[k].imag();
        }
        push(g);
    }

    // ---- state ----
    void set_basis(uint64_t key) { check(sqe_set_basis(e_, key)); }
    void load_state(const std::vector<uint64_t>& keys, const std::vector<amp_t>& amps) {
        if (keys.size() != amps.size()) throw std::string("load_state: keys and amplitudes differ in size");
        std::vector<double> re(amps.size()), im(amps.size());
        for (size_t k = 0; k < amps.size(); ++k) {
            re[k] = amps[k].real();
            im[k] = amps[k].imag();
        }
        check(sqe_load(e_, keys.data(), re.data(), im.data(), keys.size()));
    }
    uint64_t nnz() {
        uint64_t n = 0;
        check(sqe_nnz(e_, &n));
        return n;
    }
    double norm2() {
        double v = 0;
        check(sqe_norm2(e_, &v));
        return v;
    }
    void sync() { check(sqe_sync(e_)); }

    // Nonzero entries sorted by key (entries with |re| and |im| below the
    // prune threshold are left out, table.cuh:63-64); with a mask, only keys
    // with (key & mask) == value.
    void export_state(std::vector<uint64_t>& keys, std::vector<amp_t>& amps, uint64_t mask = 0, uint64_t value = 0) {
        uint64_t n = 0;
        const sqe_status s = sqe_export_where(e_, mask, value, nullptr, nullptr, nullptr, 0, &n);
        if (s != SQE_OK && s != SQE_EINVAL) check(s);
        std::vector<uint64_t> k(n);
        std::vector<double> re(n), im(n);
        check(sqe_export_where(e_, mask, value, k.data(), re.data(), im.data(), n, &n));
        std::vector<size_t> order(n);
        for (size_t j = 0; j < n; ++j) order[j] = j;
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return k[a] < k[b]; });
        keys.resize(n);
        amps.resize(n);
        for (size_t j = 0; j < n; ++j) {
            keys[j] = k[order[j]];
            amps[j] = amp_t(re[order[j]], im[order[j]]);
        }
    }

    // The reference's readout form (circuit.h:111-126): key -> amplitude.
    void get_statevector(std::unordered_map<uint64_t, amp_t>& vv) {
        std::vector<uint64_t> k;
        std::vector<amp_t> a;
        export_state(k, a);
        vv.clear();
        vv.reserve(k.size());
        for (size_t j = 0; j < k.size(); ++j) vv.emplace(k[j], a[j]);
    }

    // "key: amplitude" lines in key order (circuit.h:88-92, data_t.h:116-125).
    void print_raw_data(std::ostream& os = std::cout, int prec = 3) {
        std::vector<uint64_t> k;
        std::vector<amp_t> a;
        export_state(k, a);
        for (size_t j = 0; j < k.size(); ++j) os << k[j] << ": \t" << amp_line(a[j], prec) << "\n";
    }

    // One line per nonzero entry: the ket with qubit nq-1 first (most
    // significant bit on the left, QLib.cpp:15-28) and its amplitude.
    void print_full_statevector(std::ostream& os = std::cout, int prec = 3) {
        std::vector<uint64_t> k;
        std::vector<amp_t> a;
        export_state(k, a);
        for (size_t j = 0; j < k.size(); ++j) os << amp_line(a[j], prec) << "  |" << ket(k[j]) << ">\n";
    }

    // The reference's pattern view (circuit.h:205-257, QLib.cpp:300-346):
    // state_to_choose lists one value per qubit, qubit nq-1 first (0 or 1
    // fixes the qubit, -1 leaves it free; empty: no pattern); the matching
    // non-negligible entries are returned in `data` (bits qubit nq-1 first ->
    // amplitude, in key order) and printed into str_wv, one line per entry:
    // real and imaginary parts in scientific notation with ampl_prec digits
    // (width 9 + ampl_prec), "j", three spaces, then the ket, split into
    // groups of organize_state[i] qubits ("|..>|..>").  The pattern is applied
    // on the device (sqe_export_where).
    void get_state_full(const std::vector<uint32_t>& organize_state, std::string& str_wv,
                        std::map<std::vector<short>, amp_t>& data, const std::vector<short>& state_to_choose,
                        uint32_t ampl_prec) {
        if (state_to_choose.size() > nq_) throw std::string("get_state_full: pattern longer than the qubit count");
        uint64_t mask = 0, value = 0;
        for (size_t i = 0; i < state_to_choose.size(); ++i) {
            if (state_to_choose[i] < 0) continue;
            const uint32_t q = nq_ - 1 - (uint32_t)i;
            mask |= 1ull << q;
            if (state_to_choose[i]) value |= 1ull << q;
        }
        std::vector<uint64_t> k;
        std::vector<amp_t> a;
        export_state(k, a, mask, value);
        data.clear();
        for (size_t j = 0; j < k.size(); ++j) {
            std::vector<short> bits(nq_);
            for (uint32_t i = 0; i < nq_; ++i) bits[i] = (short)((k[j] >> (nq_ - 1 - i)) & 1ull);
            data[bits] = a[j];
        }
        std::ostringstream oss;
        oss.setf(std::ios::scientific);
        oss.precision(ampl_prec);
        str_wv.clear();
        const int w = 9 + (int)ampl_prec;
        for (const auto& kv : data) {
            std::string st = "|";
            uint32_t count_org = 0, prev = 0;
            for (uint32_t jj = 0; jj < nq_; ++jj) {
                st += std::to_string(kv.first[jj]);
                if (!organize_state.empty() && count_org < organize_state.size() &&
                    jj == prev + organize_state[count_org] - 1) {
                    prev += organize_state[count_org];
                    ++count_org;
                    if (count_org < organize_state.size()) st += ">|";
                }
            }
            st += ">";
            oss.str(std::string());
            oss << std::setw(w) << kv.second.real() << std::setw(w) << kv.second.imag() << "j";
            str_wv += oss.str() + "   " + st + "\n";
        }
    }

    // Measurement in the computational basis: `shots` keys drawn with
    // probability |a|^2 (sqe_sample, on the device).
    std::vector<uint64_t> sample(uint64_t shots, uint64_t seed = 0) {
        std::vector<uint64_t> k(shots);
        check(sqe_sample(e_, shots, seed, k.data()));
        return k;
    }

    // Restart file of the current state (sqe_checkpoint_save / _load).
    void save_restart(const std::string& path) { check(sqe_checkpoint_save(e_, path.c_str())); }
    void load_restart(const std::string& path) { check(sqe_checkpoint_load(e_, path.c_str())); }

    std::string ket(uint64_t key) const {
        std::string s(nq_, '0');
        for (uint32_t q = 0; q < nq_; ++q)
            if ((key >> q) & 1ull) s[nq_ - 1 - q] = '1';
        return s;
    }

    sqe_stats stats() {
        sqe_stats s;
        check(sqe_get_stats(e_, &s));
        return s;
    }

private:
    static void check(sqe_status s) {
        if (s != SQE_OK) throw std::string(sqe_last_error());
    }
    static std::string amp_line(const amp_t& v, int prec) {
        std::ostringstream o;
        o.setf(std::ios::scientific);
        o.precision(prec);
        o << v.real() << (v.imag() < 0 ? " - " : " + ") << std::fabs(v.imag()) << "j";
        return o.str();
    }
    sqe_gate make(uint32_t kind, uint32_t t, const std::vector<uint32_t>& controls, int64_t ctrl_val) const {
        if (t >= nq_) throw std::string("gate target out of range");
        sqe_gate g{};
        g.kind = kind;
        g.target = t;
        for (uint32_t c : controls) {
            if (c >= nq_ || c == t) throw std::string("bad control qubit");
            g.ctrl_mask |= 1ull << c;
        }
        g.ctrl_val = ctrl_val < 0 ? g.ctrl_mask : ((uint64_t)ctrl_val & g.ctrl_mask);
        return g;
    }
    void gate(uint32_t kind, uint32_t t, std::initializer_list<uint32_t> controls, double theta = 0.0) {
        sqe_gate g = make(kind, t, std::vector<uint32_t>(controls), -1);
        switch (kind) {
        case SQE_GATE_H: g.p[0] = 1. / std::sqrt(2.); break;
        case SQE_GATE_PHASE: g.p[0] = std::cos(theta); g.p[1] = std::sin(theta); break;
        case SQE_GATE_RZ:
        case SQE_GATE_RX:
        case SQE_GATE_RY: g.p[0] = std::cos(theta / 2.); g.p[1] = std::sin(theta / 2.); break;
        default: break;
        }
        push(g);
    }
    void push(const sqe_gate& g) { check(sqe_apply_gates(e_, &g, 1)); }

    uint32_t nq_;
    sqe_engine* e_ = nullptr;
};

#endif  // SPARQC_CIRCUIT_HPP
