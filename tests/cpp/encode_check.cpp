// Host-only check of the sweep-program encoder (sparqc_amd/csrc/sqe_plan.hpp):
// the coefficient side-table index of general-U records and fused blocks must
// address the right entry however long a program is.  Built and run by
// tests/test_host.py (hipcc, host pass only; no GPU needed).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sqe_plan.hpp"

using namespace sqe;

static int fails = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "FAIL line %d: %s\n", __LINE__, #c); \
            if (++fails > 20) std::exit(1);                         \
        }                                                           \
    } while (0)

int main(int argc, char** argv) {
    const uint64_t n_blocks = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 4200;
    const uint64_t n_u = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : (1ull << 19) + 300;
    QMap qm = make_qmap({3, 7, 11, 20});
    std::vector<double> coef;
    std::vector<MOp> ops;
    // blocks: each matrix carries its own number in element (0, 0)
    for (uint64_t b = 0; b < n_blocks; ++b) {
        std::array<cx, 256> m{};
        for (int r = 0; r < 16; ++r) m[r * 16 + r] = 1.0;
        m[0] = cx((double)b, 0.5);
        DOp d{};
        d.kind = K_BLOCK;
        d.tslot = -1;
        d.p[1] = 6;
        ops.push_back(encode_block(d, m, coef));
    }
    // general U records: coefficient p[0] carries the record's number
    for (uint64_t u = 0; u < n_u; ++u) {
        sqe_gate g{};
        g.kind = SQE_GATE_U;
        g.target = 7;
        g.p[0] = (double)u;
        g.p[7] = -1.0;
        ops.push_back(encode_op(plan_op(g, qm), coef, false));
    }
    // an RX record after them
    {
        sqe_gate g{};
        g.kind = SQE_GATE_RX;
        g.target = 11;
        g.p[0] = 0.25;
        g.p[1] = 0.75;
        ops.push_back(encode_op(plan_op(g, qm), coef, true));
    }
    for (uint64_t b = 0; b < n_blocks; ++b) {
        const MOp& m = ops[b];
        CHECK((m.hdr & 15u) == H_SLOT_BLOCK);
        CHECK((m.hdr >> 13) == 0);
        const uint64_t ci = coef_index_of(m.c[4]);
        CHECK(ci * 8 + 1024 <= coef.size());
        // FP64 copy of the matrix follows the 512-double MFMA operand table
        CHECK(coef[ci * 8 + 512] == (double)b);
        CHECK(coef[ci * 8 + 513] == 0.5);
        CHECK(m.c[0] == 6.0);
    }
    for (uint64_t u = 0; u < n_u; ++u) {
        const MOp& m = ops[n_blocks + u];
        CHECK((m.hdr & 15u) == H_SLOT_U);
        CHECK(((m.hdr >> 4) & 7u) == 1u);       // qubit 7 is bucket slot bit 1
        const uint64_t ci = coef_index_of(m.c[4]);
        CHECK(coef[ci * 8] == (double)u);
        CHECK(coef[ci * 8 + 7] == -1.0);
    }
    {
        const MOp& m = ops.back();
        const uint64_t ci = coef_index_of(m.c[4]);
        CHECK((m.hdr & 15u) == H_SLOT_U);
        CHECK(coef[ci * 8] == 0.25 && coef[ci * 8 + 3] == -0.75 && coef[ci * 8 + 6] == 0.25);
    }
    const uint64_t last = coef_index_of(ops.back().c[4]);
    std::printf("encode_check: %llu blocks, %llu U records, last side-table index %llu (%s 2^19)\n",
                (unsigned long long)n_blocks, (unsigned long long)n_u, (unsigned long long)last,
                last >= (1ull << 19) ? ">=" : "<");
    if (fails) return 1;
    std::printf("encode_check ok\n");
    return 0;
}
