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
        BlockMat m(256, cx(0.0, 0.0));
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
    // one 32x32 block (5 bucket qubits): MFMA operand tables of two 32-row
    // tiles, 32 steps each, then the FP64 matrix after 2048 doubles
    const size_t b32 = ops.size();
    const uint64_t c32 = coef.size() / 8;
    {
        BlockMat m(1024, cx(0.0, 0.0));
        for (int r = 0; r < 32; ++r) m[r * 32 + r] = 1.0;
        m[0] = cx(7.0, 0.25);
        m[1 * 32 + 5] = cx(-2.0, 3.0);
        DOp d{};
        d.kind = K_BLOCK;
        d.tslot = -1;
        d.p[1] = 9;
        ops.push_back(encode_block(d, m, coef));
    }
    {
        const MOp& m = ops[b32];
        CHECK((m.hdr & 15u) == H_SLOT_BLOCK && coef_index_of(m.c[4]) == c32 && m.c[0] == 9.0);
        CHECK(coef.size() >= (c32 + 512) * 8);
        const float* tab = (const float*)&coef[c32 * 8];
        // tile q, step t, lane l holds R[32 q + l % 32][2t + l / 32], R = [[Mr, -Mi], [Mi, Mr]]
        CHECK(tab[(0 * 32 + 0) * 64 + 0] == 7.0f);           // R[0][0] = Mr[0][0]
        CHECK(tab[(1 * 32 + 0) * 64 + 0] == 0.25f);          // R[32][0] = Mi[0][0]
        CHECK(tab[(0 * 32 + 16) * 64 + 0] == -0.25f);        // R[0][32] = -Mi[0][0]
        CHECK(tab[(0 * 32 + 2) * 64 + 33] == -2.0f);         // R[1][5] = Mr[1][5] (step 2, second row)
        CHECK(tab[(1 * 32 + 18) * 64 + 33] == -2.0f);        // R[33][37] = Mr[1][5]
        CHECK(tab[(1 * 32 + 2) * 64 + 33] == 3.0f);          // R[33][5] = Mi[1][5]
        CHECK(coef[c32 * 8 + 2048] == 7.0 && coef[c32 * 8 + 2049] == 0.25);
        CHECK(coef[c32 * 8 + 2048 + 2 * (1 * 32 + 5)] == -2.0);
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
        const MOp& m = ops[b32 - 1];                  // the RX record
        const uint64_t ci = coef_index_of(m.c[4]);
        CHECK((m.hdr & 15u) == H_SLOT_U);
        CHECK(coef[ci * 8] == 0.25 && coef[ci * 8 + 3] == -0.75 && coef[ci * 8 + 6] == 0.25);
    }
    const uint64_t last = coef_index_of(ops[b32 - 1].c[4]);
    std::printf("encode_check: %llu blocks, %llu U records, last side-table index %llu (%s 2^19)\n",
                (unsigned long long)n_blocks, (unsigned long long)n_u, (unsigned long long)last,
                last >= (1ull << 19) ? ">=" : "<");
    if (fails) return 1;
    std::printf("encode_check ok\n");
    return 0;
}
