// Host build of the NTT pass code (csrc/ntt_tile.hpp), driven by
// tests/test_ntt_host.py.  Every block of every pass runs in order with one
// emulated thread, over the same twiddle tables, pass geometry and block
// order as on the device.
//
// stdin: "L rows inverse target radix" then the 2^20-th root and rows * 2^L
// inputs, canonical hex; radix 0 = as the device picks, 2 or 4 = forced.
// stdout: rows * 2^L outputs, canonical hex, then one line per pass:
// "pass p mode S logU tiles blocks radix".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ntt_tile.hpp"

using namespace bn254;
using namespace dvgpu::ntt;

static void parse_words(const char* hex, uint32_t* w) {
    char buf[65];
    size_t n = strlen(hex);
    memset(buf, '0', 64);
    buf[64] = 0;
    memcpy(buf + 64 - n, hex, n);
    for (int i = 0; i < 8; ++i) {
        char chunk[9];
        memcpy(chunk, buf + 56 - 8 * i, 8);
        chunk[8] = 0;
        w[i] = uint32_t(strtoul(chunk, nullptr, 16));
    }
}

// omega^k, k < 2^19, as canonical Montgomery words (the device's power table).
static std::vector<uint32_t> power_table(const Fr& w) {
    std::vector<uint32_t> t(size_t(1) << 19 << 3);
    Fr acc = fp_one<FrParams>();
    for (size_t k = 0; k < (size_t(1) << 19); ++k) {
        pack(canon(acc), *reinterpret_cast<uint32_t(*)[8]>(&t[k * 8]));
        acc = mul(acc, w);
    }
    return t;
}

int main() {
    int L, inverse, radix;
    unsigned long long rows, target;
    if (scanf("%d %llu %d %llu %d", &L, &rows, &inverse, &target, &radix) != 5) return 2;
    char hex[128];
    uint32_t rw[8];
    if (scanf("%127s", hex) != 1) return 2;
    parse_words(hex, rw);
    const size_t n = size_t(1) << L, total = n * rows;
    std::vector<uint32_t> in(total * 8), out(total * 8, 0xdeadbeefu), work(total * 8, 0xdeadbeefu);
    for (size_t i = 0; i < total; ++i) {
        if (scanf("%127s", hex) != 1) return 2;
        parse_words(hex, &in[i * 8]);
    }
    Fr w = to_mont<FrParams>(rw);
    if (inverse) w = inv(w);
    const std::vector<uint32_t> pw = power_table(w);
    const uint32_t stage_total = (1u << STAGE_TABLE_LOG) - 1;
    std::vector<uint32_t> stage(size_t(stage_total) * LIMBS);
    for (uint32_t e = 0; e < stage_total; ++e) stage_table_entry(pw.data(), stage.data(), e);
    std::vector<uint32_t> pass1;
    if (L > SINGLE_MAX_LOG) {
        const uint32_t entries = pass1_table_entries(L);
        pass1.assign(size_t(entries) * LIMBS, 0);
        for (uint32_t e = 0; e < entries; ++e)
            pass1_table_entry(pw.data(), pass1.data(), uint32_t(L), uint32_t(L / 2), pass1_log_ct(L), e);
    }
    Fr scale = fp_one<FrParams>();
    if (inverse && L) {
        uint32_t nw[8] = {0};
        nw[L / 32] = 1u << (L % 32);
        scale = canon(inv(to_mont<FrParams>(nw)));
    }
    const LengthPlan lp = plan_length(in.data(), out.data(), work.data(), L, rows, stage.data(),
                                      pass1.empty() ? nullptr : pass1.data(), inverse && L ? &scale : nullptr,
                                      uint32_t(target));
    std::vector<uint32_t> lds(size_t(LIMBS) * TILE);
    std::vector<int> radices;
    for (int p = 0; p < lp.passes; ++p) {
        const bool r4 = radix ? radix == 4 : use_radix4(lp.blocks[p], lp.tile[p]);
        radices.push_back(r4 ? 4 : 2);
        for (uint32_t b = 0; b < lp.blocks[p]; ++b) {
            std::fill(lds.begin(), lds.end(), 0xfeedfaceu);
            const uint32_t lb = xcd_order(b, lp.blocks[p]);
            if (r4) ntt_tile<true>(lds.data(), lp.pass[p], lb);
            else ntt_tile<false>(lds.data(), lp.pass[p], lb);
        }
    }
    for (size_t i = 0; i < total; ++i) {
        for (int k = 7; k >= 0; --k) printf("%08x", out[i * 8 + k]);
        printf("\n");
    }
    for (int p = 0; p < lp.passes; ++p)
        printf("pass %d %u %u %u %u %u %d\n", p, lp.pass[p].mode, lp.pass[p].S, lp.pass[p].logU, lp.pass[p].tiles,
               lp.blocks[p], radices[p]);
    return 0;
}
