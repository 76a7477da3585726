// Host build of the device pairing header, driven over stdin/stdout by
// tests/test_pairing_host.py.  Input lines: "<op> <count> <hex...>" with
// canonical 256-bit hex integers; output: one line of hex Fq values.
//   miller  xp yp x2.c0 x2.c1 y2.c0 y2.c1   -> the Miller value (12 Fq, tower order)
//   product k  (xp yp x2.c0 x2.c1 y2.c0 y2.c1) x k -> reduced product (12 Fq)
//   fexp    12 Fq                          -> final exponentiation
//   mul     24 Fq / inv 12 Fq / frob 12 Fq -> tower arithmetic
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "pairing.hpp"
using namespace bn254;

static Fq rd(const char* hex) {
    uint32_t w[8];
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
    return to_mont<FqParams>(w);
}
static void wr(const Fq& x) {
    uint32_t w[8];
    from_mont(x, w);
    for (int i = 7; i >= 0; --i) printf("%08x", w[i]);
    printf(" ");
}
static Fq2 rd2(char** t) { return {rd(t[0]), rd(t[1])}; }
static void wr2(const Fq2& a) { wr(a.c0); wr(a.c1); }
static void wr12(const Fq12& a) {
    for (const Fq6* c : {&a.c0, &a.c1}) {
        wr2(c->a0);
        wr2(c->a1);
        wr2(c->a2);
    }
    printf("\n");
}
static Fq12 rd12(char** t) {
    return {{rd2(t), rd2(t + 2), rd2(t + 4)}, {rd2(t + 6), rd2(t + 8), rd2(t + 10)}};
}

int main() {
    static char line[1 << 16];
    while (fgets(line, sizeof line, stdin)) {
        std::vector<char*> t;
        for (char* s = strtok(line, " \n"); s; s = strtok(nullptr, " \n")) t.push_back(s);
        if (t.empty()) continue;
        const char* op = t[0];
        char** a = t.data() + 1;
        if (!strcmp(op, "miller")) {
            wr12(miller_loop(rd(a[0]), rd(a[1]), rd2(a + 2), rd2(a + 4)));
        } else if (!strcmp(op, "product")) {
            const int k = atoi(a[0]);
            std::vector<PairingInput> in;
            for (int i = 0; i < k; ++i) {
                char** p = a + 1 + 6 * i;
                in.push_back({rd(p[0]), rd(p[1]), rd2(p + 2), rd2(p + 4)});
            }
            wr12(final_exponentiation(miller_product(in.data(), k)));
        } else if (!strcmp(op, "prepared")) {  // pair 0 through prepare_lines: the Miller value
            const int k = atoi(a[0]);
            std::vector<PairingInput> in;
            for (int i = 0; i < k; ++i) {
                char** p = a + 1 + 6 * i;
                in.push_back({rd(p[0]), rd(p[1]), rd2(p + 2), rd2(p + 4)});
            }
            std::vector<Line> lines(PAIRING_LINES);
            prepare_lines(in[0].x2, in[0].y2, lines.data());
            auto src = [&](int i) { return lines[i]; };
            wr12(miller_product_impl(in.data() + 1, k - 1, src, true, in[0].xp, in[0].yp));
        } else if (!strcmp(op, "millerprod")) {  // the same product without prepared lines
            const int k = atoi(a[0]);
            std::vector<PairingInput> in;
            for (int i = 0; i < k; ++i) {
                char** p = a + 1 + 6 * i;
                in.push_back({rd(p[0]), rd(p[1]), rd2(p + 2), rd2(p + 4)});
            }
            wr12(miller_product(in.data(), k));
        } else if (!strcmp(op, "fexp")) {
            wr12(final_exponentiation(rd12(a)));
        } else if (!strcmp(op, "mul")) {
            wr12(f12_mul(rd12(a), rd12(a + 12)));
        } else if (!strcmp(op, "sqr")) {
            wr12(f12_sqr(rd12(a)));
        } else if (!strcmp(op, "inv")) {
            wr12(f12_inv(rd12(a)));
        } else if (!strcmp(op, "frob")) {
            wr12(f12_frobenius(rd12(a)));
        } else {
            printf("?\n");
        }
        fflush(stdout);
    }
    return 0;
}
