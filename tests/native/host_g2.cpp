// Host build of csrc/host_g2.hpp, driven by tests/test_host_g2.py.  Values
// cross as hex: Fq2 elements as two Fq (c0 c1) in the device's Montgomery
// form (x 2^261 mod p) on input, canonical on output.
//   "mul a0 a1 b0 b1" / "inv a0 a1"             -> canonical c0 c1
//   "dbl P(8)" / "addp P(8) Q(8)"               -> canonical affine x0 x1 y0 y1 ("inf")
//   "dbln n P(8)"                               -> the same for 2^n P
// A point is X.c0 X.c1 Y.c0 Y.c1 ZZ.c0 ZZ.c1 ZZZ.c0 ZZZ.c1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_g2.hpp"

using namespace dvgpu::hostg2;

static void parse(const char* hex, uint32_t* w) {
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
static F2 rd2(char** t) {
    uint32_t w[16];
    parse(t[0], w);
    parse(t[1], w + 8);
    return from_device(w);
}
static void pr(const uint32_t (&w)[8]) {
    for (int i = 7; i >= 0; --i) printf("%08x", w[i]);
}
static void wr2(const F2& a) {
    uint32_t w[8];
    dvgpu::hostg1::to_words(a.c0, w);
    pr(w);
    printf(" ");
    dvgpu::hostg1::to_words(a.c1, w);
    pr(w);
}
static Xyzz rdp(char** t) {
    uint32_t w[64];
    for (int k = 0; k < 8; ++k) parse(t[k], w + 8 * k);
    return from_device_xyzz(w);
}
static void wra(const Xyzz& p) {
    if (is_inf(p)) {
        printf("inf");
        return;
    }
    uint32_t w[4][8];
    to_affine_words(p, w);
    for (int k = 0; k < 4; ++k) {
        if (k) printf(" ");
        pr(w[k]);
    }
}

int main() {
    char line[8192];
    while (fgets(line, sizeof line, stdin)) {
        std::vector<char*> t;
        for (char* s = strtok(line, " \n"); s; s = strtok(nullptr, " \n")) t.push_back(s);
        if (t.empty()) continue;
        const std::string op = t[0];
        if (op == "mul") wr2(mul(rd2(&t[1]), rd2(&t[3])));
        else if (op == "inv") wr2(inv(rd2(&t[1])));
        else if (op == "dbl") wra(dbl(rdp(&t[1])));
        else if (op == "addp") wra(add(rdp(&t[1]), rdp(&t[9])));
        else if (op == "dbln") wra(dbl_n(rdp(&t[2]), atoi(t[1])));
        printf("\n");
        fflush(stdout);
    }
    return 0;
}
