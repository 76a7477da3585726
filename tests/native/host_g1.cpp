// Host build of csrc/host_g1.hpp, driven by tests/test_host_g1.py.  Values
// cross as hex: field elements and point coordinates in the device's
// Montgomery form (x 2^261 mod p) on input, canonical on output.
//   "mul a b" / "add a b" / "sub a b" / "inv a"     -> canonical result
//   "dbl X Y ZZ ZZZ" / "addp P(4) Q(4)"             -> canonical affine x y ("inf")
//   "dbln n X Y ZZ ZZZ"                             -> the same for 2^n P
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_g1.hpp"

using namespace dvgpu::hostg1;

static void parse(const char* hex, uint32_t (&w)[8]) {
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
static Fe rd(const char* hex) {
    uint32_t w[8];
    parse(hex, w);
    return from_device(w);
}
static void pr(const uint32_t (&w)[8]) {
    for (int i = 7; i >= 0; --i) printf("%08x", w[i]);
}
static void wr(const Fe& a) {
    uint32_t w[8];
    to_words(a, w);
    pr(w);
}
static Xyzz rdp(char** t) {
    uint32_t w[32];
    for (int k = 0; k < 4; ++k) {
        uint32_t q[8];
        parse(t[k], q);
        memcpy(w + 8 * k, q, 32);
    }
    return from_device_xyzz(w);
}
static void wra(const Xyzz& p) {
    if (is_inf(p)) {
        printf("inf");
        return;
    }
    uint32_t x[8], y[8];
    to_affine_words(p, x, y);
    pr(x);
    printf(" ");
    pr(y);
}

int main() {
    char line[4096];
    while (fgets(line, sizeof line, stdin)) {
        std::vector<char*> t;
        for (char* s = strtok(line, " \n"); s; s = strtok(nullptr, " \n")) t.push_back(s);
        if (t.empty()) continue;
        const std::string op = t[0];
        if (op == "mul") wr(mul(rd(t[1]), rd(t[2])));
        else if (op == "add") wr(add(rd(t[1]), rd(t[2])));
        else if (op == "sub") wr(sub(rd(t[1]), rd(t[2])));
        else if (op == "inv") wr(inv(rd(t[1])));
        else if (op == "dbl") wra(dbl(rdp(&t[1])));
        else if (op == "addp") wra(add(rdp(&t[1]), rdp(&t[5])));
        else if (op == "dbln") wra(dbl_n(rdp(&t[2]), atoi(t[1])));
        printf("\n");
        fflush(stdout);
    }
    return 0;
}
