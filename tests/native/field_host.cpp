// Host build of the device field/curve headers, driven over stdin/stdout by
// tests/test_field_host.py: every line is "<op> <hex args...>", every output
// line the hex results.  Values cross as canonical 256-bit hex integers.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "curve.hpp"
using namespace bn254;

static void parse_words(const char* hex, uint32_t (&w)[8]) {
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
static void print_words(const uint32_t (&w)[8]) {
    for (int i = 7; i >= 0; --i) printf("%08x", w[i]);
}
template <class C>
static Fp<C> rd(const char* hex) {
    uint32_t w[8];
    parse_words(hex, w);
    return to_mont<C>(w);
}
template <class C>
static void wr(const Fp<C>& x) {
    uint32_t w[8];
    from_mont(x, w);
    print_words(w);
}
static Fq2 rd2(char** t) { return {rd<FqParams>(t[0]), rd<FqParams>(t[1])}; }
static void wr2(const Fq2& a) { wr(a.c0); printf(" "); wr(a.c1); }
template <class F>
static void wr_aff(const Affine<F>& a);
template <>
void wr_aff<FqOps>(const Affine<FqOps>& a) {
    if (a.inf) { printf("inf"); return; }
    wr(a.x); printf(" "); wr(a.y);
}
template <>
void wr_aff<Fq2Ops>(const Affine<Fq2Ops>& a) {
    if (a.inf) { printf("inf"); return; }
    wr2(a.x); printf(" "); wr2(a.y);
}

int main() {
    char line[8192];
    while (fgets(line, sizeof line, stdin)) {
        std::vector<char*> t;
        for (char* s = strtok(line, " \n"); s; s = strtok(nullptr, " \n")) t.push_back(s);
        if (t.empty()) continue;
        const std::string op = t[0];
        char** a = t.data() + 1;
        if (op == "qmul") wr(mul(rd<FqParams>(a[0]), rd<FqParams>(a[1])));
        else if (op == "qsqr") wr(sqr(rd<FqParams>(a[0])));
        else if (op == "qadd") wr(add(rd<FqParams>(a[0]), rd<FqParams>(a[1])));
        else if (op == "qsub") wr(sub(rd<FqParams>(a[0]), rd<FqParams>(a[1])));
        else if (op == "qneg") wr(neg(rd<FqParams>(a[0])));
        else if (op == "qinv") wr(inv(rd<FqParams>(a[0])));
        else if (op == "rmul") wr(mul(rd<FrParams>(a[0]), rd<FrParams>(a[1])));
        else if (op == "rinv") wr(inv(rd<FrParams>(a[0])));
        else if (op == "q2mul") wr2(f2_mul(rd2(a), rd2(a + 2)));
        else if (op == "q2sqr") wr2(f2_sqr(rd2(a)));
        else if (op == "q2inv") wr2(f2_inv(rd2(a)));
        else if (op == "qiszero") printf("%d", int(is_zero(rd<FqParams>(a[0]))));
        else if (op == "g1add" || op == "g1madd" || op == "g1dbl") {
            // operands: affine points "x y" or "inf"
            auto rp = [](char**& p) -> Affine<FqOps> {
                if (!strcmp(p[0], "inf")) { ++p; return {fp_zero<FqParams>(), fp_zero<FqParams>(), true}; }
                Affine<FqOps> r{rd<FqParams>(p[0]), rd<FqParams>(p[1]), false};
                p += 2;
                return r;
            };
            char** p = a;
            const auto A = rp(p);
            if (op == "g1dbl") wr_aff(xyzz_to_affine(xyzz_dbl(xyzz_from_affine(A))));
            else {
                const auto B = rp(p);
                // Route through a non-trivial Z by doubling then re-adding.
                if (op == "g1add") wr_aff(xyzz_to_affine(xyzz_add(xyzz_from_affine(A), xyzz_from_affine(B))));
                else wr_aff(xyzz_to_affine(xyzz_madd(xyzz_from_affine(A), B)));
            }
        } else if (op == "g1chain") {
            // k_accumulate's loop: from infinity, add +-P_i (tokens "x y s",
            // s = 1 negates) with the signed madd; then the generic path.
            Xyzz<FqOps> fast = xyzz_inf<FqOps>(), plain = xyzz_inf<FqOps>();
            for (size_t i = 1; i + 2 < t.size(); i += 3) {
                const Fq x = rd<FqParams>(t[i]), y = rd<FqParams>(t[i + 1]);
                const bool s = t[i + 2][0] == '1';
                fast = xyzz_madd_signed<FqOps>(fast, x, y, s);
                plain = xyzz_madd<FqOps>(plain, x, s ? neg(y) : y);
            }
            wr_aff(xyzz_to_affine(fast));
            printf(" ");
            wr_aff(xyzz_to_affine(plain));
        } else if (op == "g1combo") {
            // u = A + B (signed madds from infinity, so ZZ != 1), then 2u, 3u,
            // 3u + 3u, 3u - 3u, 2u + 4u (equal points in different
            // coordinates) and u + B, each printed as an affine point.
            const Fq ax = rd<FqParams>(a[0]), ay = rd<FqParams>(a[1]);
            const Fq bx = rd<FqParams>(a[2]), by = rd<FqParams>(a[3]);
            const Xyzz<FqOps> u = xyzz_madd_signed<FqOps>(xyzz_madd_signed<FqOps>(xyzz_inf<FqOps>(), ax, ay, false),
                                                          bx, by, false);
            const Xyzz<FqOps> v = xyzz_dbl<FqOps>(u);
            const Xyzz<FqOps> w = xyzz_add<FqOps>(u, v);
            const Xyzz<FqOps> outs[6] = {
                v, w, xyzz_add<FqOps>(w, w), xyzz_add<FqOps>(w, xyzz_neg<FqOps>(w)),
                xyzz_add<FqOps>(v, xyzz_dbl<FqOps>(xyzz_dbl<FqOps>(u))),
                xyzz_add<FqOps>(u, xyzz_from_affine<FqOps>(Affine<FqOps>{bx, by, false}))};
            for (int k = 0; k < 6; ++k) {
                if (k) printf(" | ");
                wr_aff(xyzz_to_affine(outs[k]));
            }
        } else if (op == "mul2raw" || op == "sub3raw") {
            // 36 limbs (hex, little-endian limb order) of four raw operands;
            // prints the nine output limbs.
            Fq v[4];
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < LIMBS; ++k) v[j].l[k] = uint32_t(strtoul(a[9 * j + k], nullptr, 16));
            const Fq r = op == "mul2raw" ? mul2(v[0], v[1], v[2], v[3]) : sub3_n(v[0], v[1], v[2], v[3]);
            for (int k = 0; k < LIMBS; ++k) printf("%s%x", k ? " " : "", r.l[k]);
        } else if (op == "g1mul" || op == "g2mul") {
            // point then scalar (hex); result via double-and-add on XYZZ.
            uint32_t k[8];
            if (op == "g1mul") {
                Affine<FqOps> P{rd<FqParams>(a[0]), rd<FqParams>(a[1]), false};
                parse_words(a[2], k);
                wr_aff(xyzz_to_affine(xyzz_mul_words<FqOps, 8>(xyzz_from_affine(P), k, 256)));
            } else {
                Affine<Fq2Ops> P{rd2(a), rd2(a + 2), false};
                parse_words(a[4], k);
                wr_aff(xyzz_to_affine(xyzz_mul_words<Fq2Ops, 8>(xyzz_from_affine(P), k, 256)));
            }
        } else if (op == "g2psi") {
            Affine<Fq2Ops> P{rd2(a), rd2(a + 2), false};
            wr_aff(xyzz_to_affine(psi(xyzz_from_affine(P))));
        } else if (op == "g2oncurve") {
            Affine<Fq2Ops> P{rd2(a), rd2(a + 2), false};
            printf("%d", int(affine_on_curve(P)));
        } else { printf("?"); }
        printf("\n");
        fflush(stdout);
    }
    return 0;
}
