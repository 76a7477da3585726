/*
 * bn254_ref.c -- CPU oracle and C CPU baseline for the BN254 hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a checker and by bench.py's
 * cpu_baseline leg as the timed CPU baseline ("C CPU baseline, not
 * arkworks").  Nothing in the product library links or calls it.
 *
 * Restates, with 4 x 64-bit Montgomery arithmetic (R = 2^256) and pthreads:
 *   - G1 MSM sum s_i P_i (what crates/gpu/src/lib.rs:120-139 and arkworks
 *     VariableBaseMSM compute), by Pippenger with unsigned c-bit windows,
 *     points split across threads;
 *   - the natural-order radix-2 NTT / INTT over <w_n>, w_n = root^(2^20/n)
 *     with root the caller's 2^20-th root (crates/gpu/src/icicle.rs:429-456,
 *     arkworks Radix2EvaluationDomain), inverse scaled by 1/n;
 *   - fixed-base [s_i] G1 (crates/protocol/src/setup.rs:13-22);
 *   - the KZG opening cache (crates/protocol/src/kzg.rs:217-275): the same
 *     convolution and group DFTs, in Jacobian coordinates with 254-bit
 *     double-and-add for every twiddle.
 * Pinned by tests/test_oracle.py against the Python oracle and closed forms.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef struct { uint64_t v[4]; } fe;

typedef struct {
    uint64_t m[4];
    uint64_t inv;  /* -m^-1 mod 2^64 */
    fe r2;         /* R^2 mod m */
    fe one;        /* R mod m */
} field;

static field FQ, FR;

static void set_field(field* f, const uint64_t m[4]) {
    memcpy(f->m, m, 32);
    uint64_t x = 1;
    for (int i = 0; i < 6; ++i) x *= 2 - m[0] * x;
    f->inv = (uint64_t)0 - x;
    /* R mod m and R^2 mod m by repeated doubling of 1 */
    fe t = {{1, 0, 0, 0}};
    for (int i = 0; i < 512; ++i) {
        uint64_t carry = 0;
        for (int k = 0; k < 4; ++k) {
            uint64_t nv = (t.v[k] << 1) | carry;
            carry = t.v[k] >> 63;
            t.v[k] = nv;
        }
        /* subtract m if t >= m (t < 2m since m > 2^253) */
        int ge = carry != 0;
        if (!ge) {
            ge = 1;
            for (int k = 3; k >= 0; --k) {
                if (t.v[k] != m[k]) { ge = t.v[k] > m[k]; break; }
            }
        }
        if (ge) {
            u128 b = 0;
            for (int k = 0; k < 4; ++k) {
                u128 d = (u128)t.v[k] - m[k] - (uint64_t)b;
                t.v[k] = (uint64_t)d;
                b = (d >> 64) & 1;
            }
        }
        if (i == 255) f->one = t;
    }
    f->r2 = t;
}

static inline int geq(const uint64_t* a, const uint64_t* m) {
    for (int k = 3; k >= 0; --k)
        if (a[k] != m[k]) return a[k] > m[k];
    return 1;
}

static inline void fsub_m(uint64_t* a, const uint64_t* m) {
    u128 b = 0;
    for (int k = 0; k < 4; ++k) {
        u128 d = (u128)a[k] - m[k] - (uint64_t)b;
        a[k] = (uint64_t)d;
        b = (d >> 64) & 1;
    }
}

static inline fe fmul(const field* f, fe a, fe b) {
    uint64_t t[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) {
        u128 c = 0;
        for (int j = 0; j < 4; ++j) {
            c = (u128)a.v[j] * b.v[i] + t[j] + (uint64_t)c;
            t[j] = (uint64_t)c;
            c >>= 64;
        }
        u128 s = (u128)t[4] + (uint64_t)c;
        t[4] = (uint64_t)s;
        t[5] = (uint64_t)(s >> 64);
        uint64_t mq = t[0] * f->inv;
        c = ((u128)mq * f->m[0] + t[0]) >> 64;
        for (int j = 1; j < 4; ++j) {
            c = (u128)mq * f->m[j] + t[j] + (uint64_t)c;
            t[j - 1] = (uint64_t)c;
            c >>= 64;
        }
        s = (u128)t[4] + (uint64_t)c;
        t[3] = (uint64_t)s;
        t[4] = t[5] + (uint64_t)(s >> 64);
    }
    fe r = {{t[0], t[1], t[2], t[3]}};
    if (t[4] || geq(r.v, f->m)) fsub_m(r.v, f->m);
    return r;
}

static inline fe fadd(const field* f, fe a, fe b) {
    u128 c = 0;
    fe r;
    for (int k = 0; k < 4; ++k) {
        c = (u128)a.v[k] + b.v[k] + (uint64_t)c;
        r.v[k] = (uint64_t)c;
        c >>= 64;
    }
    if ((uint64_t)c || geq(r.v, f->m)) fsub_m(r.v, f->m);
    return r;
}

static inline fe fsub(const field* f, fe a, fe b) {
    u128 br = 0;
    fe r;
    for (int k = 0; k < 4; ++k) {
        u128 d = (u128)a.v[k] - b.v[k] - (uint64_t)br;
        r.v[k] = (uint64_t)d;
        br = (d >> 64) & 1;
    }
    if (br) {
        u128 c = 0;
        for (int k = 0; k < 4; ++k) {
            c = (u128)r.v[k] + f->m[k] + (uint64_t)c;
            r.v[k] = (uint64_t)c;
            c >>= 64;
        }
    }
    return r;
}

static inline int fzero(fe a) { return !(a.v[0] | a.v[1] | a.v[2] | a.v[3]); }
static inline int feq(fe a, fe b) { return !memcmp(&a, &b, 32); }

static fe fpow(const field* f, fe a, const uint64_t e[4]) {
    fe r = f->one;
    for (int w = 3; w >= 0; --w)
        for (int b = 63; b >= 0; --b) {
            r = fmul(f, r, r);
            if ((e[w] >> b) & 1) r = fmul(f, r, a);
        }
    return r;
}

static fe finv(const field* f, fe a) {
    uint64_t e[4];
    memcpy(e, f->m, 32);
    e[0] -= 2;
    return fpow(f, a, e);
}

static fe to_mont(const field* f, const uint8_t* b) {
    fe x;
    memcpy(x.v, b, 32);
    return fmul(f, x, f->r2);
}

static void from_mont(const field* f, fe x, uint8_t* b) {
    fe one = {{1, 0, 0, 0}};
    fe r = fmul(f, x, one);
    memcpy(b, r.v, 32);
}

/* ---- G1 Jacobian (y^2 = x^3 + 3) ------------------------------------- */
typedef struct { fe X, Y, Z; } jac;
typedef struct { fe x, y; int inf; } aff;

static const jac JINF = {{{0}}, {{0}}, {{0}}};

static jac jdbl(jac p) {
    if (fzero(p.Z)) return p;
    const field* f = &FQ;
    fe A = fmul(f, p.X, p.X), B = fmul(f, p.Y, p.Y), C = fmul(f, B, B);
    fe t = fadd(f, p.X, B);
    fe D = fsub(f, fsub(f, fmul(f, t, t), A), C);
    D = fadd(f, D, D);
    fe E = fadd(f, fadd(f, A, A), A);
    fe F = fmul(f, E, E);
    jac r;
    r.X = fsub(f, F, fadd(f, D, D));
    fe C8 = fadd(f, C, C);
    C8 = fadd(f, C8, C8);
    C8 = fadd(f, C8, C8);
    r.Y = fsub(f, fmul(f, E, fsub(f, D, r.X)), C8);
    r.Z = fmul(f, fadd(f, p.Y, p.Y), p.Z);
    return r;
}

static jac jadd(jac p, jac q) {
    const field* f = &FQ;
    if (fzero(p.Z)) return q;
    if (fzero(q.Z)) return p;
    fe Z1Z1 = fmul(f, p.Z, p.Z), Z2Z2 = fmul(f, q.Z, q.Z);
    fe U1 = fmul(f, p.X, Z2Z2), U2 = fmul(f, q.X, Z1Z1);
    fe S1 = fmul(f, fmul(f, p.Y, q.Z), Z2Z2), S2 = fmul(f, fmul(f, q.Y, p.Z), Z1Z1);
    if (feq(U1, U2)) {
        if (feq(S1, S2)) return jdbl(p);
        return JINF;
    }
    fe H = fsub(f, U2, U1), R = fsub(f, S2, S1);
    fe H2 = fmul(f, H, H), H3 = fmul(f, H2, H), U1H2 = fmul(f, U1, H2);
    jac r;
    r.X = fsub(f, fsub(f, fmul(f, R, R), H3), fadd(f, U1H2, U1H2));
    r.Y = fsub(f, fmul(f, R, fsub(f, U1H2, r.X)), fmul(f, S1, H3));
    r.Z = fmul(f, fmul(f, p.Z, q.Z), H);
    return r;
}

static jac jmadd(jac p, aff q) {
    if (q.inf) return p;
    const field* f = &FQ;
    if (fzero(p.Z)) {
        jac r = {q.x, q.y, FQ.one};
        return r;
    }
    fe Z1Z1 = fmul(f, p.Z, p.Z);
    fe U2 = fmul(f, q.x, Z1Z1);
    fe S2 = fmul(f, fmul(f, q.y, p.Z), Z1Z1);
    if (feq(p.X, U2)) {
        if (feq(p.Y, S2)) return jdbl(p);
        return JINF;
    }
    fe H = fsub(f, U2, p.X), R = fsub(f, S2, p.Y);
    fe H2 = fmul(f, H, H), H3 = fmul(f, H2, H), U1H2 = fmul(f, p.X, H2);
    jac r;
    r.X = fsub(f, fsub(f, fmul(f, R, R), H3), fadd(f, U1H2, U1H2));
    r.Y = fsub(f, fmul(f, R, fsub(f, U1H2, r.X)), fmul(f, p.Y, H3));
    r.Z = fmul(f, p.Z, H);
    return r;
}

static void jac_to_bytes(jac p, uint8_t* out) {
    if (fzero(p.Z)) { memset(out, 0, 64); return; }
    fe zi = finv(&FQ, p.Z), zi2 = fmul(&FQ, zi, zi);
    from_mont(&FQ, fmul(&FQ, p.X, zi2), out);
    from_mont(&FQ, fmul(&FQ, p.Y, fmul(&FQ, zi2, zi)), out + 32);
}

static aff aff_from_bytes(const uint8_t* b) {
    aff a;
    int any = 0;
    for (int i = 0; i < 64; ++i) any |= b[i];
    a.inf = !any;
    if (any) {
        a.x = to_mont(&FQ, b);
        a.y = to_mont(&FQ, b + 32);
    }
    return a;
}

static void init_fields(void) {
    static int done = 0;
    if (done) return;
    const uint64_t q[4] = {0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull, 0x30644e72e131a029ull};
    const uint64_t r[4] = {0x43e1f593f0000001ull, 0x2833e84879b97091ull, 0xb85045b68181585dull, 0x30644e72e131a029ull};
    set_field(&FQ, q);
    set_field(&FR, r);
    done = 1;
}

/* ---- Pippenger ------------------------------------------------------------ */
typedef struct {
    const uint8_t *points, *scalars;
    size_t begin, end;
    int c;
    jac result;
} msm_job;

static uint32_t window_digit(const uint8_t* s, int bit, int c) {
    uint32_t v = 0;
    for (int k = 0; k < c; ++k) {
        int b = bit + k;
        if (b >= 256) break;
        v |= (uint32_t)((s[b >> 3] >> (b & 7)) & 1) << k;
    }
    return v;
}

static void* msm_worker(void* arg) {
    msm_job* j = (msm_job*)arg;
    const int c = j->c, windows = (254 + c - 1) / c;
    const size_t nb = (size_t)1 << c;
    jac* buckets = (jac*)malloc(sizeof(jac) * nb);
    aff* pts = (aff*)malloc(sizeof(aff) * (j->end - j->begin));
    for (size_t i = j->begin; i < j->end; ++i) pts[i - j->begin] = aff_from_bytes(j->points + 64 * i);
    jac acc = JINF;
    for (int w = windows - 1; w >= 0; --w) {
        for (int k = 0; k < c; ++k) acc = jdbl(acc);
        for (size_t b = 0; b < nb; ++b) buckets[b] = JINF;
        for (size_t i = j->begin; i < j->end; ++i) {
            uint32_t d = window_digit(j->scalars + 32 * i, w * c, c);
            if (d) buckets[d] = jmadd(buckets[d], pts[i - j->begin]);
        }
        jac run = JINF, sum = JINF;
        for (size_t b = nb - 1; b >= 1; --b) {
            run = jadd(run, buckets[b]);
            sum = jadd(sum, run);
        }
        acc = jadd(acc, sum);
    }
    free(buckets);
    free(pts);
    j->result = acc;
    return NULL;
}

/* sum s_i P_i; points 64 B canonical (zero = infinity), scalars 32 B LE. */
int bn254ref_msm_g1(const uint8_t* points, const uint8_t* scalars, size_t n, int threads, uint8_t* out) {
    init_fields();
    if (threads < 1) threads = 1;
    if ((size_t)threads > n) threads = n ? (int)n : 1;
    msm_job* jobs = (msm_job*)calloc((size_t)threads, sizeof(msm_job));
    pthread_t* tids = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    size_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        jobs[t].points = points;
        jobs[t].scalars = scalars;
        jobs[t].begin = t * per < n ? t * per : n;
        jobs[t].end = (t + 1) * per < n ? (t + 1) * per : n;
        size_t m = jobs[t].end - jobs[t].begin;
        int c = 1;
        while (((size_t)1 << (c + 1)) < m / 4 + 2 && c < 16) ++c;
        jobs[t].c = c < 2 ? 2 : c;
        pthread_create(&tids[t], NULL, msm_worker, &jobs[t]);
    }
    jac acc = JINF;
    for (int t = 0; t < threads; ++t) {
        pthread_join(tids[t], NULL);
        acc = jadd(acc, jobs[t].result);
    }
    jac_to_bytes(acc, out);
    free(jobs);
    free(tids);
    return 0;
}

/* [s_i] G1 (canonical affine out). */
int bn254ref_fixed_base_g1(const uint8_t* scalars, size_t n, uint8_t* out) {
    init_fields();
    uint8_t g[64] = {0};
    g[0] = 1;
    g[32] = 2;
    aff G = aff_from_bytes(g);
    for (size_t i = 0; i < n; ++i) {
        jac acc = JINF;
        for (int b = 255; b >= 0; --b) {
            acc = jdbl(acc);
            if ((scalars[32 * i + (b >> 3)] >> (b & 7)) & 1) acc = jmadd(acc, G);
        }
        jac_to_bytes(acc, out + 64 * i);
    }
    return 0;
}

/* ---- NTT ------------------------------------------------------------------ */
typedef struct {
    fe* a;
    size_t n, h, lo, hi;
    const fe* tw;  /* w_{2h}^j for j < h */
} ntt_job;

static void* ntt_worker(void* arg) {
    ntt_job* j = (ntt_job*)arg;
    for (size_t start = j->lo; start < j->hi; start += 2 * j->h)
        for (size_t k = 0; k < j->h; ++k) {
            fe u = j->a[start + k];
            fe v = fmul(&FR, j->a[start + k + j->h], j->tw[k]);
            j->a[start + k] = fadd(&FR, u, v);
            j->a[start + k + j->h] = fsub(&FR, u, v);
        }
    return NULL;
}

int bn254ref_ntt(const uint8_t* in, size_t n, int inverse, const uint8_t* root, int threads, uint8_t* out) {
    init_fields();
    if (!n || (n & (n - 1)) || n > (1u << 20)) return -1;
    int logn = 0;
    while (((size_t)1 << logn) < n) ++logn;
    fe* a = (fe*)malloc(sizeof(fe) * n);
    for (size_t i = 0; i < n; ++i) {
        size_t r = 0;
        for (int b = 0; b < logn; ++b) r |= ((i >> b) & 1) << (logn - 1 - b);
        a[r] = to_mont(&FR, in + 32 * i);
    }
    fe w = to_mont(&FR, root);
    for (int k = logn; k < 20; ++k) w = fmul(&FR, w, w); /* w_n = root^(2^20/n) */
    if (inverse) w = finv(&FR, w);
    fe* tw = (fe*)malloc(sizeof(fe) * (n / 2 + 1));
    if (threads < 1) threads = 1;
    pthread_t* tids = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    ntt_job* jobs = (ntt_job*)calloc((size_t)threads, sizeof(ntt_job));
    for (size_t h = 1; h < n; h <<= 1) {
        fe wh = w;
        for (size_t s = h; s < n / 2; s <<= 1) wh = fmul(&FR, wh, wh); /* w_{2h} */
        tw[0] = FR.one;
        for (size_t k = 1; k < h; ++k) tw[k] = fmul(&FR, tw[k - 1], wh);
        size_t blocks = n / (2 * h);
        int nt = (size_t)threads < blocks ? threads : (int)blocks;
        size_t per = (blocks + nt - 1) / nt;
        for (int t = 0; t < nt; ++t) {
            jobs[t].a = a;
            jobs[t].n = n;
            jobs[t].h = h;
            jobs[t].tw = tw;
            size_t b0 = t * per < blocks ? t * per : blocks, b1 = (t + 1) * per < blocks ? (t + 1) * per : blocks;
            jobs[t].lo = b0 * 2 * h;
            jobs[t].hi = b1 * 2 * h;
            if (nt == 1) ntt_worker(&jobs[t]);
            else pthread_create(&tids[t], NULL, ntt_worker, &jobs[t]);
        }
        if (nt > 1)
            for (int t = 0; t < nt; ++t) pthread_join(tids[t], NULL);
    }
    fe scale = FR.one;
    if (inverse) {
        uint8_t nb[32] = {0};
        for (int k = 0; k < 8; ++k) nb[k] = (uint8_t)(n >> (8 * k));
        scale = finv(&FR, to_mont(&FR, nb));
    }
    for (size_t i = 0; i < n; ++i) from_mont(&FR, fmul(&FR, a[i], scale), out + 32 * i);
    free(a);
    free(tw);
    free(tids);
    free(jobs);
    return 0;
}

/* ---- KZG opening cache (crates/protocol/src/kzg.rs:217-275) ------------- */
/* [k]P by double-and-add over the 254-bit canonical scalar k. */
static jac jmul(jac p, const uint64_t k[4]) {
    jac acc = JINF;
    for (int b = 253; b >= 0; --b) {
        acc = jdbl(acc);
        if ((k[b >> 6] >> (b & 63)) & 1) acc = jadd(acc, p);
    }
    return acc;
}

static void fr_canon(fe x, uint64_t k[4]) {
    uint8_t b[32];
    from_mont(&FR, x, b);
    memcpy(k, b, 32);
}

typedef struct {
    jac* a;
    size_t h, lo, hi;        /* butterflies [lo, hi) of the stage with half-size h */
    const uint64_t* tw;      /* canonical w_{2h}^j, 4 words each */
} gfft_job;

static void* gfft_worker(void* arg) {
    gfft_job* j = (gfft_job*)arg;
    for (size_t b = j->lo; b < j->hi; ++b) {
        size_t blk = b / j->h, k = b % j->h, i = blk * 2 * j->h + k;
        jac u = j->a[i], v = j->a[i + j->h];
        if (k) v = jmul(v, j->tw + 4 * k);
        j->a[i] = jadd(u, v);
        jac nv = v;
        nv.Y = fsub(&FQ, (fe){{0}}, v.Y);
        j->a[i + j->h] = jadd(u, nv);
    }
    return NULL;
}

/* Natural-order DFT over <w_n> in G1 (radix-2 DIT, the textbook iterative
 * form of Radix2EvaluationDomain::fft); inverse uses w^-1 and scales 1/n. */
static void gfft(jac* a, size_t n, fe w_n, int inverse, int threads) {
    int logn = 0;
    while (((size_t)1 << logn) < n) ++logn;
    for (size_t i = 0; i < n; ++i) {
        size_t r = 0;
        for (int b = 0; b < logn; ++b) r |= ((i >> b) & 1) << (logn - 1 - b);
        if (r > i) { jac t = a[i]; a[i] = a[r]; a[r] = t; }
    }
    fe w = inverse ? finv(&FR, w_n) : w_n;
    uint64_t* tw = (uint64_t*)malloc(sizeof(uint64_t) * 4 * (n / 2 + 1));
    if (threads < 1) threads = 1;
    pthread_t* tids = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    gfft_job* jobs = (gfft_job*)calloc((size_t)threads, sizeof(gfft_job));
    for (size_t h = 1; h < n; h <<= 1) {
        fe wh = w;
        for (size_t s = h; s < n / 2; s <<= 1) wh = fmul(&FR, wh, wh); /* w_{2h} */
        fe t = FR.one;
        for (size_t k = 0; k < h; ++k) {
            fr_canon(t, tw + 4 * k);
            t = fmul(&FR, t, wh);
        }
        size_t total = n / 2, per = (total + threads - 1) / threads;
        for (int q = 0; q < threads; ++q) {
            jobs[q].a = a;
            jobs[q].h = h;
            jobs[q].tw = tw;
            jobs[q].lo = q * per < total ? q * per : total;
            jobs[q].hi = (q + 1) * per < total ? (q + 1) * per : total;
            if (threads == 1) gfft_worker(&jobs[q]);
            else pthread_create(&tids[q], NULL, gfft_worker, &jobs[q]);
        }
        if (threads > 1)
            for (int q = 0; q < threads; ++q) pthread_join(tids[q], NULL);
    }
    if (inverse && n > 1) {
        uint8_t nb[32] = {0};
        for (int k = 0; k < 8; ++k) nb[k] = (uint8_t)(n >> (8 * k));
        uint64_t s[4];
        fr_canon(finv(&FR, to_mont(&FR, nb)), s);
        for (size_t i = 0; i < n; ++i) a[i] = jmul(a[i], s);
    }
    free(tw);
    free(tids);
    free(jobs);
}

static fe root_of(const uint8_t* root_2_20, size_t n) {
    fe w = to_mont(&FR, root_2_20);
    int logn = 0;
    while (((size_t)1 << logn) < n) ++logn;
    for (int k = logn; k < 20; ++k) w = fmul(&FR, w, w);
    return w;
}

typedef struct { jac* a; const fe* s; size_t lo, hi; } scale_job;

static void* scale_worker(void* arg) {
    scale_job* j = (scale_job*)arg;
    for (size_t i = j->lo; i < j->hi; ++i) {
        uint64_t k[4];
        fr_canon(j->s[i], k);
        j->a[i] = jmul(j->a[i], k);
    }
    return NULL;
}

static void scale_all(jac* a, const fe* s, size_t n, int threads) {
    if (threads < 1) threads = 1;
    pthread_t* tids = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    scale_job* jobs = (scale_job*)calloc((size_t)threads, sizeof(scale_job));
    size_t per = (n + threads - 1) / threads;
    for (int q = 0; q < threads; ++q) {
        jobs[q].a = a;
        jobs[q].s = s;
        jobs[q].lo = q * per < n ? q * per : n;
        jobs[q].hi = (q + 1) * per < n ? (q + 1) * per : n;
        pthread_create(&tids[q], NULL, scale_worker, &jobs[q]);
    }
    for (int q = 0; q < threads; ++q) pthread_join(tids[q], NULL);
    free(tids);
    free(jobs);
}

/* precompute_single_openings: powers (count x 64 B affine), coefficients
 * (ncoeff x 32 B), domain of n points (coset by `offset` when not NULL).
 * out: n x 64 B affine.  Returns -1 for the reference's error cases. */
int bn254ref_kzg_openings(const uint8_t* powers, size_t count, const uint8_t* coeffs, size_t ncoeff, size_t n,
                          const uint8_t* offset, const uint8_t* root_2_20, int threads, uint8_t* out) {
    init_fields();
    if (!n || (n & (n - 1)) || n > (1u << 20) || ncoeff > n || ncoeff > count) return -1;
    size_t degree = ncoeff ? ncoeff - 1 : 0;
    if (!degree) {
        memset(out, 0, 64 * n);
        return 0;
    }
    size_t M = 1;
    while (M < 2 * degree - 1) M <<= 1;
    if (M > (1u << 20)) return -1;
    /* scalar_fft = DFT_M(f_d, ..., f_1) */
    uint8_t* rev = (uint8_t*)calloc(M, 32);
    for (size_t p = 0; p < degree; ++p) memcpy(rev + 32 * p, coeffs + 32 * (degree - p), 32);
    uint8_t* sf = (uint8_t*)malloc(32 * M);
    bn254ref_ntt(rev, M, 0, root_2_20, threads, sf);
    /* prepared = DFT_M([tau^0], ..., [tau^(d-1)]) */
    jac* conv = (jac*)malloc(sizeof(jac) * M);
    for (size_t p = 0; p < M; ++p) {
        conv[p] = JINF;
        if (p < degree) {
            aff a = aff_from_bytes(powers + 64 * p);
            if (!a.inf) { conv[p].X = a.x; conv[p].Y = a.y; conv[p].Z = FQ.one; }
        }
    }
    gfft(conv, M, root_of(root_2_20, M), 0, threads);
    fe* s = (fe*)malloc(sizeof(fe) * (M > n ? M : n));
    for (size_t p = 0; p < M; ++p) s[p] = to_mont(&FR, sf + 32 * p);
    scale_all(conv, s, M, threads);
    gfft(conv, M, root_of(root_2_20, M), 1, threads);
    /* truncate(degree), reverse, zero-pad to n */
    jac* h = (jac*)malloc(sizeof(jac) * n);
    for (size_t j = 0; j < n; ++j) h[j] = j < degree ? conv[degree - 1 - j] : JINF;
    if (offset) { /* coset fft: coefficients times offset^j */
        fe g = to_mont(&FR, offset), t = FR.one;
        for (size_t j = 0; j < n; ++j) {
            s[j] = t;
            t = fmul(&FR, t, g);
        }
        scale_all(h, s, n, threads);
    }
    gfft(h, n, root_of(root_2_20, n), 0, threads);
    for (size_t j = 0; j < n; ++j) jac_to_bytes(h[j], out + 64 * j);
    free(rev);
    free(sf);
    free(conv);
    free(s);
    free(h);
    return 0;
}
