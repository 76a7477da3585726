// Device kernels of fixed-base setup and G2 validation (templated on G1/G2).
#pragma once
#include <cstring>
#include <vector>

#include "setup.hpp"

namespace dvgpu {

using namespace bn254;

template <class F>
struct Gen;
template <>
struct Gen<FqOps> {
    static HD Affine<FqOps> g() { return {fp_const<FqParams>(G1_X), fp_const<FqParams>(G1_Y), false}; }
    static constexpr int EW = 8;  // words per field element
};
template <>
struct Gen<Fq2Ops> {
    static HD Affine<Fq2Ops> g() {
        return {{fp_const<FqParams>(G2_X_C0), fp_const<FqParams>(G2_X_C1)},
                {fp_const<FqParams>(G2_Y_C0), fp_const<FqParams>(G2_Y_C1)},
                false};
    }
    static constexpr int EW = 16;
};

// ---- element (de)serialisation shared by host and device ------------------

HD void put_fe(uint32_t* w, const Fq& v) {
    uint32_t o[8];
    pack(v, o);
    for (int i = 0; i < 8; ++i) w[i] = o[i];
}
HD void put_fe(uint32_t* w, const Fq2& v) {
    put_fe(w, v.c0);
    put_fe(w + 8, v.c1);
}
HD void get_fe(const uint32_t* w, Fq& v) {
    uint32_t o[8];
    for (int i = 0; i < 8; ++i) o[i] = w[i];
    v = unpack<FqParams>(o);
}
HD void get_fe(const uint32_t* w, Fq2& v) {
    get_fe(w, v.c0);
    get_fe(w + 8, v.c1);
}
// canonical (from Montgomery) words
HD void put_canon(uint32_t* w, const Fq& v) {
    uint32_t o[8];
    from_mont(v, o);
    for (int i = 0; i < 8; ++i) w[i] = o[i];
}
HD void put_canon(uint32_t* w, const Fq2& v) {
    put_canon(w, v.c0);
    put_canon(w + 8, v.c1);
}

// ---- host: window table -----------------------------------------------------

// table[w*16 + d] = d * 16^w * G as affine Montgomery words (zero = infinity).
template <class F>
std::vector<uint32_t> build_setup_table() {
    using E = typename F::E;
    constexpr int EW = Gen<F>::EW;
    std::vector<Xyzz<F>> pts(SETUP_TABLE);
    Xyzz<F> base = xyzz_from_affine<F>(Gen<F>::g());
    for (int w = 0; w < SETUP_WINDOWS; ++w) {
        pts[w * 16] = xyzz_inf<F>();
        for (int d = 1; d < 16; ++d) pts[w * 16 + d] = xyzz_add<F>(pts[w * 16 + d - 1], base);
        base = xyzz_add<F>(pts[w * 16 + 15], base);  // 16 * base
    }
    // Batch inversion of ZZZ over the finite entries (Montgomery's trick).
    std::vector<E> prefix(SETUP_TABLE);
    E acc = F::one();
    for (int i = 0; i < SETUP_TABLE; ++i) {
        prefix[i] = acc;
        if (!xyzz_is_inf<F>(pts[i])) acc = F::mul(acc, pts[i].ZZZ);
    }
    E inv_acc = F::inv(acc);
    std::vector<uint32_t> out(size_t(SETUP_TABLE) * 2 * EW, 0);
    for (int i = SETUP_TABLE - 1; i >= 0; --i) {
        if (xyzz_is_inf<F>(pts[i])) continue;
        const E izzz = F::mul(inv_acc, prefix[i]);
        inv_acc = F::mul(inv_acc, pts[i].ZZZ);
        const E iz = F::mul(izzz, pts[i].ZZ);
        const E izz = F::sqr(iz);
        put_fe(&out[size_t(i) * 2 * EW], F::canon(F::mul(pts[i].X, izz)));
        put_fe(&out[size_t(i) * 2 * EW + EW], F::canon(F::mul(pts[i].Y, izzz)));
    }
    return out;
}

// ---- device ------------------------------------------------------------------

template <class F>
__device__ __forceinline__ Affine<F> load_table(const uint32_t* table, int idx) {
    constexpr int EW = Gen<F>::EW;
    Affine<F> a;
    get_fe(table + size_t(idx) * 2 * EW, a.x);
    get_fe(table + size_t(idx) * 2 * EW + EW, a.y);
    a.inf = false;
    return a;
}

// Block-wide batch inversion (Montgomery's trick as a product tree in LDS).
// Every thread of the 256-thread block must call it.
template <class F>
__device__ typename F::E block_inverse(const typename F::E& v, uint32_t* tree) {
    using E = typename F::E;
    constexpr int EW = Gen<F>::EW;
    constexpr unsigned B = 256;
    const unsigned t = threadIdx.x;
    put_fe(tree + (B + t) * EW, v);
    __syncthreads();
    for (unsigned width = B / 2; width; width >>= 1) {
        if (t < width) {
            const unsigned node = width + t;
            E l, r;
            get_fe(tree + 2 * node * EW, l);
            get_fe(tree + (2 * node + 1) * EW, r);
            put_fe(tree + node * EW, F::mul(l, r));
        }
        __syncthreads();
    }
    if (t == 0) {
        E root;
        get_fe(tree + EW, root);
        put_fe(tree + EW, F::inv(root));
    }
    __syncthreads();
    for (unsigned width = 1; width < B; width <<= 1) {
        if (t < width) {
            const unsigned node = width + t;
            E iv, l, r;
            get_fe(tree + node * EW, iv);
            get_fe(tree + 2 * node * EW, l);
            get_fe(tree + (2 * node + 1) * EW, r);
            put_fe(tree + 2 * node * EW, F::mul(iv, r));
            put_fe(tree + (2 * node + 1) * EW, F::mul(iv, l));
        }
        __syncthreads();
    }
    E out;
    get_fe(tree + (B + t) * EW, out);
    return out;
}

// mode 0: double-and-add; mode 1/2: 4-bit window table.  Output affine,
// either canonical little-endian (CANON) or Montgomery words.
template <class F, bool CANON>
__global__ void __launch_bounds__(256) k_fixed_base(const uint32_t* __restrict__ scalars, size_t n,
                                                    const uint32_t* __restrict__ table, int mode,
                                                    uint32_t* __restrict__ out) {
    using E = typename F::E;
    constexpr int EW = Gen<F>::EW;
    __shared__ uint4 tree4[2 * 256 * EW / 4];
    uint32_t* tree = reinterpret_cast<uint32_t*>(tree4);
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    const bool live = i < n;
    Xyzz<F> acc = xyzz_inf<F>();
    if (live) {
        uint32_t s[8];
        for (int k = 0; k < 8; ++k) s[k] = scalars[i * 8 + k];
        if (mode == 0) {
            acc = xyzz_mul_words<F, 8>(xyzz_from_affine<F>(Gen<F>::g()), s, 256);
        } else {
            for (int w = 0; w < SETUP_WINDOWS; ++w) {
                const uint32_t d = (s[w >> 3] >> (4 * (w & 7))) & 15u;
                if (d) acc = xyzz_madd<F>(acc, load_table<F>(table, w * 16 + int(d)));
            }
        }
    }
    const bool inf = xyzz_is_inf<F>(acc);
    E izzz;
    if (mode == 2) {
        izzz = block_inverse<F>(inf ? F::one() : acc.ZZZ, tree);
    } else if (!inf) {
        izzz = F::inv(acc.ZZZ);
    }
    if (!live) return;
    uint32_t* o = out + i * 2 * EW;
    if (inf) {
        for (int k = 0; k < 2 * EW; ++k) o[k] = 0;
        return;
    }
    const E iz = F::mul(izzz, acc.ZZ);
    const E x = F::canon(F::mul(acc.X, F::sqr(iz)));
    const E y = F::canon(F::mul(acc.Y, izzz));
    if (CANON) {
        put_canon(o, x);
        put_canon(o + EW, y);
    } else {
        put_fe(o, x);
        put_fe(o + EW, y);
    }
}

}  // namespace dvgpu

namespace dvgpu {

template <class F>
const uint32_t* setup_table(SetupState& st, hipStream_t stream) {
    constexpr bool G2 = Gen<F>::EW == 16;
    DeviceBuffer& buf = G2 ? st.table_g2 : st.table_g1;
    bool& have = G2 ? st.have_g2 : st.have_g1;
    if (!have) {
        const std::vector<uint32_t> host = build_setup_table<F>();
        uint32_t* dev = buf.as<uint32_t>(host.size());
        HIP_CHECK(hipMemcpyAsync(dev, host.data(), host.size() * 4, hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipStreamSynchronize(stream));  // `host` dies here
        have = true;
        ++st.stats.table_builds[G2 ? 1 : 0];
        st.stats.table_bytes += host.size() * 4;
    }
    return static_cast<const uint32_t*>(buf.ptr);
}

template <class F>
void fixed_base_enqueue(SetupState& st, const uint32_t* scalars, size_t n, uint8_t* out_canonical,
                        hipStream_t stream) {
    if (!n) return;
    const uint32_t* table = st.mode ? setup_table<F>(st, stream) : nullptr;
    k_fixed_base<F, true><<<grid_for(n, 256), 256, 0, stream>>>(scalars, n, table, st.mode,
                                                                 reinterpret_cast<uint32_t*>(out_canonical));
    HIP_CHECK_LAUNCH();
}

template <class F>
void fixed_base_mont_enqueue(SetupState& st, const uint32_t* scalars, size_t n, uint32_t* out_mont,
                             hipStream_t stream) {
    if (!n) return;
    const uint32_t* table = setup_table<F>(st, stream);
    k_fixed_base<F, false><<<grid_for(n, 256), 256, 0, stream>>>(scalars, n, table, 2, out_mont);
    HIP_CHECK_LAUNCH();
}

}  // namespace dvgpu
