// Polynomial arithmetic over Fr with batched device products
// (include/dynaverse_poly.h; reference: crates/protocol/src/domain/fast.rs,
// `SubproductTree`, `newton_divide_exact`, `fft_multiply_with_backend`).
//
// The reference multiplies one pair at a time through backend.ntt (three
// transforms and two host round trips per product).  Here every product of
// one round -- a tree level, one Newton step of every node of a level, one
// remainder level -- goes to the device together: one upload of all padded
// operands, one fused NTT batch over every row, one pointwise kernel, one
// fused inverse batch, one download.  Polynomials live on the host as
// canonical little-endian words; the orchestration mirrors the reference's
// recursion level by level, so every result is the same polynomial or value.
// Subproduct trees over DEV_TREE_MIN_POINTS points or more are kept in HBM
// instead and run their rounds there (poly_tree.hip).
// Intermediate remainders are kept untrimmed (same polynomials, possibly
// with zero high coefficients); the outputs are trimmed as the reference's.
#include <algorithm>
#include <array>
#include <iterator>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/dynaverse_poly.h"
#include "g1mem.hpp"
#include "owner.hpp"
#include "poly_tree.hpp"

namespace dvgpu {

namespace {

using W8 = std::array<uint32_t, 8>;  // canonical Fr
using Poly = std::vector<W8>;

W8 zero8() { return W8{}; }

bool is_zero8(const W8& a) {
    for (uint32_t v : a)
        if (v) return false;
    return true;
}

Fr to_fr(const W8& a) {
    uint32_t w[8];
    std::memcpy(w, a.data(), 32);
    return to_mont<FrParams>(w);
}

W8 from_fr(const Fr& v) {
    uint32_t w[8];
    from_mont(canon(v), w);
    W8 o;
    std::memcpy(o.data(), w, 32);
    return o;
}

// Canonical-word addition, subtraction and negation mod r (no Montgomery
// round trip: these run once per coefficient per round on the host).
bool geq_r(const W8& a) {
    for (int i = 7; i >= 0; --i) {
        if (a[i] != FR_MOD32[i]) return a[i] > FR_MOD32[i];
    }
    return true;
}

W8 sub_r(W8 a) {  // a - r, for r <= a < 2r
    uint64_t borrow = 0;
    for (int i = 0; i < 8; ++i) {
        const uint64_t d = uint64_t(a[i]) - FR_MOD32[i] - borrow;
        a[i] = uint32_t(d);
        borrow = (d >> 32) & 1;
    }
    return a;
}

W8 add8(const W8& a, const W8& b) {
    W8 r;
    uint64_t carry = 0;
    for (int i = 0; i < 8; ++i) {
        const uint64_t t = uint64_t(a[i]) + b[i] + carry;
        r[i] = uint32_t(t);
        carry = t >> 32;
    }
    return (carry || geq_r(r)) ? sub_r(r) : r;  // a + b < 2r < 2^256: no carry out in practice
}

W8 sub8(const W8& a, const W8& b) {
    W8 r;
    uint64_t borrow = 0;
    for (int i = 0; i < 8; ++i) {
        const uint64_t d = uint64_t(a[i]) - b[i] - borrow;
        r[i] = uint32_t(d);
        borrow = (d >> 32) & 1;
    }
    if (borrow) {  // + r
        uint64_t carry = 0;
        for (int i = 0; i < 8; ++i) {
            const uint64_t t = uint64_t(r[i]) + FR_MOD32[i] + carry;
            r[i] = uint32_t(t);
            carry = t >> 32;
        }
    }
    return r;
}

W8 neg8(const W8& a) { return sub8(W8{}, a); }

void trim(Poly& p) {
    while (p.size() > 1 && is_zero8(p.back())) p.pop_back();
    if (p.empty()) p.push_back(zero8());
}

// a[i] <- a[i] * b[i] (canonical words in and out) over `n` words.
__global__ void k_pointwise_fr(uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x[8], y[8];
    for (int q = 0; q < 8; ++q) {
        x[q] = a[i * 8 + q];
        y[q] = b[i * 8 + q];
    }
    // (x R^-1)(y) R^-1 R^2 R^-1... : unpack both as residues, one product with
    // R^2 restores the plain product x y mod r.
    const Fr p = mul(mul(unpack<FrParams>(x), unpack<FrParams>(y)), fp_const<FrParams>(FrParams::R2));
    uint32_t o[8];
    pack(canon(p), o);
    for (int q = 0; q < 8; ++q) a[i * 8 + q] = o[q];
}

struct MulJob {
    const Poly* a;
    const Poly* b;
    Poly* out;  // receives a * b, length la + lb - 1 (untrimmed); [0] for an empty operand
};

class PolyEngine {
  public:
    explicit PolyEngine(Owner* o) : o_(o) {}

    // A round whose products are all small (at most HOST_PRODUCT_MAX
    // coefficient pairs each, HOST_ROUND_MAX in total) is formed on the host
    // by schoolbook: a device round costs a fixed few hundred microseconds,
    // which such a round does not repay.
    static constexpr size_t HOST_PRODUCT_MAX = 256, HOST_ROUND_MAX = 4096;

    static void host_mul(const Poly& a, const Poly& b, Poly& out) {
        std::vector<Fr> acc(a.size() + b.size() - 1, fp_zero<FrParams>());
        std::vector<Fr> bm(b.size());
        for (size_t j = 0; j < b.size(); ++j) bm[j] = to_fr(b[j]);
        for (size_t i = 0; i < a.size(); ++i) {
            const Fr x = to_fr(a[i]);
            for (size_t j = 0; j < b.size(); ++j) acc[i + j] = bn254::add(acc[i + j], bn254::mul(x, bm[j]));
        }
        out.resize(acc.size());
        for (size_t i = 0; i < acc.size(); ++i) out[i] = from_fr(acc[i]);
    }

    // Every device-sized job of the round in one device pass.
    void mul(std::vector<MulJob>& jobs) {
        std::vector<size_t> size(jobs.size(), 0), off(jobs.size(), 0);
        size_t total = 0, host_work = 0;
        bool all_small = true;
        for (const MulJob& j : jobs) {
            const size_t w = j.a->size() * j.b->size();
            all_small = all_small && w <= HOST_PRODUCT_MAX;
            host_work += w;
        }
        const bool on_host = all_small && host_work <= HOST_ROUND_MAX;
        for (size_t j = 0; j < jobs.size(); ++j) {
            const size_t la = jobs[j].a->size(), lb = jobs[j].b->size();
            if (!la || !lb) continue;
            if (on_host) {
                host_mul(*jobs[j].a, *jobs[j].b, *jobs[j].out);
                ++stats_.products;
                size[j] = SIZE_MAX;  // done
                continue;
            }
            size_t s = 1;
            while (s < la + lb - 1) s <<= 1;
            require(s <= DVGPU_MAX_NTT_SIZE, "polynomial product exceeds 2^20 coefficients");
            size[j] = s;
            off[j] = total;
            total += s;
        }
        for (size_t j = 0; j < jobs.size(); ++j) {
            if (!size[j]) *jobs[j].out = Poly{zero8()};
            if (size[j] == SIZE_MAX) size[j] = 0;  // host product: skip below
        }
        if (!total) return;
        require(total < (size_t(1) << 30), "polynomial batch too large");
        // pinned host staging: all padded left operands, then all padded right operands
        uint32_t* host = o_->poly_stage.as<uint32_t>(2 * total * 8);
        std::memset(host, 0, 2 * total * 32);
        for (size_t j = 0; j < jobs.size(); ++j) {
            if (!size[j]) continue;
            std::memcpy(host + off[j] * 8, jobs[j].a->data(), jobs[j].a->size() * 32);
            std::memcpy(host + (total + off[j]) * 8, jobs[j].b->data(), jobs[j].b->size() * 32);
        }
        hipStream_t st = o_->stream;
        auto* dev = o_->stage_a.as<uint32_t>(2 * total * 8);
        HIP_CHECK(hipMemcpyAsync(dev, host, 2 * total * 32, hipMemcpyHostToDevice, st));
        // one multi-row entry per run of consecutive products of one padded
        // length (their rows are adjacent in both operand halves)
        std::vector<NttBatchEntry> fwd, inv;
        size_t device_jobs = 0;
        for (size_t j = 0; j < jobs.size();) {
            if (!size[j]) {
                ++j;
                continue;
            }
            size_t e = j + 1;
            while (e < jobs.size() && size[e] == size[j] && off[e] == off[e - 1] + size[j]) ++e;
            int L = 0;
            while ((size_t(1) << L) < size[j]) ++L;
            uint32_t* a = dev + off[j] * 8;
            uint32_t* b = dev + (total + off[j]) * 8;
            fwd.push_back({a, a, L, e - j, false});
            fwd.push_back({b, b, L, e - j, false});
            inv.push_back({a, a, L, e - j, true});
            device_jobs += e - j;
            j = e;
        }
        ntt_batch_prepare(o_->domain, fwd, o_->ntt_batch, st);
        ntt_batch_enqueue(o_->domain, o_->ntt_batch, st);
        k_pointwise_fr<<<grid_for(total, 256), 256, 0, st>>>(dev, dev + total * 8, total);
        HIP_CHECK_LAUNCH();
        // its own plan: preparing it does not wait for the forward half
        ntt_batch_prepare(o_->domain, inv, o_->ntt_batch_inv, st);
        ntt_batch_enqueue(o_->domain, o_->ntt_batch_inv, st);
        HIP_CHECK(hipMemcpyAsync(host, dev, total * 32, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (size_t j = 0; j < jobs.size(); ++j) {
            if (!size[j]) continue;
            const size_t n = jobs[j].a->size() + jobs[j].b->size() - 1;
            jobs[j].out->resize(n);
            std::memcpy(jobs[j].out->data(), host + off[j] * 8, n * 32);
        }
        ++stats_.rounds;
        stats_.products += device_jobs;
        stats_.device_products += device_jobs;
    }

    // Reciprocals rev(den_j)^-1 mod x^len_j by Newton steps, all divisors in
    // lockstep (one device pass per half-step).
    std::vector<Poly> reciprocals(const std::vector<const Poly*>& dens, const std::vector<size_t>& lens) {
        const size_t k = dens.size();
        std::vector<Poly> rev(k), inv(k);
        for (size_t j = 0; j < k; ++j) {
            rev[j].assign(dens[j]->rbegin(), dens[j]->rend());
            require(!is_zero8(rev[j][0]), "cannot divide by zero polynomial");
            inv[j] = Poly{from_fr(bn254::inv(to_fr(rev[j][0])))};
        }
        const W8 two = from_fr(add(fp_one<FrParams>(), fp_one<FrParams>()));
        for (;;) {
            std::vector<size_t> act;
            for (size_t j = 0; j < k; ++j)
                if (inv[j].size() < lens[j]) act.push_back(j);
            if (act.empty()) break;
            std::vector<Poly> head(act.size()), corr(act.size()), next(act.size());
            std::vector<size_t> nlen(act.size());
            std::vector<MulJob> jobs;
            for (size_t t = 0; t < act.size(); ++t) {
                const size_t j = act[t];
                nlen[t] = std::min(2 * inv[j].size(), lens[j]);
                head[t].assign(rev[j].begin(), rev[j].begin() + std::min(rev[j].size(), nlen[t]));
                jobs.push_back({&head[t], &inv[act[t]], &corr[t]});
            }
            mul(jobs);
            for (size_t t = 0; t < act.size(); ++t) {  // corr = 2 - rev * inv (mod x^next)
                corr[t].resize(nlen[t], zero8());
                for (auto& c : corr[t]) c = neg8(c);
                corr[t][0] = add8(corr[t][0], two);
            }
            jobs.clear();
            for (size_t t = 0; t < act.size(); ++t) jobs.push_back({&inv[act[t]], &corr[t], &next[t]});
            mul(jobs);
            for (size_t t = 0; t < act.size(); ++t) {
                next[t].resize(nlen[t], zero8());
                inv[act[t]] = std::move(next[t]);
            }
        }
        return inv;
    }

    // (quotient, remainder) of num_j / den_j for a batch; den_j trimmed and
    // nonzero-led.  Remainders have length len(den) - 1 (at least 1).
    void div_rem(const std::vector<const Poly*>& num, const std::vector<const Poly*>& den,
                 const std::vector<const Poly*>* recips, std::vector<Poly>* quot, std::vector<Poly>& rem) {
        const size_t k = num.size();
        rem.assign(k, Poly{});
        if (quot) quot->assign(k, Poly{zero8()});
        std::vector<size_t> qlen(k, 0);
        std::vector<size_t> need;
        for (size_t j = 0; j < k; ++j) {
            const size_t ln = num[j]->size(), ld = den[j]->size();
            if (ln < ld) {
                rem[j] = *num[j];
                rem[j].resize(std::max<size_t>(ld - 1, 1), zero8());
            } else {
                qlen[j] = ln - ld + 1;
                need.push_back(j);
            }
        }
        if (need.empty()) return;
        std::vector<Poly> own;
        std::vector<const Poly*> rcp(k, nullptr);
        if (recips) {
            for (size_t j : need) rcp[j] = (*recips)[j];
        } else {
            std::vector<const Poly*> d;
            std::vector<size_t> l;
            for (size_t j : need) {
                d.push_back(den[j]);
                l.push_back(qlen[j]);
            }
            own = reciprocals(d, l);
            for (size_t t = 0; t < need.size(); ++t) rcp[need[t]] = &own[t];
        }
        std::vector<Poly> rnum(need.size()), rinv(need.size()), q(need.size()), qd(need.size());
        std::vector<MulJob> jobs;
        for (size_t t = 0; t < need.size(); ++t) {
            const size_t j = need[t];
            rnum[t].assign(num[j]->rbegin(), num[j]->rbegin() + qlen[j]);
            require(rcp[j]->size() >= qlen[j], "reciprocal shorter than the quotient");
            rinv[t].assign(rcp[j]->begin(), rcp[j]->begin() + qlen[j]);
            jobs.push_back({&rnum[t], &rinv[t], &q[t]});
        }
        mul(jobs);
        for (size_t t = 0; t < need.size(); ++t) {
            q[t].resize(qlen[need[t]], zero8());
            std::reverse(q[t].begin(), q[t].end());
        }
        jobs.clear();
        for (size_t t = 0; t < need.size(); ++t) jobs.push_back({&q[t], den[need[t]], &qd[t]});
        mul(jobs);
        for (size_t t = 0; t < need.size(); ++t) {
            const size_t j = need[t], ld = den[j]->size();
            Poly r(std::max<size_t>(ld - 1, 1), zero8());
            for (size_t i = 0; i < ld - 1; ++i) {
                const W8 a = i < num[j]->size() ? (*num[j])[i] : zero8();
                const W8 b = i < qd[t].size() ? qd[t][i] : zero8();
                r[i] = sub8(a, b);
            }
            // the division is exact in the high part by construction; check it
            for (size_t i = ld - 1; i < std::max(num[j]->size(), qd[t].size()); ++i) {
                const W8 a = i < num[j]->size() ? (*num[j])[i] : zero8();
                const W8 b = i < qd[t].size() ? qd[t][i] : zero8();
                require(a == b, "Newton division failed to cancel high coefficients");
            }
            rem[j] = std::move(r);
            if (quot) (*quot)[j] = std::move(q[t]);
        }
    }

    DvgpuPolyStats stats_{};

  private:
    Owner* o_;
};

Poly differentiate(const Poly& p) {
    Poly d;
    for (size_t i = 1; i < p.size(); ++i) {
        uint32_t w[8] = {uint32_t(i), uint32_t(uint64_t(i) >> 32), 0, 0, 0, 0, 0, 0};
        d.push_back(from_fr(mul(to_fr(p[i]), to_mont<FrParams>(w))));
    }
    trim(d);
    return d;
}

Poly read_poly(const uint8_t* b, size_t n) {
    Poly p(n);
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(p[i].data(), b + 32 * i, 32);
        uint32_t w[8];
        std::memcpy(w, p[i].data(), 32);
        require(is_canonical<FrParams>(w), "Fr element is not canonical");
    }
    return p;
}

void write_poly(const Poly& p, uint8_t* out, size_t cap, size_t* len) {
    require(out && len, "null output");
    require(p.size() <= cap, "output capacity too small");
    for (size_t i = 0; i < p.size(); ++i) std::memcpy(out + 32 * i, p[i].data(), 32);
    *len = p.size();
}

}  // namespace

// A balanced subproduct tree (fast.rs:9-64): level 0 holds X - x_i, pairs
// multiply upward, an unpaired node is carried.  recip[l][i] is
// rev(levels[l][i])^-1 to the length the remainder tree needs.  From
// DEV_TREE_MIN_POINTS points the tree lives in HBM instead (`dev`, built and
// used by poly_tree.hip) and the host fields other than `points` stay empty.
struct PolyTree {
    Owner* owner = nullptr;
    std::vector<W8> points;
    std::vector<std::vector<Poly>> levels;
    std::vector<std::vector<Poly>> recip;
    std::vector<W8> derivatives;          // A'(x_i)
    std::vector<W8> inverse_derivatives;  // 1 / A'(x_i)
    DevTree* dev = nullptr;
    ~PolyTree() { dev_tree_destroy(dev); }
};

void poly_tree_destroy(PolyTree* t) noexcept { delete t; }

namespace {

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

void prepare_domain(Owner* o, const uint8_t* root) {
    require(root != nullptr, "null root");
    uint32_t rw[8];
    std::memcpy(rw, root, 32);
    require(ntt_root_has_order_2_20(rw), "NTT root must have order 2^20");
    ntt_domain_prepare(o->domain, rw, o->stream);
}

PolyTree* find_tree(Owner* o, void* t) {
    auto it = std::find(o->poly_trees.begin(), o->poly_trees.end(), static_cast<PolyTree*>(t));
    require(it != o->poly_trees.end(), "unknown polynomial tree handle for this backend");
    return *it;
}

// Remainders of each polynomial modulo every leaf (fast.rs:73-84): its
// values.  Several polynomials over one tree run their remainder trees in
// lockstep (one device pass per level for all of them).
std::vector<std::vector<W8>> tree_eval_many(PolyEngine& e, const PolyTree& t, const std::vector<const Poly*>& polys) {
    const size_t P = polys.size();
    std::vector<std::vector<W8>> out(P);
    if (t.points.empty() || !P) return out;
    const size_t top = t.levels.size() - 1;
    std::vector<Poly> rem;  // P blocks of `width` remainders (one per node of the current level)
    {
        std::vector<const Poly*> num(polys.begin(), polys.end()), den(P, &t.levels[top][0]);
        e.div_rem(num, den, nullptr, nullptr, rem);
    }
    size_t width = 1;
    for (size_t l = top; l-- > 0;) {
        const size_t nodes = t.levels[l].size();
        std::vector<const Poly*> num, den, rc;
        for (size_t q = 0; q < P; ++q)
            for (size_t i = 0; i < nodes; ++i) {
                num.push_back(&rem[q * width + i / 2]);
                den.push_back(&t.levels[l][i]);
                rc.push_back(&t.recip[l][i]);
            }
        std::vector<Poly> next;
        e.div_rem(num, den, &rc, nullptr, next);
        rem = std::move(next);
        width = nodes;
    }
    for (size_t q = 0; q < P; ++q)
        for (size_t i = 0; i < width; ++i) out[q].push_back(rem[q * width + i][0]);
    return out;
}

std::vector<W8> tree_eval(PolyEngine& e, const PolyTree& t, const Poly& poly) {
    return std::move(tree_eval_many(e, t, {&poly})[0]);
}

// fast.rs:90-116
Poly tree_interpolate(PolyEngine& e, const PolyTree& t, const std::vector<W8>& values) {
    require(values.size() == t.points.size(), "interpolation value count differs from points");
    if (values.empty()) return Poly{zero8()};
    std::vector<Poly> polys;
    for (size_t i = 0; i < values.size(); ++i)
        polys.push_back(Poly{from_fr(mul(to_fr(values[i]), to_fr(t.inverse_derivatives[i])))});
    for (size_t l = 0; l + 1 < t.levels.size(); ++l) {
        const size_t pairs = polys.size() / 2;
        std::vector<Poly> left(pairs), right(pairs);
        std::vector<MulJob> jobs;
        for (size_t p = 0; p < pairs; ++p) {
            jobs.push_back({&polys[2 * p], &t.levels[l][2 * p + 1], &left[p]});
            jobs.push_back({&polys[2 * p + 1], &t.levels[l][2 * p], &right[p]});
        }
        e.mul(jobs);
        std::vector<Poly> next;
        for (size_t p = 0; p < pairs; ++p) {
            Poly c(std::max(left[p].size(), right[p].size()), zero8());
            for (size_t j = 0; j < c.size(); ++j) {
                const W8 a = j < left[p].size() ? left[p][j] : zero8();
                const W8 b = j < right[p].size() ? right[p][j] : zero8();
                c[j] = add8(a, b);
            }
            trim(c);
            next.push_back(std::move(c));
        }
        if (polys.size() % 2) next.push_back(polys.back());
        polys = std::move(next);
    }
    Poly out = polys.back();
    trim(out);
    return out;
}

void add_stats(PolyEngine& e, const DevTreeStats& st) {
    e.stats_.rounds += st.rounds;
    e.stats_.products += st.products;
    e.stats_.device_products += st.products;
}

PolyTree* tree_build(PolyEngine& e, Owner* o, const std::vector<W8>& pts) {
    auto t = std::make_unique<PolyTree>();
    t->owner = o;
    t->points = pts;
    std::unordered_set<std::string> seen;
    for (const W8& p : pts)
        require(seen.insert(std::string(reinterpret_cast<const char*>(p.data()), 32)).second,
                "duplicate interpolation point");
    if (pts.size() >= DEV_TREE_MIN_POINTS) {
        DevTreeStats st;
        t->dev = dev_tree_build(o, pts[0].data(), pts.size(), st);
        add_stats(e, st);
        return t.release();
    }
    std::vector<Poly> leaves;
    const W8 one = from_fr(fp_one<FrParams>());
    for (const W8& p : pts) leaves.push_back(Poly{neg8(p), one});
    if (leaves.empty()) leaves.push_back(Poly{one});
    t->levels.push_back(std::move(leaves));
    while (t->levels.back().size() > 1) {
        const std::vector<Poly>& prev = t->levels.back();
        std::vector<Poly> next(prev.size() / 2 + prev.size() % 2);
        std::vector<MulJob> jobs;
        for (size_t p = 0; p + 1 < prev.size(); p += 2) jobs.push_back({&prev[p], &prev[p + 1], &next[p / 2]});
        e.mul(jobs);
        if (prev.size() % 2) next.back() = prev.back();
        for (Poly& q : next) trim(q);
        t->levels.push_back(std::move(next));
    }
    // reciprocals for the remainder tree: a node divides its parent's
    // remainder (degree < deg parent), quotient length deg(parent) - deg(node).
    // They are independent, so every level's Newton steps run in lockstep
    // (rounds set by the deepest level, not summed over the levels).
    t->recip.resize(t->levels.size());
    {
        std::vector<const Poly*> dens;
        std::vector<size_t> lens, level_start;
        for (size_t l = 0; l + 1 < t->levels.size(); ++l) {
            level_start.push_back(dens.size());
            for (size_t i = 0; i < t->levels[l].size(); ++i) {
                const size_t dp = t->levels[l + 1][i / 2].size() - 1, dn = t->levels[l][i].size() - 1;
                dens.push_back(&t->levels[l][i]);
                lens.push_back(std::max<size_t>(dp > dn ? dp - dn : 0, 1));
            }
        }
        level_start.push_back(dens.size());
        std::vector<Poly> all = e.reciprocals(dens, lens);
        for (size_t l = 0; l + 1 < t->levels.size(); ++l)
            t->recip[l].assign(std::make_move_iterator(all.begin() + level_start[l]),
                               std::make_move_iterator(all.begin() + level_start[l + 1]));
    }
    if (!pts.empty()) {
        const Poly d = differentiate(t->levels.back()[0]);
        t->derivatives = tree_eval(e, *t, d);
        // batch inversion (prefix products, one inversion, back substitution)
        const size_t k = t->derivatives.size();
        std::vector<Fr> pre(k);
        Fr acc = fp_one<FrParams>();
        for (size_t i = 0; i < k; ++i) {
            require(!is_zero8(t->derivatives[i]), "repeated interpolation root");
            pre[i] = acc;
            acc = mul(acc, to_fr(t->derivatives[i]));
        }
        Fr inv_acc = bn254::inv(acc);
        t->inverse_derivatives.assign(k, zero8());
        for (size_t i = k; i-- > 0;) {
            t->inverse_derivatives[i] = from_fr(mul(inv_acc, pre[i]));
            inv_acc = mul(inv_acc, to_fr(t->derivatives[i]));
        }
    }
    return t.release();
}

// Values of p at a device tree's points; p of degree >= k is first reduced
// modulo the vanishing polynomial (the top division of fast.rs:73-84).
std::vector<W8> dev_eval(PolyEngine& e, const PolyTree& t, Poly p) {
    const size_t k = t.points.size();
    if (p.size() > k) {
        Poly A(k + 1);
        dev_tree_vanishing(t.dev, A[0].data());
        std::vector<const Poly*> num{&p}, den{&A};
        std::vector<Poly> rem;
        e.div_rem(num, den, nullptr, nullptr, rem);
        p = std::move(rem[0]);
    }
    std::vector<W8> v(k);
    const uint32_t* pp = p[0].data();
    const size_t len = p.size();
    uint32_t* vp = v[0].data();
    DevTreeStats st;
    dev_tree_eval(t.dev, &pp, &len, 1, &vp, st);
    add_stats(e, st);
    return v;
}

template <class Body>
int poly_call(void* backend, const uint8_t* root, DvgpuPolyStats* stats, char* error, size_t capacity, Body&& body) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        prepare_domain(o, root);
        PolyEngine e(o);
        body(o, e);
        if (stats) *stats = e.stats_;
        return 0;
    });
}

}  // namespace
}  // namespace dvgpu

using namespace dvgpu;

extern "C" {

int dvgpu_poly_mul(void* backend, const uint8_t* root, const uint8_t* left, size_t left_len, const uint8_t* right,
                   size_t right_len, uint8_t* out, size_t out_capacity, size_t* out_len, char* error,
                   size_t capacity) {
    return poly_call(backend, root, nullptr, error, capacity, [&](Owner*, PolyEngine& e) {
        require((left || !left_len) && (right || !right_len), "null polynomial");
        const Poly a = read_poly(left, left_len), b = read_poly(right, right_len);
        Poly r;
        std::vector<MulJob> jobs{{&a, &b, &r}};
        e.mul(jobs);
        trim(r);
        write_poly(r, out, out_capacity, out_len);
    });
}

int dvgpu_poly_divide_exact(void* backend, const uint8_t* root, const uint8_t* numerator, size_t num_len,
                            const uint8_t* denominator, size_t den_len, uint8_t* out, size_t out_capacity,
                            size_t* out_len, char* error, size_t capacity) {
    return poly_call(backend, root, nullptr, error, capacity, [&](Owner*, PolyEngine& e) {
        require((numerator || !num_len) && (denominator || !den_len), "null polynomial");
        Poly num = read_poly(numerator, num_len), den = read_poly(denominator, den_len);
        trim(num);
        trim(den);
        require(!is_zero8(den.back()), "cannot divide by zero polynomial");
        Poly q{zero8()};
        bool all_zero = true;
        for (const W8& c : num) all_zero = all_zero && is_zero8(c);
        if (num.size() >= den.size() && !all_zero) {
            std::vector<const Poly*> n{&num}, d{&den};
            std::vector<Poly> quot, rem;
            e.div_rem(n, d, nullptr, &quot, rem);
            for (const W8& c : rem[0]) require(is_zero8(c), "polynomial division has a nonzero remainder");
            q = quot[0];
            trim(q);
        } else {
            require(all_zero, "polynomial division has a nonzero remainder");
        }
        write_poly(q, out, out_capacity, out_len);
    });
}

int dvgpu_poly_tree_create(void* backend, const uint8_t* root, const uint8_t* points, size_t count, void** tree,
                           DvgpuPolyStats* stats, char* error, size_t capacity) {
    return poly_call(backend, root, stats, error, capacity, [&](Owner* o, PolyEngine& e) {
        require(tree && (points || !count), "invalid tree arguments");
        require(count <= (size_t(1) << 19), "too many interpolation points");
        const Poly pts = read_poly(points, count);
        PolyTree* t = tree_build(e, o, pts);
        o->poly_trees.push_back(t);
        *tree = t;
    });
}

int dvgpu_poly_tree_destroy(void* backend, void* tree) {
    Owner* o = as_owner(backend);
    if (!o) return -1;
    auto it = std::find(o->poly_trees.begin(), o->poly_trees.end(), static_cast<PolyTree*>(tree));
    if (it == o->poly_trees.end()) return -1;
    PolyTree* t = *it;
    o->poly_trees.erase(it);
    poly_tree_destroy(t);
    return 0;
}

int dvgpu_poly_tree_vanishing(void* backend, void* tree, uint8_t* out, size_t out_capacity, size_t* out_len,
                              char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        PolyTree* t = find_tree(o, tree);
        if (t->dev) {
            Poly v(t->points.size() + 1);
            dev_tree_vanishing(t->dev, v[0].data());
            write_poly(v, out, out_capacity, out_len);
            return 0;
        }
        write_poly(t->levels.back()[0], out, out_capacity, out_len);
        return 0;
    });
}

int dvgpu_poly_tree_inverse_derivatives(void* backend, void* tree, uint8_t* out, size_t out_capacity,
                                        size_t* out_len, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        PolyTree* t = find_tree(o, tree);
        if (t->dev) {
            require(out_len && out, "null output");
            require(t->points.size() <= out_capacity, "output capacity too small");
            Poly v(t->points.size());
            dev_tree_inverse_derivatives(t->dev, v[0].data());
            std::memcpy(out, v.data(), v.size() * 32);
            *out_len = v.size();
            return 0;
        }
        require(out_len && (out || t->inverse_derivatives.empty()), "null output");
        require(t->inverse_derivatives.size() <= out_capacity, "output capacity too small");
        for (size_t i = 0; i < t->inverse_derivatives.size(); ++i)
            std::memcpy(out + 32 * i, t->inverse_derivatives[i].data(), 32);
        *out_len = t->inverse_derivatives.size();
        return 0;
    });
}

int dvgpu_poly_tree_eval(void* backend, const uint8_t* root, void* tree, const uint8_t* poly, size_t len,
                         uint8_t* values, DvgpuPolyStats* stats, char* error, size_t capacity) {
    return poly_call(backend, root, stats, error, capacity, [&](Owner* o, PolyEngine& e) {
        PolyTree* t = find_tree(o, tree);
        require(poly || !len, "null polynomial");
        require(values || t->points.empty(), "null values");
        Poly p = read_poly(poly, len);
        trim(p);
        const std::vector<W8> v = t->dev ? dev_eval(e, *t, p) : tree_eval(e, *t, p);
        for (size_t i = 0; i < v.size(); ++i) std::memcpy(values + 32 * i, v[i].data(), 32);
    });
}

int dvgpu_poly_tree_interpolate(void* backend, const uint8_t* root, void* tree, const uint8_t* values, size_t count,
                                uint8_t* out, size_t out_capacity, size_t* out_len, DvgpuPolyStats* stats,
                                char* error, size_t capacity) {
    return poly_call(backend, root, stats, error, capacity, [&](Owner* o, PolyEngine& e) {
        PolyTree* t = find_tree(o, tree);
        require(values || !count, "null values");
        const Poly v = read_poly(values, count);
        if (t->dev) {
            require(v.size() == t->points.size(), "interpolation value count differs from points");
            Poly c(v.size());
            DevTreeStats st;
            dev_tree_interpolate(t->dev, v[0].data(), c[0].data(), st);
            add_stats(e, st);
            trim(c);
            write_poly(c, out, out_capacity, out_len);
            return;
        }
        write_poly(tree_interpolate(e, *t, v), out, out_capacity, out_len);
    });
}

// fast.rs:118-139: sum_{j != i} w_j / (x_i - x_j) = (N'(x_i) - w_i A''(x_i) / 2) / A'(x_i)
// with N the interpolant of w_j A'(x_j).
int dvgpu_poly_tree_cauchy_sums(void* backend, const uint8_t* root, void* tree, const uint8_t* weights,
                                size_t count, uint8_t* sums, DvgpuPolyStats* stats, char* error, size_t capacity) {
    return poly_call(backend, root, stats, error, capacity, [&](Owner* o, PolyEngine& e) {
        PolyTree* t = find_tree(o, tree);
        require(count == t->points.size(), "Cauchy weight count differs from points");
        require(weights || !count, "null weights");
        require(sums || !count, "null sums");
        if (!count) return;
        const Poly w = read_poly(weights, count);
        if (t->dev) {
            DevTreeStats st;
            dev_tree_cauchy(t->dev, w[0].data(), reinterpret_cast<uint32_t*>(sums), st);
            add_stats(e, st);
            return;
        }
        std::vector<W8> nv;
        for (size_t i = 0; i < count; ++i) nv.push_back(from_fr(mul(to_fr(w[i]), to_fr(t->derivatives[i]))));
        const Poly num = tree_interpolate(e, *t, nv);
        const Poly d1 = differentiate(num), d2 = differentiate(differentiate(t->levels.back()[0]));
        const std::vector<std::vector<W8>> ev = tree_eval_many(e, *t, {&d1, &d2});  // both in lockstep
        const std::vector<W8>& first = ev[0];
        const std::vector<W8>& second = ev[1];
        const Fr half = bn254::inv(add(fp_one<FrParams>(), fp_one<FrParams>()));
        for (size_t i = 0; i < count; ++i) {
            const Fr s = mul(sub(to_fr(first[i]), mul(mul(to_fr(w[i]), to_fr(second[i])), half)),
                             to_fr(t->inverse_derivatives[i]));
            const W8 o8 = from_fr(s);
            std::memcpy(sums + 32 * i, o8.data(), 32);
        }
    });
}

}  // extern "C"
