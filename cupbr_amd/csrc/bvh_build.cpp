#include "bvh_build.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace cupbr {
namespace {

struct BNode {
    float mn[3], mx[3];
    uint32_t left, right;  // inner
    uint32_t first, count; // leaf when count > 0
};

inline float area(const float* mn, const float* mx) {
    float ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
    if (!(ex >= 0.0f) || !(ey >= 0.0f) || !(ez >= 0.0f)) return 0.0f;
    return 2.0f * (ex * ey + ey * ez + ez * ex);
}

struct Bin {
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t n = 0;
    void grow(const BuildPrim& p) {
        for (int k = 0; k < 3; ++k) { mn[k] = std::min(mn[k], p.mn[k]); mx[k] = std::max(mx[k], p.mx[k]); }
        ++n;
    }
    void merge(const Bin& b) {
        for (int k = 0; k < 3; ++k) { mn[k] = std::min(mn[k], b.mn[k]); mx[k] = std::max(mx[k], b.mx[k]); }
        n += b.n;
    }
};

class BinaryBuilder {
public:
    BinaryBuilder(const std::vector<BuildPrim>& p, const BvhBuildOptions& o) : P(p), opt(o) {
        size_t n = P.size();
        cen.resize(n * 3);
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) cen[3 * i + k] = 0.5f * (P[i].mn[k] + P[i].mx[k]);
        idx.resize(n);
        for (size_t i = 0; i < n; ++i) idx[i] = (uint32_t)i;
        nodes.resize(std::max<size_t>(1, 2 * n));
        max_threads = opt.threads > 0 ? opt.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    }

    uint32_t run() {
        uint32_t root = alloc();
        build(root, 0, (uint32_t)P.size(), 0);
        return root;
    }

    const std::vector<BuildPrim>& P;
    BvhBuildOptions opt;
    std::vector<float> cen;
    std::vector<uint32_t> idx;
    std::vector<BNode> nodes;

private:
    std::atomic<uint32_t> next{0};
    std::atomic<int> active{1};
    int max_threads = 1;

    uint32_t alloc() { return next.fetch_add(1); }

    void build(uint32_t me, uint32_t lo, uint32_t hi, int depth) {
        BNode& node = nodes[me];
        Bin all, cb;
        for (uint32_t i = lo; i < hi; ++i) {
            all.grow(P[idx[i]]);
            BuildPrim c;
            for (int k = 0; k < 3; ++k) c.mn[k] = c.mx[k] = cen[3 * idx[i] + k];
            cb.grow(c);
        }
        for (int k = 0; k < 3; ++k) { node.mn[k] = all.mn[k]; node.mx[k] = all.mx[k]; }
        const uint32_t n = hi - lo;
        const float leaf_cost = opt.intersect_cost * (float)n;
        if (n <= 1) { make_leaf(node, lo, n); return; }

        constexpr int NB = CUPBR_BVH_BINS;
        int best_axis = -1, best_bin = -1;
        float best_cost = INFINITY;
        const float parent_area = std::max(area(all.mn, all.mx), 1e-30f);
        for (int a = 0; a < 3; ++a) {
            const float ext = cb.mx[a] - cb.mn[a];
            if (!(ext > 0.0f)) continue;
            const float scale = (float)NB / ext;
            Bin bins[NB];
            for (uint32_t i = lo; i < hi; ++i) {
                int b = (int)((cen[3 * idx[i] + a] - cb.mn[a]) * scale);
                b = std::min(NB - 1, std::max(0, b));
                bins[b].grow(P[idx[i]]);
            }
            Bin right[NB];
            Bin acc;
            for (int b = NB - 1; b > 0; --b) { acc.merge(bins[b]); right[b] = acc; }
            Bin left;
            for (int b = 0; b < NB - 1; ++b) {
                left.merge(bins[b]);
                const Bin& r = right[b + 1];
                if (left.n == 0 || r.n == 0) continue;
                float cost = opt.traversal_cost +
                             opt.intersect_cost * (area(left.mn, left.mx) * left.n + area(r.mn, r.mx) * r.n) / parent_area;
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_bin = b; }
            }
        }
        if (n <= opt.max_leaf && leaf_cost <= best_cost) { make_leaf(node, lo, n); return; }

        uint32_t mid = lo + n / 2;
        if (best_axis >= 0 && depth < 48) {
            const float ext = cb.mx[best_axis] - cb.mn[best_axis];
            const float scale = (float)NB / ext;
            const float lo_c = cb.mn[best_axis];
            auto it = std::partition(idx.begin() + lo, idx.begin() + hi, [&](uint32_t p) {
                int b = (int)((cen[3 * p + best_axis] - lo_c) * scale);
                b = std::min(NB - 1, std::max(0, b));
                return b <= best_bin;
            });
            uint32_t m = (uint32_t)(it - idx.begin());
            if (m > lo && m < hi) mid = m;
        } else {
            // no usable split (coincident centroids) or too deep: median along the widest axis
            int a = 0;
            float best_ext = -1.0f;
            for (int k = 0; k < 3; ++k)
                if (cb.mx[k] - cb.mn[k] > best_ext) { best_ext = cb.mx[k] - cb.mn[k]; a = k; }
            std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi,
                             [&](uint32_t x, uint32_t y) { return cen[3 * x + a] < cen[3 * y + a]; });
        }
        const uint32_t l = alloc(), r = alloc();
        node.left = l;
        node.right = r;
        node.count = 0;
        if (n > 16384 && active.load() < max_threads) {
            active.fetch_add(1);
            std::thread th([=]() { build(l, lo, mid, depth + 1); });
            build(r, mid, hi, depth + 1);
            th.join();
            active.fetch_sub(1);
        } else {
            build(l, lo, mid, depth + 1);
            build(r, mid, hi, depth + 1);
        }
    }

    void make_leaf(BNode& node, uint32_t lo, uint32_t n) {
        node.first = lo;
        node.count = std::max<uint32_t>(n, 1);
        node.left = node.right = 0;
    }
};

// SAH-optimal collapse of the binary tree into a BVH4 (the dynamic program of Ylitie, Karras and
// Laine 2017, "Efficient incoherent ray traversal on GPUs through compressed wide BVHs", for 4
// children). Cost of a tree, in units of the root's surface area: c_node x area of every BVH4 node
// (a visit tests its children's boxes) + c_prim x area x primitives of every leaf. For binary node n,
// G[n][j-1] is the least cost of representing its subtree as at most j subtrees under one parent:
//   G[n][0] = min(leaf(n) if n holds <= max_leaf primitives, c_node A(n) + S(n, 4))
//   G[n][j-1] = min(G[n][0], S(n, j)) for j = 2..4,  S(n, j) = min_k G[l][k-1] + G[r][j-k-1]
// where S distributes j slots between the two binary children.
struct Collapser {
    const std::vector<BNode>& B;
    Bvh4& out;
    uint32_t max_leaf;
    double c_node, c_prim;
    double root_area = 1.0;
    std::vector<std::array<double, 4>> G;
    std::vector<std::array<uint8_t, 4>> split;  // slots given to the left child (0: keep whole)
    std::vector<uint8_t> as_leaf;
    std::vector<uint32_t> first, count;

    double a(uint32_t n) const { return (double)area(B[n].mn, B[n].mx) / root_area; }

    // post-order over the binary tree (iterative: degenerate inputs can make it deep)
    void solve(uint32_t root) {
        const size_t nn = B.size();
        G.assign(nn, {INFINITY, INFINITY, INFINITY, INFINITY});
        split.assign(nn, {0, 0, 0, 0});
        as_leaf.assign(nn, 0);
        first.assign(nn, 0);
        count.assign(nn, 0);
        std::vector<std::pair<uint32_t, bool>> st{{root, false}};
        while (!st.empty()) {
            auto [n, done] = st.back();
            st.pop_back();
            const BNode& b = B[n];
            if (!b.count && !done) {
                st.push_back({n, true});
                st.push_back({b.right, false});
                st.push_back({b.left, false});
                continue;
            }
            if (b.count) {
                first[n] = b.first;
                count[n] = b.count;
            } else {
                first[n] = first[b.left];
                count[n] = count[b.left] + count[b.right];
            }
            const double leaf = count[n] <= max_leaf ? c_prim * a(n) * (double)count[n] : INFINITY;
            if (b.count) {
                G[n] = {leaf, leaf, leaf, leaf};
                as_leaf[n] = 1;
                continue;
            }
            double S[5];
            uint8_t K[5] = {0, 0, 0, 0, 0};
            for (int j = 2; j <= 4; ++j) {
                S[j] = INFINITY;
                for (int k = 1; k < j; ++k) {
                    const double c = G[b.left][k - 1] + G[b.right][j - k - 1];
                    if (c < S[j]) { S[j] = c; K[j] = (uint8_t)k; }
                }
            }
            const double inner = c_node * a(n) + S[4];
            as_leaf[n] = leaf <= inner;
            G[n][0] = std::min(leaf, inner);
            split[n][0] = 0;
            for (int j = 2; j <= 4; ++j) {
                if (S[j] < G[n][0]) { G[n][j - 1] = S[j]; split[n][j - 1] = K[j]; }
                else { G[n][j - 1] = G[n][0]; split[n][j - 1] = 0; }
            }
        }
    }

    // the subtrees that fill j slots of a parent
    void slots(uint32_t n, int j, std::vector<uint32_t>& outv) const {
        const uint8_t k = j > 1 ? split[n][j - 1] : 0;
        if (k == 0) { outv.push_back(n); return; }
        slots(B[n].left, k, outv);
        slots(B[n].right, j - k, outv);
    }

    uint32_t emit(uint32_t b, uint32_t depth) {
        out.depth = std::max(out.depth, depth + 1);
        std::vector<uint32_t> kids;
        const BNode& bn = B[b];
        if (bn.count) kids.push_back(b);  // a leaf root gets a wrapper node
        else {
            slots(bn.left, split_root_left(b), kids);
            slots(bn.right, 4 - split_root_left(b), kids);
        }
        const uint32_t me = (uint32_t)out.nodes.size();
        out.nodes.emplace_back();
        out.sah_cost += c_node * a(b);
        Node4 n;
        for (int i = 0; i < 4; ++i) {
            n.lo_x[i] = n.lo_y[i] = n.lo_z[i] = INFINITY;
            n.hi_x[i] = n.hi_y[i] = n.hi_z[i] = -INFINITY;
            n.child[i] = kChildEmpty;
            n.spare[i] = 0;
        }
        for (size_t i = 0; i < kids.size(); ++i) {
            const BNode& k = B[kids[i]];
            n.lo_x[i] = k.mn[0]; n.lo_y[i] = k.mn[1]; n.lo_z[i] = k.mn[2];
            n.hi_x[i] = k.mx[0]; n.hi_y[i] = k.mx[1]; n.hi_z[i] = k.mx[2];
            if (as_leaf[kids[i]]) {
                n.child[i] = leaf_word(first[kids[i]], count[kids[i]]);
                out.leaves++;
                out.sah_cost += c_prim * a(kids[i]) * (double)count[kids[i]];
            } else {
                n.child[i] = emit(kids[i], depth + 1);
            }
        }
        out.nodes[me] = n;
        return me;
    }

    // slots of a BVH4 node at binary node b given to its left binary child
    int split_root_left(uint32_t b) const {
        const BNode& bn = B[b];
        int best_k = 1;
        double best = INFINITY;
        for (int k = 1; k < 4; ++k) {
            const double c = G[bn.left][k - 1] + G[bn.right][4 - k - 1];
            if (c < best) { best = c; best_k = k; }
        }
        return best_k;
    }
};

}  // namespace

namespace {

uint16_t f16_bits(_Float16 h) {
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}
_Float16 f16_of(uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return h;
}
// the binary16 value next to b towards +inf (up) or -inf
uint16_t f16_step(uint16_t b, bool up) {
    const bool neg = b & 0x8000u;
    if ((b & 0x7FFFu) == 0) return up ? 0x0001u : 0x8001u;  // +-0: smallest subnormal
    return (neg == up) ? (uint16_t)(b - 1u) : (uint16_t)(b + 1u);
}
// binary16 rounding of x towards +inf (up) or -inf
uint16_t f16_directed(double x, bool up) {
    uint16_t b = f16_bits((_Float16)x);  // nearest (overflow to +-inf)
    const double d = (double)f16_of(b);
    if (up ? d < x : d > x) b = f16_step(b, up);
    return b;
}

}  // namespace

std::vector<NodeH> half_nodes(const std::vector<Node4>& nodes) {
    constexpr uint16_t kPosInf = 0x7C00u, kNegInf = 0xFC00u;
    std::vector<NodeH> out(nodes.size());
    for (size_t n = 0; n < nodes.size(); ++n) {
        const Node4& a = nodes[n];
        NodeH h{};
        const float* lo[3] = {a.lo_x, a.lo_y, a.lo_z};
        const float* hi[3] = {a.hi_x, a.hi_y, a.hi_z};
        uint16_t(*dst[3])[8] = {h.x, h.y, h.z};
        float c[3];
        uint32_t exps = 0;
        for (int ax = 0; ax < 3; ++ax) {
            double cmin = INFINITY, hmax = -INFINITY;
            for (int i = 0; i < 4; ++i) {
                if (std::isfinite(lo[ax][i])) cmin = std::min(cmin, (double)lo[ax][i]);
                if (std::isfinite(hi[ax][i])) hmax = std::max(hmax, (double)hi[ax][i]);
            }
            if (!std::isfinite(cmin)) cmin = 0.0;
            const double e = std::isfinite(hmax) ? std::max(0.0, hmax - cmin) : 0.0;
            // the reference point: a float near the middle, so that offsets reach only about e / 2
            // either way (binary16 keeps 11 bits of each)
            c[ax] = (float)(cmin + 0.5 * e);
            const double cref = (double)c[ax];
            // |offsets| up to about e / 2^(k+1) <= 2^15, well inside binary16 (k capped at 0: a
            // wider node rounds its outer offsets to the largest finite value or to infinity,
            // which only loosens it)
            int k = e > 0.0 ? (int)std::ceil(std::log2(e)) - 16 : 0;
            k = std::min(0, std::max(-100, k));
            const double scale = std::ldexp(1.0, -k), margin = 16.0 * std::ldexp(e, -24);
            exps |= (uint32_t)(uint8_t)(int8_t)k << (8 * ax);
            for (int i = 0; i < 4; ++i) {
                const float l = lo[ax][i], u = hi[ax][i];
                uint16_t ql, qu;
                if (std::isnan(l)) ql = kNegInf;
                else if (std::isinf(l)) ql = l > 0 ? kPosInf : kNegInf;
                else ql = f16_directed(((double)l - cref - margin) * scale, false);
                if (std::isnan(u)) qu = kPosInf;
                else if (std::isinf(u)) qu = u > 0 ? kPosInf : kNegInf;
                else qu = f16_directed(((double)u - cref + margin) * scale, true);
                dst[ax][0][i] = ql;
                dst[ax][0][4 + i] = qu;
                dst[ax][1][i] = qu;
                dst[ax][1][4 + i] = ql;
            }
        }
        h.cx = c[0];
        h.cy = c[1];
        h.cz = c[2];
        h.exps = exps;
        for (int i = 0; i < 4; ++i) h.child[i] = a.child[i];
        out[n] = h;
    }
    return out;
}

Bvh4 build_bvh4(const std::vector<BuildPrim>& prims, const BvhBuildOptions& opt_in) {
    Bvh4 out;
    if (prims.empty()) return out;
    BvhBuildOptions opt = opt_in;
    opt.max_leaf = std::min<uint32_t>(std::max<uint32_t>(opt.max_leaf, 1), kMaxLeafPrims);
    // the binary tree goes down to single primitives; the collapse chooses the leaves
    BvhBuildOptions bopt = opt;
    bopt.max_leaf = 1;
    BinaryBuilder bb(prims, bopt);
    uint32_t root = bb.run();
    Collapser c{bb.nodes, out, opt.max_leaf, opt.traversal_cost, opt.intersect_cost};
    c.root_area = std::max(1e-30f, area(bb.nodes[root].mn, bb.nodes[root].mx));
    c.solve(root);
    out.nodes.reserve(prims.size() / 2 + 16);
    c.emit(root, 0);
    out.order = std::move(bb.idx);
    return out;
}

}  // namespace cupbr
