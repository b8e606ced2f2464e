// Binned-SAH binary build over FP64 bounds, collapsed to 4-wide 128-byte nodes.
#include "bvh_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <thread>

namespace phg {
namespace {

struct Box {
    double lo[3] = {INFINITY, INFINITY, INFINITY};
    double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void grow(const double* p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    double area() const {
        const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    Box box;
    uint32_t left = 0, right = 0;  // child node indices (inner)
    uint32_t first = 0, count = 0; // triangle range (leaf when count > 0)
};

struct Builder {
    const BvhBuildParams& p;
    std::vector<Box> tbox;
    std::vector<double> cen;  // 3 per triangle
    std::vector<uint32_t> idx;
    std::vector<BNode> nodes;
    std::atomic<uint32_t> next_node{0};
    std::atomic<int> live_threads{1};
    int max_threads = 1;

    explicit Builder(const BvhBuildParams& params) : p(params) {}

    uint32_t alloc() {
        const uint32_t i = next_node.fetch_add(1, std::memory_order_relaxed);
        if (i >= nodes.size()) throw std::runtime_error("bvh: node pool exhausted");
        return i;
    }

    // Build the subtree over idx[lo, hi) into node ni.
    void build(uint32_t ni, uint32_t lo, uint32_t hi, uint32_t depth) {
        BNode& nd = nodes[ni];
        Box b, cb;
        for (uint32_t i = lo; i < hi; ++i) { b.grow(tbox[idx[i]]); cb.grow(&cen[3 * (size_t)idx[i]]); }
        nd.box = b;
        const uint32_t n = hi - lo;
        if (n == 1) { nd.first = lo; nd.count = 1; return; }

        // binned SAH over the centroid bounds
        const uint32_t B = p.sah_bins;
        int best_axis = -1; uint32_t best_split = 0; double best_cost = INFINITY;
        if (depth < 48) {
            for (int ax = 0; ax < 3; ++ax) {
                const double ext = cb.hi[ax] - cb.lo[ax];
                if (!(ext > 0.0)) continue;
                Box bb[64]; uint32_t bc[64] = {0};
                const double scale = (double)B / ext;
                for (uint32_t i = lo; i < hi; ++i) {
                    const uint32_t t = idx[i];
                    uint32_t k = (uint32_t)((cen[3 * (size_t)t + ax] - cb.lo[ax]) * scale);
                    if (k >= B) k = B - 1;
                    bb[k].grow(tbox[t]); bc[k]++;
                }
                double ra[64]; uint32_t rc[64];
                Box acc; uint32_t cnt = 0;
                for (int k = (int)B - 1; k > 0; --k) { acc.grow(bb[k]); cnt += bc[k]; ra[k] = acc.area(); rc[k] = cnt; }
                acc = Box(); cnt = 0;
                for (uint32_t k = 0; k + 1 < B; ++k) {
                    acc.grow(bb[k]); cnt += bc[k];
                    if (cnt == 0 || rc[k + 1] == 0) continue;
                    const double c = acc.area() * cnt + ra[k + 1] * rc[k + 1];
                    if (c < best_cost) { best_cost = c; best_axis = ax; best_split = k; }
                }
            }
        }
        const double area = b.area();
        const double leaf_cost = (double)n;
        const double split_cost = area > 0 ? 1.0 + best_cost / area : INFINITY;
        if (n <= p.max_leaf_tris && (best_axis < 0 || leaf_cost <= split_cost)) {
            nd.first = lo; nd.count = n; return;
        }

        uint32_t mid;
        if (best_axis >= 0) {
            const int ax = best_axis;
            const double ext = cb.hi[ax] - cb.lo[ax];
            const double scale = (double)B / ext;
            auto it = std::partition(idx.begin() + lo, idx.begin() + hi, [&](uint32_t t) {
                uint32_t k = (uint32_t)((cen[3 * (size_t)t + ax] - cb.lo[ax]) * scale);
                if (k >= B) k = B - 1;
                return k <= best_split;
            });
            mid = (uint32_t)(it - idx.begin());
        } else {
            mid = lo;
        }
        if (mid == lo || mid == hi) {
            // no usable SAH plane (coincident centroids or depth limit): median on the widest axis
            int ax = 0;
            const double e[3] = {cb.hi[0] - cb.lo[0], cb.hi[1] - cb.lo[1], cb.hi[2] - cb.lo[2]};
            if (e[1] > e[ax]) ax = 1;
            if (e[2] > e[ax]) ax = 2;
            mid = lo + n / 2;
            std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi,
                             [&](uint32_t a, uint32_t c) { return cen[3 * (size_t)a + ax] < cen[3 * (size_t)c + ax]; });
        }
        const uint32_t l = alloc(), r = alloc();
        nd.left = l; nd.right = r; nd.count = 0;
        const bool spawn = (mid - lo) > 65536 && (hi - mid) > 65536 && live_threads.load() < max_threads;
        if (spawn) {
            live_threads.fetch_add(1);
            std::thread th([&, l, lo, mid, depth] { build(l, lo, mid, depth + 1); live_threads.fetch_sub(1); });
            build(r, mid, hi, depth + 1);
            th.join();
        } else {
            build(l, lo, mid, depth + 1);
            build(r, mid, hi, depth + 1);
        }
    }
};

// Round an FP64 bound outward to FP32 (exact for bounds that are unions of FP32 vertices).
inline float down_f(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}
inline float up_f(double v) {
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, INFINITY);
    return f;
}

struct Collapser {
    const std::vector<BNode>& bn;
    std::vector<WideNode>& out;
    uint32_t width = 4;
    uint32_t depth = 0;

    // Emit the wide node for binary inner node bi at index wi; returns its stack need.
    uint32_t emit(uint32_t bi, uint32_t wi, uint32_t level) {
        depth = std::max(depth, level);
        uint32_t kids[4] = {bn[bi].left, bn[bi].right, 0, 0};
        uint32_t nk = 2;
        while (nk < width) {
            int pick = -1; double best = -1.0;
            for (uint32_t i = 0; i < nk; ++i) {
                const BNode& c = bn[kids[i]];
                if (c.count == 0 && c.box.area() > best) { best = c.box.area(); pick = (int)i; }
            }
            if (pick < 0) break;
            const uint32_t e = kids[pick];
            kids[pick] = bn[e].left;
            kids[nk++] = bn[e].right;
        }
        WideNode w{};
        for (int k = 0; k < 4; ++k) {
            w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = INFINITY;
            w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = -INFINITY;
            w.child[k] = kEmptyChild;
        }
        uint32_t inner_idx[4]; uint32_t inner_slot[4]; uint32_t ni = 0;
        for (uint32_t k = 0; k < nk; ++k) {
            const BNode& c = bn[kids[k]];
            w.lo_x[k] = down_f(c.box.lo[0]); w.hi_x[k] = up_f(c.box.hi[0]);
            w.lo_y[k] = down_f(c.box.lo[1]); w.hi_y[k] = up_f(c.box.hi[1]);
            w.lo_z[k] = down_f(c.box.lo[2]); w.hi_z[k] = up_f(c.box.hi[2]);
            if (c.count > 0) {
                w.child[k] = leaf_ref(c.first, c.count);
            } else {
                inner_idx[ni] = kids[k]; inner_slot[ni] = k; ++ni;
            }
        }
        out[wi] = w;
        uint32_t deeper = 0;
        for (uint32_t i = 0; i < ni; ++i) {
            const uint32_t cw = (uint32_t)out.size();
            out.emplace_back();
            out[wi].child[inner_slot[i]] = (int32_t)cw;
            deeper = std::max(deeper, emit(inner_idx[i], cw, level + 1));
        }
        // pushing all but one child here, plus whatever the deepest child subtree needs
        return (nk - 1) + deeper;
    }
};

}  // namespace

BvhBuildResult build_wide_bvh(const float* verts, uint32_t ntri, const BvhBuildParams& params) {
    if (params.max_leaf_tris < 1 || params.max_leaf_tris > kMaxLeafTris)
        throw std::runtime_error("bvh: max_leaf_tris must be in [1, 8]");
    if (params.sah_bins < 2 || params.sah_bins > 64) throw std::runtime_error("bvh: sah_bins must be in [2, 64]");
    if (params.width < 2 || params.width > 4) throw std::runtime_error("bvh: width must be in [2, 4]");
    if (ntri >= (1u << kLeafFirstBits)) throw std::runtime_error("bvh: too many triangles");
    BvhBuildResult res;
    if (ntri == 0) return res;

    Builder b(params);
    b.max_threads = params.nthreads ? (int)params.nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    b.tbox.resize(ntri);
    b.cen.resize(3 * (size_t)ntri);
    b.idx.resize(ntri);
    for (uint32_t t = 0; t < ntri; ++t) {
        Box bx;
        for (int v = 0; v < 3; ++v) {
            const double p[3] = {verts[9 * (size_t)t + 3 * v], verts[9 * (size_t)t + 3 * v + 1], verts[9 * (size_t)t + 3 * v + 2]};
            bx.grow(p);
        }
        b.tbox[t] = bx;
        for (int k = 0; k < 3; ++k) b.cen[3 * (size_t)t + k] = 0.5 * (bx.lo[k] + bx.hi[k]);
        b.idx[t] = t;
    }
    b.nodes.resize(2 * (size_t)ntri);
    const uint32_t root = b.alloc();
    b.build(root, 0, ntri, 0);
    res.binary_nodes = b.next_node.load();

    res.tri_order = b.idx;
    res.nodes.reserve(res.binary_nodes / 2 + 2);
    res.nodes.emplace_back();
    Collapser c{b.nodes, res.nodes, params.width};
    if (b.nodes[root].count > 0) {
        // single leaf: a root with one leaf child
        WideNode w{};
        for (int k = 0; k < 4; ++k) {
            w.lo_x[k] = w.lo_y[k] = w.lo_z[k] = INFINITY;
            w.hi_x[k] = w.hi_y[k] = w.hi_z[k] = -INFINITY;
            w.child[k] = kEmptyChild;
        }
        const Box& bx = b.nodes[root].box;
        w.lo_x[0] = down_f(bx.lo[0]); w.hi_x[0] = up_f(bx.hi[0]);
        w.lo_y[0] = down_f(bx.lo[1]); w.hi_y[0] = up_f(bx.hi[1]);
        w.lo_z[0] = down_f(bx.lo[2]); w.hi_z[0] = up_f(bx.hi[2]);
        w.child[0] = leaf_ref(b.nodes[root].first, b.nodes[root].count);
        res.nodes[0] = w;
        res.depth = 1;
        res.stack_need = 1;
    } else {
        res.stack_need = c.emit(root, 0, 1) + 1;
        res.depth = c.depth;
    }
    return res;
}

}  // namespace phg
