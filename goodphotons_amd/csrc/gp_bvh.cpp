// Host BVH builder (see gp_bvh.hpp).
//
// Subtrees are built into preassigned slot ranges (a subtree over m primitives
// needs at most 2m-1 binary nodes), so the top levels run on several threads
// and the result does not depend on the thread count or scheduling.
#include "gp_bvh.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <future>
#include <stdexcept>

#include "gp_types.hpp"

namespace gp {
namespace {

constexpr int kBins = 32;
constexpr double kPad = 1e-6;   // the reference pads primitive boxes the same way

struct Prim {
    float lo[3], hi[3], c[3];
};
struct BNode {
    float lo[3], hi[3];
    int left = -1, right = -1, first = 0, count = 0;
};
struct Box {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float* l, const float* h) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], l[a]);
            hi[a] = std::max(hi[a], h[a]);
        }
    }
    double area() const {
        double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Builder {
    const std::vector<Prim>& P;
    std::vector<int>& idx;
    std::vector<BNode>& nodes;
    int parDepth;

    void build(int slot, int b, int e, int depth) {
        BNode& nd = nodes[slot];
        Box bb, cb;
        for (int i = b; i < e; ++i) {
            const Prim& p = P[idx[i]];
            bb.grow(p.lo, p.hi);
            cb.grow(p.c, p.c);
        }
        for (int a = 0; a < 3; ++a) { nd.lo[a] = bb.lo[a]; nd.hi[a] = bb.hi[a]; }
        const int m = e - b;
        if (m <= kLeafMax) {
            nd.first = b;
            nd.count = m;
            return;
        }
        // Binned SAH over every axis with a non-degenerate centroid extent.
        double bestCost = DBL_MAX;
        int bestAxis = -1, bestSplit = -1;
        for (int a = 0; a < 3; ++a) {
            const float c0 = cb.lo[a], c1 = cb.hi[a];
            if (!(c1 - c0 > 0.0f)) continue;
            const double scale = kBins / ((double)c1 - c0);
            Box bins[kBins];
            int cnt[kBins] = {};
            for (int i = b; i < e; ++i) {
                const Prim& p = P[idx[i]];
                int k = (int)(((double)p.c[a] - c0) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                cnt[k]++;
                bins[k].grow(p.lo, p.hi);
            }
            double rArea[kBins];
            int rCnt[kBins];
            Box acc;
            int rc = 0;
            for (int k = kBins - 1; k >= 1; --k) {
                acc.grow(bins[k].lo, bins[k].hi);
                rc += cnt[k];
                rArea[k] = acc.area();
                rCnt[k] = rc;
            }
            Box lacc;
            int lc = 0;
            for (int k = 0; k < kBins - 1; ++k) {
                lacc.grow(bins[k].lo, bins[k].hi);
                lc += cnt[k];
                if (lc == 0 || rCnt[k + 1] == 0) continue;
                double cost = lc * lacc.area() + rCnt[k + 1] * rArea[k + 1];
                if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = k; }
            }
        }
        int mid = b;
        if (bestAxis >= 0) {
            const int a = bestAxis;
            const float c0 = cb.lo[a], c1 = cb.hi[a];
            const double scale = kBins / ((double)c1 - c0);
            auto it = std::partition(idx.begin() + b, idx.begin() + e, [&](int pi) {
                int k = (int)(((double)P[pi].c[a] - c0) * scale);
                k = std::min(std::max(k, 0), kBins - 1);
                return k <= bestSplit;
            });
            mid = (int)(it - idx.begin());
        }
        if (mid <= b || mid >= e) {
            // No separating plane: median split on the widest centroid axis (or by
            // index when every centroid coincides).
            int a = 0;
            float ext = -1.0f;
            for (int k = 0; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > ext) { ext = cb.hi[k] - cb.lo[k]; a = k; }
            mid = (b + e) / 2;
            std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e, [&](int x, int y) {
                if (P[x].c[a] != P[y].c[a]) return P[x].c[a] < P[y].c[a];
                return x < y;
            });
        }
        const int ml = mid - b;
        const int ls = slot + 1, rs = slot + 2 * ml;
        nd.left = ls;
        nd.right = rs;
        nd.count = 0;
        if (depth < parDepth && m > 65536) {
            auto f = std::async(std::launch::async, [&, ls, b, mid, depth] { build(ls, b, mid, depth + 1); });
            build(rs, mid, e, depth + 1);
            f.get();
        } else {
            build(ls, b, mid, depth + 1);
            build(rs, mid, e, depth + 1);
        }
    }
};

inline float down(double v) {
    float f = (float)v;
    return ((double)f > v) ? std::nextafter(f, -FLT_MAX) : f;
}
inline float up(double v) {
    float f = (float)v;
    return ((double)f < v) ? std::nextafter(f, FLT_MAX) : f;
}

// Child-pair layout in depth-first order from a binary tree (node 0 the root). Each
// pair node holds its children's boxes; a link >= 0 is a pair node, < 0 a leaf code.
void emitPairs(const std::vector<BNode>& nodes, BvhResult& out) {
    out.pairs.clear();
    out.depth = 0;
    out.leaves = 0;
    std::vector<float>& pr = out.pairs;
    auto link = [&](int slot) -> int {
        const BNode& c = nodes[slot];
        return c.count > 0 ? leafCode(c.first, c.count) : -2;   // -2: internal, patched later
    };
    auto emitPair = [&](int lslot, int rslot) -> int {
        const int pi = (int)(pr.size() / 16);
        pr.resize(pr.size() + 16);
        float* q = pr.data() + 16 * (size_t)pi;
        const BNode& L = nodes[lslot];
        const BNode& R = nodes[rslot];
        q[0] = L.lo[0]; q[1] = L.lo[1]; q[2] = L.lo[2];
        q[4] = L.hi[0]; q[5] = L.hi[1]; q[6] = L.hi[2];
        q[8] = R.lo[0]; q[9] = R.lo[1]; q[10] = R.lo[2];
        q[12] = R.hi[0]; q[13] = R.hi[1]; q[14] = R.hi[2];
        int li = link(lslot), ri = link(rslot);
        std::memcpy(&q[3], &li, 4);
        std::memcpy(&q[7], &ri, 4);
        q[11] = 0.0f;
        q[15] = 0.0f;
        return pi;
    };
    const BNode& root = nodes[0];
    if (root.count > 0) {
        emitPair(0, 0);   // a lone leaf: test it through a pair with itself on both sides
        out.depth = 1;
        out.leaves = 1;
        return;
    }
    // Iterative DFS: pair for slot s is created when s is popped; children links are
    // patched once the child pair exists.
    struct Pending { int slot; int parentPair; int side; int depth; };
    std::vector<Pending> todo;
    todo.push_back({0, -1, 0, 1});
    while (!todo.empty()) {
        Pending p = todo.back();
        todo.pop_back();
        const BNode& nd = nodes[p.slot];
        const int pi = emitPair(nd.left, nd.right);
        out.depth = std::max(out.depth, p.depth);
        if (p.parentPair >= 0) std::memcpy(&pr[16 * (size_t)p.parentPair + (p.side ? 7 : 3)], &pi, 4);
        const BNode& L = nodes[nd.left];
        const BNode& R = nodes[nd.right];
        if (L.count > 0) out.leaves++;
        if (R.count > 0) out.leaves++;
        if (R.count == 0) todo.push_back({nd.right, pi, 1, p.depth + 1});
        if (L.count == 0) todo.push_back({nd.left, pi, 0, p.depth + 1});
    }
}

}  // namespace

void buildBvh(const float* verts, int64_t n, int threads, BvhResult& out) {
    if (n <= 0) throw std::runtime_error("scene has no triangles");
    if (n >= (int64_t(1) << 28)) throw std::runtime_error("more than 2^28 triangles");
    std::vector<Prim> P((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const float* v = verts + 9 * i;
        Prim& p = P[(size_t)i];
        for (int a = 0; a < 3; ++a) {
            double lo = std::min({(double)v[a], (double)v[3 + a], (double)v[6 + a]}) - kPad;
            double hi = std::max({(double)v[a], (double)v[3 + a], (double)v[6 + a]}) + kPad;
            p.lo[a] = down(lo);
            p.hi[a] = up(hi);
            p.c[a] = (float)(0.5 * (lo + hi));
        }
    }
    std::vector<int> idx((size_t)n);
    for (int64_t i = 0; i < n; ++i) idx[(size_t)i] = (int)i;
    std::vector<BNode> nodes((size_t)(2 * n));
    int parDepth = 0;
    while ((1 << parDepth) < std::max(1, threads) && parDepth < 6) ++parDepth;
    Builder bld{P, idx, nodes, parDepth};
    bld.build(0, 0, (int)n, 0);

    out.order = idx;
    emitPairs(nodes, out);
}

void importBvh(const float* verts, int64_t nTris, const gp_bvh_node* in, int64_t nNodes,
               const int32_t* primIndex, int64_t nIndex, BvhResult& out) {
    if (!in || nNodes <= 0 || !primIndex || nIndex <= 0) throw std::runtime_error("prebuilt BVH: empty");
    if (nNodes >= (int64_t(1) << 29) || nIndex >= (int64_t(1) << 28))
        throw std::runtime_error("prebuilt BVH: too many nodes or primitive references");
    std::vector<BNode> nodes((size_t)nNodes);
    std::vector<char> seen((size_t)nNodes, 0);
    std::vector<int> refs((size_t)nTris, 0);
    auto inside = [](const gp_bvh_node& c, const gp_bvh_node& p) {
        for (int a = 0; a < 3; ++a)
            if (!(c.lo[a] >= p.lo[a] && c.hi[a] <= p.hi[a])) return false;
        return true;
    };
    struct Visit { int64_t node; int depth; };
    std::vector<Visit> todo{{0, 1}};
    while (!todo.empty()) {
        const Visit v = todo.back();
        todo.pop_back();
        if (seen[(size_t)v.node]) throw std::runtime_error("prebuilt BVH: a node is reached twice");
        seen[(size_t)v.node] = 1;
        if (v.depth > 1024) throw std::runtime_error("prebuilt BVH: deeper than 1024 levels");
        const gp_bvh_node& g = in[v.node];
        BNode& b = nodes[(size_t)v.node];
        for (int a = 0; a < 3; ++a) {
            if (!(g.lo[a] <= g.hi[a])) throw std::runtime_error("prebuilt BVH: empty or non-finite box");
            b.lo[a] = down(g.lo[a]);
            b.hi[a] = up(g.hi[a]);
        }
        if (g.count > 0) {
            if (g.count > kLeafMax) throw std::runtime_error("prebuilt BVH: a leaf holds more than 4 primitives");
            if (g.first < 0 || (int64_t)g.first + g.count > nIndex)
                throw std::runtime_error("prebuilt BVH: leaf range outside the primitive index");
            for (int k = 0; k < g.count; ++k) {
                const int32_t prim = primIndex[g.first + k];
                if (prim < 0 || prim >= nTris) throw std::runtime_error("prebuilt BVH: primitive index out of range");
                const float* t = verts + 9 * (size_t)prim;
                for (int c = 0; c < 9; ++c)
                    if (!((double)t[c] >= g.lo[c % 3] && (double)t[c] <= g.hi[c % 3]))
                        throw std::runtime_error("prebuilt BVH: a leaf box does not contain its triangle");
                refs[(size_t)prim]++;
            }
            b.first = g.first;
            b.count = g.count;
        } else {
            if (g.left < 0 || g.left >= nNodes || g.right < 0 || g.right >= nNodes)
                throw std::runtime_error("prebuilt BVH: child index out of range");
            if (!inside(in[g.left], g) || !inside(in[g.right], g))
                throw std::runtime_error("prebuilt BVH: a child box is not inside its parent's");
            b.left = g.left;
            b.right = g.right;
            todo.push_back({g.right, v.depth + 1});
            todo.push_back({g.left, v.depth + 1});
        }
    }
    for (int64_t p = 0; p < nTris; ++p)
        if (refs[(size_t)p] == 0) throw std::runtime_error("prebuilt BVH: a triangle is in no leaf");
    out.order.assign(primIndex, primIndex + nIndex);
    emitPairs(nodes, out);
}

void collapse4(const std::vector<float>& pairs, std::vector<float>& quads, int& depth4) {
    struct Slot { float lo[3], hi[3]; int link; };
    auto child = [&](int p, int side) {
        Slot c;
        const float* q = &pairs[16 * (size_t)p + (side ? 8 : 0)];
        for (int a = 0; a < 3; ++a) { c.lo[a] = q[a]; c.hi[a] = q[4 + a]; }
        std::memcpy(&c.link, &pairs[16 * (size_t)p + (side ? 7 : 3)], 4);
        return c;
    };
    struct Todo { int pair, quad, level; };
    quads.assign(32, 0.0f);
    depth4 = 0;
    std::vector<Todo> todo{{0, 0, 1}};
    while (!todo.empty()) {
        const Todo t = todo.back();
        todo.pop_back();
        depth4 = std::max(depth4, t.level);
        Slot s[4];
        int n = 0;
        for (int side = 0; side < 2; ++side) {
            const Slot c = child(t.pair, side);
            if (c.link < 0) s[n++] = c;
            else { s[n++] = child(c.link, 0); s[n++] = child(c.link, 1); }
        }
        int link[4];
        for (int k = 0; k < n; ++k) {
            link[k] = s[k].link;
            if (link[k] >= 0) {
                const int nq = (int)(quads.size() / 32);
                quads.resize(quads.size() + 32, 0.0f);
                todo.push_back({link[k], nq, t.level + 1});
                link[k] = nq;
            }
        }
        float* q = &quads[32 * (size_t)t.quad];
        for (int k = 0; k < 4; ++k) {
            if (k < n) {
                for (int a = 0; a < 3; ++a) { q[8 * a + k] = s[k].lo[a]; q[8 * a + 4 + k] = s[k].hi[a]; }
                std::memcpy(&q[24 + k], &link[k], 4);
            } else {
                // empty slot: a point box far outside any scene, linked to slot 0 (valid)
                for (int a = 0; a < 3; ++a) { q[8 * a + k] = 1e30f; q[8 * a + 4 + k] = 1e30f; }
                std::memcpy(&q[24 + k], &link[0], 4);
            }
        }
    }
}

}  // namespace gp
