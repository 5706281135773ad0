// Test infrastructure only: the oracle. A double-precision CPU restatement of the
// reference's mode-B forward light tracer for triangle scenes, quad area lights and
// a pinhole camera. It checks the GPU library and provides the "port" CPU baseline;
// the product never links or calls it. See gp_oracle.hpp for how it is pinned.
//
// Floating-point expressions keep the reference's evaluation order wherever the
// order changes bits (vector ops are component-wise left to right, divisions stay
// divisions), because the films are compared with the reference's bit for bit.
// Followed files (reference repository, src/):
//   rng.h:6-52 (Pcg32, mix64, seedUnit, cosineHemisphere), linalg.h:5-39,
//   color.h:8-39 (CIE fit, cieYIntegral), spectrum.h:80-143, 247-284 (Planck,
//   Sellmeier, Cauchy, EmissionSampler), hero.h:106-127 (sampleBundle, maxOf),
//   geometry.h:51-99 (Tri::finalize normal), :174-242 (orientedGeoN, adjoint,
//   terminator), :256-369 (watertight triangle test), bvh.h:18-250 (Aabb, binned
//   SAH build, closest/any traversal), scene.h:1306, 1491-1528, 1851-1857, 1907-1944,
//   2005-2044, 2194-2271 (lights, build, closestHit, occluded), camera.h:110-240,
//   render.h:46-64 (glossy lobe), :499-577 (connectGeom, connectHero),
//   :1612-1733 (specular interactions), :2252-2631 (tracePhotonHero + loop),
//   :2657-2694 (refractOrReflect), main.cpp:11305-11340 (thread pool + seeding),
//   main.cpp:10265-10348 + color.h:41-53 (filmToLinear, filmToRgb8, xyzToLinearSrgb,
//   srgbGamma; without the optional denoiser), main.cpp:12317-12329, 12467-12483
//   (checkpointGuard, readCheckpoint).
#include "gp_oracle.hpp"

#include <algorithm>
#include <array>
#include <cstdio>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <thread>
#include <vector>

namespace orc {

constexpr double kPi = 3.141592653589793;
constexpr double kLamMin = 360.0;
constexpr double kLamMax = 830.0;

// ---------------------------------------------------------------- vectors
struct V3 {
    double x = 0, y = 0, z = 0;
};
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, const V3& a) { return a * s; }
inline V3 operator/(const V3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 neg(const V3& a) { return {-a.x, -a.y, -a.z}; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const V3& a) { return std::sqrt(dot(a, a)); }
inline V3 unit(const V3& a) { return a / norm(a); }
inline V3 mirrorDir(const V3& d, const V3& n) { return d - n * (2.0 * dot(d, n)); }
inline double at(const V3& v, int i) { return (&v.x)[i]; }
// Branchless orthonormal basis (Duff et al. 2017).
inline void frame(const V3& n, V3& t, V3& b) {
    double sign = std::copysign(1.0, n.z);
    double a = -1.0 / (sign + n.z);
    double d = n.x * n.y * a;
    t = V3{1.0 + sign * n.x * n.x * a, sign * d, -sign * n.x};
    b = V3{d, sign + n.y * n.y * a, -n.y};
}
inline V3 v3(const double* p) { return V3{p[0], p[1], p[2]}; }

// ---------------------------------------------------------------- rng
struct Pcg {
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;
    void seed(uint64_t seq, uint64_t s) {
        state = 0;
        inc = (seq << 1u) | 1u;
        next();
        state += s;
        next();
    }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31));
    }
    double uniform() { return (next() >> 8) * (1.0 / 16777216.0); }
};
inline uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
inline void seedPhoton(Pcg& r, uint64_t k, uint64_t salt) { r.seed(splitmix(k ^ salt), splitmix(k + salt)); }

inline V3 cosineDir(const V3& n, Pcg& rng) {
    double u1 = rng.uniform(), u2 = rng.uniform();
    double r = std::sqrt(u1);
    double phi = 6.283185307179586 * u2;
    double lx = r * std::cos(phi), ly = r * std::sin(phi), lz = std::sqrt(1.0 - u1);
    V3 t, b;
    frame(n, t, b);
    return unit(lx * t + ly * b + lz * n);
}
inline V3 glossyDir(const V3& mdir, double rough, double u1, double u2) {
    if (u1 >= 1.0) return mdir;
    double rr = rough < 1e-3 ? 1e-3 : rough;
    double e = 2.0 / (rr * rr) - 2.0;
    if (e < 0) e = 0;
    double cosT = std::pow(u1, 1.0 / (e + 1.0));
    double sinT = std::sqrt(std::max(0.0, 1.0 - cosT * cosT));
    double phi = 2.0 * kPi * u2;
    V3 t, b;
    frame(mdir, t, b);
    return unit(t * (sinT * std::cos(phi)) + b * (sinT * std::sin(phi)) + mdir * cosT);
}
inline V3 glossySample(const V3& mdir, double rough, Pcg& rng) {
    double u1 = rng.uniform(), u2 = rng.uniform();
    return glossyDir(mdir, rough, u1, u2);
}

// ---------------------------------------------------------------- spectra
inline double planck(double kelvin, double lambdaNm) {
    const double h = 6.62607015e-34, c = 2.99792458e8, kb = 1.380649e-23;
    double l = lambdaNm * 1e-9;
    double e = std::exp((h * c) / (l * kb * kelvin)) - 1.0;
    return (2.0 * h * c * c) / (std::pow(l, 5.0) * e);
}
// GP_SPEC_SAMPLED: linear interpolation of n uniform samples over [lo, hi], clamped to
// the end samples outside it (the reference GPU backend's bakeSpec / specLookup,
// render_cuda.cu:12658-12663, :1641-1649, at the caller's sample count and range).
inline double sampledEval(const double* tab, int n, double lo, double hi, double w) {
    double f = (w - lo) * (double)(n - 1) / (hi - lo);   // exact at the sample wavelengths
    if (!(f > 0.0)) return tab[0];
    if (f >= (double)(n - 1)) return tab[n - 1];
    int i = (int)f;
    double fr = f - (double)i;
    return tab[i] * (1.0 - fr) + tab[i + 1] * fr;
}
// `pool` holds the samples of GP_SPEC_SAMPLED spectra (gp_scene_desc.spectrum_samples).
double specEval(const gp_spectrum& s, double w, const double* pool = nullptr) {
    if (s.band_hi > s.band_lo && !(w >= s.band_lo && w <= s.band_hi)) return 0.0;
    switch (s.kind) {
        case GP_SPEC_SAMPLED:
            return pool ? sampledEval(pool + s.table, (int)s.p[2], s.p[0], s.p[1], w) : 0.0;
        case GP_SPEC_CONST: return s.p[0];
        case GP_SPEC_STEP: return (w < s.p[2]) ? s.p[0] : s.p[1];
        case GP_SPEC_GAUSS: {
            double t = (w - s.p[2]) / s.p[3];
            return s.p[0] + s.p[1] * std::exp(-0.5 * t * t);
        }
        case GP_SPEC_BLACKBODY: return (s.p[1] == 1.0) ? planck(s.p[0], w) : s.p[1] * planck(s.p[0], w);
        case GP_SPEC_SELLMEIER: {
            double l2 = (w * 1e-3) * (w * 1e-3);
            double n2 = 1.0 + s.p[0] * l2 / (l2 - s.p[3]) + s.p[1] * l2 / (l2 - s.p[4]) +
                        s.p[2] * l2 / (l2 - s.p[5]);
            return std::sqrt(n2 > 1.0 ? n2 : 1.0);
        }
        case GP_SPEC_CAUCHY: {
            double um = w * 1e-3;
            return s.p[0] + s.p[1] / (um * um);
        }
        default: return 0.0;
    }
}

// CIE 1931 observer, multi-lobe Gaussian fit (Wyman, Sloan, Shirley 2013).
inline double lobe(double x, double mu, double s1, double s2) {
    double t = (x - mu) * ((x < mu) ? s1 : s2);
    return std::exp(-0.5 * t * t);
}
inline double xbar(double w) {
    return 0.362 * lobe(w, 442.0, 0.0624, 0.0374) + 1.056 * lobe(w, 599.8, 0.0264, 0.0323) -
           0.065 * lobe(w, 501.1, 0.0490, 0.0382);
}
inline double ybar(double w) {
    return 0.821 * lobe(w, 568.8, 0.0213, 0.0247) + 0.286 * lobe(w, 530.9, 0.0613, 0.0322);
}
inline double zbar(double w) {
    return 1.217 * lobe(w, 437.0, 0.0845, 0.0278) + 0.681 * lobe(w, 459.0, 0.0385, 0.0725);
}
double ybarIntegral() {
    double s = 0;
    for (double w = kLamMin; w <= kLamMax; w += 1.0) s += ybar(w);
    return s;
}

// Piecewise-constant 1-nm CDF over [360, 830) and its inversion.
struct LambdaCdf {
    std::vector<double> cdf;
    double integral = 0.0;
    void build(const gp_spectrum& spd, const double* pool = nullptr) {
        const double step = 1.0;
        int n = static_cast<int>((kLamMax - kLamMin) / step);
        cdf.assign(n + 1, 0.0);
        double acc = 0.0;
        for (int i = 0; i < n; ++i) {
            double w = kLamMin + (i + 0.5) * step;
            acc += std::max(0.0, specEval(spd, w, pool)) * step;
            cdf[i + 1] = acc;
        }
        integral = acc;
        if (acc > 0)
            for (auto& c : cdf) c /= acc;
    }
    double invert(double u, double& pdf) const {
        const double step = 1.0;
        int lo = 0, hi = static_cast<int>(cdf.size()) - 1;
        while (lo + 1 < hi) {
            int mid = (lo + hi) / 2;
            (cdf[mid] <= u ? lo : hi) = mid;
        }
        double c0 = cdf[lo], c1 = cdf[lo + 1];
        double frac = (c1 > c0) ? (u - c0) / (c1 - c0) : 0.5;
        double w = kLamMin + (lo + frac) * step;
        pdf = (c1 - c0) / step;
        return w;
    }
};
// Hero + stratified secondaries from one uniform.
bool drawBundle(const LambdaCdf& cdf, double u, int C, double* lam, double* pdf) {
    pdf[0] = 0.0;
    lam[0] = cdf.invert(u, pdf[0]);
    if (pdf[0] <= 0.0) return false;
    for (int i = 1; i < C; ++i) {
        double uu = u + (double)i / (double)C;
        if (uu >= 1.0) uu -= 1.0;
        pdf[i] = 0.0;
        lam[i] = cdf.invert(uu, pdf[i]);
    }
    return true;
}
inline double maxLive(const double* a, int n) {
    double m = a[0];
    for (int i = 1; i < n; ++i)
        if (a[i] > m) m = a[i];
    return m;
}
inline double clampUnit(double x) { return x < 0 ? 0 : (x > 1 ? 1 : x); }

// ---------------------------------------------------------------- geometry
struct Tri {
    V3 v0, v1, v2, gn;
    int mat = 0;
};
struct Hit {
    double t = DBL_MAX;
    bool valid = false;
    V3 p, n, ng;
    int mat = 0;
    int prim = -1;
};
struct Shear {
    int kx, ky, kz;
    double Sx, Sy, Sz;
};
Shear makeShear(const V3& d) {
    Shear s;
    double ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    if (ax >= ay && ax >= az) s.kz = 0;
    else if (ay >= az) s.kz = 1;
    else s.kz = 2;
    s.kx = s.kz + 1;
    if (s.kx == 3) s.kx = 0;
    s.ky = s.kx + 1;
    if (s.ky == 3) s.ky = 0;
    if (at(d, s.kz) < 0.0) std::swap(s.kx, s.ky);
    s.Sx = at(d, s.kx) / at(d, s.kz);
    s.Sy = at(d, s.ky) / at(d, s.kz);
    s.Sz = 1.0 / at(d, s.kz);
    return s;
}
inline double diffExt(double a, double b, double c, double d) {
    return (double)((long double)a * (long double)b - (long double)c * (long double)d);
}
// Watertight two-sided ray/triangle test (Woop, Benthin, Wald 2013).
bool hitTriangle(const Shear& sh, const V3& o, const V3& d, const Tri& tri, int prim, double tmin,
                 Hit& h) {
    const int kx = sh.kx, ky = sh.ky, kz = sh.kz;
    V3 A = tri.v0 - o, B = tri.v1 - o, C = tri.v2 - o;
    double Ax = at(A, kx) - sh.Sx * at(A, kz), Ay = at(A, ky) - sh.Sy * at(A, kz);
    double Bx = at(B, kx) - sh.Sx * at(B, kz), By = at(B, ky) - sh.Sy * at(B, kz);
    double Cx = at(C, kx) - sh.Sx * at(C, kz), Cy = at(C, ky) - sh.Sy * at(C, kz);
    double U = Cx * By - Cy * Bx;
    double V = Ax * Cy - Ay * Cx;
    double W = Bx * Ay - By * Ax;
    if (U == 0.0 || V == 0.0 || W == 0.0) {
        if (U == 0.0) U = diffExt(Cx, By, Cy, Bx);
        if (V == 0.0) V = diffExt(Ax, Cy, Ay, Cx);
        if (W == 0.0) W = diffExt(Bx, Ay, By, Ax);
    }
    if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return false;
    double det = U + V + W;
    if (det == 0.0) return false;
    double T = U * (sh.Sz * at(A, kz)) + V * (sh.Sz * at(B, kz)) + W * (sh.Sz * at(C, kz));
    double invDet = 1.0 / det;
    double t = T * invDet;
    if (t < tmin || t >= h.t) return false;
    double b0 = U * invDet, b1 = V * invDet, b2 = W * invDet;
    h.t = t;
    h.p = o + d * t;
    h.valid = true;
    h.mat = tri.mat;
    h.prim = prim;
    // Flat triangle: every vertex normal is the geometric normal, interpolated and
    // renormalised exactly as a smooth-normal mesh would be.
    V3 ns = tri.gn * b0 + tri.gn * b1 + tri.gn * b2;
    double nl = dot(ns, ns);
    ns = (nl > 1e-18) ? ns * (1.0 / std::sqrt(nl)) : tri.gn;
    V3 gn = (dot(tri.gn, ns) < 0.0) ? neg(tri.gn) : tri.gn;
    bool flipped = !(dot(d, gn) < 0.0);
    h.ng = gn;
    h.n = flipped ? neg(ns) : ns;
    return true;
}

// ---------------------------------------------------------------- BVH
struct Box {
    V3 lo{DBL_MAX, DBL_MAX, DBL_MAX};
    V3 hi{-DBL_MAX, -DBL_MAX, -DBL_MAX};
    void grow(const V3& p) {
        lo.x = std::min(lo.x, p.x); lo.y = std::min(lo.y, p.y); lo.z = std::min(lo.z, p.z);
        hi.x = std::max(hi.x, p.x); hi.y = std::max(hi.y, p.y); hi.z = std::max(hi.z, p.z);
    }
    void grow(const Box& b) { grow(b.lo); grow(b.hi); }
    V3 center() const { return (lo + hi) * 0.5; }
    double area() const {
        V3 d = hi - lo;
        if (d.x < 0 || d.y < 0 || d.z < 0) return 0.0;
        return 2.0 * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    int widest() const {
        V3 d = hi - lo;
        if (d.x >= d.y && d.x >= d.z) return 0;
        return d.y >= d.z ? 1 : 2;
    }
    // Slab test on [tmin, tmax]; one axis at a time with early out.
    bool slab(const V3& o, const V3& inv, double tmin, double tmax, double& tEnter) const {
        double te = tmin, tx = tmax;
        const double* L = &lo.x;
        const double* H = &hi.x;
        const double* O = &o.x;
        const double* I = &inv.x;
        for (int a = 0; a < 3; ++a) {
            double t0 = (L[a] - O[a]) * I[a];
            double t1 = (H[a] - O[a]) * I[a];
            if (t0 > t1) std::swap(t0, t1);
            te = t0 > te ? t0 : te;
            tx = t1 < tx ? t1 : tx;
            if (tx < te) return false;
        }
        tEnter = te;
        return true;
    }
};
struct Node {
    Box box;
    int left = -1, right = -1, first = 0, count = 0;
};
struct Bvh {
    std::vector<Node> nodes;
    std::vector<int> order;
    static constexpr int kLeaf = 4;
    static constexpr int kBins = 16;
    struct Item { Box box; V3 c; int idx; };

    void build(const std::vector<Box>& boxes) {
        const int n = (int)boxes.size();
        nodes.clear();
        order.clear();
        if (n == 0) return;
        std::vector<Item> it(n);
        for (int i = 0; i < n; ++i) { it[i].box = boxes[i]; it[i].c = boxes[i].center(); it[i].idx = i; }
        nodes.reserve(2 * n);
        split(it, 0, n);
        order.resize(n);
        for (int i = 0; i < n; ++i) order[i] = it[i].idx;
    }
    // Binned SAH on the widest centroid axis, always subdividing to kLeaf, with a
    // median split when no SAH plane separates the items.
    int split(std::vector<Item>& it, int begin, int end) {
        const int me = (int)nodes.size();
        nodes.push_back(Node{});
        Box bounds, cb;
        for (int i = begin; i < end; ++i) { bounds.grow(it[i].box); cb.grow(it[i].c); }
        const int count = end - begin;
        auto leaf = [&]() {
            Node& nd = nodes[me];
            nd.box = bounds; nd.first = begin; nd.count = count; nd.left = nd.right = -1;
        };
        if (count <= kLeaf) { leaf(); return me; }
        const int axis = cb.widest();
        const double c0 = at(cb.lo, axis), c1 = at(cb.hi, axis);
        if (c1 - c0 < 1e-12) { leaf(); return me; }
        struct Bin { Box box; int n = 0; };
        Bin bins[kBins];
        const double scale = kBins / (c1 - c0);
        auto binOf = [&](const Item& x) {
            int b = (int)((at(x.c, axis) - c0) * scale);
            if (b < 0) b = 0;
            if (b >= kBins) b = kBins - 1;
            return b;
        };
        for (int i = begin; i < end; ++i) { Bin& b = bins[binOf(it[i])]; b.n++; b.box.grow(it[i].box); }
        double rArea[kBins];
        int rCount[kBins];
        Box acc;
        int rc = 0;
        for (int b = kBins - 1; b >= 1; --b) {
            acc.grow(bins[b].box);
            rc += bins[b].n;
            rArea[b] = acc.area();
            rCount[b] = rc;
        }
        Box lacc;
        int lc = 0;
        double best = DBL_MAX;
        int bestSplit = -1;
        for (int b = 0; b < kBins - 1; ++b) {
            lacc.grow(bins[b].box);
            lc += bins[b].n;
            if (lc == 0 || rCount[b + 1] == 0) continue;
            double cost = lc * lacc.area() + rCount[b + 1] * rArea[b + 1];
            if (cost < best) { best = cost; bestSplit = b; }
        }
        int mid = begin;
        if (bestSplit >= 0) {
            auto m = std::partition(it.begin() + begin, it.begin() + end,
                                    [&](const Item& x) { return binOf(x) <= bestSplit; });
            mid = (int)(m - it.begin());
        }
        if (mid == begin || mid == end) {
            mid = (begin + end) / 2;
            std::nth_element(it.begin() + begin, it.begin() + mid, it.begin() + end,
                             [&](const Item& a, const Item& b) { return at(a.c, axis) < at(b.c, axis); });
        }
        int l = split(it, begin, mid);
        int r = split(it, mid, end);
        Node& nd = nodes[me];
        nd.box = bounds; nd.left = l; nd.right = r; nd.count = 0;
        return me;
    }

    template <class Leaf>
    void closest(const V3& o, const V3& d, double tmin, double& tMax, Leaf&& leaf, int64_t* visits,
                 int64_t* tests) const {
        if (nodes.empty()) return;
        V3 inv{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
        int stack[64];
        double tst[64];
        int sp = 0;
        double tRoot;
        if (!nodes[0].box.slab(o, inv, tmin, tMax, tRoot)) { if (visits) ++*visits; return; }
        stack[0] = 0; tst[0] = tRoot; sp = 1;
        while (sp) {
            --sp;
            if (visits) ++*visits;
            if (tst[sp] > tMax) continue;
            const Node& nd = nodes[stack[sp]];
            if (nd.count > 0) {
                if (tests) *tests += nd.count;
                for (int i = 0; i < nd.count; ++i) leaf(order[nd.first + i], tMax);
            } else {
                double tL, tR;
                bool hL = nodes[nd.left].box.slab(o, inv, tmin, tMax, tL);
                bool hR = nodes[nd.right].box.slab(o, inv, tmin, tMax, tR);
                if (hL && hR) {
                    if (tL <= tR) {
                        stack[sp] = nd.right; tst[sp] = tR; ++sp;
                        stack[sp] = nd.left; tst[sp] = tL; ++sp;
                    } else {
                        stack[sp] = nd.left; tst[sp] = tL; ++sp;
                        stack[sp] = nd.right; tst[sp] = tR; ++sp;
                    }
                } else if (hL) {
                    stack[sp] = nd.left; tst[sp] = tL; ++sp;
                } else if (hR) {
                    stack[sp] = nd.right; tst[sp] = tR; ++sp;
                }
            }
        }
    }
    // Optional counters (an extension: the reference's traverseAny has none): one
    // node visit per popped node, plus one for a root that misses, as closest()
    // counts them; one leaf test per primitive tested.
    template <class Leaf>
    bool any(const V3& o, const V3& d, double tmin, double tMax, Leaf&& leaf, int64_t* visits = nullptr,
             int64_t* tests = nullptr) const {
        if (nodes.empty()) return false;
        V3 inv{1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
        double tRoot;
        if (!nodes[0].box.slab(o, inv, tmin, tMax, tRoot)) { if (visits) ++*visits; return false; }
        int stack[64];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            if (visits) ++*visits;
            const Node& nd = nodes[stack[--sp]];
            if (nd.count > 0) {
                for (int i = 0; i < nd.count; ++i) {
                    if (tests) ++*tests;
                    if (leaf(order[nd.first + i])) return true;
                }
            } else {
                double tc;
                if (nodes[nd.left].box.slab(o, inv, tmin, tMax, tc)) stack[sp++] = nd.left;
                if (nodes[nd.right].box.slab(o, inv, tmin, tMax, tc)) stack[sp++] = nd.right;
            }
        }
        return false;
    }
};

// ---------------------------------------------------------------- scene
struct Mat {
    int type = GP_MAT_DIFFUSE;
    double roughness = 0.1;
    gp_spectrum reflect{}, ior{};
};
struct Light {
    V3 origin, u, v, normal, beam;
    double area = 0, power = 0, integral = 0;
    bool collimated = false;
    LambdaCdf cdf;
};
// A flat mirror surface: the coplanar mirror triangles of one plane, reduced to the
// plane (unit normal n with its first dominant component positive, dot(n, x) = d) and
// the padded box of its member triangles (scene.h:1606-1642).
struct MirrorPlane {
    V3 n;
    double d = 0;
    V3 lo, hi;
    double area = 0;
};
constexpr int kMaxMirrorPlanes = 64;

// Materials whose reflection is an exact mirror image (scene.h:356-360; HalfMirror is
// not among the supported types).
inline bool planarMirror(const Mat& m) {
    return m.type == GP_MAT_MIRROR || (m.type == GP_MAT_GLOSSY && m.roughness <= 0.0);
}

struct Scene {
    std::vector<Tri> tris;
    std::vector<Mat> mats;
    std::vector<double> pool;   // samples of GP_SPEC_SAMPLED spectra
    std::vector<Light> lights;
    std::vector<double> lightCdf;
    std::vector<MirrorPlane> mirrors;
    double totalPower = 0;
    Bvh bvh;

    // Group mirror triangles by plane (scene.h:1646-1724): a quantised key of the
    // canonical normal (1e-6) and offset (1e-5) picks the entry; entries keep their
    // order of first appearance. Past kMaxMirrorPlanes the largest total areas stay
    // (the reference's sort is not stable; equal areas keep their order here).
    void buildMirrorPlanes() {
        mirrors.clear();
        std::map<std::array<long long, 4>, int> byKey;
        for (const Tri& t : tris) {
            if (t.mat < 0 || t.mat >= (int)mats.size() || !planarMirror(mats[t.mat])) continue;
            V3 n = t.gn;
            const double nl = norm(n);
            if (nl < 1e-12) continue;
            n = n * (1.0 / nl);
            const double ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
            const double lead = (ax >= ay && ax >= az) ? n.x : (ay >= az ? n.y : n.z);
            if (lead < 0.0) n = neg(n);
            const double d = dot(n, t.v0);
            const std::array<long long, 4> key{std::llround(n.x * 1e6), std::llround(n.y * 1e6),
                                               std::llround(n.z * 1e6), std::llround(d * 1e5)};
            auto it = byKey.find(key);
            int idx;
            if (it == byKey.end()) {
                idx = (int)mirrors.size();
                byKey.emplace(key, idx);
                MirrorPlane mp;
                mp.n = n;
                mp.d = d;
                mp.lo = t.v0;
                mp.hi = t.v0;
                mirrors.push_back(mp);
            } else {
                idx = it->second;
            }
            MirrorPlane& mp = mirrors[(size_t)idx];
            for (const V3& v : {t.v0, t.v1, t.v2}) {
                mp.lo = V3{std::min(mp.lo.x, v.x), std::min(mp.lo.y, v.y), std::min(mp.lo.z, v.z)};
                mp.hi = V3{std::max(mp.hi.x, v.x), std::max(mp.hi.y, v.y), std::max(mp.hi.z, v.z)};
            }
            mp.area += 0.5 * norm(cross(t.v1 - t.v0, t.v2 - t.v0));
        }
        if ((int)mirrors.size() > kMaxMirrorPlanes) {
            std::stable_sort(mirrors.begin(), mirrors.end(),
                             [](const MirrorPlane& a, const MirrorPlane& b) { return a.area > b.area; });
            mirrors.resize(kMaxMirrorPlanes);
        }
        for (MirrorPlane& mp : mirrors) {
            const double pad = 1e-6 * (1.0 + norm(mp.hi - mp.lo));
            mp.lo = mp.lo - V3{pad, pad, pad};
            mp.hi = mp.hi + V3{pad, pad, pad};
        }
    }

    void build() {
        for (auto& t : tris) t.gn = unit(cross(t.v1 - t.v0, t.v2 - t.v0));
        buildMirrorPlanes();
        std::vector<Box> boxes;
        boxes.reserve(tris.size());
        const double pad = 1e-6;
        for (const auto& t : tris) {
            Box b;
            b.grow(t.v0); b.grow(t.v1); b.grow(t.v2);
            b.lo = b.lo - V3{pad, pad, pad};
            b.hi = b.hi + V3{pad, pad, pad};
            boxes.push_back(b);
        }
        bvh.build(boxes);
        totalPower = 0.0;
        lightCdf.assign(lights.size(), 0.0);
        for (size_t i = 0; i < lights.size(); ++i) {
            lights[i].power = lights[i].integral * lights[i].area * kPi;
            totalPower += lights[i].power;
            lightCdf[i] = totalPower;
        }
        if (totalPower > 0)
            for (auto& c : lightCdf) c /= totalPower;
    }
    int pickLight(Pcg& rng) const {
        if (lights.size() <= 1) return 0;
        double u = rng.uniform();
        int lo = 0, hi = (int)lightCdf.size() - 1;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (lightCdf[mid] < u) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    }
    Hit closestHit(const V3& o, const V3& d, double tmin = 1e-6, int64_t* visits = nullptr,
                   int64_t* tests = nullptr) const {
        Hit h;
        double tMax = DBL_MAX;
        const Shear sh = makeShear(d);
        bvh.closest(o, d, tmin, tMax, [&](int prim, double& tm) {
            if (hitTriangle(sh, o, d, tris[prim], prim, tmin, h)) tm = h.t;
        }, visits, tests);
        return h;
    }
    bool occluded(const V3& o, const V3& dir, double maxDist, double tmin = 1e-6, int64_t* visits = nullptr,
                  int64_t* tests = nullptr) const {
        const double seg = maxDist - tmin;
        const Shear sh = makeShear(dir);
        return bvh.any(o, dir, tmin, seg, [&](int prim) {
            Hit h;
            h.t = seg;
            return hitTriangle(sh, o, dir, tris[prim], prim, tmin, h);
        }, visits, tests);
    }
};

// ---------------------------------------------------------------- camera
struct Cam {
    V3 eye, u, v, w;
    double tanHalfX = 0, tanHalfY = 0;
    int resX = 0, resY = 0;
    void lookAt(const gp_camera_desc& c) {
        eye = v3(c.eye);
        w = unit(v3(c.target) - eye);
        u = unit(cross(w, v3(c.up)));
        v = cross(u, w);
        tanHalfY = std::tan(0.5 * c.fov_y_deg * 3.141592653589793 / 180.0);
        tanHalfX = tanHalfY * (double)c.res_x / (double)c.res_y;
        resX = c.res_x;
        resY = c.res_y;
    }
    bool project(const V3& p, int& px, int& py, double& cosCam, double& dist2) const {
        V3 d = p - eye;
        double cz = dot(d, w);
        if (cz <= 1e-9) return false;
        double cx = dot(d, u), cy = dot(d, v);
        double ix = (cx / cz) / tanHalfX - 0.0;
        double iy = (cy / cz) / tanHalfY;
        if (ix < -1 || ix >= 1 || iy < -1 || iy >= 1) return false;
        px = (int)((ix * 0.5 + 0.5) * resX);
        py = (int)((iy * 0.5 + 0.5) * resY);
        dist2 = dot(d, d);
        cosCam = cz / std::sqrt(dist2);
        return true;
    }
    double pixelArea() const { return (4.0 * tanHalfX * tanHalfY) / ((double)resX * (double)resY); }
    double pixelOmega(double cosCam) const { return pixelArea() * cosCam * cosCam * cosCam; }
};

// ---------------------------------------------------------------- transport
struct Energy {
    double emitted = 0, absorbed = 0, sensor = 0, escaped = 0, residual = 0;
};
struct Counters {
    uint64_t segments = 0, connections = 0, splats = 0;
    // traversal statistics (TraversalStats, bvh.h:79) of the segments' closest-hit
    // queries and of the camera connections' any-hit queries
    int64_t nodeVisits = 0, leafTests = 0, anyNodeVisits = 0, anyLeafTests = 0;
};
struct Film {
    int resX = 0, resY = 0;
    std::vector<V3> xyz;
    std::vector<double> hits;
    void alloc(int x, int y) {
        resX = x; resY = y;
        xyz.assign((size_t)x * y, V3{});
        hits.assign((size_t)x * y, 0.0);
    }
    void add(int px, int py, const V3& c) {
        size_t i = (size_t)py * resX + px;
        xyz[i] = xyz[i] + c;
        hits[i] += 1.0;
    }
};
struct Params {
    int heroC = 4;
    int maxBounce = 32;
    int rrDepth = 0;
};

inline double adjointCorr(const V3& wi, const V3& wo, const V3& ns, const V3& ng) {
    double denom = std::fabs(dot(wi, ng)) * std::fabs(dot(wo, ns));
    if (denom <= 1e-8) return 1.0;
    double num = std::fabs(dot(wi, ns)) * std::fabs(dot(wo, ng));
    return num / denom;
}
inline double terminatorG(const V3& wi, const V3& ns, const V3& ng) {
    double cosNgNs = dot(ng, ns);
    if (cosNgNs >= 1.0 - 1e-7) return (dot(ng, wi) > 0.0) ? 1.0 : 0.0;
    double cosNgWi = dot(ng, wi);
    if (cosNgWi <= 0.0) return 0.0;
    double denom = dot(ns, wi) * cosNgNs;
    if (denom <= 1e-8) return 1.0;
    double g = cosNgWi / denom;
    if (g >= 1.0) return 1.0;
    return g * g * (1.0 - g) + g;
}
inline V3 facingGeoN(const Hit& h) { return (dot(h.ng, h.n) >= 0.0) ? h.ng : neg(h.ng); }

struct Tracer {
    const Scene& S;
    const Cam& cam;
    Params P;
    Film& film;
    Energy& e;
    Counters& k;

    // Pinhole connection of surface point p (shading normal n, oriented geometric
    // normal ng, wi toward the previous vertex) and per-wavelength splat.
    void splatHero(const V3& p, const V3& n, const V3& ng, const V3& wi, const double* lam,
                   const double* beta, const double* rho, int nUp) {
        V3 toCam = cam.eye - p;
        double dist = norm(toCam);
        V3 wdir = toCam / dist;
        double cosSurf = dot(n, wdir);
        if (cosSurf <= 0) return;
        double stG = terminatorG(wdir, n, ng);
        if (stG <= 0.0) return;
        int px, py;
        double cosCam, dist2;
        if (!cam.project(p, px, py, cosCam, dist2)) return;
        ++k.connections;
        if (S.occluded(p + ng * 1e-6, wdir, dist - 2e-6, 1e-6, &k.anyNodeVisits, &k.anyLeafTests)) return;
        ++k.splats;
        double omega = cam.pixelOmega(cosCam);
        double corr = adjointCorr(wi, wdir, n, ng);
        double denom = dist2 * omega;
        for (int i = 0; i < nUp; ++i) {
            double f = rho[i] / kPi;
            double contrib = beta[i] * f * cosSurf * corr / denom * stG;
            film.add(px, py, V3{xbar(lam[i]), ybar(lam[i]), zbar(lam[i])} * contrib);
        }
    }

    // Pinhole connection of vertex p (normal np, per-wavelength weight w) by
    // reflection in one flat mirror (render.h:1315-1375): the eye's mirror image turns
    // p -> R -> eye into the straight segment p -> eye', R is where it crosses the
    // plane, and the specular geometry factor is 1 / D^2 with D = |p - eye'|. The
    // light-side leg must be clear, and a closest hit from the eye must land on a
    // planar mirror at R (within 1e-4 (1 + dE)); the mirror's reflectance is read off
    // the material that hit found.
    void mirrorPlane(const MirrorPlane& mp, const V3& p, const V3& np, const double* lam,
                     const double* beta, const double* w, int nUp) {
        const V3 eye = cam.eye;
        const double se = dot(mp.n, eye) - mp.d;
        const double sp = dot(mp.n, p) - mp.d;
        if (!(se * sp > 0.0)) return;
        const V3 eyeM = eye - mp.n * (2.0 * se);
        V3 wP = eyeM - p;
        const double D = norm(wP);
        if (D < 1e-9) return;
        wP = wP * (1.0 / D);
        const double cosSurf = dot(np, wP);
        const double shape = cosSurf <= 0.0 ? -1.0 : cosSurf / kPi;
        if (shape < 0.0) return;
        const double denom = dot(mp.n, wP);
        if (std::fabs(denom) < 1e-12) return;
        const double dP = (mp.d - dot(mp.n, p)) / denom;
        if (dP <= 1e-7 || dP >= D - 1e-7) return;
        const V3 R = p + wP * dP;
        if (R.x < mp.lo.x || R.x > mp.hi.x || R.y < mp.lo.y || R.y > mp.hi.y || R.z < mp.lo.z ||
            R.z > mp.hi.z)
            return;
        int px, py;
        double cosCam, dist2e;
        if (!cam.project(R, px, py, cosCam, dist2e)) return;
        const double omega = cam.pixelOmega(cosCam);
        if (omega <= 0.0) return;
        const double dE = D - dP;
        const V3 wRE = (eye - R) * (1.0 / dE);
        ++k.connections;
        if (S.occluded(p + wP * 1e-6, wP, dP - 2e-6)) return;
        ++k.connections;
        const Hit hm = S.closestHit(eye, neg(wRE));
        if (!hm.valid) return;
        if (std::fabs(hm.t - dE) > 1e-4 * (1.0 + dE)) return;
        const Mat& mm = S.mats[hm.mat];
        if (!planarMirror(mm)) return;
        ++k.splats;
        const double invD2Omega = 1.0 / (D * D * omega);
        for (int i = 0; i < nUp; ++i) {
            const double refl = clampUnit(specEval(mm.reflect, lam[i], S.pool.data()));
            const double contrib = beta[i] * (w[i] * shape) * refl * invD2Omega;
            if (contrib <= 0.0) continue;
            film.add(px, py, V3{xbar(lam[i]), ybar(lam[i]), zbar(lam[i])} * contrib);
        }
    }
    void splatMirrors(const V3& p, const V3& np, const double* lam, const double* beta,
                      const double* w, int nUp) {
        for (const MirrorPlane& mp : S.mirrors) mirrorPlane(mp, p, np, lam, beta, w, nUp);
    }

    // Dielectric interface: Fresnel-weighted reflect or Snell refract (exterior
    // index 1), optionally frosted by a glossy lobe.
    void dielectric(const Mat& m, const Hit& h, V3& ro, V3& rd, double lambda, Pcg& rng) {
        double ng = specEval(m.ior, lambda, S.pool.data());
        bool entering = dot(rd, h.ng) < 0.0;
        V3 nl = entering ? h.ng : neg(h.ng);
        double n1 = entering ? 1.0 : ng;
        double n2 = entering ? ng : 1.0;
        double eta = n1 / n2;
        double cosI = -dot(rd, nl);
        double sin2t = eta * eta * (1.0 - cosI * cosI);
        V3 out;
        bool refracted = false;
        if (sin2t > 1.0) {
            out = mirrorDir(rd, nl);
        } else {
            double cosT = std::sqrt(1.0 - sin2t);
            double rs = (n1 * cosI - n2 * cosT) / (n1 * cosI + n2 * cosT);
            double rp = (n1 * cosT - n2 * cosI) / (n1 * cosT + n2 * cosI);
            double R = 0.5 * (rs * rs + rp * rp);
            if (rng.uniform() < R) out = mirrorDir(rd, nl);
            else { out = eta * rd + nl * (eta * cosI - cosT); refracted = true; }
        }
        out = unit(out);
        double rough = m.roughness;
        if (rough > 1e-3) {
            V3 pert = glossySample(out, rough, rng);
            bool ok = refracted ? (dot(pert, nl) < 0.0) : (dot(pert, nl) > 0.0);
            if (ok) out = pert;
        }
        ro = h.p + out * 1e-6;
        rd = out;
    }

    void trace(Pcg& rng) {
        if (S.lights.empty()) return;
        const int C = P.heroC;
        const int li = S.pickLight(rng);
        const Light& L = S.lights[li];
        double u1 = rng.uniform(), u2 = rng.uniform();
        V3 origin = L.origin + L.u * u1 + L.v * u2;
        V3 emitN = L.normal;
        V3 dir = L.collimated ? L.beam : cosineDir(emitN, rng);
        double lam[8], pdf[8];
        if (!drawBundle(L.cdf, rng.uniform(), C, lam, pdf)) return;
        double base = (S.lights.size() == 1) ? L.power : S.totalPower;
        base *= 1.0;
        double beta[8] = {};
        for (int i = 0; i < C; ++i) beta[i] = base / C;
        bool secAlive = (C > 1);
        auto activeSum = [&]() {
            double s = 0.0;
            int n = secAlive ? C : 1;
            for (int i = 0; i < n; ++i) s += beta[i];
            return s;
        };
        e.emitted += activeSum();
        {
            double one[8];
            for (int i = 0; i < C; ++i) one[i] = 1.0;
            splatHero(origin, emitN, emitN, emitN, lam, beta, one, C);
            splatMirrors(origin, emitN, lam, beta, one, C);
        }
        V3 ro = origin + dir * 1e-6, rd = dir;
        for (int bounce = 0; bounce < P.maxBounce; ++bounce) {
            int nUp = secAlive ? C : 1;
            ++k.segments;
            Hit h = S.closestHit(ro, rd, 1e-6, &k.nodeVisits, &k.leafTests);
            if (!h.valid) { e.escaped += activeSum(); return; }
            const Mat& m = S.mats[h.mat];
            const bool roulette = bounce >= P.rrDepth;
            if (m.type == GP_MAT_DIELECTRIC) {
                if (secAlive) { beta[0] *= (double)C; secAlive = false; }
                dielectric(m, h, ro, rd, lam[0], rng);
                continue;
            }
            if (m.type == GP_MAT_MIRROR || m.type == GP_MAT_GLOSSY) {
                double c[8] = {};
                for (int i = 0; i < nUp; ++i) c[i] = clampUnit(specEval(m.reflect, lam[i], S.pool.data()));
                const double q = maxLive(c, nUp);
                if (roulette) {
                    if (rng.uniform() >= q) { e.absorbed += activeSum(); return; }
                    for (int i = 0; i < nUp; ++i) {
                        double w = c[i] / q;
                        e.absorbed += beta[i] * (1.0 - w);
                        beta[i] *= w;
                    }
                } else {
                    if (q <= 0.0) { e.absorbed += activeSum(); return; }
                    for (int i = 0; i < nUp; ++i) {
                        e.absorbed += beta[i] * (1.0 - c[i]);
                        beta[i] *= c[i];
                    }
                }
                if (m.type == GP_MAT_MIRROR) {
                    V3 o = mirrorDir(rd, h.n);
                    ro = h.p + h.n * 1e-6;
                    rd = o;
                } else {
                    V3 o = glossySample(mirrorDir(rd, h.n), m.roughness, rng);
                    if (dot(o, h.n) <= 0) { e.absorbed += activeSum(); return; }
                    ro = h.p + h.n * 1e-6;
                    rd = o;
                }
                continue;
            }
            // Diffuse
            double rho[8];
            for (int i = 0; i < nUp; ++i) rho[i] = clampUnit(specEval(m.reflect, lam[i], S.pool.data()));
            V3 ngo = facingGeoN(h);
            V3 wi = neg(rd);
            splatHero(h.p, h.n, ngo, wi, lam, beta, rho, nUp);
            splatMirrors(h.p, h.n, lam, beta, rho, nUp);
            const double q = maxLive(rho, nUp);
            if (roulette) {
                if (rng.uniform() >= q) { e.absorbed += activeSum(); return; }
                for (int i = 0; i < nUp; ++i) {
                    double w = rho[i] / q;
                    e.absorbed += beta[i] * (1.0 - w);
                    beta[i] *= w;
                }
            } else {
                if (q <= 0.0) { e.absorbed += activeSum(); return; }
                for (int i = 0; i < nUp; ++i) {
                    e.absorbed += beta[i] * (1.0 - rho[i]);
                    beta[i] *= rho[i];
                }
            }
            V3 wo = cosineDir(h.n, rng);
            double corr = adjointCorr(wi, wo, h.n, ngo);
            for (int i = 0; i < nUp; ++i) beta[i] *= corr;
            ro = h.p + h.n * 1e-6;
            rd = wo;
        }
        e.residual += activeSum();
    }
};

Scene* makeScene(const gp_scene_desc* d) {
    auto* s = new Scene();
    if (d->spectrum_samples && d->n_spectrum_samples > 0)
        s->pool.assign(d->spectrum_samples, d->spectrum_samples + d->n_spectrum_samples);
    for (int i = 0; i < d->n_materials; ++i) {
        Mat m;
        m.type = d->materials[i].type;
        m.roughness = d->materials[i].roughness;
        m.reflect = d->materials[i].reflect;
        m.ior = d->materials[i].ior;
        s->mats.push_back(m);
    }
    s->tris.resize((size_t)d->n_tris);
    for (int64_t i = 0; i < d->n_tris; ++i) {
        const float* v = d->verts + 9 * i;
        Tri& t = s->tris[(size_t)i];
        t.v0 = V3{v[0], v[1], v[2]};
        t.v1 = V3{v[3], v[4], v[5]};
        t.v2 = V3{v[6], v[7], v[8]};
        t.mat = d->mat_ids[i];
    }
    for (int i = 0; i < d->n_emitters; ++i) {
        const gp_emitter& ge = d->emitters[i];
        Light L;
        L.origin = v3(ge.origin); L.u = v3(ge.u); L.v = v3(ge.v); L.normal = v3(ge.normal);
        L.area = ge.area; L.collimated = ge.collimated != 0; L.beam = v3(ge.beam_dir);
        L.cdf.build(ge.spd, s->pool.data());
        L.integral = L.cdf.integral;
        s->lights.push_back(std::move(L));
    }
    s->build();
    return s;
}

}  // namespace orc

using namespace orc;

extern "C" {

void* or_scene_create(const gp_scene_desc* d) { return makeScene(d); }
void or_scene_destroy(void* s) { delete static_cast<Scene*>(s); }
int64_t or_scene_bvh_nodes(void* s) { return (int64_t) static_cast<Scene*>(s)->bvh.nodes.size(); }
double or_scene_total_power(void* s) { return static_cast<Scene*>(s)->totalPower; }
int or_scene_mirror_planes(void* s, double* out10, int cap) {
    const auto& m = static_cast<Scene*>(s)->mirrors;
    for (int i = 0; i < (int)m.size() && i < cap; ++i) {
        const MirrorPlane& p = m[(size_t)i];
        const double v[10] = {p.n.x, p.n.y, p.n.z, p.d, p.lo.x, p.lo.y, p.lo.z, p.hi.x, p.hi.y, p.hi.z};
        std::memcpy(out10 + 10 * i, v, sizeof v);
    }
    return (int)m.size();
}

int or_render(void* sp, const gp_camera_desc* cd, int64_t N, uint64_t seedBase, uint64_t salt,
              int heroC, int maxBounce, int rrDepth, int nThreads, double* xyz, double* hits,
              gp_energy* eOut, uint64_t* stats7) {
    const Scene& S = *static_cast<Scene*>(sp);
    Cam cam;
    cam.lookAt(*cd);
    if (nThreads < 1) nThreads = 1;
    Params P;
    P.heroC = heroC < 1 ? 1 : heroC;
    P.maxBounce = maxBounce >= 1 ? maxBounce : 32;
    P.rrDepth = rrDepth < 0 ? 0 : rrDepth;
    std::vector<Film> films(nThreads);
    std::vector<Energy> energies(nThreads);
    std::vector<Counters> counts(nThreads);
    for (auto& f : films) f.alloc(cd->res_x, cd->res_y);
    auto worker = [&](int tid) {
        Tracer tr{S, cam, P, films[tid], energies[tid], counts[tid]};
        Pcg rng;
        long long lo = N * tid / nThreads, hi = N * (tid + 1) / nThreads;
        for (long long i = lo; i < hi; ++i) {
            seedPhoton(rng, seedBase + (uint64_t)i, salt);
            tr.trace(rng);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nThreads; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
    Film out;
    out.alloc(cd->res_x, cd->res_y);
    for (int t = 0; t < nThreads; ++t)
        for (size_t i = 0; i < out.xyz.size(); ++i) {
            out.xyz[i] = out.xyz[i] + films[t].xyz[i];
            out.hits[i] += films[t].hits[i];
        }
    for (size_t i = 0; i < out.xyz.size(); ++i) {
        if (xyz) { xyz[3 * i] = out.xyz[i].x; xyz[3 * i + 1] = out.xyz[i].y; xyz[3 * i + 2] = out.xyz[i].z; }
        if (hits) hits[i] = out.hits[i];
    }
    if (eOut)
        for (auto& r : energies) {
            eOut->emitted += r.emitted; eOut->absorbed += r.absorbed; eOut->sensor += r.sensor;
            eOut->escaped += r.escaped; eOut->residual += r.residual;
        }
    if (stats7) {
        for (int i = 0; i < 7; ++i) stats7[i] = 0;
        for (auto& c : counts) {
            stats7[0] += c.segments; stats7[1] += c.connections; stats7[2] += c.splats;
            stats7[3] += (uint64_t)c.nodeVisits; stats7[4] += (uint64_t)c.leafTests;
            stats7[5] += (uint64_t)c.anyNodeVisits; stats7[6] += (uint64_t)c.anyLeafTests;
        }
    }
    return 0;
}

int or_closest_hit(void* sp, const double* rays, int64_t n, double tmin, int32_t* prim, double* t,
                   int64_t* nodeVisits, int64_t* leafTests) {
    const Scene& S = *static_cast<Scene*>(sp);
    for (int64_t i = 0; i < n; ++i) {
        V3 o = v3(rays + 6 * i), d = v3(rays + 6 * i + 3);
        int64_t nv = 0, lt = 0;
        Hit h = S.closestHit(o, d, tmin, &nv, &lt);
        prim[i] = h.valid ? h.prim : -1;
        t[i] = h.valid ? h.t : INFINITY;
        if (nodeVisits) nodeVisits[i] = nv;
        if (leafTests) leafTests[i] = lt;
    }
    return 0;
}

int or_closest_hit_linear(void* sp, const double* rays, int64_t n, double tmin, int32_t* prim,
                          double* t) {
    const Scene& S = *static_cast<Scene*>(sp);
    for (int64_t i = 0; i < n; ++i) {
        V3 o = v3(rays + 6 * i), d = v3(rays + 6 * i + 3);
        Hit h;
        const Shear sh = makeShear(d);
        for (size_t k = 0; k < S.tris.size(); ++k) hitTriangle(sh, o, d, S.tris[k], (int)k, tmin, h);
        prim[i] = h.valid ? h.prim : -1;
        t[i] = h.valid ? h.t : INFINITY;
    }
    return 0;
}

// The triangle test with a margin (an extension of this repository, no reference
// counterpart): the edge functions of the sheared frame (the watertight test's U, V,
// W) compared against +-relMargin times their magnitude scale
//   scale = (|Cx|+|Cy|)(|Bx|+|By|) + (|Ax|+|Ay|)(|Cx|+|Cy|) + (|Bx|+|By|)(|Ax|+|Ay|),
// which bounds |U| + |V| + |W|. `grown` = hit with every edge moved outwards by the
// margin, `shrunk` = hit with every edge moved inwards (a robust hit); tt = the
// plane distance. A triangle that is grown- but not shrunk-hit is grazed by the ray.
struct MarginHit {
    bool grown = false, shrunk = false;
    double tt = INFINITY;
};
MarginHit marginTest(const Shear& sh, const V3& o, const Tri& tri, double tmin, double relMargin) {
    MarginHit r;
    const int kx = sh.kx, ky = sh.ky, kz = sh.kz;
    V3 A = tri.v0 - o, B = tri.v1 - o, C = tri.v2 - o;
    double Ax = at(A, kx) - sh.Sx * at(A, kz), Ay = at(A, ky) - sh.Sy * at(A, kz);
    double Bx = at(B, kx) - sh.Sx * at(B, kz), By = at(B, ky) - sh.Sy * at(B, kz);
    double Cx = at(C, kx) - sh.Sx * at(C, kz), Cy = at(C, ky) - sh.Sy * at(C, kz);
    double U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    double scale = (std::fabs(Cx) + std::fabs(Cy)) * (std::fabs(Bx) + std::fabs(By)) +
                   (std::fabs(Ax) + std::fabs(Ay)) * (std::fabs(Cx) + std::fabs(Cy)) +
                   (std::fabs(Bx) + std::fabs(By)) * (std::fabs(Ax) + std::fabs(Ay));
    double mg = relMargin * scale;
    double det = U + V + W;
    if (det == 0.0) return r;
    double sgn = det > 0 ? 1.0 : -1.0;
    if (!((sgn * U >= -mg) && (sgn * V >= -mg) && (sgn * W >= -mg))) return r;
    double T = U * (sh.Sz * at(A, kz)) + V * (sh.Sz * at(B, kz)) + W * (sh.Sz * at(C, kz));
    double tt = T / det;
    if (tt < tmin * (1.0 - relMargin) - relMargin) return r;
    r.grown = true;
    r.tt = tt;
    r.shrunk = (sgn * U >= mg) && (sgn * V >= mg) && (sgn * W >= mg) && tt >= tmin;
    return r;
}

// Ambiguity report for the closest-hit parity probe (an extension of this
// repository, no reference counterpart). For each ray, brute force over all
// triangles:
//   prim, t   the exact closest hit (the watertight test, as or_closest_hit_linear);
//   t2        the nearest grown-hit distance of any OTHER triangle than `prim`;
//   edge      `prim` is grown- but not shrunk-hit (the ray grazes its boundary);
//   tRobust   the nearest shrunk-hit distance of any triangle (inf if none);
//   tq, qGraze  for the queried triangle query[i] (the other side's answer; -1 =
//             none): its grown-hit distance (inf if not grown-hit) and whether it
//             is grazed.
// Every output pointer after `edge` may be null. A float32 closest hit may differ
// from this double one only by a near-tie (t2 within a relative tie tolerance of t)
// or by a grazed triangle; the tests classify each mismatch from these fields.
int or_closest_hit_ambiguity(void* sp, const double* rays, const int32_t* query, int64_t n,
                             double tmin, double relMargin, int32_t* prim, double* t, double* t2,
                             uint8_t* edge, double* tRobust, double* tq, uint8_t* qGraze) {
    const Scene& S = *static_cast<Scene*>(sp);
    auto one = [&](int64_t i) {
        V3 o = v3(rays + 6 * i), d = v3(rays + 6 * i + 3);
        Hit h;
        const Shear sh = makeShear(d);
        for (size_t k = 0; k < S.tris.size(); ++k) hitTriangle(sh, o, d, S.tris[k], (int)k, tmin, h);
        prim[i] = h.valid ? h.prim : -1;
        t[i] = h.valid ? h.t : INFINITY;
        const int q = query ? query[i] : -1;
        double best2 = INFINITY, robust = INFINITY, qt = INFINITY;
        bool grazing = false, qg = false;
        for (size_t k = 0; k < S.tris.size(); ++k) {
            const MarginHit m = marginTest(sh, o, S.tris[k], tmin, relMargin);
            if (!m.grown) continue;
            if (m.shrunk && m.tt < robust) robust = m.tt;
            if ((int)k == q) {
                qt = m.tt;
                qg = !m.shrunk;
            }
            if ((int)k == h.prim) {
                if (!m.shrunk) grazing = true;
                continue;
            }
            if (m.tt < best2) best2 = m.tt;
        }
        t2[i] = best2;
        edge[i] = grazing ? 1 : 0;
        if (tRobust) tRobust[i] = robust;
        if (tq) tq[i] = qt;
        if (qGraze) qGraze[i] = qg ? 1 : 0;
    };
    // rays are independent: split them over the host's threads (large scenes)
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(n, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int64_t w = 0; w < nt; ++w)
        pool.emplace_back([&, w] {
            for (int64_t i = w; i < n; i += nt) one(i);
        });
    for (auto& th : pool) th.join();
    return 0;
}

int or_closest_hit_margin(void* sp, const double* rays, int64_t n, double tmin, double relMargin,
                          int32_t* prim, double* t, double* t2, uint8_t* edge) {
    return or_closest_hit_ambiguity(sp, rays, nullptr, n, tmin, relMargin, prim, t, t2, edge, nullptr,
                                    nullptr, nullptr);
}

int or_occluded(void* sp, const double* rays, const double* maxDist, int64_t n, uint8_t* occ) {
    const Scene& S = *static_cast<Scene*>(sp);
    for (int64_t i = 0; i < n; ++i) occ[i] = S.occluded(v3(rays + 6 * i), v3(rays + 6 * i + 3), maxDist[i]) ? 1 : 0;
    return 0;
}

void or_pcg_stream(uint64_t seq, uint64_t seed, int n, uint32_t* out) {
    Pcg r;
    r.seed(seq, seed);
    for (int i = 0; i < n; ++i) out[i] = r.next();
}
void or_seed_unit_uniforms(uint64_t k, uint64_t salt, int n, double* out) {
    Pcg r;
    seedPhoton(r, k, salt);
    for (int i = 0; i < n; ++i) out[i] = r.uniform();
}
uint64_t or_mix64(uint64_t x) { return splitmix(x); }
void or_spectrum_eval(const gp_spectrum* s, const double* pool, const double* lam, int n, double* out) {
    for (int i = 0; i < n; ++i) out[i] = specEval(*s, lam[i], pool);
}
void or_cie(const double* lam, int n, double* xyz) {
    for (int i = 0; i < n; ++i) { xyz[3 * i] = xbar(lam[i]); xyz[3 * i + 1] = ybar(lam[i]); xyz[3 * i + 2] = zbar(lam[i]); }
}
double or_cie_y_integral() { return ybarIntegral(); }
void or_emission_sample(const gp_spectrum* s, const double* pool, const double* u, int n, double* lam,
                        double* pdf, double* integral) {
    LambdaCdf c;
    c.build(*s, pool);
    for (int i = 0; i < n; ++i) { pdf[i] = 0.0; lam[i] = c.invert(u[i], pdf[i]); }
    if (integral) *integral = c.integral;
}
int or_sample_bundle(const gp_spectrum* s, const double* pool, double u, int C, double* lam, double* pdf) {
    LambdaCdf c;
    c.build(*s, pool);
    return drawBundle(c, u, C, lam, pdf) ? 1 : 0;
}
void or_cosine_hemisphere(const double* nrm, uint64_t seq, uint64_t seed, int n, double* out) {
    Pcg r;
    r.seed(seq, seed);
    V3 nn = v3(nrm);
    for (int i = 0; i < n; ++i) {
        V3 d = cosineDir(nn, r);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}
void or_sample_glossy(const double* mdir, double rough, uint64_t seq, uint64_t seed, int n,
                      double* out) {
    Pcg r;
    r.seed(seq, seed);
    V3 m = v3(mdir);
    for (int i = 0; i < n; ++i) {
        V3 d = glossySample(m, rough, r);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}
void or_camera_basis(const gp_camera_desc* cd, double* out14) {
    Cam c;
    c.lookAt(*cd);
    const double v[14] = {c.eye.x, c.eye.y, c.eye.z, c.u.x, c.u.y, c.u.z, c.v.x, c.v.y, c.v.z,
                          c.w.x, c.w.y, c.w.z, c.tanHalfX, c.tanHalfY};
    std::memcpy(out14, v, sizeof v);
}
int or_camera_project(const gp_camera_desc* cd, const double* pts, int n, int32_t* px, int32_t* py,
                      double* cosCam, double* dist2, double* omega) {
    Cam c;
    c.lookAt(*cd);
    for (int i = 0; i < n; ++i) {
        int x = -1, y = -1;
        double cc = 0, d2 = 0;
        bool ok = c.project(v3(pts + 3 * i), x, y, cc, d2);
        px[i] = ok ? x : -1; py[i] = ok ? y : -1;
        cosCam[i] = ok ? cc : 0.0; dist2[i] = ok ? d2 : 0.0;
        omega[i] = ok ? c.pixelOmega(cc) : 0.0;
    }
    return 0;
}


// ---- film -> image (src/main.cpp:10282-10348, src/color.h:41-53)
// filmToLinear without the optional denoiser: norm = 1 / (N * cieYIntegral()), each
// pixel lin = xyzToLinearSrgb(xyz * norm), in double.
void or_film_to_linear(const double* xyz, int64_t npix, double N, double* lin) {
    const double norm = 1.0 / (N * ybarIntegral());
    for (int64_t i = 0; i < npix; ++i) {
        const double x = xyz[3 * i + 0] * norm, y = xyz[3 * i + 1] * norm, z = xyz[3 * i + 2] * norm;
        lin[3 * i + 0] = 3.2406 * x - 1.5372 * y - 0.4986 * z;
        lin[3 * i + 1] = -0.9689 * x + 1.8758 * y + 0.0415 * z;
        lin[3 * i + 2] = 0.0557 * x - 0.2040 * y + 1.0570 * z;
    }
}

static double srgbGammaD(double c) {
    c = std::max(0.0, c);
    return (c <= 0.0031308) ? 12.92 * c : 1.055 * std::pow(c, 1.0 / 2.4) - 0.055;
}

// filmToRgb8: exposure = eAuto * (expComp > 0 ? expComp : 1) with eAuto = 6 (absolute
// EV), the locked anchor, or 0.9 / p99 where p99 is the element at (size_t)(0.99 (n-1))
// of max(r, g, b, 0) over the pixels (1 when p99 <= 0); 8-bit sRGB, rows flipped so
// row 0 is the image top.
void or_film_to_rgb8(const double* xyz, int W, int H, double N, double expComp, int absolute,
                     double* lockAnchor, uint8_t* img, double* eAutoOut, double* exposureOut) {
    const size_t n = (size_t)W * H;
    std::vector<double> lin(3 * n);
    or_film_to_linear(xyz, (int64_t)n, N, lin.data());
    double eAuto, exposure;
    if (absolute) {
        eAuto = 6.0;
        exposure = eAuto * (expComp > 0.0 ? expComp : 1.0);
    } else {
        if (lockAnchor && *lockAnchor > 0.0) {
            eAuto = *lockAnchor;
        } else {
            std::vector<double> lum(n);
            for (size_t i = 0; i < n; ++i)
                lum[i] = std::max({lin[3 * i], lin[3 * i + 1], lin[3 * i + 2], 0.0});
            const size_t k = (size_t)(0.99 * (double)(n - 1));
            std::nth_element(lum.begin(), lum.begin() + (long)k, lum.end());
            const double p99 = lum[k];
            eAuto = (p99 > 0) ? 0.9 / p99 : 1.0;
            if (lockAnchor) *lockAnchor = eAuto;
        }
        exposure = eAuto * (expComp > 0.0 ? expComp : 1.0);
    }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t src = (size_t)(H - 1 - y) * W + x;
            const size_t dst = ((size_t)y * W + x) * 3;
            for (int c = 0; c < 3; ++c) {
                const double v = lin[3 * src + c] * exposure;
                img[dst + c] = (uint8_t)std::clamp(srgbGammaD(v) * 255.0 + 0.5, 0.0, 255.0);
            }
        }
    if (eAutoOut) *eAutoOut = eAuto;
    if (exposureOut) *exposureOut = exposure;
}


// ---- .ftbuf checkpoints (src/main.cpp:12317-12329 checkpointGuard, :12467-12483
// readCheckpoint). Returns 0 on success, 1 unreadable / not FTBUF01, 2 mismatch of
// resolution, mode or guard (readCheckpoint's "starting fresh"), 3 truncated. On
// success fills the header fields, xyz (rx*ry*3) and hits (rx*ry) when non-null, and
// *trailing = number of bytes after the film blobs (optional sections).
uint64_t or_checkpoint_guard(int64_t nTris, int64_t nSpheres, int64_t nEmitters, double totalPower,
                             int mode, int res, int resY) {
    uint64_t h = 14695981039346656037ULL;
    auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ULL; };
    mix((uint64_t)nTris);
    mix((uint64_t)nSpheres);
    mix((uint64_t)nEmitters);
    uint64_t tp;
    std::memcpy(&tp, &totalPower, sizeof tp);
    mix(tp);
    mix((uint64_t)(unsigned char)mode);
    mix((uint64_t)(unsigned)res);
    mix((uint64_t)(unsigned)resY);
    return h;
}

int or_read_checkpoint(const char* path, int res, int resY, uint64_t guard, int mode, long long* N,
                       double* energy5, double* xyz, double* hits, long long* trailing) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return 1;
    auto rd = [&](void* p, size_t n) { return std::fread(p, 1, n, f) == n; };
    char magic[8];
    int rc = 0;
    int32_t rx = 0, ry = 0, m = 0;
    long long n = 0;
    double en[5] = {0, 0, 0, 0, 0};
    uint64_t g = 0;
    if (!rd(magic, 8) || std::memcmp(magic, "FTBUF01\n", 8) != 0) rc = 1;
    else if (!rd(&rx, 4) || !rd(&ry, 4) || !rd(&m, 4) || !rd(&n, 8) || !rd(en, sizeof en) || !rd(&g, 8)) rc = 3;
    else if (rx != res || ry != resY || m != (int32_t)(unsigned char)mode || g != guard) rc = 2;
    else {
        const size_t np = (size_t)rx * ry;
        std::vector<double> x(np * 3), h(np);
        if (!rd(x.data(), x.size() * sizeof(double)) || !rd(h.data(), h.size() * sizeof(double))) rc = 3;
        else {
            if (xyz) std::memcpy(xyz, x.data(), x.size() * sizeof(double));
            if (hits) std::memcpy(hits, h.data(), h.size() * sizeof(double));
            if (N) *N = n;
            if (energy5) std::memcpy(energy5, en, sizeof en);
            const long pos = std::ftell(f);
            std::fseek(f, 0, SEEK_END);
            if (trailing) *trailing = std::ftell(f) - pos;
        }
    }
    std::fclose(f);
    return rc;
}

}  // extern "C"
