// Flat mirror planes of a scene: host side of the camera connection through mirrors.
//
// The connection (k_mirror) needs planes, not triangles: mirroring the eye in a plane
// makes the reflection point a line/plane intersection. Mirror triangles are grouped
// by plane as the reference's Scene::buildMirrorPlanes does (src/scene.h:1646-1724),
// so both sides connect through the same set of surfaces:
//   - normal: the triangle's geometric normal (normalize(cross(v1 - v0, v2 - v0)) in
//     double, as Tri::finalize), renormalised, then flipped so that its largest
//     component (first of x, y, z on ties) is positive: the two faces of one plane
//     fall into one entry;
//   - key: (round(n * 1e6), round(d * 1e5)) with d = dot(n, v0); triangles with equal
//     keys share an entry; entries are kept in order of first appearance;
//   - box: the bounds of the member vertices, padded by 1e-6 (1 + |hi - lo|);
//   - past kMaxMirrorPlanes entries the largest total triangle areas are kept (a
//     stable sort: equal areas keep their order; the reference's sort is not stable).
// The arithmetic runs in double without contraction so the planes are bit-identical
// to the reference's for the same vertices.
#pragma clang fp contract(off)

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "goodphotons.h"
#include "gp_runtime.hpp"

namespace gp {

namespace {

struct Vd {
    double x, y, z;
};
inline Vd sub(const Vd& a, const Vd& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vd scale(const Vd& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dotd(const Vd& a, const Vd& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vd crossd(const Vd& a, const Vd& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double lengthd(const Vd& a) { return std::sqrt(dotd(a, a)); }

struct Group {
    Vd n, lo, hi;
    double d = 0.0;
    double area = 0.0;
};

}  // namespace

void buildMirrorPlanes(const gp_scene_desc* d, std::vector<DevMirrorPlane>& out) {
    out.clear();
    std::vector<Group> groups;
    std::map<std::array<long long, 4>, int> byKey;
    for (int64_t t = 0; t < d->n_tris; ++t) {
        const int m = d->mat_ids[t];
        if (m < 0 || m >= d->n_materials || d->materials[m].type != GP_MAT_MIRROR) continue;
        const float* f = d->verts + 9 * t;
        const Vd v0{f[0], f[1], f[2]}, v1{f[3], f[4], f[5]}, v2{f[6], f[7], f[8]};
        const Vd c = crossd(sub(v1, v0), sub(v2, v0));
        const double cl = lengthd(c);
        const Vd gn{c.x / cl, c.y / cl, c.z / cl};
        const double nl = lengthd(gn);
        if (!(nl >= 1e-12)) continue;   // degenerate triangle (also catches NaN)
        Vd n = scale(gn, 1.0 / nl);
        const double ax = std::fabs(n.x), ay = std::fabs(n.y), az = std::fabs(n.z);
        const double lead = (ax >= ay && ax >= az) ? n.x : (ay >= az ? n.y : n.z);
        if (lead < 0.0) n = Vd{-n.x, -n.y, -n.z};
        const double off = dotd(n, v0);
        const std::array<long long, 4> key{std::llround(n.x * 1e6), std::llround(n.y * 1e6),
                                           std::llround(n.z * 1e6), std::llround(off * 1e5)};
        auto it = byKey.find(key);
        int g;
        if (it == byKey.end()) {
            g = (int)groups.size();
            byKey.emplace(key, g);
            Group gr;
            gr.n = n;
            gr.d = off;
            gr.lo = v0;
            gr.hi = v0;
            groups.push_back(gr);
        } else {
            g = it->second;
        }
        Group& gr = groups[(size_t)g];
        for (const Vd& v : {v0, v1, v2}) {
            gr.lo = Vd{std::min(gr.lo.x, v.x), std::min(gr.lo.y, v.y), std::min(gr.lo.z, v.z)};
            gr.hi = Vd{std::max(gr.hi.x, v.x), std::max(gr.hi.y, v.y), std::max(gr.hi.z, v.z)};
        }
        gr.area += 0.5 * lengthd(crossd(sub(v1, v0), sub(v2, v0)));
    }
    if ((int)groups.size() > kMaxMirrorPlanes) {
        std::stable_sort(groups.begin(), groups.end(), [](const Group& a, const Group& b) { return a.area > b.area; });
        groups.resize(kMaxMirrorPlanes);
    }
    out.resize(groups.size());
    for (size_t i = 0; i < groups.size(); ++i) {
        const Group& gr = groups[i];
        const double pad = 1e-6 * (1.0 + lengthd(sub(gr.hi, gr.lo)));
        DevMirrorPlane& p = out[i];
        p.n[0] = gr.n.x; p.n[1] = gr.n.y; p.n[2] = gr.n.z;
        p.d = gr.d;
        p.lo[0] = gr.lo.x - pad; p.lo[1] = gr.lo.y - pad; p.lo[2] = gr.lo.z - pad;
        p.hi[0] = gr.hi.x + pad; p.hi[1] = gr.hi.y + pad; p.hi[2] = gr.hi.z + pad;
    }
}

}  // namespace gp

extern "C" int32_t gp_mirror_planes(const gp_scene_desc* d, double* out10, int32_t cap) {
    if (!d || d->n_tris < 0 || (d->n_tris > 0 && (!d->verts || !d->mat_ids)) ||
        (d->n_materials > 0 && !d->materials) || cap < 0 || (cap > 0 && !out10))
        return -(int32_t)gp::fail(GP_ERR_INVALID, "gp_mirror_planes: bad arguments");
    std::vector<gp::DevMirrorPlane> planes;
    gp::buildMirrorPlanes(d, planes);
    for (int i = 0; i < (int)planes.size() && i < cap; ++i) {
        const gp::DevMirrorPlane& p = planes[(size_t)i];
        const double v[10] = {p.n[0], p.n[1], p.n[2], p.d, p.lo[0], p.lo[1], p.lo[2], p.hi[0], p.hi[1], p.hi[2]};
        std::memcpy(out10 + 10 * i, v, sizeof v);
    }
    return (int32_t)planes.size();
}
