// Test infrastructure only: a checker of the drop-in adapter
// (goodphotons_amd/host/ref_adapter.hpp) against the reference renderer itself,
// compiled from the reference headers where they lie (src/render.h and what it
// includes) by oracle/Makefile into oracle/_ref/libref_adapter.so, linked with the
// HIP library.
//
// Scenes are built with the reference's own API, as its scene builders do: Scene::tris
// with materials carrying the reference's Spectrum closures (whiteWall, redWall,
// greenWall, sellmeier, constantSpectrum; src/spectrum.h:74-143), Scene::addAreaLight
// with a blackbody SPD (src/scene.h:1306-1315), Scene::build (finalize, BVH,
// emitters), Camera::lookAt. The adapter then sees exactly the objects
// renderForwardCuda would receive (src/render_cuda.h:79-82). Exports:
//   adapter_scene / adapter_scene_free   build scene 0 (box) or 1 (box + glass, mirror,
//                                       glossy) or a variant outside the supported set
//   adapter_supported                    gpref::forwardSupported and its reason
//   adapter_desc                         the gp_scene_desc the adapter hands over
//   adapter_render_gpu                   gpref::renderForward (the HIP library)
//   adapter_render_reference             the reference's CPU forward path (the loop of
//                                       renderForward's CPU branch, src/main.cpp:11305-11340)
//   adapter_shared_session               gpref::SharedSession over two cameras, as the
//                                       progressive loop drives sharedForwardGpu*
//                                       (src/main.cpp:19041-19236): uninterrupted, and
//                                       stopped after k batches then resumed from the
//                                       downloaded totals as seedFilms / seedEnergy
// glibc declares `extern int daylight` in <time.h>, which collides with the
// reference's `Spectrum daylight(double)` (src/lights.h:32); including <ctime> first
// and renaming the reference's symbol avoids the clash without touching any
// reference file.
#include <ctime>
#define daylight ft_daylight
#include "render.h"
#undef daylight

#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ref_adapter.hpp"

namespace {

struct Handle {
    Scene scene;
    gpref::Upload upload;
    bool described = false;
    std::string why;
};

void quad(Scene& s, Vec3 a, Vec3 b, Vec3 c, Vec3 d, int mat) {
    s.tris.push_back(Tri{a, b, c, mat, -1, {}});
    s.tris.push_back(Tri{a, c, d, mat, -1, {}});
}

// Axis-aligned box of size (sx, sy, sz) at corner (x, y, z), outward-facing quads.
void box(Scene& s, double x, double y, double z, double sx, double sy, double sz, int mat) {
    const Vec3 p000{x, y, z}, p100{x + sx, y, z}, p010{x, y + sy, z}, p110{x + sx, y + sy, z};
    const Vec3 p001{x, y, z + sz}, p101{x + sx, y, z + sz}, p011{x, y + sy, z + sz}, p111{x + sx, y + sy, z + sz};
    quad(s, p000, p010, p110, p100, mat);   // -z
    quad(s, p001, p101, p111, p011, mat);   // +z
    quad(s, p000, p001, p011, p010, mat);   // -x
    quad(s, p100, p110, p111, p101, mat);   // +x
    quad(s, p010, p011, p111, p110, mat);   // +y
    quad(s, p000, p100, p101, p001, mat);   // -y
}

// Octahedron subdivided `level` times and pushed onto a sphere of radius r.
void sphere(Scene& s, Vec3 c, double r, int level, int mat) {
    std::vector<std::array<Vec3, 3>> f;
    const Vec3 X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1}, mX{-1, 0, 0}, mY{0, -1, 0}, mZ{0, 0, -1};
    f = {{Y, Z, X}, {Y, X, mZ}, {Y, mZ, mX}, {Y, mX, Z}, {mY, X, Z}, {mY, mZ, X}, {mY, mX, mZ}, {mY, Z, mX}};
    for (int l = 0; l < level; ++l) {
        std::vector<std::array<Vec3, 3>> g;
        for (auto& t : f) {
            const Vec3 a = normalize(t[0] + t[1]), b = normalize(t[1] + t[2]), d = normalize(t[2] + t[0]);
            g.push_back({t[0], a, d});
            g.push_back({a, t[1], b});
            g.push_back({d, b, t[2]});
            g.push_back({a, b, d});
        }
        f.swap(g);
    }
    for (auto& t : f) s.tris.push_back(Tri{c + t[0] * r, c + t[1] * r, c + t[2] * r, mat, -1, {}});
}

void buildScene(Scene& s, int which) {
    // materials: 0 white, 1 red, 2 green, 3 light geometry (absorbing), 4 N-BK7 glass,
    // 5 glossy, 6 mirror
    Material white, red, green, lightGeo, glass, glossy, mirror;
    white.reflect = whiteWall(0.75);
    red.reflect = redWall();
    green.reflect = greenWall();
    lightGeo.reflect = constantSpectrum(0.0);
    glass.type = MatType::Dielectric;
    glass.ior = sellmeier(1.03961212, 0.231792344, 1.01046945, 0.00600069867, 0.0200179144, 103.560653);
    glossy.type = MatType::Glossy;
    glossy.reflect = whiteWall(0.8);
    glossy.roughness = 0.15;
    mirror.type = MatType::Mirror;
    mirror.reflect = constantSpectrum(0.9);
    s.mats = {white, red, green, lightGeo, glass, glossy, mirror};
    quad(s, {0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, 0);   // floor
    quad(s, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}, 0);   // ceiling
    quad(s, {0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, 1);   // left, red
    quad(s, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, 2);   // right, green
    quad(s, {0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, 0);   // back
    box(s, 0.55, 0.0, 0.45, 0.25, 0.3, 0.25, 0);
    const double lx0 = 0.375, lx1 = 0.625, lz0 = 0.375, lz1 = 0.625, ly = 0.999;
    quad(s, {lx0, ly, lz0}, {lx0, ly, lz1}, {lx1, ly, lz1}, {lx1, ly, lz0}, 3);
    if (which >= 1) {
        sphere(s, {0.3, 0.2, 0.5}, 0.18, 3, 4);
        box(s, 0.2, 0.0, 0.15, 0.2, 0.12, 0.2, 5);
        quad(s, {0.95, 0.2, 0.3}, {0.95, 0.7, 0.3}, {0.95, 0.7, 0.8}, {0.95, 0.2, 0.8}, 6);
    }
    s.addAreaLight({lx0, ly, lz0}, {lx1 - lx0, 0, 0}, {0, 0, lz1 - lz0}, {0, -1, 0},
                   (lx1 - lx0) * (lz1 - lz0), blackbody(6500.0), 1.0);
    // variants outside the supported set (adapter_supported's reasons)
    switch (which) {
        case 10: s.mats[0].reflectTex = 0; break;                               // texture
        case 11: s.mats[4].priority = 1; break;                                 // priority
        case 12: s.mats[4].absorb = constantSpectrum(0.5); break;               // absorption
        case 13: s.mats[5].roughness = 0.0; break;                              // planar Glossy
        case 14: s.mats[0].type = MatType::HalfMirror; break;                   // type
        case 15: s.tris[0].n0 = Vec3{0, 1, 0.1}; break;                         // smooth normals
        case 16: s.spheres.push_back(Sphere{}); break;                          // analytic sphere
        default: break;
    }
    s.build();
}

Camera camera(int res) {
    Camera cam;
    cam.lookAt(Vec3{0.5, 0.5, 2.7}, Vec3{0.5, 0.5, 0.5}, Vec3{0, 1, 0}, 40.0, res, res);
    return cam;
}

void copyFilm(const Film& f, double* xyz, double* hits) {
    for (size_t i = 0; i < f.hits.size(); ++i) {
        if (xyz) { xyz[3 * i] = f.xyz[i].x; xyz[3 * i + 1] = f.xyz[i].y; xyz[3 * i + 2] = f.xyz[i].z; }
        if (hits) hits[i] = f.hits[i];
    }
}

}  // namespace

extern "C" {

void* adapter_scene(int which) {
    auto* h = new Handle();
    buildScene(h->scene, which);
    h->described = gpref::describe(h->scene, h->upload, &h->why);
    return h;
}
void adapter_scene_free(void* p) { delete static_cast<Handle*>(p); }

// 1 if supported; else 0 and the reason in why (cap bytes)
int adapter_supported(void* p, char* why, int cap) {
    auto* h = static_cast<Handle*>(p);
    std::string r;
    const bool ok = gpref::forwardSupported(h->scene, &r);
    if (why && cap > 0) {
        std::strncpy(why, r.c_str(), (size_t)cap - 1);
        why[cap - 1] = 0;
    }
    return ok ? 1 : 0;
}

const gp_scene_desc* adapter_desc(void* p) {
    auto* h = static_cast<Handle*>(p);
    return h->described ? &h->upload.desc : nullptr;
}

void adapter_camera_desc(int res, gp_camera_desc* out) { *out = gpref::describeCamera(camera(res), res, res); }

int64_t adapter_n_tris(void* p) { return (int64_t) static_cast<Handle*>(p)->scene.tris.size(); }

// gpref::renderForward on the HIP library: the film and the accumulated ledger.
void adapter_render_gpu(void* p, int res, long long N, unsigned long long seedBase, int heroC, int maxDepth,
                        int bvh, double* xyz, double* hits, double* e5) {
    auto* h = static_cast<Handle*>(p);
    gpref::gMaxDepth = maxDepth;
    gpref::gBvh = bvh;
    EnergyReport e;
    Film f = gpref::renderForward(h->scene, camera(res), res, res, N, e, true, 'B', seedBase, true, heroC);
    copyFilm(f, xyz, hits);
    const double v[5] = {e.emitted, e.absorbed, e.sensor, e.escaped, e.residual};
    std::memcpy(e5, v, sizeof v);
}

// The reference's CPU forward path on the same scene and photons (renderForward's CPU
// branch: static photon slices per thread, per-thread films merged in order).
void adapter_render_reference(void* p, int res, long long N, unsigned long long seedBase, int heroC,
                              int maxDepth, int nThreads, double* xyz, double* hits, double* e5) {
    auto* h = static_cast<Handle*>(p);
    const Camera cam = camera(res);
    if (nThreads < 1) nThreads = 1;
    std::vector<Film> films(nThreads);
    std::vector<EnergyReport> reports(nThreads);
    for (auto& f : films) { f.resX = res; f.resY = res; f.alloc(); }
    auto worker = [&](int tid) {
        Renderer r;
        r.forwardCatch = false; r.lensMode = false; r.diffraction = true;
        r.useHero = heroC > 1; r.heroC = heroC;
        r.maxBounce = maxDepth;
        Pcg32 rng;
        long long lo = N * tid / nThreads, hi = N * (tid + 1) / nThreads;
        for (long long i = lo; i < hi; ++i) {
            seedUnit(rng, seedBase + (uint64_t)i, gpref::gSalt);
            r.tracePhoton(h->scene, &cam, nullptr, &films[tid], rng, reports[tid]);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nThreads; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
    Film out; out.resX = res; out.resY = res; out.alloc();
    for (int t = 0; t < nThreads; ++t) out.merge(films[t]);
    copyFilm(out, xyz, hits);
    EnergyReport e;
    for (auto& r : reports) {
        e.emitted += r.emitted; e.absorbed += r.absorbed; e.sensor += r.sensor;
        e.escaped += r.escaped; e.residual += r.residual;
    }
    const double v[5] = {e.emitted, e.absorbed, e.sensor, e.escaped, e.residual};
    std::memcpy(e5, v, sizeof v);
}

// Two cameras (the standard view at res, a second view at res2 x res2), `batches`
// batches of N photons. out0/out1: the uninterrupted session's running totals of each
// camera (xyz then hits, 4 doubles per pixel), hits0: its camera-0 poll after the
// last batch, e: its ledger; rout0/rout1 and re: the same from a session stopped after
// k batches, downloaded, and a new session seeded with those totals that ran the
// rest. Returns 1 if every session began, else 0.
int adapter_shared_session(void* p, int res, int res2, long long N, int batches, int k, int heroC, int maxDepth,
                           double* out0, double* out1, double* hits0, double* e, double* rout0, double* rout1,
                           double* re) {
    auto* h = static_cast<Handle*>(p);
    gpref::gMaxDepth = maxDepth;
    Camera c2;
    c2.lookAt(Vec3{0.2, 0.6, 2.5}, Vec3{0.5, 0.4, 0.5}, Vec3{0, 1, 0}, 45.0, res2, res2);
    const std::vector<Camera> cams{camera(res), c2};
    const std::vector<int> rx{res, res2}, ry{res, res2};
    auto films = [&] {
        std::vector<Film> f(2);
        for (int c = 0; c < 2; ++c) { f[c].resX = rx[c]; f[c].resY = ry[c]; f[c].alloc(); }
        return f;
    };
    auto put = [&](const std::vector<Film>& f, double* o0, double* o1) {
        double* o[2] = {o0, o1};
        for (int c = 0; c < 2; ++c) {
            const size_t np = f[c].hits.size();
            copyFilm(f[c], o[c], o[c] + 3 * np);
        }
    };
    auto putE = [](const EnergyReport& r, double* o) {
        o[0] = r.emitted; o[1] = r.absorbed; o[2] = r.sensor; o[3] = r.escaped; o[4] = r.residual;
    };
    {
        gpref::SharedSession s;
        if (!s.begin(h->scene, cams, rx, ry, 'B', true, heroC, false)) return 0;
        for (int b = 0; b < batches; ++b) s.batch(N, (unsigned long long)(N * b), true);
        std::vector<double> hh((size_t)res * res, -1.0);
        s.hits0(hh);
        std::memcpy(hits0, hh.data(), hh.size() * sizeof(double));
        std::vector<Film> f = films();
        EnergyReport r;
        s.download(f, r);
        put(f, out0, out1);
        putE(r, e);
    }
    std::vector<Film> seed = films();
    EnergyReport seedE;
    {
        gpref::SharedSession s;
        if (!s.begin(h->scene, cams, rx, ry, 'B', true, heroC, false)) return 0;
        for (int b = 0; b < k; ++b) s.batch(N, (unsigned long long)(N * b), true);
        s.download(seed, seedE);
    }
    {
        gpref::SharedSession s;
        if (!s.begin(h->scene, cams, rx, ry, 'B', true, heroC, false, &seed, &seedE)) return 0;
        for (int b = k; b < batches; ++b) s.batch(N, (unsigned long long)(N * b), true);
        std::vector<Film> f = films();
        EnergyReport r;
        s.download(f, r);
        put(f, rout0, rout1);
        putE(r, re);
    }
    return 1;
}

}  // extern "C"
