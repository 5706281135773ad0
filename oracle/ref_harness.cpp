// Test infrastructure only: a C-ABI harness around the REFERENCE renderer's own
// CPU forward path, compiled from the reference headers where they lie
// (src/render.h and what it includes) by oracle/Makefile into oracle/_ref/.
// It is a checker for the oracle restatement and the CPU baseline; nothing in the
// product links, loads or calls it.
//
// The harness builds a reference `Scene` from the same gp_scene_desc the GPU
// library consumes, then runs exactly the reference's code:
//   renderForward CPU branch          src/main.cpp:11289-11341 (loop restated here
//                                     because the function is file-static)
//   Renderer::tracePhoton             src/render.h:1811
//   Scene::closestHit / occluded      src/scene.h:2194, :2249
//   Bvh::traverseClosest with stats   src/bvh.h:180-224
// glibc declares `extern int daylight` in <time.h>, which collides with the
// reference's `Spectrum daylight(double)` (src/lights.h:32); including <ctime>
// first and renaming the reference's symbol avoids the clash without touching
// any reference file.
#include <ctime>
#define daylight ft_daylight
#include "render.h"
#undef daylight

#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/goodphotons.h"

namespace {

// `pool`: the samples of GP_SPEC_SAMPLED spectra (gp_scene_desc.spectrum_samples);
// a sampled spectrum becomes a closure interpolating its own copy of them, as the
// reference's GPU backend interpolates its baked tables (specLookup,
// src/render_cuda.cu:1641-1649).
Spectrum toSpectrum(const gp_spectrum& s, const double* pool = nullptr) {
    Spectrum base;
    switch (s.kind) {
        case GP_SPEC_SAMPLED: {
            const int n = (int)s.p[2];
            const double lo = s.p[0], hi = s.p[1];
            std::vector<double> tab(pool ? pool + s.table : nullptr, pool ? pool + s.table + n : nullptr);
            if (tab.empty()) { base = constantSpectrum(0.0); break; }
            base = [tab, n, lo, hi](double w) {
                double f = (w - lo) * (double)(n - 1) / (hi - lo);
                if (!(f > 0.0)) return tab[0];
                if (f >= (double)(n - 1)) return tab[(size_t)n - 1];
                int i = (int)f;
                double fr = f - (double)i;
                return tab[(size_t)i] * (1.0 - fr) + tab[(size_t)i + 1] * fr;
            };
            break;
        }
        case GP_SPEC_CONST: base = constantSpectrum(s.p[0]); break;
        case GP_SPEC_STEP: {
            const double lo = s.p[0], hi = s.p[1], edge = s.p[2];
            base = [=](double w) { return (w < edge) ? lo : hi; };
            break;
        }
        case GP_SPEC_GAUSS: {
            const double b = s.p[0], a = s.p[1], mu = s.p[2], sg = s.p[3];
            base = [=](double w) { double t = (w - mu) / sg; return b + a * std::exp(-0.5 * t * t); };
            break;
        }
        case GP_SPEC_BLACKBODY:
            if (s.p[1] == 1.0) base = blackbody(s.p[0]);
            else {
                const double T = s.p[0], k = s.p[1];
                base = [=](double w) { return k * blackbodyRadiance(T, w); };
            }
            break;
        case GP_SPEC_SELLMEIER:
            base = sellmeier(s.p[0], s.p[1], s.p[2], s.p[3], s.p[4], s.p[5]);
            break;
        case GP_SPEC_CAUCHY: base = cauchy(s.p[0], s.p[1]); break;
        default: base = constantSpectrum(0.0); break;
    }
    if (s.band_hi > s.band_lo) {
        const double lo = s.band_lo, hi = s.band_hi;
        Spectrum inner = base;
        return [=](double w) { return (w >= lo && w <= hi) ? inner(w) : 0.0; };
    }
    return base;
}

struct RefScene {
    Scene scene;
};

Camera toCamera(const gp_camera_desc& c) {
    Camera cam;
    cam.lookAt(Vec3{c.eye[0], c.eye[1], c.eye[2]}, Vec3{c.target[0], c.target[1], c.target[2]},
               Vec3{c.up[0], c.up[1], c.up[2]}, c.fov_y_deg, c.res_x, c.res_y);
    return cam;
}

}  // namespace

extern "C" {

void* ref_scene_create(const gp_scene_desc* d) {
    auto* rs = new RefScene();
    Scene& s = rs->scene;
    for (int i = 0; i < d->n_materials; ++i) {
        const gp_material& gm = d->materials[i];
        Material m;
        m.type = static_cast<MatType>(gm.type);
        m.reflect = toSpectrum(gm.reflect, d->spectrum_samples);
        m.ior = toSpectrum(gm.ior, d->spectrum_samples);
        m.roughness = gm.roughness;
        s.mats.push_back(m);
    }
    s.tris.reserve((size_t)d->n_tris);
    for (int64_t i = 0; i < d->n_tris; ++i) {
        const float* v = d->verts + 9 * i;
        Tri t{Vec3{v[0], v[1], v[2]}, Vec3{v[3], v[4], v[5]}, Vec3{v[6], v[7], v[8]},
              d->mat_ids[i], -1, {}};
        s.tris.push_back(t);
    }
    for (int i = 0; i < d->n_emitters; ++i) {
        const gp_emitter& e = d->emitters[i];
        s.addAreaLight(Vec3{e.origin[0], e.origin[1], e.origin[2]},
                       Vec3{e.u[0], e.u[1], e.u[2]}, Vec3{e.v[0], e.v[1], e.v[2]},
                       Vec3{e.normal[0], e.normal[1], e.normal[2]}, e.area, toSpectrum(e.spd, d->spectrum_samples),
                       1.0, e.collimated != 0,
                       Vec3{e.beam_dir[0], e.beam_dir[1], e.beam_dir[2]});
    }
    s.build();
    return rs;
}

void ref_scene_destroy(void* p) { delete static_cast<RefScene*>(p); }

int64_t ref_scene_bvh_nodes(void* p) { return (int64_t) static_cast<RefScene*>(p)->scene.bvh.nodes.size(); }
double ref_scene_total_power(void* p) { return static_cast<RefScene*>(p)->scene.totalPower; }
// Scene::mirrorPlanes (src/scene.h:1617-1624) as n.xyz, d, lo.xyz, hi.xyz per plane.
int ref_scene_mirror_planes(void* p, double* out10, int cap) {
    const auto& m = static_cast<RefScene*>(p)->scene.mirrorPlanes;
    for (int i = 0; i < (int)m.size() && i < cap; ++i) {
        const auto& q = m[(size_t)i];
        const double v[10] = {q.n.x, q.n.y, q.n.z, q.d, q.lo.x, q.lo.y, q.lo.z, q.hi.x, q.hi.y, q.hi.z};
        std::memcpy(out10 + 10 * i, v, sizeof v);
    }
    return (int)m.size();
}

// The CPU branch of renderForward (src/main.cpp:11305-11340): T threads, static
// contiguous slices, per-thread films merged in thread order, photon i seeded by
// seedUnit(seedBase + i, salt). ray_queries gets segments + connections (the
// -raystats count, src/scene.h:23-37).
int ref_render(void* p, const gp_camera_desc* cd, int64_t N, uint64_t seedBase, uint64_t salt,
               int heroC, int maxBounce, int nThreads, double* xyz, double* hits,
               gp_energy* eOut, uint64_t* rayQueries) {
    const Scene& scene = static_cast<RefScene*>(p)->scene;
    Camera cam = toCamera(*cd);
    const int resX = cd->res_x, resY = cd->res_y;
    if (nThreads < 1) nThreads = 1;
    std::vector<Film> films(nThreads);
    std::vector<EnergyReport> reports(nThreads);
    std::vector<unsigned long long> rq(nThreads, 0);
    for (auto& f : films) { f.resX = resX; f.resY = resY; f.alloc(); }
    auto worker = [&](int tid) {
        raystats::tls = 0;
        Renderer r;
        r.forwardCatch = false; r.lensMode = false; r.diffraction = true;
        r.useHero = heroC > 1; r.heroC = heroC;
        if (maxBounce >= 1) r.maxBounce = maxBounce;
        Pcg32 rng;
        long long lo = N * tid / nThreads, hi = N * (tid + 1) / nThreads;
        for (long long i = lo; i < hi; ++i) {
            seedUnit(rng, seedBase + (uint64_t)i, salt);
            r.tracePhoton(scene, &cam, nullptr, &films[tid], rng, reports[tid]);
        }
        rq[tid] = raystats::tls;
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nThreads; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
    Film out; out.resX = resX; out.resY = resY; out.alloc();
    for (int t = 0; t < nThreads; ++t) out.merge(films[t]);
    const size_t np = (size_t)resX * resY;
    for (size_t i = 0; i < np; ++i) {
        if (xyz) { xyz[3 * i] = out.xyz[i].x; xyz[3 * i + 1] = out.xyz[i].y; xyz[3 * i + 2] = out.xyz[i].z; }
        if (hits) hits[i] = out.hits[i];
    }
    if (eOut) {
        for (auto& r : reports) {
            eOut->emitted += r.emitted; eOut->absorbed += r.absorbed; eOut->sensor += r.sensor;
            eOut->escaped += r.escaped; eOut->residual += r.residual;
        }
    }
    if (rayQueries) { unsigned long long s = 0; for (auto v : rq) s += v; *rayQueries = s; }
    return 0;
}

// Closest hit with the primitive id recovered through a custom leaf callback on
// the public Bvh::traverseClosest template (the Hit struct carries no prim id).
// The leaf body is Scene::closestHit's triangle branch (src/scene.h:2219).
int ref_closest_hit(void* p, const double* rays, int64_t n, double tmin, int32_t* prim,
                    double* tOut, int64_t* nodeVisits, int64_t* leafTests) {
    const Scene& scene = static_cast<RefScene*>(p)->scene;
    for (int64_t i = 0; i < n; ++i) {
        Ray r{Vec3{rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]},
              Vec3{rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]}};
        Hit h;
        double tMax = DBL_MAX;
        int best = -1;
        const TriShear sh = makeTriShear(r.d);
        TraversalStats st;
        scene.bvh.traverseClosest(r, tmin, tMax, [&](int pr, double& tm) {
            if (intersectTri(sh, r, scene.tris[pr], tmin, h)) { tm = h.t; best = pr; }
        }, &st);
        prim[i] = h.valid ? best : -1;
        tOut[i] = h.valid ? h.t : INFINITY;
        if (nodeVisits) nodeVisits[i] = st.nodeVisits;
        if (leafTests) leafTests[i] = st.leafTests;
    }
    return 0;
}

// Linear-scan reference (Scene::closestHitLinear, src/scene.h:2470) with prim id.
int ref_closest_hit_linear(void* p, const double* rays, int64_t n, double tmin, int32_t* prim,
                           double* tOut) {
    const Scene& scene = static_cast<RefScene*>(p)->scene;
    for (int64_t i = 0; i < n; ++i) {
        Ray r{Vec3{rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]},
              Vec3{rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]}};
        Hit h;
        int best = -1;
        const TriShear sh = makeTriShear(r.d);
        for (size_t k = 0; k < scene.tris.size(); ++k)
            if (intersectTri(sh, r, scene.tris[k], tmin, h)) best = (int)k;
        prim[i] = h.valid ? best : -1;
        tOut[i] = h.valid ? h.t : INFINITY;
    }
    return 0;
}

int ref_occluded(void* p, const double* rays, const double* maxDist, int64_t n, uint8_t* occ) {
    const Scene& scene = static_cast<RefScene*>(p)->scene;
    for (int64_t i = 0; i < n; ++i) {
        Vec3 o{rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]};
        Vec3 d{rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]};
        occ[i] = scene.occluded(o, d, maxDist[i]) ? 1 : 0;
    }
    return 0;
}

void ref_pcg_stream(uint64_t seq, uint64_t seed, int n, uint32_t* out) {
    Pcg32 r; r.seed(seq, seed);
    for (int i = 0; i < n; ++i) out[i] = r.next();
}
void ref_seed_unit_uniforms(uint64_t k, uint64_t salt, int n, double* out) {
    Pcg32 r; seedUnit(r, k, salt);
    for (int i = 0; i < n; ++i) out[i] = r.uniform();
}
uint64_t ref_mix64(uint64_t x) { return mix64(x); }

void ref_spectrum_eval(const gp_spectrum* s, const double* pool, const double* lam, int n, double* out) {
    Spectrum f = toSpectrum(*s, pool);
    for (int i = 0; i < n; ++i) out[i] = f(lam[i]);
}
void ref_cie(const double* lam, int n, double* xyz) {
    for (int i = 0; i < n; ++i) { xyz[3 * i] = cieX(lam[i]); xyz[3 * i + 1] = cieY(lam[i]); xyz[3 * i + 2] = cieZ(lam[i]); }
}
double ref_cie_y_integral() { return cieYIntegral(); }

void ref_emission_sample(const gp_spectrum* s, const double* pool, const double* u, int n, double* lam,
                         double* pdf, double* integral) {
    EmissionSampler es; es.build(toSpectrum(*s, pool), 1.0);
    for (int i = 0; i < n; ++i) { pdf[i] = 0.0; lam[i] = es.sampleAt(u[i], pdf[i]); }
    if (integral) *integral = es.integral;
}
int ref_sample_bundle(const gp_spectrum* s, const double* pool, double u, int C, double* lam, double* pdf) {
    EmissionSampler es; es.build(toSpectrum(*s, pool), 1.0);
    return hero::sampleBundle(es, u, C, lam, pdf) ? 1 : 0;
}
// n cosine-weighted directions about `nrm` (src/rng.h:45-52), each drawn from a
// Pcg32 seeded (seq, seed) in sequence.
void ref_cosine_hemisphere(const double* nrm, uint64_t seq, uint64_t seed, int n, double* out) {
    Pcg32 r; r.seed(seq, seed);
    Vec3 nn{nrm[0], nrm[1], nrm[2]};
    for (int i = 0; i < n; ++i) {
        Vec3 d = cosineHemisphere(nn, r);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}
// Glossy lobe directions (src/render.h:46-64) from a seeded Pcg32.
void ref_sample_glossy(const double* mdir, double rough, uint64_t seq, uint64_t seed, int n,
                       double* out) {
    Pcg32 r; r.seed(seq, seed);
    Vec3 m{mdir[0], mdir[1], mdir[2]};
    for (int i = 0; i < n; ++i) {
        Vec3 d = sampleGlossy(m, rough, r);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}
// Camera basis and projection (src/camera.h:110-190).
void ref_camera_basis(const gp_camera_desc* cd, double* out14) {
    Camera c = toCamera(*cd);
    const double v[14] = {c.eye.x, c.eye.y, c.eye.z, c.u.x, c.u.y, c.u.z, c.v.x, c.v.y, c.v.z,
                          c.w.x, c.w.y, c.w.z, c.tanHalfX, c.tanHalfY};
    std::memcpy(out14, v, sizeof v);
}
int ref_camera_project(const gp_camera_desc* cd, const double* pts, int n, int32_t* px, int32_t* py,
                       double* cosCam, double* dist2, double* omega) {
    Camera c = toCamera(*cd);
    for (int i = 0; i < n; ++i) {
        int x = -1, y = -1; double cc = 0, d2 = 0;
        bool ok = c.project(Vec3{pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]}, x, y, cc, d2);
        px[i] = ok ? x : -1; py[i] = ok ? y : -1;
        cosCam[i] = ok ? cc : 0.0; dist2[i] = ok ? d2 : 0.0;
        omega[i] = ok ? c.pixelSolidAngle(cc) : 0.0;
    }
    return 0;
}

}  // extern "C"
