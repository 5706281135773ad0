// The reference renderer's GPU forward entry point over this core, for code that
// holds the reference's own types: include this header after the reference's
// src/render.h (which brings Scene, Camera, Film, EnergyReport, Material, Emitter,
// Bvh) and link libgoodphotons.so. A translation unit that takes the place of
// src/render_cuda.cu forwards the reference's names to these functions:
//
//   renderForwardCuda(scene, cam, resX, resY, N, eOut, diffraction, camMode,
//                     seedBase, wavefront, heroC)       (src/render_cuda.h:79-82)
//       -> gpref::renderForward. A NEW film per call (XYZ un-normalised sums +
//          per-pixel contribution counts, as the CPU renderForward), eOut
//          ACCUMULATED (+=), seedBase = photons already traced; an EMPTY (zeroed,
//          allocated) film when the scene, camera model or hero count is outside the
//          supported set or there is no device (the caller gates with
//          cudaForwardSupported first, as the reference does).
//   cudaForwardSupported(scene)  (src/render_cuda.h:56)  -> gpref::forwardSupported
//   cudaAvailable / cudaDeviceName / cudaMemInfo / cudaGracefulShutdown
//                                (src/render_cuda.h:24-52) -> gpref::available ...
//   sharedForwardGpuBegin / Batch / Hits0 / Download / End (src/render_cuda.h:117-143)
//       -> gpref::SharedSession: scene and 1..16 camera films resident in HBM, each
//          batch one shared photon flight, the camera-0 hit poll, running totals
//          (seed included) on download; begin fails (the caller then stays on the
//          CPU path) outside the supported set.
//
// What crosses the boundary (the role of buildUploadScene, src/render_cuda.cu:12692):
//   - triangles: Scene::tris vertices (double -> float) and material ids, in order;
//   - materials: Diffuse / Dielectric / Mirror / Glossy with their `reflect` and `ior`
//     Spectrum functions sampled at the 470 bin centres 360.5 .. 829.5 nm
//     (GP_SPEC_SAMPLED, interpolated linearly in between, as the reference's GPU
//     backend bakes every Spectrum into a table, bakeSpec / specLookup,
//     src/render_cuda.cu:12658-12663, :1641-1649); at the bin centres the emitter
//     wavelength CDF is then built from exactly the SPD values the reference's
//     EmissionSampler reads, so both draw the same wavelengths;
//   - emitters: quad area lights (origin, u, v, normal, area, collimation, SPD);
//   - the scene's own BVH (Scene::bvh nodes and primIdx, src/bvh.h:70-84), handed
//     over as built (GP_BVH_PREBUILT; its binary nodes are collapsed into the core's
//     four-wide nodes, boxes rounded outwards to float), as the reference copies it
//     to the device verbatim (src/render_cuda.cu:12764-12772);
//   - the camera: eye and orthonormal basis as lookAt left them, tanHalfY.
// The supported set is narrower than the reference's CUDA backend: triangles only,
// flat shading, the four material types without textures, patterns, records,
// priorities or interior absorption, quad emitters without emission patterns, no
// media, pinhole camera model 'B' with the rectilinear projection, hero C <= 4.
// forwardSupported reports the first reason a scene falls outside it.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/goodphotons.h"
#include "gp_forward.hpp"

namespace gpref {

inline int gMaxDepth = 32;                       // Renderer::maxBounce (the CUDA backend hard-codes 32)
inline uint64_t gSalt = 0x9E3779B97F4A7C15ULL;   // the reference's seeding salt (main.cpp:11326)
inline int gBvh = GP_BVH_PREBUILT;               // or GP_BVH_HOST_SAH / GP_BVH_DEVICE_LBVH to rebuild

// Wavelength samples of a sampled spectrum: the emission CDF's bin centres.
constexpr int kSpecSamples = 470;
constexpr double kSpecLo = 360.5, kSpecHi = 829.5;

// Everything a gp_scene_desc points at, kept alive for the call.
struct Upload {
    std::vector<float> verts;
    std::vector<int32_t> matIds;
    std::vector<gp_material> mats;
    std::vector<gp_emitter> ems;
    std::vector<double> samples;
    std::vector<gp_bvh_node> nodes;
    std::vector<int32_t> primIndex;
    gp_scene_desc desc{};
};

inline gp_spectrum sampled(const Spectrum& f, std::vector<double>& pool) {
    gp_spectrum s{};
    s.kind = GP_SPEC_SAMPLED;
    s.table = (int32_t)pool.size();
    s.p[0] = kSpecLo;
    s.p[1] = kSpecHi;
    s.p[2] = kSpecSamples;
    for (int i = 0; i < kSpecSamples; ++i) pool.push_back(f ? f(kSpecLo + (double)i) : 0.0);
    return s;
}

inline bool materialSupported(const Material& m, std::string* why) {
    auto no = [&](const char* r) { if (why) *why = r; return false; };
    if (m.type != MatType::Diffuse && m.type != MatType::Dielectric && m.type != MatType::Mirror &&
        m.type != MatType::Glossy)
        return no("material type outside Diffuse / Dielectric / Mirror / Glossy");
    if (m.type == MatType::Glossy && !(m.roughness > 0.0)) return no("Glossy with roughness <= 0");
    if (m.hasPriority()) return no("dielectric priority");
    if (m.reflectTex >= 0 || m.roughnessTex >= 0 || m.normalTex >= 0 || m.filmThicknessTex >= 0)
        return no("textured material");
    if (m.reflectPat >= 0 || m.roughnessPat >= 0 || m.transmitPat >= 0 || m.mixWeightPat >= 0 ||
        m.filmThicknessPat >= 0)
        return no("pattern-driven material");
    if (!m.recBindings.empty()) return no("record-driven material");
    if (m.type == MatType::Dielectric && m.absorb)
        for (int i = 0; i < kSpecSamples; ++i)
            if (m.absorb(kSpecLo + (double)i) != 0.0) return no("interior absorption");
    return true;
}

// cudaForwardSupported for this core.
inline bool forwardSupported(const Scene& s, std::string* why = nullptr) {
    auto no = [&](const char* r) { if (why) *why = r; return false; };
    if (s.tris.empty()) return no("no triangles");
    if (!s.spheres.empty() || !s.implicits.empty() || !s.curves.empty() || !s.curveSegs.empty() ||
        !s.instances.empty())
        return no("primitives other than triangles");
    if (!s.media.empty()) return no("participating media");
    for (const Tri& t : s.tris) {
        if (t.matId < 0 || t.matId >= (int)s.mats.size()) return no("triangle material id out of range");
        if (t.sensorId >= 0) return no("sensor triangles");
        auto sameAsGn = [&](const Vec3& n) {
            return n.x == t.gn.x && n.y == t.gn.y && n.z == t.gn.z;
        };
        if (!sameAsGn(t.n0) || !sameAsGn(t.n1) || !sameAsGn(t.n2)) return no("smooth shading normals");
    }
    for (const Material& m : s.mats)
        if (!materialSupported(m, why)) return false;
    if (s.emitters.empty()) return no("no emitters");
    for (const Emitter& e : s.emitters) {
        if (e.shape != EmitterShape::Quad) return no("emitter shape other than a quad");
        if (e.emitPat >= 0) return no("emission pattern");
    }
    return true;
}

inline bool describe(const Scene& s, Upload& u, std::string* why = nullptr) {
    if (!forwardSupported(s, why)) return false;
    u = Upload{};
    u.verts.reserve(9 * s.tris.size());
    for (const Tri& t : s.tris) {
        for (const Vec3* v : {&t.v0, &t.v1, &t.v2}) {
            u.verts.push_back((float)v->x);
            u.verts.push_back((float)v->y);
            u.verts.push_back((float)v->z);
        }
        u.matIds.push_back(t.matId);
    }
    for (const Material& m : s.mats) {
        gp_material g{};
        g.type = m.type == MatType::Diffuse      ? GP_MAT_DIFFUSE
                 : m.type == MatType::Dielectric ? GP_MAT_DIELECTRIC
                 : m.type == MatType::Mirror     ? GP_MAT_MIRROR
                                                 : GP_MAT_GLOSSY;
        g.roughness = m.roughness;
        g.reflect = sampled(m.reflect, u.samples);
        g.ior = sampled(m.ior, u.samples);
        u.mats.push_back(g);
    }
    for (const Emitter& e : s.emitters) {
        gp_emitter g{};
        const Vec3* src[4] = {&e.origin, &e.u, &e.v, &e.normal};
        double* dst[4] = {g.origin, g.u, g.v, g.normal};
        for (int k = 0; k < 4; ++k) { dst[k][0] = src[k]->x; dst[k][1] = src[k]->y; dst[k][2] = src[k]->z; }
        g.area = e.area;
        g.collimated = e.collimated ? 1 : 0;
        g.beam_dir[0] = e.beamDir.x; g.beam_dir[1] = e.beamDir.y; g.beam_dir[2] = e.beamDir.z;
        g.spd = sampled(e.spdFn, u.samples);
        u.ems.push_back(g);
    }
    for (const BvhNode& n : s.bvh.nodes) {
        gp_bvh_node g{};
        g.lo[0] = n.box.lo.x; g.lo[1] = n.box.lo.y; g.lo[2] = n.box.lo.z;
        g.hi[0] = n.box.hi.x; g.hi[1] = n.box.hi.y; g.hi[2] = n.box.hi.z;
        g.left = n.left; g.right = n.right; g.first = n.first; g.count = n.count;
        u.nodes.push_back(g);
    }
    u.primIndex.assign(s.bvh.primIdx.begin(), s.bvh.primIdx.end());
    gp_scene_desc& d = u.desc;
    d.n_tris = (int64_t)s.tris.size();
    d.n_materials = (int32_t)u.mats.size();
    d.n_emitters = (int32_t)u.ems.size();
    d.verts = u.verts.data();
    d.mat_ids = u.matIds.data();
    d.materials = u.mats.data();
    d.emitters = u.ems.data();
    d.spectrum_samples = u.samples.data();
    d.n_spectrum_samples = (int64_t)u.samples.size();
    d.bvh_nodes = u.nodes.data();
    d.n_bvh_nodes = (int64_t)u.nodes.size();
    d.bvh_prim_index = u.primIndex.data();
    d.n_bvh_prim_index = (int64_t)u.primIndex.size();
    return true;
}

// The reference's Camera as lookAt left it: target = eye + w, up = v (orthogonal to w,
// so lookAt rebuilds the same u), fov_y from tanHalfY.
inline gp_camera_desc describeCamera(const Camera& c, int resX, int resY) {
    gp_camera_desc d{};
    d.eye[0] = c.eye.x; d.eye[1] = c.eye.y; d.eye[2] = c.eye.z;
    d.target[0] = c.eye.x + c.w.x; d.target[1] = c.eye.y + c.w.y; d.target[2] = c.eye.z + c.w.z;
    d.up[0] = c.v.x; d.up[1] = c.v.y; d.up[2] = c.v.z;
    d.fov_y_deg = 2.0 * std::atan(c.tanHalfY) * 180.0 / 3.141592653589793;
    d.res_x = resX;
    d.res_y = resY;
    return d;
}

inline bool available() {
    gp_device_info info{};
    return gp_device_query(&info) == GP_OK;
}
inline const char* deviceName() {
    static gp_device_info info{};
    return gp_device_query(&info) == GP_OK ? info.arch : "";
}
inline bool memInfo(size_t* freeB, size_t* totalB) {
    gp_device_info info{};
    if (gp_device_query(&info) != GP_OK) return false;
    if (freeB) *freeB = (size_t)info.free_mem;
    if (totalB) *totalB = (size_t)info.total_mem;
    return true;
}
// Nothing to drain: every call below is synchronous and frees what it allocated.
inline void gracefulShutdown() {}

inline Film emptyFilm(int resX, int resY) {
    Film f;
    f.resX = resX;
    f.resY = resY;
    f.alloc();
    return f;
}

inline Film renderForward(const Scene& scene, const Camera& cam, int resX, int resY, long long N,
                          EnergyReport& eOut, bool diffraction, char camMode, unsigned long long seedBase = 0,
                          bool wavefront = false, int heroC = 1) {
    (void)diffraction;   // only gratings read it (outside the supported set)
    (void)wavefront;     // the core always runs as a wavefront
    Film out = emptyFilm(resX, resY);
    if (camMode != 'B' || heroC < 1 || heroC > 4 || N < 0) return out;
    Upload u;
    std::string why;
    if (!describe(scene, u, &why)) return out;
    gp_scene* gs = nullptr;
    if (gp_scene_create_ex(&u.desc, gBvh, &gs) != GP_OK) {
        std::fprintf(stderr, "gpref: scene upload failed: %s\n", gp_last_error());
        return out;
    }
    gp_film* film = nullptr;
    if (gp_film_create(resX, resY, N > 0 ? N : 1, &film) != GP_OK) {
        std::fprintf(stderr, "gpref: film allocation failed: %s\n", gp_last_error());
        gp_scene_destroy(gs);
        return out;
    }
    const gp_camera_desc cd = describeCamera(cam, resX, resY);
    gp_trace_params p{};
    p.n_photons = N;
    p.seed_base = seedBase;
    p.salt = gSalt;
    p.hero_c = heroC;
    p.max_depth = gMaxDepth;
    p.rr_depth = 0;
    gp_energy e{};
    const gp_status st = gp_trace(gs, &cd, film, &p, &e, nullptr);
    if (st == GP_OK) {
        std::vector<double> xyz((size_t)resX * resY * 3), hits((size_t)resX * resY);
        if (gp_film_download(film, xyz.data(), hits.data()) == GP_OK) {
            for (size_t i = 0; i < hits.size(); ++i) {
                out.xyz[i] = Vec3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
                out.hits[i] = hits[i];
            }
            eOut.emitted += e.emitted;
            eOut.absorbed += e.absorbed;
            eOut.sensor += e.sensor;
            eOut.escaped += e.escaped;
            eOut.residual += e.residual;
        }
    } else {
        std::fprintf(stderr, "gpref: trace failed: %s\n", gp_last_error());
    }
    gp_film_destroy(film);
    gp_scene_destroy(gs);
    return out;
}

// The reference's resident shared session (src/render_cuda.h:103-143). Its films
// accumulate without a photon limit; the core's fixed-point films are sized for
// gSessionMaxPhotons photons in all (2^36 by default: the largest single splat is then
// 2^20 units of the film's scale at the reference's maxBounce of 32).
inline long long gSessionMaxPhotons = 1LL << 36;

class SharedSession {
public:
    // seedFilms / seedEnergy: the running totals of a resumed render (a checkpoint the
    // reference wrote); downloads then return totals including them. The photons they
    // hold, counted in the films' budget, follow from the emitted energy (each photon
    // carries the scene's total power).
    bool begin(const Scene& scene, const std::vector<Camera>& cams, const std::vector<int>& resX,
               const std::vector<int>& resY, char camMode, bool wavefront, int heroC, bool beamGather,
               const std::vector<Film>* seedFilms = nullptr, const EnergyReport* seedEnergy = nullptr) {
        (void)wavefront;    // the core always runs as a wavefront
        (void)beamGather;   // collimated-beam gathering is outside the supported set
        if (camMode != 'B' || heroC < 1 || heroC > 4 || cams.empty() || cams.size() > 16 ||
            resX.size() != cams.size() || resY.size() != cams.size())
            return false;
        Upload u;
        if (!describe(scene, u)) return false;
        std::vector<gp_camera_desc> cd;
        for (size_t c = 0; c < cams.size(); ++c) cd.push_back(describeCamera(cams[c], resX[c], resY[c]));
        if (!core_.open(u.desc, gBvh, cd, gSessionMaxPhotons)) return false;
        heroC_ = heroC;
        if (seedFilms) {
            const double P = core_.totalPower();
            const long long nSeed = (seedEnergy && P > 0.0) ? std::llround(seedEnergy->emitted / P) : 0;
            if (seedFilms->size() != cams.size()) { core_.close(); return false; }
            for (size_t c = 0; c < cams.size(); ++c) {
                const Film& f = (*seedFilms)[c];
                const size_t np = (size_t)resX[c] * resY[c];
                if (f.hits.size() != np || f.xyz.size() != np) { core_.close(); return false; }
                std::vector<double> xyz(np * 3);
                for (size_t i = 0; i < np; ++i) {
                    xyz[3 * i] = f.xyz[i].x; xyz[3 * i + 1] = f.xyz[i].y; xyz[3 * i + 2] = f.xyz[i].z;
                }
                if (!core_.seed(c, xyz.data(), f.hits.data(), nSeed, gMaxDepth)) { core_.close(); return false; }
            }
        }
        if (seedEnergy)
            core_.seedEnergy(gp_energy{seedEnergy->emitted, seedEnergy->absorbed, seedEnergy->sensor,
                                       seedEnergy->escaped, seedEnergy->residual});
        return true;
    }
    // N more photons (absolute indices seedBase ..) into every camera's film.
    void batch(long long N, unsigned long long seedBase, bool diffraction) {
        (void)diffraction;
        core_.batch(N, seedBase, gSalt, heroC_, gMaxDepth, 0);
    }
    // Camera 0's running hit counts (hits must already hold its pixel count).
    void hits0(std::vector<double>& hits) {
        if (hits.size() != (size_t)core_.camera(0).res_x * core_.camera(0).res_y) return;
        core_.hits0(hits.data());
    }
    // Running totals into films[c] (allocated at each camera's resolution); e is
    // REPLACED with the total ledger.
    void download(std::vector<Film>& films, EnergyReport& e) {
        for (size_t c = 0; c < core_.cameras() && c < films.size(); ++c) {
            const size_t np = (size_t)core_.camera(c).res_x * core_.camera(c).res_y;
            if (films[c].hits.size() != np || films[c].xyz.size() != np) continue;
            std::vector<double> xyz(np * 3);
            if (!core_.download(c, xyz.data(), films[c].hits.data())) continue;
            for (size_t i = 0; i < np; ++i) films[c].xyz[i] = Vec3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        }
        const gp_energy& g = core_.energy();
        e.emitted = g.emitted;
        e.absorbed = g.absorbed;
        e.sensor = g.sensor;
        e.escaped = g.escaped;
        e.residual = g.residual;
    }
    void end() { core_.close(); }

private:
    gp::SessionCore core_;
    int heroC_ = 1;
};

}  // namespace gpref
