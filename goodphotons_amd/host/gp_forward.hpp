// C++20 host entry point of the forward light-tracing core, over the C ABI
// (include/goodphotons.h). It keeps the contract of the reference renderer's
// GPU forward mode (src/render_cuda.h):
//
//   Film renderForwardGpu(scene, cam, resX, resY, N, eOut, diffraction, camMode,
//                         seedBase = 0, wavefront = false, heroC = 1)
//       <- renderForwardCuda (src/render_cuda.h:79-82): a NEW film per call (XYZ
//          un-normalised sums + per-pixel hit counts), eOut ACCUMULATED (+=),
//          seedBase = photons already traced, an EMPTY (zeroed, allocated) film when
//          the scene or camera model is unsupported or there is no device.
//   gpuAvailable / gpuDeviceName / gpuMemInfo / gpuForwardSupported
//       <- cudaAvailable / cudaDeviceName / cudaMemInfo / cudaForwardSupported
//          (src/render_cuda.h:24-56)
//   ForwardSession (begin / batch / download / end)
//       <- sharedForwardGpuBegin / Batch / Download / End (src/render_cuda.h:121-143),
//          one camera: scene and film resident in HBM between batches.
//   SharedSession (begin / batch / hits0 / download / end)
//       <- the same session with 1..16 cameras sharing each batch's photon flight,
//          the camera-0 noise poll (sharedForwardGpuHits0) and the resume seed
//          (begin's seedFilms / seedEnergy). SessionCore holds its device state for
//          any host types (ref_adapter.hpp wraps it for the reference's).
//
// The scene types mirror the subset of the reference's Scene the core supports
// (triangles, Diffuse / Dielectric / Mirror / Glossy materials with analytic or sampled
// spectra, quad area lights, pinhole camera). A caller holding the reference's own
// Scene/Camera uses ref_adapter.hpp instead, which converts them itself.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/goodphotons.h"

namespace gp {

struct Vec3d {
    double x = 0, y = 0, z = 0;
};

struct Tri {
    Vec3d v0, v1, v2;
    int matId = 0;
};

struct Scene {
    std::vector<Tri> tris;
    std::vector<gp_material> mats;
    std::vector<gp_emitter> emitters;
    std::vector<double> spectrumSamples;   // samples of GP_SPEC_SAMPLED spectra
};

struct Camera {
    Vec3d eye, target, up{0, 1, 0};
    double fovYDeg = 40.0;
};

struct Film {
    int resX = 0, resY = 0;
    std::vector<Vec3d> xyz;     // accumulated XYZ per pixel (un-normalised)
    std::vector<double> hits;   // contributions per pixel
    void alloc() {
        xyz.assign((size_t)resX * resY, Vec3d{});
        hits.assign((size_t)resX * resY, 0.0);
    }
};

struct EnergyReport {
    double emitted = 0, absorbed = 0, sensor = 0, escaped = 0, residual = 0;
};

inline uint64_t kReferenceSalt = 0x9E3779B97F4A7C15ULL;
inline int gMaxDepth = 32;      // Renderer::maxBounce default
inline int gRrDepth = 0;        // Russian roulette from the first vertex, as the reference

namespace detail {
struct SceneHandle {
    gp_scene* s = nullptr;
    ~SceneHandle() { if (s) gp_scene_destroy(s); }
};
struct FilmHandle {
    gp_film* f = nullptr;
    ~FilmHandle() { if (f) gp_film_destroy(f); }
};
inline gp_scene_desc describe(const Scene& sc, std::vector<float>& verts, std::vector<int32_t>& mats) {
    verts.resize(sc.tris.size() * 9);
    mats.resize(sc.tris.size());
    for (size_t i = 0; i < sc.tris.size(); ++i) {
        const Tri& t = sc.tris[i];
        const double v[9] = {t.v0.x, t.v0.y, t.v0.z, t.v1.x, t.v1.y, t.v1.z, t.v2.x, t.v2.y, t.v2.z};
        for (int k = 0; k < 9; ++k) verts[9 * i + k] = (float)v[k];
        mats[i] = t.matId;
    }
    gp_scene_desc d{};
    d.n_tris = (int64_t)sc.tris.size();
    d.n_materials = (int32_t)sc.mats.size();
    d.n_emitters = (int32_t)sc.emitters.size();
    d.verts = verts.data();
    d.mat_ids = mats.data();
    d.materials = sc.mats.data();
    d.emitters = sc.emitters.data();
    d.spectrum_samples = sc.spectrumSamples.empty() ? nullptr : sc.spectrumSamples.data();
    d.n_spectrum_samples = (int64_t)sc.spectrumSamples.size();
    return d;
}
inline gp_camera_desc camera(const Camera& c, int resX, int resY) {
    gp_camera_desc d{};
    const Vec3d* v[3] = {&c.eye, &c.target, &c.up};
    double* o[3] = {d.eye, d.target, d.up};
    for (int k = 0; k < 3; ++k) { o[k][0] = v[k]->x; o[k][1] = v[k]->y; o[k][2] = v[k]->z; }
    d.fov_y_deg = c.fovYDeg;
    d.res_x = resX;
    d.res_y = resY;
    return d;
}
inline void report(const char* what) {
    std::fprintf(stderr, "[goodphotons] %s: %s\n", what, gp_last_error());
}
inline bool download(gp_film* f, Film& out) {
    std::vector<double> xyz((size_t)out.resX * out.resY * 3);
    if (gp_film_download(f, xyz.data(), out.hits.data()) != GP_OK) return false;
    for (size_t i = 0; i < out.xyz.size(); ++i) out.xyz[i] = Vec3d{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    return true;
}
}  // namespace detail

inline bool gpuAvailable() {
    gp_device_info i;
    return gp_device_query(&i) == GP_OK;
}
inline std::string gpuDeviceName() {
    gp_device_info i;
    if (gp_device_query(&i) != GP_OK) return "none";
    return std::string(i.name[0] ? i.name : i.arch);
}
inline bool gpuMemInfo(size_t* freeB, size_t* totalB) {
    gp_device_info i;
    if (gp_device_query(&i) != GP_OK) return false;
    if (freeB) *freeB = (size_t)i.free_mem;
    if (totalB) *totalB = (size_t)i.total_mem;
    return true;
}
// The support gate (the reference's cudaForwardSupported): validates by building.
inline bool gpuForwardSupported(const Scene& sc) {
    std::vector<float> v;
    std::vector<int32_t> m;
    gp_scene_desc d = detail::describe(sc, v, m);
    detail::SceneHandle h;
    return gp_scene_create(&d, &h.s) == GP_OK;
}

inline Film renderForwardGpu(const Scene& scene, const Camera& cam, int resX, int resY, long long N,
                             EnergyReport& eOut, bool diffraction, char camMode,
                             unsigned long long seedBase = 0, bool wavefront = false, int heroC = 1) {
    (void)diffraction;   // gratings are outside the supported material set
    (void)wavefront;     // the core always runs the wavefront
    Film film;
    film.resX = resX;
    film.resY = resY;
    film.alloc();
    if (camMode != 'B') {
        std::fprintf(stderr, "[goodphotons] camera model %c not supported (pinhole 'B' only)\n", camMode);
        return film;
    }
    std::vector<float> v;
    std::vector<int32_t> m;
    gp_scene_desc d = detail::describe(scene, v, m);
    detail::SceneHandle sh;
    if (gp_scene_create(&d, &sh.s) != GP_OK) { detail::report("scene"); return film; }
    detail::FilmHandle fh;
    if (gp_film_create(resX, resY, N > 0 ? N : 1, &fh.f) != GP_OK) { detail::report("film"); return film; }
    gp_camera_desc cd = detail::camera(cam, resX, resY);
    gp_trace_params p{};
    p.n_photons = N;
    p.seed_base = seedBase;
    p.salt = kReferenceSalt;
    p.hero_c = heroC < 1 ? 1 : (heroC > 4 ? 4 : heroC);
    p.max_depth = gMaxDepth;
    p.rr_depth = gRrDepth;
    gp_energy e{};
    if (gp_trace(sh.s, &cd, fh.f, &p, &e, nullptr) != GP_OK) { detail::report("trace"); return film; }
    if (!detail::download(fh.f, film)) { detail::report("download"); film.alloc(); return film; }
    eOut.emitted += e.emitted;
    eOut.absorbed += e.absorbed;
    eOut.sensor += e.sensor;
    eOut.escaped += e.escaped;
    eOut.residual += e.residual;
    return film;
}

// Shared photon flight into several cameras, as renderForwardSharedCuda
// (src/render_cuda.h, src/render_cuda.cu:13873-13897): each returned film equals the
// single-camera renderForwardGpu of the same photons. Up to 16 cameras.
inline std::vector<Film> renderForwardSharedGpu(const Scene& scene, const std::vector<Camera>& cams,
                                                const std::vector<int>& resX, const std::vector<int>& resY,
                                                long long N, EnergyReport& eOut, bool diffraction,
                                                char camMode, unsigned long long seedBase = 0,
                                                bool wavefront = false, int heroC = 1) {
    (void)diffraction;
    (void)wavefront;
    const size_t nc = cams.size();
    std::vector<Film> out(nc);
    for (size_t c = 0; c < nc; ++c) {
        out[c].resX = resX[c];
        out[c].resY = resY[c];
        out[c].alloc();
    }
    if (camMode != 'B' || nc == 0 || nc > 16) {
        std::fprintf(stderr, "[goodphotons] shared render needs 1..16 pinhole ('B') cameras\n");
        return out;
    }
    std::vector<float> v;
    std::vector<int32_t> m;
    gp_scene_desc d = detail::describe(scene, v, m);
    detail::SceneHandle sh;
    if (gp_scene_create(&d, &sh.s) != GP_OK) { detail::report("scene"); return out; }
    std::vector<detail::FilmHandle> fh(nc);
    std::vector<gp_film*> films(nc);
    std::vector<gp_camera_desc> cd(nc);
    for (size_t c = 0; c < nc; ++c) {
        if (gp_film_create(resX[c], resY[c], N > 0 ? N : 1, &fh[c].f) != GP_OK) { detail::report("film"); return out; }
        films[c] = fh[c].f;
        cd[c] = detail::camera(cams[c], resX[c], resY[c]);
    }
    gp_trace_params p{};
    p.n_photons = N;
    p.seed_base = seedBase;
    p.salt = kReferenceSalt;
    p.hero_c = heroC < 1 ? 1 : (heroC > 4 ? 4 : heroC);
    p.max_depth = gMaxDepth;
    p.rr_depth = gRrDepth;
    gp_energy e{};
    if (gp_trace_shared(sh.s, (int32_t)nc, cd.data(), films.data(), &p, &e, nullptr) != GP_OK) {
        detail::report("shared trace");
        return out;
    }
    for (size_t c = 0; c < nc; ++c)
        if (!detail::download(films[c], out[c])) { detail::report("download"); out[c].alloc(); }
    eOut.emitted += e.emitted;
    eOut.absorbed += e.absorbed;
    eOut.sensor += e.sensor;
    eOut.escaped += e.escaped;
    eOut.residual += e.residual;
    return out;
}

// Resident forward session (one camera): the scene and film stay in HBM; every
// batch traces N more photons into the running totals.
// Device state of a resident multi-camera session, independent of the host types:
// the scene, one film per camera (fixed-point, sized for maxPhotons photons in all)
// and the running ledger.
class SessionCore {
public:
    bool open(const gp_scene_desc& d, int32_t builder, const std::vector<gp_camera_desc>& cams,
              long long maxPhotons) {
        close();
        if (cams.empty() || cams.size() > 16) { std::fprintf(stderr, "[goodphotons] session: 1..16 cameras\n"); return false; }
        if (gp_scene_create_ex(&d, builder, &scene_) != GP_OK) { detail::report("session scene"); return false; }
        cams_ = cams;
        films_.assign(cams.size(), nullptr);
        for (size_t c = 0; c < cams.size(); ++c)
            if (gp_film_create(cams[c].res_x, cams[c].res_y, maxPhotons, &films_[c]) != GP_OK) {
                detail::report("session film");
                close();
                return false;
            }
        return true;
    }
    // Running totals of camera c (pixel-major doubles) from an earlier session or
    // checkpoint: seedPhotons photons traced with maxDepth.
    bool seed(size_t c, const double* xyz, const double* hits, long long seedPhotons, int maxDepth) {
        gp_trace_params p{};
        p.n_photons = seedPhotons;
        p.hero_c = 1;
        p.max_depth = maxDepth;
        if (gp_film_seed(films_[c], scene_, &cams_[c], &p, xyz, hits) != GP_OK) { detail::report("session seed"); return false; }
        return true;
    }
    void seedEnergy(const gp_energy& e) { energy_ = e; }
    bool batch(long long N, unsigned long long seedBase, uint64_t salt, int heroC, int maxDepth, int rrDepth) {
        gp_trace_params p{};
        p.n_photons = N;
        p.seed_base = seedBase;
        p.salt = salt;
        p.hero_c = heroC < 1 ? 1 : (heroC > 4 ? 4 : heroC);
        p.max_depth = maxDepth;
        p.rr_depth = rrDepth;
        if (gp_trace_shared(scene_, (int32_t)films_.size(), cams_.data(), films_.data(), &p, &energy_, nullptr) != GP_OK) {
            detail::report("session batch");
            return false;
        }
        return true;
    }
    bool hits0(double* hits) const { return gp_film_download_hits(films_[0], hits) == GP_OK; }
    bool download(size_t c, double* xyz, double* hits) const { return gp_film_download(films_[c], xyz, hits) == GP_OK; }
    const gp_energy& energy() const { return energy_; }
    size_t cameras() const { return films_.size(); }
    const gp_camera_desc& camera(size_t c) const { return cams_[c]; }
    double totalPower() const {
        gp_scene_info i{};
        return gp_scene_get_info(scene_, &i) == GP_OK ? i.total_power : 0.0;
    }
    bool isOpen() const { return scene_ != nullptr; }
    void close() {
        for (gp_film* f : films_)
            if (f) gp_film_destroy(f);
        films_.clear();
        if (scene_) gp_scene_destroy(scene_);
        scene_ = nullptr;
        energy_ = gp_energy{};
    }
    ~SessionCore() { close(); }

private:
    gp_scene* scene_ = nullptr;
    std::vector<gp_film*> films_;
    std::vector<gp_camera_desc> cams_;
    gp_energy energy_{};
};

// Resident shared session for the core's own types, as sharedForwardGpuBegin / Batch /
// Hits0 / Download / End. maxPhotons bounds the photons the session may hold in all
// (seed included): it fixes the films' fixed-point scale.
class SharedSession {
public:
    bool begin(const Scene& scene, const std::vector<Camera>& cams, const std::vector<int>& resX,
               const std::vector<int>& resY, char camMode, int heroC, long long maxPhotons,
               const std::vector<Film>* seedFilms = nullptr, const EnergyReport* seedEnergy = nullptr,
               long long seedPhotons = 0) {
        if (camMode != 'B' || cams.size() != resX.size() || cams.size() != resY.size()) return false;
        std::vector<float> v;
        std::vector<int32_t> m;
        const gp_scene_desc d = detail::describe(scene, v, m);
        std::vector<gp_camera_desc> cd;
        for (size_t c = 0; c < cams.size(); ++c) cd.push_back(detail::camera(cams[c], resX[c], resY[c]));
        if (!core_.open(d, GP_BVH_HOST_SAH, cd, maxPhotons)) return false;
        heroC_ = heroC;
        if (seedFilms) {
            if (seedFilms->size() != cams.size()) { core_.close(); return false; }
            for (size_t c = 0; c < cams.size(); ++c) {
                const Film& f = (*seedFilms)[c];
                std::vector<double> xyz((size_t)resX[c] * resY[c] * 3);
                for (size_t i = 0; i < f.xyz.size() && 3 * i + 2 < xyz.size(); ++i) {
                    xyz[3 * i] = f.xyz[i].x; xyz[3 * i + 1] = f.xyz[i].y; xyz[3 * i + 2] = f.xyz[i].z;
                }
                if (f.hits.size() != (size_t)resX[c] * resY[c] ||
                    !core_.seed(c, xyz.data(), f.hits.data(), seedPhotons, gMaxDepth)) {
                    core_.close();
                    return false;
                }
            }
        }
        if (seedEnergy)
            core_.seedEnergy(gp_energy{seedEnergy->emitted, seedEnergy->absorbed, seedEnergy->sensor,
                                       seedEnergy->escaped, seedEnergy->residual});
        return true;
    }
    bool batch(long long N, unsigned long long seedBase) {
        return core_.batch(N, seedBase, kReferenceSalt, heroC_, gMaxDepth, gRrDepth);
    }
    // Camera 0's running hit counts only (a quarter of its film's bytes).
    bool hits0(std::vector<double>& hits) {
        hits.resize((size_t)core_.camera(0).res_x * core_.camera(0).res_y);
        return core_.hits0(hits.data());
    }
    // Running totals, seed included (REPLACE the caller's accumulators).
    bool download(std::vector<Film>& films, EnergyReport& e) {
        films.resize(core_.cameras());
        for (size_t c = 0; c < core_.cameras(); ++c) {
            films[c].resX = core_.camera(c).res_x;
            films[c].resY = core_.camera(c).res_y;
            films[c].alloc();
            std::vector<double> xyz(films[c].hits.size() * 3);
            if (!core_.download(c, xyz.data(), films[c].hits.data())) return false;
            for (size_t i = 0; i < films[c].xyz.size(); ++i)
                films[c].xyz[i] = Vec3d{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        }
        const gp_energy& g = core_.energy();
        e = EnergyReport{g.emitted, g.absorbed, g.sensor, g.escaped, g.residual};
        return true;
    }
    void end() { core_.close(); }

private:
    SessionCore core_;
    int heroC_ = 4;
};

class ForwardSession {
public:
    bool begin(const Scene& scene, const Camera& cam, int resX, int resY, int heroC,
               long long maxPhotons = 0) {
        end();
        std::vector<float> v;
        std::vector<int32_t> m;
        gp_scene_desc d = detail::describe(scene, v, m);
        if (gp_scene_create(&d, &scene_) != GP_OK) { detail::report("scene"); return false; }
        if (gp_film_create(resX, resY, maxPhotons, &film_) != GP_OK) { detail::report("film"); end(); return false; }
        cam_ = detail::camera(cam, resX, resY);
        heroC_ = heroC;
        return true;
    }
    bool batch(long long N, unsigned long long seedBase) {
        gp_trace_params p{};
        p.n_photons = N;
        p.seed_base = seedBase;
        p.salt = kReferenceSalt;
        p.hero_c = heroC_;
        p.max_depth = gMaxDepth;
        p.rr_depth = gRrDepth;
        if (gp_trace(scene_, &cam_, film_, &p, &energy_, nullptr) != GP_OK) { detail::report("batch"); return false; }
        return true;
    }
    // Running totals (REPLACES the caller's accumulators).
    bool download(Film& film, EnergyReport& e) {
        film.resX = cam_.res_x;
        film.resY = cam_.res_y;
        film.alloc();
        if (!detail::download(film_, film)) return false;
        e = EnergyReport{energy_.emitted, energy_.absorbed, energy_.sensor, energy_.escaped, energy_.residual};
        return true;
    }
    // Resumable checkpoint of the session film (writeCheckpoint's .ftbuf, plus the
    // exact fixed-point words). `photons` = photons traced so far; guard as
    // checkpointGuard(scene, mode, res, resY).
    bool saveCheckpoint(const std::string& path, long long photons, uint64_t guard) const {
        if (gp_film_save_checkpoint(film_, path.c_str(), photons, &energy_, 'B', guard) != GP_OK) {
            detail::report("checkpoint");
            return false;
        }
        return true;
    }
    // Replace the session film and ledger with a checkpoint's; returns the photons it
    // holds (the seedBase of the next batch), or -1.
    long long resume(const std::string& path, uint64_t guard) {
        gp_film* f = nullptr;
        int64_t n = 0;
        gp_energy e{};
        if (gp_film_load_checkpoint(path.c_str(), 'B', guard, &f, &n, &e) != GP_OK) {
            detail::report("resume");
            return -1;
        }
        if (film_) gp_film_destroy(film_);
        film_ = f;
        energy_ = e;
        return n;
    }
    void end() {
        if (film_) gp_film_destroy(film_);
        if (scene_) gp_scene_destroy(scene_);
        film_ = nullptr;
        scene_ = nullptr;
        energy_ = gp_energy{};
    }
    ~ForwardSession() { end(); }

private:
    gp_scene* scene_ = nullptr;
    gp_film* film_ = nullptr;
    gp_camera_desc cam_{};
    gp_energy energy_{};
    int heroC_ = 4;
};

}  // namespace gp
