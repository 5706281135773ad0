// C ABI of the forward light-tracing core (include/goodphotons.h): scenes,
// films, gp_trace (one job through the wavefront engine), probes and tuning.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "gp_bvh.hpp"
#include "gp_bvh_device.hpp"
#include "gp_runtime.hpp"

namespace gp {
const char* lastError();

namespace {

// Host evaluation of a descriptor spectrum in double: the same closed forms the
// reference's Spectrum closures compute (src/spectrum.h:80-143), used for the
// emitter wavelength CDFs. A sampled spectrum interpolates its samples in `pool`
// linearly, clamped to the end samples (the reference GPU backend's specLookup,
// src/render_cuda.cu:1641-1649).
double specEvalHost(const gp_spectrum& s, double w, const double* pool) {
    if (s.band_hi > s.band_lo && !(w >= s.band_lo && w <= s.band_hi)) return 0.0;
    switch (s.kind) {
        case GP_SPEC_SAMPLED: {
            const double* tab = pool + s.table;
            const int n = (int)s.p[2];
            // (multiplied before the division: exact at the sample wavelengths)
            const double f = (w - s.p[0]) * (double)(n - 1) / (s.p[1] - s.p[0]);
            if (!(f > 0.0)) return tab[0];
            if (f >= (double)(n - 1)) return tab[n - 1];
            const int i = (int)f;
            const double fr = f - (double)i;
            return tab[i] * (1.0 - fr) + tab[i + 1] * fr;
        }
        case GP_SPEC_CONST: return s.p[0];
        case GP_SPEC_STEP: return (w < s.p[2]) ? s.p[0] : s.p[1];
        case GP_SPEC_GAUSS: {
            double t = (w - s.p[2]) / s.p[3];
            return s.p[0] + s.p[1] * std::exp(-0.5 * t * t);
        }
        case GP_SPEC_BLACKBODY: {
            const double h = 6.62607015e-34, c = 2.99792458e8, kb = 1.380649e-23;
            double l = w * 1e-9;
            double e = std::exp((h * c) / (l * kb * s.p[0])) - 1.0;
            double v = (2.0 * h * c * c) / (std::pow(l, 5.0) * e);
            return (s.p[1] == 1.0) ? v : s.p[1] * v;
        }
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

// 1-nm emission CDF over [360, 830) (EmissionSampler::build, src/spectrum.h:253-265).
double buildCdf(const gp_spectrum& spd, double* cdf, const double* pool) {
    double acc = 0.0;
    cdf[0] = 0.0;
    for (int i = 0; i < kLambdaBins; ++i) {
        double w = 360.0 + (i + 0.5) * 1.0;
        acc += std::max(0.0, specEvalHost(spd, w, pool)) * 1.0;
        cdf[i + 1] = acc;
    }
    if (acc > 0)
        for (int i = 0; i <= kLambdaBins; ++i) cdf[i] /= acc;
    return acc;
}

// `dPool`: the scene's device copy (float) of gp_scene_desc.spectrum_samples.
DevSpectrum toDev(const gp_spectrum& s, const float* dPool) {
    DevSpectrum d;
    std::memset(&d, 0, sizeof d);
    d.kind = s.kind;
    for (int i = 0; i < 6; ++i) d.p[i] = (float)s.p[i];
    if (s.kind == GP_SPEC_SAMPLED) {
        d.n = (int32_t)s.p[2];
        d.tab = dPool + s.table;
        d.p[0] = (float)s.p[0];
        d.p[1] = (float)((double)(d.n - 1) / (s.p[1] - s.p[0]));
        d.p[2] = 0.0f;
    }
    d.band_lo = (float)s.band_lo;
    d.band_hi = (float)s.band_hi;
    return d;
}

bool finite3(const double* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

gp_status validateSpectrum(const gp_spectrum& s, const char* what, const gp_scene_desc* d) {
    if (s.kind < GP_SPEC_CONST || s.kind > GP_SPEC_SAMPLED)
        return fail(GP_ERR_UNSUPPORTED, std::string(what) + ": unknown spectrum kind");
    for (double v : s.p)
        if (!std::isfinite(v)) return fail(GP_ERR_INVALID, std::string(what) + ": non-finite spectrum parameter");
    if (s.kind == GP_SPEC_GAUSS && !(s.p[3] != 0.0)) return fail(GP_ERR_INVALID, "gaussian spectrum with zero width");
    if (s.kind == GP_SPEC_SAMPLED) {
        const double n = s.p[2];
        if (!(n >= 2.0 && n <= 1e6 && n == std::floor(n)) || !(s.p[1] > s.p[0]))
            return fail(GP_ERR_INVALID, std::string(what) + ": sampled spectrum needs >= 2 samples over lo < hi");
        if (!d->spectrum_samples || s.table < 0 || (int64_t)s.table + (int64_t)n > d->n_spectrum_samples)
            return fail(GP_ERR_INVALID, std::string(what) + ": sampled spectrum outside spectrum_samples");
        for (int64_t k = 0; k < (int64_t)n; ++k)
            if (!std::isfinite(d->spectrum_samples[s.table + k]))
                return fail(GP_ERR_INVALID, std::string(what) + ": non-finite spectrum sample");
    }
    return GP_OK;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};

}  // namespace
}  // namespace gp

using namespace gp;

gp_scene::~gp_scene() {
    void* ptrs[] = {dNodes, dTris, dTrin, dMats, dEms, dCdfs, dEmCdf, dGuides, dMirrors, dSpec};
    for (void* p : ptrs)
        if (p) hipFree(p);
}

gp_film::~gp_film() {
    if (d) hipFree(d);
}

extern "C" {

int32_t gp_abi_version(void) { return GP_ABI_VERSION; }

const char* gp_status_string(gp_status s) {
    switch (s) {
        case GP_OK: return "GP_OK";
        case GP_ERR_INVALID: return "GP_ERR_INVALID";
        case GP_ERR_UNSUPPORTED: return "GP_ERR_UNSUPPORTED";
        case GP_ERR_NO_DEVICE: return "GP_ERR_NO_DEVICE";
        case GP_ERR_HIP: return "GP_ERR_HIP";
        case GP_ERR_OOM: return "GP_ERR_OOM";
        case GP_ERR_CAPACITY: return "GP_ERR_CAPACITY";
        case GP_ERR_REJECTED: return "GP_ERR_REJECTED";
        case GP_ERR_STATE: return "GP_ERR_STATE";
    }
    return "unknown";
}

const char* gp_last_error(void) { return gp::lastError(); }

gp_status gp_device_query(gp_device_info* out) {
    if (!out) return fail(GP_ERR_INVALID, "null output");
    std::memset(out, 0, sizeof *out);
    DeviceState& d = device();
    std::snprintf(out->name, sizeof out->name, "%s", d.name.c_str());
    std::snprintf(out->arch, sizeof out->arch, "%s", d.arch.c_str());
    out->compute_units = d.cus;
    out->wave_size = 64;
    out->lds_per_block = (int32_t)d.ldsPerBlock;
    if (!d.ok) return fail(GP_ERR_NO_DEVICE, d.why);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
        out->free_mem = (int64_t)fr;
        out->total_mem = (int64_t)tot;
    }
    return GP_OK;
}

// ------------------------------------------------------------------ scenes
static gp_status createScene(const gp_scene_desc* d, int builder, gp_scene** out) {
    if (!d || !out) return fail(GP_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->n_tris <= 0 || !d->verts || !d->mat_ids) return fail(GP_ERR_INVALID, "scene needs triangles");
    if (d->n_materials <= 0 || !d->materials) return fail(GP_ERR_INVALID, "scene needs materials");
    if (d->n_emitters <= 0 || !d->emitters) return fail(GP_ERR_INVALID, "scene needs at least one emitter");
    if (d->n_tris >= (int64_t(1) << 28)) return fail(GP_ERR_CAPACITY, "more than 2^28 triangles");
    for (int i = 0; i < d->n_materials; ++i) {
        const gp_material& m = d->materials[i];
        if (m.type == GP_MAT_GLOSSY && !(m.roughness > 0.0))
            return fail(GP_ERR_UNSUPPORTED, "Glossy with roughness <= 0 is a planar mirror");
        if (m.type != GP_MAT_DIFFUSE && m.type != GP_MAT_DIELECTRIC && m.type != GP_MAT_MIRROR &&
            m.type != GP_MAT_GLOSSY)
            return fail(GP_ERR_UNSUPPORTED, "material type outside {Diffuse, Dielectric, Mirror, Glossy}");
        gp_status st;
        if ((st = validateSpectrum(m.reflect, "reflect", d)) != GP_OK) return st;
        if ((st = validateSpectrum(m.ior, "ior", d)) != GP_OK) return st;
        if (!std::isfinite(m.roughness)) return fail(GP_ERR_INVALID, "non-finite roughness");
    }
    for (int64_t i = 0; i < d->n_tris; ++i) {
        if (d->mat_ids[i] < 0 || d->mat_ids[i] >= d->n_materials)
            return fail(GP_ERR_INVALID, "triangle material id out of range");
    }
    for (int64_t i = 0; i < 9 * d->n_tris; ++i)
        if (!std::isfinite(d->verts[i])) return fail(GP_ERR_INVALID, "non-finite vertex");
    for (int i = 0; i < d->n_emitters; ++i) {
        const gp_emitter& e = d->emitters[i];
        if (!finite3(e.origin) || !finite3(e.u) || !finite3(e.v) || !finite3(e.normal) || !(e.area > 0.0))
            return fail(GP_ERR_INVALID, "bad emitter geometry");
        gp_status st = validateSpectrum(e.spd, "emitter spd", d);
        if (st != GP_OK) return st;
    }
    if (builder == GP_BVH_PREBUILT && (!d->bvh_nodes || d->n_bvh_nodes <= 0 || !d->bvh_prim_index ||
                                       d->n_bvh_prim_index <= 0))
        return fail(GP_ERR_INVALID, "GP_BVH_PREBUILT needs bvh_nodes and bvh_prim_index");
    gp_status st;
    // the device builder needs the device now; the host trees are built (and a
    // caller's tree checked) first
    if (builder == GP_BVH_DEVICE_LBVH && (st = requireDevice()) != GP_OK) return st;

    std::unique_ptr<gp_scene> s(new gp_scene());
    s->nTris = d->n_tris;
    s->builder = builder;
    const size_t n = (size_t)d->n_tris;
    std::vector<float4> tris(3 * n), trin(n);
    auto fillTri = [&](size_t k, int prim) {
        const float* v = d->verts + 9 * (size_t)prim;
        const int mat = d->mat_ids[prim];
        tris[3 * k + 0] = make_float4(v[0], v[1], v[2], 0.f);
        tris[3 * k + 1] = make_float4(v[3], v[4], v[5], 0.f);
        tris[3 * k + 2] = make_float4(v[6], v[7], v[8], 0.f);
        std::memcpy(&tris[3 * k + 0].w, &prim, 4);
        std::memcpy(&tris[3 * k + 1].w, &mat, 4);
        double e1[3] = {(double)v[3] - v[0], (double)v[4] - v[1], (double)v[5] - v[2]};
        double e2[3] = {(double)v[6] - v[0], (double)v[7] - v[1], (double)v[8] - v[2]};
        double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        if (l > 0) { c[0] /= l; c[1] /= l; c[2] /= l; } else { c[0] = 0; c[1] = 0; c[2] = 1; }
        trin[k] = make_float4((float)c[0], (float)c[1], (float)c[2], 0.f);
        std::memcpy(&trin[k].w, &mat, 4);
    };
    {
        // ray-binning grid: kBinCellsPerAxis cells per axis over the centroid box
        float blo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (size_t k = 0; k < n; ++k) {
            const float* v = d->verts + 9 * k;
            for (int a = 0; a < 3; ++a) {
                const float c = 0.5f * (std::min({v[a], v[3 + a], v[6 + a]}) + std::max({v[a], v[3 + a], v[6 + a]}));
                blo[a] = std::min(blo[a], c);
                bhi[a] = std::max(bhi[a], c);
            }
        }
        for (int a = 0; a < 3; ++a) {
            const float ext = bhi[a] - blo[a];
            s->binLo[a] = blo[a];
            s->binScale[a] = (ext > 0.f && std::isfinite(ext)) ? (float)kBinCellsPerAxis / ext : 0.f;
        }
    }
    DeviceBvh dbvh;
    std::vector<float> quads;   // four-wide nodes of the host-built tree
    if (builder == GP_BVH_DEVICE_LBVH) {
        // input order; the device sorts, builds and gathers
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (size_t k = 0; k < n; ++k) {
            fillTri(k, (int)k);
            const float* v = d->verts + 9 * k;
            for (int a = 0; a < 3; ++a) {
                const float c = 0.5f * (std::min({v[a], v[3 + a], v[6 + a]}) + std::max({v[a], v[3 + a], v[6 + a]}));
                clo[a] = std::min(clo[a], c);
                chi[a] = std::max(chi[a], c);
            }
        }
        DevBuf dIn, dInN;
        GP_HIP(hipMalloc(&dIn.p, tris.size() * sizeof(float4)));
        GP_HIP(hipMalloc(&dInN.p, trin.size() * sizeof(float4)));
        GP_HIP(hipMemcpy(dIn.p, tris.data(), tris.size() * sizeof(float4), hipMemcpyHostToDevice));
        GP_HIP(hipMemcpy(dInN.p, trin.data(), trin.size() * sizeof(float4), hipMemcpyHostToDevice));
        if ((st = buildBvhDevice((const float4*)dIn.p, (const float4*)dInN.p, d->n_tris, clo, chi, dbvh)) != GP_OK)
            return st;
        s->nNodes = dbvh.nPairs;
        s->depth = dbvh.depth;
        s->stackNeed = dbvh.stackNeed;
        s->buildMs = dbvh.buildMs;
        s->dNodes = dbvh.nodes;
        s->dTris = dbvh.tris;
        s->dTrin = dbvh.trin;
        s->deviceBytes += (int64_t)(dbvh.nPairs * 128 + (int64_t)n * 64);
    } else {
        auto t0 = std::chrono::steady_clock::now();
        BvhResult bvh;
        try {
            if (builder == GP_BVH_PREBUILT)
                importBvh(d->verts, d->n_tris, d->bvh_nodes, d->n_bvh_nodes, d->bvh_prim_index,
                          d->n_bvh_prim_index, bvh);
            else
                buildBvh(d->verts, d->n_tris, 8, bvh);
        } catch (const std::exception& ex) {
            return fail(GP_ERR_INVALID, ex.what());
        }
        auto t1 = std::chrono::steady_clock::now();
        s->depth = bvh.depth;
        s->buildMs = std::chrono::duration<double, std::milli>(t1 - t0).count();
        // a triangle may be referenced from several leaves (one entry per reference)
        const size_t m = bvh.order.size();
        tris.resize(3 * m);
        trin.resize(m);
        s->leafTris.resize(9 * m);
        for (size_t k = 0; k < m; ++k) {
            const int prim = bvh.order[k];
            fillTri(k, prim);
            std::memcpy(&s->leafTris[9 * k], d->verts + 9 * (size_t)prim, 9 * sizeof(float));
        }
        s->pairs = std::move(bvh.pairs);
        int depth4 = 0;
        collapse4(s->pairs, quads, depth4);
        s->nNodes = (int64_t)(quads.size() / 32);
        s->stackNeed = 3 * depth4;
        if ((st = requireDevice()) != GP_OK) return st;
    }

    auto upload = [&](const void* src, size_t bytes, void** dst) -> gp_status {
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) {
            *dst = nullptr;
            return fail(GP_ERR_OOM, "scene upload: hipMalloc failed");
        }
        e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(GP_ERR_HIP, std::string("scene upload: ") + hipGetErrorString(e));
        s->deviceBytes += (int64_t)bytes;
        return GP_OK;
    };
    if (d->spectrum_samples && d->n_spectrum_samples > 0) {
        std::vector<float> pool((size_t)d->n_spectrum_samples);
        for (size_t k = 0; k < pool.size(); ++k) pool[k] = (float)d->spectrum_samples[k];
        if ((st = upload(pool.data(), pool.size() * sizeof(float), (void**)&s->dSpec)) != GP_OK) return st;
    }
    std::vector<DevMaterial> mats((size_t)d->n_materials);
    for (int i = 0; i < d->n_materials; ++i) {
        DevMaterial& m = mats[(size_t)i];
        std::memset(&m, 0, sizeof m);
        m.type = d->materials[i].type;
        m.roughness = (float)d->materials[i].roughness;
        m.reflect = toDev(d->materials[i].reflect, s->dSpec);
        m.ior = toDev(d->materials[i].ior, s->dSpec);
    }
    const int nEm = d->n_emitters;
    std::vector<double> cdfs((size_t)nEm * (kLambdaBins + 1));
    std::vector<DevEmitter> ems((size_t)nEm);
    std::vector<double> powers((size_t)nEm);
    double total = 0.0;
    std::vector<double> emCdf((size_t)nEm);
    for (int i = 0; i < nEm; ++i) {
        const gp_emitter& e = d->emitters[i];
        double integral = buildCdf(e.spd, &cdfs[(size_t)i * (kLambdaBins + 1)], d->spectrum_samples);
        powers[(size_t)i] = integral * e.area * 3.141592653589793;
        total += powers[(size_t)i];
        emCdf[(size_t)i] = total;
        s->emitQuads.push_back({e.origin[0], e.origin[1], e.origin[2], e.u[0], e.u[1], e.u[2],
                                e.v[0], e.v[1], e.v[2], e.normal[0], e.normal[1], e.normal[2]});
    }
    if (!(total > 0.0)) return fail(GP_ERR_INVALID, "emitters carry no power");
    for (auto& c : emCdf) c /= total;
    s->totalPower = total;
    s->nEm = nEm;
    s->basePower = (nEm == 1) ? powers[0] : total;

    if (builder != GP_BVH_DEVICE_LBVH) {
        if ((st = upload(quads.data(), quads.size() * sizeof(float), (void**)&s->dNodes)) != GP_OK) return st;
        if ((st = upload(tris.data(), tris.size() * sizeof(float4), (void**)&s->dTris)) != GP_OK) return st;
        if ((st = upload(trin.data(), trin.size() * sizeof(float4), (void**)&s->dTrin)) != GP_OK) return st;
    }
    if ((st = upload(mats.data(), mats.size() * sizeof(DevMaterial), (void**)&s->dMats)) != GP_OK) return st;
    if ((st = upload(cdfs.data(), cdfs.size() * sizeof(double), (void**)&s->dCdfs)) != GP_OK) return st;
    if ((st = upload(emCdf.data(), emCdf.size() * sizeof(double), (void**)&s->dEmCdf)) != GP_OK) return st;
    // guide[k] = last bin i in [0, kLambdaBins) with cdf[i] <= k / kCdfGuide
    std::vector<int32_t> guides((size_t)nEm * (kCdfGuide + 1));
    for (int i = 0; i < nEm; ++i) {
        const double* c = &cdfs[(size_t)i * (kLambdaBins + 1)];
        for (int k = 0; k <= kCdfGuide; ++k) {
            const double target = (double)k / (double)kCdfGuide;
            guides[(size_t)i * (kCdfGuide + 1) + (size_t)k] =
                (int32_t)(std::upper_bound(c, c + kLambdaBins, target) - c) - 1;
        }
    }
    if ((st = upload(guides.data(), guides.size() * sizeof(int32_t), (void**)&s->dGuides)) != GP_OK) return st;
    for (int i = 0; i < nEm; ++i) {
        const gp_emitter& e = d->emitters[i];
        DevEmitter& de = ems[(size_t)i];
        std::memset(&de, 0, sizeof de);
        for (int k = 0; k < 3; ++k) {
            de.origin[k] = (float)e.origin[k];
            de.u[k] = (float)e.u[k];
            de.v[k] = (float)e.v[k];
            de.normal[k] = (float)e.normal[k];
            de.beam[k] = (float)e.beam_dir[k];
        }
        de.collimated = e.collimated ? 1 : 0;
        de.power = (float)powers[(size_t)i];
        de.cdf = s->dCdfs + (size_t)i * (kLambdaBins + 1);
        de.guide = s->dGuides + (size_t)i * (kCdfGuide + 1);
    }
    if ((st = upload(ems.data(), ems.size() * sizeof(DevEmitter), (void**)&s->dEms)) != GP_OK) return st;
    std::vector<DevMirrorPlane> planes;
    buildMirrorPlanes(d, planes);
    if (!planes.empty()) {
        if ((st = upload(planes.data(), planes.size() * sizeof(DevMirrorPlane), (void**)&s->dMirrors)) != GP_OK) return st;
        s->nMirrors = (int)planes.size();
    }
    *out = s.release();
    return GP_OK;
}

gp_status gp_scene_create(const gp_scene_desc* d, gp_scene** out) {
    return createScene(d, GP_BVH_HOST_SAH, out);
}

gp_status gp_scene_create_ex(const gp_scene_desc* d, int32_t builder, gp_scene** out) {
    if (builder != GP_BVH_HOST_SAH && builder != GP_BVH_DEVICE_LBVH && builder != GP_BVH_PREBUILT)
        return fail(GP_ERR_INVALID, "unknown BVH builder");
    return createScene(d, builder, out);
}

gp_status gp_scene_get_info(const gp_scene* s, gp_scene_info* out) {
    if (!s || !out) return fail(GP_ERR_INVALID, "null argument");
    out->n_tris = s->nTris;
    out->n_nodes = s->nNodes;
    out->bvh_depth = s->depth;
    out->n_emitters = s->nEm;
    out->device_bytes = s->deviceBytes;
    out->build_ms = s->buildMs;
    out->total_power = s->totalPower;
    return GP_OK;
}

void gp_scene_destroy(gp_scene* s) { delete s; }

// ------------------------------------------------------------------ films
gp_status gp_film_create(int32_t rx, int32_t ry, int64_t maxPhotons, gp_film** out) {
    if (!out) return fail(GP_ERR_INVALID, "null output");
    *out = nullptr;
    if (rx <= 0 || ry <= 0 || (int64_t)rx * ry > (int64_t(1) << 28)) return fail(GP_ERR_INVALID, "bad film size");
    gp_status st = requireDevice();
    if (st != GP_OK) return st;
    std::unique_ptr<gp_film> f(new gp_film());
    f->rx = rx;
    f->ry = ry;
    f->maxPhotons = maxPhotons > 0 ? maxPhotons : (int64_t(1) << 32);
    const size_t bytes = (size_t)rx * ry * 4 * sizeof(unsigned long long);
    if (hipMalloc(&f->d, bytes) != hipSuccess) {
        f->d = nullptr;
        return fail(GP_ERR_OOM, "film allocation failed");
    }
    if (hipMemset(f->d, 0, bytes) != hipSuccess) return fail(GP_ERR_HIP, "film clear failed");
    *out = f.release();
    return GP_OK;
}

gp_status gp_film_reset(gp_film* f) {
    if (!f) return fail(GP_ERR_INVALID, "null film");
    GP_HIP(hipMemset(f->d, 0, (size_t)f->rx * f->ry * 4 * sizeof(unsigned long long)));
    f->scale = 0.0;
    f->qmax = 0;
    f->splatBudget = 0.0;
    f->splatUsed = 0.0;
    return GP_OK;
}

gp_status gp_film_download_raw(gp_film* f, int64_t* words, double* scale) {
    if (!f) return fail(GP_ERR_INVALID, "null film");
    const size_t n = (size_t)f->rx * f->ry * 4;
    if (words) GP_HIP(hipMemcpy(words, f->d, n * sizeof(int64_t), hipMemcpyDeviceToHost));
    if (scale) *scale = f->scale;
    return GP_OK;
}

gp_status gp_film_download(gp_film* f, double* xyz, double* hits) {
    if (!f) return fail(GP_ERR_INVALID, "null film");
    const size_t np = (size_t)f->rx * f->ry;
    std::vector<int64_t> w(np * 4);
    GP_HIP(hipMemcpy(w.data(), f->d, w.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    const double inv = f->scale > 0 ? 1.0 / f->scale : 0.0;
    for (size_t i = 0; i < np; ++i) {
        if (xyz) {
            xyz[3 * i + 0] = (double)w[4 * i + 0] * inv;
            xyz[3 * i + 1] = (double)w[4 * i + 1] * inv;
            xyz[3 * i + 2] = (double)w[4 * i + 2] * inv;
        }
        if (hits) hits[i] = (double)w[4 * i + 3];
    }
    return GP_OK;
}

gp_status gp_film_download_hits(gp_film* f, double* hits) {
    if (!f || !hits) return fail(GP_ERR_INVALID, "null argument");
    const size_t np = (size_t)f->rx * f->ry;
    std::vector<int64_t> h(np);
    // the fourth word of every pixel: a strided copy of 8 of each 32 bytes
    GP_HIP(hipMemcpy2D(h.data(), sizeof(int64_t), f->d + 3, 4 * sizeof(int64_t), sizeof(int64_t), np,
                       hipMemcpyDeviceToHost));
    for (size_t i = 0; i < np; ++i) hits[i] = (double)h[i];
    return GP_OK;
}

gp_status gp_film_seed(gp_film* f, gp_scene* s, const gp_camera_desc* cam, const gp_trace_params* p,
                       const double* xyz, const double* hits) {
    if (!f || !s || !cam || !p || !xyz || !hits) return fail(GP_ERR_INVALID, "null argument");
    if (f->scale > 0.0) return fail(GP_ERR_STATE, "film already holds a trace (seed a fresh film)");
    if (p->n_photons < 0 || p->max_depth < 1) return fail(GP_ERR_INVALID, "invalid trace parameters");
    if (cam->res_x != f->rx || cam->res_y != f->ry) return fail(GP_ERR_INVALID, "camera resolution differs from the film's");
    CamBasis cb;
    if (!cameraBasis(*cam, cb)) return fail(GP_ERR_INVALID, "degenerate camera");
    const size_t np = (size_t)f->rx * f->ry;
    // check everything before the film's scale is fixed
    const double lim = std::ldexp(1.0, 62);
    for (size_t i = 0; i < np; ++i) {
        if (!(hits[i] >= 0.0 && hits[i] < lim && hits[i] == std::floor(hits[i])))
            return fail(GP_ERR_INVALID, "seed hit counts must be whole numbers in [0, 2^62)");
        for (int c = 0; c < 3; ++c)
            if (!(xyz[3 * i + c] >= 0.0 && std::isfinite(xyz[3 * i + c])))
                return fail(GP_ERR_INVALID, "seed sums must be finite and non-negative");
    }
    const double sScale = f->scale, sBudget = f->splatBudget, sUsed = f->splatUsed;
    const unsigned long long sQmax = f->qmax;
    auto restore = [&] {
        f->scale = sScale; f->qmax = sQmax; f->splatBudget = sBudget; f->splatUsed = sUsed;
    };
    gp_status st = prepareFilm(f, s, cb, *p);
    if (st != GP_OK) { restore(); return st; }
    std::vector<int64_t> w(np * 4);
    for (size_t i = 0; i < np; ++i) {
        for (int c = 0; c < 3; ++c) {
            const double q = std::nearbyint(xyz[3 * i + c] * f->scale);
            if (!(q < lim)) { restore(); return fail(GP_ERR_INVALID, "seed sum beyond the film's range"); }
            w[4 * i + c] = (int64_t)q;
        }
        w[4 * i + 3] = (int64_t)hits[i];
    }
    hipError_t e = hipMemcpy(f->d, w.data(), w.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) { restore(); return fail(GP_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)); }
    return GP_OK;
}

void gp_film_destroy(gp_film* f) { delete f; }

// ------------------------------------------------------------------ trace
gp_status gp_trace_shared(gp_scene* s, int32_t n_cams, const gp_camera_desc* cams, gp_film* const* films,
                          const gp_trace_params* p, gp_energy* energy, gp_trace_stats* stats) {
    if (!s || !cams || !films || !p) return fail(GP_ERR_INVALID, "null argument");
    if (n_cams < 1 || n_cams > kMaxCams) return fail(GP_ERR_INVALID, "n_cams must be 1..16");
    if (p->n_photons < 0) return fail(GP_ERR_INVALID, "negative photon count");
    if (p->hero_c < 1 || p->hero_c > kMaxHero) return fail(GP_ERR_INVALID, "hero_c must be 1..4");
    if (p->max_depth < 1 || p->max_depth > 250) return fail(GP_ERR_INVALID, "max_depth must be 1..250");
    if (p->rr_depth < 0) return fail(GP_ERR_INVALID, "negative rr_depth");
    if (p->flags & ~GP_TRACE_TRAVERSAL_STATS) return fail(GP_ERR_INVALID, "unknown trace flags");
    gp_status st = requireDevice();
    if (st != GP_OK) return st;
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    if (E.owner)
        return fail(GP_ERR_STATE, "a server has jobs in flight on this device; step it until they finish first");
    if (!E.tunSet) { E.tun = defaultTuning(); E.tunSet = true; }
    ++E.epoch;
    const int lds = E.tun.lds_stack;
    if ((st = E.ensure((unsigned)E.tun.queue_capacity, lds, s->stackNeed, n_cams, s->nMirrors > 0, s->nTris)) != GP_OK)
        return st;
    E.mirrorActive = s->nMirrors > 0;
    // the statistics variants run for this call only (plain launches, no graph); set
    // after ensure(), which may rebuild the queue descriptor
    struct StatsScope {
        Engine& E;
        ~StatsScope() { E.q.stats = 0; }
    } statsScope{E};
    E.q.stats = (p->flags & GP_TRACE_TRAVERSAL_STATS) ? 1 : 0;
    hipStream_t strm = device().stream;

    gp_trace_stats acc{};
    gp_energy eacc{};
    E.resetProfile();
    auto t0 = std::chrono::steady_clock::now();
    const int64_t chunkMax = int64_t(1) << 31;
    int64_t done = 0;
    while (done < p->n_photons || (p->n_photons == 0 && done == 0)) {
        const int64_t n = std::min(chunkMax, p->n_photons - done);
        if (n <= 0) break;
        gp_trace_params cp = *p;
        cp.n_photons = n;
        cp.seed_base = p->seed_base + (uint64_t)done;
        DevJob J;
        if ((st = fillJob(s, cams, films, n_cams, cp, 0, J, E.hJobCams)) != GP_OK) return st;
        E.hJobs[0] = J;
        Ctrl& C = *E.hCtrl;
        E.initCtrl();
        C.nJobs = 1;
        C.nOrder = 1;
        C.order[0] = 0;
        C.nextPhoton[0] = 0;
        C.endPhoton[0] = (unsigned long long)n;
        C.active[0] = 1;
        C.lastEmitIter[0] = -1;
        GP_HIP(hipMemcpyAsync(E.dJobs, E.hJobs, sizeof(DevJob), hipMemcpyHostToDevice, strm));
        GP_HIP(hipMemcpyAsync(E.dJobCams, E.hJobCams, sizeof(DevCamera) * (size_t)n_cams, hipMemcpyHostToDevice, strm));
        GP_HIP(hipMemcpyAsync(E.dCtrl, E.hCtrl, sizeof(Ctrl), hipMemcpyHostToDevice, strm));
        if ((st = E.resetJobAcc(0)) != GP_OK) return st;
        E.parity = 0;
        const int K = chunkIterations(E.tun);
        // Every iteration retires or advances each live path by one segment, so the
        // loop needs at most ceil(n / capacity) refills times maxDepth iterations.
        const int64_t maxIters = ((n + E.capacity - 1) / E.capacity + 2) * (int64_t)(p->max_depth + 2) * 2 + 64;
        int64_t iters = 0;
        while (true) {
            if ((st = E.runIterations(K)) != GP_OK) return st;
            iters += K;
            if ((st = E.readCtrl()) != GP_OK) return st;
            unsigned live = 0;
            for (int g = 0; g < kSegs; ++g) live += C.pathCount[E.parity][g][0];
            if (C.nextPhoton[0] >= (unsigned long long)n && live == 0) break;
            if (iters > maxIters) return fail(GP_ERR_STATE, "wavefront did not drain");
        }
        unsigned long long ev[5], cv[kAccCounters];
        if ((st = E.readJobAcc(0, ev, cv)) != GP_OK) return st;
        eacc.emitted += (double)ev[0] / J.energyScale;
        eacc.absorbed += (double)ev[1] / J.energyScale;
        eacc.sensor += (double)ev[2] / J.energyScale;
        eacc.escaped += (double)ev[3] / J.energyScale;
        eacc.residual += (double)ev[4] / J.energyScale;
        addCounters(cv, acc);
        acc.iterations += C.iter;
        done += n;
    }
    auto t1 = std::chrono::steady_clock::now();
    acc.wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    for (int k = 0; k < kKernelSlots; ++k) {
        acc.kernel_ms[k] = E.kernelMs[k];
        acc.kernel_launches[k] = E.kernelLaunches[k];
    }
    if (energy) {
        energy->emitted += eacc.emitted;
        energy->absorbed += eacc.absorbed;
        energy->sensor += eacc.sensor;
        energy->escaped += eacc.escaped;
        energy->residual += eacc.residual;
    }
    if (stats) *stats = acc;
    return GP_OK;
}

gp_status gp_trace(gp_scene* s, const gp_camera_desc* cam, gp_film* film, const gp_trace_params* p,
                   gp_energy* energy, gp_trace_stats* stats) {
    if (!cam || !film) return fail(GP_ERR_INVALID, "null argument");
    return gp_trace_shared(s, 1, cam, &film, p, energy, stats);
}

// ------------------------------------------------------------------ probes
static gp_status probeCommon(gp_scene* s, int64_t n, int& blocks, DevBuf& spill, int& spillDepth, int lds) {
    if (!s || n < 0) return fail(GP_ERR_INVALID, "bad probe arguments");
    gp_status st = requireDevice();
    if (st != GP_OK) return st;
    blocks = std::max(1, device().cus * 8);
    spillDepth = std::max(0, s->stackNeed - lds);
    if (spillDepth > 0) {
        hipError_t e = hipMalloc(&spill.p, (size_t)blocks * 256 * spillDepth * sizeof(int2));
        if (e != hipSuccess) return fail(GP_ERR_OOM, "probe spill allocation failed");
    }
    return GP_OK;
}

gp_status gp_probe_closest_hit(gp_scene* s, const float* rays, int64_t n, float tmin, int32_t* prim,
                               float* t) {
    if (!rays || !prim || !t) return fail(GP_ERR_INVALID, "null argument");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    if (!E.tunSet) { E.tun = defaultTuning(); E.tunSet = true; }
    const int lds = E.tun.lds_stack;
    int blocks = 0, spillDepth = 0;
    DevBuf spill, dr, dp, dt;
    gp_status st = probeCommon(s, n, blocks, spill, spillDepth, lds);
    if (st != GP_OK) return st;
    if (n == 0) return GP_OK;
    GP_HIP(hipMalloc(&dr.p, (size_t)n * 6 * sizeof(float)));
    GP_HIP(hipMalloc(&dp.p, (size_t)n * sizeof(int)));
    GP_HIP(hipMalloc(&dt.p, (size_t)n * sizeof(float)));
    hipStream_t strm = device().stream;
    GP_HIP(hipMemcpyAsync(dr.p, rays, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice, strm));
    launchProbeClosest(s->dNodes, s->dTris, (const float*)dr.p, n, tmin, (int*)dp.p, (float*)dt.p,
                       (int2*)spill.p, spillDepth, lds, blocks, 256, strm);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(prim, dp.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, strm));
    GP_HIP(hipMemcpyAsync(t, dt.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, strm));
    GP_HIP(hipStreamSynchronize(strm));
    return GP_OK;
}

gp_status gp_probe_any_hit(gp_scene* s, const float* rays, const float* tmax, int64_t n, float tmin,
                           uint8_t* occ) {
    if (!rays || !tmax || !occ) return fail(GP_ERR_INVALID, "null argument");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    if (!E.tunSet) { E.tun = defaultTuning(); E.tunSet = true; }
    const int lds = E.tun.lds_stack;
    int blocks = 0, spillDepth = 0;
    DevBuf spill, dr, dm, dpo;
    gp_status st = probeCommon(s, n, blocks, spill, spillDepth, lds);
    if (st != GP_OK) return st;
    if (n == 0) return GP_OK;
    GP_HIP(hipMalloc(&dr.p, (size_t)n * 6 * sizeof(float)));
    GP_HIP(hipMalloc(&dm.p, (size_t)n * sizeof(float)));
    GP_HIP(hipMalloc(&dpo.p, (size_t)n));
    hipStream_t strm = device().stream;
    GP_HIP(hipMemcpyAsync(dr.p, rays, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice, strm));
    GP_HIP(hipMemcpyAsync(dm.p, tmax, (size_t)n * sizeof(float), hipMemcpyHostToDevice, strm));
    launchProbeAny(s->dNodes, s->dTris, (const float*)dr.p, (const float*)dm.p, n, tmin,
                   (unsigned char*)dpo.p, (int2*)spill.p, spillDepth, lds, blocks, 256, strm);
    GP_HIP(hipGetLastError());
    GP_HIP(hipMemcpyAsync(occ, dpo.p, (size_t)n, hipMemcpyDeviceToHost, strm));
    GP_HIP(hipStreamSynchronize(strm));
    return GP_OK;
}

// ------------------------------------------------------------------ tuning
gp_status gp_get_tuning(gp_tuning* out) {
    if (!out) return fail(GP_ERR_INVALID, "null output");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    if (!E.tunSet) { E.tun = defaultTuning(); E.tunSet = true; }
    *out = E.tun;
    return GP_OK;
}

gp_status gp_set_tuning(const gp_tuning* in) {
    if (!in) return fail(GP_ERR_INVALID, "null input");
    const int q = in->queue_capacity;
    if (q < 4096 || q > (1 << 24) || (q & (q - 1)) != 0) return fail(GP_ERR_INVALID, "queue_capacity must be a power of two in [2^12, 2^24]");
    if (in->grid_waves_per_cu < 4 || in->grid_waves_per_cu > 64) return fail(GP_ERR_INVALID, "grid_waves_per_cu must be 4..64");
    if (in->iters_per_graph < 1 || in->iters_per_graph > 64 || (in->use_graph && in->iters_per_graph < 2))
        return fail(GP_ERR_INVALID, "iters_per_graph must be 2..64 (1..64 without graphs)");
    if (in->lds_stack != 8 && in->lds_stack != 12 && in->lds_stack != 16 && in->lds_stack != 24 &&
        in->lds_stack != 32)
        return fail(GP_ERR_INVALID, "lds_stack must be 8, 12, 16, 24 or 32");
    if ((in->block_extend != 0 && in->block_extend != 256) || (in->block_shade != 0 && in->block_shade != 256))
        return fail(GP_ERR_INVALID, "block_extend and block_shade are fixed at 256 (0 or 256)");
    if (in->ray_sort < 0 || in->ray_sort > 2) return fail(GP_ERR_INVALID, "ray_sort must be 0, 1 or 2");
    Engine& E = engine();
    std::lock_guard<std::mutex> lock(E.mu);
    if (E.owner) return fail(GP_ERR_STATE, "a server has jobs in flight on this device");
    E.tun = *in;
    E.tun.profile = in->profile ? 1 : 0;
    E.tun.block_extend = 256;
    E.tun.block_shade = 256;
    E.tunSet = true;
    if (E.graph) { hipGraphExecDestroy(E.graph); E.graph = nullptr; }
    return GP_OK;
}

}  // extern "C"
