/*
 * goodphotons.h - C ABI of the MI355X-native forward light-tracing core.
 *
 * Plain C: fixed-width integers, doubles, pointers and sizes. No HIP, CUDA or
 * torch types appear here, so the renderer's host (C++20, or Python through
 * ctypes) can bind it without a device compiler.
 *
 * What each entry point replaces in the reference (inhahe/GoodPhotons, ftrace
 * v0.190.4, paths relative to its repository root):
 *
 *   gp_device_query        <- cudaAvailable / cudaDeviceName / cudaMemInfo
 *                             (src/render_cuda.h:24, :27, :32)
 *   gp_scene_create        <- buildUpload / buildUploadScene + cudaForwardSupported
 *                             (src/render_cuda.cu:13600, :12692-13548, :12504-12631)
 *   gp_scene_create_ex     <- the same, with the BVH built on the device instead of
 *                             Bvh::build (src/bvh.h:89-176) on the host
 *   gp_scene_destroy       <- freeUpload (src/render_cuda.cu:12683)
 *   gp_mirror_planes       <- Scene::buildMirrorPlanes / Scene::mirrorPlanes
 *                             (src/scene.h:1606-1724), which the device upload
 *                             copies verbatim (src/render_cuda.cu:12735-12745)
 *   gp_film_*              <- the device film inside renderForwardCuda
 *                             (src/render_cuda.cu:13722-13746) and the resident
 *                             films of SharedGpuSession (src/render_cuda.h:121-143)
 *   gp_trace               <- renderForwardCuda (src/render_cuda.h:79-82,
 *                             src/render_cuda.cu:13712-13756) and
 *                             sharedForwardGpuBatch (src/render_cuda.h:136-137)
 *   gp_trace_shared        <- renderForwardSharedCuda / sharedForwardGpuBatch with
 *                             several cameras (src/render_cuda.h:121-143)
 *   gp_probe_closest_hit   <- Scene::closestHit / Bvh::traverseClosest
 *                             (src/scene.h:2194, src/bvh.h:180-224) exposed as a
 *                             batch probe in the style of -checklattice
 *                             (src/render_cuda.h:446-464)
 *   gp_probe_any_hit       <- Scene::occluded / Bvh::traverseAny
 *                             (src/scene.h:2249, src/bvh.h:228-249)
 *   gp_film_save/load_checkpoint, gp_checkpoint_guard
 *                          <- writeCheckpoint / readCheckpoint / checkpointGuard
 *                             (src/main.cpp:12317-12329, :12443-12483)
 *   gp_film_to_linear      <- filmToLinear (src/main.cpp:10282-10294) feeding
 *                             writePfm (src/main.cpp:10372-10387)
 *   gp_film_tonemap        <- filmToRgb8 (src/main.cpp:10296-10348): p99
 *                             auto-exposure and 8-bit sRGB, computed on the device
 *   gp_server_*            <- no single counterpart: the resident session
 *                             (sharedForwardGpuBegin/Batch/Download/End,
 *                             src/render_cuda.h:121-143) generalised to many
 *                             scenes, films and jobs in flight at once.
 *
 * Errors: every entry point returns gp_status. GP_OK is 0. On failure
 * gp_last_error() returns a message for the calling thread. Nothing here
 * calls exit(); the reference aborts the process instead
 * (src/render_cuda.cu:99-121).
 */
#ifndef GOODPHOTONS_H
#define GOODPHOTONS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_ABI_VERSION 1

typedef enum gp_status {
    GP_OK = 0,
    GP_ERR_INVALID = 1,      /* bad argument (null pointer, size out of range) */
    GP_ERR_UNSUPPORTED = 2,  /* scene feature outside the supported set */
    GP_ERR_NO_DEVICE = 3,    /* no usable gfx950 device */
    GP_ERR_HIP = 4,          /* a HIP runtime call failed */
    GP_ERR_OOM = 5,          /* device memory exhausted */
    GP_ERR_CAPACITY = 6,     /* a fixed limit (stack depth, jobs, scratch) exceeded */
    GP_ERR_REJECTED = 7,     /* admission control refused a job */
    GP_ERR_STATE = 8         /* call not valid in the object's current state */
} gp_status;

/* ---------------------------------------------------------------------------
 * Spectra. Analytic forms only (no sampled tables), so a step edge such as the
 * red wall's 600 nm switch is exact on both host and device.
 * Wavelengths are in nm. A band [band_lo, band_hi] multiplies the curve by an
 * indicator when band_hi > band_lo (used to restrict an emitter's SPD).
 * ------------------------------------------------------------------------- */
enum {
    GP_SPEC_CONST = 0,      /* v = p[0]                                            */
    GP_SPEC_STEP = 1,       /* v = (lam < p[2]) ? p[0] : p[1]                      */
    GP_SPEC_GAUSS = 2,      /* v = p[0] + p[1]*exp(-0.5*((lam-p[2])/p[3])^2)       */
    GP_SPEC_BLACKBODY = 3,  /* v = p[1] * Planck(T = p[0] K, lam)  (SI, per metre) */
    GP_SPEC_SELLMEIER = 4,  /* n = sqrt(max(1, 1 + sum Bi l^2/(l^2-Ci))), l in um,
                               B = p[0..2], C = p[3..5] (um^2)                      */
    GP_SPEC_CAUCHY = 5,     /* n = p[0] + p[1]/l^2, l in um                         */
    GP_SPEC_SAMPLED = 6     /* n = (int)p[2] >= 2 samples at lam_k = p[0] + k (p[1]-p[0])/(n-1),
                               gp_scene_desc.spectrum_samples[table .. table + n); linear
                               interpolation between samples, the end samples outside
                               [p[0], p[1]] (the reference GPU backend's baked tables,
                               bakeSpec / specLookup, src/render_cuda.cu:12658-12663,
                               :1641-1649). Any Spectrum function of the reference's scene
                               reaches the core this way.                          */
};

typedef struct gp_spectrum {
    int32_t kind;
    int32_t table;           /* GP_SPEC_SAMPLED: first sample in spectrum_samples, else 0 */
    double p[6];
    double band_lo, band_hi;
} gp_spectrum;

/* Material types use the reference's MatType numbering (src/scene.h:42). */
enum {
    GP_MAT_DIFFUSE = 0,
    GP_MAT_DIELECTRIC = 1,
    GP_MAT_MIRROR = 2,
    GP_MAT_GLOSSY = 4
};

/* Mirror is a perfect (delta) reflector. A forward light tracer cannot reach the
 * camera from a delta lobe, so, as the reference does, every diffuse or emitter
 * vertex is also connected to the camera THROUGH each flat mirror surface: the
 * coplanar Mirror triangles of a plane (at most 64 planes, gp_mirror_planes). One
 * specular vertex per connection (a mirror seen in a mirror stays dark), pinhole
 * camera only. Glossy with roughness <= 0 and the reference's HalfMirror are not
 * supported. */
typedef struct gp_material {
    int32_t type;
    int32_t reserved;
    double roughness;        /* Glossy lobe width; on a Dielectric it frosts both lobes */
    gp_spectrum reflect;     /* diffuse albedo, or mirror/glossy tint                    */
    gp_spectrum ior;         /* Dielectric index of refraction                           */
} gp_material;

/* Quad area emitter (Scene::addAreaLight, src/scene.h:1306). The point is
 * origin + u*s + v*t for s,t in [0,1); emission is cosine-weighted about
 * `normal` unless `collimated`, in which case every photon leaves along
 * beam_dir. Power = integral(SPD) * area * pi over 1-nm bins in [360, 830). */
typedef struct gp_emitter {
    double origin[3], u[3], v[3], normal[3];
    double area;
    int32_t collimated;
    int32_t reserved;
    double beam_dir[3];
    gp_spectrum spd;
} gp_emitter;

/* Triangle-only scene. Global primitive id = index into the triangle list,
 * exactly as the reference numbers prims (tris first, src/scene.h:2005-2044). */
/* A node of a caller's binary BVH, as the reference's BvhNode (src/bvh.h:70-75). */
typedef struct gp_bvh_node {
    double lo[3], hi[3];         /* box                                                */
    int32_t left, right;         /* inner node (count == 0): children (node indices)   */
    int32_t first, count;        /* leaf (count 1..4): bvh_prim_index[first .. +count) */
} gp_bvh_node;

typedef struct gp_scene_desc {
    int64_t n_tris;
    int32_t n_materials;
    int32_t n_emitters;
    const float* verts;          /* n_tris * 9 floats: v0.xyz v1.xyz v2.xyz */
    const int32_t* mat_ids;      /* n_tris                                    */
    const gp_material* materials;
    const gp_emitter* emitters;
    /* Samples of the GP_SPEC_SAMPLED spectra (may be null when there are none). */
    const double* spectrum_samples;
    int64_t n_spectrum_samples;
    /* The caller's own binary BVH over the triangles (the reference's Bvh::nodes and
     * primIdx, src/bvh.h:82-84; node 0 the root), read by gp_scene_create_ex with
     * GP_BVH_PREBUILT and ignored otherwise (may be null). */
    const gp_bvh_node* bvh_nodes;
    int64_t n_bvh_nodes;
    const int32_t* bvh_prim_index;
    int64_t n_bvh_prim_index;
} gp_scene_desc;

/* Pinhole camera as Camera::lookAt builds it (src/camera.h:110-129). */
typedef struct gp_camera_desc {
    double eye[3], target[3], up[3];
    double fov_y_deg;
    int32_t res_x, res_y;
} gp_camera_desc;

/* Energy ledger, same fields as EnergyReport (src/render.h:29-31). */
typedef struct gp_energy {
    double emitted, absorbed, sensor, escaped, residual;
} gp_energy;

typedef struct gp_trace_params {
    int64_t n_photons;       /* photons to trace in this call                        */
    uint64_t seed_base;      /* absolute index of the first photon (resume offset)   */
    uint64_t salt;           /* seeding salt; 0x9E3779B97F4A7C15 is the reference's  */
    int32_t hero_c;          /* wavelengths per path, 1..4                           */
    int32_t max_depth;       /* closest-hit segments per path (Renderer::maxBounce)  */
    int32_t rr_depth;        /* first bounce with Russian roulette; 0 = every vertex */
    int32_t flags;           /* 0, or GP_TRACE_TRAVERSAL_STATS                        */
} gp_trace_params;

/* gp_trace_params.flags: count the traversal statistics of the trace (the reference's
 * TraversalStats / -bvhstats, src/bvh.h:79, src/main.cpp:2599) into
 * gp_trace_stats.node_visits ... shadow_tri_tests. Kernel variants with counters run
 * instead of the default ones, as plain launches: slower, for measurement. */
#define GP_TRACE_TRAVERSAL_STATS 1

typedef struct gp_trace_stats {
    uint64_t photons;        /* photons emitted                                      */
    uint64_t segments;       /* closest-hit queries (one per bounce-loop iteration)  */
    uint64_t connections;    /* connection ray queries: camera shadow rays, plus the
                                light-side any hit and eye-side closest hit of each
                                connection through a flat mirror                     */
    uint64_t splats;         /* connections that reached the film                    */
    uint64_t iterations;     /* wavefront iterations run                             */
    double wall_ms;          /* host wall time of the trace loop                     */
    /* With gp_tuning.profile = 1: summed device time and launch count per kernel,
     * index 0 schedule, 1 emit, 2 extend (closest hit), 3 shade, 4 connect (and the
     * flat-mirror connections), 5 ray binning (its four kernels, gp_tuning.ray_sort). */
    double kernel_ms[6];
    uint64_t kernel_launches[6];
    /* With GP_TRACE_TRAVERSAL_STATS (else 0), over the four-wide tree: nodes loaded and
     * slab-tested (a leaf is a link of its parent, not a node) and triangles
     * intersected by the segments' closest hits, and the same for the camera shadow
     * rays' any hits (up to the first occluder). Flat-mirror connection legs are not
     * counted. The closest-hit counts depend slightly on which rays share a wave (a
     * lane keeps visiting nodes while its wave-mates look for a leaf); the any-hit
     * counts are a function of the rays alone. */
    uint64_t node_visits;
    uint64_t tri_tests;
    uint64_t shadow_node_visits;
    uint64_t shadow_tri_tests;
    uint64_t direct_connections; /* of `connections`: those of the emission points (the
                                    light seen directly), queued by the emission kernel */
} gp_trace_stats;

typedef struct gp_device_info {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t wave_size;
    int64_t total_mem;
    int64_t free_mem;
    int32_t lds_per_block;
    int32_t reserved;
} gp_device_info;

typedef struct gp_scene_info {
    int64_t n_tris;
    int64_t n_nodes;         /* four-wide nodes the traversal reads                  */
    int32_t bvh_depth;       /* longest root-to-leaf path of the binary tree built   */
    int32_t n_emitters;
    int64_t device_bytes;    /* BVH + triangles + tables resident in HBM             */
    double build_ms;         /* host BVH build time                                  */
    double total_power;
} gp_scene_info;

typedef struct gp_scene gp_scene;
typedef struct gp_film gp_film;
typedef struct gp_server gp_server;

/* ---- library ------------------------------------------------------------ */
int32_t gp_abi_version(void);
const char* gp_status_string(gp_status s);
const char* gp_last_error(void);
gp_status gp_device_query(gp_device_info* out);

/* ---- scenes ------------------------------------------------------------- */
/* Validates the descriptor, builds the BVH on the host and uploads BVH,
 * triangles, materials and emitter tables. Fails with GP_ERR_UNSUPPORTED for
 * features outside the set above (the reference's cudaForwardSupported gate). */
gp_status gp_scene_create(const gp_scene_desc* desc, gp_scene** out);
/* BVH builders for gp_scene_create_ex. GP_BVH_HOST_SAH is gp_scene_create's: binned
 * SAH on the host (3 axes x 32 bins), the higher-quality tree. GP_BVH_DEVICE_LBVH
 * builds a linear BVH on the device (Morton order, Karras 2012) in milliseconds,
 * for large or per-frame geometry; closest-hit results are the same either way.
 * GP_BVH_PREBUILT takes the caller's tree (desc->bvh_nodes, bvh_prim_index), as the
 * reference hands its Scene::bvh to the device verbatim (src/render_cuda.cu:12764-12772):
 * no build at all. It is checked as it is read (every node reached once from node 0,
 * child boxes inside their parent's, every leaf's triangles inside its box, leaves of
 * at most 4 triangles, every triangle in some leaf) and GP_ERR_INVALID names the first
 * failed check. */
#define GP_BVH_HOST_SAH 0
#define GP_BVH_DEVICE_LBVH 1
#define GP_BVH_PREBUILT 2
gp_status gp_scene_create_ex(const gp_scene_desc* desc, int32_t bvh_builder, gp_scene** out);
gp_status gp_scene_get_info(const gp_scene* scene, gp_scene_info* out);
/* Flat mirror planes the camera connection goes through, as gp_scene_create
 * derives them from the descriptor (host only, no device needed): Mirror
 * triangles grouped by plane, 10 doubles per plane written to out10 (unit normal
 * n.xyz with its largest component positive, offset d with dot(n, x) = d, padded
 * box lo.xyz, hi.xyz) for up to `cap` planes. Returns the number of planes (0..64),
 * or -GP_ERR_INVALID. */
int32_t gp_mirror_planes(const gp_scene_desc* desc, double* out10, int32_t cap);
/* A scene must outlive every trace and server job that uses it (a server job uses
 * its scene until the job has finished). */
void gp_scene_destroy(gp_scene* scene);

/* ---- films -------------------------------------------------------------- */
/* A film is res_x*res_y pixels of int64 fixed-point (X, Y, Z, hits). The
 * fixed-point scale is fixed by the first gp_trace into the film; it is
 * chosen so that max_photons photons cannot overflow any pixel (pass 0 for
 * the default of 2^32). Integer accumulation makes the film bit-identical
 * run to run and independent of scheduling. */
gp_status gp_film_create(int32_t res_x, int32_t res_y, int64_t max_photons, gp_film** out);
gp_status gp_film_reset(gp_film* film);
/* xyz: res_x*res_y*3 doubles (un-normalised sums, pixel-major, row py*res_x+px);
 * hits: res_x*res_y doubles. Either may be null. */
gp_status gp_film_download(gp_film* film, double* xyz, double* hits);
/* Raw fixed-point words, res_x*res_y*4 int64 (X, Y, Z, hits); scale_out gets the
 * factor that maps a double contribution to one integer unit. */
gp_status gp_film_download_raw(gp_film* film, int64_t* words, double* scale_out);
/* The hit counts alone, res_x*res_y doubles: a quarter of the film's bytes cross
 * PCIe. Replaces sharedForwardGpuHits0 (src/render_cuda.h:135-137), the progressive
 * loop's per-batch noise poll of camera 0. */
gp_status gp_film_download_hits(gp_film* film, double* hits);
/* Seeds a fresh film (nothing traced into it yet) with running totals held as
 * doubles, such as a checkpoint the reference wrote: replaces the seedFilms of
 * sharedForwardGpuBegin (src/render_cuda.h:117-130, src/render_cuda.cu:13806-13825).
 * The film's fixed-point scale is fixed as a trace of params->n_photons photons
 * (the photons the seed holds, counted in the film's budget) of `scene` seen by
 * `cam` would fix it; each X, Y, Z sum is stored rounded to the nearest integer unit
 * (an error of at most half a unit per pixel and channel, where the largest single
 * splat is 2^K units, K >= 8) and each hit count exactly. Sums that are negative,
 * not finite or beyond the film's 2^62 range, and hit counts that are not whole
 * numbers, are refused with GP_ERR_INVALID; a film already traced into with
 * GP_ERR_STATE. Sums downloaded from a film of the same scene, camera and
 * max_photons come back to the same words, so tracing on continues bit-exactly. */
gp_status gp_film_seed(gp_film* film, gp_scene* scene, const gp_camera_desc* cam,
                       const gp_trace_params* params, const double* xyz, const double* hits);
void gp_film_destroy(gp_film* film);

/* ---- checkpoints (.ftbuf) ------------------------------------------------ */
/* Identity guard of a checkpoint, as checkpointGuard (src/main.cpp:12317-12329):
 * the scene's triangle, sphere and emitter counts, the bits of its total emitted
 * power, the render mode character and the resolution. */
uint64_t gp_checkpoint_guard(int64_t n_tris, int64_t n_spheres, int64_t n_emitters, double total_power,
                             int32_t mode, int32_t res_x, int32_t res_y);
/* Write `film` as an FTBUF01 checkpoint (src/main.cpp:12443-12462) that the
 * reference's readCheckpoint accepts (n_photons = cumulative photons in the film,
 * energy = cumulative ledger, may be null), followed by a tagged section with the
 * exact fixed-point words, so gp_film_load_checkpoint resumes bit-exactly. Replaces
 * writeCheckpoint for a film resident on the device. */
gp_status gp_film_save_checkpoint(gp_film* film, const char* path, int64_t n_photons,
                                  const gp_energy* energy, int32_t mode, uint64_t guard);
/* Recreate a device film from a checkpoint written by gp_film_save_checkpoint.
 * GP_ERR_STATE if mode or guard differ from the file's (readCheckpoint refuses the
 * same files); GP_ERR_UNSUPPORTED for a file without the fixed-point section (one
 * the reference wrote: read it with readCheckpoint and seed a film with
 * gp_film_seed).
 * n_photons_out: the cumulative photon count, i.e. the seed_base of the next trace. */
gp_status gp_film_load_checkpoint(const char* path, int32_t mode, uint64_t guard, gp_film** film_out,
                                  int64_t* n_photons_out, gp_energy* energy_out);

/* ---- film -> image ------------------------------------------------------- */
/* Scene-linear sRGB radiance of a film that holds n_photons photons:
 * rgb = xyzToLinearSrgb(xyz / (n_photons * integral(ybar))) per pixel, in double,
 * stored as float32 triples in the film's row order (row 0 = bottom, the PFM
 * raster order). rgb holds res_x*res_y*3 floats. Replaces filmToLinear
 * (src/main.cpp:10282-10294; the optional denoiser is not part of this core). */
gp_status gp_film_to_linear(gp_film* film, int64_t n_photons, float* rgb);
/* 8-bit sRGB image (res_x*res_y*3 bytes, row 0 = image top) as filmToRgb8
 * (src/main.cpp:10296-10348): exposure = e_auto * (exposure_comp > 0 ? exposure_comp
 * : 1) with e_auto = 6 when `absolute`, else *anchor when anchor && *anchor > 0, else
 * 0.9 / p99 of max(r, g, b, 0) over the pixels (the element at index
 * (size_t)(0.99 (n - 1)) in order; 1 if p99 <= 0), which is then stored to *anchor
 * when anchor is non-null. The reduction runs on the device; only the bytes come
 * back. exposure_out (optional) receives the exposure used. */
gp_status gp_film_tonemap(gp_film* film, int64_t n_photons, double exposure_comp, int32_t absolute,
                          double* anchor, uint8_t* rgb8, double* exposure_out);
/* gp_film_tonemap without any device allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments and result, bit for
 * bit; its working memory is a fixed scratch array in the library (44 MiB per GPU,
 * held from the library's load until the process ends), so films are limited to
 * 4194304 pixels (2048 x 2048; larger ones return GP_ERR_CAPACITY). All calls on one
 * GPU share that array: calls that overlap on one GPU (from several host threads)
 * race and must be serialised by the caller. */
gp_status gp_film_tonemap_noalloc(gp_film* film, int64_t n_photons, double exposure_comp, int32_t absolute,
                                  double* anchor, uint8_t* rgb8, double* exposure_out);

/* ---- tracing ------------------------------------------------------------ */
/* Trace params->n_photons photons of `scene` through `cam` into `film`
 * (blocking). Energy is ACCUMULATED into *energy (+=), as renderForwardCuda
 * accumulates into eOut. stats may be null. */
gp_status gp_trace(gp_scene* scene, const gp_camera_desc* cam, gp_film* film,
                   const gp_trace_params* params, gp_energy* energy,
                   gp_trace_stats* stats);
/* One photon flight splatted into several cameras, as renderForwardShared /
 * renderForwardSharedCuda (src/main.cpp:11354-11405, src/render_cuda.cu:13873-13897):
 * every diffuse vertex (and the emission point) connects to each camera; films[c]
 * (cams[c]'s resolution, one film per camera) receives exactly what a single-camera
 * gp_trace of the same photons would splat into it, bit for bit. n_cams 1..16. The
 * ledger and the photon/segment counts are those of the one flight; connections and
 * splats are summed over the cameras. */
gp_status gp_trace_shared(gp_scene* scene, int32_t n_cams, const gp_camera_desc* cams, gp_film* const* films,
                          const gp_trace_params* params, gp_energy* energy, gp_trace_stats* stats);

/* Closest hit for n rays (host arrays). rays: n*6 floats (origin.xyz, dir.xyz).
 * prim_out: global primitive id or -1; t_out: hit distance (inf on miss). */
gp_status gp_probe_closest_hit(gp_scene* scene, const float* rays, int64_t n,
                               float tmin, int32_t* prim_out, float* t_out);
/* Any hit on [tmin, tmax_i) for n rays: rays n*6 floats, tmax n floats,
 * occluded_out n bytes (0/1). */
gp_status gp_probe_any_hit(gp_scene* scene, const float* rays, const float* tmax,
                           int64_t n, float tmin, uint8_t* occluded_out);

/* ---- wavefront tuning --------------------------------------------------- */
typedef struct gp_tuning {
    int32_t queue_capacity;  /* paths in flight (power of two, default 2^23)         */
    int32_t block_extend;    /* threads per block of the traversal kernels: fixed at
                                256 (the LDS stacks, splat tables and per-job tables
                                are laid out for 256-lane workgroups); 0 or 256     */
    int32_t block_shade;     /* threads per block of emit/shade: fixed at 256 (0/256) */
    int32_t grid_waves_per_cu; /* waves per CU the fixed grids launch (workgroups =
                                  CUs x this / 4), 4..64; beyond what fits at once they
                                  run in later rounds, which evens out the grid-stride
                                  loops (default 64); 16 = the resident wave slots of
                                  the traversal kernels (persistent threads)         */
    int32_t iters_per_graph; /* iterations per host read of the wavefront state (per
                                captured hipGraph: even, 2..64; 1..64 without graphs) */
    int32_t use_graph;       /* 1 = replay captured graphs, 0 = plain launches       */
    int32_t lds_stack;       /* traversal stack entries per lane held in LDS: 8, 12,
                                16, 24 or 32 (default 12; deeper entries spill to
                                global memory)                                       */
    int32_t profile;         /* 1 = time every kernel launch with HIP events (no graphs) */
    int32_t ray_sort;        /* 1 = reorder each iteration's paths by origin cell
                                (8 x 8 x 8 over the scene) and direction octant before
                                the closest-hit kernel, a counting sort into a third
                                path queue, so that the lanes of a wave trace similar
                                rays; 0 = queue order; 2 = automatic (default): 1 while
                                the largest traced scene has 4096 to 65535 triangles.
                                Results are identical in every mode; gp_autotune tries
                                0 and 1. */
} gp_tuning;

gp_status gp_get_tuning(gp_tuning* out);
gp_status gp_set_tuning(const gp_tuning* in);
/* Times candidate settings on `scene`/`cam` with a short trace and installs
 * the fastest. On failure the tuning in place before the call is restored. */
gp_status gp_autotune(gp_scene* scene, const gp_camera_desc* cam,
                      const gp_trace_params* params, gp_tuning* best_out);

/* ---- serving: many jobs in one wavefront -------------------------------- */
typedef struct gp_server_config {
    int64_t hbm_budget_bytes; /* admission limit for films + queues of live jobs   */
    int32_t max_live_jobs;    /* job slots in the device job table (<= 256)        */
    int32_t order_policy;     /* GP_ORDER_SHORTEST_FIRST (0) or GP_ORDER_ADMISSION */
} gp_server_config;

/* Which live job's photons fill free path slots first. Admission itself is first come
 * first admitted either way. SHORTEST_FIRST: the job with the fewest photons left to
 * emit (a small job admitted behind large ones does not wait for them); ADMISSION:
 * the order of admission. */
#define GP_ORDER_SHORTEST_FIRST 0
#define GP_ORDER_ADMISSION 1

typedef struct gp_job_desc {
    const gp_scene* scene;
    gp_camera_desc camera;
    gp_trace_params params;
    int64_t user_tag;
} gp_job_desc;

typedef struct gp_job_result {
    int64_t job_id;
    int64_t user_tag;
    int32_t status;          /* gp_status of the job                                 */
    int32_t res_x, res_y;
    gp_energy energy;
    gp_trace_stats stats;
    double submit_ms, start_ms, finish_ms;   /* server clock, ms                     */
} gp_job_result;

/* A server is driven from one host thread at a time: submit, step, poll and take
 * may be called in any order but not concurrently on the same server. All servers and
 * traces of a process share one device engine. While a server holds live jobs or
 * chunks in flight the engine is that server's: gp_trace, gp_trace_shared,
 * gp_set_tuning, gp_autotune and other servers' steps return GP_ERR_STATE until it has
 * drained (a step with max_iterations = 0) or is destroyed. A step that fails (e.g.
 * GP_ERR_CAPACITY on a queue overflow) leaves the server's live jobs undefined:
 * destroy the server, which releases the engine. */
gp_status gp_server_create(const gp_server_config* cfg, gp_server** out);
/* Queue a job. It is admitted into the running wavefront as soon as a job slot
 * and its HBM footprint are free. Returns its id in *job_id. */
gp_status gp_server_submit(gp_server* srv, const gp_job_desc* job, int64_t* job_id);
/* Advance the wavefront: queue at least max_iterations more iterations and return
 * (0 = run until every queued job has finished and its results are on the host).
 * With max_iterations > 0 the call returns while up to two chunks of iterations are
 * still running, so the caller's work overlaps the device; the next step, poll or
 * take picks up where it left. Returns the number of jobs finished so far. */
gp_status gp_server_step(gp_server* srv, int32_t max_iterations, int32_t* finished_out);
/* Ids of finished jobs not yet taken, oldest first (at most `capacity`), without
 * waiting for the device. n_out receives the count written. */
gp_status gp_server_poll(gp_server* srv, int64_t* ids_out, int32_t capacity, int32_t* n_out);
/* Fetch a finished job's result and film (xyz res*3, hits res, either may be
 * null). Returns GP_ERR_STATE if the job is not finished. The job's device
 * film is released after this call. */
gp_status gp_server_take_result(gp_server* srv, int64_t job_id, gp_job_result* res,
                                double* xyz, double* hits);
/* Jobs waiting for admission, jobs admitted and not yet retired, and the HBM
 * footprint the admitted ones hold against hbm_budget_bytes (any pointer may be null). */
gp_status gp_server_pending(gp_server* srv, int32_t* queued, int32_t* live, int64_t* live_bytes);
void gp_server_destroy(gp_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* GOODPHOTONS_H */
