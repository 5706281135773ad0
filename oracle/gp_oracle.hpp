// Test infrastructure only: declarations of the oracle's C ABI.
//
// The oracle is a CPU restatement, in double precision, of the reference's mode-B
// forward light tracer for the triangle/quad-light/pinhole subset the GPU core
// supports. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
// use it, and only as the checker; the product never links or calls it.
//
// Pinning (how each part is checked, see tests/test_oracle_*.py):
//   rng, spectra, CIE, emission sampler, hero bundle, cosine/glossy lobes,
//   camera basis/projection, BVH closest/any hit, full films + energy ledgers:
//     bit-exact against golden vectors written by the reference itself
//     (oracle/_ref, tests/golden/make_golden.py), plus closed forms
//     (PCG32/SplitMix64 published constants, N-BK7 n_d, Fresnel at normal
//     incidence, energy ledger closure, brute-force closest hit).
//   rr_depth > 0 and the closest-hit ambiguity margin are this repository's own
//     extensions with no reference counterpart: rr_depth is pinned by unbiasedness
//     (same expectation as rr_depth = 0) and the ledger; the margin report by
//     constructed rays whose answers the geometry fixes (exact ties, shared edges,
//     silhouette offsets on either side of the margin; tests/test_oracle_margin.py).
#pragma once
#include <cstdint>
#include "../include/goodphotons.h"

extern "C" {
void* or_scene_create(const gp_scene_desc* d);
void or_scene_destroy(void* s);
int64_t or_scene_bvh_nodes(void* s);
double or_scene_total_power(void* s);
// Flat mirror planes of the scene (n.xyz, d, lo.xyz, hi.xyz per plane, up to cap);
// returns their number.
int or_scene_mirror_planes(void* s, double* out10, int cap);

// stats7: segments, connections, splats, then the traversal statistics
// (TraversalStats) of the segments' closest hits (node visits, leaf tests) and of the
// camera connections' any hits (node visits, leaf tests).
int or_render(void* s, const gp_camera_desc* cam, int64_t N, uint64_t seedBase, uint64_t salt,
              int heroC, int maxBounce, int rrDepth, int nThreads, double* xyz, double* hits,
              gp_energy* eOut, uint64_t* stats7);

int or_closest_hit(void* s, const double* rays, int64_t n, double tmin, int32_t* prim,
                   double* t, int64_t* nodeVisits, int64_t* leafTests);
int or_closest_hit_linear(void* s, const double* rays, int64_t n, double tmin, int32_t* prim,
                          double* t);
int or_closest_hit_margin(void* s, const double* rays, int64_t n, double tmin, double relMargin,
                          int32_t* prim, double* t, double* t2, uint8_t* edge);
// Per ray: exact closest hit (prim, t); nearest grown-hit distance of another
// triangle (t2); winner grazed (edge); nearest robust (shrunk) hit (tRobust); and,
// for the queried triangle query[i], its grown-hit distance (tq) and graze flag.
int or_closest_hit_ambiguity(void* s, const double* rays, const int32_t* query, int64_t n,
                             double tmin, double relMargin, int32_t* prim, double* t, double* t2,
                             uint8_t* edge, double* tRobust, double* tq, uint8_t* qGraze);
int or_occluded(void* s, const double* rays, const double* maxDist, int64_t n, uint8_t* occ);

void or_pcg_stream(uint64_t seq, uint64_t seed, int n, uint32_t* out);
void or_seed_unit_uniforms(uint64_t k, uint64_t salt, int n, double* out);
uint64_t or_mix64(uint64_t x);
// pool: samples of GP_SPEC_SAMPLED spectra (may be null for the analytic kinds)
void or_spectrum_eval(const gp_spectrum* s, const double* pool, const double* lam, int n, double* out);
void or_cie(const double* lam, int n, double* xyz);
double or_cie_y_integral();
uint64_t or_checkpoint_guard(int64_t nTris, int64_t nSpheres, int64_t nEmitters, double totalPower,
                             int mode, int res, int resY);
int or_read_checkpoint(const char* path, int res, int resY, uint64_t guard, int mode, long long* N,
                       double* energy5, double* xyz, double* hits, long long* trailing);
void or_film_to_linear(const double* xyz, int64_t npix, double N, double* lin);
void or_film_to_rgb8(const double* xyz, int W, int H, double N, double expComp, int absolute,
                     double* lockAnchor, uint8_t* img, double* eAutoOut, double* exposureOut);
void or_emission_sample(const gp_spectrum* s, const double* pool, const double* u, int n, double* lam, double* pdf,
                        double* integral);
int or_sample_bundle(const gp_spectrum* s, const double* pool, double u, int C, double* lam, double* pdf);
void or_cosine_hemisphere(const double* nrm, uint64_t seq, uint64_t seed, int n, double* out);
void or_sample_glossy(const double* mdir, double rough, uint64_t seq, uint64_t seed, int n,
                      double* out);
void or_camera_basis(const gp_camera_desc* cd, double* out14);
int or_camera_project(const gp_camera_desc* cd, const double* pts, int n, int32_t* px,
                      int32_t* py, double* cosCam, double* dist2, double* omega);
}
