// Resumable film checkpoints in the reference's .ftbuf sidecar format.
//
// Layout written (src/main.cpp:12443-12462, "FTBUF01"): 8-byte magic, int32 res_x,
// res_y, mode, int64 cumulative photons, 5 doubles of energy (emitted, absorbed,
// sensor, escaped, residual), uint64 identity guard, then the film's X, Y, Z sums as
// res_x*res_y packed double triples and the hit counts as res_x*res_y doubles, all
// little-endian. The reference's reader stops there and probes for optional tagged
// sections after it (it skips a tag it does not know), so a file written here resumes
// in the reference's CPU renderer unchanged.
//
// Appended after the film blobs is one section of our own, tag "GPFILM01": the film's
// fixed-point scale, saturation value and photon budget, followed by the raw int64
// words. gp_film_load_checkpoint restores the device film from it exactly, so a
// render saved and resumed is bit-identical to one that never stopped. A file without
// the section (written by the reference) holds only doubles; resuming it is the
// reference's own host-side merge of the downloaded film, which this core does not
// replace.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "gp_runtime.hpp"

using namespace gp;

namespace {

constexpr char kMagic[8] = {'F', 'T', 'B', 'U', 'F', '0', '1', '\n'};
constexpr char kFixedTag[8] = {'G', 'P', 'F', 'I', 'L', 'M', '0', '1'};

struct FileCloser {
    void operator()(FILE* f) const { if (f) std::fclose(f); }
};
using File = std::unique_ptr<FILE, FileCloser>;

bool put(FILE* f, const void* p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
bool get(FILE* f, void* p, size_t n) { return std::fread(p, 1, n, f) == n; }

}  // namespace

extern "C" {

// checkpointGuard (src/main.cpp:12317-12329): FNV-1a style mix of whole 64-bit words.
uint64_t gp_checkpoint_guard(int64_t n_tris, int64_t n_spheres, int64_t n_emitters, double total_power,
                             int32_t mode, int32_t res_x, int32_t res_y) {
    uint64_t h = 14695981039346656037ULL;
    auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ULL; };
    mix((uint64_t)n_tris);
    mix((uint64_t)n_spheres);
    mix((uint64_t)n_emitters);
    uint64_t tp;
    std::memcpy(&tp, &total_power, sizeof tp);
    mix(tp);
    mix((uint64_t)(unsigned char)mode);
    mix((uint64_t)(unsigned)res_x);
    mix((uint64_t)(unsigned)res_y);
    return h;
}

gp_status gp_film_save_checkpoint(gp_film* film, const char* path, int64_t n_photons, const gp_energy* energy,
                                  int32_t mode, uint64_t guard) {
    if (!film || !path) return fail(GP_ERR_INVALID, "null argument");
    if (n_photons < 0) return fail(GP_ERR_INVALID, "negative photon count");
    gp_status st = requireDevice();
    if (st != GP_OK) return st;
    const size_t np = (size_t)film->rx * film->ry;
    std::vector<long long> words(np * 4);
    GP_HIP(hipMemcpy(words.data(), film->d, words.size() * sizeof(long long), hipMemcpyDeviceToHost));
    const double inv = film->scale > 0 ? 1.0 / film->scale : 0.0;
    std::vector<double> xyz(np * 3), hits(np);
    for (size_t i = 0; i < np; ++i) {
        xyz[3 * i + 0] = (double)words[4 * i + 0] * inv;
        xyz[3 * i + 1] = (double)words[4 * i + 1] * inv;
        xyz[3 * i + 2] = (double)words[4 * i + 2] * inv;
        hits[i] = (double)words[4 * i + 3];
    }
    File f(std::fopen(path, "wb"));
    if (!f) return fail(GP_ERR_INVALID, std::string("cannot open ") + path + " for writing");
    const int32_t rx = film->rx, ry = film->ry, m = (int32_t)(unsigned char)mode;
    const long long N = n_photons;
    const double en[5] = {energy ? energy->emitted : 0.0, energy ? energy->absorbed : 0.0,
                          energy ? energy->sensor : 0.0, energy ? energy->escaped : 0.0,
                          energy ? energy->residual : 0.0};
    bool ok = put(f.get(), kMagic, 8) && put(f.get(), &rx, 4) && put(f.get(), &ry, 4) && put(f.get(), &m, 4) &&
              put(f.get(), &N, 8) && put(f.get(), en, sizeof en) && put(f.get(), &guard, 8) &&
              put(f.get(), xyz.data(), xyz.size() * sizeof(double)) &&
              put(f.get(), hits.data(), hits.size() * sizeof(double));
    // fixed-point section
    const double fscale = film->scale, budget = film->splatBudget, used = film->splatUsed;
    const uint64_t qmax = film->qmax;
    const int64_t maxPh = film->maxPhotons;
    ok = ok && put(f.get(), kFixedTag, 8) && put(f.get(), &fscale, 8) && put(f.get(), &qmax, 8) &&
         put(f.get(), &budget, 8) && put(f.get(), &used, 8) && put(f.get(), &maxPh, 8) &&
         put(f.get(), words.data(), words.size() * sizeof(long long));
    if (!ok) return fail(GP_ERR_INVALID, std::string("write to ") + path + " failed");
    if (std::fclose(f.release()) != 0) return fail(GP_ERR_INVALID, std::string("closing ") + path + " failed");
    return GP_OK;
}

gp_status gp_film_load_checkpoint(const char* path, int32_t mode, uint64_t guard, gp_film** out,
                                  int64_t* n_photons_out, gp_energy* energy_out) {
    if (!path || !out) return fail(GP_ERR_INVALID, "null argument");
    *out = nullptr;
    File f(std::fopen(path, "rb"));
    if (!f) return fail(GP_ERR_INVALID, std::string("cannot open ") + path);
    char magic[8];
    int32_t rx = 0, ry = 0, m = 0;
    long long N = 0;
    double en[5];
    uint64_t g = 0;
    if (!get(f.get(), magic, 8) || std::memcmp(magic, kMagic, 8) != 0)
        return fail(GP_ERR_INVALID, std::string(path) + " is not an FTBUF01 checkpoint");
    if (!get(f.get(), &rx, 4) || !get(f.get(), &ry, 4) || !get(f.get(), &m, 4) || !get(f.get(), &N, 8) ||
        !get(f.get(), en, sizeof en) || !get(f.get(), &g, 8))
        return fail(GP_ERR_INVALID, std::string(path) + " is truncated");
    if (m != (int32_t)(unsigned char)mode || g != guard)
        return fail(GP_ERR_STATE, "checkpoint does not match this render (scene/mode/resolution differ)");
    if (rx <= 0 || ry <= 0 || (int64_t)rx * ry > (int64_t(1) << 28) || N < 0)
        return fail(GP_ERR_INVALID, std::string(path) + " has an invalid header");
    const size_t np = (size_t)rx * ry;
    if (std::fseek(f.get(), (long)(np * 4 * sizeof(double)), SEEK_CUR) != 0)
        return fail(GP_ERR_INVALID, std::string(path) + " is truncated");
    char tag[8];
    if (!get(f.get(), tag, 8) || std::memcmp(tag, kFixedTag, 8) != 0)
        return fail(GP_ERR_UNSUPPORTED,
                    "checkpoint has no fixed-point film section: resume it on the host (merge the downloaded film)");
    double fscale, budget, used;
    uint64_t qmax;
    int64_t maxPh;
    if (!get(f.get(), &fscale, 8) || !get(f.get(), &qmax, 8) || !get(f.get(), &budget, 8) ||
        !get(f.get(), &used, 8) || !get(f.get(), &maxPh, 8))
        return fail(GP_ERR_INVALID, std::string(path) + " is truncated");
    // a film that has been traced into has a positive scale, a power-of-two saturation
    // value of at most 2^52 and a budget covering what it used; a fresh one has none
    const bool fresh = fscale == 0.0 && qmax == 0 && budget == 0.0 && used == 0.0;
    const bool traced = std::isfinite(fscale) && fscale > 0.0 && qmax != 0 && (qmax & (qmax - 1)) == 0 &&
                        qmax <= (1ull << 52) && std::isfinite(budget) && used >= 0.0 && used <= budget * (1.0 + 1e-12);
    if (!(fresh || traced) || maxPh <= 0)
        return fail(GP_ERR_INVALID, std::string(path) + " has an invalid fixed-point film section");
    std::vector<long long> words(np * 4);
    if (!get(f.get(), words.data(), words.size() * sizeof(long long)))
        return fail(GP_ERR_INVALID, std::string(path) + " is truncated");
    gp_film* film = nullptr;
    gp_status st = gp_film_create(rx, ry, maxPh, &film);
    if (st != GP_OK) return st;
    std::unique_ptr<gp_film, void (*)(gp_film*)> guardFilm(film, gp_film_destroy);
    film->scale = fscale;
    film->qmax = qmax;
    film->splatBudget = budget;
    film->splatUsed = used;
    GP_HIP(hipMemcpy(film->d, words.data(), words.size() * sizeof(long long), hipMemcpyHostToDevice));
    *out = guardFilm.release();
    if (n_photons_out) *n_photons_out = N;
    if (energy_out) *energy_out = gp_energy{en[0], en[1], en[2], en[3], en[4]};
    return GP_OK;
}

}  // extern "C"
