// C++ caller of the forward core through gp_forward.hpp: renders an empty Cornell
// box (5 wall quads + ceiling light quad) with renderForwardGpu, then again through
// a ForwardSession in two batches, and once more checkpointed after the first batch
// and resumed in a second session; then a two-camera SharedSession in two batches,
// and again stopped after the first and resumed from the downloaded totals (seeded
// films); writes the films and ledgers as raw doubles and prints the session checks.
// Usage: forward_demo <out-prefix> <res> <photons>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gp_forward.hpp"

static void quad(gp::Scene& s, gp::Vec3d a, gp::Vec3d b, gp::Vec3d c, gp::Vec3d d, int m) {
    s.tris.push_back({a, b, c, m});
    s.tris.push_back({a, c, d, m});
}

static gp_spectrum spec(int kind, double p0, double p1 = 0, double p2 = 0, double p3 = 0) {
    gp_spectrum s{};
    s.kind = kind;
    s.p[0] = p0; s.p[1] = p1; s.p[2] = p2; s.p[3] = p3;
    return s;
}

static gp_material diffuse(gp_spectrum r) {
    gp_material m{};
    m.type = GP_MAT_DIFFUSE;
    m.roughness = 0.1;
    m.reflect = r;
    m.ior = spec(GP_SPEC_CONST, 1.5);
    return m;
}

static bool save(const std::string& path, const gp::Film& f, const gp::EnergyReport& e) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const double hdr[7] = {(double)f.resX, (double)f.resY, e.emitted, e.absorbed, e.sensor, e.escaped, e.residual};
    std::fwrite(hdr, sizeof(double), 7, fp);
    std::fwrite(f.xyz.data(), sizeof(gp::Vec3d), f.xyz.size(), fp);
    std::fwrite(f.hits.data(), sizeof(double), f.hits.size(), fp);
    std::fclose(fp);
    return true;
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <out-prefix> <res> <photons>\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const int res = std::atoi(argv[2]);
    const long long N = std::atoll(argv[3]);
    if (!gp::gpuAvailable()) {
        std::fprintf(stderr, "no gfx950 device\n");
        return 3;
    }
    gp::Scene s;
    s.mats = {diffuse(spec(GP_SPEC_CONST, 0.75)), diffuse(spec(GP_SPEC_STEP, 0.05, 0.8, 600.0)),
              diffuse(spec(GP_SPEC_GAUSS, 0.05, 0.75, 540.0, 30.0)), diffuse(spec(GP_SPEC_CONST, 0.0))};
    quad(s, {0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, 0);
    quad(s, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}, 0);
    quad(s, {0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, 0);
    quad(s, {0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, 1);
    quad(s, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, 2);
    quad(s, {0.375, 0.999, 0.375}, {0.625, 0.999, 0.375}, {0.625, 0.999, 0.625}, {0.375, 0.999, 0.625}, 3);
    gp_emitter e{};
    e.origin[0] = 0.375; e.origin[1] = 0.999; e.origin[2] = 0.375;
    e.u[0] = 0.25; e.v[2] = 0.25;
    e.normal[1] = -1.0;
    e.area = 0.0625;
    e.beam_dir[0] = 1.0;
    e.spd = spec(GP_SPEC_BLACKBODY, 6500.0, 1.0);
    e.spd.band_lo = 380.0;
    e.spd.band_hi = 780.0;
    s.emitters.push_back(e);
    gp::Camera cam;
    cam.eye = {0.5, 0.5, 2.7};
    cam.target = {0.5, 0.5, 0.5};
    cam.fovYDeg = 40.0;
    gp::gMaxDepth = 8;

    std::printf("device %s supported %d\n", gp::gpuDeviceName().c_str(), (int)gp::gpuForwardSupported(s));
    gp::EnergyReport e1;
    gp::Film f1 = gp::renderForwardGpu(s, cam, res, res, N, e1, true, 'B', 0, true, 4);
    gp::ForwardSession sess;
    if (!sess.begin(s, cam, res, res, 4, N)) return 1;
    if (!sess.batch(N / 2, 0) || !sess.batch(N - N / 2, (unsigned long long)(N / 2))) return 1;
    gp::Film f2;
    gp::EnergyReport e2;
    if (!sess.download(f2, e2)) return 1;
    // checkpoint after the first half, resume in a new session, finish there
    gp::ForwardSession a, b;
    const uint64_t guard = gp_checkpoint_guard((int64_t)s.tris.size(), 0, (int64_t)s.emitters.size(), 1.0, 'B', res, res);
    const std::string ck = out + ".ftbuf";
    if (!a.begin(s, cam, res, res, 4, N) || !a.batch(N / 2, 0) || !a.saveCheckpoint(ck, N / 2, guard)) return 1;
    if (!b.begin(s, cam, res, res, 4, N)) return 1;
    const long long done = b.resume(ck, guard);
    if (done != N / 2 || !b.batch(N - done, (unsigned long long)done)) return 1;
    gp::Film f4;
    gp::EnergyReport e4;
    if (!b.download(f4, e4)) return 1;
    // shared flight: the first camera is the same view, the second a different one
    gp::Camera cam2 = cam;
    cam2.eye = {1.3, 0.7, 2.2};
    gp::EnergyReport e5;
    std::vector<gp::Film> fs = gp::renderForwardSharedGpu(s, {cam, cam2}, {res, res / 2}, {res, res / 2}, N, e5,
                                                         true, 'B', 0, true, 4);
    // resident two-camera session: uninterrupted, and resumed from downloaded totals
    std::vector<gp::Film> g1, g2, seed;
    gp::EnergyReport eg1, eg2, eseed;
    std::vector<double> poll;
    {
        gp::SharedSession ss;
        if (!ss.begin(s, {cam, cam2}, {res, res / 2}, {res, res / 2}, 'B', 4, 4 * N)) return 1;
        if (!ss.batch(N / 2, 0) || !ss.batch(N - N / 2, (unsigned long long)(N / 2)) || !ss.hits0(poll) ||
            !ss.download(g1, eg1))
            return 1;
    }
    {
        gp::SharedSession ss;
        if (!ss.begin(s, {cam, cam2}, {res, res / 2}, {res, res / 2}, 'B', 4, 4 * N) || !ss.batch(N / 2, 0) ||
            !ss.download(seed, eseed))
            return 1;
    }
    {
        gp::SharedSession ss;
        if (!ss.begin(s, {cam, cam2}, {res, res / 2}, {res, res / 2}, 'B', 4, 4 * N, &seed, &eseed, N / 2) ||
            !ss.batch(N - N / 2, (unsigned long long)(N / 2)) || !ss.download(g2, eg2))
            return 1;
    }
    bool same = g1.size() == 2 && g2.size() == 2;
    for (size_t c = 0; same && c < 2; ++c) {
        same = g1[c].hits == g2[c].hits;
        for (size_t i = 0; same && i < g1[c].xyz.size(); ++i)
            same = g1[c].xyz[i].x == g2[c].xyz[i].x && g1[c].xyz[i].y == g2[c].xyz[i].y && g1[c].xyz[i].z == g2[c].xyz[i].z;
    }
    std::printf("shared_session_resume_exact %d\n", (int)same);
    std::printf("shared_session_hits0_equal %d\n", (int)(g1.size() == 2 && poll == g1[0].hits));
    std::printf("shared_session_hits_equal_single %d\n", (int)(g1.size() == 2 && g1[0].hits == f1.hits));
    gp::EnergyReport e3;
    gp::Film f3 = gp::renderForwardGpu(s, cam, res, res, N, e3, true, 'A');   // unsupported model
    double sum3 = 0;
    for (auto& v : f3.xyz) sum3 += v.x + v.y + v.z;
    std::printf("emitted %.6e absorbed %.6e escaped %.6e residual %.6e\n", e1.emitted, e1.absorbed,
                e1.escaped, e1.residual);
    std::printf("unsupported_model_empty %d\n", (int)(sum3 == 0.0 && (int)f3.xyz.size() == res * res));
    if (!save(out + "_single.bin", f1, e1) || !save(out + "_session.bin", f2, e2) ||
        !save(out + "_resumed.bin", f4, e4) || fs.size() != 2 || !save(out + "_shared0.bin", fs[0], e5))
        return 1;
    return 0;
}
