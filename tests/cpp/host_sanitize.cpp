// Host-only code of the projector library under AddressSanitizer + UndefinedBehaviorSanitizer
// (or ThreadSanitizer): the geometry encoder and decoder, DEN/NRRD file I/O, the phantom generator,
// the serving arena's best-fit allocator, and the CPU oracle's threaded projections. Built and run
// by tests/test_host_sanitizers.py on the CPU. Exit code 0 and no sanitizer report = pass.
#include "arena.h"
#include "ctlproj.h"
#include "geometry.h"
#include "phantom.h"

extern "C" {
#include "ctl_oracle.h"
}

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <thread>
#include <vector>

static int g_fail = 0;
#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++g_fail;                                                             \
        }                                                                         \
    } while (0)

static void geometry(uint32_t threads)
{
    // every trajectory type on both gantries, then decode -> re-encode, from several threads
    auto work = [](uint32_t seed) {
        ctlgeo::Detector flat;
        flat.channels = 64;
        flat.rows = 48;
        flat.pixel_w = flat.pixel_h = 0.8;
        ctlgeo::Detector cyl = flat;
        cyl.channels = 16;
        cyl.rows = 32;
        cyl.modules = 12;
        cyl.cylindrical = true;
        cyl.angulation = 0.015;
        for (int type = 0; type < 5; ++type) {
            ctlgeo::Trajectory t;
            t.type = ctlgeo::Trajectory::Type(type);
            const double p[5][6] = { { 750, 0.1, 3.0 }, { 3.0, 750, 0.0, 0.2, 1.0 }, { 3.0, 750, 100, 0.3, 0 },
                                     { 0.05, 1.0, -10, 0 }, { 0.3 } };
            for (int k = 0; k < 6; ++k) t.p[k] = p[type][k];
            for (int gt = 0; gt < 2; ++gt) {
                ctlgeo::Gantry g;
                g.type = ctlgeo::Gantry::Type(gt);
                g.carm_span = 1200;
                g.sdd = 1085;
                g.sid = 595;
                const ctlgeo::Detector& d = gt ? cyl : flat;
                const uint32_t views = 7 + seed % 5;
                for (uint32_t v = 0; v < views; ++v) {
                    if (!ctlgeo::prepareView(t, v, views, d, g)) break;
                    std::vector<ctlgeo::PMat> P(d.modules);
                    ctlgeo::encodeSingleView(g, d, ctlgeo::Vec3{ { 0.1 * seed, 0.0, 0.0 } }, P.data());
                    for (auto& m : P) ctlgeo::normalize(m);
                    ctlgeo::DecodedView dv;
                    CHECK(ctlgeo::decodeView(P.data(), d.modules, d.channels, d.rows, ctlgeo::DecodePixelSize,
                                             d.pixel_w, d.pixel_h, &dv));
                    std::vector<ctlgeo::PMat> Q(d.modules);
                    ctlgeo::encodeModules(dv.source, dv.modules.data(), dv.modules.size(), d.channels, d.rows,
                                          dv.pixel_w, dv.pixel_h, dv.skew, Q.data());
                    for (uint32_t m = 0; m < d.modules; ++m) {
                        ctlgeo::normalize(Q[m]);
                        double err = 0, scale = 0;
                        for (int i = 0; i < 12; ++i) {
                            err = std::fmax(err, std::fabs(Q[m].p[i] - P[m].p[i]));
                            scale = std::fmax(scale, std::fabs(P[m].p[i]));
                        }
                        CHECK(err <= 1e-8 * scale);
                    }
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (uint32_t k = 0; k < threads; ++k) th.emplace_back(work, k);
    for (auto& t : th) t.join();
}

static void files(const std::string& dir)
{
    // DEN and NRRD round trips of every element size, plus malformed inputs
    std::vector<double> d(3 * 4 * 5);
    for (size_t i = 0; i < d.size(); ++i) d[i] = 0.5 * double(i) - 7.0;
    const std::string den = dir + "/a.den";
    CHECK(ctlproj_den_write(den.c_str(), d.data(), 8, 3, 4, 5) == CTLPROJ_OK);
    uint32_t r = 0, c = 0, n = 0, eb = 0;
    CHECK(ctlproj_den_info(den.c_str(), &r, &c, &n, &eb) == CTLPROJ_OK && r == 3 && c == 4 && n == 5 && eb == 8);
    std::vector<double> back(d.size());
    CHECK(ctlproj_den_read(den.c_str(), back.data(), back.size() * 8) == CTLPROJ_OK && back == d);
    CHECK(ctlproj_den_read(den.c_str(), back.data(), 8) == CTLPROJ_ERR_INVALID_ARGUMENT);
    CHECK(ctlproj_den_info((dir + "/missing.den").c_str(), &r, &c, &n, &eb) != CTLPROJ_OK);
    std::vector<float> f(6 * 5 * 4);
    for (size_t i = 0; i < f.size(); ++i) f[i] = float(i) * 0.25f;
    const uint64_t sizes[3] = { 6, 5, 4 };
    const double sp[3] = { 0.5, 0.6, 0.7 }, org[3] = { 1.0, -2.0, 3.0 };
    const std::string nrrd = dir + "/a.nrrd";
    CHECK(ctlproj_nrrd_write(nrrd.c_str(), f.data(), CTLPROJ_ELEM_F32, 3, sizes, "\"x\" \"y\" \"z\"", sp, org,
                             "key:=value\n") == CTLPROJ_OK);
    ctlproj_nrrd_header h{};
    char kv[256];
    CHECK(ctlproj_nrrd_read_header(nrrd.c_str(), &h, kv, sizeof kv) == CTLPROJ_OK && h.count == f.size());
    std::vector<float> fb(f.size());
    CHECK(ctlproj_nrrd_read(nrrd.c_str(), fb.data(), fb.size() * 4) == CTLPROJ_OK && fb == f);
    // a truncated file: header promises more data than the file holds
    const std::string cut = dir + "/cut.nrrd";
    {
        FILE* in = std::fopen(nrrd.c_str(), "rb");
        FILE* out = std::fopen(cut.c_str(), "wb");
        std::vector<char> buf(1 << 16);
        const size_t got = std::fread(buf.data(), 1, buf.size(), in);
        std::fwrite(buf.data(), 1, got > 100 ? got - 100 : got, out);
        std::fclose(in);
        std::fclose(out);
    }
    CHECK(ctlproj_nrrd_read_header(cut.c_str(), &h, nullptr, 0) == CTLPROJ_OK && !h.data_complete);
    CHECK(ctlproj_nrrd_read(cut.c_str(), fb.data(), fb.size() * 4) != CTLPROJ_OK);
}

static void phantom()
{
    ctlphantom::PhantomSpec s{ { 33, 29, 31 }, { 1.1f, 0.9f, 1.3f }, 9, 77, 2.0f, 9.0f, 0.005f, 0.03f, 0.02f };
    std::vector<float> v(size_t(33) * 29 * 31);
    ctlphantom::ellipsoids(s, v.data());
    double sum = 0;
    for (float x : v) sum += x;
    CHECK(sum > 0 && std::isfinite(sum));
}

static void arena()
{
    // random allocations and releases against a model of the occupied blocks
    ctlsrv::Arena a;
    const size_t size = 1 << 20;
    a.init(size);
    std::mt19937_64 rng(5);
    std::vector<std::pair<size_t, size_t>> live;
    for (int it = 0; it < 20000; ++it) {
        if (live.empty() || rng() % 3) {
            const size_t bytes = 1 + rng() % 40000;
            size_t off = 0;
            if (a.alloc(bytes, &off)) {
                CHECK(off % ctlsrv::kAlign == 0 && off + ctlsrv::alignUp(bytes) <= size);
                for (const auto& l : live) CHECK(off + ctlsrv::alignUp(bytes) <= l.first || l.first + ctlsrv::alignUp(l.second) <= off);
                live.emplace_back(off, bytes);
            }
        } else {
            const size_t k = rng() % live.size();
            a.release(live[k].first, live[k].second);
            live.erase(live.begin() + long(k));
        }
        size_t used = 0;
        for (const auto& l : live) used += ctlsrv::alignUp(l.second);
        CHECK(a.used() == used);
    }
    for (const auto& l : live) a.release(l.first, l.second);
    CHECK(a.used() == 0 && a.freeBlocks().size() == 1 && a.freeBlocks().begin()->second == size);
}

static void oracle(int threads)
{
    // ref32, ref64 and refcpu on a small phantom with the oracle's worker threads
    const uint32_t dims[3] = { 20, 18, 16 };
    const float vs[3] = { 1.0f, 1.2f, 0.9f }, off[3] = { 1.0f, -1.0f, 0.5f };
    std::vector<float> vol(size_t(20) * 18 * 16);
    ctlphantom::PhantomSpec s{ { 20, 18, 16 }, { 1.0f, 1.2f, 0.9f }, 5, 3, 2.0f, 6.0f, 0.005f, 0.03f, 0.02f };
    ctlphantom::ellipsoids(s, vol.data());
    ctlgeo::Detector d;
    d.channels = 24;
    d.rows = 20;
    ctlgeo::Gantry g;
    g.carm_span = 1200;
    ctlgeo::Trajectory t;
    t.p[0] = 750;
    t.p[2] = 2.0;
    const uint32_t views = 3;
    std::vector<double> pm(views * 12);
    for (uint32_t v = 0; v < views; ++v) {
        ctlgeo::prepareView(t, v, views, d, g);
        ctlgeo::PMat P;
        ctlgeo::encodeSingleView(g, d, ctlgeo::Vec3{ { 0, 0, 0 } }, &P);
        ctlgeo::normalize(P);
        for (int i = 0; i < 12; ++i) pm[v * 12 + i] = P.p[i];
    }
    oracle_settings os{ { 2, 1 }, 0.2f, 1, 1 };
    for (int mode = 0; mode < 3; ++mode) {
        std::vector<float> out(size_t(views) * 24 * 20);
        uint64_t n = 0;
        CHECK(oracle_project(vol.data(), dims, vs, off, pm.data(), views, 1, 24, 20, &os, mode, threads, nullptr,
                             0, out.data(), nullptr, &n) == 0);
        double sum = 0;
        for (float x : out) sum += x;
        CHECK(n > 0 && sum > 0 && std::isfinite(sum));
    }
}

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    const bool threaded_only = argc > 2 && std::string(argv[2]) == "threads";
    geometry(4);
    oracle(4);
    if (!threaded_only) {
        files(dir);
        phantom();
        arena();
    }
    std::printf("host checks: %d failed\n", g_fail);
    return g_fail ? 1 : 0;
}
