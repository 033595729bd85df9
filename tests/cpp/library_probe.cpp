// C++ lifecycle of the OpenCorr-shaped API over libfasticgn, in the shape of
// the reference's smoke/library_boundary_probe.cpp:41-62 (LCG images, known
// (+2, -1) shift, radius guards, cached layout) followed by ICGN2D1.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "fasticgn_opencorr.hpp"

using namespace opencorr;

static int failures = 0;
static void check(const char* label, bool ok)
{
    failures += !ok;
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", label);
}

int main()
{
    const int W = 192, H = 128;
    std::vector<float> ref(W * H), tar(W * H);
    unsigned state = 12345;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            state = 1664525u * state + 1013904223u;
            ref[y * W + x] = float(state >> 24);
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) tar[y * W + x] = (x >= 2 && y + 1 < H) ? ref[(y + 1) * W + x - 2] : 127.f;
    Image2D R(ref.data(), W, H), T(tar.data(), W, H);

    for (auto r : std::vector<std::pair<int, int>>{ { 7, 15 }, { 15, 7 }, { 0, 0 }, { -1, -1 }, { 23171, 23171 } }) {
        bool rejected = false;
        try { FFTCC2D s(r.first, r.second); } catch (const std::invalid_argument&) { rejected = true; }
        check("invalid-or-rectangular-radius", rejected);
    }

    std::vector<POI2D> layout;
    for (int y = 32; y < 96; y += 16)
        for (int x = 32; x < 160; x += 16) layout.emplace_back(x, y);
    FFTCC2D fftcc(7, 7);
    fftcc.setImages(R, T);
    fftcc.prepareRef(layout);
    check("prepared-layout", fftcc.hasPreparedReference(R));
    auto q = layout;
    fftcc.compute(q);
    int correct = 0;
    for (const auto& p : q) correct += p.deformation.u == 2 && p.deformation.v == -1 && p.result.zncc > .5f;
    check("square-translation-truth", correct == (int)layout.size());
    bool rejected = false;
    try { auto bad = layout; fftcc.compute(bad, std::vector<size_t>(layout.size(), 0)); }
    catch (const std::invalid_argument&) { rejected = true; }
    check("indexed-layout-mismatch-rejected", rejected);

    ICGN2D1 icgn(7, 7, 0.001f, 10.f);
    icgn.setImages(R, T);
    icgn.prepare();
    icgn.compute(q);
    int close = 0;
    for (const auto& p : q) close += std::fabs(p.deformation.u - 2.f) < 0.01f && std::fabs(p.deformation.v + 1.f) < 0.01f;
    check("icgn-integer-shift", close == (int)q.size());

    // multi-frame driver and learned seeder (namespace fasticgn)
    fasticgn::FrameDriver drv(7);
    drv.setReference(R);
    drv.setPois(layout);
    drv.setGraph(true);
    std::vector<POI2D> res;
    const fi_frame_stats s0 = drv.process(T, res);
    const fi_frame_stats s1 = drv.process(T, res);
    close = 0;
    for (const auto& p : res) close += std::fabs(p.deformation.u - 2.f) < 0.01f && std::fabs(p.deformation.v + 1.f) < 0.01f;
    check("frame-driver-sequence", close == (int)layout.size() && s0.skip_fftcc_used == 0 && s1.skip_fftcc_used == 1);
    fasticgn::LearnedSeeder seeder;
    const std::vector<float> params = seeder.params();
    check("seeder-param-count", params.size() == fasticgn::LearnedSeeder::paramCount() && params.size() == 44534);
    drv.setSeeder(&seeder);
    drv.resetSequence();
    const fi_frame_stats s2 = drv.process(T, res, FI_POLICY_LEARNED_SEED);
    close = 0;
    for (const auto& p : res) close += std::fabs(p.deformation.u - 2.f) < 0.01f && std::fabs(p.deformation.v + 1.f) < 0.01f;
    check("frame-driver-learned-seed", close == (int)layout.size() && s2.seed_pois == (int64_t)layout.size());
    drv.setSeedPasses(2);
    drv.resetSequence();
    const fi_frame_stats s3 = drv.process(T, res, FI_POLICY_LEARNED_SEED);
    close = 0;
    for (const auto& p : res) close += std::fabs(p.deformation.u - 2.f) < 0.01f && std::fabs(p.deformation.v + 1.f) < 0.01f;
    check("frame-driver-two-seed-passes", close == (int)layout.size() && s3.seed_pois == (int64_t)layout.size());
    bool refused = false;
    try {
        drv.setSeedPasses(5);
    } catch (const std::invalid_argument&) {
        refused = true;
    }
    check("seed-passes-range", refused);
    drv.setSeeder(nullptr);
    return failures ? 1 : 0;
}
