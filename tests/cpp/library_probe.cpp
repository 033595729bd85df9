// C++ lifecycle of the OpenCorr-shaped API over libfasticgn, written in the
// call shapes of the reference's callers:
//   - opencorr_cli.cpp:117-133: new FFTCC2D(r, r, threads), setImages, compute;
//     new ICGN2D1(r, r, conv, max_iteration, threads), setImages, prepare, compute
//   - smoke/library_boundary_probe.cpp:41-70, 125-134: Image2D(w, h) filled
//     through eg_mat, the LCG known answer (+2, -1), the cached-layout
//     lifecycle, radius and thread-count guards, and DIC::setSubset on the
//     base class rejected by FFTCC2D::prepareRef
// followed by Image2D(file) loading and the multi-frame driver.
#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fasticgn_opencorr.hpp"

using namespace opencorr;

static int failures = 0;
static void check(const char* label, bool ok)
{
    failures += !ok;
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", label);
}
template <class F>
static void rejects(const char* label, F op)
{
    bool ok = false;
    try { op(); } catch (const std::invalid_argument&) { ok = true; }
    check(label, ok);
}
static bool same(const std::vector<POI2D>& a, const std::vector<POI2D>& b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].deformation.u != b[i].deformation.u || a[i].deformation.v != b[i].deformation.v ||
            a[i].result.zncc != b[i].result.zncc)
            return false;
    return true;
}

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int W = 192, H = 128;
    Image2D ref(W, H), tar(W, H);
    unsigned state = 12345;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            state = 1664525u * state + 1013904223u;
            ref.eg_mat(y, x) = float(state >> 24);
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) tar.eg_mat(y, x) = (x >= 2 && y + 1 < H) ? ref.eg_mat(y + 1, x - 2) : 127.f;

    for (auto r : std::vector<std::pair<int, int>>{ { 7, 15 }, { 15, 7 }, { 0, 0 }, { -1, -1 }, { 23171, 23171 } })
        rejects("invalid-or-rectangular-radius", [&] { FFTCC2D solver(r.first, r.second, 2); });
    rejects("invalid-thread-count", [] { FFTCC2D solver(7, 7, 0); });
    rejects("icgn-invalid-thread-count", [] { ICGN2D1 solver(7, 7, 0.001f, 10, 0); });

    std::vector<POI2D> layout;
    for (int y = 32; y < 96; y += 16)
        for (int x = 32; x < 160; x += 16) layout.emplace_back(x, y);
    auto fresh = [] { return std::make_unique<FFTCC2D>(7, 7, 2); };
    auto solver = fresh();
    solver->setImages(ref, tar);
    solver->prepareRef(layout);
    check("prepared-original-layout", solver->hasPreparedReference(ref));
    solver->setRefImage(ref);
    auto different = layout;
    different[0].x += 1;
    auto expected = different;
    auto uncached = fresh();
    uncached->setImages(ref, tar);
    uncached->compute(expected);
    solver->compute(different);
    check("mismatched-queue-does-not-rebuild-old-cache", !solver->hasPreparedReference(ref));
    check("mismatched-queue-matches-fresh", same(different, expected));
    auto matching = layout;
    solver->compute(matching);
    check("matching-queue-rebuilds-retained-layout", solver->hasPreparedReference(ref));
    int correct = 0;
    for (const auto& p : matching) correct += p.deformation.u == 2 && p.deformation.v == -1 && p.result.zncc > .5f;
    check("square-translation-truth", correct == (int)layout.size());
    solver->setRefImage(ref);
    auto indexed = layout;
    rejects("indexed-layout-mismatch-rejected",
            [&] { solver->compute(indexed, std::vector<size_t>(layout.size(), 0)); });
    {
        FFTCC2D s(7, 7, 2);
        DIC& base = s;
        base.setSubset(15, 7);
        rejects("base-subset-mutation", [&] { s.prepareRef(); });
        base.setSubset(15, 15);
        rejects("square-subset-resize", [&] { s.prepareRef(); });
    }

    // opencorr_cli.cpp:117-133
    const int subset_radius = 7, num_threads = 4;
    std::vector<POI2D> poi_queue = layout;
    FFTCC2D* fftcc = new FFTCC2D(subset_radius, subset_radius, num_threads);
    fftcc->setImages(ref, tar);
    fftcc->compute(poi_queue);
    int max_iteration = 10;
    float max_deformation_norm = 0.001f;
    ICGN2D1* icgn = new ICGN2D1(subset_radius, subset_radius, max_deformation_norm, max_iteration, num_threads);
    icgn->setImages(ref, tar);
    icgn->prepare();
    icgn->compute(poi_queue);
    int close = 0;
    for (const auto& p : poi_queue)
        close += std::fabs(p.deformation.u - 2.f) < 0.01f && std::fabs(p.deformation.v + 1.f) < 0.01f;
    check("cli-fftcc-icgn-integer-shift", close == (int)poi_queue.size());
    // DIC::setSubset / setSelfAdaptive through the base reach the device
    DIC& base = *icgn;
    base.setSubset(9, 9);
    auto q9 = layout;
    for (auto& p : q9) { p.deformation.u = 2.f; p.deformation.v = -1.f; }
    icgn->compute(q9);
    close = 0;
    for (const auto& p : q9) close += std::fabs(p.deformation.u - 2.f) < 0.01f && p.result.zncc > 0.9f;
    check("icgn-base-setsubset", close == (int)q9.size());
    base.setSelfAdaptive(true);
    auto qa = layout;
    for (size_t i = 0; i < qa.size(); i++) {
        qa[i].subset_radius.x = qa[i].subset_radius.y = (float)(5 + 2 * (i % 3));
        qa[i].deformation.u = 2.f;
        qa[i].deformation.v = -1.f;
    }
    icgn->compute(qa);
    close = 0;
    for (const auto& p : qa) close += std::fabs(p.deformation.v + 1.f) < 0.01f && p.result.zncc > 0.9f;
    check("icgn-base-self-adaptive", close == (int)qa.size());
    delete fftcc;
    delete icgn;

    // Image2D(file): a 16-bit PGM holding 12-bit data, loaded as the reference
    // loads it by default (8-bit: the high byte) and with options
    {
        const std::string path = dir + "/probe16.pgm";
        std::FILE* f = std::fopen(path.c_str(), "wb");
        std::fprintf(f, "P5\n# 12-bit speckle\n%d %d\n65535\n", W, H);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const unsigned v = (unsigned)ref.eg_mat(y, x) * 16u;  // 0..4080
                std::fputc((int)(v >> 8), f);
                std::fputc((int)(v & 255u), f);
            }
        std::fclose(f);
        Image2D def(path);
        check("pgm-default-load-is-8bit", def.pixel == FI_PIXEL_U8 && def.width == W && def.height == H &&
                                               def.eg_mat(5, 7) == (float)(((unsigned)ref.eg_mat(5, 7) * 16u) >> 8));
        Image2D img(W, H);
        ImageLoadOptions opt;
        opt.depth = ImageInputDepthMode::Depth12;
        opt.alignment = ImageInputAlignment::Right;
        ImageLoadInfo info;
        img.loadWithOptions(path, opt, &info);
        check("pgm-12bit-right", info.effective_depth == ImageInputDepthMode::Depth12 &&
                                     info.effective_alignment == ImageInputAlignment::Right && info.source_max <= 4095 &&
                                     img.eg_mat(3, 4) == ref.eg_mat(3, 4) * 16.f);
        opt.depth = ImageInputDepthMode::Auto;
        opt.normalize = ImageInputNormalize::Unit;
        img.loadWithOptions(path, opt, &info);
        check("pgm-auto-unit", info.effective_depth == ImageInputDepthMode::Depth12 &&
                                   std::fabs(img.eg_mat(3, 4) - ref.eg_mat(3, 4) * 16.f / 4095.f) < 1e-6f);
        bool thrown = false;
        try { Image2D missing(dir + "/no-such-file.pgm"); } catch (const std::string&) { thrown = true; }
        check("missing-file-throws-string", thrown);
        // the loaded image through the solvers: same peaks as the float image
        Image2D r12(W, H);
        opt.depth = ImageInputDepthMode::Depth12;
        opt.normalize = ImageInputNormalize::None;
        r12.loadWithOptions(path, opt);
        auto q12 = layout;
        FFTCC2D f12(7, 7, 1);
        f12.setImages(r12, tar);
        f12.compute(q12);
        correct = 0;
        for (const auto& p : q12) correct += p.deformation.u == 2 && p.deformation.v == -1 && p.result.zncc > .5f;
        check("loaded-image-through-fftcc", correct == (int)layout.size());
    }

    // multi-frame driver and learned seeder (namespace fasticgn)
    fasticgn::FrameDriver drv(7);
    drv.setReference(ref);
    drv.setPois(layout);
    drv.setGraph(true);
    std::vector<POI2D> res;
    const fi_frame_stats s0 = drv.process(tar, res);
    const fi_frame_stats s1 = drv.process(tar, res);
    close = 0;
    for (const auto& p : res) close += std::fabs(p.deformation.u - 2.f) < 0.01f && std::fabs(p.deformation.v + 1.f) < 0.01f;
    check("frame-driver-sequence", close == (int)layout.size() && s0.skip_fftcc_used == 0 && s1.skip_fftcc_used == 1);
    fasticgn::LearnedSeeder seeder;
    const std::vector<float> params = seeder.params();
    check("seeder-param-count", params.size() == fasticgn::LearnedSeeder::paramCount() && params.size() == 44534);
    drv.setSeeder(&seeder);
    drv.resetSequence();
    const fi_frame_stats s2 = drv.process(tar, res, FI_POLICY_LEARNED_SEED);
    close = 0;
    for (const auto& p : res) close += std::fabs(p.deformation.u - 2.f) < 0.01f && std::fabs(p.deformation.v + 1.f) < 0.01f;
    check("frame-driver-learned-seed", close == (int)layout.size() && s2.seed_pois == (int64_t)layout.size());
    drv.setSeedPasses(2);
    drv.resetSequence();
    const fi_frame_stats s3 = drv.process(tar, res, FI_POLICY_LEARNED_SEED);
    close = 0;
    for (const auto& p : res) close += std::fabs(p.deformation.u - 2.f) < 0.01f && std::fabs(p.deformation.v + 1.f) < 0.01f;
    check("frame-driver-two-seed-passes", close == (int)layout.size() && s3.seed_pois == (int64_t)layout.size());
    rejects("seed-passes-range", [&] { drv.setSeedPasses(5); });
    drv.setSeeder(nullptr);
    std::printf("library boundary failures=%d\n", failures);
    return failures ? 1 : 0;
}
