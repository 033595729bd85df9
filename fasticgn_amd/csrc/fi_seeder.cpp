// Learned-seeder engine behind the C ABI: parameters (FP32 master copy, Adam
// moments, BF16 transposed copies for the MFMA kernels), the device-resident
// training set, and one training step = forward, MSE loss, backward, Adam,
// all on the gfx950 kernels of k_seeder.hip.
#include "fi_seeder_engine.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <random>

using namespace fi;

namespace fi {
void launch_gemm_bf16(const uint16_t* A, long sai, long sak, const uint16_t* B, long sbk, long sbj, float* C, int M,
                      int N, int K, int splits, float* parts, hipStream_t s);
void launch_bias_relu_patchify(const float* Z, const float* b, int rows, int n, int g, float* zb, uint16_t* xn,
                               hipStream_t s);
void launch_relu_bwd(const float* dxn, const float* zb, int rows, int n, int g, uint16_t* dz, float* dz32,
                     hipStream_t s);
void launch_colsum(const float* x, int rows, int n, float* out, hipStream_t s);
void launch_head_fwd(const float* zb4, const float* w5, const float* b5, const float* t, int bsz, float* dout,
                     float* loss_parts, hipStream_t s);
void launch_head_bwd(const float* zb4, const float* dout, const float* w5, int bsz, float* gw5, float* gb5,
                     float* dh4, hipStream_t s);
void launch_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float c1,
                 float c2, float eps, hipStream_t s);
void launch_transpose_bf16(const float* w, int K, int N, uint16_t* wt, hipStream_t s);
void launch_gather_rows16(const uint16_t* src, const uint32_t* idx, int rows, int width, uint16_t* dst,
                          hipStream_t s);
void launch_gather_rows32(const float* src, const uint32_t* idx, int rows, int width, float* dst, hipStream_t s);
}  // namespace fi

namespace {
constexpr int kWidth1 = 32;  // layer-1 row width
constexpr int kGrid[kSeedLayers] = { 8, 4, 2, 1 };

size_t offset_w(int l)
{
    size_t o = 0;
    for (int i = 0; i < l; i++) o += (size_t)kSeedK[i] * kSeedN[i] + kSeedN[i];
    return o;
}
size_t offset_b(int l) { return offset_w(l) + (size_t)kSeedK[l] * kSeedN[l]; }
size_t offset_w5() { return offset_w(kSeedLayers); }
size_t offset_b5() { return offset_w5() + 128 * kSeedOut; }
}  // namespace

fi_seeder::fi_seeder(int dev) : device(dev)
{
    DeviceGuard g(device);
    FI_HIP(hipStreamCreate(&stream));
    const size_t np = seed_param_count();
    params.reserve(np);
    grads.reserve(np);
    adam_m.reserve(np);
    adam_v.reserve(np);
    for (int l = 0; l < kSeedLayers; l++) wt[l].reserve((size_t)kSeedK[l] * kSeedN[l]);
    frag.reserve(seed_frag_elems());
    w5f.reserve(seed_w5f_elems());
    FI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    // He-style initialisation, seeded (a fresh seeder is usable but untrained)
    std::vector<float> host(np, 0.f);
    std::mt19937 rng(20260);
    for (int l = 0; l <= kSeedLayers; l++) {
        const int K = l < kSeedLayers ? kSeedK[l] : 128, N = l < kSeedLayers ? kSeedN[l] : kSeedOut;
        std::normal_distribution<float> d(0.f, std::sqrt(2.f / (float)K) * (l < kSeedLayers ? 1.f : 0.1f));
        const size_t o = l < kSeedLayers ? offset_w(l) : offset_w5();
        for (size_t i = 0; i < (size_t)K * N; i++) host[o + i] = d(rng);
    }
    set_params(host.data(), np);
}

fi_seeder::~fi_seeder()
{
    if (stream) (void)hipStreamDestroy(stream);
}

void fi_seeder::refresh_bf16()
{
    const float* w[kSeedLayers];
    for (int l = 0; l < kSeedLayers; l++) {
        launch_transpose_bf16(params.p + offset_w(l), kSeedK[l], kSeedN[l], wt[l].p, stream);
        w[l] = params.p + offset_w(l);
    }
    launch_pack_infer(w, params.p + offset_w5(), frag.p, w5f.p, stream);
    FI_HIP(hipGetLastError());
}

SeedWeights fi_seeder::weights() const
{
    SeedWeights w{};
    for (int l = 0; l < kSeedLayers; l++) {
        w.wt[l] = wt[l].p;
        w.b[l] = params.p + offset_b(l);
    }
    w.frag = frag.p;
    w.w5f = w5f.p;
    w.w5 = params.p + offset_w5();
    w.b5 = params.p + offset_b5();
    return w;
}

void fi_seeder::set_params(const float* host, size_t n)
{
    if (n != seed_param_count())
        fail(FI_E_INVALID_ARGUMENT, "seeder expects " + std::to_string(seed_param_count()) + " parameters, got " +
                                        std::to_string(n));
    DeviceGuard g(device);
    FI_HIP(hipMemcpyAsync(params.p, host, n * sizeof(float), hipMemcpyHostToDevice, stream));
    FI_HIP(hipMemsetAsync(adam_m.p, 0, n * sizeof(float), stream));
    FI_HIP(hipMemsetAsync(adam_v.p, 0, n * sizeof(float), stream));
    adam_t = 0;
    refresh_bf16();
    FI_HIP(hipStreamSynchronize(stream));
}

void fi_seeder::get_params(float* host, size_t n)
{
    if (n != seed_param_count()) fail(FI_E_INVALID_ARGUMENT, "parameter buffer size mismatch");
    DeviceGuard g(device);
    FI_HIP(hipMemcpyAsync(host, params.p, n * sizeof(float), hipMemcpyDeviceToHost, stream));
    FI_HIP(hipStreamSynchronize(stream));
}

void fi_seeder::set_training_data(const uint16_t* x1, const float* targets, size_t n)
{
    if (n == 0 || x1 == nullptr || targets == nullptr) fail(FI_E_INVALID_ARGUMENT, "empty training set");
    DeviceGuard g(device);
    data_x.reserve(n * 64 * kWidth1);
    data_t.reserve(n * kSeedOut);
    FI_HIP(hipMemcpyAsync(data_x.p, x1, n * 64 * kWidth1 * sizeof(uint16_t), hipMemcpyHostToDevice, stream));
    FI_HIP(hipMemcpyAsync(data_t.p, targets, n * kSeedOut * sizeof(float), hipMemcpyHostToDevice, stream));
    FI_HIP(hipStreamSynchronize(stream));
    data_n = n;
}

void fi_seeder::reserve_batch(int bsz)
{
    if (bsz <= batch_cap) return;
    for (int l = 0; l < kSeedLayers; l++) {
        const size_t rows = (size_t)bsz * kSeedRows[l];
        x[l].reserve(rows * kSeedK[l]);
        z[l].reserve(rows * kSeedN[l]);
        zb[l].reserve(rows * kSeedN[l]);
        dz[l].reserve(rows * kSeedN[l]);
        dz32[l].reserve(rows * kSeedN[l]);
        dx[l].reserve(rows * kSeedK[l]);
    }
    x_last.reserve((size_t)bsz * 128);
    tgt.reserve((size_t)bsz * kSeedOut);
    dout.reserve((size_t)bsz * kSeedOut);
    dh4.reserve((size_t)bsz * 128);
    loss_parts.reserve((size_t)(bsz + 255) / 256);
    idx.reserve(bsz);
    // split-K partial sums for the weight gradients (K = rows up to bsz * 64)
    size_t parts = 0;
    for (int l = 0; l < kSeedLayers; l++) parts = std::max(parts, (size_t)kSplits * kSeedK[l] * kSeedN[l]);
    gemm_parts.reserve(parts);
    batch_cap = bsz;
}

// forward of the current batch x[0] (bsz samples) through layer 4 (zb[3] holds
// layer 4's pre-activation)
void fi_seeder::forward_batch(int bsz)
{
    for (int l = 0; l < kSeedLayers; l++) {
        const int rows = bsz * kSeedRows[l], K = kSeedK[l], N = kSeedN[l];
        // Z = X W: A(i,k) = X[i K + k]; B(k,n) = Wt[n K + k]
        launch_gemm_bf16(x[l].p, K, 1, wt[l].p, 1, K, z[l].p, rows, N, K, 1, nullptr, stream);
        uint16_t* next = l + 1 < kSeedLayers ? x[l + 1].p : x_last.p;
        launch_bias_relu_patchify(z[l].p, params.p + offset_b(l), rows, N, kGrid[l], zb[l].p, next, stream);
    }
}

void fi_seeder::train_epoch(int bsz, float lr, uint64_t seed, float* mean_loss)
{
    if (data_n == 0) fail(FI_E_LOGIC, "seeder training data not set");
    if (bsz <= 0 || bsz % 16 != 0) fail(FI_E_INVALID_ARGUMENT, "batch must be a positive multiple of 16");
    if ((size_t)bsz > data_n) fail(FI_E_INVALID_ARGUMENT, "batch larger than the training set");
    if (((size_t)bsz * 64) % kSplits != 0) fail(FI_E_INVALID_ARGUMENT, "batch incompatible with the K split");
    DeviceGuard g(device);
    reserve_batch(bsz);
    std::vector<uint32_t> order(data_n);
    std::iota(order.begin(), order.end(), 0u);
    std::mt19937_64 rng(seed);
    std::shuffle(order.begin(), order.end(), rng);
    const size_t steps = data_n / (size_t)bsz;
    const size_t np = seed_param_count();
    PinnedArray<float> h_loss;
    const int nloss = (bsz + 255) / 256;
    h_loss.reserve((size_t)nloss * steps);
    DevArray<float> d_loss;
    d_loss.reserve((size_t)nloss * steps);
    DevArray<uint32_t> d_order;
    d_order.reserve(data_n);
    FI_HIP(hipMemcpyAsync(d_order.p, order.data(), data_n * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (size_t st = 0; st < steps; st++) {
        const uint32_t* bidx = d_order.p + st * (size_t)bsz;
        // gather the batch: 64 rows of 32 per sample
        launch_gather_rows16(data_x.p, bidx, bsz, 64 * kWidth1, x[0].p, stream);
        launch_gather_rows32(data_t.p, bidx, bsz, kSeedOut, tgt.p, stream);
        loss_and_grad(bsz, d_loss.p + st * nloss);
        adam_t++;
        const float c1 = 1.f / (1.f - std::pow(b1, (float)adam_t));
        const float c2 = 1.f / (1.f - std::pow(b2, (float)adam_t));
        launch_adam(params.p, grads.p, adam_m.p, adam_v.p, np, lr, b1, b2, c1, c2, eps, stream);
        refresh_bf16();
    }
    FI_HIP(hipGetLastError());
    FI_HIP(hipMemcpyAsync(h_loss.p, d_loss.p, (size_t)nloss * steps * sizeof(float), hipMemcpyDeviceToHost, stream));
    FI_HIP(hipStreamSynchronize(stream));
    if (mean_loss) {
        double s = 0.0;
        for (size_t i = 0; i < (size_t)nloss * steps; i++) s += h_loss.p[i];
        *mean_loss = steps ? (float)(s / ((double)steps * bsz * kSeedOut)) : 0.f;
    }
}

// MSE loss partial sums (loss_parts) and the parameter gradient (grads) of
// the batch in x[0] / tgt (bsz samples)
void fi_seeder::loss_and_grad(int bsz, float* loss_parts_out)
{
    forward_batch(bsz);
    launch_head_fwd(zb[3].p, params.p + offset_w5(), params.p + offset_b5(), tgt.p, bsz, dout.p, loss_parts_out,
                    stream);
    launch_head_bwd(zb[3].p, dout.p, params.p + offset_w5(), bsz, grads.p + offset_w5(), grads.p + offset_b5(),
                    dh4.p, stream);
    const float* dnext = dh4.p;
    for (int l = kSeedLayers - 1; l >= 0; l--) {
        const int rows = bsz * kSeedRows[l], K = kSeedK[l], N = kSeedN[l];
        launch_relu_bwd(dnext, zb[l].p, rows, N, kGrid[l], dz[l].p, dz32[l].p, stream);
        // dW = X^T dZ: A(i,k) = X[k K + i]; B(k,n) = dZ[k N + n]; K_gemm = rows
        const int splits = rows >= 16 * kSplits && rows % (16 * kSplits) == 0 ? kSplits : 1;
        launch_gemm_bf16(x[l].p, 1, K, dz[l].p, N, 1, grads.p + offset_w(l), K, N, rows, splits, gemm_parts.p,
                         stream);
        launch_colsum(dz32[l].p, rows, N, grads.p + offset_b(l), stream);
        if (l > 0) {
            // dX = dZ W^T: A(i,n) = dZ[i N + n]; B(n,k) = Wt[n K + k]
            launch_gemm_bf16(dz[l].p, N, 1, wt[l].p, K, 1, dx[l].p, rows, K, N, 1, nullptr, stream);
            dnext = dx[l].p;
        }
    }
}

void fi_seeder::gradient(const uint16_t* x1, const float* targets, size_t n, float* g, float* loss)
{
    if (x1 == nullptr || targets == nullptr || g == nullptr) fail(FI_E_INVALID_ARGUMENT, "null buffer");
    if (n == 0 || n % 16 != 0 || n > (size_t)INT_MAX / 64)
        fail(FI_E_INVALID_ARGUMENT, "batch must be a positive multiple of 16");
    const int bsz = (int)n;
    if (((size_t)bsz * 64) % kSplits != 0) fail(FI_E_INVALID_ARGUMENT, "batch incompatible with the K split");
    DeviceGuard gd(device);
    reserve_batch(bsz);
    const int nloss = (bsz + 255) / 256;
    DevArray<float> d_loss;
    d_loss.reserve((size_t)nloss);
    FI_HIP(hipMemcpyAsync(x[0].p, x1, n * 64 * kWidth1 * sizeof(uint16_t), hipMemcpyHostToDevice, stream));
    FI_HIP(hipMemcpyAsync(tgt.p, targets, n * kSeedOut * sizeof(float), hipMemcpyHostToDevice, stream));
    loss_and_grad(bsz, d_loss.p);
    FI_HIP(hipGetLastError());
    const size_t np = seed_param_count();
    std::vector<float> h_loss((size_t)nloss);
    FI_HIP(hipMemcpyAsync(g, grads.p, np * sizeof(float), hipMemcpyDeviceToHost, stream));
    FI_HIP(hipMemcpyAsync(h_loss.data(), d_loss.p, (size_t)nloss * sizeof(float), hipMemcpyDeviceToHost, stream));
    FI_HIP(hipStreamSynchronize(stream));
    if (loss) {
        double t = 0.0;
        for (float v : h_loss) t += v;
        *loss = (float)(t / ((double)bsz * kSeedOut));
    }
}

void fi_seeder::forward(const uint16_t* x1, size_t n, float* out)
{
    if (n == 0) return;
    if (x1 == nullptr || out == nullptr) fail(FI_E_INVALID_ARGUMENT, "null buffer");
    DeviceGuard g(device);
    const int bsz = 256;
    reserve_batch(bsz);
    DevArray<float> d_out;
    d_out.reserve(n * kSeedOut);
    for (size_t b0 = 0; b0 < n; b0 += bsz) {
        const int cnt = (int)std::min<size_t>(bsz, n - b0);
        const int pad = (cnt + 15) / 16 * 16;
        FI_HIP(hipMemsetAsync(x[0].p, 0, (size_t)pad * 64 * kWidth1 * sizeof(uint16_t), stream));
        FI_HIP(hipMemcpyAsync(x[0].p, x1 + b0 * 64 * kWidth1, (size_t)cnt * 64 * kWidth1 * sizeof(uint16_t),
                              hipMemcpyHostToDevice, stream));
        forward_batch(pad);
        FI_HIP(hipMemsetAsync(tgt.p, 0, (size_t)pad * kSeedOut * sizeof(float), stream));
        launch_head_fwd(zb[3].p, params.p + offset_w5(), params.p + offset_b5(), tgt.p, pad, dout.p, loss_parts.p,
                        stream);
        // dout = 2 (out - 0) / (pad 6): recover out
        FI_HIP(hipMemcpyAsync(d_out.p + b0 * kSeedOut, dout.p, (size_t)cnt * kSeedOut * sizeof(float),
                              hipMemcpyDeviceToDevice, stream));
        FI_HIP(hipGetLastError());
        FI_HIP(hipStreamSynchronize(stream));
        const float scale = (float)(pad * kSeedOut) / 2.f;
        std::vector<float> tmp((size_t)cnt * kSeedOut);
        FI_HIP(hipMemcpy(tmp.data(), d_out.p + b0 * kSeedOut, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < tmp.size(); i++) out[b0 * kSeedOut + i] = tmp[i] * scale;
    }
}

void fi_seeder::seed(const float* ref, const float* tar, int w, int h, fi_poi* pois, size_t n, uint32_t* ok,
                     float radius, float tau, int passes, int sample)
{
    if (n == 0) return;
    if (!ref || !tar || !pois) fail(FI_E_INVALID_ARGUMENT, "null buffer");
    if (w < kSeedCrop || h < kSeedCrop) fail(FI_E_INVALID_ARGUMENT, "images smaller than the seeder crop");
    if (n > (size_t)INT_MAX) fail(FI_E_INVALID_ARGUMENT, "queue too long");
    DeviceGuard g(device);
    DevArray<float> d_ref, d_tar;
    DevArray<fi_poi> d_pois;
    DevArray<uint32_t> d_ok;
    const size_t px = (size_t)w * h;
    d_ref.reserve(px);
    d_tar.reserve(px);
    d_pois.reserve(n);
    d_ok.reserve(n);
    FI_HIP(hipMemcpyAsync(d_ref.p, ref, px * sizeof(float), hipMemcpyHostToDevice, stream));
    FI_HIP(hipMemcpyAsync(d_tar.p, tar, px * sizeof(float), hipMemcpyHostToDevice, stream));
    FI_HIP(hipMemcpyAsync(d_pois.p, pois, n * sizeof(fi_poi), hipMemcpyHostToDevice, stream));
    DevSeedCells cells;
    DevArray<float4> raw;
    if (radius > 0.f) {
        std::vector<float> xs(n), ys(n);
        for (size_t i = 0; i < n; i++) {
            xs[i] = pois[i].x;
            ys[i] = pois[i].y;
        }
        cells.upload(build_seed_cells(xs.data(), ys.data(), n, radius, sample), radius, tau, stream);
        raw.reserve(2 * n);
    }
    // the network on every POI, or on the sample (the consensus seeds the rest)
    auto network = [&](uint32_t* okp) {
        if (!cells.view.sample)
            run_seed(d_ref.p, d_tar.p, w, h, d_pois.p, nullptr, (int)n, okp, stream);
        else if (cells.sampled > 0)
            run_seed(d_ref.p, d_tar.p, w, h, d_pois.p, cells.sample_idx.p, (int)cells.sampled, okp, stream);
    };
    network(d_ok.p);
    if (radius > 0.f) {
        launch_seed_consensus(d_pois.p, d_ok.p, raw.p, (int)n, cells.view, stream);
        for (int pass = 1; pass < passes; pass++) {
            network(nullptr);
            launch_seed_consensus(d_pois.p, d_ok.p, raw.p, (int)n, cells.view, stream);
        }
        FI_HIP(hipGetLastError());
    }
    FI_HIP(hipMemcpyAsync(pois, d_pois.p, n * sizeof(fi_poi), hipMemcpyDeviceToHost, stream));
    std::vector<uint32_t> h_ok(n);
    FI_HIP(hipMemcpyAsync(h_ok.data(), d_ok.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    FI_HIP(hipStreamSynchronize(stream));
    if (ok) std::copy(h_ok.begin(), h_ok.end(), ok);
}

void fi_seeder::run_seed(const float* ref, const float* tar, int w, int h, fi_poi* pois, const uint32_t* index,
                         int n, uint32_t* ok, hipStream_t s) const
{
    SeedInferParams sp{};
    sp.ref = ref;
    sp.tar = tar;
    sp.w = w;
    sp.h = h;
    sp.pois = pois;
    sp.index = index;
    sp.n = n;
    sp.wts = weights();
    sp.ok = ok;
    sp.cus = cus;
    launch_seed_infer(sp, s);
    FI_HIP(hipGetLastError());
}
