// Learned-seeder engine (fi_seeder.cpp) behind the C ABI.
#pragma once

#include "fi_host.h"
#include "fi_seeder.h"

struct fi_seeder {
    static constexpr int kSplits = 16;  // split-K factor of the weight-gradient GEMMs
    int device = 0;
    hipStream_t stream = nullptr;
    fi::DevArray<float> params, grads, adam_m, adam_v;
    int adam_t = 0;
    fi::DevArray<uint16_t> wt[fi::kSeedLayers], frag;
    fi::DevArray<float> w5f;
    int cus = 0;
    // training set
    fi::DevArray<uint16_t> data_x;
    fi::DevArray<float> data_t;
    size_t data_n = 0;
    // batch workspace
    int batch_cap = 0;
    fi::DevArray<uint16_t> x[fi::kSeedLayers], dz[fi::kSeedLayers], x_last;
    fi::DevArray<float> z[fi::kSeedLayers], zb[fi::kSeedLayers], dz32[fi::kSeedLayers], dx[fi::kSeedLayers];
    fi::DevArray<float> tgt, dout, dh4, loss_parts, gemm_parts;
    fi::DevArray<uint32_t> idx;

    explicit fi_seeder(int dev);
    ~fi_seeder();
    void set_params(const float* host, size_t n);
    void get_params(float* host, size_t n);
    void set_training_data(const uint16_t* x1, const float* targets, size_t n);
    void train_epoch(int bsz, float lr, uint64_t seed, float* mean_loss);
    void forward(const uint16_t* x1, size_t n, float* out);
    void gradient(const uint16_t* x1, const float* targets, size_t n, float* g, float* loss);
    void loss_and_grad(int bsz, float* loss_parts_out);
    // radius > 0: the neighbourhood consensus of the predictions follows
    // (launch_seed_consensus, neighbours within +-radius agreeing within tau);
    // passes > 1: further passes, each cropping the target at the previous
    // seeds and followed by the consensus over the first pass's ok, as the
    // frame driver runs them
    // sample (FI_SEED_SAMPLE_*): with the consensus, the POIs the network
    // seeds; the others take the consensus of their neighbours
    void seed(const float* ref, const float* tar, int w, int h, fi_poi* pois, size_t n, uint32_t* ok,
              float radius = 0.f, float tau = 0.f, int passes = 1, int sample = 0);
    // device-side seeding for the frame driver (stream-ordered, no sync)
    void run_seed(const float* ref, const float* tar, int w, int h, fi_poi* pois, const uint32_t* index, int n,
                  uint32_t* ok, hipStream_t s) const;
    fi::SeedWeights weights() const;

private:
    void refresh_bf16();
    void reserve_batch(int bsz);
    void forward_batch(int bsz);
};
