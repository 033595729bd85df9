// Engines behind the C ABI.  Each mirrors one reference class's lifecycle:
//   FftccEngine  <- opencorr::FFTCC2D       src/oc_fftcc.h:56-100
//   IcgnEngine   <- opencorr::ICGN2D1       src/oc_icgn.h:45-77
//   FrameEngine  <- the fixed-reference sequence driver,
//                   opencorr_cli_cuda_mvp.cpp:3024-3210 and :3834-4509
#pragma once

#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fi_host.h"

namespace fi {
// Validates an explicit IC-GN lane-to-POI shape (0, 0 = default).
void check_shape(int waves, int spl);
}

struct fi_fftcc {
    fi::RecycleScope recycle_;  // first member: see fi_frame
    int device = 0;
    int r = 0;
    hipStream_t stream = nullptr;
    fi::DeviceImage ref, tar;
    fi::DevArray<float> ref_f, tar_f;  // f32 copies the kernels read
    // prepared reference layout (prepared_ref_, src/oc_fftcc.h:63-70)
    std::vector<int32_t> lay_x, lay_y;
    bool ready = false;
    fi::DevArray<float> spectra;
    fi::DevArray<fi::FftccSlot> slots;
    fi::DevArray<float> tw;
    fi::DevArray<fi_poi> d_pois;
    fi::DevArray<uint32_t> d_slot;
    fi::PinnedArray<fi_poi> h_pois;

    fi_fftcc(int rx, int ry, int dev);
    ~fi_fftcc();
    void set_ref(const void* px, int w, int h, size_t stride, int pixel, int mode);
    void set_tar(const void* px, int w, int h, size_t stride, int pixel, int mode);
    void prepare_ref_clear();
    void prepare_ref(const fi_poi* layout, size_t n);
    void compute(fi_poi* pois, size_t n);
    void compute_indexed(fi_poi* pois, size_t n, const size_t* idx);

private:
    void build_cache(const fi_poi* layout_dev, size_t n);
    void ensure_cache();
    bool layout_matches(const fi_poi* pois, size_t n) const;
    void run(fi_poi* pois, size_t n, int mode, const uint32_t* slot_host);
    void check_sizes() const;
};

struct fi_icgn {
    fi::RecycleScope recycle_;  // first member: see fi_frame
    int device = 0;
    int rx = 0, ry = 0;
    float conv = 0.001f, stop = 10.f;
    bool self_adaptive = false;
    int shape_waves = 0, shape_spl = 0;
    int sampling = FI_SAMPLING_AUTO;
    hipStream_t stream = nullptr;
    fi::DeviceImage ref_raw, tar_raw;
    fi::DevArray<float> ref_f, gx, gy, coef, tar_f;
    bool grads_ready = false, coef_ready = false;
    uint64_t ref_version = 0;
    // per-POI reference records, keyed by (x, y, rx, ry) per POI
    std::vector<float> prep_key;
    uint64_t prep_version = ~0ull;
    fi::DevArray<fi::PoiPrep> prep;
    fi::DevArray<fi_poi> d_pois;
    fi::DevArray<uint32_t> d_index;
    fi::DevArray<float> d_offset;  // (x, y) subset centre offsets of the last compute
    fi::PinnedArray<fi_poi> h_pois;

    fi_icgn(int rx, int ry, float conv, float stop, int dev);
    ~fi_icgn();
    void set_ref(const void* px, int w, int h, size_t stride, int pixel, int mode);
    void set_tar(const void* px, int w, int h, size_t stride, int pixel, int mode);
    void prepare_ref();
    void prepare_tar();
    // offsets: n (x, y) pairs (compute(POI2D*, Point2D&)), or nullptr
    void compute(fi_poi* pois, size_t n, const float* offsets = nullptr);
    size_t prep_n = 0;  // records of the last compute
    // compute on POIs in device memory with the records in the library's
    // static scratch: no allocation (fi_icgn_compute_noalloc)
    void compute_noalloc(fi_poi* d_pois, size_t n);
    // Copies a device-side array out (fi_icgn_export).
    void export_data(int what, void* out, size_t bytes);
};

// One radius class of the frame driver's template: POIs [off, off + n) of the
// radius-sorted device layout, with their FFTCC state and IC-GN launch
// configuration (lane shape, sampling, persistent queue).
struct FrameClass {
    int r = 0;
    size_t off = 0, n = 0;
    int waves = 0, spl = 0;  // IC-GN lane shape (0, 0 = the driver's)
    int sampling = -1;       // -1 = the driver's sampling
    int queue_blocks = -1;   // persistent work-queue blocks (0 = one block per POI, -1 = the driver's)
    fi::DevArray<float> tw, spectra;
    fi::DevArray<fi::FftccSlot> slots;
};

struct fi_frame {
    // first member, destroyed last: armed by the destructor once the streams
    // are idle, so the buffers and streams go back to the block pool
    fi::RecycleScope recycle_;
    int device = 0;
    int r = 0;                     // subset radius of every POI unless self_adaptive
    bool self_adaptive = false;    // per-POI subset_radius (square) instead of r
    float conv = 0.001f, stop = 10.f;
    int shape_waves = 0, shape_spl = 0;
    int sampling = FI_SAMPLING_AUTO;
    bool use_graph = false;        // capture each frame's device work as a hipGraph
    int persistent_blocks = 0;     // IC-GN persistent work-queue blocks (0 = one block per POI)
    // One IC-GN launch over every radius class (per-POI radius, a shape that
    // covers all of them) instead of one launch per class
    bool fuse = false;
    FrameClass fused;              // its launch configuration (r = largest radius)
    fi_seeder* seeder = nullptr;   // learned seeder (FI_POLICY_LEARNED_SEED)
    float seed_threshold = 0.8f;   // ZNCC below which a learned seed falls back to FFTCC
    int seed_passes = 1;           // seeder passes, each re-centred on the previous one's rounded u, v
    // Neighbourhood consensus of the seeder's predictions after each pass
    // (launch_seed_consensus) within +-r, neighbours agreeing within
    // consensus_tau px; its cells are built on the host the first time a
    // learned frame runs on a layout.
    bool consensus = true;
    float consensus_tau = 0.5f;
    int seed_sample = FI_SEED_SAMPLE_CHECKER;  // the POIs the network seeds (FI_SEED_SAMPLE_*)
    bool cells_ready = false;
    fi::DevSeedCells cells;
    fi::DevArray<float4> seed_raw;
    hipStream_t stream = nullptr;
    fi::DeviceImage ref_raw;
    // Targets: two slots filled by a staging thread (host copy into pinned
    // memory, DMA on copy_stream) while frames run; run() takes the oldest
    // uploaded slot, or re-runs the last one when none is queued.
    struct TargetSlot {
        fi::DeviceImage img;
        hipEvent_t ready = nullptr;  // DMA done (copy_stream)
        hipEvent_t freed = nullptr;  // the last frame that took the slot finished (compute stream)
        int state = 0;               // 0 idle, 1 staging, 2 staged
        int err_code = 0;
        std::string err;
    };
    TargetSlot tslot[2];
    std::deque<int> pending;
    int next_slot = 0, last_slot = -1;
    hipStream_t copy_stream = nullptr;
    hipStream_t aux_stream = nullptr;      // B-spline table beside the seeding
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    hipEvent_t tmpl_ev = nullptr;  // template upload done (copy_stream)
    std::thread stager;
    std::mutex stage_mu;
    std::condition_variable stage_cv;
    struct StageJob {
        int slot;
        const void* px;
        int w, h;
        size_t stride;
        int pixel, mode;
    };
    std::deque<StageJob> stage_jobs;
    bool stage_stop = false;
    fi::DevArray<float> ref_f, gx, gy, coef, tar_f;
    fi::DevArray<fi::PoiPrep> prep;
    fi::PinnedArray<fi_poi> tmpl_host;    // template in the caller's order (n POIs)
    std::vector<uint32_t> order;          // layout position -> template index (empty: identity)
    std::vector<std::unique_ptr<FrameClass>> classes;
    // cur: the frame being run; prev, prev2: the two frames before it
    // (rotated after each frame); save: the skip attempt before repairs
    fi::DevArray<fi_poi> tmpl, cur, prev, prev2, save, repair;
    fi::DevArray<uint32_t> cand, seed_ok;
    // counters[c] for the scans and repairs of class c, counters[classes] for QA
    fi::DevArray<fi::FrameCounters> counters;
    fi::PinnedArray<fi::FrameCounters> h_counters;
    fi::DevArray<uint32_t> iters;  // executed iterations per IC-GN launch entry, 3 x n (see icgn_pass)
    fi::DevArray<int> queue;       // persistent-queue claim counters, kQueueWords per class
    fi::PinnedArray<fi_poi> h_out;
    fi::DevArray<uint8_t> redo;    // per POI: re-run from FFTCC in the last frame
    fi::DevArray<unsigned> class_n;  // class sizes (k_decide)
    fi::DevArray<float> threshold;   // baseline bad ratio + margin (k_decide)
    fi::DevArray<uint32_t> d_order;  // layout position -> template index, on the device
    // Two-frame temporal extrapolation of the seeds (seedPoiQueueFromPrevious
    // with --temporal-extrapolation), steps up to temporal_max_delta px.
    bool temporal = false;
    float temporal_max_delta = 0.f;
    // Raw records of each frame (36 B, template order): packed on the device
    // after the frame and copied to pinned memory on rec_stream, two slots.
    bool records_on = false;
    fi::DevArray<uint8_t> d_rec[2];
    fi::PinnedArray<uint8_t> h_rec[2];
    hipEvent_t rec_ev[2] = {}, pack_ev = nullptr;
    hipStream_t rec_stream = nullptr;
    int rec_next = 0;
    int rec_held[2] = { -1, -1 };    // frame index whose records each slot holds
    size_t n = 0;
    bool ref_set = false, pois_set = false, prepared = false, tar_ready = false;
    bool have_prev = false, have_prev2 = false, baseline_ready = false;
    float baseline_bad = NAN;
    float threshold_host = NAN;      // the value last written to `threshold`
    fi_prepare_stats last_prepare{}; // timing of the last reference-side preparation
    int frame_index = 0;
    hipEvent_t ev[8] = {};
    // Frames in flight (fi_frame_run_async): at most two, each with the set
    // of per-frame buffers it was launched on (scratch_set, the members
    // below or their *_alt twins, exchanged by swap_scratch), so a frame
    // launched behind another leaves the first one's candidates, counters,
    // iteration counts and converted target intact for its completion.
    struct InFlight {
        fi_frame_stats st;
        unsigned policy;
        int mode, rs, ts, set, index;
        bool collect, skip, table, graph;
        fi_poi* out;  // the frame's results
    };
    std::deque<InFlight> inflight;
    std::deque<fi_frame_stats> done;  // completed, not yet returned by wait()
    int scratch_set = 0;
    hipEvent_t fev[2][2] = {};  // per buffer set: graph launch start, end
    fi::DevArray<fi_poi> repair_alt;
    fi::DevArray<uint32_t> cand_alt, seed_ok_alt, iters_alt;
    fi::DevArray<uint8_t> redo_alt;
    fi::DevArray<fi::FrameCounters> counters_alt;
    fi::PinnedArray<fi::FrameCounters> h_counters_alt;
    fi::DevArray<float> tar_f_alt;
    struct Graph {
        uint64_t key[12];
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;     // instantiated frame graphs (cleared on any change)

    fi_frame(int radius, float conv, float stop, int dev);
    ~fi_frame();
    void set_ref(const void* px, int w, int h, size_t stride, int pixel, int mode);
    void set_pois(const fi_poi* pois, size_t n);
    void set_self_adaptive(bool on);
    void upload_target(const void* px, int w, int h, size_t stride, int pixel, int mode);
    void run(unsigned policy, fi_frame_stats* stats);
    // run() split in two: launch a frame (behind one still running when both
    // are independent frames replayed as graphs) and return; wait() returns
    // the oldest launched frame's statistics once it has finished.
    void run_async(unsigned policy);
    void wait(fi_frame_stats* stats);
    // completes every frame in flight (their statistics stay for wait())
    void drain();
    void download(fi_poi* out, size_t n);
    // Raw records of the frame `age` runs back (0: the last) in pinned host
    // memory, valid until the second run after that frame; the last frame's
    // are packed and copied now if its run did not queue them.
    void records(int age, const void** out, size_t* count);
    void set_records(bool on);
    void download_redo(uint8_t* out, size_t n);
    void reset_sequence();
    void invalidate_graphs();
    // Prepares the reference side now if it is not prepared; the timing of
    // the last preparation.
    fi_prepare_stats prepare_now();
    // Times the IC-GN lane shapes, sampling modes and launch modes on the
    // uploaded target per radius class and keeps the fastest; returns the
    // number of configurations timed.
    int autotune(fi_tune_entry* report, size_t report_len, size_t* report_n);

private:
    void prepare();
    void launch(unsigned policy);
    void complete_oldest();
    void finish(const InFlight& f);
    void swap_scratch();
    void queue_records(const fi_poi* src);
    int records_slot();
    void pack_records(const fi_poi* src, int s);
    void copy_records(int s, int index);
    void fftcc_pass(const FrameClass& c, fi_poi* queue, int count, const uint32_t* slot_dev,
                    const unsigned* count_dev = nullptr);
    // region: which third of `iters` the launch writes its per-entry
    // iteration counts to (0 first pass, 1 POI-level repair, 2 whole-frame
    // fallback), from entry `first` of the class on (a repair's candidates
    // past repair_cap follow the ones repaired inside the frame); k_qa sums
    // all three
    void icgn_pass(const FrameClass& c, fi_poi* queue, int count, const uint32_t* prep_index,
                   const unsigned* count_dev = nullptr, int region = 0, int first = 0);
    void icgn_fused(fi_poi* base, int region = 0);
    bool fused_shape(int& waves, int& spl) const;
    void icgn_all(fi_poi* base, int region = 0);
    void fftcc_all(fi_poi* base);
    // seeded: the learned seeder's flags (null for temporal repairs), see
    // launch_merge_repair
    void repair_class(fi_poi* out, const FrameClass& c, int m, int lo = 0, const unsigned* count_dev = nullptr,
                      const uint32_t* seeded = nullptr);
    int repair_cap(const FrameClass& c) const;
    void prep_target(bool table);
    bool table_needed() const;
    fi::FrameCounters read_counters(fi::FrameCounters* per_class);
    void zero_counters();
    size_t nclass() const { return classes.size(); }
    void stage_loop();
    int take_target(bool consume);   // slot of the next frame's target, staged
    const fi::DevImage& target(int slot) const { return tslot[slot].img.view; }
};

