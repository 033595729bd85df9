/* The multi-frame C ABI as a C caller drives it (the loop of
 * opencorr_cli_cuda_mvp.cpp:3834-4509): a reference, a POI template, frames
 * uploaded from host memory, FFTCC then skip-FFTCC frames, graph replay, the
 * autotuner, results downloaded and written as a raw frame file, and a
 * learned-seed frame.  The images are the LCG pattern of
 * smoke/library_boundary_probe.cpp:41-62 (target = reference shifted by
 * (+2, -1), fill 127), so every valid POI must come out at u = 2, v = -1.
 * Prints PASS/FAIL lines; exit status 0 when all pass. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fasticgn.h"

#define W 256
#define H 192

static int failures = 0;

static void check(int ok, const char* what)
{
    printf("%s %s\n", ok ? "PASS" : "FAIL", what);
    if (!ok) failures++;
}

static int call(int rc, const char* what)
{
    if (rc != FI_OK) {
        char msg[512];
        fi_last_error(msg, sizeof msg);
        printf("FAIL %s: status %d: %s\n", what, rc, msg);
        failures++;
    }
    return rc;
}

int main(void)
{
    static uint8_t ref[H][W], tar[H][W];
    uint32_t state = 12345u;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            state = 1664525u * state + 1013904223u;
            ref[y][x] = (uint8_t)(state >> 24);
        }
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const int sx = x - 2, sy = y + 1;
            tar[y][x] = (sx >= 0 && sx < W && sy >= 0 && sy < H) ? ref[sy][sx] : 127;
        }

    /* POI template: grid 8 at margin 24, radius 10 */
    enum { R = 10, STEP = 8, M = 24 };
    fi_poi pois[((W - 2 * M) / STEP + 1) * ((H - 2 * M) / STEP + 1)];
    size_t n = 0;
    for (int y = M; y < H - M; y += STEP)
        for (int x = M; x < W - M; x += STEP) {
            memset(&pois[n], 0, sizeof(fi_poi));
            pois[n].x = (float)x;
            pois[n].y = (float)y;
            n++;
        }

    fi_frame* f = NULL;
    if (call(fi_frame_create(R, 0.001f, 10.f, 0, &f), "fi_frame_create")) return 1;
    call(fi_frame_set_ref(f, ref, W, H, W, FI_PIXEL_U8, FI_DEPTH_NATIVE), "set_ref");
    call(fi_frame_set_pois(f, pois, n), "set_pois");
    call(fi_frame_set_graph(f, 1), "set_graph");

    fi_frame_stats qa;
    fi_poi* out = (fi_poi*)calloc(n, sizeof(fi_poi));
    uint8_t* redo = (uint8_t*)calloc(n, 1);
    for (int frame = 0; frame < 3; frame++) {
        call(fi_frame_upload_target(f, tar, W, H, W, FI_PIXEL_U8, FI_DEPTH_NATIVE), "upload_target");
        call(fi_frame_run(f, FI_POLICY_DEFAULT, &qa), "run");
        call(fi_frame_download(f, out, n), "download");
        int good = 0;
        for (size_t i = 0; i < n; i++)
            good += fabsf(out[i].deformation[0] - 2.f) < 0.01f && fabsf(out[i].deformation[6] + 1.f) < 0.01f &&
                    out[i].result[2] > 0.99f;
        char what[128];
        snprintf(what, sizeof what, "frame %d: %d of %zu POIs at (2, -1), skip_fftcc %d", frame, good, n,
                 (int)qa.skip_fftcc_used);
        check(good == (int)n && qa.skip_fftcc_used == (frame > 0), what);
    }
    call(fi_frame_download_redo(f, redo, n), "download_redo");
    size_t redone = 0;
    for (size_t i = 0; i < n; i++) redone += redo[i];
    check(redone == (size_t)qa.poi_fallback_count, "redo flags match the repair count");

    fi_tune_entry rep[128];
    size_t m = 0;
    call(fi_frame_upload_target(f, tar, W, H, W, FI_PIXEL_U8, FI_DEPTH_NATIVE), "upload_target (tune)");
    call(fi_frame_autotune(f, rep, 128, &m), "autotune");
    int chosen = 0;
    for (size_t i = 0; i < m; i++) chosen += rep[i].chosen;
    check(m >= 4 && chosen == 1, "autotune keeps one configuration");

    size_t bytes = 0;
    call(fi_write_raw(out, n, "frame_probe_000000.bin", &bytes), "write_raw");
    check(bytes == 36 * n, "raw frame file: 36 bytes per POI");
    remove("frame_probe_000000.bin");

    /* learned seed (untrained weights): whatever it proposes, the FFTCC
     * fallback and IC-GN must still land every POI on the shift */
    fi_seeder* s = NULL;
    call(fi_seeder_create(0, &s), "seeder_create");
    call(fi_frame_set_seeder(f, s, 0.8f), "set_seeder");
    call(fi_frame_reset_sequence(f), "reset_sequence");
    call(fi_frame_upload_target(f, tar, W, H, W, FI_PIXEL_U8, FI_DEPTH_NATIVE), "upload_target (learned)");
    call(fi_frame_run(f, FI_POLICY_LEARNED_SEED, &qa), "run (learned)");
    call(fi_frame_download(f, out, n), "download (learned)");
    int good = 0;
    for (size_t i = 0; i < n; i++)
        good += fabsf(out[i].deformation[0] - 2.f) < 0.01f && fabsf(out[i].deformation[6] + 1.f) < 0.01f;
    check(good == (int)n && qa.seed_pois == (int64_t)n, "learned-seed frame with FFTCC fallback");

    call(fi_frame_destroy(f), "destroy");
    call(fi_seeder_destroy(s), "seeder_destroy");
    free(out);
    free(redo);
    printf("%s\n", failures ? "FAIL" : "PASS all");
    return failures ? 1 : 0;
}
