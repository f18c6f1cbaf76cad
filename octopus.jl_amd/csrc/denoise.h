// denoise.h -- the residual denoiser inside the collision (k_denoise.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"
#include "octopic.h"

namespace octopic {

// One batch's field planes through the denoiser: source moments of every interaction, the
// Gaussian-subtracted residual of every field plane as network input, and, after the network, the
// plane rebuilt as Gaussian + denoised residual (its total charge kept) where the plane's monitor
// allows it.
struct DnCollideArgs {
    const Inter *inters;
    int ninters;
    const Plane *planes;
    int nplanes;
    States st;
    unsigned long long *charge;   // fixed-point field planes (row y, column x), 2^-k units
    int64_t charge_stride;
    double scale, inv_scale;
    int method, nx, ny;
    int chunks;                   // moment chunks per interaction
    double *mom_part;             // [ninters][chunks][10]
    double *meta;                 // [nplanes][4]: amp, peak, monitor, use
    double *prof;                 // [nplanes][nx + ny]: gx, gy
    float *input;                 // [nplanes][ny][nx][2]
    double limit;                 // monitor limit
    unsigned long long *counters; // [2]: planes denoised, planes left to plain PIC
    const ErrorInfo *err;
};
constexpr int kDnMomentChunks = 16;
hipError_t launch_dn_prepare(const DnCollideArgs &a, hipStream_t s);
hipError_t launch_dn_apply(const DnCollideArgs &a, const float *out, hipStream_t s);

// The network on `batch` planes of `input`, enqueued on stream s; returns its output buffer.
const float *denoiser_run(octopic_denoiser *d, const float *input, int batch, hipStream_t s, hipError_t *e);
void denoiser_shape(const octopic_denoiser *d, int *height, int *width, int *max_batch);

}  // namespace octopic
