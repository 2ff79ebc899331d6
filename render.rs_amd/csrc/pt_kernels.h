// Launch interface of the wavefront path-integrator kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "pt_types.h"

namespace rrhip {

// Traversal stack entries available per lane to the instance path.
constexpr uint32_t kPtTraceStack = 48;

// Path slots [0, total) with slot = s * npix + pixel: camera rays of
// samples prm.sample0 + s; fills the live queue (q0) and resets counters.
hipError_t pt_launch_raygen(const PtSceneView& S, const PathBuffers& P, const PtParams& prm,
                            uint32_t total, hipStream_t st);
// One bounce over the live queue `q`: closest hit, shading (emission + MIS,
// NEE shadow rays, BSDF continuation, Russian roulette) into `qnext`,
// shadow rays, counter rotation. `capacity` bounds every queue.
hipError_t pt_launch_bounce(const PtSceneView& S, const PathBuffers& P, const PtParams& prm,
                            const uint32_t* q, uint32_t* qnext, uint32_t capacity, hipStream_t st);
// Adds each pixel's pass samples (firefly-clamped) into the f64
// accumulator in sample order (deterministic, no atomics).
hipError_t pt_launch_resolve(const PathBuffers& P, const PtParams& prm, double* accum, hipStream_t st);

// Per-block partial sums of relMSE between two accumulators (`blocks`
// doubles into `partial`, <= 1024 blocks); the caller adds them.
hipError_t pt_launch_relmse(const double* accum, const double* ref_accum, uint32_t npix, double eps,
                            double* partial, uint32_t blocks, hipStream_t st);

}  // namespace rrhip
