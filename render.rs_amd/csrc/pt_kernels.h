// Launch interface of the wavefront path-integrator kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "pt_types.h"

namespace rrhip {

// Traversal stack entries available per lane to the instance path.
constexpr uint32_t kPtTraceStack = 48;
// Film accumulator per pixel (doubles): diffuse sums rgb [0, 3), albedo
// sums [3, 6), normal sums [6, 9), depth sum [9], samples with a first
// real hit [10], kFilmIds vote ids [11, 15), their counts [15, 19),
// samples whose id found no slot [19].
constexpr uint32_t kFilmIds = 4;
constexpr uint32_t kFilmStride = 12 + 2 * kFilmIds;
// Hull crossings and cutout pass-throughs per path (pt/mod.rs:32): they do
// not count as bounces, so transparent scenes run this many extra rounds.
constexpr uint32_t kPtMaxSkips = 16;

// Path slots [0, total) with slot = s * npix + column: camera rays of
// samples prm.sample0 + s of the column's pixel (prm.live[column], or the
// column itself); fills the live queue (q0) and resets counters.
// With N.ntrain > 0 also training paths in slots [total, total + ntrain).
// sample0_dev (may be null) overrides prm.sample0 from device memory.
hipError_t pt_launch_raygen(const PtSceneView& S, const PathBuffers& P, const PtParams& prm,
                            uint32_t total, const NrcPathBuffers& N, const uint64_t* sample0_dev, hipStream_t st);
// *dst = v, in stream order (a kernel, so it never becomes a memset node).
hipError_t pt_launch_set_u64(uint64_t* dst, uint64_t v, hipStream_t st);
// One bounce over the live queue `q`: closest hit, shading (emission + MIS,
// NEE shadow rays, BSDF continuation, Russian roulette) into `qnext`,
// shadow rays, counter rotation. `capacity` bounds every queue. With
// N.q set, paths terminate into radiance-cache queries (see NrcPathBuffers).
// ev (profiling, may be null): four events recorded before extend, before
// shade (the queue sort included), before shadow, and after it.
hipError_t pt_launch_bounce(const PtSceneView& S, const PathBuffers& P, const PtParams& prm,
                            const uint32_t* q, uint32_t* qnext, uint32_t capacity, const NrcPathBuffers& N,
                            hipStream_t st, const hipEvent_t* ev = nullptr);
// One pass of a batch: raygen over every job's slots (jobs sorted by
// slot_begin and pix_begin), `bounces` rounds of extend / shade / shadow
// with each path reading its own job's scene and parameters, and the
// resolve of every job's pixels into its accumulator. `quadrics`: some
// job's scene holds quadrics (the trace kernels then carry their solvers).
hipError_t pt_launch_batch_pass(const PtJob* jobs, uint32_t nj, const PathBuffers& P, uint32_t total,
                                 uint32_t total_pix, uint32_t bounces, bool quadrics, hipStream_t st);
// Every job of a batch in one persistent launch: `lanes` path slots (a
// multiple of the block), each pulling (job, pixel) items from *next_item
// (zero on entry) until `total_items`; a job's prm.spp_pass is all of its
// samples and pix_begin its first item. job_left[j] counts the job's
// pixels down; the last one writes the device clock to job_clock[j], and
// job_clock[nj] takes the earliest block start (set it to ~0 first).
// flags: bit 0 some job is transparent, bit 1 some job moves, bit 2 some
// job holds quadrics (error-bounded ray origins only).
hipError_t pt_launch_batch_persistent(const PtJob* jobs, uint32_t nj, const PathBuffers& P, uint32_t total_items,
                                      uint32_t lanes, uint32_t flags, uint32_t* next_item, uint32_t* job_left,
                                      unsigned long long* job_clock, hipStream_t st);
// Path slots of one round of resident persistent blocks (0 if unknown).
uint32_t pt_batch_persistent_lanes(uint32_t flags);
// Adds each pixel's pass samples (firefly-clamped) into the f64
// accumulator in sample order (deterministic, no atomics).
hipError_t pt_launch_resolve(const PathBuffers& P, const PtParams& prm, double* accum, double* film,
                             hipStream_t st);

// out[i * count + k] = draw k of Pcg32::for_pixel_sample(pixel[i],
// sample[i]), the stream a camera path of that pixel and sample runs on.
// The device fiber BSDF on n queries (f, pdf towards wi; a sample from u).
hipError_t pt_launch_hair_bsdf(const PtMaterial& m, uint32_t n, const float* wo, const float* wi, const float* h,
                               const float* u, float* f, float* pdf, float* ws, float* fs, float* pdfs,
                               hipStream_t st);
hipError_t pt_launch_pcg32_draws(const uint64_t* pixel, const uint64_t* sample, uint32_t n, uint32_t count,
                                 uint32_t* out, hipStream_t st);

// Per-block partial sums of relMSE between two accumulators (`blocks`
// doubles into `partial`, <= 1024 blocks); the caller adds them.
hipError_t pt_launch_relmse(const double* accum, const double* ref_accum, uint32_t npix, double eps,
                            double* partial, uint32_t blocks, hipStream_t st);

// Per-block partial sums of the accumulators' sample counts (`blocks`
// doubles into `partial`); the caller adds them.
hipError_t pt_launch_count_sum(const double* accum, uint32_t npix, double* partial, uint32_t blocks,
                               hipStream_t st);
// Adaptive sampling: the pixels of live list `in` (n columns; null: pixel
// = column) whose accumulators have not met the stopping rule of
// pt/mod.rs:319-325 at tolerance `tol`, in column order, into `out`; their
// number into *nlive. bcount holds ceil(n / 256) u32. n >= 1.
hipError_t pt_launch_adaptive_prune(const double* accum, const uint32_t* in, uint32_t n, double tol,
                                    uint32_t* bcount, uint32_t* out, uint32_t* nlive, hipStream_t st);

// Training samples of the pass's training paths, in (path, vertex) order
// (deterministic): inputs into bq, targets into bt (3 floats each), at
// most `cap`; the count into N.counts[1]. `scratch` holds ntrain u32.
hipError_t pt_launch_nrc_samples(const PathBuffers& P, const NrcPathBuffers& N, float clamp, uint32_t* scratch,
                                 uint32_t cap, NrcQuery* bq, float* bt, hipStream_t st);
// *n_out = the size of optimiser batch k (of `batch` samples) of the pass.
hipError_t pt_launch_nrc_set_batch(const NrcPathBuffers& N, uint32_t k, uint32_t batch, uint32_t* n_out,
                                   hipStream_t st);

}  // namespace rrhip
