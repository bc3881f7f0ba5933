/*
 * ope_hip.h -- C ABI of the MI355X (gfx950) processing library
 * (libope_hip.so).  This is the drop-in boundary behind which the HIP
 * backend replaces OCTproEngine's CUDA backend: the C++ class
 * ope::HipBackend (an implementation of the reference's
 * ProcessingBackend, src/backends/backend_interface.h:12-165) calls only
 * these functions, and the Python tests call them through ctypes.
 *
 * Conventions
 *  - Every function returns an int status: OPE_HIP_OK (0) or a negative
 *    OPE_HIP_E* code.  ope_hip_last_error(ctx) returns the message of the
 *    last failure on that context (thread-unsafe by design, like errno on one
 *    thread).  No C++ exception crosses this boundary.
 *  - Plain pointers and sizes only.  "device" pointers are HIP device
 *    memory; "host" pointers are ordinary or pinned host memory.
 *  - One context drives one GPU (the device id given at creation).
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef OPE_HIP_H
#define OPE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__) || defined(__clang__)
#define OPE_HIP_API __attribute__((visibility("default")))
#else
#define OPE_HIP_API
#endif

/* Status codes. */
#define OPE_HIP_OK 0
#define OPE_HIP_EINVAL -1    /* bad argument (null pointer, wrong length)  */
#define OPE_HIP_EUNSUP -2    /* configuration this build does not support  */
#define OPE_HIP_ERUNTIME -3  /* HIP runtime error (message has the detail)  */
#define OPE_HIP_ENOMEM -4    /* device or pinned-host allocation failed     */
#define OPE_HIP_ESTATE -5    /* call not valid in the context's state       */

/* Input sample types: the numeric values of ope::DataType
 * (include/types.h:6-19).  UINT8/INT8 are read as 8-bit unsigned, UINT16/
 * INT16 as 16-bit unsigned, UINT32/INT32 as 32-bit unsigned, exactly as the
 * reference converts by bit depth (cpu_kernels.tpp:24-54).  FLOAT32 is read
 * as IEEE float (the reference reinterprets its bits as uint32). */
#define OPE_HIP_UINT8 1
#define OPE_HIP_UINT16 2
#define OPE_HIP_UINT32 3
#define OPE_HIP_INT8 5
#define OPE_HIP_INT16 6
#define OPE_HIP_INT32 7
#define OPE_HIP_FLOAT32 9

/* Interpolation methods, ope::InterpolationMethod values
 * (include/processorconfiguration.h:17-21). */
#define OPE_HIP_INTERP_LINEAR 0
#define OPE_HIP_INTERP_CUBIC 1
#define OPE_HIP_INTERP_LANCZOS 2
/* Lanczos with a = 4 (8 taps, n0-3 .. n0+4), the kernel of
 * cpu_kernels.tpp:148-162 at the other window size it is written for.  No
 * counterpart in ope::InterpolationMethod, whose LANCZOS is a = 8
 * (cpu_kernels.tpp:174): reachable through this C ABI only. */
#define OPE_HIP_INTERP_LANCZOS4 3

/* Plain-old-data snapshot of the ProcessorConfiguration fields the chain
 * reads (include/processorconfiguration.h:33-172; read by
 * CpuBackend::Impl::processData, cpu_backend.cpp:169-302).  Booleans are
 * 0/1 ints. */
typedef struct ope_hip_params {
    int32_t signal_length;      /* N: power of two, 256..4096            */
    int32_t ascans_per_bscan;
    int32_t bscans_per_buffer;
    int32_t input_type;         /* OPE_HIP_UINT16, ...                    */
    int32_t resampling_enabled;
    int32_t interpolation;      /* OPE_HIP_INTERP_*                       */
    int32_t windowing_enabled;
    int32_t dispersion_enabled;
    int32_t dc_removal_enabled; /* rolling-average background removal     */
    int32_t dc_window_size;
    int32_t log_scaling;
    float grayscale_min;
    float grayscale_max;
    float multiplicator;
    float addend;
    int32_t fpn_removal;        /* fixed-pattern-noise subtraction        */
    int32_t fpn_bscan_count;
    int32_t fpn_continuous;
    int32_t bg_removal;         /* post-process background subtraction    */
    float bg_weight;
    float bg_offset;
    /* The CUDA backend's extra steps (the CPU backend ignores them):
     * bitshift: samples >> 4 (8/16-bit) or * 2^-32 (32-bit integer input)
     *   before conversion (cuda_kernels.cu:45-70); no effect on FLOAT32;
     * bscan_flip: reverse the A-scan order of B-scans 0, 2, 4, ... of a frame
     *   (cuda_kernels.cu:553-571; the last B-scan of an odd count stays);
     * sinusoidal_correction: resample each B-scan from sinusoidal to linear
     *   scan positions (cuda_kernels.cu:408-437). */
    int32_t bitshift;
    int32_t bscan_flip;
    int32_t sinusoidal_correction;
} ope_hip_params;

typedef struct ope_hip_ctx ope_hip_ctx;

/* Completion callback of a host frame: called on a HIP runtime thread after
 * the output has landed in the host output buffer. */
typedef void (*ope_hip_frame_cb)(void* user, float* host_out);

/* Returns a static string naming the library build. */
OPE_HIP_API const char* ope_hip_version(void);

/* Number of visible HIP devices (0 when none). */
OPE_HIP_API int ope_hip_device_count(int* count);

/* Creates a context on `device` and allocates the device tables and
 * `n_slots` (1..8) in-flight frame slots.
 * Replaces CudaBackend::initialize (cuda_backend.cu:179-292). */
OPE_HIP_API int ope_hip_create(int device, const ope_hip_params* params, int n_slots,
                               ope_hip_ctx** out_ctx);

/* Waits for in-flight work and frees every resource.
 * Replaces CudaBackend::cleanup (cuda_backend.cu:294-385). */
OPE_HIP_API int ope_hip_destroy(ope_hip_ctx* ctx);

/* Message of the last failure on ctx ("" if none); ctx may be NULL for
 * failures of ope_hip_create. */
OPE_HIP_API const char* ope_hip_last_error(const ope_hip_ctx* ctx);

/* Replaces the scalar/flag snapshot.  Dimensions (signal_length,
 * ascans_per_bscan, bscans_per_buffer, input_type) must equal the ones the
 * context was created with.  Replaces CudaBackend::updateConfig
 * (cuda_backend.cu:776-782). */
OPE_HIP_API int ope_hip_set_params(ope_hip_ctx* ctx, const ope_hip_params* params);

/* Curve hot-swap.  Each call waits for in-flight frames, so a new curve
 * never applies to half a frame.
 *   resampling: N float positions (CpuBackend::updateResamplingCurve,
 *               cpu_backend.cpp:485-490; CUDA cuda_backend.cu:784-797)
 *   window:     N floats (cpu_backend.cpp:503-508; cuda_backend.cu:815-828)
 *   dispersion: 2N floats, interleaved cos/sin (cpu_backend.cpp:492-501;
 *               cuda_backend.cu:799-813).  Another length disables the
 *               dispersion step, as the CPU backend does
 *               (cpu_kernels.tpp:205-208). */
OPE_HIP_API int ope_hip_set_resampling_curve(ope_hip_ctx* ctx, const float* curve, size_t length);
OPE_HIP_API int ope_hip_set_window_curve(ope_hip_ctx* ctx, const float* curve, size_t length);
OPE_HIP_API int ope_hip_set_dispersion_curve(ope_hip_ctx* ctx, const float* cos_sin,
                                             size_t length);

/* Fixed-pattern-noise profile, interleaved (re, im) pairs, k < N/2.
 * set: CpuBackend::setFixedPatternNoiseProfile (cpu_backend.cpp:817-822);
 *      pairs beyond N/2 are ignored, missing pairs subtract 0.  pairs == 0
 *      clears the profile.
 * get: writes min(pairs_capacity, stored) pairs; *pairs_out = stored pairs
 *      (0 when no profile exists).  (cpu_backend.cpp:824-826)
 * request: one-shot determination on the next frames
 *      (cpu_backend.cpp:813-815, 305-357). */
OPE_HIP_API int ope_hip_set_fpn_profile(ope_hip_ctx* ctx, const float* interleaved, size_t pairs);
OPE_HIP_API int ope_hip_get_fpn_profile(ope_hip_ctx* ctx, float* interleaved, size_t pairs_capacity,
                                        size_t* pairs_out);
OPE_HIP_API int ope_hip_request_fpn(ope_hip_ctx* ctx);

/* Post-process background profile, N/2 floats.
 * set: CpuBackend::setPostProcessBackgroundProfile (cpu_backend.cpp:793-806)
 *      (length must be N/2, else OPE_HIP_EINVAL)
 * get: as for the FPN profile (cpu_backend.cpp:808-810)
 * request: record from the next frame (cpu_backend.cpp:789-791,105-118). */
OPE_HIP_API int ope_hip_set_bg_profile(ope_hip_ctx* ctx, const float* profile, size_t length);
OPE_HIP_API int ope_hip_get_bg_profile(ope_hip_ctx* ctx, float* profile, size_t capacity,
                                       size_t* length_out);
OPE_HIP_API int ope_hip_request_bg_recording(ope_hip_ctx* ctx);

/* ---- Sample offset.  No counterpart in the reference, whose chain
 * converts the samples to float and carries their constant level (the
 * fringe's DC, ~2e4 for 16-bit cameras) through every float operation
 * (cpu_kernels.tpp:24-54, 84-267).  Its rounding then reaches every bin of
 * the float FFT; once an FPN profile removes the DC bins
 * (cpu_backend.cpp:246-281) weak bins within 60 dB of the peak carry it.
 * With an offset m0 the chain runs on x - m0 (exact for 8/16-bit samples,
 * m0 an integer) and the result is corrected by m0 * IFFT(S * phasor), S[j]
 * the sum of output j's tap weights, evaluated in FP64 and folded into the
 * epilogue's FPN constant: the outputs are those of the same chain on x,
 * with less rounding.  The FPN statistics add the term back, so profiles are
 * those of the chain on x.  Inactive while DC removal is enabled (that
 * removes the level itself).
 *   OPE_HIP_OFFSET_OFF:   m0 = 0 (the reference's arithmetic);
 *   OPE_HIP_OFFSET_AUTO:  default; on the next frame m0 becomes the mean of
 *                         its first (up to 32) A-scans, rounded to an
 *                         integer for integer samples, and is kept;
 *   OPE_HIP_OFFSET_FIXED: m0 = value (an integer for integer samples).
 * Frames on different caller streams are not ordered with the frame that
 * determines m0: submit that first frame before others that run
 * concurrently on another stream.  A ragged batch built while an automatic
 * offset is still undetermined runs without it.
 * get: *mode and the offset in use (0 while undetermined); waits for
 * in-flight frames. */
#define OPE_HIP_OFFSET_OFF 0
#define OPE_HIP_OFFSET_AUTO 1
#define OPE_HIP_OFFSET_FIXED 2
OPE_HIP_API int ope_hip_set_sample_offset(ope_hip_ctx* ctx, int mode, float value);
OPE_HIP_API int ope_hip_get_sample_offset(ope_hip_ctx* ctx, int* mode, float* value);

/* Device-resident processing: runs the whole chain over `n_ascans` A-scans
 * (a multiple of ascans_per_bscan is not required) from d_in (raw samples,
 * input_type, [n_ascans][N]) to d_out (float, [n_ascans][N/2]) on `stream`
 * (a hipStream_t; NULL = the HIP null stream).  Asynchronous.
 * FPN accumulation/determination and background recording follow the
 * frame semantics of CpuBackend::Impl::processData (cpu_backend.cpp:
 * 169-357), treating the call as one frame.
 * This is the GPU counterpart of the per-frame body of CudaBackend::process
 * (cuda_backend.cu:486-743) without the host copies. */
OPE_HIP_API int ope_hip_process_device(ope_hip_ctx* ctx, const void* d_in, float* d_out,
                                       int64_t n_ascans, void* stream);

/* Host frame: copies one frame (ascans_per_bscan * bscans_per_buffer
 * A-scans) from host_in to the device, processes it and copies the result to
 * host_out (N/2 floats per A-scan), all on the next free slot's stream.
 * Returns once the work is enqueued.  `on_input_free(user_in)` (may be
 * NULL) fires when host_in may be reused; `on_done(user, host_out)` fires
 * after host_out is complete.  Replaces CudaBackend::process
 * (cuda_backend.cu:442-773): H2D, chain, D2H, host callbacks. */
OPE_HIP_API int ope_hip_submit_host(ope_hip_ctx* ctx, const void* host_in, float* host_out,
                                    void (*on_input_free)(void* user_in), void* user_in,
                                    ope_hip_frame_cb on_done, void* user);

/* Blocks until all work submitted on the context has finished. */
OPE_HIP_API int ope_hip_synchronize(ope_hip_ctx* ctx);

/* Pinned host memory for input/output IOBuffers (hipHostMalloc; replaces
 * the cudaHostRegister of cuda_backend.cu:220-233). */
OPE_HIP_API int ope_hip_host_alloc(size_t bytes, void** ptr);
OPE_HIP_API int ope_hip_host_free(void* ptr);

/* Individual operation getMinimumVarianceMean (ProcessingBackend,
 * backend_interface.h:117-122; CPU semantics cpu_backend.cpp:682-699 and
 * cpu_kernels.tpp:402-458): for each of `width` columns of the interleaved
 * complex rows input[height][width][2], the complex mean of the segment (of
 * height/segments rows) with the smallest variance.  Host pointers; out has
 * 2*width floats.  Synchronous. */
OPE_HIP_API int ope_hip_op_min_variance_mean(ope_hip_ctx* ctx, const float* input, int width,
                                             int height, int segments, float* out);

/* The same operation on device memory, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_input: device, interleaved complex rows [height][width][2] (8-byte
 * aligned); d_out: device, 2*width floats.  Asynchronous on `stream` (NULL:
 * the HIP null stream).  Its segment sums live in a static device array of
 * 2 MiB per GPU, so segments * width must not exceed 87381 (else
 * OPE_HIP_EINVAL); the array is shared by every call on that GPU, so calls
 * that overlap on one GPU race.  The context supplies the device (its
 * creation takes the array's address there). */
OPE_HIP_API int ope_hip_op_min_variance_mean_device(ope_hip_ctx* ctx, const float* d_input, int width,
                                                    int height, int segments, float* d_out, void* stream);

/* ---- Launch autotuning.  No counterpart in the reference (its CUDA
 * backend launches fixed grids, cuda_backend.cu:486-743).  Times the chain's
 * launch variants for the context's current configuration (interpolation,
 * DC removal, input width): 1 or 2 A-scans per lane (with dispersion off
 * and no DC removal the two A-scans of a lane are real and share one complex
 * FFT; that is also the default there), the position of the FFT's
 * small-radix pass (position = number of passes: the swap schedule of
 * N = 1024 and 2048, radices 8, 8, N/512, 8 with the exchange into the
 * small-radix pass made by lane-half register swaps instead of LDS),
 * 1 .. resident workgroups per CU; each variant
 * runs once untimed and 3 times timed on d_in -> d_out (n_ascans A-scans;
 * d_out is overwritten).  The fastest is used by later frames of that
 * configuration on this context.  With d_in and d_out NULL the context's own
 * first frame slot is used.  choice (may be NULL) receives {A-scans per
 * lane, small-radix pass position, workgroups per CU}; best_us (may be
 * NULL) the fastest variant's time per launch.  Synchronous. */
OPE_HIP_API int ope_hip_autotune(ope_hip_ctx* ctx, const void* d_in, float* d_out, int64_t n_ascans,
                                 void* stream, int* choice, float* best_us);

/* Pins the launch variant of the context's current configuration instead
 * of tuning it: A-scans per lane (1 or 2), small-radix pass position
 * (0 .. passes-1; the last pass is the default; `passes` selects the swap
 * schedule, built for N = 1024 and 2048; other positions and the swap
 * schedule exist for one A-scan per lane only: OPE_HIP_EUNSUP), workgroups
 * per CU (0: all that are resident). */
OPE_HIP_API int ope_hip_set_launch_variant(ope_hip_ctx* ctx, int ascans_per_lane, int small_radix_pass,
                                           int workgroups_per_cu);

/* ---- Graph-captured frames.  No counterpart in the reference (one
 * process() call per frame, cuda_backend.cu:442-773).  For callers that
 * replay a fixed ring of device frames (volume acquisition into
 * preallocated buffers): n_frames launches of the chain on d_in[f] ->
 * d_out[f] captured into one HIP graph, replayed by one call.
 * Only valid in steady state (no FPN determination, FPN accumulation or
 * background recording pending: OPE_HIP_ESTATE otherwise); the graph keeps
 * the context's tables, so curve or parameter changes after capture need a
 * new graph.  Errors of create go to ope_hip_graph_last_error(NULL). */
typedef struct ope_hip_graph ope_hip_graph;
OPE_HIP_API int ope_hip_graph_create(ope_hip_ctx* ctx, const void* const* d_in, float* const* d_out,
                                     int n_frames, int64_t ascans_per_frame, ope_hip_graph** out);
OPE_HIP_API int ope_hip_graph_launch(ope_hip_graph* graph, void* stream);
OPE_HIP_API const char* ope_hip_graph_last_error(const ope_hip_graph* graph);
OPE_HIP_API int ope_hip_graph_destroy(ope_hip_graph* graph);

/* ---- Individual operations (ProcessingBackend, backend_interface.h:40-165)
 * with the CPU backend's semantics (cpu_backend.cpp:514-733); the CUDA
 * backend returns nothing for them (cuda_backend.cu:1133-1296).  Host
 * pointers in and out, synchronous, run on the context's GPU.
 *  convert_input: samples as the chain converts them (bitshift honoured;
 *    FLOAT32 read as IEEE) -- convertInput;
 *  klinearization: one line of line_width samples resampled at curve[j]
 *    (LINEAR/CUBIC/LANCZOS), first `samples` outputs -- kLinearization;
 *  windowing: input * window -- windowing;
 *  dispersion: real part of input * (cos, sin), cos_sin interleaved --
 *    dispersionCompensation;
 *  ifft_magnitude: |unnormalised inverse FFT| of `lines` real lines of
 *    line_width (a power of two 256..4096) samples -- ifft;
 *  fpn_removal: |x - mean| of line_width * lines interleaved complex values,
 *    the mean subtracted from the first min(total/2, line_width) values as
 *    the reference does on the flattened lines -- fixedPatternNoiseRemoval. */
OPE_HIP_API int ope_hip_op_convert_input(ope_hip_ctx* ctx, const void* input, int input_type, int64_t samples,
                                         int bitshift, float* out);
OPE_HIP_API int ope_hip_op_klinearization(ope_hip_ctx* ctx, const float* input, const float* curve, int method,
                                          int line_width, int samples, float* out);
OPE_HIP_API int ope_hip_op_windowing(ope_hip_ctx* ctx, const float* input, const float* window, int line_width,
                                     int samples, float* out);
OPE_HIP_API int ope_hip_op_dispersion(ope_hip_ctx* ctx, const float* input, const float* cos_sin, int line_width,
                                      int samples, float* out);
OPE_HIP_API int ope_hip_op_ifft_magnitude(ope_hip_ctx* ctx, const float* input, int line_width, int lines,
                                          float* out);
OPE_HIP_API int ope_hip_op_fpn_removal(ope_hip_ctx* ctx, const float* input, const float* mean, int line_width,
                                       int lines, float* out);

/* ---- Ragged batches.  No counterpart in the reference, whose backends
 * process one fixed-size frame per call (ProcessingBackend::process,
 * backend_interface.h:27).  A ragged batch runs many requests of different
 * signal lengths and A-scan counts in one call: one persistent launch per
 * signal length present, each that length's own chain kernel (its tuned
 * launch variant and occupancy) over a row map of the requests' A-scans,
 * for serving callers that queue requests from several scan protocols.
 * Every request's output equals that of a process_device call over its
 * A-scans alone, bit for bit (without dispersion and DC removal, where two
 * A-scans share one FFT, each request starts a new pair).  Within a signal
 * length the requests are processed shortest first (fewest A-scans), which
 * minimises the mean request latency. */

/* One request: n_ascans A-scans of signal_length samples from d_in (device,
 * [n_ascans][signal_length] in the contexts' input type, 16-byte aligned)
 * to d_out (device, [n_ascans][signal_length/2] float, 4-byte aligned). */
typedef struct ope_hip_request {
    const void* d_in;
    float* d_out;
    int64_t n_ascans;
    int32_t signal_length;
    int32_t reserved; /* 0 */
} ope_hip_request;

typedef struct ope_hip_ragged ope_hip_ragged;

/* Builds a batch over `reqs` (copied).  ctxs[0..n_ctx) supply the chain of
 * each signal length (tables, epilogue, FPN profile and background as they
 * are at this call); their lengths must differ, and they must agree on the
 * interpolation method, DC removal and input sample width, and live on one
 * device.  A context with an FPN determination or background recording
 * pending is rejected (OPE_HIP_ESTATE): run it through process_device
 * first.  chunk_ascans: unused, must be >= 0 (the launches stride over
 * single A-scans).  Errors are reported through ope_hip_last_error(NULL). */
OPE_HIP_API int ope_hip_ragged_create(ope_hip_ctx* const* ctxs, int n_ctx, const ope_hip_request* reqs,
                                      int n_reqs, int chunk_ascans, ope_hip_ragged** out);

/* The whole batch on `stream` (NULL: the HIP null stream): the largest
 * signal length on `stream`, the second largest partly beside it on a
 * stream of the batch's own (forked from `stream` and joined back into it by
 * events) and partly after it, the others after them; work queued on
 * `stream` afterwards sees the whole batch.  Asynchronous, allocates nothing
 * and may be captured in a graph.  Runs of one batch must not overlap (they
 * share its stream, its schedule counters and its completion stamps). */
OPE_HIP_API int ope_hip_ragged_run(ope_hip_ragged* batch, void* stream);

/* Per-request latency of the last completed run: microseconds from the
 * start of the run's first launch to the moment the request's last A-scan
 * had its outputs issued (0 for empty requests).  Call after the run has
 * finished.  latency_us has n_reqs entries. */
OPE_HIP_API int ope_hip_ragged_latencies(ope_hip_ragged* batch, double* latency_us);

OPE_HIP_API const char* ope_hip_ragged_last_error(const ope_hip_ragged* batch);
OPE_HIP_API int ope_hip_ragged_destroy(ope_hip_ragged* batch);

/* ---- Autofocus.  No counterpart in the reference, which takes dispersion
 * and resampling coefficients as given (setDispersionCoefficients,
 * setResamplingCoefficients, processor.h).  A differentiable form of the
 * chain for callers that tune those coefficients by gradient descent on
 * image sharpness.
 *
 * theta[6] = (d2, d3, c0, c1, c2, c3):
 *   resampling curve r[j] = clamp(c0 + c1 u + c2 u^2 + c3 u^3, 7, N-9),
 *   u = j/(N-1) (the polynomial of generateResamplingCurve,
 *   processorconfiguration.cpp:438-459); dispersion phase
 *   phi[j] = d2 (j-N/2)^2 + d3 (j-N/2)^3.
 * The chain is the context's: input type (16- or 32-bit), DC removal,
 * window, and linear or cubic interpolation (resampling must be enabled);
 * the context's curves and FPN/background state are not used.  The loss is
 * the Shannon entropy of the intensity |X[k]|^2, k < N/2, normalised over
 * all n_ascans A-scans of d_in (device, 16-byte aligned).
 * Both calls run on `stream` and return after the result is on the host. */

/* out[3] = (loss, S = sum of intensities, sum of I log I). */
OPE_HIP_API int ope_hip_autofocus_loss(ope_hip_ctx* ctx, const void* d_in, int64_t n_ascans,
                                       const double* theta, double* out, void* stream);

/* grad[6] = dloss/dtheta at theta, given S and the loss of the same theta
 * from ope_hip_autofocus_loss (the hand-written backward of the chain). */
OPE_HIP_API int ope_hip_autofocus_grad(ope_hip_ctx* ctx, const void* d_in, int64_t n_ascans,
                                       const double* theta, double sum_i, double loss, double* grad,
                                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OPE_HIP_H */
