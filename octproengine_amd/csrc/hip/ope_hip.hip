// Host side of the C ABI declared in include/ope_hip.h.
//
// Owns one GPU context: device tables (banded resampling taps with the window
// folded in, dispersion phasor, FFT twiddles, FPN profile, background
// profile), the FPN segment-statistics accumulator, and `n_slots` frame
// slots for host frames.  Host frames run H2D on a copy-in stream, all
// kernels in submission order on one compute stream, and D2H on a copy-out
// stream, chained with events so copies of one frame overlap compute of
// another while the chain state (FPN accumulation) stays in frame order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <complex>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/ope_hip.h"
#include "chain_dispatch.h"
#include "frame_kernels.hpp"

using namespace ope_hip;

// Device scratch of ope_hip_op_min_variance_mean_device: the FP64 segment
// sums [segments][width][3] of one call, so that the entry point allocates
// nothing.  Each GPU holds its own copy for the life of the process; calls
// on one GPU share it.  In host code `scratch` names a host-side stand-in:
// its device address on each GPU comes from hipGetSymbolAddress, taken once
// per device when a context is created there (scratch_init).
constexpr size_t kScratchBytes = 2u << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // device address of `scratch`, per device id

// Takes the address of `scratch` on `device` (which must be current) once.
hipError_t scratch_init(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0) return hipErrorInvalidDevice;
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (g_scratch_ptr[device]) return hipSuccess;
    double* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&ptr), HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

double* scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && (size_t)device < g_scratch_ptr.size()) ? g_scratch_ptr[device] : nullptr;
}

constexpr int kFpnSegments = 8;  // cpu_backend.cpp:306

struct Slot {
    void* d_in = nullptr;
    float* d_out = nullptr;
    hipEvent_t h2d_done = nullptr;
    hipEvent_t comp_done = nullptr;
    hipEvent_t d2h_done = nullptr;
    bool used = false;
};

struct HostCall {
    void (*fn)(void*, float*);
    void* user;
    float* out;
    void (*free_fn)(void*);
    void* free_user;
};

}  // namespace

struct ope_hip_ctx {
    int device = 0;
    ope_hip_params p{};
    int n = 0;
    int taps = 1;
    int cu_count = 0;
    std::string err;

    // host copies of the curves as last given
    std::vector<float> curve_resample, curve_window, curve_phasor;
    bool phasor_valid = false;

    // device tables
    int* d_tap_base = nullptr;
    float* d_tap_w = nullptr;
    size_t tap_w_cap = 0;
    float2* d_phasor = nullptr;
    float2* d_twiddle = nullptr;
    float2* d_fpn = nullptr;
    float* d_bg = nullptr;
    double* d_acc = nullptr;
    bool have_fpn = false;
    bool have_bg = false;
    size_t fpn_pairs_stored = 0;

    // FPN accumulation state (cpu_backend.cpp:57-59,305-357)
    long long fpn_count = 0;
    bool fpn_requested = false;
    bool bg_requested = false;

    // streams and frame slots
    hipStream_t s_in = nullptr, s_comp = nullptr, s_out = nullptr;
    std::vector<Slot> slots;
    int next_slot = 0;
    std::vector<HostCall> calls;  // per slot
    std::mutex submit_mu;

    // resident workgroups per CU of each chain kernel used so far (the
    // occupancy query is a runtime call; frames reuse one or two kernels)
    std::vector<std::pair<const void*, int>> occupancy;

    // launch variant chosen by ope_hip_autotune for a kernel configuration
    struct Tuning {
        int taps, mode, iw;      // configuration
        int as, sched, per_cu;   // A-scans per lane, small-radix pass, workgroups per CU
    };
    std::vector<Tuning> tuning;

    // sinusoidal scan correction: positions per A-scan of a B-scan and the
    // chain's output before the correction
    float* d_sin_curve = nullptr;
    int sin_curve_len = 0;
    float* d_sin_tmp = nullptr;
    size_t sin_tmp_cap = 0;  // floats

    // autofocus: window used by the differentiable chain, FP64 sums
    float* d_af_window = nullptr;
    double* d_af_acc = nullptr;

    // sample offset (ope_hip_set_sample_offset): mode, whether an automatic
    // offset is still to be determined, the device scalar m0, the FP64 table
    // D = IFFT(S * phasor) for k < N/2 and the epilogue constant
    // fl(fpn - m0 * D); S[j] = sum of output j's tap weights
    int offset_mode = OPE_HIP_OFFSET_AUTO;
    bool offset_pending = true;
    float* d_offset = nullptr;
    double2* d_offset_tab = nullptr;
    float2* d_ep = nullptr;
    std::vector<double> tap_sums;

    // set when work on caller streams (process_device, graphs, ragged
    // batches) may still read the context's tables: the table setters then
    // wait for the whole device, not only the context's own streams
    std::atomic<bool> foreign_work{false};
};

// A ragged batch: per signal length present, the chain arguments of its
// context (snapshot) with the row map of that length's A-scans; the
// per-request completion stamps.
struct ope_hip_ragged {
    int device = 0;
    std::vector<ope_hip_ctx*> ctxs;  // whose tables the batch reads
    std::string err;
    int n_reqs = 0;
    double tick_hz = 0.0;
    std::vector<long long> req_ascans;
    struct Class {
        ope_hip_ctx* ctx;
        ChainArgs args;      // the row map of this length's A-scans
        long long ascans;
        ChainLaunch launch;  // the length's default launch (row-map kernel)
        long long grid_side = 0, grid_main = 0;  // workgroups beside the largest length and after it
        unsigned side_lds = 0;                   // dynamic LDS of the side launch (bytes)
        unsigned* d_tickets = nullptr;           // [2] the dynamic schedule's counters
    };
    std::vector<Class> classes;  // in the order of their first request
    RowRef* d_rows = nullptr;
    unsigned long long* d_finish = nullptr;
    unsigned long long* d_start = nullptr;
    float* d_discard = nullptr;  // output of the rows that pad paired requests
    // the lengths after the first run concurrently with it on streams of
    // their own, forked from and joined back into the caller's stream
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
};

// Frames of one context captured into a HIP graph.
struct ope_hip_graph {
    int device = 0;
    ope_hip_ctx* ctx = nullptr;  // whose tables the frames read
    std::string err;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int fail(ope_hip_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

thread_local std::string g_create_err;

#define HIP_TRY(ctx, expr)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((ctx), OPE_HIP_ERUNTIME, "%s failed at %s:%d: %s", #expr, __FILE__,    \
                        __LINE__, hipGetErrorString(e_));                                      \
    } while (0)

bool supported_n(int n) { return n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096; }

int input_class(int t) {
    switch (t) {
        case OPE_HIP_UINT8:
        case OPE_HIP_INT8:
            return IN_U8;
        case OPE_HIP_UINT16:
        case OPE_HIP_INT16:
            return IN_U16;
        case OPE_HIP_UINT32:
        case OPE_HIP_INT32:
            return IN_U32;
        case OPE_HIP_FLOAT32:
            return IN_F32;
        default:
            return 0;
    }
}

size_t input_bytes(int t) {
    switch (input_class(t)) {
        case IN_U8: return 1;
        case IN_U16: return 2;
        default: return 4;
    }
}

int taps_for(const ope_hip_params& p) {
    if (!p.resampling_enabled) return 1;
    switch (p.interpolation) {
        case OPE_HIP_INTERP_LINEAR: return 2;
        case OPE_HIP_INTERP_CUBIC: return 4;
        case OPE_HIP_INTERP_LANCZOS4: return 8;
        default: return 16;
    }
}

// C static_cast<int>(float): truncation toward zero.
inline int trunc_int(float v) { return static_cast<int>(v); }

// cpu_kernels.tpp:148-162 evaluated in double, window size a
double lanczos(double d, int a) {
    if (std::fabs(d) < 1e-5) return 1.0;
    if (std::fabs(d) >= (double)a) return 0.0;
    const double pd = M_PI * d;
    const double pda = pd / (double)a;
    return (std::sin(pd) / pd) * (std::sin(pda) / pda);
}

// Banded tap table: y[j] = sum_i w[j][i] * x[base[j] + i], with the window
// folded in.  Index clamping follows cpu_kernels.tpp:84-196 (upper clamp to
// N-1 everywhere, lower clamp to 0 for Lanczos); indices below 0, which the
// reference reads out of bounds, are clamped to 0 here.
// Tap table of `count` outputs over an input line of n samples with taps T
// (1: identity) at positions curve[j], window (may be null: ones) folded in.
int build_taps_for(ope_hip_ctx* c, int n, int count, int T, const float* curve, const float* window,
                   std::vector<int>& base, std::vector<float>& w) {
    const bool win = window != nullptr;
    base.assign(count, 0);
    w.assign((size_t)count * T, 0.f);
    std::vector<double> acc(T);
    std::vector<int> idx(16);
    std::vector<double> cw(16);
    for (int j = 0; j < count; ++j) {
        int cnt = 0;
        if (T == 1) {
            idx[0] = j;
            cw[0] = 1.0;
            cnt = 1;
        } else {
            const float r = curve[j];
            if (!std::isfinite(r)) return fail(c, OPE_HIP_EINVAL, "non-finite resampling position at %d", j);
            if (T == 2) {
                int n1 = trunc_int(r), n2 = n1 + 1;
                n1 = std::min(n1, n - 1);
                n2 = std::min(n2, n - 1);
                const double t = (double)r - (double)n1;
                idx[0] = n1; cw[0] = 1.0 - t;
                idx[1] = n2; cw[1] = t;
                cnt = 2;
            } else if (T == 4) {
                int n1 = trunc_int(r);
                int n0 = std::abs(n1 - 1), n2 = n1 + 1, n3 = n2 + 1;
                n0 = std::min(n0, n - 1);
                n1 = std::min(n1, n - 1);
                n2 = std::min(n2, n - 1);
                n3 = std::min(n3, n - 1);
                const double t = (double)r - (double)n1;
                const double t2 = t * t, t3 = t2 * t;
                idx[0] = n0; cw[0] = -0.5 * t3 + t2 - 0.5 * t;
                idx[1] = n1; cw[1] = 1.5 * t3 - 2.5 * t2 + 1.0;
                idx[2] = n2; cw[2] = -1.5 * t3 + 2.0 * t2 + 0.5 * t;
                idx[3] = n3; cw[3] = 0.5 * t3 - 0.5 * t2;
                cnt = 4;
            } else {
                // Lanczos, a = T/2: taps n0-a+1 .. n0+a (cpu_kernels.tpp:164-196)
                const int a = T / 2;
                const int n0 = trunc_int(r);
                for (int i = -a + 1; i <= a; ++i) {
                    int k = n0 + i;
                    if (k < 0) k = 0;
                    if (k >= n) k = n - 1;
                    idx[cnt] = k;
                    cw[cnt] = lanczos((double)r - (double)k, a);
                    ++cnt;
                }
            }
        }
        int lo = n, hi = -1;
        for (int i = 0; i < cnt; ++i) {
            idx[i] = std::max(0, std::min(idx[i], n - 1));
            lo = std::min(lo, idx[i]);
            hi = std::max(hi, idx[i]);
        }
        int b = std::max(0, std::min(lo, n - T));
        if (hi - b >= T) return fail(c, OPE_HIP_EINVAL, "tap window overflow at %d", j);
        std::fill(acc.begin(), acc.end(), 0.0);
        for (int i = 0; i < cnt; ++i) acc[idx[i] - b] += cw[i];
        const double wj = win ? (double)window[j] : 1.0;
        base[j] = b;
        for (int i = 0; i < T; ++i) w[(size_t)j * T + i] = (float)(acc[i] * wj);
    }
    return OPE_HIP_OK;
}

int build_taps(ope_hip_ctx* c, std::vector<int>& base, std::vector<float>& w) {
    return build_taps_for(c, c->n, c->n, c->taps, c->curve_resample.data(),
                          c->p.windowing_enabled ? c->curve_window.data() : nullptr, base, w);
}

int sync_all(ope_hip_ctx* c);

int upload_taps(ope_hip_ctx* c) {
    c->taps = taps_for(c->p);
    std::vector<int> base;
    std::vector<float> w;
    int rc = build_taps(c, base, w);
    if (rc) return rc;
    rc = sync_all(c);
    if (rc) return rc;
    if (w.size() > c->tap_w_cap) {
        if (c->d_tap_w) HIP_TRY(c, hipFree(c->d_tap_w));
        c->d_tap_w = nullptr;
        HIP_TRY(c, hipMalloc(&c->d_tap_w, w.size() * sizeof(float)));
        c->tap_w_cap = w.size();
    }
    HIP_TRY(c, hipMemcpy(c->d_tap_base, base.data(), base.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_tap_w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    // the offset table's S[j]: the float weights as uploaded, summed in FP64
    c->tap_sums.assign(c->n, 0.0);
    for (int j = 0; j < c->n; ++j)
        for (int i = 0; i < c->taps; ++i) c->tap_sums[j] += (double)w[(size_t)j * c->taps + i];
    return OPE_HIP_OK;
}

// ---------------------------------------------------------------- sample offset

// Unnormalised inverse DFT in FP64, X[k] = sum_j a[j] exp(+2 pi i j k / N),
// in place (iterative radix 2; N a power of two).  Twiddles come from cos
// and sin of each angle, so the error stays within a few ulp of |a|_1.
void idft64(std::vector<std::complex<double>>& a) {
    const int n = (int)a.size();
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    std::vector<std::complex<double>> tw(n / 2);
    for (int k = 0; k < n / 2; ++k) {
        const double ang = 2.0 * M_PI * (double)k / (double)n;
        tw[k] = std::complex<double>(std::cos(ang), std::sin(ang));
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int h = len / 2, stride = n / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < h; ++k) {
                const std::complex<double> u = a[i + k], v = a[i + k + h] * tw[(size_t)k * stride];
                a[i + k] = u + v;
                a[i + k + h] = u - v;
            }
    }
}

bool offset_active(const ope_hip_ctx* c) {
    return c->offset_mode != OPE_HIP_OFFSET_OFF && !c->p.dc_removal_enabled;
}

bool fpn_applied(const ope_hip_ctx* c) { return c->p.fpn_removal && c->have_fpn; }

// Frames with the offset: samples minus m0, the epilogue constant in place
// of the FPN profile, D for the statistics mode.
void apply_offset(ope_hip_ctx* c, ChainArgs& a) {
    a.flags |= F_OFFSET;
    a.offset = c->d_offset;
    a.offset_tab = c->d_offset_tab;
    a.fpn = c->d_ep;
}

// Rebuilds the offset's tables after a change of the taps, the phasor, the
// dispersion flag, the FPN profile or m0: D (when `table`) on the host in
// FP64, then the epilogue constant fl(fpn - m0 * D) on the device.  Waits
// for it, so frames on any stream see the new constant.
int refresh_offset(ope_hip_ctx* c, bool table) {
    if (c->offset_mode == OPE_HIP_OFFSET_OFF) return OPE_HIP_OK;
    const int n = c->n, half = n / 2;
    HIP_TRY(c, hipSetDevice(c->device));
    if (table) {
        const bool disp = c->p.dispersion_enabled && c->phasor_valid;
        std::vector<std::complex<double>> a(n);
        for (int j = 0; j < n; ++j) {
            const std::complex<double> ph =
                disp ? std::complex<double>(c->curve_phasor[2 * (size_t)j], c->curve_phasor[2 * (size_t)j + 1])
                     : std::complex<double>(1.0, 0.0);
            a[j] = c->tap_sums[j] * ph;
        }
        idft64(a);
        std::vector<double2> d(half);
        for (int k = 0; k < half; ++k) d[k] = make_double2(a[k].real(), a[k].imag());
        HIP_TRY(c, hipMemcpy(c->d_offset_tab, d.data(), half * sizeof(double2), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(offset_prepare_kernel, dim3(1), dim3(256), 0, c->s_comp, nullptr, n, 0, 2, 0, 0, 0,
                       c->d_offset, c->d_fpn, fpn_applied(c) ? 1 : 0, c->d_offset_tab, c->d_ep, half);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->s_comp));
    return OPE_HIP_OK;
}

// Determines an automatic offset from the first (up to 32) A-scans of d_in
// on stream s, and rebuilds the epilogue constant with it.
int determine_offset(ope_hip_ctx* c, const void* d_in, long long n_ascans, hipStream_t s) {
    const int n = c->n;
    const int rows = (int)std::min<long long>(n_ascans, 32);
    const int cls = input_class(c->p.input_type);
    hipLaunchKernelGGL(offset_prepare_kernel, dim3(1), dim3(256), 0, s, d_in, n, rows,
                       (int)input_bytes(c->p.input_type), cls == IN_F32 ? 1 : 0, c->p.bitshift ? 1 : 0, 1,
                       c->d_offset, c->d_fpn, fpn_applied(c) ? 1 : 0, c->d_offset_tab, c->d_ep, n / 2);
    HIP_TRY(c, hipGetLastError());
    c->offset_pending = false;
    return OPE_HIP_OK;
}

int validate_params(ope_hip_ctx* c, const ope_hip_params* p) {
    if (!p) return fail(c, OPE_HIP_EINVAL, "params is null");
    if (!supported_n(p->signal_length))
        return fail(c, OPE_HIP_EUNSUP, "signal_length %d not supported (powers of two 256..4096)",
                    p->signal_length);
    if (p->ascans_per_bscan <= 0 || p->bscans_per_buffer <= 0)
        return fail(c, OPE_HIP_EINVAL, "ascans_per_bscan and bscans_per_buffer must be > 0");
    if (!input_class(p->input_type))
        return fail(c, OPE_HIP_EUNSUP, "input type %d not supported", p->input_type);
    if (p->resampling_enabled && (p->interpolation < 0 || p->interpolation > OPE_HIP_INTERP_LANCZOS4))
        return fail(c, OPE_HIP_EINVAL, "unknown interpolation method %d", p->interpolation);
    if (p->dc_removal_enabled && p->dc_window_size < 0)
        return fail(c, OPE_HIP_EINVAL, "rolling-average window size must be >= 0");
    if (p->fpn_bscan_count < 1) return fail(c, OPE_HIP_EINVAL, "fpn_bscan_count must be >= 1");
    return OPE_HIP_OK;
}

// ---------------------------------------------------------------- launches

// Resident workgroups per CU of `kernel` (at least 1), queried once per
// context and kernel.
hipError_t cached_occupancy(ope_hip_ctx* c, const void* kernel, int wg, int* per_cu) {
    for (const auto& e : c->occupancy)
        if (e.first == kernel) {
            *per_cu = e.second;
            return hipSuccess;
        }
    int n = 0;
    const hipError_t err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, wg, 0);
    if (err != hipSuccess) return err;
    *per_cu = n < 1 ? 1 : n;
    c->occupancy.emplace_back(kernel, *per_cu);
    return hipSuccess;
}

template <int AS>
ChainLaunch pick_kernel_as(int n, int taps, int mode, int iw, int sched) {
    switch (n) {
        case 256: return pick_chain<256, AS>(taps, mode, iw, sched);
        case 512: return pick_chain<512, AS>(taps, mode, iw, sched);
        case 1024: return pick_chain<1024, AS>(taps, mode, iw, sched);
        case 2048: return pick_chain<2048, AS>(taps, mode, iw, sched);
        default: return pick_chain<4096, AS>(taps, mode, iw, sched);
    }
}

ChainLaunch pick_kernel(int n, int taps, int mode, int iw, int as = 1, int sched = -1) {
    return as == 2 ? pick_kernel_as<2>(n, taps, mode, iw, sched) : pick_kernel_as<1>(n, taps, mode, iw, sched);
}

int fft_passes(int n) { return (ilog2(n) + 2) / 3; }

ChainArgs base_args(ope_hip_ctx* c) {
    ChainArgs a{};
    const ope_hip_params& p = c->p;
    const int n = c->n, half = n / 2;
    a.tap_base = c->d_tap_base;
    a.tap_w = c->d_tap_w;
    a.phasor = c->d_phasor;
    a.twiddle = c->d_twiddle;
    a.fpn = c->d_fpn;
    a.bg = c->d_bg;
    a.in_type = input_class(p.input_type);
    a.dc_window = p.dc_window_size;
    int flags = 0;
    if (p.dispersion_enabled && c->phasor_valid) flags |= F_DISP;
    if (p.log_scaling) flags |= F_LOG;
    a.flags = flags;
    // epilogue constants (cpu_kernels.tpp:301-378), folded in double:
    //   log: out = coeff*((10 log10(m2/(N/2)) - min)/R + addend)
    //   lin: out = coeff*((sqrt(m2)/(N/2) - min)/R + addend)
    const double coeff = p.multiplicator, gmin = p.grayscale_min, add = p.addend;
    double range = (double)(float)(p.grayscale_max - p.grayscale_min);
    if (range < 1e-6) range = 1.0;
    const double c10 = 10.0 * std::log10(2.0);
    const double h = std::log2((double)half);
    if (p.log_scaling) {
        a.ep_a = (float)(coeff * c10 / range);
        a.ep_b = (float)(coeff * (add - (c10 * h + gmin) / range));
        if (p.grayscale_min == p.grayscale_max) a.flags |= F_AUTOMM;
    } else {
        a.ep_a = (float)(coeff / (range * (double)half));
        a.ep_b = (float)(coeff * (add - gmin / range));
    }
    a.coeff = p.multiplicator;
    a.addend = p.addend;
    a.log_c = (float)c10;
    a.log_h = (float)h;
    a.bg_w = p.bg_weight;
    a.bg_off = p.bg_offset;
    if (p.bitshift) a.flags |= F_BITSHIFT;
    a.ascans_per_bscan = p.ascans_per_bscan;
    a.acc = c->d_acc;
    return a;
}

// Without dispersion the A-scans are real: two A-scans per lane then share
// one complex FFT (the real and the imaginary part) instead of taking one
// each.  Pairs exist for the throughput mode without DC removal: with the
// DC term removed, the bins within 60 dB of the reflector peak include weak
// ones, whose float32 rounding the shared transform roughly doubles (the
// error bar of 1e-3 dB was exceeded in tests).
bool pairable(const ChainArgs& a, int mode) { return !(a.flags & F_DISP) && (mode & (M_DC | M_STATS)) == 0; }

int variant_mode(const ChainArgs& a, int mode, int as) { return (as == 2 && pairable(a, mode)) ? mode | M_PAIR : mode; }

// Launches one chain variant: `as` A-scans per lane (paired into one FFT
// where pairable), small-radix pass at `sched`, `per_cu` workgroups per CU
// (0: as many as are resident).
int launch_variant(ope_hip_ctx* c, const ChainArgs& a, int mode, int as, int sched, int per_cu_want,
                   long long work_ascans, hipStream_t s) {
    mode = variant_mode(a, mode, as);
    const ChainLaunch l = pick_kernel(c->n, c->taps, mode, (int)input_bytes(c->p.input_type), as, sched);
    if (!l.fn) return fail(c, OPE_HIP_EUNSUP, "chain variant not built (mode %d, %d A-scans per lane)", mode, as);
    int per_cu = 0;
    HIP_TRY(c, cached_occupancy(c, reinterpret_cast<const void*>(l.fn), l.wg, &per_cu));
    if (per_cu_want > 0 && per_cu_want < per_cu) per_cu = per_cu_want;
    long long groups = (work_ascans + l.slots - 1) / l.slots;
    long long grid = std::min<long long>(groups, (long long)per_cu * c->cu_count);
    if (grid < 1) return OPE_HIP_OK;
    hipLaunchKernelGGL(l.fn, dim3((unsigned)grid), dim3(l.wg), 0, s, a);
    HIP_TRY(c, hipGetLastError());
    return OPE_HIP_OK;
}

// The untuned launch: pairs where pairable; the swap schedule where it is
// built (N = 1024, 2048: faster in every configuration measured).
int default_sched(int n) { return (n == 1024 || n == 2048) ? fft_passes(n) : -1; }

int launch_chain(ope_hip_ctx* c, const ChainArgs& a, bool stats, long long work_ascans, hipStream_t s) {
    const int mode = (c->p.dc_removal_enabled ? M_DC : 0) | (stats ? M_STATS : 0);
    const int iw = (int)input_bytes(c->p.input_type);
    if (!stats)
        for (const auto& t : c->tuning)
            if (t.taps == c->taps && t.mode == mode && t.iw == iw)
                return launch_variant(c, a, mode, t.as, t.sched, t.per_cu, work_ascans, s);
    return launch_variant(c, a, mode, pairable(a, mode) ? 2 : 1, default_sched(c->n), 0, work_ascans, s);
}

int zero_acc(ope_hip_ctx* c, hipStream_t s) {
    const long long n = (long long)kFpnSegments * (c->n / 2) * 3;
    hipLaunchKernelGGL(fill_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, c->d_acc, n);
    HIP_TRY(c, hipGetLastError());
    return OPE_HIP_OK;
}

// Positions of the sinusoidal scan correction for the context's B-scan
// length (fillSinusoidalScanCorrectionCurve, cuda_kernels.cu:433-437:
// acos in float, the product in double).
int ensure_sin_curve(ope_hip_ctx* c) {
    const int a = c->p.ascans_per_bscan;
    if (c->d_sin_curve && c->sin_curve_len == a) return OPE_HIP_OK;
    std::vector<float> curve(a);
    for (int k = 0; k < a; ++k) {
        const float arg = (float)(1.0 - (2.0 * (double)(float)k) / (double)(float)a);
        curve[k] = (float)(((double)(float)a / M_PI) * (double)std::acos(arg));
    }
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->d_sin_curve) HIP_TRY(c, hipFree(c->d_sin_curve));
    c->d_sin_curve = nullptr;
    HIP_TRY(c, hipMalloc(&c->d_sin_curve, a * sizeof(float)));
    HIP_TRY(c, hipMemcpy(c->d_sin_curve, curve.data(), a * sizeof(float), hipMemcpyHostToDevice));
    c->sin_curve_len = a;
    return OPE_HIP_OK;
}

// The chain's output before the sinusoidal correction: n_ascans rows.
int ensure_sin_tmp(ope_hip_ctx* c, long long n_ascans) {
    const size_t need = (size_t)n_ascans * (c->n / 2);
    if (need <= c->sin_tmp_cap) return OPE_HIP_OK;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipDeviceSynchronize());  // frames on any stream may still read it
    if (c->d_sin_tmp) HIP_TRY(c, hipFree(c->d_sin_tmp));
    c->d_sin_tmp = nullptr;
    c->sin_tmp_cap = 0;
    HIP_TRY(c, hipMalloc(&c->d_sin_tmp, need * sizeof(float)));
    c->sin_tmp_cap = need;
    return OPE_HIP_OK;
}

int launch_sinusoidal(ope_hip_ctx* c, const float* in, float* out, long long n_ascans, bool use_bg, hipStream_t s) {
    const int half = c->n / 2;
    const long long total = n_ascans * half;
    hipLaunchKernelGGL(sinusoidal_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, total, half,
                       c->p.ascans_per_bscan, c->d_sin_curve, c->d_bg, c->p.bg_weight, c->p.bg_offset,
                       use_bg ? 1 : 0);
    HIP_TRY(c, hipGetLastError());
    return OPE_HIP_OK;
}

// One frame of the chain on stream s (cpu_backend.cpp:169-357).
int run_frame(ope_hip_ctx* c, const void* d_in, float* d_out, long long n_ascans, hipStream_t s) {
    if (n_ascans <= 0) return OPE_HIP_OK;
    const ope_hip_params& p = c->p;
    const int half = c->n / 2;
    ChainArgs a = base_args(c);
    a.in = d_in;
    a.out = d_out;
    a.n_ascans = n_ascans;
    if (offset_active(c)) {
        if (c->offset_pending) {
            int rc = determine_offset(c, d_in, n_ascans, s);
            if (rc) return rc;
        }
        apply_offset(c, a);
    }

    // FPN accumulation: only the first `required` accumulated A-scans ever
    // enter a determination (cpu_backend.cpp:324-330), and of those only
    // the 8 full segments (cpu_kernels.tpp:414-440).
    const long long required = (long long)p.fpn_bscan_count * p.ascans_per_bscan;
    const int seg_w = (int)(required / kFpnSegments);
    const long long rows_used = (long long)seg_w * kFpnSegments;
    if (seg_w > 0 && c->fpn_count < rows_used) {
        ChainArgs st = a;
        st.flags = a.flags & (F_DISP | F_OFFSET);
        st.row0 = c->fpn_count;
        st.seg_w = seg_w;
        st.stat_rows = (int)std::min<long long>(n_ascans, rows_used - c->fpn_count);
        int rc = launch_chain(c, st, true, st.stat_rows, s);
        if (rc) return rc;
    }
    c->fpn_count += n_ascans;
    if ((c->fpn_requested || p.fpn_continuous) && c->fpn_count >= required) {
        if (seg_w > 0) {
            // with an offset mode the epilogue constant follows the profile
            const bool ep = c->offset_mode != OPE_HIP_OFFSET_OFF;
            hipLaunchKernelGGL(fpn_select_kernel, dim3((half + 255) / 256), dim3(256), 0, s, c->d_acc, half,
                               kFpnSegments, seg_w, c->d_fpn, c->d_offset, c->d_offset_tab, p.fpn_removal ? 1 : 0,
                               ep ? c->d_ep : nullptr);
            HIP_TRY(c, hipGetLastError());
        } else {
            // fewer than 8 rows per segment: the reference divides by zero;
            // its profile is then NaN.
            std::vector<float2> nanv(half, make_float2(NAN, NAN));
            HIP_TRY(c, hipMemcpyAsync(c->d_fpn, nanv.data(), half * sizeof(float2), hipMemcpyHostToDevice, s));
            HIP_TRY(c, hipStreamSynchronize(s));
            if (c->offset_mode != OPE_HIP_OFFSET_OFF) {
                hipLaunchKernelGGL(offset_prepare_kernel, dim3(1), dim3(256), 0, s, nullptr, c->n, 0, 2, 0, 0, 0,
                                   c->d_offset, c->d_fpn, p.fpn_removal ? 1 : 0, c->d_offset_tab, c->d_ep, half);
                HIP_TRY(c, hipGetLastError());
            }
        }
        c->have_fpn = true;
        c->fpn_pairs_stored = half;
        if (p.fpn_continuous) {
            c->fpn_count = std::min(c->fpn_count, required);
        } else {
            c->fpn_count = 0;
            int rc = zero_acc(c, s);
            if (rc) return rc;
        }
        c->fpn_requested = false;
    }
    if (p.fpn_removal && c->have_fpn) a.flags |= F_FPN;
    if (p.bscan_flip) {
        a.flags |= F_FLIP;
        a.flip_bscans = n_ascans / p.ascans_per_bscan;
    }

    // CUDA order after the grayscale map: B-scan flip (in the chain's
    // epilogue), sinusoidal correction, background (cuda_backend.cu:689-743)
    const bool sinus = p.sinusoidal_correction != 0;
    const bool record_bg = p.bg_removal && c->bg_requested;
    const bool subtract_bg = p.bg_removal && c->have_bg && !record_bg;
    if (subtract_bg && !sinus) a.flags |= F_BG;
    if (sinus) {
        int rc = ensure_sin_curve(c);
        if (!rc) rc = ensure_sin_tmp(c, n_ascans);
        if (rc) return rc;
        a.out = c->d_sin_tmp;
    }
    int rc = launch_chain(c, a, false, n_ascans, s);
    if (rc) return rc;
    if (sinus) {
        rc = launch_sinusoidal(c, c->d_sin_tmp, d_out, n_ascans, subtract_bg, s);
        if (rc) return rc;
    }
    if (record_bg) {
        hipLaunchKernelGGL(bg_record_kernel, dim3((half + 255) / 256), dim3(256), 0, s, d_out, n_ascans, half,
                           c->d_bg);
        HIP_TRY(c, hipGetLastError());
        const long long total = n_ascans * half;
        hipLaunchKernelGGL(bg_subtract_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d_out,
                           total, half, c->d_bg, p.bg_weight, p.bg_offset);
        HIP_TRY(c, hipGetLastError());
        c->have_bg = true;
        c->bg_requested = false;
    }
    return OPE_HIP_OK;
}

void free_ctx(ope_hip_ctx* c) {
    if (!c) return;
    // teardown: errors are ignored, every resource is released regardless
    (void)hipSetDevice(c->device);
    for (hipStream_t s : {c->s_comp, c->s_in, c->s_out})
        if (s) (void)hipStreamSynchronize(s);
    for (auto& s : c->slots) {
        if (s.d_in) (void)hipFree(s.d_in);
        if (s.d_out) (void)hipFree(s.d_out);
        if (s.h2d_done) (void)hipEventDestroy(s.h2d_done);
        if (s.comp_done) (void)hipEventDestroy(s.comp_done);
        if (s.d2h_done) (void)hipEventDestroy(s.d2h_done);
    }
    for (void* p : {(void*)c->d_tap_base, (void*)c->d_tap_w, (void*)c->d_phasor, (void*)c->d_twiddle,
                    (void*)c->d_fpn, (void*)c->d_bg, (void*)c->d_acc, (void*)c->d_af_window,
                    (void*)c->d_af_acc, (void*)c->d_sin_curve, (void*)c->d_sin_tmp, (void*)c->d_offset,
                    (void*)c->d_offset_tab, (void*)c->d_ep})
        if (p) (void)hipFree(p);
    for (hipStream_t s : {c->s_in, c->s_comp, c->s_out})
        if (s) (void)hipStreamDestroy(s);
    delete c;
}

void host_input_free_cb(void* u) {
    HostCall* hc = static_cast<HostCall*>(u);
    if (hc->free_fn) hc->free_fn(hc->free_user);
}

void host_done_cb(void* u) {
    HostCall* hc = static_cast<HostCall*>(u);
    if (hc->fn) hc->fn(hc->user, hc->out);
}

int sync_all(ope_hip_ctx* c) {
    HIP_TRY(c, hipSetDevice(c->device));
    // frames on caller streams read the tables too
    if (c->foreign_work.exchange(false)) HIP_TRY(c, hipDeviceSynchronize());
    HIP_TRY(c, hipStreamSynchronize(c->s_in));
    HIP_TRY(c, hipStreamSynchronize(c->s_comp));
    HIP_TRY(c, hipStreamSynchronize(c->s_out));
    return OPE_HIP_OK;
}

AutofocusLaunch pick_autofocus(int n, int taps, int iw, bool dc, bool bwd) {
    switch (n) {
        case 256: return pick_autofocus_n<256>(taps, iw, dc, bwd);
        case 512: return pick_autofocus_n<512>(taps, iw, dc, bwd);
        case 1024: return pick_autofocus_n<1024>(taps, iw, dc, bwd);
        case 2048: return pick_autofocus_n<2048>(taps, iw, dc, bwd);
        default: return pick_autofocus_n<4096>(taps, iw, dc, bwd);
    }
}

// One pass of the differentiable chain over d_in; copies the `n_out` FP64
// sums to host_out and waits for them.  theta = (d2, d3, c0, c1, c2, c3).
int autofocus_pass(ope_hip_ctx* c, const void* d_in, long long n_ascans, const double* theta, bool bwd,
                   float inv_s, float log_inv_s_plus_l, hipStream_t s, double* host_out) {
    const ope_hip_params& p = c->p;
    if (!p.resampling_enabled || (p.interpolation != OPE_HIP_INTERP_LINEAR && p.interpolation != OPE_HIP_INTERP_CUBIC))
        return fail(c, OPE_HIP_EUNSUP, "autofocus needs linear or cubic resampling enabled");
    const int iw = (int)input_bytes(p.input_type);
    if (iw == 1) return fail(c, OPE_HIP_EUNSUP, "autofocus supports 16- and 32-bit input");
    if (n_ascans <= 0 || !d_in) return fail(c, OPE_HIP_EINVAL, "no input");
    if (reinterpret_cast<uintptr_t>(d_in) % 16) return fail(c, OPE_HIP_EINVAL, "d_in must be 16-byte aligned");
    if (!theta) return fail(c, OPE_HIP_EINVAL, "theta is null");
    const int n = c->n;
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->d_af_window) HIP_TRY(c, hipMalloc(&c->d_af_window, n * sizeof(float)));
    if (!c->d_af_acc) HIP_TRY(c, hipMalloc(&c->d_af_acc, 8 * sizeof(double)));
    std::vector<float> win(n, 1.0f);
    if (p.windowing_enabled) win = c->curve_window;
    HIP_TRY(c, hipMemcpyAsync(c->d_af_window, win.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(fill_zero_f64, dim3(1), dim3(64), 0, s, c->d_af_acc, 8LL);
    HIP_TRY(c, hipGetLastError());
    AutofocusArgs a{};
    a.in = d_in;
    a.n_ascans = n_ascans;
    a.window = c->d_af_window;
    a.twiddle = c->d_twiddle;
    a.in_type = input_class(p.input_type);
    a.bitshift = p.bitshift;
    a.dc_window = p.dc_window_size;
    a.d2 = (float)theta[0];
    a.d3 = (float)theta[1];
    a.c0 = theta[2];
    a.c1 = theta[3];
    a.c2 = theta[4];
    a.c3 = theta[5];
    a.inv_s = inv_s;
    a.log_inv_s_plus_l = log_inv_s_plus_l;
    a.acc = c->d_af_acc;
    const AutofocusLaunch l = pick_autofocus(n, c->taps, iw, p.dc_removal_enabled != 0, bwd);
    int per_cu = 0;
    HIP_TRY(c, cached_occupancy(c, reinterpret_cast<const void*>(l.fn), l.wg, &per_cu));
    const long long groups = (n_ascans + l.slots - 1) / l.slots;
    const long long grid = std::min<long long>(groups, (long long)per_cu * c->cu_count);
    hipLaunchKernelGGL(l.fn, dim3((unsigned)grid), dim3(l.wg), 0, s, a);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(host_out, c->d_af_acc, (bwd ? 6 : 2) * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return OPE_HIP_OK;
}

void free_ragged(ope_hip_ragged* r) {
    if (!r) return;
    (void)hipSetDevice(r->device);
    for (const auto& k : r->classes)
        if (k.d_tickets) (void)hipFree(k.d_tickets);
    for (hipStream_t q : r->side) (void)hipStreamDestroy(q);
    for (hipEvent_t e : r->joins) (void)hipEventDestroy(e);
    if (r->fork) (void)hipEventDestroy(r->fork);
    for (void* q : {(void*)r->d_rows, (void*)r->d_finish, (void*)r->d_start, (void*)r->d_discard})
        if (q) (void)hipFree(q);
    delete r;
}

int ragged_fail(ope_hip_ragged* r, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    r->err = buf;
    return code;
}

// Chain arguments of a context for frames that need no FPN statistics or
// background recording (the flags run_frame sets in that state).
ChainArgs steady_args(ope_hip_ctx* c) {
    ChainArgs a = base_args(c);
    if (c->p.fpn_removal && c->have_fpn) a.flags |= F_FPN;
    if (c->p.bg_removal && c->have_bg) a.flags |= F_BG;
    if (offset_active(c) && !c->offset_pending) apply_offset(c, a);
    return a;
}

int build_ragged(ope_hip_ragged* r, ope_hip_ctx* const* ctxs, int n_ctx, const ope_hip_request* reqs,
                 int n_reqs) {
    if (!ctxs || n_ctx < 1 || n_ctx > 5) return ragged_fail(r, OPE_HIP_EINVAL, "n_ctx must be in 1..5");
    if (n_reqs < 0 || (n_reqs > 0 && !reqs)) return ragged_fail(r, OPE_HIP_EINVAL, "bad request array");
    for (int i = 0; i < n_ctx; ++i) {
        ope_hip_ctx* c = ctxs[i];
        if (!c) return ragged_fail(r, OPE_HIP_EINVAL, "context %d is null", i);
        const ope_hip_ctx* c0 = ctxs[0];
        if (c->device != c0->device || c->taps != c0->taps ||
            c->p.dc_removal_enabled != c0->p.dc_removal_enabled ||
            input_bytes(c->p.input_type) != input_bytes(c0->p.input_type) ||
            input_class(c->p.input_type) != input_class(c0->p.input_type))
            return ragged_fail(r, OPE_HIP_EINVAL,
                               "contexts must share device, interpolation, DC removal and input type");
        for (int j = 0; j < i; ++j)
            if (ctxs[j]->n == c->n) return ragged_fail(r, OPE_HIP_EINVAL, "two contexts of length %d", c->n);
        if ((c->fpn_requested || c->p.fpn_continuous) && c->p.fpn_removal)
            return ragged_fail(r, OPE_HIP_ESTATE, "context of length %d has an FPN determination pending", c->n);
        if (c->bg_requested && c->p.bg_removal)
            return ragged_fail(r, OPE_HIP_ESTATE, "context of length %d has a background recording pending",
                               c->n);
        if (c->p.bscan_flip || c->p.sinusoidal_correction)
            return ragged_fail(r, OPE_HIP_EUNSUP,
                               "B-scan flip and sinusoidal correction act on frames, not on ragged requests");
    }
    const ope_hip_ctx* c0 = ctxs[0];
    r->device = c0->device;
    r->ctxs.assign(ctxs, ctxs + n_ctx);
    r->n_reqs = n_reqs;
    r->req_ascans.assign(n_reqs, 0);
    // requests by length class, in request order; classes in the order of
    // their first request
    std::vector<int> req_cls(n_reqs, -1), order;
    std::vector<long long> cls_ascans(n_ctx, 0);
    // Without dispersion and DC removal two A-scans share one FFT (rows 2k
    // and 2k+1 of a launch).  Each request of such a length then starts on
    // an even row, and one of odd count ends with its last A-scan again,
    // written to a discard row: its A-scans pair as in a frame launch of
    // the request alone (whose last odd A-scan pairs with itself), so the
    // outputs agree bit for bit.
    std::vector<char> paired(n_ctx, 0);
    for (int i = 0; i < n_ctx; ++i)
        paired[i] = pairable(steady_args(ctxs[i]), ctxs[i]->p.dc_removal_enabled ? M_DC : 0);
    for (int q = 0; q < n_reqs; ++q) {
        const ope_hip_request& rq = reqs[q];
        int cls = -1;
        for (int i = 0; i < n_ctx; ++i)
            if (ctxs[i]->n == rq.signal_length) cls = i;
        if (cls < 0) return ragged_fail(r, OPE_HIP_EINVAL, "request %d: no context of length %d", q, rq.signal_length);
        if (rq.n_ascans < 0) return ragged_fail(r, OPE_HIP_EINVAL, "request %d: n_ascans < 0", q);
        if (rq.n_ascans > 0 && (!rq.d_in || !rq.d_out))
            return ragged_fail(r, OPE_HIP_EINVAL, "request %d: null device pointer", q);
        if (reinterpret_cast<uintptr_t>(rq.d_in) % 16 || reinterpret_cast<uintptr_t>(rq.d_out) % 4)
            return ragged_fail(r, OPE_HIP_EINVAL, "request %d: d_in must be 16-byte and d_out 4-byte aligned", q);
        req_cls[q] = cls;
        r->req_ascans[q] = rq.n_ascans;
        if (rq.n_ascans > 0 && cls_ascans[cls] == 0) order.push_back(cls);
        cls_ascans[cls] += rq.n_ascans + (paired[cls] ? (rq.n_ascans & 1) : 0);
    }
    // the row map: each class's A-scans contiguous
    long long total = 0;
    std::vector<long long> cls_first(n_ctx, 0);
    for (int cls : order) {
        cls_first[cls] = total;
        total += cls_ascans[cls];
    }
    std::vector<RowRef> rows((size_t)total);
    std::vector<long long> fill(cls_first);
    std::vector<size_t> pads;  // rows whose output is the discard row
    int pad_n = 0;
    // within a length, the shortest requests first (the launches work
    // through the row map roughly in order, so this minimises the mean
    // request latency; the outputs do not depend on the order)
    std::vector<int> by_size(n_reqs);
    for (int q = 0; q < n_reqs; ++q) by_size[q] = q;
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](int x, int y) { return reqs[x].n_ascans < reqs[y].n_ascans; });
    for (int q : by_size) {
        const ope_hip_request& rq = reqs[q];
        const int cls = req_cls[q];
        const size_t iw = input_bytes(ctxs[cls]->p.input_type);
        for (long long a = 0; a < rq.n_ascans; ++a) {
            RowRef& e = rows[(size_t)fill[cls]++];
            e.in = static_cast<const char*>(rq.d_in) + (size_t)a * rq.signal_length * iw;
            e.out = rq.d_out + (size_t)a * (rq.signal_length / 2);
            e.req = q;
        }
        if (paired[cls] && (rq.n_ascans & 1)) {
            const size_t i = (size_t)fill[cls]++;
            rows[i] = rows[i - 1];
            pads.push_back(i);
            pad_n = std::max(pad_n, rq.signal_length);
        }
    }

#define RAGGED_TRY(expr)                                                                         \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return ragged_fail(r, e_ == hipErrorOutOfMemory ? OPE_HIP_ENOMEM : OPE_HIP_ERUNTIME,  \
                               "%s: %s", #expr, hipGetErrorString(e_));                          \
    } while (0)
    RAGGED_TRY(hipSetDevice(r->device));
    const size_t nq = (size_t)std::max(n_reqs, 1);
    RAGGED_TRY(hipMalloc(&r->d_rows, std::max<size_t>(rows.size(), 1) * sizeof(RowRef)));
    RAGGED_TRY(hipMalloc(&r->d_finish, nq * sizeof(unsigned long long)));
    RAGGED_TRY(hipMalloc(&r->d_start, sizeof(unsigned long long)));
    if (!pads.empty()) {
        RAGGED_TRY(hipMalloc(&r->d_discard, (size_t)(pad_n / 2) * sizeof(float)));
        for (size_t i : pads) rows[i].out = r->d_discard;
    }
    if (!rows.empty())
        RAGGED_TRY(hipMemcpy(r->d_rows, rows.data(), rows.size() * sizeof(RowRef), hipMemcpyHostToDevice));
    std::vector<unsigned long long> zt(nq, 0ull);
    RAGGED_TRY(hipMemcpy(r->d_finish, zt.data(), nq * sizeof(unsigned long long), hipMemcpyHostToDevice));
    RAGGED_TRY(hipMemcpy(r->d_start, zt.data(), sizeof(unsigned long long), hipMemcpyHostToDevice));
    for (int cls : order) {
        ope_hip_ragged::Class k{ctxs[cls], steady_args(ctxs[cls]), cls_ascans[cls]};
        k.args.in = nullptr;
        k.args.out = nullptr;
        k.args.n_ascans = k.ascans;
        k.args.rows = r->d_rows + cls_first[cls];
        k.args.finish = r->d_finish;
        r->classes.push_back(k);
    }
    // the largest length first: its launch takes the CUs first, and the
    // smaller lengths' workgroups fill the room it leaves on each CU
    std::stable_sort(r->classes.begin(), r->classes.end(),
                     [](const ope_hip_ragged::Class& x, const ope_hip_ragged::Class& y) { return x.ctx->n > y.ctx->n; });
    // the launches: the largest length on the caller's stream with as many
    // workgroups as are resident, striding over its A-scans; the second
    // largest twice on one dynamic schedule: beside it on a side stream with
    // one workgroup per CU (so that whichever the GPU dispatches first, each
    // CU keeps room for a workgroup of the largest length), and after it on
    // the caller's stream with as many workgroups as are resident, which
    // take what the side launch has left; the others after those, striding.
    // (Side launches of the third length too, after the second's, slowed
    // the largest length's launch more than they gained: 1.73-1.76 ms
    // against 1.61-1.65 ms for the 512-request batch.)
    for (size_t i = 0; i < r->classes.size(); ++i) {
        auto& k = r->classes[i];
        const int mode0 = M_ROWS | (k.ctx->p.dc_removal_enabled ? M_DC : 0);
        const int as = pairable(k.args, mode0) ? 2 : 1;
        k.launch = pick_kernel(k.ctx->n, k.ctx->taps, variant_mode(k.args, mode0, as),
                               (int)input_bytes(k.ctx->p.input_type), as, default_sched(k.ctx->n));
        if (!k.launch.fn) return ragged_fail(r, OPE_HIP_EUNSUP, "length %d: row-map kernel not built", k.ctx->n);
        int per_cu = 0;
        RAGGED_TRY(cached_occupancy(k.ctx, reinterpret_cast<const void*>(k.launch.fn), k.launch.wg, &per_cu));
        const long long groups = (k.ascans + k.launch.slots - 1) / k.launch.slots;
        k.grid_main = std::min<long long>(groups, (long long)per_cu * k.ctx->cu_count);
        if (i == 1) {
            k.grid_side = std::min<long long>(groups, k.ctx->cu_count);
            // One side workgroup per CU whichever launch the GPU dispatches
            // first: padded with unused LDS to more than half a CU's, two
            // cannot share a CU (two there would leave no room for a
            // workgroup of the largest length), while one still fits beside
            // the largest length's resident workgroups.  Where those leave
            // half a CU or less, no pad.
            const auto& big = r->classes[0];
            int lds_cu = 0, big_cu = 0;
            hipFuncAttributes fa_big{}, fa_side{};
            RAGGED_TRY(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                             r->device));
            RAGGED_TRY(cached_occupancy(big.ctx, reinterpret_cast<const void*>(big.launch.fn), big.launch.wg, &big_cu));
            RAGGED_TRY(hipFuncGetAttributes(&fa_big, reinterpret_cast<const void*>(big.launch.fn)));
            RAGGED_TRY(hipFuncGetAttributes(&fa_side, reinterpret_cast<const void*>(k.launch.fn)));
            const long long room = (long long)lds_cu - (long long)big_cu * (long long)fa_big.sharedSizeBytes;
            const long long half = lds_cu / 2;
            if (room > half && room >= (long long)fa_side.sharedSizeBytes) {
                const long long want = half + (room - half) / 2;  // between half a CU and the room left
                if (want > (long long)fa_side.sharedSizeBytes)
                    k.side_lds = (unsigned)(want - (long long)fa_side.sharedSizeBytes);
            }
            RAGGED_TRY(hipMalloc(&k.d_tickets, 2 * sizeof(unsigned)));
            RAGGED_TRY(hipMemset(k.d_tickets, 0, 2 * sizeof(unsigned)));
            k.args.tickets = k.d_tickets;
            k.args.ticket_wgs = (unsigned)(k.grid_side + k.grid_main);
        }
    }
    RAGGED_TRY(hipEventCreateWithFlags(&r->fork, hipEventDisableTiming));
    if (r->classes.size() > 1) {
        hipStream_t q = nullptr;
        hipEvent_t e = nullptr;
        RAGGED_TRY(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
        r->side.push_back(q);
        RAGGED_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        r->joins.push_back(e);
    }
    int khz = 0;
    RAGGED_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, r->device));
    r->tick_hz = (double)khz * 1e3;
#undef RAGGED_TRY
    return OPE_HIP_OK;
}

// ---- individual operations: device scratch per call, synchronous

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

#define OP_TRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(c, OPE_HIP_ERUNTIME, "%s: %s", #expr, hipGetErrorString(e_));         \
    } while (0)

int op_scale(ope_hip_ctx* c, const float* input, const float* s, int stride, int line_width, int samples, float* out) {
    if (!c || !input || !s || !out || line_width < 1 || samples < 0 || samples > line_width) return OPE_HIP_EINVAL;
    if (samples == 0) return OPE_HIP_OK;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    OP_TRY(hipSetDevice(c->device));
    DevBuf dx, ds, dy;
    OP_TRY(hipMalloc(&dx.p, samples * sizeof(float)));
    OP_TRY(hipMalloc(&ds.p, (size_t)samples * stride * sizeof(float)));
    OP_TRY(hipMalloc(&dy.p, samples * sizeof(float)));
    OP_TRY(hipMemcpy(dx.p, input, samples * sizeof(float), hipMemcpyHostToDevice));
    OP_TRY(hipMemcpy(ds.p, s, (size_t)samples * stride * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(op_scale_kernel, dim3((samples + 255) / 256), dim3(256), 0, c->s_comp,
                       static_cast<const float*>(dx.p), static_cast<const float*>(ds.p), stride, samples,
                       static_cast<float*>(dy.p));
    OP_TRY(hipGetLastError());
    OP_TRY(hipStreamSynchronize(c->s_comp));
    OP_TRY(hipMemcpy(out, dy.p, samples * sizeof(float), hipMemcpyDeviceToHost));
    return OPE_HIP_OK;
}

template <int N>
void launch_ifft_mag(const float* in, long long lines, const float2* tw, float* out, hipStream_t s) {
    using L = ChainLds<N, 0, 4, 1, Geo<N>::WG>;
    hipLaunchKernelGGL(op_ifft_mag_kernel<N>, dim3((unsigned)((lines + L::SLOTS - 1) / L::SLOTS)), dim3(Geo<N>::WG),
                       0, s, in, lines, tw, out);
}

}  // namespace

}  // namespace

// ====================================================================== C ABI

extern "C" {

const char* ope_hip_version(void) { return "octproengine_amd libope_hip 0.1 (gfx950)"; }

int ope_hip_device_count(int* count) {
    if (!count) return OPE_HIP_EINVAL;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return OPE_HIP_OK;
}

const char* ope_hip_last_error(const ope_hip_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int ope_hip_create(int device, const ope_hip_params* params, int n_slots, ope_hip_ctx** out_ctx) {
    g_create_err.clear();
    if (!out_ctx) {
        g_create_err = "out_ctx is null";
        return OPE_HIP_EINVAL;
    }
    *out_ctx = nullptr;
    ope_hip_ctx* c = new (std::nothrow) ope_hip_ctx();
    if (!c) {
        g_create_err = "out of host memory";
        return OPE_HIP_ENOMEM;
    }
    int rc = validate_params(c, params);
    if (rc) {
        g_create_err = c->err;
        delete c;
        return rc;
    }
    if (n_slots < 1 || n_slots > 8) {
        g_create_err = "n_slots must be in 1..8";
        delete c;
        return OPE_HIP_EINVAL;
    }
    c->device = device;
    c->p = *params;
    c->n = params->signal_length;
    c->taps = taps_for(c->p);
    auto bail = [&](int code) {
        g_create_err = c->err;
        free_ctx(c);
        return code;
    };
#define CREATE_TRY(expr)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            fail(c, e_ == hipErrorOutOfMemory ? OPE_HIP_ENOMEM : OPE_HIP_ERUNTIME, "%s: %s", #expr, \
                 hipGetErrorString(e_));                                                        \
            return bail(e_ == hipErrorOutOfMemory ? OPE_HIP_ENOMEM : OPE_HIP_ERUNTIME);         \
        }                                                                                       \
    } while (0)
    CREATE_TRY(hipSetDevice(device));
    CREATE_TRY(scratch_init(device));
    hipDeviceProp_t prop;
    CREATE_TRY(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    CREATE_TRY(hipStreamCreateWithFlags(&c->s_in, hipStreamNonBlocking));
    CREATE_TRY(hipStreamCreateWithFlags(&c->s_comp, hipStreamNonBlocking));
    CREATE_TRY(hipStreamCreateWithFlags(&c->s_out, hipStreamNonBlocking));
    const int n = c->n, half = n / 2;
    CREATE_TRY(hipMalloc(&c->d_tap_base, n * sizeof(int)));
    CREATE_TRY(hipMalloc(&c->d_tap_w, (size_t)n * 16 * sizeof(float)));
    c->tap_w_cap = (size_t)n * 16;
    CREATE_TRY(hipMalloc(&c->d_phasor, n * sizeof(float2)));
    CREATE_TRY(hipMalloc(&c->d_twiddle, n * sizeof(float2)));
    CREATE_TRY(hipMalloc(&c->d_fpn, half * sizeof(float2)));
    CREATE_TRY(hipMalloc(&c->d_bg, half * sizeof(float)));
    CREATE_TRY(hipMalloc(&c->d_acc, (size_t)kFpnSegments * half * 3 * sizeof(double)));
    CREATE_TRY(hipMalloc(&c->d_offset, sizeof(float)));
    CREATE_TRY(hipMalloc(&c->d_offset_tab, half * sizeof(double2)));
    CREATE_TRY(hipMalloc(&c->d_ep, half * sizeof(float2)));
    CREATE_TRY(hipMemset(c->d_offset, 0, sizeof(float)));
    CREATE_TRY(hipMemset(c->d_offset_tab, 0, half * sizeof(double2)));
    CREATE_TRY(hipMemset(c->d_ep, 0, half * sizeof(float2)));
    {
        std::vector<float2> tw(n);
        for (int k = 0; k < n; ++k) {
            const double ang = 2.0 * M_PI * (double)k / (double)n;
            tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
        CREATE_TRY(hipMemcpy(c->d_twiddle, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice));
        std::vector<float2> zero2(half, make_float2(0.f, 0.f));
        CREATE_TRY(hipMemcpy(c->d_fpn, zero2.data(), half * sizeof(float2), hipMemcpyHostToDevice));
        std::vector<float> zero1(half, 0.f);
        CREATE_TRY(hipMemcpy(c->d_bg, zero1.data(), half * sizeof(float), hipMemcpyHostToDevice));
        std::vector<double> zacc((size_t)kFpnSegments * half * 3, 0.0);
        CREATE_TRY(hipMemcpy(c->d_acc, zacc.data(), zacc.size() * sizeof(double), hipMemcpyHostToDevice));
        std::vector<float2> ph(n, make_float2(1.f, 0.f));
        CREATE_TRY(hipMemcpy(c->d_phasor, ph.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    }
    const size_t frame_ascans = (size_t)params->ascans_per_bscan * params->bscans_per_buffer;
    c->slots.resize(n_slots);
    c->calls.resize(n_slots);
    for (auto& s : c->slots) {
        CREATE_TRY(hipMalloc(&s.d_in, frame_ascans * n * input_bytes(params->input_type)));
        CREATE_TRY(hipMalloc(&s.d_out, frame_ascans * half * sizeof(float)));
        CREATE_TRY(hipMemset(s.d_in, 0, frame_ascans * n * input_bytes(params->input_type)));
        CREATE_TRY(hipEventCreateWithFlags(&s.h2d_done, hipEventDisableTiming));
        CREATE_TRY(hipEventCreateWithFlags(&s.comp_done, hipEventDisableTiming));
        CREATE_TRY(hipEventCreateWithFlags(&s.d2h_done, hipEventDisableTiming));
    }
#undef CREATE_TRY
    // Curves start as zeros, as the CPU backend's do after initialize
    // (cpu_backend.cpp:423-425), until the Processor uploads real ones.
    c->curve_resample.assign(n, 0.f);
    c->curve_window.assign(n, 0.f);
    c->curve_phasor.assign(2 * (size_t)n, 0.f);
    c->phasor_valid = true;
    {
        hipError_t e = hipMemcpy(c->d_phasor, c->curve_phasor.data(), 2 * (size_t)n * sizeof(float),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            fail(c, OPE_HIP_ERUNTIME, "hipMemcpy phasor: %s", hipGetErrorString(e));
            return bail(OPE_HIP_ERUNTIME);
        }
    }
    rc = upload_taps(c);
    if (!rc) rc = refresh_offset(c, true);
    if (rc) return bail(rc);
    *out_ctx = c;
    return OPE_HIP_OK;
}

int ope_hip_destroy(ope_hip_ctx* ctx) {
    free_ctx(ctx);
    return OPE_HIP_OK;
}

int ope_hip_set_params(ope_hip_ctx* c, const ope_hip_params* params) {
    if (!c) return OPE_HIP_EINVAL;
    int rc = validate_params(c, params);
    if (rc) return rc;
    if (params->signal_length != c->p.signal_length || params->ascans_per_bscan != c->p.ascans_per_bscan ||
        params->bscans_per_buffer != c->p.bscans_per_buffer || params->input_type != c->p.input_type)
        return fail(c, OPE_HIP_EINVAL, "set_params cannot change dimensions or input type");
    std::lock_guard<std::mutex> lk(c->submit_mu);
    rc = sync_all(c);
    if (rc) return rc;
    ope_hip_params old = c->p;
    c->p = *params;
    if (old.resampling_enabled != params->resampling_enabled || old.interpolation != params->interpolation ||
        old.windowing_enabled != params->windowing_enabled) {
        rc = upload_taps(c);
        if (rc) return rc;
    }
    // an automatic offset belongs to the samples as converted
    if (old.bitshift != params->bitshift && c->offset_mode == OPE_HIP_OFFSET_AUTO) c->offset_pending = true;
    // the dispersion flag and the FPN switch enter the offset's tables
    return refresh_offset(c, true);
}

int ope_hip_set_resampling_curve(ope_hip_ctx* c, const float* curve, size_t length) {
    if (!c) return OPE_HIP_EINVAL;
    if (!curve && length > 0) return fail(c, OPE_HIP_EINVAL, "Invalid resampling curve pointer");
    if (length != (size_t)c->n)
        return fail(c, OPE_HIP_EINVAL, "resampling curve has %zu values, expected %d", length, c->n);
    std::lock_guard<std::mutex> lk(c->submit_mu);
    c->curve_resample.assign(curve, curve + length);
    int rc = sync_all(c);
    if (!rc) rc = upload_taps(c);
    if (!rc) rc = refresh_offset(c, true);
    return rc;
}

int ope_hip_set_window_curve(ope_hip_ctx* c, const float* curve, size_t length) {
    if (!c) return OPE_HIP_EINVAL;
    if (!curve && length > 0) return fail(c, OPE_HIP_EINVAL, "Invalid window curve pointer");
    if (length != (size_t)c->n)
        return fail(c, OPE_HIP_EINVAL, "window curve has %zu values, expected %d", length, c->n);
    std::lock_guard<std::mutex> lk(c->submit_mu);
    c->curve_window.assign(curve, curve + length);
    int rc = sync_all(c);
    if (!rc) rc = upload_taps(c);
    if (!rc) rc = refresh_offset(c, true);
    return rc;
}

int ope_hip_set_dispersion_curve(ope_hip_ctx* c, const float* cos_sin, size_t length) {
    if (!c) return OPE_HIP_EINVAL;
    if (!cos_sin && length > 0) return fail(c, OPE_HIP_EINVAL, "Invalid dispersion curve pointer");
    std::lock_guard<std::mutex> lk(c->submit_mu);
    int rc = sync_all(c);
    if (rc) return rc;
    c->curve_phasor.assign(cos_sin, cos_sin + length);
    c->phasor_valid = (length == 2 * (size_t)c->n);
    if (c->phasor_valid)
        HIP_TRY(c, hipMemcpy(c->d_phasor, cos_sin, length * sizeof(float), hipMemcpyHostToDevice));
    return refresh_offset(c, true);
}

int ope_hip_set_fpn_profile(ope_hip_ctx* c, const float* interleaved, size_t pairs) {
    if (!c) return OPE_HIP_EINVAL;
    if (!interleaved && pairs > 0) return fail(c, OPE_HIP_EINVAL, "Invalid fixed-pattern noise profile pointer");
    std::lock_guard<std::mutex> lk(c->submit_mu);
    int rc = sync_all(c);
    if (rc) return rc;
    const int half = c->n / 2;
    std::vector<float2> prof(half, make_float2(0.f, 0.f));
    const size_t use = std::min<size_t>(pairs, half);
    for (size_t k = 0; k < use; ++k) prof[k] = make_float2(interleaved[2 * k], interleaved[2 * k + 1]);
    HIP_TRY(c, hipMemcpy(c->d_fpn, prof.data(), half * sizeof(float2), hipMemcpyHostToDevice));
    c->have_fpn = pairs > 0;
    c->fpn_pairs_stored = pairs;
    return refresh_offset(c, false);
}

int ope_hip_get_fpn_profile(ope_hip_ctx* c, float* interleaved, size_t pairs_capacity, size_t* pairs_out) {
    if (!c) return OPE_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    int rc = sync_all(c);
    if (rc) return rc;
    const size_t half = c->n / 2;
    const size_t stored = c->have_fpn ? std::min(c->fpn_pairs_stored, half) : 0;
    if (pairs_out) *pairs_out = stored;
    if (interleaved && stored) {
        std::vector<float2> prof(half);
        HIP_TRY(c, hipMemcpy(prof.data(), c->d_fpn, half * sizeof(float2), hipMemcpyDeviceToHost));
        const size_t use = std::min(pairs_capacity, stored);
        for (size_t k = 0; k < use; ++k) {
            interleaved[2 * k] = prof[k].x;
            interleaved[2 * k + 1] = prof[k].y;
        }
    }
    return OPE_HIP_OK;
}

int ope_hip_request_fpn(ope_hip_ctx* c) {
    if (!c) return OPE_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    c->fpn_requested = true;
    return OPE_HIP_OK;
}

int ope_hip_set_bg_profile(ope_hip_ctx* c, const float* profile, size_t length) {
    if (!c) return OPE_HIP_EINVAL;
    if (!profile) return fail(c, OPE_HIP_EINVAL, "Background curve data is null");
    if (length != (size_t)(c->n / 2))
        return fail(c, OPE_HIP_EINVAL, "Invalid background buffer size. Expected %d but got %zu", c->n / 2, length);
    std::lock_guard<std::mutex> lk(c->submit_mu);
    int rc = sync_all(c);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(c->d_bg, profile, length * sizeof(float), hipMemcpyHostToDevice));
    c->have_bg = true;
    return OPE_HIP_OK;
}

int ope_hip_get_bg_profile(ope_hip_ctx* c, float* profile, size_t capacity, size_t* length_out) {
    if (!c) return OPE_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    int rc = sync_all(c);
    if (rc) return rc;
    const size_t half = c->n / 2;
    const size_t stored = c->have_bg ? half : 0;
    if (length_out) *length_out = stored;
    if (profile && stored) {
        std::vector<float> tmp(half);
        HIP_TRY(c, hipMemcpy(tmp.data(), c->d_bg, half * sizeof(float), hipMemcpyDeviceToHost));
        std::copy(tmp.begin(), tmp.begin() + std::min(capacity, stored), profile);
    }
    return OPE_HIP_OK;
}

int ope_hip_request_bg_recording(ope_hip_ctx* c) {
    if (!c) return OPE_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    c->bg_requested = true;
    return OPE_HIP_OK;
}

int ope_hip_set_sample_offset(ope_hip_ctx* c, int mode, float value) {
    if (!c) return OPE_HIP_EINVAL;
    if (mode != OPE_HIP_OFFSET_OFF && mode != OPE_HIP_OFFSET_AUTO && mode != OPE_HIP_OFFSET_FIXED)
        return fail(c, OPE_HIP_EINVAL, "unknown sample offset mode %d", mode);
    if (mode == OPE_HIP_OFFSET_FIXED) {
        if (!std::isfinite(value)) return fail(c, OPE_HIP_EINVAL, "sample offset must be finite");
        if (input_class(c->p.input_type) != IN_F32 && value != std::rint(value))
            return fail(c, OPE_HIP_EINVAL, "the sample offset of integer samples must be an integer");
    }
    std::lock_guard<std::mutex> lk(c->submit_mu);
    int rc = sync_all(c);
    if (rc) return rc;
    c->offset_mode = mode;
    c->offset_pending = mode == OPE_HIP_OFFSET_AUTO;
    const float m0 = mode == OPE_HIP_OFFSET_FIXED ? value : 0.f;
    HIP_TRY(c, hipMemcpy(c->d_offset, &m0, sizeof(float), hipMemcpyHostToDevice));
    return refresh_offset(c, true);
}

int ope_hip_get_sample_offset(ope_hip_ctx* c, int* mode, float* value) {
    if (!c) return OPE_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    int rc = sync_all(c);
    if (rc) return rc;
    float m0 = 0.f;
    if (c->offset_mode != OPE_HIP_OFFSET_OFF && !c->offset_pending)
        HIP_TRY(c, hipMemcpy(&m0, c->d_offset, sizeof(float), hipMemcpyDeviceToHost));
    if (mode) *mode = c->offset_mode;
    if (value) *value = m0;
    return OPE_HIP_OK;
}

int ope_hip_process_device(ope_hip_ctx* c, const void* d_in, float* d_out, int64_t n_ascans, void* stream) {
    if (!c) return OPE_HIP_EINVAL;
    if (n_ascans < 0) return fail(c, OPE_HIP_EINVAL, "n_ascans must be >= 0");
    if (n_ascans > 0 && (!d_in || !d_out)) return fail(c, OPE_HIP_EINVAL, "null device pointer");
    if (reinterpret_cast<uintptr_t>(d_in) % 16 || reinterpret_cast<uintptr_t>(d_out) % 4)
        return fail(c, OPE_HIP_EINVAL, "d_in must be 16-byte and d_out 4-byte aligned");
    std::lock_guard<std::mutex> lk(c->submit_mu);
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the HIP null stream
    c->foreign_work = true;
    return run_frame(c, d_in, d_out, n_ascans, s);
}

int ope_hip_submit_host(ope_hip_ctx* c, const void* host_in, float* host_out, void (*on_input_free)(void*),
                        void* user_in, ope_hip_frame_cb on_done, void* user) {
    if (!c) return OPE_HIP_EINVAL;
    if (!host_in || !host_out) return fail(c, OPE_HIP_EINVAL, "null host buffer");
    std::lock_guard<std::mutex> lk(c->submit_mu);
    HIP_TRY(c, hipSetDevice(c->device));
    // the slot advances only when the frame is enqueued, as the caller's
    // output-buffer index does (HipBackend::process)
    const int si = c->next_slot;
    Slot& s = c->slots[si];
    if (s.used) HIP_TRY(c, hipEventSynchronize(s.d2h_done));  // slot buffers and callback record free
    s.used = true;
    HostCall& hc = c->calls[si];
    hc = HostCall{on_done, user, host_out, on_input_free, user_in};
    const long long frame = (long long)c->p.ascans_per_bscan * c->p.bscans_per_buffer;
    const size_t in_bytes = (size_t)frame * c->n * input_bytes(c->p.input_type);
    const size_t out_bytes = (size_t)frame * (c->n / 2) * sizeof(float);
    HIP_TRY(c, hipMemcpyAsync(s.d_in, host_in, in_bytes, hipMemcpyHostToDevice, c->s_in));
    // events are recorded after the host callbacks so that a slot is only
    // reused once both of its callbacks have returned
    HIP_TRY(c, hipLaunchHostFunc(c->s_in, host_input_free_cb, &hc));
    HIP_TRY(c, hipEventRecord(s.h2d_done, c->s_in));
    HIP_TRY(c, hipStreamWaitEvent(c->s_comp, s.h2d_done, 0));
    int rc = run_frame(c, s.d_in, s.d_out, frame, c->s_comp);
    if (rc) return rc;
    HIP_TRY(c, hipEventRecord(s.comp_done, c->s_comp));
    HIP_TRY(c, hipStreamWaitEvent(c->s_out, s.comp_done, 0));
    HIP_TRY(c, hipMemcpyAsync(host_out, s.d_out, out_bytes, hipMemcpyDeviceToHost, c->s_out));
    HIP_TRY(c, hipLaunchHostFunc(c->s_out, host_done_cb, &hc));
    HIP_TRY(c, hipEventRecord(s.d2h_done, c->s_out));
    c->next_slot = (si + 1) % (int)c->slots.size();
    return OPE_HIP_OK;
}

int ope_hip_synchronize(ope_hip_ctx* c) {
    if (!c) return OPE_HIP_EINVAL;
    return sync_all(c);
}

int ope_hip_op_min_variance_mean(ope_hip_ctx* c, const float* input, int width, int height, int segments,
                                 float* out) {
    if (!c) return OPE_HIP_EINVAL;
    if (!input || !out) return fail(c, OPE_HIP_EINVAL, "null pointer");
    if (width <= 0 || height <= 0 || segments <= 0)
        return fail(c, OPE_HIP_EINVAL, "width, height and segments must be > 0");
    const int seg_w = height / segments;
    if (seg_w <= 0) return fail(c, OPE_HIP_EINVAL, "height must be >= segments");
    if (segments > 65535) return fail(c, OPE_HIP_EINVAL, "at most 65535 segments");
    std::lock_guard<std::mutex> lk(c->submit_mu);
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t in_bytes = (size_t)seg_w * segments * width * 2 * sizeof(float);
    void* d_in = nullptr;
    double* d_acc = nullptr;
    float2* d_out = nullptr;
    int rc = OPE_HIP_OK;
    auto cleanup = [&]() {
        if (d_in) (void)hipFree(d_in);
        if (d_acc) (void)hipFree(d_acc);
        if (d_out) (void)hipFree(d_out);
    };
    hipError_t e = hipMalloc(&d_in, in_bytes);
    if (e == hipSuccess) e = hipMalloc(&d_acc, (size_t)segments * width * 3 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_out, (size_t)width * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(d_in, input, in_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(segment_sums_kernel, dim3((width + 255) / 256, segments), dim3(256), 0, c->s_comp,
                           static_cast<const float2*>(d_in), width, seg_w, segments, d_acc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fpn_select_kernel, dim3((width + 255) / 256), dim3(256), 0, c->s_comp, d_acc, width,
                           segments, seg_w, d_out, nullptr, nullptr, 0, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->s_comp);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, (size_t)width * sizeof(float2), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(c, OPE_HIP_ERUNTIME, "min_variance_mean: %s", hipGetErrorString(e));
    cleanup();
    return rc;
}

int ope_hip_op_min_variance_mean_device(ope_hip_ctx* c, const float* d_input, int width, int height, int segments,
                                        float* d_out, void* stream) {
    if (!c) return OPE_HIP_EINVAL;
    if (!d_input || !d_out) return fail(c, OPE_HIP_EINVAL, "null pointer");
    if (width <= 0 || height <= 0 || segments <= 0)
        return fail(c, OPE_HIP_EINVAL, "width, height and segments must be > 0");
    const int seg_w = height / segments;
    if (seg_w <= 0) return fail(c, OPE_HIP_EINVAL, "height must be >= segments");
    if (segments > 65535) return fail(c, OPE_HIP_EINVAL, "at most 65535 segments");
    if ((size_t)segments * (size_t)width * 3 > kScratchDoubles)
        return fail(c, OPE_HIP_EINVAL, "segments * width = %lld exceeds the scratch limit of %zu",
                    (long long)segments * width, kScratchDoubles / 3);
    if (reinterpret_cast<uintptr_t>(d_input) % 8 || reinterpret_cast<uintptr_t>(d_out) % 8)
        return fail(c, OPE_HIP_EINVAL, "d_input and d_out must be 8-byte aligned");
    double* acc = scratch_ptr(c->device);
    if (!acc) return fail(c, OPE_HIP_ESTATE, "no scratch address for device %d", c->device);
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // every sum the selection reads is written here first: nothing to clear
    hipLaunchKernelGGL(segment_sums_kernel, dim3((width + 255) / 256, segments), dim3(256), 0, s,
                       reinterpret_cast<const float2*>(d_input), width, seg_w, segments, acc);
    HIP_TRY(c, hipGetLastError());
    hipLaunchKernelGGL(fpn_select_kernel, dim3((width + 255) / 256), dim3(256), 0, s, acc, width, segments, seg_w,
                       reinterpret_cast<float2*>(d_out), nullptr, nullptr, 0, nullptr);
    HIP_TRY(c, hipGetLastError());
    return OPE_HIP_OK;
}

int ope_hip_op_convert_input(ope_hip_ctx* c, const void* input, int input_type, int64_t samples, int bitshift,
                             float* out) {
    if (!c || (samples > 0 && (!input || !out)) || samples < 0) return OPE_HIP_EINVAL;
    const int cls = input_class(input_type);
    if (!cls) return fail(c, OPE_HIP_EUNSUP, "input type %d not supported", input_type);
    if (samples == 0) return OPE_HIP_OK;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    OP_TRY(hipSetDevice(c->device));
    const size_t iw = input_bytes(input_type);
    DevBuf din, dout;
    OP_TRY(hipMalloc(&din.p, samples * iw));
    OP_TRY(hipMalloc(&dout.p, samples * sizeof(float)));
    OP_TRY(hipMemcpy(din.p, input, samples * iw, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(op_convert_kernel, dim3((unsigned)((samples + 255) / 256)), dim3(256), 0, c->s_comp, din.p,
                       static_cast<float*>(dout.p), (long long)samples, (int)iw, cls == IN_F32 ? 1 : 0, bitshift ? 1 : 0);
    OP_TRY(hipGetLastError());
    OP_TRY(hipStreamSynchronize(c->s_comp));
    OP_TRY(hipMemcpy(out, dout.p, samples * sizeof(float), hipMemcpyDeviceToHost));
    return OPE_HIP_OK;
}

int ope_hip_op_klinearization(ope_hip_ctx* c, const float* input, const float* curve, int method, int line_width,
                              int samples, float* out) {
    if (!c || !input || !curve || !out || line_width < 1 || samples < 0 || samples > line_width) return OPE_HIP_EINVAL;
    const int T = method == OPE_HIP_INTERP_LINEAR    ? 2
                  : method == OPE_HIP_INTERP_CUBIC   ? 4
                  : method == OPE_HIP_INTERP_LANCZOS ? 16
                  : method == OPE_HIP_INTERP_LANCZOS4 ? 8
                                                     : 0;
    if (!T) return fail(c, OPE_HIP_EINVAL, "unknown interpolation method %d", method);
    if (line_width < T) return fail(c, OPE_HIP_EINVAL, "line shorter than the interpolation kernel");
    if (samples == 0) return OPE_HIP_OK;
    std::vector<int> base;
    std::vector<float> w;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    int rc = build_taps_for(c, line_width, samples, T, curve, nullptr, base, w);
    if (rc) return rc;
    OP_TRY(hipSetDevice(c->device));
    DevBuf dx, db, dw, dy;
    OP_TRY(hipMalloc(&dx.p, line_width * sizeof(float)));
    OP_TRY(hipMalloc(&db.p, samples * sizeof(int)));
    OP_TRY(hipMalloc(&dw.p, w.size() * sizeof(float)));
    OP_TRY(hipMalloc(&dy.p, samples * sizeof(float)));
    OP_TRY(hipMemcpy(dx.p, input, line_width * sizeof(float), hipMemcpyHostToDevice));
    OP_TRY(hipMemcpy(db.p, base.data(), samples * sizeof(int), hipMemcpyHostToDevice));
    OP_TRY(hipMemcpy(dw.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(op_taps_kernel, dim3((samples + 255) / 256), dim3(256), 0, c->s_comp,
                       static_cast<const float*>(dx.p), static_cast<const int*>(db.p), static_cast<const float*>(dw.p),
                       T, samples, static_cast<float*>(dy.p));
    OP_TRY(hipGetLastError());
    OP_TRY(hipStreamSynchronize(c->s_comp));
    OP_TRY(hipMemcpy(out, dy.p, samples * sizeof(float), hipMemcpyDeviceToHost));
    return OPE_HIP_OK;
}

int ope_hip_op_windowing(ope_hip_ctx* c, const float* input, const float* window, int line_width, int samples,
                         float* out) {
    return op_scale(c, input, window, 1, line_width, samples, out);
}

int ope_hip_op_dispersion(ope_hip_ctx* c, const float* input, const float* cos_sin, int line_width, int samples,
                          float* out) {
    return op_scale(c, input, cos_sin, 2, line_width, samples, out);
}

int ope_hip_op_ifft_magnitude(ope_hip_ctx* c, const float* input, int line_width, int lines, float* out) {
    if (!c || !input || !out || lines < 0) return OPE_HIP_EINVAL;
    if (!supported_n(line_width))
        return fail(c, OPE_HIP_EUNSUP, "line width %d not supported (powers of two 256..4096)", line_width);
    if (lines == 0) return OPE_HIP_OK;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    OP_TRY(hipSetDevice(c->device));
    const size_t bytes = (size_t)line_width * lines * sizeof(float);
    DevBuf dx, dy, dt;
    OP_TRY(hipMalloc(&dx.p, bytes));
    OP_TRY(hipMalloc(&dy.p, bytes));
    std::vector<float2> tw(line_width);
    for (int k = 0; k < line_width; ++k) {
        const double ang = 2.0 * M_PI * (double)k / (double)line_width;
        tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    OP_TRY(hipMalloc(&dt.p, line_width * sizeof(float2)));
    OP_TRY(hipMemcpy(dt.p, tw.data(), line_width * sizeof(float2), hipMemcpyHostToDevice));
    OP_TRY(hipMemcpy(dx.p, input, bytes, hipMemcpyHostToDevice));
    const float* x = static_cast<const float*>(dx.p);
    const float2* t = static_cast<const float2*>(dt.p);
    float* y = static_cast<float*>(dy.p);
    switch (line_width) {
        case 256: launch_ifft_mag<256>(x, lines, t, y, c->s_comp); break;
        case 512: launch_ifft_mag<512>(x, lines, t, y, c->s_comp); break;
        case 1024: launch_ifft_mag<1024>(x, lines, t, y, c->s_comp); break;
        case 2048: launch_ifft_mag<2048>(x, lines, t, y, c->s_comp); break;
        default: launch_ifft_mag<4096>(x, lines, t, y, c->s_comp); break;
    }
    OP_TRY(hipGetLastError());
    OP_TRY(hipStreamSynchronize(c->s_comp));
    OP_TRY(hipMemcpy(out, dy.p, bytes, hipMemcpyDeviceToHost));
    return OPE_HIP_OK;
}

int ope_hip_op_fpn_removal(ope_hip_ctx* c, const float* input, const float* mean, int line_width, int lines,
                           float* out) {
    if (!c || !input || !out || line_width < 1 || lines < 0) return OPE_HIP_EINVAL;
    const long long total = (long long)line_width * lines;
    if (total == 0) return OPE_HIP_OK;
    const int use = mean ? (int)std::min<long long>(total / 2, line_width) : 0;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    OP_TRY(hipSetDevice(c->device));
    DevBuf dx, dm, dy;
    OP_TRY(hipMalloc(&dx.p, total * sizeof(float2)));
    OP_TRY(hipMalloc(&dm.p, std::max(use, 1) * sizeof(float2)));
    OP_TRY(hipMalloc(&dy.p, total * sizeof(float)));
    OP_TRY(hipMemcpy(dx.p, input, total * sizeof(float2), hipMemcpyHostToDevice));
    if (use > 0) OP_TRY(hipMemcpy(dm.p, mean, use * sizeof(float2), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(op_fpn_mag_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->s_comp,
                       static_cast<const float2*>(dx.p), static_cast<const float2*>(dm.p), total, use,
                       static_cast<float*>(dy.p));
    OP_TRY(hipGetLastError());
    OP_TRY(hipStreamSynchronize(c->s_comp));
    OP_TRY(hipMemcpy(out, dy.p, total * sizeof(float), hipMemcpyDeviceToHost));
    return OPE_HIP_OK;
}

#undef OP_TRY

int ope_hip_host_alloc(size_t bytes, void** ptr) {
    if (!ptr) return OPE_HIP_EINVAL;
    *ptr = nullptr;
    if (bytes == 0) return OPE_HIP_OK;
    if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
        *ptr = nullptr;
        return OPE_HIP_ENOMEM;
    }
    std::memset(*ptr, 0, bytes);
    return OPE_HIP_OK;
}

int ope_hip_host_free(void* ptr) {
    if (ptr) (void)hipHostFree(ptr);
    return OPE_HIP_OK;
}

int ope_hip_autofocus_loss(ope_hip_ctx* c, const void* d_in, int64_t n_ascans, const double* theta, double* out,
                           void* stream) {
    if (!c || !out) return OPE_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    double acc[2] = {0.0, 0.0};
    const int rc = autofocus_pass(c, d_in, n_ascans, theta, false, 0.f, 0.f, static_cast<hipStream_t>(stream), acc);
    if (rc) return rc;
    const double S = acc[0], Q = acc[1];
    out[0] = std::log(S) - Q / S;
    out[1] = S;
    out[2] = Q;
    return OPE_HIP_OK;
}

int ope_hip_autofocus_grad(ope_hip_ctx* c, const void* d_in, int64_t n_ascans, const double* theta, double sum_i,
                           double loss, double* grad, void* stream) {
    if (!c || !grad) return OPE_HIP_EINVAL;
    if (!(sum_i > 0.0)) return fail(c, OPE_HIP_EINVAL, "sum_i must be > 0");
    std::lock_guard<std::mutex> lk(c->submit_mu);
    const double inv_s = 1.0 / sum_i;
    return autofocus_pass(c, d_in, n_ascans, theta, true, (float)inv_s, (float)(std::log(inv_s) + loss),
                          static_cast<hipStream_t>(stream), grad);
}

int ope_hip_graph_create(ope_hip_ctx* c, const void* const* d_in, float* const* d_out, int n_frames,
                         int64_t ascans_per_frame, ope_hip_graph** out) {
    g_create_err.clear();
    if (!out) {
        g_create_err = "out is null";
        return OPE_HIP_EINVAL;
    }
    *out = nullptr;
    if (!c || !d_in || !d_out || n_frames < 1 || ascans_per_frame < 1) {
        g_create_err = "bad arguments";
        return OPE_HIP_EINVAL;
    }
    std::lock_guard<std::mutex> lk(c->submit_mu);
    const ope_hip_params& p = c->p;
    // steady state: the frames would launch the chain kernel only
    const long long required = (long long)p.fpn_bscan_count * p.ascans_per_bscan;
    const long long rows_used = (required / kFpnSegments) * kFpnSegments;
    if (c->fpn_requested || p.fpn_continuous) {
        g_create_err = "an FPN determination is pending";
        return OPE_HIP_ESTATE;
    }
    if (rows_used > 0 && c->fpn_count < rows_used) {
        g_create_err = "FPN statistics are still accumulating: process frames first";
        return OPE_HIP_ESTATE;
    }
    if (c->bg_requested && p.bg_removal) {
        g_create_err = "a background recording is pending";
        return OPE_HIP_ESTATE;
    }
    for (int f = 0; f < n_frames; ++f) {
        if (!d_in[f] || !d_out[f] || reinterpret_cast<uintptr_t>(d_in[f]) % 16 ||
            reinterpret_cast<uintptr_t>(d_out[f]) % 4) {
            g_create_err = "frame buffers must be non-null, d_in 16-byte and d_out 4-byte aligned";
            return OPE_HIP_EINVAL;
        }
    }
    if (p.sinusoidal_correction) {
        // allocations cannot happen inside the capture
        int rc0 = ensure_sin_curve(c);
        if (!rc0) rc0 = ensure_sin_tmp(c, ascans_per_frame);
        if (rc0) {
            g_create_err = c->err;
            return rc0;
        }
    }
    ope_hip_graph* g = new (std::nothrow) ope_hip_graph();
    if (!g) {
        g_create_err = "out of host memory";
        return OPE_HIP_ENOMEM;
    }
    g->device = c->device;
    hipStream_t s = nullptr;
    auto cleanup = [&](int code, const char* what, hipError_t e) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
        g_create_err = buf;
        if (s) (void)hipStreamDestroy(s);
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return code;
    };
    g->ctx = c;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(OPE_HIP_ERUNTIME, "stream", e);
    // an automatic offset is determined before the capture, from the first
    // frame, so that every replay and later eager frame use the same one
    if (offset_active(c) && c->offset_pending) {
        if (determine_offset(c, d_in[0], ascans_per_frame, c->s_comp) != OPE_HIP_OK) {
            const std::string msg = c->err;
            cleanup(OPE_HIP_ERUNTIME, "offset", hipSuccess);
            g_create_err = msg;
            return OPE_HIP_ERUNTIME;
        }
        e = hipStreamSynchronize(c->s_comp);
        if (e != hipSuccess) return cleanup(OPE_HIP_ERUNTIME, "offset", e);
    }
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return cleanup(OPE_HIP_ERUNTIME, "hipStreamBeginCapture", e);
    // the frame bookkeeping of run_frame is host state: keep it as it was
    const long long fpn_count = c->fpn_count;
    int rc = OPE_HIP_OK;
    for (int f = 0; f < n_frames && rc == OPE_HIP_OK; ++f) rc = run_frame(c, d_in[f], d_out[f], ascans_per_frame, s);
    c->fpn_count = fpn_count;
    e = hipStreamEndCapture(s, &g->graph);
    if (rc != OPE_HIP_OK) {
        const std::string msg = c->err;
        cleanup(rc, "capture", hipSuccess);
        g_create_err = msg;
        return rc;
    }
    if (e != hipSuccess) return cleanup(OPE_HIP_ERUNTIME, "hipStreamEndCapture", e);
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return cleanup(OPE_HIP_ERUNTIME, "hipGraphInstantiate", e);
    (void)hipStreamDestroy(s);
    *out = g;
    return OPE_HIP_OK;
}

int ope_hip_graph_launch(ope_hip_graph* g, void* stream) {
    if (!g) return OPE_HIP_EINVAL;
    g->ctx->foreign_work = true;
    hipError_t e = hipSetDevice(g->device);
    if (e == hipSuccess) e = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        g->err = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
        return OPE_HIP_ERUNTIME;
    }
    return OPE_HIP_OK;
}

const char* ope_hip_graph_last_error(const ope_hip_graph* g) { return g ? g->err.c_str() : g_create_err.c_str(); }

int ope_hip_graph_destroy(ope_hip_graph* g) {
    if (!g) return OPE_HIP_EINVAL;
    (void)hipSetDevice(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return OPE_HIP_OK;
}

// Stores the launch variant of the context's current configuration.
void store_tuning(ope_hip_ctx* c, int mode, int iw, int as, int sched, int per_cu) {
    for (auto& t : c->tuning)
        if (t.taps == c->taps && t.mode == mode && t.iw == iw) {
            t.as = as;
            t.sched = sched;
            t.per_cu = per_cu;
            return;
        }
    c->tuning.push_back({c->taps, mode, iw, as, sched, per_cu});
}

int ope_hip_set_launch_variant(ope_hip_ctx* c, int ascans_per_lane, int small_radix_pass, int workgroups_per_cu) {
    if (!c) return OPE_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    if (ascans_per_lane != 1 && ascans_per_lane != 2)
        return fail(c, OPE_HIP_EINVAL, "A-scans per lane must be 1 or 2");
    if (workgroups_per_cu < 0) return fail(c, OPE_HIP_EINVAL, "workgroups per CU must be >= 0");
    const int np = fft_passes(c->n);
    // np: the swap schedule (N = 1024, 2048; built where pick_kernel finds it)
    if (small_radix_pass < 0 || small_radix_pass > np)
        return fail(c, OPE_HIP_EINVAL, "small-radix pass must be in 0..%d", np);
    const int mode = c->p.dc_removal_enabled ? M_DC : 0;
    const int iw = (int)input_bytes(c->p.input_type);
    const int sched = small_radix_pass == np - 1 ? -1 : small_radix_pass;
    const ChainLaunch l = pick_kernel(c->n, c->taps, mode, iw, ascans_per_lane, sched);
    // placements other than the last (and the swap schedule) exist for one
    // A-scan per lane only
    if (!l.fn || (ascans_per_lane == 2 && sched != -1 && sched != np))
        return fail(c, OPE_HIP_EUNSUP, "launch variant not built");
    store_tuning(c, mode, iw, ascans_per_lane, sched, workgroups_per_cu);
    return OPE_HIP_OK;
}

int ope_hip_autotune(ope_hip_ctx* c, const void* d_in, float* d_out, int64_t n_ascans, void* stream, int* choice,
                     float* best_us) {
    if (!c) return OPE_HIP_EINVAL;
    std::lock_guard<std::mutex> lk(c->submit_mu);
    HIP_TRY(c, hipSetDevice(c->device));
    const ope_hip_params& p = c->p;
    if (!d_in || !d_out) {
        // the context's own first frame slot
        if (c->slots.empty()) return fail(c, OPE_HIP_ESTATE, "no frame slot");
        d_in = c->slots[0].d_in;
        d_out = c->slots[0].d_out;
        n_ascans = (long long)p.ascans_per_bscan * p.bscans_per_buffer;
    }
    if (n_ascans < 1) return fail(c, OPE_HIP_EINVAL, "n_ascans must be >= 1");
    if (reinterpret_cast<uintptr_t>(d_in) % 16 || reinterpret_cast<uintptr_t>(d_out) % 4)
        return fail(c, OPE_HIP_EINVAL, "d_in must be 16-byte and d_out 4-byte aligned");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the launch a steady-state frame makes (run_frame)
    ChainArgs a = base_args(c);
    a.in = d_in;
    a.out = d_out;
    a.n_ascans = n_ascans;
    if (p.fpn_removal && c->have_fpn) a.flags |= F_FPN;
    if (p.bscan_flip) {
        a.flags |= F_FLIP;
        a.flip_bscans = n_ascans / p.ascans_per_bscan;
    }
    if (p.bg_removal && c->have_bg && !p.sinusoidal_correction) a.flags |= F_BG;
    if (offset_active(c) && !c->offset_pending) apply_offset(c, a);
    c->foreign_work = true;
    const int mode = p.dc_removal_enabled ? M_DC : 0;
    const int iw = (int)input_bytes(p.input_type);
    struct Cand {
        int as, sched, per_cu;
    };
    std::vector<Cand> cands;
    const int np = fft_passes(c->n);
    for (int as = 1; as <= 2; ++as) {
        // every placement of the small-radix pass (one A-scan per lane), and
        // the swap schedule (sched = np) where it is built
        const bool small = (c->n >> (3 * (np - 1))) != 8;
        std::vector<int> scheds;
        if (as == 1 && small)
            for (int sc = 0; sc < np - 1; ++sc) scheds.push_back(sc);
        scheds.push_back(-1);
        if (small || c->n == 4096) scheds.push_back(np);
        for (const int sched : scheds) {
            const ChainLaunch l = pick_kernel(c->n, c->taps, variant_mode(a, mode, as), iw, as, sched);
            if (!l.fn) continue;
            int occ = 0;
            HIP_TRY(c, cached_occupancy(c, reinterpret_cast<const void*>(l.fn), l.wg, &occ));
            for (int k = 1; k <= occ; ++k) cands.push_back({as, sched, k});
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    float best = 0.f;
    Cand pick{1, -1, 0};
    int rc = OPE_HIP_OK;
    constexpr int kReps = 3;
    for (const Cand& cd : cands) {
        rc = launch_variant(c, a, mode, cd.as, cd.sched, cd.per_cu, n_ascans, s);  // warm-up
        if (rc) break;
        hipError_t e = hipEventRecord(e0, s);
        for (int r = 0; r < kReps && e == hipSuccess && !rc; ++r)
            rc = launch_variant(c, a, mode, cd.as, cd.sched, cd.per_cu, n_ascans, s);
        if (!rc && e == hipSuccess) e = hipEventRecord(e1, s);
        if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (!rc && e != hipSuccess) rc = fail(c, OPE_HIP_ERUNTIME, "autotune timing: %s", hipGetErrorString(e));
        if (rc) break;
        const float us = ms * 1000.f / kReps;
        if (best == 0.f || us < best) {
            best = us;
            pick = cd;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    store_tuning(c, mode, iw, pick.as, pick.sched, pick.per_cu);
    if (choice) {
        choice[0] = pick.as;
        choice[1] = pick.sched < 0 ? np - 1 : pick.sched;
        choice[2] = pick.per_cu;
    }
    if (best_us) *best_us = best;
    return OPE_HIP_OK;
}

int ope_hip_ragged_create(ope_hip_ctx* const* ctxs, int n_ctx, const ope_hip_request* reqs, int n_reqs,
                          int chunk_ascans, ope_hip_ragged** out) {
    g_create_err.clear();
    if (!out) {
        g_create_err = "out is null";
        return OPE_HIP_EINVAL;
    }
    *out = nullptr;
    if (chunk_ascans < 0) {
        g_create_err = "chunk_ascans must be >= 0";
        return OPE_HIP_EINVAL;
    }
    ope_hip_ragged* r = new (std::nothrow) ope_hip_ragged();
    if (!r) {
        g_create_err = "out of host memory";
        return OPE_HIP_ENOMEM;
    }
    const int rc = build_ragged(r, ctxs, n_ctx, reqs, n_reqs);
    if (rc) {
        g_create_err = r->err;
        free_ragged(r);
        return rc;
    }
    *out = r;
    return OPE_HIP_OK;
}

int ope_hip_ragged_run(ope_hip_ragged* r, void* stream) {
    if (!r) return OPE_HIP_EINVAL;
    if (r->classes.empty()) return OPE_HIP_OK;
    hipError_t e = hipSetDevice(r->device);
    if (e != hipSuccess) return ragged_fail(r, OPE_HIP_ERUNTIME, "hipSetDevice: %s", hipGetErrorString(e));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // the run's start (the latencies' origin), then the fork: the side
    // streams start after it and after the work the caller queued before
    hipLaunchKernelGGL(stamp_clock_kernel, dim3(1), dim3(1), 0, s, r->d_start);
    e = hipEventRecord(r->fork, s);
    if (e != hipSuccess) return ragged_fail(r, OPE_HIP_ERUNTIME, "fork: %s", hipGetErrorString(e));
    // (grids and the split between the launches: build_ragged)
    for (const auto& k : r->classes) k.ctx->foreign_work = true;
    // The largest length first, before the side stream's work is queued:
    // the GPU then dispatches its workgroups first even while this call is
    // still queueing the rest, and their waves, the older ones on each
    // SIMD, keep the issue priority (the largest length is the batch's
    // critical path).  The side launch also waits 20 us behind the fork.
    const auto& big = r->classes[0];
    hipLaunchKernelGGL(big.launch.fn, dim3((unsigned)big.grid_main), dim3(big.launch.wg), 0, s, big.args);
    for (size_t i = 0; e == hipSuccess && i < r->side.size(); ++i) e = hipStreamWaitEvent(r->side[i], r->fork, 0);
    if (e != hipSuccess) return ragged_fail(r, OPE_HIP_ERUNTIME, "fork: %s", hipGetErrorString(e));
    if (r->classes.size() > 1) {
        hipLaunchKernelGGL(hold_clock_kernel, dim3(1), dim3(64), 0, r->side[0],
                           (unsigned long long)(20e-6 * r->tick_hz));
        const auto& k = r->classes[1];
        hipLaunchKernelGGL(k.launch.fn, dim3((unsigned)k.grid_side), dim3(k.launch.wg), k.side_lds, r->side[0],
                           k.args);
    }
    for (size_t i = 1; i < r->classes.size(); ++i) {
        const auto& k = r->classes[i];
        hipLaunchKernelGGL(k.launch.fn, dim3((unsigned)k.grid_main), dim3(k.launch.wg), 0, s, k.args);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return ragged_fail(r, OPE_HIP_ERUNTIME, "ragged launch: %s", hipGetErrorString(e));
    // join: the caller's stream continues once every length is done
    for (size_t i = 0; e == hipSuccess && i < r->side.size(); ++i) {
        e = hipEventRecord(r->joins[i], r->side[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, r->joins[i], 0);
    }
    if (e != hipSuccess) return ragged_fail(r, OPE_HIP_ERUNTIME, "join: %s", hipGetErrorString(e));
    return OPE_HIP_OK;
}

int ope_hip_ragged_latencies(ope_hip_ragged* r, double* latency_us) {
    if (!r || (!latency_us && r->n_reqs > 0)) return OPE_HIP_EINVAL;
    if (r->n_reqs == 0) return OPE_HIP_OK;
    std::vector<unsigned long long> fin(r->n_reqs);
    unsigned long long start = 0;
    hipError_t e = hipSetDevice(r->device);
    if (e == hipSuccess) e = hipMemcpy(fin.data(), r->d_finish, fin.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&start, r->d_start, sizeof(start), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return ragged_fail(r, OPE_HIP_ERUNTIME, "latencies: %s", hipGetErrorString(e));
    for (int q = 0; q < r->n_reqs; ++q) {
        if (r->req_ascans[q] == 0 || fin[q] < start) {
            latency_us[q] = 0.0;
            continue;
        }
        latency_us[q] = (double)(fin[q] - start) / r->tick_hz * 1e6;
    }
    return OPE_HIP_OK;
}

const char* ope_hip_ragged_last_error(const ope_hip_ragged* r) { return r ? r->err.c_str() : g_create_err.c_str(); }

int ope_hip_ragged_destroy(ope_hip_ragged* r) {
    if (!r) return OPE_HIP_EINVAL;
    free_ragged(r);
    return OPE_HIP_OK;
}

}  // extern "C"
