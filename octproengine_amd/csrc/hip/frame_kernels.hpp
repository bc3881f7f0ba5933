// Per-frame helper kernels around the chain: FPN segment selection, the
// segment sums of the individual minimum-variance operation, accumulator
// clearing and the post-process background.  Included by the C ABI host
// translation unit only.
#pragma once

#include <hip/hip_runtime.h>

#include "chain_kernels.hpp"

namespace ope_hip {

// ------------------------------------------------------------------ FPN select

// Epilogue constant of the sample offset: fl(fpn[k] - m0 * D[k]) in FP64
// (fpn taken as 0 when the profile is not applied).
__device__ __forceinline__ float2 offset_epilogue(float2 fpn, bool use_fpn, float m0, double2 d) {
    const double fx = use_fpn ? (double)fpn.x : 0.0, fy = use_fpn ? (double)fpn.y : 0.0;
    return make_float2((float)(fx - (double)m0 * d.x), (float)(fy - (double)m0 * d.y));
}

// Minimum-variance segment mean per bin, the float/double mixture of
// getMinimumVarianceMean (cpu_kernels.tpp:402-458).  acc is [segs][width][3].
// With `ep` set, the offset's epilogue constant of the new profile is
// written too (m0 = *offset, D = offset_tab).
__global__ void fpn_select_kernel(const double* __restrict__ acc, int width, int segments,
                                  int seg_w, float2* __restrict__ out, const float* __restrict__ offset,
                                  const double2* __restrict__ offset_tab, int use_fpn, float2* __restrict__ ep) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= width) return;
    const float factor = 1.0f / (float)seg_w;
    float min_var = INFINITY, bx = 0.f, by = 0.f;
    for (int s = 0; s < segments; ++s) {
        const double* d = acc + ((size_t)s * width + k) * 3;
        const float mx = (float)(d[0] * (double)factor);
        const float my = (float)(d[1] * (double)factor);
        const float var = (float)(d[2] * (double)factor - ((double)mx * (double)mx + (double)my * (double)my));
        if (var < min_var) {
            min_var = var;
            bx = mx;
            by = my;
        }
    }
    out[k] = make_float2(bx, by);
    if (ep) ep[k] = offset_epilogue(make_float2(bx, by), use_fpn != 0, *offset, offset_tab[k]);
}

// Sample offset of a context, one workgroup of 256 lanes.  With
// `determine` set, m0 becomes the mean of the first `rows` A-scans of `in`
// (converted as the chain converts them), rounded to an integer for
// integer samples (so that x - m0 is exact), and is stored to *offset;
// otherwise *offset is kept.  Then the epilogue constant of every bin k <
// width is rebuilt from the current profile: ep[k] = fl(fpn[k] - m0*D[k]).
__global__ void __launch_bounds__(256) offset_prepare_kernel(const void* __restrict__ in, int n, int rows, int iw,
                                                             int ieee, int shift, int determine,
                                                             float* __restrict__ offset,
                                                             const float2* __restrict__ fpn, int use_fpn,
                                                             const double2* __restrict__ offset_tab,
                                                             float2* __restrict__ ep, int width) {
    __shared__ double part[256];
    __shared__ float m0s;
    if (determine) {
        const long long total = (long long)rows * n;
        double s = 0.0;
        for (long long i = threadIdx.x; i < total; i += 256) s += (double)convert_sample(in, i, iw, ieee != 0, shift != 0);
        part[threadIdx.x] = s;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const double mean = total > 0 ? part[0] / (double)total : 0.0;
            float m0 = ieee ? (float)mean : (float)rint(mean);
            if (!isfinite(m0)) m0 = 0.f;
            *offset = m0;
            m0s = m0;
        }
    } else if (threadIdx.x == 0) {
        m0s = *offset;
    }
    __syncthreads();
    const float m0 = m0s;
    for (int k = threadIdx.x; k < width; k += 256) ep[k] = offset_epilogue(fpn[k], use_fpn != 0, m0, offset_tab[k]);
}

// Segment sums over arbitrary interleaved complex rows [height][width]
// (for the individual getMinimumVarianceMean operation).
__global__ void segment_sums_kernel(const float2* __restrict__ in, int width, int seg_w,
                                    int segments, double* __restrict__ acc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (k >= width || s >= segments) return;
    double sx = 0.0, sy = 0.0, sxx = 0.0;
    for (int r = 0; r < seg_w; ++r) {
        const float2 c = in[(size_t)(s * seg_w + r) * width + k];
        const double dx = c.x, dy = c.y;
        sx += dx;
        sy += dy;
        sxx += dx * dx + dy * dy;
    }
    double* d = acc + ((size_t)s * width + k) * 3;
    d[0] = sx;
    d[1] = sy;
    d[2] = sxx;
}

// ------------------------------------------------------------------ small kernels

__global__ void fill_zero_f64(double* __restrict__ p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// The wall clock into *p (one lane): a stream-ordered time stamp.
__global__ void stamp_clock_kernel(unsigned long long* p) { *p = wall_clock64(); }

// Holds its stream for `ticks` of the wall clock (one wave, sleeping), so
// that work queued behind it reaches the GPU after work queued at the same
// time on another stream.  Ends by itself.
__global__ void hold_clock_kernel(unsigned long long ticks) {
    const unsigned long long t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
}

// Column mean of the processed frame (cpu_backend.cpp:105-118), accumulated
// in double.
__global__ void bg_record_kernel(const float* __restrict__ out, long long rows, int width,
                                 float* __restrict__ bg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= width) return;
    double s = 0.0;
    for (long long r = 0; r < rows; ++r) s += out[r * width + k];
    bg[k] = (float)(s / (double)rows);
}

// out -= weight*bg + offset (cpu_kernels.tpp:380-400)
__global__ void bg_subtract_kernel(float* __restrict__ out, long long total, int width,
                                   const float* __restrict__ bg, float w, float off) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float b = __fadd_rn(__fmul_rn(w, bg[i % width]), off);
    out[i] = out[i] - b;
}

// Sinusoidal scan correction (cuda_kernels.cu:408-431) from `in` to `out`
// (total floats, rows of `width`): A-scan k of B-scan l becomes the linear
// interpolation at scan position curve[k] between A-scans floor(curve[k])
// and floor(curve[k]) + 1 of the flat buffer; the buffer's last A-scan is
// copied.  Then, when use_bg is set, the post-process background
// w*bg + off is subtracted (cuda_backend.cu:710-743 runs after this step).
__global__ void sinusoidal_kernel(const float* __restrict__ in, float* __restrict__ out, long long total, int width,
                                  int ascans, const float* __restrict__ curve, const float* __restrict__ bg, float bg_w,
                                  float bg_off, int use_bg) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const long long row = i / width;
    const int j = (int)(i - row * width);
    float v;
    if (i < total - width) {
        const int k = (int)(row % ascans);
        const long long l = row / ascans;
        const float x = curve[k];
        const int xi = (int)x;
        // rows clamped to the buffer: only B-scans shorter than 3 A-scans or
        // a partial last B-scan reach past it
        const long long rows = total / width;
        const long long r0 = min((long long)xi + l * ascans, rows - 1);
        const long long r1 = min(r0 + 1, rows - 1);
        const float f0 = in[r0 * width + j], f1 = in[r1 * width + j];
        v = f0 + (f1 - f0) * (x - (float)xi);
    } else {
        v = in[i];
    }
    if (use_bg) v = v - __fadd_rn(__fmul_rn(bg_w, bg[j]), bg_off);
    out[i] = v;
}

// ------------------------------------------------------------------ individual operations

// Samples to float, as the chain converts them (one sample per thread).
__global__ void op_convert_kernel(const void* __restrict__ in, float* __restrict__ out, long long samples, int iw,
                                  int ieee, int shift) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= samples) return;
    out[i] = convert_sample(in, i, iw, ieee != 0, shift != 0);
}

// y[j] = sum_i w[j][i] x[base[j] + i], j < count (banded tap table).
__global__ void op_taps_kernel(const float* __restrict__ x, const int* __restrict__ base, const float* __restrict__ w,
                               int taps, int count, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    float acc = 0.f;
    for (int i = 0; i < taps; ++i) acc = fmaf(w[j * taps + i], x[base[j] + i], acc);
    out[j] = acc;
}

// out[i] = in[i] * s[i * stride], i < count (window: stride 1; real part of
// the dispersion product with an interleaved cos/sin curve: stride 2).
__global__ void op_scale_kernel(const float* __restrict__ in, const float* __restrict__ s, int stride, int count,
                                float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = in[i] * s[(long long)i * stride];
}

// |unnormalised inverse FFT| of real lines of N samples: all N bins.
template <int N>
__global__ void __launch_bounds__(Geo<N>::WG) op_ifft_mag_kernel(const float* __restrict__ in, long long lines,
                                                                 const float2* __restrict__ twiddle,
                                                                 float* __restrict__ out) {
    using G = Geo<N>;
    using L = ChainLds<N, 0, 4, 1, G::WG>;
    constexpr int TPA = G::TPA;
    __shared__ float2 bufA[L::BUF];
    __shared__ float2 bufB[L::BUF];
    const int slot = threadIdx.x / TPA;
    const int t = threadIdx.x - slot * TPA;
    float2* const pa[1] = {bufA + slot * G::PADN};
    float2* const pb[1] = {bufB + slot * G::PADN};
    cf tw[G::NPASS][8];
    load_twiddles<N>(twiddle, t, tw);
    const long long line = (long long)blockIdx.x * L::SLOTS + slot;
    const bool active = line < lines;
    cf v[1][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[0][e] = {active ? in[line * N + t + e * TPA] : 0.f, 0.f};
    stockham_inverse<N, 1>(v, pa, pb, tw);
    if (active) {
#pragma unroll
        for (int e = 0; e < 8; ++e)
            out[line * N + t + e * TPA] = __fsqrt_rn(fmaf(v[0][e].x, v[0][e].x, v[0][e].y * v[0][e].y));
    }
}

// |x - mean| over interleaved complex x of `total` values, where the mean is
// subtracted from the first `use` values only (meanALineSubtraction applied
// to the flattened lines, cpu_kernels.tpp:461-479 via
// cpu_backend.cpp:708-733).
__global__ void op_fpn_mag_kernel(const float2* __restrict__ in, const float2* __restrict__ mean, long long total,
                                  int use, float* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    float2 z = in[i];
    if (i < use) {
        z.x -= mean[i].x;
        z.y -= mean[i].y;
    }
    out[i] = __fsqrt_rn(z.x * z.x + z.y * z.y);
}

}  // namespace ope_hip
