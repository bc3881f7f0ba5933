// Differentiable chain for dispersion / k-linearisation autofocus.
//
// Parameters theta = (d2, d3, c0, c1, c2, c3):
//   resampling curve  r[j] = clamp(c0 + c1 u + c2 u^2 + c3 u^3, 7, N-9),
//                     u = j/(N-1)  (the polynomial of the reference's
//                     generateResamplingCurve, processorconfiguration.cpp:438-459)
//   dispersion phase  phi[j] = d2 (j-N/2)^2 + d3 (j-N/2)^3
//   chain             x (DC-removed) -> y[j] = interp(x, r[j]) (linear or
//                     Catmull-Rom, cpu_kernels.tpp:84-146) -> s = y w e^{i phi}
//                     -> X = unnormalised inverse FFT(s) -> I[k] = |X[k]|^2,
//                     k < N/2
//   loss              Shannon entropy of p = I / S over the whole frame,
//                     S = sum I:  L = log S - (sum I log I) / S
//
// Forward kernel: accumulates S and sum I log I (FP64).
// Backward kernel (given S and L): recomputes the forward per A-scan, then
//   dL/dI      = -(log(I/S) + L) / S              (k < N/2, 0 above)
//   g_X        = 2 dL/dI X
//   g_s[j]     = sum_k g_X[k] e^{-2 pi i jk/N} = conj(IFFT(conj(g_X)))[j]
//   dL/dphi[j] = -Im(conj(g_s) s),   dL/dy[j] = Re(conj(g_s) w e^{i phi})
//   dL/dr[j]   = dL/dy dy/dt (0 where the clamp is active)
// and accumulates the six sums dL/dd2 = sum dphi (j-N/2)^2, dL/dd3 =
// sum dphi (j-N/2)^3, dL/dc_m = sum dL/dr u^m (FP64).
#pragma once

#include "chain_kernels.hpp"

namespace ope_hip {

struct AutofocusArgs {
    const void* in;          // [n_ascans][N] samples
    long long n_ascans;
    const float* window;     // [N]
    const float2* twiddle;   // [N] exp(+2 pi i k / N)
    int in_type;
    int bitshift;
    int dc_window;
    double c0, c1, c2, c3;   // resampling polynomial (evaluated in FP64: a float32
                             // position is off by ~1e-4 samples at N = 2048,
                             // enough to move a linear tap across an integer)
    float d2, d3;            // dispersion about N/2
    float inv_s;             // backward: 1/S
    float log_inv_s_plus_l;  // backward: log(1/S) + L
    double* acc;             // forward: [2] S, sum I log I; backward: [6] gradient
};

template <int N, int T, int IW, bool DC, bool BWD>
__global__ void __launch_bounds__(Geo<N>::WG) autofocus_kernel(AutofocusArgs A) {
    using G = Geo<N>;
    constexpr int TPA = G::TPA;
    constexpr int WG = G::WG;
    constexpr int NP = G::NPASS;
    using L = ChainLds<N, DC ? M_DC : 0, IW, 1, WG>;
    constexpr int SLOTS = L::SLOTS;
    constexpr int NW = L::NW;
    static_assert(T == 2 || T == 4, "linear or cubic interpolation");
    __shared__ float2 bufA[L::BUF];
    __shared__ float2 bufB[L::BUF];
    __shared__ float xsep[L::X];
    __shared__ typename L::DcSum scan_part[L::SCAN];
    __shared__ double red[NW][6];

    const int slot = threadIdx.x / TPA;
    const int t = threadIdx.x - slot * TPA;
    float2* const pa[1] = {bufA + slot * G::PADN};
    float2* const pb[1] = {bufB + slot * G::PADN};
    float* const xs = L::X_ALIAS_B ? reinterpret_cast<float*>(pb[0]) : (xsep + slot * N);
    const bool ieee = A.in_type == IN_F32;

    cf tw[NP][8];
    load_twiddles<N>(A.twiddle, t, tw);
    float w[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) w[e] = gmem(A.window)[t + e * TPA];

    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const float k0 = (float)(N / 2);

    for (long long a0 = (long long)blockIdx.x * SLOTS; a0 < A.n_ascans; a0 += (long long)gridDim.x * SLOTS) {
        const long long a = a0 + slot;
        const bool active = a < A.n_ascans;
        uint32_t raw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (active) load_raw<IW>(A.in, N, a, t, raw);
        float xv[8];
        convert_raw<IW>(ieee, A.bitshift != 0, 0.f, raw, xv);
        __syncthreads();  // the previous A-scan's LDS reads are done
        if constexpr (DC) {
            dc_remove_stage<N, TPA, typename L::DcSum>(xv, t, A.dc_window,
                                                       reinterpret_cast<typename L::DcSum*>(pa[0]), scan_part, xs);
        } else {
            float4* dst = reinterpret_cast<float4*>(xs + 8 * t);
            dst[0] = make_float4(xv[0], xv[1], xv[2], xv[3]);
            dst[1] = make_float4(xv[4], xv[5], xv[6], xv[7]);
        }
        __syncthreads();

        // resample (with the derivative of the interpolant), window, phase
        cf v[1][8];
        cf s[8], ph[8];
        float dydt[8], mask[8], uu[8], dj[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int j = t + e * TPA;
            const double ud = (double)j / (double)(N - 1);
            const double r = fma(fma(fma(A.c3, ud, A.c2), ud, A.c1), ud, A.c0);
            const double rc = fmin(fmax(r, 7.0), (double)(N - 9));
            mask[e] = (r > 7.0 && r < (double)(N - 9)) ? 1.0f : 0.0f;
            const int n1 = (int)rc;
            const float tt = (float)(rc - (double)n1);
            const float u = (float)ud;
            float y, dy;
            if constexpr (T == 2) {
                const float x1 = xs[n1], x2 = xs[n1 + 1];
                y = fmaf(x2 - x1, tt, x1);
                dy = x2 - x1;
            } else {
                const float y0 = xs[n1 - 1], y1 = xs[n1], y2 = xs[n1 + 1], y3 = xs[n1 + 2];
                const float ca = -y0 + 3.0f * (y1 - y2) + y3;
                const float cb = 2.0f * y0 - 5.0f * y1 + 4.0f * y2 - y3;
                const float cc = -y0 + y2;
                y = 0.5f * tt * (ca * tt * tt + cb * tt + cc) + y1;
                dy = 0.5f * (3.0f * ca * tt * tt + 2.0f * cb * tt + cc);
            }
            const float d = (float)j - k0;
            const float phi = d * d * fmaf(A.d3, d, A.d2);
            float sn, cs;
            sincosf(phi, &sn, &cs);
            ph[e] = {cs, sn};
            const float yw = y * w[e];
            s[e] = {yw * cs, yw * sn};
            v[0][e] = s[e];
            dydt[e] = dy;
            uu[e] = u;
            dj[e] = d;
        }

        stockham_inverse<N, 1>(v, pa, pb, tw);
        // v[0][e] = X[t + e*TPA]; bins e < 4 are k < N/2

        if constexpr (!BWD) {
            if (active) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float I = fmaf(v[0][e].x, v[0][e].x, v[0][e].y * v[0][e].y);
                    acc[0] += (double)I;
                    if (I > 0.0f) acc[1] += (double)I * (double)logf(I);
                }
            }
        } else {
            // g_X = 2 dL/dI X on k < N/2, conjugated for the second inverse FFT
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                if (e < 4) {
                    const float I = fmaf(v[0][e].x, v[0][e].x, v[0][e].y * v[0][e].y);
                    const float g = I > 0.0f ? -2.0f * A.inv_s * (logf(I) + A.log_inv_s_plus_l) : 0.0f;
                    v[0][e] = {g * v[0][e].x, -g * v[0][e].y};
                } else {
                    v[0][e] = {0.0f, 0.0f};
                }
            }
            __syncthreads();  // the forward transform's last reads of pa/pb are done
            stockham_inverse<N, 1>(v, pa, pb, tw);
            if (active) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const cf gs = {v[0][e].x, -v[0][e].y};  // g_s[t + e*TPA]
                    const float dphi = gs.y * s[e].x - gs.x * s[e].y;
                    const float dy = w[e] * (gs.x * ph[e].x + gs.y * ph[e].y);
                    const float dr = dy * dydt[e] * mask[e];
                    const float d2 = dj[e] * dj[e];
                    acc[0] += (double)(dphi * d2);
                    acc[1] += (double)(dphi * d2 * dj[e]);
                    acc[2] += (double)dr;
                    acc[3] += (double)(dr * uu[e]);
                    acc[4] += (double)(dr * uu[e] * uu[e]);
                    acc[5] += (double)(dr * uu[e] * uu[e] * uu[e]);
                }
            }
        }
    }

    // workgroup reduction, then one FP64 atomic per sum
    constexpr int NS = BWD ? 6 : 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        double x = acc[i];
        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
        if (lane == 0) red[wave][i] = x;
    }
    __syncthreads();
    if (threadIdx.x < NS) {
        double x = 0.0;
        for (int k = 0; k < NW; ++k) x += red[k][threadIdx.x];
        unsafeAtomicAdd(A.acc + threadIdx.x, x);
    }
}

}  // namespace ope_hip
