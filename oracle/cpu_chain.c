/*
 * FP32 C restatement of OCTproEngine's CPU backend chain.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a second checker (the float32
 * arithmetic the reference performs) and by bench.py as the timed CPU
 * baseline ("port").  No product code path links or loads it.
 *
 * Each step follows the reference CPU backend in float32, in the same order:
 *   convertInputData          cpu_kernels.tpp:24-54
 *   rollingAverageDCRemoval   cpu_kernels.tpp:56-82  (float cumulative sum)
 *   kLinearization{Linear,Cubic,Lanczos}  cpu_kernels.tpp:84-196
 *   applyWindow               cpu_kernels.tpp:227-235
 *   dispersionCompensation    cpu_kernels.tpp:198-225
 *   computeIFFT<float>        cpu_kernels.tpp:238-267 (unnormalised backward)
 *   meanALineSubtraction      cpu_kernels.tpp:460-479
 *   logScaleAndTruncate       cpu_kernels.tpp:301-341
 *   linearScaleAndTruncate    cpu_kernels.tpp:343-378
 *   getMinimumVarianceMean    cpu_kernels.tpp:402-458
 * Orchestration: CpuBackend::Impl::processData, cpu_backend.cpp:169-302.
 *
 * The reference links FFTW3f (FFTW_BACKWARD | FFTW_ESTIMATE,
 * cpu_backend.cpp:385-391).  FFTW is not installed in this image, so the
 * inverse DFT here is a self-written float32 Stockham radix-4/2 FFT with
 * twiddles rounded from double.  Every report that times this file says so.
 *
 * FLOAT32 input: the reference reinterprets it as uint32 bits; this file
 * reads IEEE floats when dtype == OPE_CPU_F32_IEEE (the HIP backend's
 * semantics) and uint32 bits when dtype == OPE_CPU_U32.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { OPE_CPU_U8 = 1, OPE_CPU_U16 = 2, OPE_CPU_U32 = 3, OPE_CPU_F32_IEEE = 9 };

typedef struct {
    int32_t signal_length;
    int32_t n_ascans;
    int32_t dtype;
    int32_t resampling;
    int32_t interpolation; /* 0 linear, 1 cubic, 2 lanczos (a = 8), 3 lanczos (a = 4) */
    int32_t windowing;
    int32_t dispersion;    /* applied only if phasor_len == 2*N */
    int32_t dc_removal;
    int32_t dc_window;
    int32_t log_scaling;
    float grayscale_min;
    float grayscale_max;
    float multiplicator;
    float addend;
    int32_t fpn_pairs;     /* 0: no subtraction */
    int32_t phasor_len;    /* number of floats in phasor */
} ope_cpu_params;

/* ---------------------------------------------------------------- FFT */

typedef struct {
    int n;
    float *twr, *twi; /* exp(+2 pi i k / n), k < n */
} fft_plan;

static void plan_init(fft_plan *p, int n) {
    p->n = n;
    p->twr = (float *)malloc(sizeof(float) * (size_t)n);
    p->twi = (float *)malloc(sizeof(float) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double a = 2.0 * 3.14159265358979323846 * (double)k / (double)n;
        p->twr[k] = (float)cos(a);
        p->twi[k] = (float)sin(a);
    }
}

static void plan_free(fft_plan *p) {
    free(p->twr);
    free(p->twi);
}

/* Unnormalised inverse DFT, X[k] = sum_j x[j] exp(+2 pi i j k / n), for
 * any power-of-two n (falls back to a direct O(n^2) sum otherwise).
 * Stockham autosort: radix-4 passes, one radix-2 pass if log2(n) is odd.
 * (xr, xi) in; result written back to (xr, xi); (yr, yi) scratch. */
static void ifft_inplace(const fft_plan *p, float *xr, float *xi, float *yr, float *yi) {
    const int n = p->n;
    if (n & (n - 1)) {
        for (int k = 0; k < n; ++k) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; ++j) {
                long e = ((long)j * k) % n;
                sr += (double)xr[j] * p->twr[e] - (double)xi[j] * p->twi[e];
                si += (double)xr[j] * p->twi[e] + (double)xi[j] * p->twr[e];
            }
            yr[k] = (float)sr;
            yi[k] = (float)si;
        }
        memcpy(xr, yr, sizeof(float) * n);
        memcpy(xi, yi, sizeof(float) * n);
        return;
    }
    float *ar = xr, *ai = xi, *br = yr, *bi = yi;
    int ns = 1;
    while (ns < n) {
        int radix = ((n / ns) % 4 == 0) ? 4 : 2;
        int m = n / radix;          /* butterflies */
        int tstep = n / (ns * radix); /* twiddle index step */
        if (radix == 4) {
            for (int b = 0; b < m; ++b) {
                int k = b % ns;
                int o = (b / ns) * ns * 4 + k;
                float r0 = ar[b], i0 = ai[b];
                float r1 = ar[b + m], i1 = ai[b + m];
                float r2 = ar[b + 2 * m], i2 = ai[b + 2 * m];
                float r3 = ar[b + 3 * m], i3 = ai[b + 3 * m];
                if (k) {
                    int e1 = k * tstep, e2 = 2 * e1, e3 = 3 * e1;
                    float t;
                    t = r1 * p->twr[e1] - i1 * p->twi[e1]; i1 = r1 * p->twi[e1] + i1 * p->twr[e1]; r1 = t;
                    t = r2 * p->twr[e2] - i2 * p->twi[e2]; i2 = r2 * p->twi[e2] + i2 * p->twr[e2]; r2 = t;
                    t = r3 * p->twr[e3] - i3 * p->twi[e3]; i3 = r3 * p->twi[e3] + i3 * p->twr[e3]; r3 = t;
                }
                float s02r = r0 + r2, s02i = i0 + i2, d02r = r0 - r2, d02i = i0 - i2;
                float s13r = r1 + r3, s13i = i1 + i3, d13r = r1 - r3, d13i = i1 - i3;
                br[o] = s02r + s13r;           bi[o] = s02i + s13i;
                br[o + ns] = d02r - d13i;      bi[o + ns] = d02i + d13r;      /* + i*d13 */
                br[o + 2 * ns] = s02r - s13r;  bi[o + 2 * ns] = s02i - s13i;
                br[o + 3 * ns] = d02r + d13i;  bi[o + 3 * ns] = d02i - d13r;  /* - i*d13 */
            }
        } else {
            for (int b = 0; b < m; ++b) {
                int k = b % ns;
                int o = (b / ns) * ns * 2 + k;
                float r0 = ar[b], i0 = ai[b], r1 = ar[b + m], i1 = ai[b + m];
                if (k) {
                    int e = k * tstep;
                    float t = r1 * p->twr[e] - i1 * p->twi[e];
                    i1 = r1 * p->twi[e] + i1 * p->twr[e];
                    r1 = t;
                }
                br[o] = r0 + r1;       bi[o] = i0 + i1;
                br[o + ns] = r0 - r1;  bi[o + ns] = i0 - i1;
            }
        }
        float *tr = ar, *ti = ai;
        ar = br; ai = bi; br = tr; bi = ti;
        ns *= radix;
    }
    if (ar != xr) {
        memcpy(xr, ar, sizeof(float) * n);
        memcpy(xi, ai, sizeof(float) * n);
    }
}

/* ---------------------------------------------------------------- stages */

static float cubic_hermite(float y0, float y1, float y2, float y3, float pos) {
    const float a = -y0 + 3.0f * (y1 - y2) + y3;
    const float b = 2.0f * y0 - 5.0f * y1 + 4.0f * y2 - y3;
    const float c = -y0 + y2;
    const float pos2 = pos * pos;
    return 0.5f * pos * (a * pos2 + b * pos + c) + y1;
}

static float lanczos(float x, int a) {
    if (fabsf(x) < 1e-5f) return 1.0f;
    if (fabsf(x) >= (float)a) return 0.0f;
    const float pi = (float)3.14159265358979323846;
    float px = pi * x;
    float pxa = px / (float)a;
    return (sinf(px) / px) * (sinf(pxa) / pxa);
}

static int imin(int a, int b) { return a < b ? a : b; }

static void chain_one(const ope_cpu_params *P, const fft_plan *plan, const void *raw_ascan,
                      const float *curve, const float *window, const float *phasor,
                      const float *fpn, float *out, float *x, float *cs, float *yr, float *yi,
                      float *sr, float *si) {
    const int n = P->signal_length;
    /* convertInputData */
    if (P->dtype == OPE_CPU_U8) {
        const uint8_t *in = (const uint8_t *)raw_ascan;
        for (int i = 0; i < n; ++i) x[i] = (float)in[i];
    } else if (P->dtype == OPE_CPU_U16) {
        const uint16_t *in = (const uint16_t *)raw_ascan;
        for (int i = 0; i < n; ++i) x[i] = (float)in[i];
    } else if (P->dtype == OPE_CPU_U32) {
        const uint32_t *in = (const uint32_t *)raw_ascan;
        for (int i = 0; i < n; ++i) x[i] = (float)in[i];
    } else {
        const float *in = (const float *)raw_ascan;
        for (int i = 0; i < n; ++i) x[i] = in[i];
    }
    /* rollingAverageDCRemoval, float cumulative sum */
    if (P->dc_removal) {
        long w = P->dc_window;
        cs[0] = 0.0f;
        for (int i = 0; i < n; ++i) cs[i + 1] = cs[i] + x[i];
        for (long i = 0; i < n; ++i) {
            long start = (w >= 1) ? ((i >= w - 1) ? i - (w - 1) : 0) : 0;
            long end = (w >= 1) ? ((i + w < n - 1) ? i + w : n - 1) : i;
            float sum = cs[end + 1] - cs[start];
            x[i] -= sum / (float)(end - start + 1);
        }
    }
    /* kLinearization */
    if (P->resampling) {
        for (int j = 0; j < n; ++j) {
            float nx = curve[j];
            float v;
            if (P->interpolation == 0) {
                int n1 = (int)nx;
                int n2 = n1 + 1;
                n1 = imin(n1, n - 1);
                n2 = imin(n2, n - 1);
                float y1 = x[n1], y2 = x[n2];
                v = y1 + (y2 - y1) * (nx - n1);
            } else if (P->interpolation == 1) {
                int n1 = (int)nx;
                int n0 = abs(n1 - 1);
                int n2 = n1 + 1;
                int n3 = n2 + 1;
                n0 = imin(n0, n - 1); n1 = imin(n1, n - 1);
                n2 = imin(n2, n - 1); n3 = imin(n3, n - 1);
                v = cubic_hermite(x[n0], x[n1], x[n2], x[n3], nx - n1);
            } else {
                /* Lanczos: a = 8 (method 2, the reference's), a = 4 (method 3) */
                const int a = P->interpolation == 3 ? 4 : 8;
                int n0 = (int)nx;
                float sum = 0.0f;
                for (int i = -a + 1; i <= a; ++i) {
                    int idx = n0 + i;
                    if (idx < 0) idx = 0;
                    if (idx >= n) idx = n - 1;
                    sum += x[idx] * lanczos(nx - idx, a);
                }
                v = sum;
            }
            yr[j] = v;
        }
    } else {
        memcpy(yr, x, sizeof(float) * n);
    }
    /* applyWindow */
    if (P->windowing)
        for (int j = 0; j < n; ++j) yr[j] *= window[j];
    /* dispersionCompensation: real x complex */
    if (P->dispersion && P->phasor_len == 2 * n) {
        for (int j = 0; j < n; ++j) {
            float re = yr[j];
            yr[j] = re * phasor[2 * j];
            yi[j] = re * phasor[2 * j + 1];
        }
    } else {
        memset(yi, 0, sizeof(float) * n);
    }
    /* computeIFFT */
    ifft_inplace(plan, yr, yi, sr, si);
    /* meanALineSubtraction */
    const int half = n / 2;
    if (fpn && P->fpn_pairs > 0) {
        int use = imin(half, P->fpn_pairs);
        for (int k = 0; k < use; ++k) {
            yr[k] -= fpn[2 * k];
            yi[k] -= fpn[2 * k + 1];
        }
    }
    /* log / linear scale and truncate */
    const float coeff = P->multiplicator, addend = P->addend;
    float vmin = P->grayscale_min, vmax = P->grayscale_max;
    const float len = (float)half;
    if (P->log_scaling) {
        for (int k = 0; k < half; ++k) {
            float m2 = yr[k] * yr[k] + yi[k] * yi[k];
            sr[k] = 10.0f * log10f(m2 / len);
        }
        if (P->grayscale_min == P->grayscale_max) {
            vmin = sr[0]; vmax = sr[0];
            for (int k = 1; k < half; ++k) {
                if (sr[k] < vmin) vmin = sr[k];
                if (sr[k] > vmax) vmax = sr[k];
            }
        }
        float range = vmax - vmin;
        if (range < 1e-6f) range = 1.0f;
        for (int k = 0; k < half; ++k) out[k] = coeff * ((sr[k] - vmin) / range + addend);
    } else {
        float range = vmax - vmin;
        if (range < 1e-6f) range = 1.0f;
        for (int k = 0; k < half; ++k) {
            float mag = sqrtf(yr[k] * yr[k] + yi[k] * yi[k]) / len;
            out[k] = coeff * ((mag - vmin) / range + addend);
        }
    }
}

static size_t dtype_bytes(int dtype) {
    return dtype == OPE_CPU_U8 ? 1 : dtype == OPE_CPU_U16 ? 2 : 4;
}

/* Process a frame of P->n_ascans A-scans.  out is [n_ascans][N/2].
 * nthreads <= 1 runs the single worker thread of the reference
 * (cpu_backend.cpp:428-431); nthreads > 1 splits the A-scan loop. */
int ope_cpu_process(const ope_cpu_params *P, const void *raw, const float *curve,
                    const float *window, const float *phasor, const float *fpn, float *out,
                    int nthreads) {
    const int n = P->signal_length;
    if (n <= 1 || P->n_ascans < 0) return -1;
    fft_plan plan;
    plan_init(&plan, n);
    const size_t stride = (size_t)n * dtype_bytes(P->dtype);
    int nt = nthreads > 1 ? nthreads : 1;
#pragma omp parallel num_threads(nt)
    {
        float *buf = (float *)malloc(sizeof(float) * (size_t)(6 * n + 8));
        float *x = buf, *cs = buf + n, *yr = cs + n + 1, *yi = yr + n, *sr = yi + n, *si = sr + n;
#pragma omp for schedule(static)
        for (int a = 0; a < P->n_ascans; ++a) {
            chain_one(P, &plan, (const uint8_t *)raw + (size_t)a * stride, curve, window, phasor,
                      fpn, out + (size_t)a * (n / 2), x, cs, yr, yi, sr, si);
        }
        free(buf);
    }
    plan_free(&plan);
    return 0;
}

/* Complex IFFT outputs (before FPN) for the first n_rows A-scans, written as
 * interleaved float [n_rows][N][2]; the input to getMinimumVarianceMean. */
int ope_cpu_spectra(const ope_cpu_params *P, const void *raw, const float *curve,
                    const float *window, const float *phasor, int n_rows, float *spec) {
    const int n = P->signal_length;
    fft_plan plan;
    plan_init(&plan, n);
    const size_t stride = (size_t)n * dtype_bytes(P->dtype);
    float *buf = (float *)malloc(sizeof(float) * (size_t)(7 * n + 8));
    float *x = buf, *cs = buf + n, *yr = cs + n + 1, *yi = yr + n, *sr = yi + n, *si = sr + n;
    float *dummy = si + n;
    ope_cpu_params Q = *P;
    Q.fpn_pairs = 0;
    for (int a = 0; a < n_rows; ++a) {
        /* re-run the chain and capture the spectrum before scaling */
        const void *src = (const uint8_t *)raw + (size_t)a * stride;
        Q.log_scaling = 0;
        chain_one(&Q, &plan, src, curve, window, phasor, NULL, dummy, x, cs, yr, yi, sr, si);
        for (int k = 0; k < n; ++k) {
            spec[((size_t)a * n + k) * 2] = yr[k];
            spec[((size_t)a * n + k) * 2 + 1] = yi[k];
        }
    }
    free(buf);
    plan_free(&plan);
    return 0;
}

/* getMinimumVarianceMean (cpu_kernels.tpp:402-458): input interleaved
 * complex [height][width]; output interleaved [width][2]. */
int ope_cpu_min_variance_mean(const float *in, int width, int height, int segments, float *out) {
    if (height <= 0 || width <= 0 || segments <= 0) return -1;
    int seg_w = height / segments;
    float factor = 1.0f / (float)seg_w;
    for (int col = 0; col < width; ++col) {
        float min_var = INFINITY, mx_best = 0.0f, my_best = 0.0f;
        for (int s = 0; s < segments; ++s) {
            double sx = 0.0, sy = 0.0, sxx = 0.0;
            for (int r = 0; r < seg_w; ++r) {
                const float *c = in + ((size_t)(s * seg_w + r) * width + col) * 2;
                double dx = c[0], dy = c[1];
                sx += dx;
                sy += dy;
                sxx += dx * dx + dy * dy;
            }
            float mx = (float)(sx * (double)factor);
            float my = (float)(sy * (double)factor);
            float var = (float)((sxx * (double)factor) - ((double)mx * (double)mx + (double)my * (double)my));
            if (var < min_var) {
                min_var = var;
                mx_best = mx;
                my_best = my;
            }
        }
        out[2 * col] = mx_best;
        out[2 * col + 1] = my_best;
    }
    return 0;
}

/* Standalone unnormalised inverse DFT of interleaved complex data, n_lines
 * lines of length n (for checking the FFT against the DFT definition). */
int ope_cpu_ifft(const float *in, int n, int n_lines, float *out) {
    fft_plan plan;
    plan_init(&plan, n);
    float *buf = (float *)malloc(sizeof(float) * 4 * (size_t)n);
    for (int l = 0; l < n_lines; ++l) {
        for (int k = 0; k < n; ++k) {
            buf[k] = in[((size_t)l * n + k) * 2];
            buf[n + k] = in[((size_t)l * n + k) * 2 + 1];
        }
        ifft_inplace(&plan, buf, buf + n, buf + 2 * n, buf + 3 * n);
        for (int k = 0; k < n; ++k) {
            out[((size_t)l * n + k) * 2] = buf[k];
            out[((size_t)l * n + k) * 2 + 1] = buf[n + k];
        }
    }
    free(buf);
    plan_free(&plan);
    return 0;
}
