// Fused OCT processing chain for CDNA4 (gfx950).
//
// One kernel runs the whole per-A-scan chain of the reference's
// CpuBackend::Impl::processData (cpu_backend.cpp:184-281) in a single pass
// over HBM: raw samples are read once (128-bit loads), staged in LDS,
// DC-removed, resampled by a banded tap table, windowed, multiplied by the
// dispersion phasor, transformed by a radix-8 Stockham inverse FFT that lives
// in registers and LDS, FPN-corrected and mapped to log/linear grayscale; only
// the N/2 output floats are written.
//
// Work layout: an A-scan of N samples is owned by TPA = N/8 lanes; lane t
// holds the 8 complex values at positions t + e*TPA (e = 0..7) throughout the
// FFT.  Every Stockham pass reads those positions from LDS, so all LDS reads
// are unit-stride across lanes; pass outputs are written to padded layouts
// (pad_slots, pad_pass1) that keep the strided writes free of bank
// conflicts.  Because a lane always handles the same positions, the
// per-lane tables (taps, phasor, twiddles, FPN profile, background) are read
// once per kernel launch and stay in VGPRs while the workgroup loops over
// A-scans (persistent grid-stride loop).  Without dispersion two real
// A-scans share one complex FFT (M_PAIR).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ope_hip {

// ------------------------------------------------------------------ complex

struct cf {
    float x, y;
};

__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) {
    return {fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x)};
}
// +i*a and -i*a
__device__ __forceinline__ cf cmul_pi(cf a) { return {-a.y, a.x}; }
__device__ __forceinline__ cf cmul_ni(cf a) { return {a.y, -a.x}; }

// Inverse (sign +) DFTs of size 2, 4, 8 in registers.
__device__ __forceinline__ void idft2(cf& a0, cf& a1) {
    cf s = cadd(a0, a1), d = csub(a0, a1);
    a0 = s;
    a1 = d;
}

__device__ __forceinline__ void idft4(cf& a0, cf& a1, cf& a2, cf& a3) {
    cf s02 = cadd(a0, a2), d02 = csub(a0, a2);
    cf s13 = cadd(a1, a3), d13 = cmul_pi(csub(a1, a3));
    a0 = cadd(s02, s13);
    a1 = cadd(d02, d13);
    a2 = csub(s02, s13);
    a3 = csub(d02, d13);
}

__device__ __forceinline__ void idft8(cf* v) {
    // even / odd halves, then combine with w^m, w = exp(+i*pi/4)
    cf e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cf o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    idft4(e0, e1, e2, e3);
    idft4(o0, o1, o2, o3);
    // w * o1 = c (o1.x - o1.y, o1.x + o1.y), w^3 * o3 = c (-(o3.x + o3.y), o3.x - o3.y),
    // c = 1/sqrt(2); the scaling by c is folded into the combining FMAs.
    const float c = 0.70710678118654752440f;
    const cf u1 = {o1.x - o1.y, o1.x + o1.y};
    const cf u3 = {o3.x + o3.y, o3.x - o3.y};
    const cf t2 = cmul_pi(o2);                          // w^2 * o2
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = {fmaf(c, u1.x, e1.x), fmaf(c, u1.y, e1.y)};
    v[5] = {fmaf(-c, u1.x, e1.x), fmaf(-c, u1.y, e1.y)};
    v[2] = cadd(e2, t2);
    v[6] = csub(e2, t2);
    v[3] = {fmaf(-c, u3.x, e3.x), fmaf(c, u3.y, e3.y)};
    v[7] = {fmaf(c, u3.x, e3.x), fmaf(-c, u3.y, e3.y)};
}

// ------------------------------------------------------------------ geometry

__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// Orders the LDS accesses of the lanes that own one A-scan (TPA of them).
// Up to 64 lanes they are one wavefront, whose LDS instructions execute in
// order, and every LDS array of that A-scan is private to it: a wavefront
// barrier (compiler fences, no hardware barrier) then suffices and the
// workgroup's other wavefronts are not held up.  Otherwise a workgroup
// barrier.
template <int TPA>
__device__ __forceinline__ void ascan_sync() {
    if constexpr (TPA <= 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Padded LDS slot of complex element i (one float2 of padding per 32
// elements, one more per 128).  Every shift is >= 3, so the Stockham read
// (t + e*TPA) and write ((b/ns)*ns*R + b%ns + m*ns) addresses of a lane
// split into one base register plus compile-time offsets: the terms share no
// bits, and the padding is additive over such sums.  Among paddings of this
// family it gave the fewest SQ_LDS_BANK_CONFLICT cycles on gfx950.
__host__ __device__ constexpr int pad_slots(int i) { return i + 2 * (i >> 5) + (i >> 7); }

// Padding of the buffer written by Stockham pass 1 (and read by pass 2).
// Pass 1 writes runs of 8 consecutive slots 64 apart; with pad_slots two of
// every eight such runs meet in the same LDS banks (64 banks of 4 bytes on
// gfx950), with this padding none do.  Same family, so the same base-plus-
// immediate split applies.  Chosen by counting bank conflicts of the
// kernel's access pattern for every length and pass.
__host__ __device__ constexpr int pad_pass1(int i) { return i + 2 * (i >> 5) + 4 * (i >> 6); }

// padding of the buffer written by pass p
__host__ __device__ constexpr int pad_of_pass(int p, int i) { return p == 1 ? pad_pass1(i) : pad_slots(i); }

template <int N>
struct Geo {
    static constexpr int LOG2N = ilog2(N);
    static constexpr int TPA = N / 8;                       // lanes per A-scan
    static constexpr int WG = TPA >= 256 ? TPA : 256;       // workgroup of the autofocus and operation kernels
    static constexpr int APW = WG / TPA;                    // A-scans per chain_kernel workgroup
    static constexpr int NPASS = (LOG2N + 2) / 3;           // Stockham passes
    static constexpr int LAST_R = N >> (3 * (NPASS - 1));   // radix of the one non-radix-8 pass
    // padded float2 slots (the larger of the two paddings); at least N+1 so
    // buffer A also holds the N+1 prefix sums (int or double) of the
    // DC-removal prologue
    static constexpr int PADN_0 = pad_slots(N - 1) + 1 > pad_pass1(N - 1) + 1 ? pad_slots(N - 1) + 1
                                                                               : pad_pass1(N - 1) + 1;
    static constexpr int PADN = PADN_0 > N + 1 ? PADN_0 : N + 1;
    static constexpr int HALF = N / 2;
    static_assert((1 << LOG2N) == N, "N must be a power of two");
    static_assert(N >= 256 && N <= 4096, "N outside 256..4096");
};

// Workgroup of chain_kernel: one A-scan from N = 1024 on (a workgroup
// barrier then holds up only the wavefronts that share the A-scan's LDS
// arrays), 256 lanes below (where the syncs are per wavefront anyway).
template <int N>
constexpr int chain_wg() {
    return Geo<N>::TPA >= 128 ? Geo<N>::TPA : 256;
}


// ------------------------------------------------------------------ arguments

enum : int {
    F_DISP = 1,      // multiply by the dispersion phasor
    F_FPN = 4,       // subtract the FPN profile before the magnitude
    F_BG = 8,        // subtract the post-process background
    F_LOG = 16,      // log scaling (else linear)
    F_AUTOMM = 32,   // log scaling with per-A-scan min/max
    F_BITSHIFT = 64, // input samples >> 4 (8/16-bit) or * 2^-32 (32-bit integers)
    F_FLIP = 128,    // reverse the A-scan order of B-scans 0, 2, 4, ...
    F_OFFSET = 256,  // samples minus *offset; the epilogue constant `fpn` then
                     // holds fpn - offset * IFFT(S * phasor) (see ChainArgs)
};

// One A-scan of a ragged batch (M_ROWS): where it is read and written, and
// the request it belongs to.
struct RowRef {
    const void* in;          // N samples
    float* out;              // N/2 outputs
    long long req;
};

struct ChainArgs {
    const void* in;          // raw samples [n_ascans][N]
    float* out;              // [n_ascans][N/2]
    long long n_ascans;
    const int* tap_base;     // [N]   first input index of the tap window
    const float* tap_w;      // [N][T] weights, window folded in
    const float2* phasor;    // [N]   (cos, sin)
    const float2* twiddle;   // [N]   exp(+2 pi i k / N)
    const float2* fpn;       // [N/2]
    const float* bg;         // [N/2]
    int in_type;             // OPE_HIP_UINT8 / UINT16 / UINT32 / FLOAT32 classes
    int flags;
    int dc_window;
    float ep_a, ep_b;        // fixed epilogue: out = ep_a * f(|X|^2) + ep_b
    float coeff, addend;     // auto min/max epilogue
    float log_c, log_h;      // v = log_c * (log2(m2) - log_h) = 10 log10(m2/(N/2))
    float bg_w, bg_off;
    int ascans_per_bscan;    // B-scan flip: B-scan length
    long long flip_bscans;   // B-scans b with b even and b + 1 < flip_bscans are flipped
    // Sample offset (F_OFFSET): the chain runs on x - m0, m0 = *offset, so
    // that the float arithmetic does not carry the fringe's constant level;
    // by linearity X = X' + m0 * D with D[k] = IFFT(S * phasor)[k], S[j] the
    // sum of output j's tap weights.  The host folds -m0 * D (FP64) into the
    // epilogue constant `fpn`; the statistics mode adds m0 * D back itself.
    const float* offset;     // device scalar m0
    const double2* offset_tab;  // [N/2] D (statistics mode)
    // FPN statistics mode
    double* acc;             // [8][N/2][3] (sum re, sum im, sum |X|^2)
    long long row0;          // accumulated row index of A-scan 0 of this call
    int seg_w;               // rows per segment
    int stat_rows;           // A-scans of this call that enter the statistics
    // Row map (M_ROWS kernels): A-scan a of the call reads
    // rows[a].in and writes rows[a].out instead of rows a of in and out.
    // Once an A-scan's outputs are issued, its first lane raises
    // finish[req] to the wall clock (an atomic max without return, so the
    // wavefront does not wait on it; the clock only grows, so no reset is
    // needed between runs).
    const RowRef* rows;
    unsigned long long* finish;
    // Dynamic schedule of a row map (tickets not null): workgroups take
    // chunks of kRowChunk consecutive groups of A-scans from tickets[0],
    // three at the start and one more each time they enter a new chunk, so
    // that several launches on the same map (one beside other kernels, one
    // after them) share the work.  The last of the ticket_wgs workgroups of
    // those launches to finish clears tickets[0..1] for the next run.
    unsigned* tickets;
    unsigned ticket_wgs;
};

// groups of A-scans per ticket of a dynamic row-map schedule (at least 4:
// the pipeline looks up to three groups ahead while it knows two chunks)
constexpr int kRowChunk = 4;

// input classes (numeric values match ope::DataType)
enum : int { IN_U8 = 1, IN_U16 = 2, IN_U32 = 3, IN_F32 = 9 };

// ------------------------------------------------------------------ global memory

// Pointers that live in argument structs are generic; every HBM access of
// the chain goes through these casts so that it is a global (not flat)
// instruction also where the body is not inlined into a kernel.
#define OPE_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ const OPE_GLOBAL T* gmem(const T* p) {
    return (const OPE_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ OPE_GLOBAL T* gmem(T* p) {
    return (OPE_GLOBAL T*)p;
}

typedef float f32x2 __attribute__((ext_vector_type(2)));

// p[i] of a float2 table through a global load
__device__ __forceinline__ float2 gload2(const float2* p, int i) {
    const f32x2 v = *gmem(reinterpret_cast<const f32x2*>(p) + i);
    return make_float2(v.x, v.y);
}

// ------------------------------------------------------------------ loads

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Loads the 8 consecutive raw samples [8t, 8t+8) of A-scan a (streaming,
// non-temporal: every input byte is read exactly once).  IW is the sample
// width in bytes; it is a compile-time constant so that the loaded words
// stay in VGPRs and their first use is the only wait on the load.
template <int IW>
__device__ __forceinline__ void load_raw(const void* in, int n, long long a, int t, uint32_t (&raw)[8]) {
    if constexpr (IW == 2) {
        const OPE_GLOBAL u32x4* p = gmem(reinterpret_cast<const u32x4*>(static_cast<const uint16_t*>(in) + a * n) + t);
        u32x4 q = __builtin_nontemporal_load(p);
        raw[0] = q.x; raw[1] = q.y; raw[2] = q.z; raw[3] = q.w;
    } else if constexpr (IW == 1) {
        const OPE_GLOBAL u32x2* p = gmem(reinterpret_cast<const u32x2*>(static_cast<const uint8_t*>(in) + a * n) + t);
        u32x2 q = __builtin_nontemporal_load(p);
        raw[0] = q.x; raw[1] = q.y;
    } else {
        const OPE_GLOBAL u32x4* p = gmem(reinterpret_cast<const u32x4*>(static_cast<const uint32_t*>(in) + a * n) + 2 * t);
        u32x4 q0 = __builtin_nontemporal_load(p);
        u32x4 q1 = __builtin_nontemporal_load(p + 1);
        raw[0] = q0.x; raw[1] = q0.y; raw[2] = q0.z; raw[3] = q0.w;
        raw[4] = q1.x; raw[5] = q1.y; raw[6] = q1.z; raw[7] = q1.w;
    }
}

// Samples to float (cpu_kernels.tpp:33-52): unsigned integers by width;
// 4-byte words are IEEE floats when `ieee` is set (FLOAT32 input).  `shift`
// is the CUDA backend's bitshift option (cuda_kernels.cu:45-70): 8/16-bit
// samples >> 4, 32-bit integers * 2^-32 (exact: a power-of-two scaling of
// the rounded value); IEEE floats are not shifted.  Then the sample offset
// m0 is subtracted (0 without F_OFFSET): exact for 8/16-bit samples, whose
// m0 is an integer.
template <int IW>
__device__ __forceinline__ void convert_raw(bool ieee, bool shift, float m0, const uint32_t (&raw)[8],
                                            float (&x)[8]) {
    // 8/16-bit fields convert with one sub-dword-select conversion each; the
    // shift then is an exact float division by 16 and floor (v < 2^16).
    if constexpr (IW == 2) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            x[2 * i] = (float)(uint16_t)raw[i];
            x[2 * i + 1] = (float)(uint16_t)(raw[i] >> 16);
        }
        if (shift) {
#pragma unroll
            for (int i = 0; i < 8; ++i) x[i] = floorf(x[i] * 0.0625f);
        }
    } else if constexpr (IW == 1) {
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = (float)(uint8_t)(raw[i >> 2] >> (8 * (i & 3)));
        if (shift) {
#pragma unroll
            for (int i = 0; i < 8; ++i) x[i] = floorf(x[i] * 0.0625f);
        }
    } else {
        const float scale = (shift && !ieee) ? 2.3283064365386963e-10f : 1.0f;  // 2^-32
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = ieee ? __uint_as_float(raw[i]) : (float)raw[i] * scale;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] -= m0;
}

// One sample as convert_raw converts it, before the offset.
__device__ __forceinline__ float convert_sample(const void* in, long long i, int iw, bool ieee, bool shift) {
    if (iw == 1) {
        const unsigned x = static_cast<const uint8_t*>(in)[i];
        return (float)(shift ? x >> 4 : x);
    }
    if (iw == 2) {
        const unsigned x = static_cast<const uint16_t*>(in)[i];
        return (float)(shift ? x >> 4 : x);
    }
    const uint32_t x = static_cast<const uint32_t*>(in)[i];
    return ieee ? __uint_as_float(x) : (shift ? (float)x * 2.3283064365386963e-10f : (float)x);
}

// ------------------------------------------------------------------ scans

// Lane v of a wavefront moved by the DPP control CTRL (rows of 16 lanes,
// ROW_MASK selects the rows written); lanes without a source get 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp(double v) {
    // the two 32-bit halves move alike: exact
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)dpp<CTRL, ROW_MASK>((int)b);
    const unsigned hi = (unsigned)dpp<CTRL, ROW_MASK>((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Inclusive prefix sum over aligned groups of G lanes (G = 32 or 64) of a
// wavefront: row shifts by 1, 2, 4, 8 within rows of 16 lanes, then the
// row totals through row broadcasts (VALU only, no LDS).
template <int G, typename T>
__device__ __forceinline__ T group_inclusive_scan(T v) {
    static_assert(G == 32 || G == 64, "groups of 32 or 64 lanes");
    v += dpp<0x111, 0xf>(v);                        // row_shr:1
    v += dpp<0x112, 0xf>(v);                        // row_shr:2
    v += dpp<0x114, 0xf>(v);                        // row_shr:4
    v += dpp<0x118, 0xf>(v);                        // row_shr:8
    v += dpp<0x142, 0xa>(v);                        // row_bcast:15 into rows 1 and 3
    if constexpr (G == 64) v += dpp<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// min/max reduction over the TPA lanes of one A-scan (result in all lanes).
template <int TPA>
__device__ __forceinline__ void ascan_minmax(float& mn, float& mx, float* part_mn, float* part_mx) {
    constexpr int W = TPA < 64 ? TPA : 64;
#pragma unroll
    for (int d = 1; d < W; d <<= 1) {
        mn = fminf(mn, __shfl_xor(mn, d, 64));
        mx = fmaxf(mx, __shfl_xor(mx, d, 64));
    }
    if constexpr (TPA > 64) {
        const int wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            part_mn[wave] = mn;
            part_mx[wave] = mx;
        }
        __syncthreads();
        const int first = (threadIdx.x / TPA) * (TPA / 64);
        mn = part_mn[first];
        mx = part_mx[first];
        for (int w = 1; w < TPA / 64; ++w) {
            mn = fminf(mn, part_mn[first + w]);
            mx = fmaxf(mx, part_mx[first + w]);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ FFT

// Radix schedule of the Stockham passes: every pass is radix 8 except, when
// N is not a power of 8, one pass of radix N / 8^(NPASS-1) at position
// SMALL_AT (SCHED < 0: the last pass).  Any position gives the same
// natural-order result; they differ in twiddle count and LDS access pattern.
template <int N, int SCHED>
struct FftPlan {
    static constexpr int NP = Geo<N>::NPASS;
    static constexpr int SMALL = Geo<N>::LAST_R;
    static constexpr int SMALL_AT = (SMALL == 8 || SCHED < 0 || SCHED >= NP) ? NP - 1 : SCHED;
    static constexpr int radix(int p) { return p == SMALL_AT ? SMALL : 8; }
    // product of the radices of passes 0..p-1
    static constexpr int ns(int p) {
        int v = 1;
        for (int q = 0; q < p; ++q) v *= radix(q);
        return v;
    }
};

// Twiddles of Stockham passes 1..NPASS-1 for lane t (pass 0 has Ns = 1: all
// ones); elements with r == 0 need none.  table = exp(+2 pi i k / N).
template <int N, int SCHED = -1, int P0 = 1>
__device__ __forceinline__ void load_twiddles(const float2* table, int t, cf (&tw)[Geo<N>::NPASS][8]) {
    using G = Geo<N>;
    using P = FftPlan<N, SCHED>;
    constexpr int NP = G::NPASS;
    constexpr int TPA = G::TPA;
#pragma unroll
    for (int p = 1; p < P0; ++p) {
#pragma unroll
        for (int e = 0; e < 8; ++e) tw[p][e] = {1.f, 0.f};  // passes read from an LDS table instead
    }
#pragma unroll
    for (int p = P0; p < NP; ++p) {
        const int R = P::radix(p);
        const int RQ = 8 / R;
        const int ns = P::ns(p);
        const int step = N / (ns * R);
#pragma unroll
        for (int e = RQ; e < 8; ++e) {
            const int q = e % RQ, r = e / RQ;
            const int b = t + q * TPA;
            const int ex = (r * (b & (ns - 1)) * step) & (N - 1);
            float2 w = gload2(table, ex);
            tw[p][e] = {w.x, w.y};
        }
    }
}

// N = 4096, one A-scan per lane, Stockham schedule: the twiddles of passes 1
// and 2 depend on t mod 8 and t mod 64 only, so a workgroup keeps them once
// in LDS (7 x 8 + 7 x 64 entries, the same floats as the global table) and
// every lane reads its 7 per pass there, next to the pass's data reads.
// That frees 28 VGPRs of the lane's loop-invariant tables: with the
// background constant read in the epilogue and the tap-window starts kept as
// 16-bit halves, the kernels with up to 2 taps fit 128 VGPRs, i.e. two
// 512-lane workgroups (four waves per SIMD) per CU instead of one
// (4096-point frames without resampling 64.1 -> 55-56 us, linear 67.4 -> 57).
// Cubic taps need 138: squeezed to 128 they spilled 40 bytes per lane and
// ran no faster (71.8 us against 71.6 for two A-scans per lane), so they
// keep the register tables.
template <int N, int T, int AS, int SCHED>
constexpr bool twl_fft() {
    return N == 4096 && T <= 2 && AS == 1 && !(SCHED == Geo<N>::NPASS);
}
constexpr int kTwlEntries = 7 * 8 + 7 * 64;

// entry i of the table: pass 1 (i < 56): w^(e*m*64), e = 1 + i/8, m = i%8;
// pass 2: w^(e*m*8), e = 1 + (i-56)/64, m = (i-56)%64 (w = exp(2 pi i/4096))
__device__ __forceinline__ int twl_exponent(int i) {
    return i < 56 ? (1 + i / 8) * (i % 8) * 64 : (1 + (i - 56) / 64) * ((i - 56) % 64) * 8;
}

// Work a caller runs inside an FFT once the last pass's LDS reads are
// issued: after the FFT's last workgroup barrier, so every wavefront has
// finished the FFT's first exchange and every LDS read that came before it.
// The default does nothing.
struct NoHook {
    __device__ void operator()() const {}
};

// Unnormalised inverse FFT (sign +) of the AS sequences whose elements
// t + e*N/8 lane t holds in v[u][e]; on return v[u][e] = X[t + e*N/8].
// Radix-8 Stockham passes (where N is not a power of 8, one of radix 4 or
// 2 at the position FftPlan gives) through the padded LDS buffers pa/pb
// (pass p writes pa for even p, pb for odd p).
// Ends without a barrier: the last pass's reads of pa/pb may still be in
// flight in other wavefronts.
template <int N, int AS, int SCHED = -1, typename Hook = NoHook, bool TWL = false>
__device__ __forceinline__ void stockham_inverse(cf (&v)[AS][8], float2* const (&pa)[AS], float2* const (&pb)[AS],
                                                 const cf (&tw)[Geo<N>::NPASS][8], Hook&& after_last_reads = {},
                                                 const float2* twl = nullptr) {
    using G = Geo<N>;
    using P = FftPlan<N, SCHED>;
    constexpr int NP = G::NPASS;
    constexpr int TPA = G::TPA;
    static_assert(!TWL || twl_fft<N, 1, AS, SCHED>(), "LDS twiddles: N = 4096, one A-scan per lane");
    const int t = threadIdx.x % TPA;
    // lane bases of the unit-stride reads, per padding
    const int rbase0 = pad_slots(t), rbase1 = pad_pass1(t);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const int R = P::radix(p);
        const int RQ = 8 / R;
        const int ns = P::ns(p);
        if (p > 0) {
#pragma unroll
            for (int u = 0; u < AS; ++u) {
                // pass p>0 reads what pass p-1 wrote
                const float2* rp = ((p % 2 == 1) ? pa[u] : pb[u]) + (p == 2 ? rbase1 : rbase0);
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    // t < TPA and e*TPA share no bits, so the padded
                    // index splits into a lane base plus a constant
                    float2 w = rp[pad_of_pass(p - 1, e * TPA)];
                    v[u][e] = {w.x, w.y};
                }
            }
            cf twp[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) twp[e] = tw[p][e];
            if (TWL && (p == 1 || p == 2)) {
                // this pass's twiddles from the workgroup's LDS table
                const float2* tp = p == 1 ? twl + (t & 7) - 8 : twl + 56 + (t & 63) - 64;
#pragma unroll
                for (int e = 1; e < 8; ++e) {
                    const float2 w = tp[e * (p == 1 ? 8 : 64)];
                    twp[e] = {w.x, w.y};
                }
            }
            if (p == NP - 1) after_last_reads();
#pragma unroll
            for (int u = 0; u < AS; ++u) {
#pragma unroll
                for (int e = RQ; e < 8; ++e) v[u][e] = cmul(v[u][e], twp[e]);
            }
        }
#pragma unroll
        for (int u = 0; u < AS; ++u) {
            if (R == 8) {
                idft8(v[u]);
            } else if (R == 4) {
                idft4(v[u][0], v[u][2], v[u][4], v[u][6]);
                idft4(v[u][1], v[u][3], v[u][5], v[u][7]);
            } else {
                idft2(v[u][0], v[u][4]);
                idft2(v[u][1], v[u][5]);
                idft2(v[u][2], v[u][6]);
                idft2(v[u][3], v[u][7]);
            }
        }
        if (p < NP - 1) {
            // Stockham placement of output m of butterfly b:
            //   (b / ns) * ns * R + b % ns + m * ns, padded
#pragma unroll
            for (int u = 0; u < AS; ++u) {
                float2* dst = (p % 2 == 0) ? pa[u] : pb[u];
#pragma unroll
                for (int q = 0; q < RQ; ++q) {
                    const int b = t + q * TPA;
                    float2* wp = dst + pad_of_pass(p, (b / ns) * ns * R + (b & (ns - 1)));
#pragma unroll
                    for (int m = 0; m < R; ++m) {
                        const int e = q + RQ * m;
                        // base has no bits where m*ns has any: the padded
                        // index splits into base and a constant
                        wp[pad_of_pass(p, m * ns)] = make_float2(v[u][e].x, v[u][e].y);
                    }
                }
            }
            ascan_sync<TPA>();
        }
    }
}

// ------------------------------------------------------------------ FFT with a register exchange
//
// For N = 1024 and 2048 (radices 8, 8, S, 8 with S = N/512) the exchange
// into the small-radix pass can stay inside each wavefront.  Four-step form,
// with n = TPA d0 + r0, r0 = (TPA/8) d1 + r1, r1 = 8 d2 + r2 and
// k = k0 + 8 k1 + 64 k2 + 64 S K3:
//   pass 0: DFT-8 over d0 (registers, lane t = r0), twiddle w_N^(r0 k0);
//   exchange 1 through LDS: registers d1; lane bits 0-2 hold k0, the top
//     in-wavefront lane bits (5, and 4 for S = 4) hold d2, the others r2;
//   pass 1: DFT-8 over d1, twiddle w_TPA^(r1 k1);
//   exchange 2 inside the wavefront: v_permlane32_swap transposes register
//     bit 2 with lane bit 5 and (S = 4) v_permlane16_swap register bit 1
//     with lane bit 4, so that d2 is in the registers and k1's top bits in
//     those lanes (one VALU instruction per dword pair, no LDS, no barrier);
//   pass 2: DFT-S over d2, twiddle w_(8S)^(r2 k2);
//   exchange 3 through LDS: registers r2, lane t = k0 + 8 k1 + 64 k2;
//   pass 3: DFT-8 over r2: v[e] = X[t + e*TPA], the layout stockham_inverse
//     leaves, so everything after the FFT is shared.
// The LDS addresses are sums of one constant per index bit (lane bits give
// the lane's base, register bits a compile-time offset); the constants make
// every 32-lane half of each 8-byte access hit 32 distinct bank pairs.
// One exchange fewer through LDS than stockham_inverse (whose writes cost
// about twice its reads), 8 or 16 swap instructions more per lane.

// the swap schedule is launch variant "small-radix pass = NPASS"
template <int N, int SCHED>
constexpr bool swap_sched() {
    return (N == 1024 || N == 2048 || N == 4096) && SCHED == Geo<N>::NPASS;
}

template <int N>
struct SwapFft {
    static_assert(N == 1024 || N == 2048, "swap schedule: N = 1024 or 2048");
    static constexpr int TPA = N / 8;
    static constexpr int S = N / 512;   // radix of pass 2
    static constexpr int RQ = 8 / S;    // pass-2 butterflies per lane
    static constexpr int E3 = N == 2048 ? 264 : 136;   // exchange-3 stride of r2
    // lane fields of the exchange-1 reader (and of the exchange-3 writer,
    // whose d2 lanes then hold k1's top bits)
    __device__ static int r2_of(int t) { return N == 2048 ? ((t >> 3) & 1) | ((t >> 5) & 6) : ((t >> 3) & 3) | ((t >> 4) & 4); }
    __device__ static int d2_of(int t) { return N == 2048 ? (t >> 4) & 3 : (t >> 5) & 1; }
    __device__ static int k1hi_of(int t) { return N == 2048 ? (t >> 3) & 6 : (t >> 3) & 4; }
    // exchange 1: address of (k0, d1, d2, r2) = r2 + 8 d2 + skew(k0) + c(d1)
    static constexpr int skew(int k0) {
        return N == 2048 ? 34 * (k0 & 1) + 68 * ((k0 >> 1) & 1) + 144 * (k0 >> 2)
                         : 36 * (k0 & 1) + 72 * ((k0 >> 1) & 1) + 144 * (k0 >> 2);
    }
    static constexpr int c_d1(int d1) { return N == 2048 ? 288 * d1 : 16 * (d1 & 1) + 284 * (d1 >> 1); }
    // writer lane t = r0: d1 = r0 / (TPA/8), r1 = r0 % (TPA/8)
    __device__ static int e1_write_base(int t) {
        return N == 2048 ? (t & 31) + 288 * (t >> 5) : (t & 31) + 284 * (t >> 5);
    }
    __device__ static int e1_read_base(int t) { return r2_of(t) + 8 * d2_of(t) + skew(t & 7); }
    // exchange 3: address of (k0, k1, k2, r2) = k0 + 8 k1 + 64 k2 + E3 r2;
    // the writer's register e holds k1 % RQ = e % RQ and k2 = e / RQ
    __device__ static int e3_write_base(int t) { return (t & 7) + 8 * k1hi_of(t) + E3 * r2_of(t); }
    static constexpr int e3_write_off(int e) { return 8 * (e % RQ) + 64 * (e / RQ); }
};

// The swap schedule at N = 1024 and 2048 with 8 taps (Lanczos-4), one A-scan
// per lane (and cubic with DC removal of 8/16-bit samples, where the freed
// registers take the banded DC path): the twiddles of its passes 1 and 2
// depend on r1 = 8 d2 + r2 (64 values) and r2 (8 values) only, so the
// workgroup keeps them once in LDS (the same floats as the global table, so
// results are bit-identical) and each lane reads its 7 + (8 - RQ) there
// next to the passes' data reads.  That frees the 2 x 7 complex registers
// those passes held: with the tap-window starts kept as 16-bit halves and
// the background constant read in the epilogue, as for N = 4096, the
// Lanczos-4 kernels fit three waves per SIMD instead of two (2048-point
// frames 65.5 -> 63.0 us, with DC removal 95.5 -> 85.4, the 2048 x 1024 x
// 256 volume 1.42-1.48 -> 1.27-1.30 ms).  Lanczos-8 stays at two waves with
// the registers freed and ran slower with the table (80.7 -> 87.6 us), so
// it keeps its register twiddles.
template <int N, int T, int AS, int SCHED>
constexpr bool twl_swap() {
    return (N == 1024 || N == 2048) && SCHED == Geo<N>::NPASS && AS == 1 && T == 8;
}

// entry i of that table: pass 1 (i < 448): w^(8 r1 e), e = 1 + i/64,
// r1 = i%64; pass 2: w^(64 r2 (e/RQ)), e = RQ + (i-448)/8, r2 = (i-448)%8
// (w = exp(2 pi i/N); entries past pass 2's are not read)
template <int N>
__device__ __forceinline__ int twl_swap_exponent(int i) {
    constexpr int RQ = SwapFft<N>::RQ;
    if (i < 448) return (8 * (i % 64) * (1 + i / 64)) & (N - 1);
    const int j = i - 448;
    return (64 * (j % 8) * ((RQ + j / 8) / RQ)) & (N - 1);
}

// ------------------------------------------------------------------ N = 4096 with a register exchange
//
// Radices 8, 8, 8, 8 with n = 512 d0 + 64 d1 + 8 d2 + r2 and
// k = k0 + 8 k1 + 64 k2 + 512 k3 (lane t of 512, one A-scan per workgroup):
//   pass 0: DFT-8 over d0 (registers, lane t = r0 = 64 d1 + 8 d2 + r2),
//     twiddle w_N^(r0 k0);
//   exchange 1 through LDS at (t & 63) + 64 k0 + 512 (t >> 6), read at
//     t + 512 d1: registers d1, lane t = r2 + 8 d2 + 64 k0, so d2 sits in
//     lane bits 3-5 of every wavefront;
//   pass 1: DFT-8 over d1, twiddle w_512^(r1 k1), r1 = t & 63;
//   exchange 2 inside the wavefront: register bits 2, 1, 0 trade places
//     with lane bits 5, 4, 3 (v_permlane32_swap, v_permlane16_swap, and a
//     DPP row rotation by 8 with a select): registers d2, lane bits 3-5 k1;
//   pass 2: DFT-8 over d2, twiddle w_64^(r2 k2), r2 = t & 7;
//   exchange 3 through LDS at k0 + 8 k1 + 64 k2 + 513 r2, read at
//     t + 513 r2: registers r2, lane t = k0 + 8 k1 + 64 k2;
//   pass 3: DFT-8 over r2: v[e] = X[t + 512 e], the layout stockham_inverse
//     leaves.
// Every 32-lane half of each 8-byte LDS access hits 32 distinct bank pairs
// (consecutive addresses; the odd stride 513 spreads the exchange-3
// writes), without padding; the highest address, 4102, fits the buffer.
struct Swap4096 {
    static constexpr int E3 = 513;
    __device__ static int e1_write_base(int t) { return (t & 63) + 512 * (t >> 6); }
    __device__ static int e3_write_base(int t) { return (t >> 6) + 8 * ((t >> 3) & 7) + E3 * (t & 7); }
};
static_assert(7 + 8 * 7 + 64 * 7 + Swap4096::E3 * 7 < Geo<4096>::PADN, "exchange 3 fits the buffer");

// Register bit 0 of the pair (a, b) trades places with lane bit 3: lanes
// whose bit 3 is set take b from lane t ^ 8 into a, the others a from
// lane t ^ 8 into b (DPP row_ror:8, which within a row of 16 lanes reads
// lane t ^ 8).
__device__ __forceinline__ void swap_lane_bit3(cf& a, cf& b) {
    const bool hi = (threadIdx.x & 8) != 0;
    const float pax = __int_as_float(dpp<0x128, 0xf>(__float_as_int(a.x)));
    const float pay = __int_as_float(dpp<0x128, 0xf>(__float_as_int(a.y)));
    const float pbx = __int_as_float(dpp<0x128, 0xf>(__float_as_int(b.x)));
    const float pby = __int_as_float(dpp<0x128, 0xf>(__float_as_int(b.y)));
    const cf na = {hi ? pbx : a.x, hi ? pby : a.y};
    const cf nb = {hi ? b.x : pax, hi ? b.y : pay};
    a = na;
    b = nb;
}

// Per-lane twiddles of the swap schedule: tw[0][e] = w_N^(t e) (pass 0, lane
// t = r0), tw[1][e] = w_TPA^(r1 e) (pass 1), tw[2][e] = w_(8S)^(r2 (e / RQ))
// (pass 2); for N = 4096 tw[1][e] = w_512^((t & 63) e), tw[2][e] =
// w_64^((t & 7) e); table = exp(+2 pi i k / N).
template <int N>
__device__ __forceinline__ void load_twiddles_swap(const float2* table, int t, cf (&tw)[Geo<N>::NPASS][8]) {
    if constexpr (N == 4096) {
#pragma unroll
        for (int e = 1; e < 8; ++e) {
            float2 w = gload2(table, (t * e) & (N - 1));
            tw[0][e] = {w.x, w.y};
            w = gload2(table, (8 * (t & 63) * e) & (N - 1));
            tw[1][e] = {w.x, w.y};
            w = gload2(table, (64 * (t & 7) * e) & (N - 1));
            tw[2][e] = {w.x, w.y};
        }
        return;
    }
    using F = SwapFft<N < 4096 ? N : 2048>;
    const int r1 = 8 * F::d2_of(t) + F::r2_of(t);
    const int r2 = F::r2_of(t);
#pragma unroll
    for (int e = 1; e < 8; ++e) {
        float2 w = gload2(table, (t * e) & (N - 1));
        tw[0][e] = {w.x, w.y};
        w = gload2(table, (8 * r1 * e) & (N - 1));
        tw[1][e] = {w.x, w.y};
        w = gload2(table, (64 * r2 * (e / F::RQ)) & (N - 1));
        tw[2][e] = {w.x, w.y};
    }
}

// v_permlane32_swap / v_permlane16_swap on the two dwords of a complex pair:
// a (register bit 0) keeps lanes whose bit 5 (4) is 0 and gains b's, b the
// rest: the register bit and the lane bit trade places.
__device__ __forceinline__ void swap_lane_bit5(cf& a, cf& b) {
    const auto x = __builtin_amdgcn_permlane32_swap(__float_as_uint(a.x), __float_as_uint(b.x), false, false);
    const auto y = __builtin_amdgcn_permlane32_swap(__float_as_uint(a.y), __float_as_uint(b.y), false, false);
    a = {__uint_as_float(x[0]), __uint_as_float(y[0])};
    b = {__uint_as_float(x[1]), __uint_as_float(y[1])};
}
__device__ __forceinline__ void swap_lane_bit4(cf& a, cf& b) {
    const auto x = __builtin_amdgcn_permlane16_swap(__float_as_uint(a.x), __float_as_uint(b.x), false, false);
    const auto y = __builtin_amdgcn_permlane16_swap(__float_as_uint(a.y), __float_as_uint(b.y), false, false);
    a = {__uint_as_float(x[0]), __uint_as_float(y[0])};
    b = {__uint_as_float(x[1]), __uint_as_float(y[1])};
}

// The swap schedule of N = 4096 (see Swap4096).
template <int AS, typename Hook>
__device__ __forceinline__ void fft_swap4096(cf (&v)[AS][8], float2* const (&pa)[AS], float2* const (&pb)[AS],
                                             const cf (&tw)[4][8], Hook&& after_last_reads) {
    constexpr int TPA = 512;
    const int t = threadIdx.x % TPA;
    // pass 0 and exchange 1
    const int w1 = Swap4096::e1_write_base(t);
#pragma unroll
    for (int u = 0; u < AS; ++u) {
        idft8(v[u]);
#pragma unroll
        for (int e = 1; e < 8; ++e) v[u][e] = cmul(v[u][e], tw[0][e]);
#pragma unroll
        for (int e = 0; e < 8; ++e) pa[u][w1 + 64 * e] = make_float2(v[u][e].x, v[u][e].y);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < AS; ++u) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float2 w = pa[u][t + 512 * e];
            v[u][e] = {w.x, w.y};
        }
    }
    // pass 1, exchange 2 in registers, pass 2
#pragma unroll
    for (int u = 0; u < AS; ++u) {
        idft8(v[u]);
#pragma unroll
        for (int e = 1; e < 8; ++e) v[u][e] = cmul(v[u][e], tw[1][e]);
#pragma unroll
        for (int e = 0; e < 4; ++e) swap_lane_bit5(v[u][e], v[u][e + 4]);
        swap_lane_bit4(v[u][0], v[u][2]);
        swap_lane_bit4(v[u][1], v[u][3]);
        swap_lane_bit4(v[u][4], v[u][6]);
        swap_lane_bit4(v[u][5], v[u][7]);
#pragma unroll
        for (int e = 0; e < 8; e += 2) swap_lane_bit3(v[u][e], v[u][e + 1]);
        idft8(v[u]);
#pragma unroll
        for (int e = 1; e < 8; ++e) v[u][e] = cmul(v[u][e], tw[2][e]);
    }
    // exchange 3 and pass 3
    const int w3 = Swap4096::e3_write_base(t);
#pragma unroll
    for (int u = 0; u < AS; ++u) {
#pragma unroll
        for (int e = 0; e < 8; ++e) pb[u][w3 + 64 * e] = make_float2(v[u][e].x, v[u][e].y);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < AS; ++u) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float2 w = pb[u][t + Swap4096::E3 * e];
            v[u][e] = {w.x, w.y};
        }
    }
    after_last_reads();
#pragma unroll
    for (int u = 0; u < AS; ++u) idft8(v[u]);
}

// The swap schedule of N = 1024 and 2048 (see SwapFft); TWLS: the twiddles
// of passes 1 and 2 from the workgroup's LDS table twl (twl_swap).
template <int N, int AS, typename Hook, bool TWLS = false>
__device__ __forceinline__ void fft_swap_small(cf (&v)[AS][8], float2* const (&pa)[AS], float2* const (&pb)[AS],
                                               const cf (&tw)[Geo<N>::NPASS][8], Hook&& after_last_reads,
                                               const float2* twl = nullptr) {
    using F = SwapFft<N>;
    const int t = threadIdx.x % F::TPA;
    static_assert(!TWLS || AS == 1, "LDS twiddles of the swap schedule: one A-scan per lane");
    // the twiddles of passes 1 and 2: registers, or the workgroup's table
    auto tw1 = [&](int e) -> cf {
        if constexpr (TWLS) {
            const float2 w = twl[(e - 1) * 64 + 8 * F::d2_of(t) + F::r2_of(t)];
            return {w.x, w.y};
        } else {
            return tw[1][e];
        }
    };
    auto tw2 = [&](int e) -> cf {
        if constexpr (TWLS) {
            const float2 w = twl[448 + (e - F::RQ) * 8 + F::r2_of(t)];
            return {w.x, w.y};
        } else {
            return tw[2][e];
        }
    };
    // pass 0 and exchange 1
    const int w1 = F::e1_write_base(t);
#pragma unroll
    for (int u = 0; u < AS; ++u) {
        idft8(v[u]);
#pragma unroll
        for (int e = 1; e < 8; ++e) v[u][e] = cmul(v[u][e], tw[0][e]);
#pragma unroll
        for (int e = 0; e < 8; ++e) pa[u][w1 + F::skew(e)] = make_float2(v[u][e].x, v[u][e].y);
    }
    __syncthreads();
    const int r1 = F::e1_read_base(t);
#pragma unroll
    for (int u = 0; u < AS; ++u) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float2 w = pa[u][r1 + F::c_d1(e)];
            v[u][e] = {w.x, w.y};
        }
    }
    // pass 1, exchange 2 in registers, pass 2
#pragma unroll
    for (int u = 0; u < AS; ++u) {
        idft8(v[u]);
#pragma unroll
        for (int e = 1; e < 8; ++e) v[u][e] = cmul(v[u][e], tw1(e));
#pragma unroll
        for (int e = 0; e < 4; ++e) swap_lane_bit5(v[u][e], v[u][e + 4]);
        if constexpr (F::S == 4) {
            swap_lane_bit4(v[u][0], v[u][2]);
            swap_lane_bit4(v[u][1], v[u][3]);
            swap_lane_bit4(v[u][4], v[u][6]);
            swap_lane_bit4(v[u][5], v[u][7]);
            idft4(v[u][0], v[u][2], v[u][4], v[u][6]);
            idft4(v[u][1], v[u][3], v[u][5], v[u][7]);
        } else {
            idft2(v[u][0], v[u][4]);
            idft2(v[u][1], v[u][5]);
            idft2(v[u][2], v[u][6]);
            idft2(v[u][3], v[u][7]);
        }
#pragma unroll
        for (int e = F::RQ; e < 8; ++e) v[u][e] = cmul(v[u][e], tw2(e));
    }
    // exchange 3 and pass 3
    const int w3 = F::e3_write_base(t);
#pragma unroll
    for (int u = 0; u < AS; ++u) {
#pragma unroll
        for (int e = 0; e < 8; ++e) pb[u][w3 + F::e3_write_off(e)] = make_float2(v[u][e].x, v[u][e].y);
    }
    __syncthreads();
    // raised priority while the reads are issued, as for the tap gather
    __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int u = 0; u < AS; ++u) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float2 w = pb[u][t + F::E3 * e];
            v[u][e] = {w.x, w.y};
        }
    }
    __builtin_amdgcn_s_setprio(0);
    after_last_reads();
#pragma unroll
    for (int u = 0; u < AS; ++u) idft8(v[u]);
}

// Unnormalised inverse FFT of the AS sequences whose elements t + e*N/8
// lane t holds in v[u][e], by the swap schedule; on return v[u][e] =
// X[t + e*N/8] as after stockham_inverse.  Exchange 1 goes through pa,
// exchange 3 through pb; ends without a barrier (other wavefronts may still
// read pb).  Every lane of each wavefront must be active.
template <int N, int AS, typename Hook = NoHook, bool TWLS = false>
__device__ __forceinline__ void fft_swap(cf (&v)[AS][8], float2* const (&pa)[AS], float2* const (&pb)[AS],
                                         const cf (&tw)[Geo<N>::NPASS][8], Hook&& after_last_reads = {},
                                         const float2* twl = nullptr) {
    if constexpr (N == 4096) {
        fft_swap4096<AS>(v, pa, pb, tw, after_last_reads);
    } else {
        fft_swap_small<N, AS, Hook, TWLS>(v, pa, pb, tw, static_cast<Hook&&>(after_last_reads), twl);
    }
}

// Compile-time modes of the chain kernel.
enum : int {
    M_DC = 1,      // rolling-average DC removal prologue
    M_STATS = 2,   // FPN determination: accumulate segment statistics, no output
    M_PAIR = 4,    // no dispersion: the lane's two A-scans as the real and
                   // imaginary part of one complex FFT (AS = 2)
    M_ROWS = 8,    // A-scans through the row map of ChainArgs (ragged
                   // batches), not rows of one frame
};

// Rolling-average DC removal of one A-scan (cpu_kernels.tpp:56-82), staged
// into xs (N floats in LDS): sample n becomes x[n] minus the mean of
// x[start .. end], start = max(n - (W-1), 0), end = min(n + W, N-1) for
// W >= 1, and start = 0, end = n for W = 0.  Lane t brings samples
// n = 8t .. 8t+7 in xv.  Window sums are differences of exact inclusive
// prefix sums Q[n] = x[0] + ... + x[n] (int for 8/16-bit samples, double
// for 32-bit ones), built in the lane's own 8 consecutive samples; `part`
// holds one S per wavefront.
//
// The banded path (BANDED, int sums, TPA >= 128, windows of 1 .. DC_BAND
// samples): lane t keeps its own samples.  The prefix sums go to Q in
// blocks of 8 at a stride of 9 ints (block b at 9b: the stores of a
// wavefront, and its loads of 8 consecutive sums from any offset, hit 64
// distinct banks), with DC_BAND/8 blocks of zeros (Q[-DC_BAND .. -1])
// before the A-scan's and as many blocks of the total (Q[N .. N+DC_BAND-1])
// after it, so that clipped windows need no clamping.  Lane t loads
// Q[8t+i+W] and Q[8t+i-W] (i = 0..7) with offsets fixed per W mod 8 and
// takes every sample as an interior one (the exact numerator x*2W - sum
// times the reciprocal of 2W); the 2W-1 samples whose windows clip are then
// redone one per lane (the exact numerator x*cnt - sum times the reciprocal
// of cnt with one correction step, within one ulp).
//
// The general path (W = 0, W > DC_BAND, double sums): samples to xs, the
// prefix sums to Q[0 .. N-1], then the means with lane t on samples
// t + i*TPA, so that every LDS access of that step is lane-consecutive and
// the clipped windows at both ends fall on one sample per lane (i = 0 and
// i = 7) as long as W <= TPA - 2.  For i = 1..6 the windows lie inside the
// A-scan (cnt = 2W): the exact integer numerator x*2W - sum is scaled by
// the reciprocal of 2W (within one ulp); the end samples take the reciprocal
// of cnt with one correction step (within one ulp).
//
// Q holds N entries of S, and dc_prefix_ints ints where the banded path can
// run.  The prefix half contains two workgroup barriers (one for
// TPA <= 64), the means half none; every earlier read of Q, part and xs
// must be ordered before the prefix half.
constexpr int DC_BAND = 64;
template <int TPA>
constexpr int dc_prefix_ints() { return 9 * (TPA + DC_BAND / 4); }

// The banded path's samples x - (hi[.] - lo[.]) / 2W for W mod 8 = B, to
// dst (the lane's 8 staging floats), in two halves of 4: the block pad sits
// after the first 8-B sums from hi and after the first B from lo (none for
// B = 0).  The numerator x*2W - sum is exact in float (below 2^24).
template <int B>
__device__ __forceinline__ void dc_band_means(const int* hi, const int* lo, const float (&xv)[8], float w2, float inv,
                                              float4* dst) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float r[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = 4 * h + k;
            const float sum = (float)(hi[i + ((B != 0 && i >= 8 - B) ? 1 : 0)] - lo[i + ((B != 0 && i >= B) ? 1 : 0)]);
            r[k] = fmaf(xv[i], w2, -sum) * inv;
        }
        dst[h] = make_float4(r[0], r[1], r[2], r[3]);
    }
}

// the window the stages use: wider windows cover the whole A-scan alike;
// the host rejects W < 0
template <int N>
__device__ __forceinline__ int dc_window(int window) { return window > N ? N : (window < 0 ? 0 : window); }

template <int N, int TPA, typename S>
constexpr bool dc_banded() { return sizeof(S) == 4 && TPA >= 128 && dc_prefix_ints<TPA>() <= 2 * N; }

// First half: the prefix sums (the samples too on the general path), and
// the barrier behind them.
template <int N, int TPA, typename S, bool BANDED>
__device__ __forceinline__ void dc_prefix_stage(const float (&xv)[8], int t, int window, S* Q, S* part, float* xs) {
    const int W = dc_window<N>(window);
    constexpr bool BAND = BANDED && dc_banded<N, TPA, S>();
    // without the banded path the samples go out first (fewer registers
    // held through the scan)
    if constexpr (!BAND) {
        float4* dst = reinterpret_cast<float4*>(xs + 8 * t);
        dst[0] = make_float4(xv[0], xv[1], xv[2], xv[3]);
        dst[1] = make_float4(xv[4], xv[5], xv[6], xv[7]);
    }
    S loc[8];
    S s = S(0);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        s += (S)xv[i];  // exact: integers below 2^16, or any float/uint32 in double
        loc[i] = s;
    }
    S incl = group_inclusive_scan<(TPA < 64 ? TPA : 64)>(s);
    if constexpr (TPA > 64) {
        constexpr int WPA = TPA / 64;  // wavefronts per A-scan
        const int wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 63) part[wave] = incl;
        __syncthreads();
        const int first = (threadIdx.x / TPA) * WPA;
        S off = S(0);
#pragma unroll
        for (int w = 0; w < WPA - 1; ++w) {
            const S pw = part[first + w];
            off += (first + w < wave) ? pw : S(0);
        }
        incl += off;
    }
    const S excl = incl - s;

    constexpr int GB = DC_BAND / 8;  // guard blocks at each end
    if constexpr (BAND) {
        if (W >= 1 && W <= DC_BAND) {
            int* const qp = reinterpret_cast<int*>(Q);
            // the A-scan's total: its last lane's inclusive sum
            int total;
            if constexpr (TPA >= 64)
                total = __builtin_amdgcn_readlane(incl, 63);  // read by the last wavefront only
            else
                total = __shfl(incl, ((int)threadIdx.x & 63) | (TPA - 1), 64);
#pragma unroll
            for (int i = 0; i < 8; ++i) qp[9 * (t + GB) + i] = excl + loc[i];
            if (t < GB) {
#pragma unroll
                for (int i = 0; i < 8; ++i) qp[9 * t + i] = 0;
            }
            if (t >= TPA - GB) {
#pragma unroll
                for (int i = 0; i < 8; ++i) qp[9 * (t + 2 * GB) + i] = total;
            }
            ascan_sync<TPA>();
            return;
        }
    }

    if constexpr (BAND) {
        float4* dst = reinterpret_cast<float4*>(xs + 8 * t);
        dst[0] = make_float4(xv[0], xv[1], xv[2], xv[3]);
        dst[1] = make_float4(xv[4], xv[5], xv[6], xv[7]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) Q[8 * t + i] = excl + loc[i];
    ascan_sync<TPA>();
}

// Second half: the window means into xs.  The same xv, window and Q as
// the first half; any code may run between the two that leaves Q, part and
// xs alone.
template <int N, int TPA, typename S, bool BANDED>
__device__ __forceinline__ void dc_means_stage(const float (&xv)[8], int t, int window, S* Q, float* xs) {
    const int W = dc_window<N>(window);
    constexpr int GB = DC_BAND / 8;
    if constexpr (BANDED && dc_banded<N, TPA, S>()) {
        if (W >= 1 && W <= DC_BAND) {
            int* const qp = reinterpret_cast<int*>(Q);
            // Q[8t + W] and Q[8t - W] in the padded layout
            const int* hi = qp + 9 * (t + GB) + W + (W >> 3);
            const int* lo = qp + 9 * t + (DC_BAND - W) + ((DC_BAND - W) >> 3);
            // The 2W-1 clipped samples are redone one per lane: n = t
            // (n <= W-2, start clipped) in the first wavefront and
            // n = N - TPA + t (n >= N-W, end clipped) in the last, the same
            // wavefronts as those samples' owners, whose stores below these
            // follow.  x[n] = Q[n] - Q[n-1]; the guard blocks stand in for
            // Q[-1] = 0 and Q[N-1 ..] = total.  Computed first, so that its
            // loads and division overlap the band's loads.
            const int n = t < 64 ? t : N - TPA + t;
            const bool fix = (t < 64 && n <= W - 2) || (t >= TPA - 64 && n >= N - W);
            float fixed = 0.f;
            if (fix) {
                auto qat = [&](int m) {  // Q[m], -DC_BAND <= m < N + DC_BAND
                    const int j = m + DC_BAND;
                    return qp[9 * (j >> 3) + (j & 7)];
                };
                const int x = qat(n) - qat(n - 1);
                const int sum = qat(n + W) - qat(n - W);
                const int start = n >= W - 1 ? n - (W - 1) : 0;
                const int end = n + W < N - 1 ? n + W : N - 1;
                const float c = (float)(end - start + 1);
                // x - sum/cnt from the exact numerator: reciprocal and one
                // correction step (within one ulp)
                const float num = (float)(__mul24(x, end - start + 1) - sum);
                const float rc = __builtin_amdgcn_rcpf(c);
                const float q0 = num * rc;
                fixed = fmaf(fmaf(-q0, c, num), rc, q0);
            }
            // every sample as if its window were inside the A-scan
            const float w2 = (float)(2 * W);
            const float inv = 1.0f / w2;
            float4* dst = reinterpret_cast<float4*>(xs + 8 * t);
            switch (W & 7) {
                case 0: dc_band_means<0>(hi, lo, xv, w2, inv, dst); break;
                case 1: dc_band_means<1>(hi, lo, xv, w2, inv, dst); break;
                case 2: dc_band_means<2>(hi, lo, xv, w2, inv, dst); break;
                case 3: dc_band_means<3>(hi, lo, xv, w2, inv, dst); break;
                case 4: dc_band_means<4>(hi, lo, xv, w2, inv, dst); break;
                case 5: dc_band_means<5>(hi, lo, xv, w2, inv, dst); break;
                case 6: dc_band_means<6>(hi, lo, xv, w2, inv, dst); break;
                default: dc_band_means<7>(hi, lo, xv, w2, inv, dst); break;
            }
            if (fix) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                xs[n] = fixed;
            }
            return;
        }
    }

    // samples t + i*TPA.  With 1 <= W <= TPA - 2, i = 1..6 are always
    // inside the A-scan, i = 0 is inside iff t >= W, i = 7 iff
    // t <= TPA - 2 - W.
    const bool banded = W >= 1 && W <= TPA - 2;
    const S* hi = Q + t + W;
    const S* lo = Q + t - W;  // read only where t + i*TPA >= W
    if (banded) {
#pragma unroll
        for (int i = 1; i < 7; ++i) {
            const int n = t + i * TPA;
            const S sum = hi[i * TPA] - lo[i * TPA];
            if constexpr (sizeof(S) == 4)
                xs[n] = (float)(__mul24((S)xs[n], 2 * W) - sum) * (1.0f / (float)(2 * W));
            else
                xs[n] = (float)((S)xs[n] - sum * (1.0 / (double)(2 * W)));
        }
    }
    // the end samples (all samples without the band): one copy of the code,
    // not unrolled, so that it adds few registers
#pragma unroll 1
    for (int k = 0; k < (banded ? 2 : 8); ++k) {
        const int i = banded ? 7 * k : k;
        const int n = t + i * TPA;
        const bool inside = banded && (i > 0 || t >= W) && (i < 7 || t <= TPA - 2 - W);
        int start, end;
        if (inside) {
            start = n - W + 1;
            end = n + W;
        } else if (W >= 1) {
            start = (n >= W - 1) ? n - (W - 1) : 0;
            end = (n + W < N - 1) ? n + W : N - 1;
        } else {
            start = 0;
            end = n;
        }
        const S sum = Q[end] - (start > 0 ? Q[start - 1] : S(0));
        const int cnt = end - start + 1;
        const S x = (S)xs[n];
        // x - sum/cnt from the exact numerator: the reciprocal of cnt and one
        // correction step (within one ulp), the same code for clipped and
        // inside windows
        if constexpr (sizeof(S) == 4) {
            const float c = (float)cnt;
            const float num = (float)(__mul24(x, cnt) - sum);
            const float rc = __builtin_amdgcn_rcpf(c);
            const float q0 = num * rc;
            xs[n] = fmaf(fmaf(-q0, c, num), rc, q0);
        } else {
            xs[n] = (float)(x - sum / (double)cnt);
        }
    }
}

template <int N, int TPA, typename S, bool BANDED = false>
__device__ __forceinline__ void dc_remove_stage(const float (&xv)[8], int t, int window, S* Q, S* part, float* xs) {
    dc_prefix_stage<N, TPA, S, BANDED>(xv, t, window, Q, part, xs);
    dc_means_stage<N, TPA, S, BANDED>(xv, t, window, Q, xs);
}

#ifndef OPE_WAVES_PER_EU
#define OPE_WAVES_PER_EU 1
#endif

// LDS arrays of one chain body with a workgroup of WG lanes, in elements:
// two padded complex buffers per A-scan slot, the float staging buffer when
// it cannot alias buffer B, and the per-wavefront partials of the DC scan
// and of the auto min/max.  The kernels declare them as separate
// __shared__ arrays: offsets into them then fold into the LDS instructions'
// immediate fields, which a single byte array carved at run time loses.
template <int N, int MODE, int IW, int AS, int WG, int SCHED = -1>
struct ChainLds {
    using G = Geo<N>;
    static_assert(WG % G::TPA == 0 && WG % 64 == 0, "workgroup must hold whole A-scans");
    static constexpr int SLOTS = (WG / G::TPA) * AS;        // A-scans per workgroup iteration
    static constexpr int NW = WG / 64;                      // wavefronts per workgroup
    static constexpr bool PAIR = (MODE & M_PAIR) != 0;
    static_assert(!PAIR || (AS == 2 && (MODE & (M_DC | M_STATS)) == 0), "pairs: two A-scans per lane, no DC/stats");
    static constexpr int FFTS = PAIR ? SLOTS / 2 : SLOTS;               // complex FFTs per iteration
    // x staging can share buffer B; not with pairs, whose buffer B also
    // carries the spectrum exchange that splits the pair after the FFT.  The
    // swap schedule's last reads are of buffer B: with one A-scan per lane
    // the staging gets its own array, with two (where three arrays would
    // leave room for one workgroup per CU) it shares B behind a barrier at
    // the top of each iteration (SWAP_TOP_BARRIER)
    static constexpr bool SWAP = swap_sched<N, SCHED>();
    static constexpr bool SWAP_TOP_BARRIER = SWAP && AS == 2 && !PAIR;
    static constexpr bool X_ALIAS_B = ((G::NPASS % 2) == 0 && !PAIR && !SWAP) || SWAP_TOP_BARRIER;
    using DcSum = typename std::conditional<(IW <= 2), int, double>::type;
    static constexpr int BUF = FFTS * G::PADN;                           // float2 per buffer
    static constexpr int SCAN = (MODE & M_DC) ? AS * NW : 1;            // DcSum
    static constexpr int MM = (MODE & M_STATS) ? 1 : 2 * AS * NW;       // floats
    static constexpr int X = X_ALIAS_B ? 1 : SLOTS * N;                 // floats
    // DC removal staged inside the FFT (one A-scan per lane, a staging array
    // of its own): the prefix sums then go to buffer A, which every
    // wavefront has finished reading once the last pass's reads (of buffer
    // B: the swap schedule's, and the Stockham schedule's with an odd number
    // of passes, the only one with a staging array of its own) are issued
    static constexpr bool DC_PIPE = (MODE & M_DC) != 0 && AS == 1 && !X_ALIAS_B;
    // twiddles of FFT passes 1 and 2 in LDS (the kernels of T taps that
    // then fit four waves per SIMD; the DC prologue needs more registers)
    template <int T>
    static constexpr bool TWL = twl_fft<N, T, AS, SCHED>() && !PAIR && (MODE & (M_DC | M_STATS)) == 0;
    // the swap schedule's passes 1 and 2 from LDS (twl_swap)
    template <int T>
    static constexpr bool TWLS = (twl_swap<N, T, AS, SCHED>() || (T == 4 && twl_swap<N, 8, AS, SCHED>() &&
                                                                    (MODE & M_DC) && IW <= 2)) &&
                                 !PAIR && (MODE & M_STATS) == 0;
};

// The chain over A-scans a_begin + k*a_step ... (< a_end) of A, each
// workgroup iteration taking SLOTS consecutive A-scans.
//
// T = taps per output sample: 1 (no resampling), 2 (linear), 4 (cubic),
// 8 (Lanczos-4), 16 (Lanczos-8).  MODE: M_* bits.  IW: input sample width in bytes (1, 2,
// 4); DC removal sums 1- and 2-byte samples in int, 4-byte ones in double.
// AS: A-scans each lane carries through the chain at once (instruction-level
// parallelism: the LDS traffic of one overlaps the arithmetic of the other,
// and each barrier covers AS A-scans; the per-position tables are shared).
// WG: workgroup size, a multiple of N/8.  All lanes of the workgroup call
// this with the same range.
template <int N, int T, int MODE, int IW, int AS, int WG, int SCHED = -1>
__device__ __forceinline__ void chain_body(const ChainArgs A, float2* bufA, float2* bufB, float* xsep,
                                           typename ChainLds<N, MODE, IW, AS, WG, SCHED>::DcSum* scan_part,
                                           float* mm_part, float2* twl, long long a_begin, long long a_end,
                                           long long a_step) {
    using G = Geo<N>;
    using L = ChainLds<N, MODE, IW, AS, WG, SCHED>;
    constexpr int TPA = G::TPA;
    constexpr int NP = G::NPASS;
    constexpr int NW = L::NW;
    constexpr bool DC = (MODE & M_DC) != 0;
    constexpr bool STATS = (MODE & M_STATS) != 0;
    constexpr bool X_ALIAS_B = L::X_ALIAS_B;
    constexpr bool PAIR = L::PAIR;
    constexpr bool SWAP = swap_sched<N, SCHED>();
    using DcSum = typename L::DcSum;

    // one A-scan per workgroup iteration: the A-scan index is uniform
    const int slot = (WG == TPA) ? 0 : (int)threadIdx.x / TPA;
    const int t = threadIdx.x - slot * TPA;
    // Row map on a dynamic schedule: the workgroup's first three chunks are
    // taken before anything else (the ticket passes from lane 0 through the
    // min/max partials' first slot), and a workgroup that finds none (the
    // schedule's other launch took them all) counts itself out at once
    // instead of reading per-lane tables it would not use.
    unsigned first_ticket = 0;
    if constexpr ((MODE & M_ROWS) != 0) {
        if (A.tickets != nullptr) {
            unsigned* const sh = reinterpret_cast<unsigned*>(mm_part);
            if (threadIdx.x == 0) *sh = atomicAdd(A.tickets, 3u);
            __syncthreads();
            first_ticket = *sh;
            __syncthreads();
            if ((long long)first_ticket * kRowChunk * ChainLds<N, MODE, IW, AS, WG, SCHED>::SLOTS >= a_end) {
                if (threadIdx.x == 0 && atomicAdd(A.tickets + 1, 1u) == A.ticket_wgs - 1) {
                    atomicExch(A.tickets, 0u);
                    atomicExch(A.tickets + 1, 0u);
                }
                return;
            }
        }
    }
    float2* pa[AS];
    float2* pb[AS];
    float* xs[AS];
#pragma unroll
    for (int u = 0; u < AS; ++u) {
        // pairs share one FFT (and its buffers) per slot
        const int f = PAIR ? slot : slot * AS + u;
        pa[u] = bufA + f * G::PADN;
        pb[u] = bufB + f * G::PADN;
        xs[u] = X_ALIAS_B ? reinterpret_cast<float*>(pb[u]) : (xsep + (slot * AS + u) * N);
    }

    const int flags = A.flags;
    const bool ieee = A.in_type == IN_F32;
    const bool shift = (flags & F_BITSHIFT) != 0;
    const float m0 = (flags & F_OFFSET) ? *gmem(A.offset) : 0.f;

    // ---- per-lane tables, read once per call
    // tap windows as LDS pointers into the lane's first staging buffer (the
    // other A-scans of the lane sit at constant offsets from it)
    // (with the LDS twiddles the window starts are kept as 16-bit halves,
    // two per register: 4 VGPRs fewer for one extraction per output)
    // (the kernels with LDS twiddles, where registers are scarce)
    constexpr bool LEAN = L::template TWL<T> || L::template TWLS<T>;
    constexpr bool PACK_TB = LEAN;
    const float* tb[8];
    unsigned tbp[4];
    float tw8[8][T];
    cf ph[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int j = t + e * TPA;
        const int b0 = gmem(A.tap_base)[j];
        if constexpr (PACK_TB) {
            if (e & 1)
                tbp[e >> 1] |= (unsigned)b0 << 16;
            else
                tbp[e >> 1] = (unsigned)b0;
        } else {
            tb[e] = xs[0] + b0;
        }
#pragma unroll
        for (int i = 0; i < T; ++i) tw8[e][i] = gmem(A.tap_w)[j * T + i];
        if (!PAIR && (flags & F_DISP)) {
            float2 p = gload2(A.phasor, j);
            ph[e] = {p.x, p.y};
        } else {
            ph[e] = {1.f, 0.f};
        }
    }
    cf tw[NP][8];
    constexpr bool TWL = L::template TWL<T>;
    constexpr bool TWLS = L::template TWLS<T>;
    if constexpr (SWAP) {
        // (with the LDS table the registers of passes 1 and 2 go unread)
        if constexpr (TWLS)
            for (int i = threadIdx.x; i < kTwlEntries; i += WG) twl[i] = gload2(A.twiddle, twl_swap_exponent<N>(i));
        load_twiddles_swap<N>(A.twiddle, t, tw);
    } else if constexpr (TWL) {
        // passes 1 and 2 from the workgroup's LDS table, filled here; the
        // first FFT reads it after the barrier at the top of the loop
        for (int i = threadIdx.x; i < kTwlEntries; i += WG) twl[i] = gload2(A.twiddle, twl_exponent(i));
        load_twiddles<N, SCHED, 3>(A.twiddle, t, tw);
    } else {
        load_twiddles<N, SCHED>(A.twiddle, t, tw);
    }
    // the background constant: per lane for the whole call, or (with the
    // LDS twiddles, where registers are scarce) read in the epilogue
    auto bg_const = [&](int k) { return __fadd_rn(__fmul_rn(A.bg_w, gmem(A.bg)[k]), A.bg_off); };
    cf fp[4];
    float bgv[4];
    if constexpr (!STATS) {
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const int k = t + o * TPA;
            if (flags & (F_FPN | F_OFFSET)) {
                // pairs split into twice the spectrum (exact scaling)
                float2 f = gload2(A.fpn, k);
                fp[o] = PAIR ? cf{2.f * f.x, 2.f * f.y} : cf{f.x, f.y};
            } else {
                fp[o] = {0.f, 0.f};
            }
            bgv[o] = (!LEAN && (flags & F_BG)) ? bg_const(k) : 0.f;
        }
    }

    // LDS base: every pass reads positions t + e*TPA (base + constant)

    // statistics of X = X' + m0 * D, the spectrum without the offset
    double2 dsh[4];
    if constexpr (STATS) {
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            dsh[o] = make_double2(0.0, 0.0);
            if (flags & F_OFFSET) {
                const OPE_GLOBAL double* d = gmem(reinterpret_cast<const double*>(A.offset_tab)) + 2 * (t + o * TPA);
                dsh[o] = make_double2((double)m0 * d[0], (double)m0 * d[1]);
            }
        }
    }

    double sx[AS][4], sy[AS][4], sxx[AS][4];
    int cur_seg[AS];
#pragma unroll
    for (int u = 0; u < AS; ++u) {
        cur_seg[u] = -1;
#pragma unroll
        for (int o = 0; o < 4; ++o) sx[u][o] = sy[u][o] = sxx[u][o] = 0.0;
    }

    // Row map (ragged batches): nin[u] is the input row of the lane's next
    // A-scan to be loaded, read one group ahead of the sample loads that
    // use it.  Its loads are unconditional (indices past the range clamp
    // to its last A-scan, which is then read again and not used): with the
    // same number of loads in flight on every path, the epilogue waits only
    // for its own row entry, not for the prefetches issued after it.
    constexpr bool rowmap = (MODE & M_ROWS) != 0;
    static_assert(!rowmap || !STATS, "row maps are for the throughput mode");
    auto clamp_row = [&](long long a) { return a < a_end ? a : a_end - 1; };
    // The group of A-scans after group g (workgroup-uniform): g + a_step, or
    // under a row map's dynamic schedule the next group of g's chunk, else
    // the first group of the workgroup's next chunk.  cur, nxt and nn are
    // the workgroup's current chunk and the two it holds tickets for;
    // callers look at most three groups ahead of the current one.
    const bool dyn = rowmap && A.tickets != nullptr;
    constexpr long long CG = (long long)kRowChunk * L::SLOTS;
    long long cur = 0, nxt = 0, nn = 0;
    // a ticket passes from lane 0 to the workgroup through the min/max
    // partials' first slot (free outside the epilogue's reduction, which
    // ends with a barrier)
    unsigned* const tk_sh = reinterpret_cast<unsigned*>(mm_part);
    if constexpr (rowmap) {
        if (dyn) {
            cur = first_ticket;
            nxt = cur + 1;
            nn = cur + 2;
            a_begin = cur * CG;
            a_step = L::SLOTS;
        }
    }
    auto adv = [&](long long g) -> long long {
        if (!dyn) return g + a_step;
        const long long ch = g / CG;
        if (g + a_step < (ch + 1) * CG) return g + a_step;  // a_step = SLOTS here
        return (ch == cur ? nxt : nn) * CG;
    };
    const void* nin[AS];
    uint32_t raw[AS][8];
    // loads group g's samples into raw (and reads the input rows of group
    // adv(g) under a row map)
    auto load_next = [&](int u, long long g) {
        const long long an = g + slot * AS + u;
        if constexpr (rowmap) {
            load_raw<IW>(nin[u], N, 0, t, raw[u]);
            nin[u] = gmem(A.rows)[clamp_row(adv(g) + slot * AS + u)].in;
        } else {
            // pairs: a slot whose second A-scan lies past the end pairs its
            // first with itself (as a row map pads a request of odd count)
            const long long src = (PAIR && u == 1 && an == a_end) ? an - 1 : an;
            if (src < a_end) load_raw<IW>(A.in, N, src, t, raw[u]);
        }
    };
#pragma unroll
    for (int u = 0; u < AS; ++u) {
#pragma unroll
        for (int i = 0; i < 8; ++i) raw[u][i] = 0;
        nin[u] = nullptr;
        if constexpr (rowmap) nin[u] = gmem(A.rows)[clamp_row(a_begin + slot * AS + u)].in;
        load_next(u, a_begin);
    }

    // Staging of A-scan group a: convert the raw words (loaded one group
    // ahead) into the staging array, then prefetch group a + a_step.
    // Pipelined (PIPE): group a0 + a_step is staged inside the FFT of group
    // a0, once its last LDS reads are issued: the staging array is free by
    // then (every wavefront has passed the barrier after the tap gather),
    // and the stores overlap the last pass instead of preceding a barrier.
    // With DC removal the whole prologue (prefix sums, window means, their
    // barriers) moves with it where every array it writes is free by then:
    // a staging array of its own, and the prefix sums in buffer A (read for
    // the last time before the last pass's barrier).  Not where the staging
    // shares buffer B with the last reads.  (Pairs have a staging array of their own; their split exchange
    // after the FFT uses the FFT buffers.)
    // The banded DC path (windows of 1 .. DC_BAND samples) needs ~29 more
    // registers at the staging point: the kernels with up to 2 taps take it
    // (linear spills 8 bytes per lane at N = 2048 and still gains 9 %), and
    // cubic where the swap schedule's LDS twiddles free the registers
    // (N = 1024 and 2048, 8/16-bit samples: 2048-point cubic + DC 75.6 ->
    // 73.5 us); with register twiddles cubic and Lanczos-8 would spill
    // 50-100 bytes and run slower.  Staged
    // inside the FFT, it splits: the prefix sums go in there, the window
    // means after the output stores, where the FFT's registers are free (the
    // raw words of the staged group stay in registers in between).  The
    // general path stages whole inside the FFT.
    constexpr bool BANDED = T <= 2 || (T == 4 && L::template TWLS<T>);
    constexpr bool SPLIT = DC && BANDED && dc_banded<N, TPA, DcSum>();
    constexpr bool PIPE = DC ? L::DC_PIPE : !L::SWAP_TOP_BARRIER;
    auto stage_prefix = [&]() {
#pragma unroll
        for (int u = 0; u < AS; ++u) {
            float x8[8];
            convert_raw<IW>(ieee, shift, m0, raw[u], x8);
            dc_prefix_stage<N, TPA, DcSum, BANDED>(x8, t, A.dc_window, reinterpret_cast<DcSum*>(pa[u]),
                                                   scan_part + u * NW, xs[u]);
        }
    };
    auto stage_rest = [&](long long a) {
#pragma unroll
        for (int u = 0; u < AS; ++u) {
            float x8[8];
            convert_raw<IW>(ieee, shift, m0, raw[u], x8);
            if constexpr (SPLIT) {
                dc_means_stage<N, TPA, DcSum, BANDED>(x8, t, A.dc_window, reinterpret_cast<DcSum*>(pa[u]), xs[u]);
            } else if constexpr (DC) {
                dc_remove_stage<N, TPA, DcSum, BANDED>(x8, t, A.dc_window, reinterpret_cast<DcSum*>(pa[u]),
                                                       scan_part + u * NW, xs[u]);
            } else {
                float4* dst = reinterpret_cast<float4*>(xs[u] + 8 * t);
                dst[0] = make_float4(x8[0], x8[1], x8[2], x8[3]);
                dst[1] = make_float4(x8[4], x8[5], x8[6], x8[7]);
            }
            load_next(u, adv(a));
        }
    };
    if constexpr (PIPE) {
        if constexpr (SPLIT) stage_prefix();
        stage_rest(a_begin);
    }

    unsigned tk = 0;        // lane 0: the ticket requested on entering the current chunk
    int since_ticket = -1;  // iterations since that request (-1: none pending)
    // the next group; on entering a new chunk the workgroup moves to the
    // chunks it holds tickets for and lane 0 requests one more
    auto adv_and_step = [&](long long a0) -> long long {
        const long long na = adv(a0);
        if constexpr (rowmap) {
            if (dyn && na / CG != a0 / CG) {
                cur = nxt;
                nxt = nn;
                if (threadIdx.x == 0) tk = atomicAdd(A.tickets, 1u);
                since_ticket = 0;
            }
        }
        return na;
    };
    for (long long a0 = a_begin; a0 < a_end; a0 = adv_and_step(a0)) {
        if constexpr (rowmap) {
            // the ticket requested on entering this chunk: all lanes read it
            // one iteration later (barriers around the slot: the waves of an
            // A-scan of up to 64 lanes meet at no other workgroup barrier)
            if (dyn && since_ticket >= 0 && ++since_ticket == 2) {
                if (threadIdx.x == 0) *tk_sh = tk;
                __syncthreads();
                nn = *tk_sh;
                __syncthreads();
                since_ticket = -1;
            }
        }
        // the staging shares buffer B, which the previous A-scan's last FFT
        // pass may still be reading in other wavefronts
        if constexpr (L::SWAP_TOP_BARRIER) __syncthreads();
        // row map: this group's output rows and requests, read here so that
        // the epilogue's stores do not wait on them
        RowRef rr[AS];
#pragma unroll
        for (int u = 0; u < AS; ++u) {
            rr[u] = RowRef{nullptr, nullptr, 0};
            if constexpr (rowmap) {
                const OPE_GLOBAL RowRef* e = gmem(A.rows) + clamp_row(a0 + slot * AS + u);
                rr[u] = RowRef{nullptr, e->out, e->req};
            }
        }
        // ---- stage converted samples in LDS
        float xv[AS][8];
        if constexpr (!PIPE) {
#pragma unroll
            for (int u = 0; u < AS; ++u) {
                convert_raw<IW>(ieee, shift, m0, raw[u], xv[u]);
                if constexpr (!DC) {
                    float4* dst = reinterpret_cast<float4*>(xs[u] + 8 * t);
                    dst[0] = make_float4(xv[u][0], xv[u][1], xv[u][2], xv[u][3]);
                    dst[1] = make_float4(xv[u][4], xv[u][5], xv[u][6], xv[u][7]);
                }
                // prefetch the next A-scan while this one is processed
                load_next(u, adv(a0));
            }
        }

        if constexpr (DC && !PIPE) {
            // Prefix sums aliasing buffer A: the previous A-scan's last pass
            // may still be reading A.  With TPA > 64 the first barrier inside
            // dc_remove_stage already orders those reads before the writes.
            if constexpr (TPA <= 64) ascan_sync<TPA>();
#pragma unroll
            for (int u = 0; u < AS; ++u) {
                dc_remove_stage<N, TPA, DcSum, BANDED>(xv[u], t, A.dc_window, reinterpret_cast<DcSum*>(pa[u]),
                                                       scan_part + u * NW, xs[u]);
            }
        }
        ascan_sync<TPA>();
        // Raised wave priority until the tap gather is done: the waves that
        // leave the barrier get their LDS reads out ahead of the other
        // workgroups' arithmetic (2048-point cubic frames ~2 % faster,
        // Lanczos-8 ~9 %; the same after the first FFT exchange measured
        // slower, and N = 4096, one workgroup per CU, 1-2 % slower)
        constexpr bool PRIO = N < 4096;
        if constexpr (PRIO) __builtin_amdgcn_s_setprio(2);

        // ---- resample (banded taps, window folded in) and dispersion
        cf v[AS][8];
#pragma unroll
        for (int u = 0; u < AS; ++u) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float* w0 = PACK_TB ? xs[0] + ((tbp[e >> 1] >> (16 * (e & 1))) & 0xffffu) : tb[e];
                const float* src = w0 + (xs[u] - xs[0]);
                float acc = 0.f;
#pragma unroll
                for (int i = 0; i < T; ++i) acc = fmaf(tw8[e][i], src[i], acc);
                v[u][e] = {acc * ph[e].x, acc * ph[e].y};
            }
        }
        if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);

        // ---- inverse FFT: NP Stockham passes
        // (a row map stages past the range too, the clamped last A-scan
        // again, so that its loads are unconditional)
        auto next = [&]() {
            if constexpr (PIPE) {
                if (rowmap || adv(a0) < a_end) {
                    if constexpr (SPLIT)
                        stage_prefix();
                    else
                        stage_rest(adv(a0));
                }
            }
        };
        if constexpr (PAIR) {
            // z = y0 + i y1: Z = X0 + i X1 with X0, X1 the spectra of the two
            // real sequences, so X0[k] = (Z[k] + conj Z[N-k]) / 2 and
            // X1[k] = (Z[k] - conj Z[N-k]) / 2i.
            cf z[1][8];
#pragma unroll
            for (int e = 0; e < 8; ++e) z[0][e] = {v[0][e].x, v[1][e].x};
            float2* const za[1] = {pa[0]};
            float2* const zb[1] = {pb[0]};
            if constexpr (SWAP)
                fft_swap<N, 1>(z, za, zb, tw, next);
            else
                stockham_inverse<N, 1, SCHED>(z, za, zb, tw, next);
            // Z[t + e*TPA] to the buffer the last pass did not read (its
            // readers passed the previous pass's barrier), then Z[N - k]
            float2* zx = (SWAP || (NP - 1) % 2 == 0) ? pa[0] : pb[0];
            const int base_w = pad_slots(t);
#pragma unroll
            for (int e = 0; e < 8; ++e) zx[base_w + pad_slots(e * TPA)] = make_float2(z[0][e].x, z[0][e].y);
            ascan_sync<TPA>();
            // N - (t + e*TPA) = (TPA - t) + (7 - e)*TPA for t > 0; lane 0's
            // partners Z[(8 - e)*TPA mod N] are its own elements
            const int base_r = pad_slots(t > 0 ? TPA - t : 1);
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                const float2 q = zx[base_r + pad_slots((7 - o) * TPA)];
                const cf w = z[0][(8 - o) & 7];
                const cf p = t > 0 ? cf{q.x, q.y} : w;
                v[0][o] = {z[0][o].x + p.x, z[0][o].y - p.y};  // 2 X0[k]
                v[1][o] = {z[0][o].y + p.y, p.x - z[0][o].x};  // 2 X1[k]
            }
        } else {
            if constexpr (SWAP)
                fft_swap<N, AS, decltype(next)&, TWLS>(v, pa, pb, tw, next, twl);
            else
                stockham_inverse<N, AS, SCHED, decltype(next)&, TWL>(v, pa, pb, tw, next, twl);
        }
        // after the last pass v[u][e] holds X[t + e*TPA] (twice X for
        // pairs); e < 4 are the kept bins

#pragma unroll
        for (int u = 0; u < AS; ++u) {
            const long long a = a0 + slot * AS + u;
            const bool active = a < a_end;
            if constexpr (STATS) {
                if (active) {
                    const long long row = A.row0 + a;
                    const int seg = (int)(row / A.seg_w);
                    if (seg != cur_seg[u]) {
                        if (cur_seg[u] >= 0) {
#pragma unroll
                            for (int o = 0; o < 4; ++o) {
                                double* d = A.acc + ((size_t)cur_seg[u] * G::HALF + t + o * TPA) * 3;
                                unsafeAtomicAdd(d, sx[u][o]);
                                unsafeAtomicAdd(d + 1, sy[u][o]);
                                unsafeAtomicAdd(d + 2, sxx[u][o]);
                                sx[u][o] = sy[u][o] = sxx[u][o] = 0.0;
                            }
                        }
                        cur_seg[u] = seg;
                    }
#pragma unroll
                    for (int o = 0; o < 4; ++o) {
                        const double re = (double)v[u][o].x + dsh[o].x, im = (double)v[u][o].y + dsh[o].y;
                        sx[u][o] += re;
                        sy[u][o] += im;
                        sxx[u][o] += re * re + im * im;
                    }
                }
            } else {
                // ---- epilogue: FPN, magnitude, grayscale map, background
                float m2[4];
#pragma unroll
                for (int o = 0; o < 4; ++o) {
                    const cf z = csub(v[u][o], fp[o]);
                    m2[o] = fmaf(z.x, z.x, z.y * z.y);
                }
                float res[4];
                // pairs: m2 = 4 |X|^2, i.e. log2 one higher by 2, sqrt twice
                const float log_h = PAIR ? A.log_h + 2.0f : A.log_h;
                const float ep_a = (PAIR && !(flags & F_LOG)) ? 0.5f * A.ep_a : A.ep_a;
                const float ep_b = (PAIR && (flags & F_LOG)) ? fmaf(-2.0f, A.ep_a, A.ep_b) : A.ep_b;
                if (flags & F_LOG) {
                    if (flags & F_AUTOMM) {
                        float val[4];
                        float mn = INFINITY, mx = -INFINITY;
#pragma unroll
                        for (int o = 0; o < 4; ++o) {
                            val[o] = A.log_c * (__log2f(m2[o]) - log_h);
                            mn = fminf(mn, val[o]);
                            mx = fmaxf(mx, val[o]);
                        }
                        ascan_minmax<TPA>(mn, mx, mm_part + 2 * u * NW, mm_part + (2 * u + 1) * NW);
                        float range = mx - mn;
                        if (range < 1e-6f) range = 1.0f;
#pragma unroll
                        for (int o = 0; o < 4; ++o) res[o] = A.coeff * ((val[o] - mn) / range + A.addend);
                    } else {
#pragma unroll
                        for (int o = 0; o < 4; ++o) res[o] = fmaf(ep_a, __log2f(m2[o]), ep_b);
                    }
                } else {
#pragma unroll
                    for (int o = 0; o < 4; ++o) res[o] = fmaf(ep_a, __fsqrt_rn(m2[o]), ep_b);
                }
                if (active) {
                    OPE_GLOBAL float* dst;
                    if (rowmap) {
                        dst = gmem(rr[u].out) + t;
                    } else {
                        long long row = a;
                        if (flags & F_FLIP) {
                            // cuda_kernels.cu:553-571: B-scans 0, 2, 4, ... reversed
                            const long long b = a / A.ascans_per_bscan;
                            if ((b & 1) == 0 && b + 1 < A.flip_bscans)
                                row = b * A.ascans_per_bscan + (A.ascans_per_bscan - 1 - (a - b * A.ascans_per_bscan));
                        }
                        dst = gmem(A.out) + row * G::HALF + t;
                    }
#pragma unroll
                    for (int o = 0; o < 4; ++o) {
                        // bgv is 0 without background subtraction: x - 0 == x exactly
                        float b = bgv[o];
                        if constexpr (LEAN) {
                            if (flags & F_BG) b = bg_const(t + o * TPA);
                        }
                        __builtin_nontemporal_store(res[o] - b, dst + o * TPA);
                    }
                    // row map: the request's completion stamp, once the
                    // stores are issued (not waited for: the stamp is a
                    // latency figure, not a completion signal)
                    if (rowmap && t == 0) atomicMax(A.finish + rr[u].req, (unsigned long long)wall_clock64());
                }
            }
        }
        if constexpr (PIPE && SPLIT) {
            if (rowmap || adv(a0) < a_end) stage_rest(adv(a0));
        }
    }

    if constexpr (rowmap) {
        // the last workgroup of the launches sharing the tickets to stop
        // taking them clears the counters for the next run (its own last
        // request is complete before it counts itself)
        if (dyn && threadIdx.x == 0) {
            __threadfence();
            if (atomicAdd(A.tickets + 1, 1u) == A.ticket_wgs - 1) {
                atomicExch(A.tickets, 0u);
                atomicExch(A.tickets + 1, 0u);
            }
        }
    }

    if constexpr (STATS) {
#pragma unroll
        for (int u = 0; u < AS; ++u) {
            if (cur_seg[u] >= 0) {
#pragma unroll
                for (int o = 0; o < 4; ++o) {
                    double* d = A.acc + ((size_t)cur_seg[u] * G::HALF + t + o * TPA) * 3;
                    unsafeAtomicAdd(d, sx[u][o]);
                    unsafeAtomicAdd(d + 1, sy[u][o]);
                    unsafeAtomicAdd(d + 2, sxx[u][o]);
                }
            }
        }
    }
}

// One launch over a frame: persistent workgroups (the grid is sized to the
// resident capacity) stride over the A-scans; in M_STATS mode each
// workgroup takes one contiguous chunk so that segment runs are long.
// Occupancy targets: left to itself the compiler often lands a few
// registers above 168 and so at two waves per SIMD where three fit without
// spills; three is asked for wherever it needs no spills (every
// throughput-mode kernel up to 4 taps, and DC removal of 8/16-bit samples).
// Lanczos (T = 16) holds 128 tap weights per lane: two waves per SIMD with a
// few spills measured faster than the one the compiler picks.  Lanczos-4
// (T = 8, 64 weights) fits three on the swap schedule with its LDS
// twiddles (twl_swap; with DC removal of 8/16-bit samples 8-20 bytes of
// spill per lane), two elsewhere.  The FPN statistics modes and the
// double-precision DC prologue keep the default.
template <int N, int T, int MODE, int IW, int AS, int SCHED>
constexpr int chain_waves_per_eu() {
    if (T >= 8)
        return (T == 8 && ChainLds<N, MODE, IW, AS, chain_wg<N>(), SCHED>::template TWLS<T> &&
                !(IW == 4 && (MODE & M_DC)))
                   ? 3
                   : 2;
    if ((MODE & M_STATS) || (IW == 4 && (MODE & M_DC))) return OPE_WAVES_PER_EU;
    // N = 4096 with the LDS twiddles: two 512-lane workgroups per CU
    if (ChainLds<N, MODE, IW, AS, chain_wg<N>(), SCHED>::template TWL<T>) return 4;
    return 3;
}

template <int N, int T, int MODE, int IW, int AS = 1, int SCHED = -1>
__global__ void __launch_bounds__(chain_wg<N>())
__attribute__((amdgpu_waves_per_eu(chain_waves_per_eu<N, T, MODE, IW, AS, SCHED>())))
chain_kernel(ChainArgs A) {
    constexpr int WG = chain_wg<N>();
    using L = ChainLds<N, MODE, IW, AS, WG, SCHED>;
    __shared__ float2 bufA[L::BUF];
    __shared__ float2 bufB[L::BUF];
    __shared__ float xsep[L::X];
    __shared__ typename L::DcSum scan_part[L::SCAN];
    __shared__ float mm_part[L::MM];
    __shared__ float2 twl[(L::template TWL<T> || L::template TWLS<T>) ? kTwlEntries : 1];
    constexpr int SLOTS = L::SLOTS;
    long long a_begin, a_step, a_end;
    if constexpr ((MODE & M_STATS) != 0) {
        const long long n_work = A.stat_rows;
        const long long groups = (n_work + SLOTS - 1) / SLOTS;
        const long long per = (groups + gridDim.x - 1) / gridDim.x;
        a_begin = (long long)blockIdx.x * per * SLOTS;
        a_end = a_begin + per * SLOTS;
        if (a_end > n_work) a_end = n_work;
        a_step = SLOTS;
    } else {
        a_begin = (long long)blockIdx.x * SLOTS;
        a_end = A.n_ascans;
        a_step = (long long)gridDim.x * SLOTS;
    }
    chain_body<N, T, MODE, IW, AS, WG, SCHED>(A, bufA, bufB, xsep, scan_part, mm_part, twl, a_begin, a_end, a_step);
}

}  // namespace ope_hip
