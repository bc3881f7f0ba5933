// Ragged batches: requests of different signal lengths and A-scan counts in
// one persistent launch.
//
// A batch is cut into tasks (consecutive A-scans of one request).  Every
// workgroup of 512 lanes takes the next task from a device counter, runs the
// chain body of that request's length on it (N/8 lanes per A-scan, so
// 512/(N/8) A-scans per workgroup iteration), and takes the next task until
// none is left.  Tasks are taken in request order.
//
// The launch leaves its counters as it found them, so a batch can be run
// again (or replayed from a graph) without clearing anything: the last
// workgroup to finish resets the task counter, and the task that completes a
// request stamps the request's finish time and restores its task count.
#pragma once

#include "chain_kernels.hpp"

namespace ope_hip {

struct RaggedRequest {
    const void* in;      // [n_ascans][N] samples
    float* out;          // [n_ascans][N/2]
    long long n_ascans;
};

struct RaggedTask {
    int req;             // request index
    int cls;             // length class (index into the class arguments)
    int n;               // signal length of the class
    int ntasks;          // tasks of this request
    long long a0, a1;    // A-scans [a0, a1) of the request
};

struct RaggedArgs {
    const ChainArgs* cls;            // per length class: tables and epilogue constants
    const RaggedRequest* reqs;
    const RaggedTask* tasks;
    int n_tasks;
    unsigned* counters;              // [2]: next task, workgroups done (both 0 between launches)
    unsigned* remaining;             // [n_reqs]: tasks of the request not yet finished
    unsigned long long* finish;      // [n_reqs]: wall-clock tick when the request completed
    unsigned long long* start;       // [1]: wall-clock tick when the first task was taken
};

constexpr int kRaggedWG = 512;

// FFT schedule of a ragged body: the swap schedule where it is built
// (N = 1024, 2048), as in the chain's default launch; its separate staging
// array is no larger than the one the 256- and 512-point bodies need anyway
template <int N>
constexpr int ragged_sched() {
    return (N == 1024 || N == 2048) ? Geo<N>::NPASS : -1;
}

template <int MODE, int IW>
struct RaggedLds {
    template <int N>
    using L = ChainLds<N, MODE, IW, 1, kRaggedWG, ragged_sched<N>()>;
    static constexpr int cmax(int a, int b) { return a > b ? a : b; }
    static constexpr int BUF = cmax(cmax(cmax(L<256>::BUF, L<512>::BUF), cmax(L<1024>::BUF, L<2048>::BUF)), L<4096>::BUF);
    static constexpr int X = cmax(cmax(cmax(L<256>::X, L<512>::X), cmax(L<1024>::X, L<2048>::X)), L<4096>::X);
    static constexpr int SCAN = L<256>::SCAN;  // depends on the workgroup only
    static constexpr int MM = L<256>::MM;
};

// One task of one length.  Not inlined: with the five bodies inlined into
// one kernel the register allocator spills heavily (over 800 B per lane);
// as separate functions each body is allocated on its own.
template <int N, int T, int MODE, int IW>
__device__ __attribute__((noinline)) void ragged_task(const ChainArgs& A, float2* bufA, float2* bufB, float* xsep,
        typename ChainLds<N, MODE, IW, 1, kRaggedWG>::DcSum* scan_part, float* mm_part, long long a0, long long a1) {
    constexpr int SCHED = ragged_sched<N>();
    using L = ChainLds<N, MODE, IW, 1, kRaggedWG, SCHED>;
    static_assert(L::X <= RaggedLds<MODE, IW>::X, "staging array");
    chain_body<N, T, MODE, IW, 1, kRaggedWG, SCHED>(A, bufA, bufB, xsep, scan_part, mm_part, a0, a1,
                                                    L::SLOTS);
}

template <int T, int MODE, int IW>
__global__ void __launch_bounds__(kRaggedWG) ragged_kernel(RaggedArgs R) {
    using S = RaggedLds<MODE, IW>;
    __shared__ float2 bufA[S::BUF];
    __shared__ float2 bufB[S::BUF];
    __shared__ float xsep[S::X];
    __shared__ typename ChainLds<256, MODE, IW, 1, kRaggedWG>::DcSum scan_part[S::SCAN];
    __shared__ float mm_part[S::MM];
    __shared__ int task_sh;

    for (;;) {
        if (threadIdx.x == 0) {
            const int ti = (int)atomicAdd(&R.counters[0], 1u);
            if (ti == 0) *R.start = wall_clock64();
            task_sh = ti;
        }
        __syncthreads();
        // uniform over the workgroup: scalar loads of the task and its
        // arguments, and a scalar branch to the body of its length
        const int ti = __builtin_amdgcn_readfirstlane(task_sh);
        __syncthreads();  // every lane has read task_sh before lane 0 rewrites it
        if (ti >= R.n_tasks) break;
        const RaggedTask tk = R.tasks[ti];
        ChainArgs A = R.cls[tk.cls];
        const RaggedRequest rq = R.reqs[tk.req];
        A.in = rq.in;
        A.out = rq.out;
        A.n_ascans = rq.n_ascans;
        switch (tk.n) {
            case 256:
                ragged_task<256, T, MODE, IW>(A, bufA, bufB, xsep, scan_part, mm_part, tk.a0, tk.a1);
                break;
            case 512:
                ragged_task<512, T, MODE, IW>(A, bufA, bufB, xsep, scan_part, mm_part, tk.a0, tk.a1);
                break;
            case 1024:
                ragged_task<1024, T, MODE, IW>(A, bufA, bufB, xsep, scan_part, mm_part, tk.a0, tk.a1);
                break;
            case 2048:
                ragged_task<2048, T, MODE, IW>(A, bufA, bufB, xsep, scan_part, mm_part, tk.a0, tk.a1);
                break;
            default:
                ragged_task<4096, T, MODE, IW>(A, bufA, bufB, xsep, scan_part, mm_part, tk.a0, tk.a1);
                break;
        }
        __syncthreads();  // all outputs of the task are written
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicSub(&R.remaining[tk.req], 1u) == 1u) {
                R.finish[tk.req] = wall_clock64();
                R.remaining[tk.req] = (unsigned)tk.ntasks;
            }
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&R.counters[1], 1u) == gridDim.x - 1) {
            // every workgroup has left the task loop: no more increments of
            // counters[0] in this launch
            atomicExch(&R.counters[0], 0u);
            atomicExch(&R.counters[1], 0u);
        }
    }
}

}  // namespace ope_hip
