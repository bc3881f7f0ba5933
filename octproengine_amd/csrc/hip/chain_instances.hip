// Chain-kernel instantiations of one signal length, OPE_CHAIN_N (set on the
// compiler line).  The table covers taps {1, 2, 4, 16} x modes {plain, DC,
// statistics, DC + statistics} x input widths {1, 2, 4} bytes.
#include "chain_dispatch.h"

#ifndef OPE_CHAIN_N
#error "compile with -DOPE_CHAIN_N=<signal length>"
#endif

namespace ope_hip {
namespace {

// A-scans per lane.  Two (instruction-level parallelism) measured slower
// than one on gfx950 for every mode: the extra VGPRs and LDS cost a wave of
// occupancy per SIMD, which hides more latency than the second A-scan does.
constexpr int kAscansPerLane = 1;

template <int N, int MODE, int IW>
KernelFn pick_taps(int taps) {
    constexpr int AS = kAscansPerLane;
    switch (taps) {
        case 1: return chain_kernel<N, 1, MODE, IW, AS>;
        case 2: return chain_kernel<N, 2, MODE, IW, AS>;
        case 4: return chain_kernel<N, 4, MODE, IW, AS>;
        default: return chain_kernel<N, 16, MODE, IW, AS>;
    }
}

template <int N, int MODE>
KernelFn pick_width(int taps, int iw) {
    switch (iw) {
        case 1: return pick_taps<N, MODE, 1>(taps);
        case 2: return pick_taps<N, MODE, 2>(taps);
        default: return pick_taps<N, MODE, 4>(taps);
    }
}

}  // namespace

template <int N>
ChainLaunch pick_chain_n(int taps, int mode, int iw) {
    ChainLaunch l;
    l.wg = Geo<N>::WG;
    l.slots = Geo<N>::APW * kAscansPerLane;
    switch (mode) {
        case 0: l.fn = pick_width<N, 0>(taps, iw); break;
        case M_DC: l.fn = pick_width<N, M_DC>(taps, iw); break;
        case M_STATS: l.fn = pick_width<N, M_STATS>(taps, iw); break;
        default: l.fn = pick_width<N, M_DC | M_STATS>(taps, iw); break;
    }
    return l;
}

template ChainLaunch pick_chain_n<OPE_CHAIN_N>(int, int, int);

}  // namespace ope_hip
