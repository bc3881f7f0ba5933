// Chain-kernel instantiations of one signal length and one number of
// A-scans per lane, OPE_CHAIN_N and OPE_CHAIN_AS (set on the compiler line).
//  AS = 1: taps {1, 2, 4, 8, 16} x modes {plain, DC} x input widths {1, 2, 4}
//          bytes x every radix placement, plus the FPN-statistics modes with
//          the default placement;
//  AS = 2: taps x modes {plain, DC, pairs} x widths, default placement;
//  both:   for N = 1024 and 2048 also the swap schedule (sched = passes);
//  row maps (ragged batches): the default launch only, i.e. modes {plain,
//          DC} at AS = 1 and pairs at AS = 2, with the swap schedule where
//          it is built and the last-pass placement elsewhere.
#include "chain_dispatch.h"

#if !defined(OPE_CHAIN_N) || !defined(OPE_CHAIN_AS)
#error "compile with -DOPE_CHAIN_N=<signal length> -DOPE_CHAIN_AS=<1|2>"
#endif

namespace ope_hip {
namespace {

template <int N, int MODE, int IW, int AS, int SCHED>
KernelFn pick_taps(int taps) {
    switch (taps) {
        case 1: return chain_kernel<N, 1, MODE, IW, AS, SCHED>;
        case 2: return chain_kernel<N, 2, MODE, IW, AS, SCHED>;
        case 4: return chain_kernel<N, 4, MODE, IW, AS, SCHED>;
        case 8: return chain_kernel<N, 8, MODE, IW, AS, SCHED>;
        default: return chain_kernel<N, 16, MODE, IW, AS, SCHED>;
    }
}

template <int N, int MODE, int AS, int SCHED>
KernelFn pick_width(int taps, int iw) {
    switch (iw) {
        case 1: return pick_taps<N, MODE, 1, AS, SCHED>(taps);
        case 2: return pick_taps<N, MODE, 2, AS, SCHED>(taps);
        default: return pick_taps<N, MODE, 4, AS, SCHED>(taps);
    }
}

// the placements of the small-radix pass that differ for this N
template <int N, int MODE, int AS>
KernelFn pick_sched(int taps, int iw, int sched) {
    constexpr int NP = FftPlan<N, -1>::NP;
    constexpr int LAST = NP - 1;
    // NP: the swap schedule (exchange into the small-radix pass in registers)
    if (sched == NP) {
        if constexpr (swap_sched<N, NP>()) return pick_width<N, MODE, AS, NP>(taps, iw);
        return nullptr;
    }
    if constexpr (AS == 1 && FftPlan<N, -1>::SMALL != 8) {
        switch (sched) {
            case 0: return pick_width<N, MODE, AS, 0>(taps, iw);
            case 1: return pick_width<N, MODE, AS, (1 < LAST ? 1 : LAST)>(taps, iw);
            case 2: return pick_width<N, MODE, AS, (2 < LAST ? 2 : LAST)>(taps, iw);
            default: break;
        }
    }
    return pick_width<N, MODE, AS, LAST>(taps, iw);
}

}  // namespace

template <int N, int AS>
ChainLaunch pick_chain(int taps, int mode, int iw, int sched) {
    ChainLaunch l;
    l.wg = chain_wg<N>();
    l.slots = chain_wg<N>() / Geo<N>::TPA * AS;
    constexpr int LAST = FftPlan<N, -1>::NP - 1;
    l.fn = nullptr;
    switch (mode) {
        case 0: l.fn = pick_sched<N, 0, AS>(taps, iw, sched); break;
        case M_DC: l.fn = pick_sched<N, M_DC, AS>(taps, iw, sched); break;
        case M_STATS:
            if constexpr (AS == 1) l.fn = pick_width<N, M_STATS, 1, LAST>(taps, iw);
            break;
        case M_PAIR:
            if constexpr (AS == 2) {
                constexpr int NP = LAST + 1;
                if (sched == NP) {
                    if constexpr (swap_sched<N, NP>()) l.fn = pick_width<N, M_PAIR, 2, NP>(taps, iw);
                } else {
                    l.fn = pick_width<N, M_PAIR, 2, LAST>(taps, iw);
                }
            }
            break;
        case M_DC | M_STATS:
            if constexpr (AS == 1) l.fn = pick_width<N, M_DC | M_STATS, 1, LAST>(taps, iw);
            break;
        default: {
            // row maps: the default schedule only (sched < 0 where it is
            // the last-pass placement)
            constexpr int DEF = (N == 1024 || N == 2048) ? LAST + 1 : LAST;
            if (sched != DEF && !(sched < 0 && DEF == LAST)) break;
            if constexpr (AS == 1) {
                if (mode == M_ROWS) l.fn = pick_width<N, M_ROWS, 1, DEF>(taps, iw);
                if (mode == (M_ROWS | M_DC)) l.fn = pick_width<N, M_ROWS | M_DC, 1, DEF>(taps, iw);
            } else {
                if (mode == (M_ROWS | M_PAIR)) l.fn = pick_width<N, M_ROWS | M_PAIR, 2, DEF>(taps, iw);
            }
            break;
        }
    }
    return l;
}

template ChainLaunch pick_chain<OPE_CHAIN_N, OPE_CHAIN_AS>(int, int, int, int);

}  // namespace ope_hip
