// Autofocus kernel instantiations of one signal length, OPE_AF_N (set on the
// compiler line): taps {2, 4} x input widths {2, 4} bytes x DC removal
// {off, on} x {forward, backward}.
#include "chain_dispatch.h"

#ifndef OPE_AF_N
#error "compile with -DOPE_AF_N=<signal length>"
#endif

namespace ope_hip {

namespace {

template <int N, int T, int IW, bool DC>
AutofocusFn pick_dir(bool bwd) {
    return bwd ? autofocus_kernel<N, T, IW, DC, true> : autofocus_kernel<N, T, IW, DC, false>;
}

template <int N, int T, int IW>
AutofocusFn pick_dc(bool dc, bool bwd) {
    return dc ? pick_dir<N, T, IW, true>(bwd) : pick_dir<N, T, IW, false>(bwd);
}

template <int N, int T>
AutofocusFn pick_iw(int iw, bool dc, bool bwd) {
    return iw == 2 ? pick_dc<N, T, 2>(dc, bwd) : pick_dc<N, T, 4>(dc, bwd);
}

}  // namespace

template <int N>
AutofocusLaunch pick_autofocus_n(int taps, int iw, bool dc, bool bwd) {
    AutofocusLaunch l;
    l.wg = Geo<N>::WG;
    l.slots = Geo<N>::APW;
    l.fn = taps == 2 ? pick_iw<N, 2>(iw, dc, bwd) : pick_iw<N, 4>(iw, dc, bwd);
    return l;
}

template AutofocusLaunch pick_autofocus_n<OPE_AF_N>(int, int, bool, bool);

}  // namespace ope_hip
