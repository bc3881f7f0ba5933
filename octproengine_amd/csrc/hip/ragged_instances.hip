// Ragged-batch kernel instantiations for one tap count, OPE_RAGGED_T (set on
// the compiler line): modes {plain, DC} x input widths {1, 2, 4} bytes, each
// holding the chain bodies of N = 256 ... 4096.
#include "chain_dispatch.h"

#ifndef OPE_RAGGED_T
#error "compile with -DOPE_RAGGED_T=<taps>"
#endif

namespace ope_hip {

namespace {

template <int T, int MODE>
RaggedFn pick_width(int iw) {
    switch (iw) {
        case 1: return ragged_kernel<T, MODE, 1>;
        case 2: return ragged_kernel<T, MODE, 2>;
        default: return ragged_kernel<T, MODE, 4>;
    }
}

}  // namespace

template <int T>
RaggedFn pick_ragged_t(int mode, int iw) {
    return mode == M_DC ? pick_width<T, M_DC>(iw) : pick_width<T, 0>(iw);
}

template RaggedFn pick_ragged_t<OPE_RAGGED_T>(int, int);

}  // namespace ope_hip
