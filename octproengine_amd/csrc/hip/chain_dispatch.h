// Selection of a chain-kernel instantiation.  Every signal length N has its
// own translation unit (chain_instances.hip compiled with -DOPE_CHAIN_N=N),
// so the instantiations over taps x mode x input width compile in parallel.
#pragma once

#include "chain_kernels.hpp"
#include "autofocus_kernels.hpp"

namespace ope_hip {

using KernelFn = void (*)(ChainArgs);

struct ChainLaunch {
    KernelFn fn;
    int wg;      // workgroup size
    int slots;   // A-scans per workgroup iteration
};

// taps: 1, 2, 4, 8 or 16; mode: M_DC / M_STATS bits; iw: input sample width in
// bytes (1, 2, 4); AS: A-scans per lane (1, or 2 for the modes without
// statistics); sched: position of the small-radix FFT pass (0 ..
// passes - 1; other values: the last pass).  fn is null for a combination
// that is not instantiated.  Explicitly instantiated for N = 256 ... 4096
// and AS = 1, 2 (chain_instances.hip).
template <int N, int AS>
ChainLaunch pick_chain(int taps, int mode, int iw, int sched);

using AutofocusFn = void (*)(AutofocusArgs);

struct AutofocusLaunch {
    AutofocusFn fn;
    int wg;
    int slots;
};

// Autofocus forward (bwd = false) or backward kernel of length N for taps 2
// (linear) or 4 (cubic), input width iw (2 or 4 bytes) and DC removal on or
// off.  Explicitly instantiated for N = 256 ... 4096 (autofocus_instances.hip
// compiled with -DOPE_AF_N=N).
template <int N>
AutofocusLaunch pick_autofocus_n(int taps, int iw, bool dc, bool bwd);

}  // namespace ope_hip
