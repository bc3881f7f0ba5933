// Selection of a chain-kernel instantiation.  Every signal length N has its
// own translation unit (chain_instances.hip compiled with -DOPE_CHAIN_N=N),
// so the instantiations over taps x mode x input width compile in parallel.
#pragma once

#include "chain_kernels.hpp"

namespace ope_hip {

using KernelFn = void (*)(ChainArgs);

struct ChainLaunch {
    KernelFn fn;
    int wg;      // workgroup size
    int slots;   // A-scans per workgroup iteration
};

// taps: 1, 2, 4 or 16; mode: M_DC / M_STATS bits; iw: input sample width in
// bytes (1, 2, 4).  Explicitly instantiated for N = 256 ... 4096.
template <int N>
ChainLaunch pick_chain_n(int taps, int mode, int iw);

}  // namespace ope_hip
