// Driver for the reference's own curve generators.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's unmodified src/core/processorconfiguration.cpp and
// src/core/types.cpp (read in place under the reference tree) into
// oracle/_ref/ref_curves.  tests/golden/make_curve_fixtures.py runs it to
// write golden curves that pin the oracle's and the C++ host's curve
// generators bit-exactly.
//
// Usage:
//   ref_curves resample N c0 c1 c2 c3
//   ref_curves dispersion N d0 d1 d2 d3
//   ref_curves window N type center fill
// Writes N float32 values, little endian, to stdout.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "processorconfiguration.h"

int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: ref_curves {resample|dispersion|window} N a b c [d]\n");
        return 2;
    }
    ope::ProcessorConfiguration cfg;
    const int n = std::atoi(argv[2]);
    cfg.dataParams.signalLength = n;
    const float* data = nullptr;
    size_t size = 0;
    if (std::strcmp(argv[1], "resample") == 0 && argc >= 7) {
        for (int i = 0; i < 4; ++i) cfg.resamplingParams.coefficients[i] = std::strtof(argv[3 + i], nullptr);
        data = cfg.getGeneratedResamplingCurve();
        size = cfg.getGeneratedResamplingCurveSize();
    } else if (std::strcmp(argv[1], "dispersion") == 0 && argc >= 7) {
        for (int i = 0; i < 4; ++i) cfg.dispersionParams.coefficients[i] = std::strtof(argv[3 + i], nullptr);
        data = cfg.getGeneratedDispersionCurve();
        size = cfg.getGeneratedDispersionCurveSize();
    } else if (std::strcmp(argv[1], "window") == 0 && argc >= 6) {
        cfg.windowingParams.windowType = static_cast<ope::WindowType>(std::atoi(argv[3]));
        cfg.windowingParams.windowCenterPosition = std::strtof(argv[4], nullptr);
        cfg.windowingParams.windowFillFactor = std::strtof(argv[5], nullptr);
        data = cfg.getGeneratedWindowCurve();
        size = cfg.getGeneratedWindowCurveSize();
    } else {
        std::fprintf(stderr, "unknown curve kind\n");
        return 2;
    }
    if (std::fwrite(data, sizeof(float), size, stdout) != size) return 1;
    return 0;
}
