// The backend plugin contract behind ope::Processor: the same pure virtual
// interface as OCTproEngine's src/backends/backend_interface.h:12-165, so a
// backend written for one builds against the other.
#ifndef OPE_BACKEND_INTERFACE_H
#define OPE_BACKEND_INTERFACE_H

#include <functional>
#include <vector>

#include "iobuffer.h"
#include "processorconfiguration.h"

namespace ope {

class ProcessingBackend {
public:
    virtual ~ProcessingBackend() = default;

    // lifecycle
    virtual void initialize(const ProcessorConfiguration& config) = 0;
    virtual void cleanup() = 0;
    virtual void setOutputCallback(std::function<void(const IOBuffer&)> callback) = 0;

    // asynchronous processing; output goes to the callback
    virtual void process(IOBuffer& input) = 0;
    virtual void updateConfig(const ProcessorConfiguration& config) = 0;

    // curve hot-swap (resampling and window: N floats; dispersion: 2N
    // floats, interleaved cos/sin)
    virtual void updateResamplingCurve(const float* curve, size_t length) = 0;
    virtual void updateDispersionCurve(const float* curve, size_t length) = 0;
    virtual void updateWindowCurve(const float* curve, size_t length) = 0;

    // input buffer pool
    virtual IOBuffer& getInputBuffer(int index) = 0;
    virtual IOBuffer& getNextAvailableInputBuffer() = 0;
    virtual int getNumInputBuffers() const = 0;

    // post-process background
    virtual void requestPostProcessBackgroundRecording() = 0;
    virtual void setPostProcessBackgroundProfile(const float* background, size_t length) = 0;
    virtual const std::vector<float>& getPostProcessBackgroundProfile() const = 0;

    // fixed-pattern noise (interleaved re, im)
    virtual void requestFixedPatternNoiseDetermination() = 0;
    virtual void setFixedPatternNoiseProfile(const float* profileInterleaved, size_t complexPairs) = 0;
    virtual const std::vector<float>& getFixedPatternNoiseProfile() const = 0;

    // individual operations (testing)
    virtual std::vector<float> convertInput(const void* input, IOBuffer::DataType inputType, int bitDepth,
                                            int samples, bool applyBitshift) = 0;
    virtual std::vector<float> rollingAverageBackgroundRemoval(const float* input, int windowSize, int lineWidth,
                                                               int numLines) = 0;
    virtual std::vector<float> kLinearization(const float* input, const float* resampleCurve,
                                              InterpolationMethod method, int lineWidth, int samples) = 0;
    virtual std::vector<float> windowing(const float* input, const float* windowCurve, int lineWidth,
                                         int samples) = 0;
    virtual std::vector<float> dispersionCompensation(const float* input, const float* phaseComplex,
                                                      int lineWidth, int samples) = 0;
    virtual std::vector<float> kLinearizationAndWindowing(const float* input, const float* resampleCurve,
                                                          const float* windowCurve, InterpolationMethod method,
                                                          int lineWidth, int samples) = 0;
    virtual std::vector<float> kLinearizationAndWindowingAndDispersion(const float* input,
                                                                       const float* resampleCurve,
                                                                       const float* windowCurve,
                                                                       const float* phaseComplex,
                                                                       InterpolationMethod method, int lineWidth,
                                                                       int samples) = 0;
    virtual std::vector<float> dispersionCompensationAndWindowing(const float* input, const float* phaseComplex,
                                                                  const float* windowCurve, int lineWidth,
                                                                  int samples) = 0;
    virtual std::vector<float> fft(const float* input, int lineWidth, int samples) = 0;
    virtual std::vector<float> ifft(const float* input, int lineWidth, int samples) = 0;
    virtual std::vector<float> getMinimumVarianceMean(const float* input, int width, int height,
                                                      int segments) = 0;
    virtual std::vector<float> fixedPatternNoiseRemoval(const float* input, const float* meanALine, int lineWidth,
                                                        int numLines) = 0;
    virtual std::vector<float> postProcessTruncate(const float* input, bool logScaling, float grayscaleMax,
                                                   float grayscaleMin, float addend, float multiplicator,
                                                   int lineWidth, int samples) = 0;
    virtual std::vector<float> bscanFlip(const float* input, int lineWidth, int linesPerBscan, int numBscans) = 0;
    virtual std::vector<float> sinusoidalScanCorrection(const float* input, const float* resampleCurve,
                                                        int lineWidth, int linesPerBscan, int numBscans) = 0;
    virtual std::vector<float> postProcessBackgroundSubtraction(const float* input, const float* backgroundLine,
                                                                float weight, float offset, int lineWidth,
                                                                int samples) = 0;
};

}  // namespace ope

#endif  // OPE_BACKEND_INTERFACE_H
