// ProcessingBackend on AMD Instinct MI355X through the C ABI of
// libope_hip.so (include/ope_hip.h).  Takes the place of OCTproEngine's
// CudaBackend (src/backends/cuda/cuda_backend.h) for Backend::CUDA.
#ifndef OPE_HIP_BACKEND_H
#define OPE_HIP_BACKEND_H

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <vector>

#include "backend_interface.h"

struct ope_hip_ctx;

namespace ope {

class HipBackend final : public ProcessingBackend {
public:
    HipBackend();
    ~HipBackend() override;

    void initialize(const ProcessorConfiguration& config) override;
    void cleanup() override;
    void setOutputCallback(std::function<void(const IOBuffer&)> callback) override;
    void process(IOBuffer& input) override;
    void updateConfig(const ProcessorConfiguration& config) override;

    void updateResamplingCurve(const float* curve, size_t length) override;
    void updateDispersionCurve(const float* curve, size_t length) override;
    void updateWindowCurve(const float* curve, size_t length) override;

    IOBuffer& getInputBuffer(int index) override;
    IOBuffer& getNextAvailableInputBuffer() override;
    int getNumInputBuffers() const override;

    void requestPostProcessBackgroundRecording() override;
    void setPostProcessBackgroundProfile(const float* background, size_t length) override;
    const std::vector<float>& getPostProcessBackgroundProfile() const override;

    void requestFixedPatternNoiseDetermination() override;
    void setFixedPatternNoiseProfile(const float* profileInterleaved, size_t complexPairs) override;
    const std::vector<float>& getFixedPatternNoiseProfile() const override;

    std::vector<float> convertInput(const void* input, IOBuffer::DataType inputType, int bitDepth, int samples,
                                    bool applyBitshift) override;
    std::vector<float> rollingAverageBackgroundRemoval(const float* input, int windowSize, int lineWidth,
                                                       int numLines) override;
    std::vector<float> kLinearization(const float* input, const float* resampleCurve, InterpolationMethod method,
                                      int lineWidth, int samples) override;
    std::vector<float> windowing(const float* input, const float* windowCurve, int lineWidth,
                                 int samples) override;
    std::vector<float> dispersionCompensation(const float* input, const float* phaseComplex, int lineWidth,
                                              int samples) override;
    std::vector<float> kLinearizationAndWindowing(const float* input, const float* resampleCurve,
                                                  const float* windowCurve, InterpolationMethod method,
                                                  int lineWidth, int samples) override;
    std::vector<float> kLinearizationAndWindowingAndDispersion(const float* input, const float* resampleCurve,
                                                               const float* windowCurve, const float* phaseComplex,
                                                               InterpolationMethod method, int lineWidth,
                                                               int samples) override;
    std::vector<float> dispersionCompensationAndWindowing(const float* input, const float* phaseComplex,
                                                          const float* windowCurve, int lineWidth,
                                                          int samples) override;
    std::vector<float> fft(const float* input, int lineWidth, int samples) override;
    std::vector<float> ifft(const float* input, int lineWidth, int samples) override;
    std::vector<float> getMinimumVarianceMean(const float* input, int width, int height, int segments) override;
    std::vector<float> fixedPatternNoiseRemoval(const float* input, const float* meanALine, int lineWidth,
                                                int numLines) override;
    std::vector<float> postProcessTruncate(const float* input, bool logScaling, float grayscaleMax,
                                           float grayscaleMin, float addend, float multiplicator, int lineWidth,
                                           int samples) override;
    std::vector<float> bscanFlip(const float* input, int lineWidth, int linesPerBscan, int numBscans) override;
    std::vector<float> sinusoidalScanCorrection(const float* input, const float* resampleCurve, int lineWidth,
                                                int linesPerBscan, int numBscans) override;
    std::vector<float> postProcessBackgroundSubtraction(const float* input, const float* backgroundLine,
                                                        float weight, float offset, int lineWidth,
                                                        int samples) override;

    // Frames in flight, i.e. output buffers lent to callbacks in turn.
    static constexpr int kSlots = 2;
    // Input buffers in the pool (cpu_backend.cpp:82, cuda_backend.cu:66).
    static constexpr int kInputBuffers = 2;

private:
    // Callback contexts: stable for the lifetime of the context, so a
    // callback still in flight never reads a context that is being reused.
    struct SlotCtx {  // one per output buffer
        HipBackend* self = nullptr;
        IOBuffer* output = nullptr;
    };
    struct InputCtx {  // one per pooled input buffer
        HipBackend* self = nullptr;
        IOBuffer* input = nullptr;
    };

    void check(int rc, const char* what) const;
    void requireContext(const char* what) const;
    static void onInputFree(void* user);
    static void onFrameDone(void* user, float* hostOut);

    ProcessorConfiguration config_;
    ope_hip_ctx* ctx_ = nullptr;
    std::function<void(const IOBuffer&)> callback_;
    std::mutex callbackMu_;

    std::vector<std::unique_ptr<IOBuffer>> inputs_;
    std::vector<std::unique_ptr<IOBuffer>> outputs_;
    SlotCtx slots_[kSlots];
    InputCtx inputCtx_[kInputBuffers];
    int nextSlot_ = 0;

    std::deque<IOBuffer*> free_;
    std::mutex freeMu_;
    std::condition_variable freeCv_;

    // requests made before initialize(), applied when the context exists
    bool pendingFpnRequest_ = false;
    bool pendingBgRequest_ = false;

    mutable std::vector<float> fpnCache_;
    mutable std::vector<float> bgCache_;
};

}  // namespace ope

#endif  // OPE_HIP_BACKEND_H
