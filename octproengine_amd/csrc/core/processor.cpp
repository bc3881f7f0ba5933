// ope::Processor.  Behaviour follows OCTproEngine's src/core/processor.cpp:
//  - lazy initialisation and re-initialisation when the data dimensions
//    change (:127-187);
//  - curve selection (custom vs generated) and the phase -> interleaved
//    (cos, sin) conversion sent to the backend (:53-125);
//  - scalar toggles change only the Processor's configuration; the backend
//    sees them at (re)initialisation or on updateConfig (:420-592), exactly
//    like the reference;
//  - profile CSV formats "Sample Number;Real;Imag" and
//    "Sample Number;Sample Value" (:628-694, :764-843).
// Backend::CUDA is served by the HIP backend.  Backend::CPU (the reference's
// FFTW backend) is not part of this library.
#include "processor.h"

#include <cmath>
#include <fstream>
#include <stdexcept>
#include <string>

#include "../backends/backend_interface.h"
#include "../backends/hip_backend.h"
#include "callback_manager.h"

namespace ope {

class Processor::Impl {
public:
    explicit Impl(Backend type) { createBackend(type); }

    ~Impl() {
        if (backend && initialized) backend->cleanup();
    }

    void createBackend(Backend type) {
        if (type == Backend::CUDA) {
            backend.reset(new HipBackend());
        } else if (type == Backend::CPU) {
            throw std::runtime_error(
                "Backend::CPU is not part of this library; Backend::CUDA runs on the HIP backend");
        } else {
            throw std::runtime_error("Unknown backend type");
        }
        backendType = type;
    }

    std::vector<float> resamplingCurve() const {
        if (config.resamplingParams.useCustomCurve && config.hasCustomResamplingCurve()) {
            const float* d = config.getCustomResamplingCurve();
            return std::vector<float>(d, d + config.getCustomResamplingCurveSize());
        }
        const float* d = config.getGeneratedResamplingCurve();
        return std::vector<float>(d, d + config.getGeneratedResamplingCurveSize());
    }

    std::vector<float> windowCurve() const {
        if (config.windowingParams.useCustomCurve && config.hasCustomWindowCurve()) {
            const float* d = config.getCustomWindowCurve();
            return std::vector<float>(d, d + config.getCustomWindowCurveSize());
        }
        const float* d = config.getGeneratedWindowCurve();
        return std::vector<float>(d, d + config.getGeneratedWindowCurveSize());
    }

    // Phase (custom or generated) as interleaved float (cos, sin).
    std::vector<float> dispersionPhasor() const {
        const float* phase;
        size_t n;
        if (config.dispersionParams.useCustomCurve && config.hasCustomDispersionCurve()) {
            phase = config.getCustomDispersionCurve();
            n = config.getCustomDispersionCurveSize();
        } else {
            phase = config.getGeneratedDispersionCurve();
            n = config.getGeneratedDispersionCurveSize();
        }
        std::vector<float> out(2 * n);
        for (size_t i = 0; i < n; ++i) {
            out[2 * i] = std::cos(phase[i]);
            out[2 * i + 1] = std::sin(phase[i]);
        }
        return out;
    }

    void pushResamplingCurve() {
        if (!initialized) return;
        std::vector<float> c = resamplingCurve();
        backend->updateResamplingCurve(c.data(), c.size());
    }
    void pushWindowCurve() {
        if (!initialized) return;
        std::vector<float> c = windowCurve();
        backend->updateWindowCurve(c.data(), c.size());
    }
    void pushDispersionCurve() {
        if (!initialized) return;
        std::vector<float> c = dispersionPhasor();
        backend->updateDispersionCurve(c.data(), c.size());
    }
    void pushAllCurves() {
        pushResamplingCurve();
        pushWindowCurve();
        pushDispersionCurve();
    }

    bool dimensionsDiffer(const ProcessorConfiguration::DataParameters& a,
                          const ProcessorConfiguration::DataParameters& b) const {
        return a.signalLength != b.signalLength || a.samplesPerBuffer != b.samplesPerBuffer ||
               a.ascansPerBscan != b.ascansPerBscan || a.bscansPerBuffer != b.bscansPerBuffer;
    }

    void ensureInitialized() {
        if (!initialized)
            initialize();
        else if (dimensionsDiffer(config.dataParams, lastInitData))
            reinitialize();
    }

    void initialize() {
        if (!config.validate()) throw std::runtime_error("Invalid processor configuration");
        if (initialized) backend->cleanup();
        backend->initialize(config);
        backend->setOutputCallback([this](const IOBuffer& out) { callbacks.invokeAll(out); });
        initialized = true;
        pushAllCurves();
        lastInitData = config.dataParams;
    }

    void reinitialize() {
        if (!config.validate()) throw std::runtime_error("Invalid processor configuration");
        backend->cleanup();
        backend->initialize(config);
        pushAllCurves();
        lastInitData = config.dataParams;
    }

    void cleanup() {
        if (initialized) {
            backend->cleanup();
            initialized = false;
        }
    }

    ProcessorConfiguration config;
    std::unique_ptr<ProcessingBackend> backend;
    Backend backendType = Backend::CUDA;
    bool initialized = false;
    ProcessorConfiguration::DataParameters lastInitData;
    CallbackManager callbacks;
};

// ---------------------------------------------------------------- lifecycle

Processor::Processor(Backend backend) : impl(new Impl(backend)) {}
Processor::~Processor() = default;

void Processor::initialize() { impl->ensureInitialized(); }
void Processor::cleanup() { impl->cleanup(); }
bool Processor::isInitialized() const { return impl->initialized; }

void Processor::loadConfigurationFromFile(const std::string& filepath) {
    if (!impl->config.loadFromFile(filepath)) throw std::runtime_error("Failed to load configuration from: " + filepath);
}

void Processor::saveConfigurationToFile(const std::string& filepath) const {
    if (!impl->config.saveToFile(filepath)) throw std::runtime_error("Failed to save configuration to: " + filepath);
}

const ProcessorConfiguration& Processor::getConfig() const { return impl->config; }

void Processor::setConfig(const ProcessorConfiguration& config) {
    const bool dimsChanged = impl->dimensionsDiffer(impl->config.dataParams, config.dataParams);
    impl->config = config;
    if (!impl->initialized) return;
    if (dimsChanged)
        impl->reinitialize();
    else
        impl->pushAllCurves();
}

void Processor::setInputParameters(int samplesPerRawAscan, int ascansPerBscan, int bscansPerBuffer, DataType type) {
    const int oldLength = impl->config.dataParams.signalLength;
    auto& d = impl->config.dataParams;
    d.signalLength = samplesPerRawAscan;
    d.ascansPerBscan = ascansPerBscan;
    d.bscansPerBuffer = bscansPerBuffer;
    d.samplesPerBuffer = samplesPerRawAscan * ascansPerBscan * bscansPerBuffer;
    d.inputDataType = type;
    if (samplesPerRawAscan != oldLength) {
        impl->config.adjustAllCustomCurves();
        if (impl->initialized) impl->pushAllCurves();
    }
}

void Processor::setBackend(Backend backend) {
    if (impl->backendType == backend) return;
    if (impl->initialized) {
        impl->backend->cleanup();
        impl->initialized = false;
    }
    impl->createBackend(backend);
}

Backend Processor::getBackend() const { return impl->backendType; }

// ---------------------------------------------------------------- processing and callbacks

Processor::CallbackId Processor::addOutputCallback(OutputCallback callback) {
    return impl->callbacks.addCallback(std::move(callback));
}
bool Processor::removeOutputCallback(CallbackId id) { return impl->callbacks.removeCallback(id); }
void Processor::clearOutputCallbacks() { impl->callbacks.clear(); }
size_t Processor::getCallbackCount() const { return impl->callbacks.getCallbackCount(); }

void Processor::setOutputCallback(OutputCallback callback) {
    impl->callbacks.clear();
    impl->callbacks.addCallback(std::move(callback));
}

void Processor::process(IOBuffer& input) {
    impl->ensureInitialized();
    impl->backend->process(input);
}

IOBuffer& Processor::getInputBuffer(int index) { return impl->backend->getInputBuffer(index); }

int Processor::getNumInputBuffers() const { return impl->initialized ? impl->backend->getNumInputBuffers() : 0; }

IOBuffer& Processor::getNextAvailableInputBuffer() { return impl->backend->getNextAvailableInputBuffer(); }

// ---------------------------------------------------------------- hot-swap: resampling

void Processor::setResamplingCoefficients(const float coefficients[4]) {
    auto& r = impl->config.resamplingParams;
    for (int i = 0; i < 4; ++i) r.coefficients[i] = coefficients[i];
    r.useCoefficients = true;
    r.useCustomCurve = false;
    impl->pushResamplingCurve();
}

void Processor::setCustomResamplingCurve(const float* curve, size_t length) {
    if (!curve || length == 0) throw std::invalid_argument("Invalid custom resampling curve");
    impl->config.setCustomResamplingCurve(curve, length);
    impl->config.resamplingParams.useCustomCurve = true;
    impl->config.resamplingParams.useCoefficients = false;
    impl->pushResamplingCurve();
}

void Processor::useCustomResamplingCurve(bool useCustom) {
    auto& r = impl->config.resamplingParams;
    if (useCustom && !impl->config.hasCustomResamplingCurve())
        throw std::runtime_error("No custom resampling curve set. Call setCustomResamplingCurve() first.");
    r.useCustomCurve = useCustom;
    r.useCoefficients = !useCustom;
    impl->pushResamplingCurve();
}

void Processor::enableResampling(bool enable) { impl->config.resamplingParams.enabled = enable; }
void Processor::setInterpolationMethod(InterpolationMethod method) {
    impl->config.resamplingParams.interpolationMethod = method;
}

// ---------------------------------------------------------------- hot-swap: dispersion

void Processor::setDispersionCoefficients(const float coefficients[4], float factor) {
    auto& d = impl->config.dispersionParams;
    for (int i = 0; i < 4; ++i) d.coefficients[i] = coefficients[i];
    d.factor = factor;  // stored, not applied (as the reference)
    d.useCoefficients = true;
    d.useCustomCurve = false;
    impl->pushDispersionCurve();
}

void Processor::setCustomDispersionCurve(const float* curve, size_t length) {
    if (!curve || length == 0) throw std::invalid_argument("Invalid custom dispersion curve");
    impl->config.setCustomDispersionCurve(curve, length);
    impl->config.dispersionParams.useCustomCurve = true;
    impl->config.dispersionParams.useCoefficients = false;
    impl->pushDispersionCurve();
}

void Processor::useCustomDispersionCurve(bool useCustom) {
    auto& d = impl->config.dispersionParams;
    if (useCustom && !impl->config.hasCustomDispersionCurve())
        throw std::runtime_error("No custom dispersion curve set. Call setCustomDispersionCurve() first.");
    d.useCustomCurve = useCustom;
    d.useCoefficients = !useCustom;
    impl->pushDispersionCurve();
}

void Processor::enableDispersionCompensation(bool enable) { impl->config.dispersionParams.enabled = enable; }

// ---------------------------------------------------------------- hot-swap: windowing

void Processor::setWindowParameters(WindowType type, float centerPosition, float fillFactor) {
    auto& w = impl->config.windowingParams;
    w.windowType = type;
    w.windowCenterPosition = centerPosition;
    w.windowFillFactor = fillFactor;
    w.useCustomCurve = false;  // a stored custom curve is kept for later toggling
    impl->pushWindowCurve();
}

void Processor::setCustomWindowCurve(const float* curve, size_t length) {
    if (!curve || length == 0) throw std::invalid_argument("Invalid custom window curve");
    impl->config.setCustomWindowCurve(curve, length);
    impl->config.windowingParams.useCustomCurve = true;
    impl->pushWindowCurve();
}

void Processor::useCustomWindowCurve(bool useCustom) {
    if (useCustom && !impl->config.hasCustomWindowCurve())
        throw std::runtime_error("No custom window curve set. Call setCustomWindowCurve() first.");
    impl->config.windowingParams.useCustomCurve = useCustom;
    impl->pushWindowCurve();
}

void Processor::enableWindowing(bool enable) { impl->config.windowingParams.enabled = enable; }

// ---------------------------------------------------------------- scalar toggles

void Processor::setGrayscaleRange(float min, float max) {
    impl->config.postProcessingParams.grayscaleMin = min;
    impl->config.postProcessingParams.grayscaleMax = max;
}

void Processor::setSignalMultiplicatorAndAddend(float multiplicator, float addend) {
    impl->config.postProcessingParams.multiplicator = multiplicator;
    impl->config.postProcessingParams.addend = addend;
}

void Processor::enableLogScaling(bool enable) { impl->config.postProcessingParams.logScaling = enable; }
void Processor::enableBackgroundRemoval(bool enable) { impl->config.backgroundRemovalParams.enabled = enable; }
void Processor::setBackgroundRemovalWindowSize(int windowSize) {
    impl->config.backgroundRemovalParams.rollingAverageWindowSize = windowSize;
}
void Processor::enableBscanFlip(bool enable) { impl->config.postProcessingParams.bscanFlip = enable; }
void Processor::enableSinusoidalScanCorrection(bool enable) {
    impl->config.postProcessingParams.sinusoidalScanCorrection = enable;
}
void Processor::enableFixedPatternNoiseRemoval(bool enable) {
    impl->config.postProcessingParams.fixedPatternNoiseRemoval = enable;
}
void Processor::enablePostProcessBackgroundSubtraction(bool enable) {
    impl->config.postProcessingParams.backgroundRemoval = enable;
}

// ---------------------------------------------------------------- post-process background

void Processor::requestPostProcessBackgroundRecording() { impl->backend->requestPostProcessBackgroundRecording(); }

void Processor::setPostProcessBackgroundWeight(float weight) {
    impl->config.postProcessingParams.backgroundWeight = weight;
    if (impl->initialized) impl->backend->updateConfig(impl->config);
}

void Processor::setPostProcessBackgroundOffset(float offset) {
    impl->config.postProcessingParams.backgroundOffset = offset;
    if (impl->initialized) impl->backend->updateConfig(impl->config);
}

const float* Processor::getPostProcessBackgroundProfile() const {
    if (!impl->initialized) return impl->config.getCustomPostProcessBackgroundProfile();
    const std::vector<float>& v = impl->backend->getPostProcessBackgroundProfile();
    return v.empty() ? nullptr : v.data();
}

size_t Processor::getPostProcessBackgroundProfileSize() const {
    if (!impl->initialized) return impl->config.getCustomPostProcessBackgroundProfileSize();
    return impl->backend->getPostProcessBackgroundProfile().size();
}

bool Processor::hasPostProcessBackgroundProfile() const {
    if (!impl->initialized) return impl->config.hasCustomPostProcessBackgroundProfile();
    return !impl->backend->getPostProcessBackgroundProfile().empty();
}

void Processor::setPostProcessBackgroundProfile(const float* data, size_t size) {
    if (!data || size == 0) throw std::invalid_argument("Invalid post-process background curve data");
    impl->config.setCustomPostProcessBackgroundProfile(data, size);
    if (impl->initialized) impl->backend->setPostProcessBackgroundProfile(data, size);
}

void Processor::savePostProcessBackgroundProfileToFile(const std::string& filepath) const {
    std::vector<float> curve;
    if (impl->initialized) {
        curve = impl->backend->getPostProcessBackgroundProfile();
    } else if (const float* d = impl->config.getCustomPostProcessBackgroundProfile()) {
        curve.assign(d, d + impl->config.getCustomPostProcessBackgroundProfileSize());
    }
    if (curve.empty()) throw std::runtime_error("No post-process background curve to save");
    std::ofstream f(filepath);
    if (!f.is_open()) throw std::runtime_error("Failed to open file for writing: " + filepath);
    f << "Sample Number;Sample Value\n";
    for (size_t i = 0; i < curve.size(); ++i) f << i << ";" << curve[i] << "\n";
    f.close();
    if (!f.good()) throw std::runtime_error("Error writing to file: " + filepath);
}

void Processor::loadPostProcessBackgroundProfileFromFile(const std::string& filepath) {
    std::ifstream f(filepath);
    if (!f.is_open()) throw std::runtime_error("Failed to open file for reading: " + filepath);
    std::string line;
    if (!std::getline(f, line)) throw std::runtime_error("Empty file: " + filepath);
    std::vector<float> curve;
    int lineNumber = 1;
    while (std::getline(f, line)) {
        ++lineNumber;
        if (line.empty()) continue;
        const size_t sep = line.find(';');
        if (sep == std::string::npos)
            throw std::runtime_error("Invalid format at line " + std::to_string(lineNumber) + ": missing semicolon");
        const std::string value = line.substr(sep + 1);
        try {
            curve.push_back(std::stof(value));
        } catch (const std::exception&) {
            throw std::runtime_error("Invalid number at line " + std::to_string(lineNumber) + ": " + value);
        }
    }
    if (curve.empty()) throw std::runtime_error("No data found in file: " + filepath);
    setPostProcessBackgroundProfile(curve.data(), curve.size());
}

// ---------------------------------------------------------------- fixed-pattern noise

void Processor::requestFixedPatternNoiseDetermination() { impl->backend->requestFixedPatternNoiseDetermination(); }

void Processor::setFixedPatternNoiseBscanCount(int numberOfBscans) {
    if (numberOfBscans < 1) throw std::invalid_argument("numberOfBscans must be >= 1");
    impl->config.postProcessingParams.fixedPatternNoiseBscanCount = numberOfBscans;
    if (impl->initialized) impl->backend->updateConfig(impl->config);
}

void Processor::enableContinuousFixedPatternNoiseDetermination(bool enable) {
    impl->config.postProcessingParams.continuousFixedPatternNoiseDetermination = enable;
    if (impl->initialized) impl->backend->updateConfig(impl->config);
}

void Processor::setFixedPatternNoiseProfile(const float* data, size_t complexPairs) {
    if (!data || complexPairs == 0) throw std::invalid_argument("Invalid fixed pattern noise profile");
    impl->config.setCustomFixedPatternNoiseProfile(data, complexPairs);
    if (impl->initialized) impl->backend->setFixedPatternNoiseProfile(data, complexPairs);
}

const float* Processor::getFixedPatternNoiseProfile() const {
    if (!impl->initialized) return nullptr;
    const std::vector<float>& v = impl->backend->getFixedPatternNoiseProfile();
    return v.empty() ? nullptr : v.data();
}

size_t Processor::getFixedPatternNoiseProfileSize() const {
    if (!impl->initialized) return 0;
    return impl->backend->getFixedPatternNoiseProfile().size() / 2;
}

bool Processor::hasFixedPatternNoiseProfile() const {
    if (!impl->initialized) return false;
    return !impl->backend->getFixedPatternNoiseProfile().empty();
}

void Processor::saveFixedPatternNoiseProfileToFile(const std::string& filepath) const {
    std::vector<float> prof;
    if (impl->initialized) {
        prof = impl->backend->getFixedPatternNoiseProfile();
    } else if (const float* d = impl->config.getCustomFixedPatternNoiseProfile()) {
        prof.assign(d, d + 2 * impl->config.getCustomFixedPatternNoiseProfileSize());
    }
    if (prof.empty()) throw std::runtime_error("No fixed pattern noise profile to save");
    std::ofstream f(filepath);
    if (!f.is_open()) throw std::runtime_error("Failed to open file for writing: " + filepath);
    f << "Sample Number;Real;Imag\n";
    for (size_t i = 0; i < prof.size() / 2; ++i) f << i << ";" << prof[2 * i] << ";" << prof[2 * i + 1] << "\n";
    f.close();
    if (!f.good()) throw std::runtime_error("Error writing to file: " + filepath);
}

void Processor::loadFixedPatternNoiseProfileFromFile(const std::string& filepath) {
    std::ifstream f(filepath);
    if (!f.is_open()) throw std::runtime_error("Failed to open file for reading: " + filepath);
    std::string line;
    if (!std::getline(f, line)) throw std::runtime_error("Empty file: " + filepath);
    std::vector<float> prof;
    int lineNumber = 1;
    while (std::getline(f, line)) {
        ++lineNumber;
        if (line.empty()) continue;
        const size_t p1 = line.find(';');
        const size_t p2 = p1 == std::string::npos ? std::string::npos : line.find(';', p1 + 1);
        if (p2 == std::string::npos) throw std::runtime_error("Invalid format at line " + std::to_string(lineNumber));
        try {
            prof.push_back(std::stof(line.substr(p1 + 1, p2 - p1 - 1)));
            prof.push_back(std::stof(line.substr(p2 + 1)));
        } catch (const std::exception&) {
            throw std::runtime_error("Invalid number at line " + std::to_string(lineNumber));
        }
    }
    if (prof.empty()) throw std::runtime_error("No data found in file: " + filepath);
    if (!impl->initialized)
        throw std::runtime_error("Processor must be initialized before loading fixed pattern noise profile");
    impl->backend->setFixedPatternNoiseProfile(prof.data(), prof.size() / 2);
}

// ---------------------------------------------------------------- individual operations

std::vector<float> Processor::convertInput(const void* input, IOBuffer::DataType inputType, int bitDepth,
                                           int samples, bool applyBitshift) {
    impl->ensureInitialized();
    return impl->backend->convertInput(input, inputType, bitDepth, samples, applyBitshift);
}
std::vector<float> Processor::rollingAverageBackgroundRemoval(const float* input, int windowSize, int lineWidth,
                                                              int numLines) {
    impl->ensureInitialized();
    return impl->backend->rollingAverageBackgroundRemoval(input, windowSize, lineWidth, numLines);
}
std::vector<float> Processor::kLinearization(const float* input, const float* resampleCurve,
                                             InterpolationMethod method, int lineWidth, int samples) {
    impl->ensureInitialized();
    return impl->backend->kLinearization(input, resampleCurve, method, lineWidth, samples);
}
std::vector<float> Processor::windowing(const float* input, const float* windowCurve, int lineWidth, int samples) {
    impl->ensureInitialized();
    return impl->backend->windowing(input, windowCurve, lineWidth, samples);
}
std::vector<float> Processor::dispersionCompensation(const float* input, const float* phaseComplex, int lineWidth,
                                                     int samples) {
    impl->ensureInitialized();
    return impl->backend->dispersionCompensation(input, phaseComplex, lineWidth, samples);
}
std::vector<float> Processor::kLinearizationAndWindowing(const float* input, const float* resampleCurve,
                                                         const float* windowCurve, InterpolationMethod method,
                                                         int lineWidth, int samples) {
    impl->ensureInitialized();
    return impl->backend->kLinearizationAndWindowing(input, resampleCurve, windowCurve, method, lineWidth, samples);
}
std::vector<float> Processor::kLinearizationAndWindowingAndDispersion(const float* input, const float* resampleCurve,
                                                                      const float* windowCurve,
                                                                      const float* phaseComplex,
                                                                      InterpolationMethod method, int lineWidth,
                                                                      int samples) {
    impl->ensureInitialized();
    return impl->backend->kLinearizationAndWindowingAndDispersion(input, resampleCurve, windowCurve, phaseComplex,
                                                                  method, lineWidth, samples);
}
std::vector<float> Processor::dispersionCompensationAndWindowing(const float* input, const float* phaseComplex,
                                                                 const float* windowCurve, int lineWidth,
                                                                 int samples) {
    impl->ensureInitialized();
    return impl->backend->dispersionCompensationAndWindowing(input, phaseComplex, windowCurve, lineWidth, samples);
}
std::vector<float> Processor::fft(const float* input, int lineWidth, int samples) {
    impl->ensureInitialized();
    return impl->backend->fft(input, lineWidth, samples);
}
std::vector<float> Processor::ifft(const float* input, int lineWidth, int samples) {
    impl->ensureInitialized();
    return impl->backend->ifft(input, lineWidth, samples);
}
std::vector<float> Processor::getMinimumVarianceMean(const float* input, int width, int height, int segments) {
    impl->ensureInitialized();
    return impl->backend->getMinimumVarianceMean(input, width, height, segments);
}
std::vector<float> Processor::fixedPatternNoiseRemoval(const float* input, const float* meanALine, int lineWidth,
                                                       int numLines) {
    impl->ensureInitialized();
    return impl->backend->fixedPatternNoiseRemoval(input, meanALine, lineWidth, numLines);
}
std::vector<float> Processor::postProcessTruncate(const float* input, bool logScaling, float grayscaleMax,
                                                  float grayscaleMin, float addend, float multiplicator,
                                                  int lineWidth, int samples) {
    impl->ensureInitialized();
    return impl->backend->postProcessTruncate(input, logScaling, grayscaleMax, grayscaleMin, addend, multiplicator,
                                              lineWidth, samples);
}
std::vector<float> Processor::bscanFlip(const float* input, int lineWidth, int linesPerBscan, int numBscans) {
    impl->ensureInitialized();
    return impl->backend->bscanFlip(input, lineWidth, linesPerBscan, numBscans);
}
std::vector<float> Processor::sinusoidalScanCorrection(const float* input, const float* resampleCurve, int lineWidth,
                                                       int linesPerBscan, int numBscans) {
    impl->ensureInitialized();
    return impl->backend->sinusoidalScanCorrection(input, resampleCurve, lineWidth, linesPerBscan, numBscans);
}
std::vector<float> Processor::postProcessBackgroundSubtraction(const float* input, const float* backgroundLine,
                                                               float weight, float offset, int lineWidth,
                                                               int samples) {
    impl->ensureInitialized();
    return impl->backend->postProcessBackgroundSubtraction(input, backgroundLine, weight, offset, lineWidth,
                                                           samples);
}

}  // namespace ope
