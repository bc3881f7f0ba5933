// ProcessorConfiguration: custom-curve storage with pad/truncate to the
// signal length, and host generation of the resampling, dispersion-phase and
// window curves.  Semantics follow OCTproEngine's
// src/core/processorconfiguration.cpp (custom curves :134-432, generators
// :438-589, validate :609-617).  The generators evaluate in float in the same
// order as the reference (compiled with -ffp-contract=off), so generated
// curves are bit-identical to the reference's; tests compare them with
// curves written by the reference's own code.
#include <cmath>
#include <stdexcept>
#include <vector>

#include "processorconfiguration.h"

namespace ope {

namespace {

constexpr double kPi = 3.14159265358979323846;

// Zero-pad or truncate `original` to `target` values (processorconfiguration.cpp:268-326).
void fit_to(const std::vector<float>& original, size_t target, std::vector<float>& out) {
    if (original.empty()) {
        out.clear();
        return;
    }
    out = original;
    out.resize(target, 0.0f);
}

void check_curve(const float* data, size_t size, const char* what) {
    if (!data) throw std::invalid_argument(std::string("Custom ") + what + " data is null");
    if (size == 0) throw std::invalid_argument(std::string("Custom ") + what + " size is zero");
}

}  // namespace

struct ProcessorConfiguration::Impl {
    // as given by the user, and fitted to the current signal length
    std::vector<float> resampleOrig, windowOrig, dispersionOrig, backgroundOrig, fpnOrig;
    std::vector<float> resample, window, dispersion, background, fpn;
    // regenerated on every request
    mutable std::vector<float> genResample, genWindow, genDispersion;
};

ProcessorConfiguration::ProcessorConfiguration() : impl(new Impl()) {}

ProcessorConfiguration::~ProcessorConfiguration() { delete this->impl; }

ProcessorConfiguration::ProcessorConfiguration(const ProcessorConfiguration& o)
    : dataParams(o.dataParams), processingParams(o.processingParams), resamplingParams(o.resamplingParams),
      windowingParams(o.windowingParams), dispersionParams(o.dispersionParams),
      backgroundRemovalParams(o.backgroundRemovalParams), postProcessingParams(o.postProcessingParams),
      impl(new Impl(*o.impl)) {}

ProcessorConfiguration::ProcessorConfiguration(ProcessorConfiguration&& o) noexcept
    : dataParams(o.dataParams), processingParams(o.processingParams), resamplingParams(o.resamplingParams),
      windowingParams(o.windowingParams), dispersionParams(o.dispersionParams),
      backgroundRemovalParams(o.backgroundRemovalParams), postProcessingParams(o.postProcessingParams),
      impl(o.impl) {
    o.impl = new Impl();
}

ProcessorConfiguration& ProcessorConfiguration::operator=(const ProcessorConfiguration& o) {
    if (this != &o) {
        Impl* copy = new Impl(*o.impl);
        this->dataParams = o.dataParams;
        this->processingParams = o.processingParams;
        this->resamplingParams = o.resamplingParams;
        this->windowingParams = o.windowingParams;
        this->dispersionParams = o.dispersionParams;
        this->backgroundRemovalParams = o.backgroundRemovalParams;
        this->postProcessingParams = o.postProcessingParams;
        delete this->impl;
        this->impl = copy;
    }
    return *this;
}

ProcessorConfiguration& ProcessorConfiguration::operator=(ProcessorConfiguration&& o) noexcept {
    if (this != &o) {
        this->dataParams = o.dataParams;
        this->processingParams = o.processingParams;
        this->resamplingParams = o.resamplingParams;
        this->windowingParams = o.windowingParams;
        this->dispersionParams = o.dispersionParams;
        this->backgroundRemovalParams = o.backgroundRemovalParams;
        this->postProcessingParams = o.postProcessingParams;
        std::swap(this->impl, o.impl);
    }
    return *this;
}

// ---------------------------------------------------------------- custom curves

void ProcessorConfiguration::setCustomResamplingCurve(const float* data, size_t size) {
    check_curve(data, size, "resampling curve");
    impl->resampleOrig.assign(data, data + size);
    adjustCustomResamplingCurve();
}

void ProcessorConfiguration::setCustomWindowCurve(const float* data, size_t size) {
    check_curve(data, size, "window curve");
    impl->windowOrig.assign(data, data + size);
    adjustCustomWindowCurve();
}

void ProcessorConfiguration::setCustomDispersionCurve(const float* data, size_t size) {
    check_curve(data, size, "dispersion curve");
    impl->dispersionOrig.assign(data, data + size);  // phase values
    adjustCustomDispersionCurve();
}

void ProcessorConfiguration::setCustomPostProcessBackgroundProfile(const float* data, size_t size) {
    check_curve(data, size, "post-process background curve");
    impl->backgroundOrig.assign(data, data + size);
    adjustCustomPostProcessBackgroundProfile();
}

void ProcessorConfiguration::setCustomFixedPatternNoiseProfile(const float* data, size_t complexPairs) {
    check_curve(data, complexPairs, "fixed-pattern noise profile");
    impl->fpnOrig.assign(data, data + 2 * complexPairs);
    adjustCustomFixedPatternNoiseProfile();
}

static const float* data_or_null(const std::vector<float>& v) { return v.empty() ? nullptr : v.data(); }

const float* ProcessorConfiguration::getCustomResamplingCurve() const { return data_or_null(impl->resample); }
const float* ProcessorConfiguration::getCustomWindowCurve() const { return data_or_null(impl->window); }
const float* ProcessorConfiguration::getCustomDispersionCurve() const { return data_or_null(impl->dispersion); }
const float* ProcessorConfiguration::getCustomPostProcessBackgroundProfile() const {
    return data_or_null(impl->background);
}
const float* ProcessorConfiguration::getCustomFixedPatternNoiseProfile() const { return data_or_null(impl->fpn); }

size_t ProcessorConfiguration::getCustomResamplingCurveSize() const { return impl->resample.size(); }
size_t ProcessorConfiguration::getCustomWindowCurveSize() const { return impl->window.size(); }
size_t ProcessorConfiguration::getCustomDispersionCurveSize() const { return impl->dispersion.size(); }
size_t ProcessorConfiguration::getCustomPostProcessBackgroundProfileSize() const { return impl->background.size(); }
size_t ProcessorConfiguration::getCustomFixedPatternNoiseProfileSize() const { return impl->fpn.size() / 2; }

bool ProcessorConfiguration::hasCustomResamplingCurve() const { return !impl->resampleOrig.empty(); }
bool ProcessorConfiguration::hasCustomWindowCurve() const { return !impl->windowOrig.empty(); }
bool ProcessorConfiguration::hasCustomDispersionCurve() const { return !impl->dispersionOrig.empty(); }
bool ProcessorConfiguration::hasCustomPostProcessBackgroundProfile() const { return !impl->backgroundOrig.empty(); }
bool ProcessorConfiguration::hasCustomFixedPatternNoiseProfile() const { return !impl->fpnOrig.empty(); }

void ProcessorConfiguration::adjustCustomResamplingCurve() {
    fit_to(impl->resampleOrig, static_cast<size_t>(dataParams.signalLength), impl->resample);
}
void ProcessorConfiguration::adjustCustomWindowCurve() {
    fit_to(impl->windowOrig, static_cast<size_t>(dataParams.signalLength), impl->window);
}
void ProcessorConfiguration::adjustCustomDispersionCurve() {
    fit_to(impl->dispersionOrig, static_cast<size_t>(dataParams.signalLength), impl->dispersion);
}
void ProcessorConfiguration::adjustCustomPostProcessBackgroundProfile() {
    fit_to(impl->backgroundOrig, static_cast<size_t>(dataParams.signalLength / 2), impl->background);
}
void ProcessorConfiguration::adjustCustomFixedPatternNoiseProfile() {
    fit_to(impl->fpnOrig, static_cast<size_t>(dataParams.signalLength / 2) * 2, impl->fpn);
}

void ProcessorConfiguration::adjustAllCustomCurves() {
    adjustCustomResamplingCurve();
    adjustCustomWindowCurve();
    adjustCustomDispersionCurve();
    adjustCustomPostProcessBackgroundProfile();
    adjustCustomFixedPatternNoiseProfile();
}

// ---------------------------------------------------------------- generated curves

const float* ProcessorConfiguration::getGeneratedResamplingCurve() const {
    generateResamplingCurve();
    return impl->genResample.data();
}
const float* ProcessorConfiguration::getGeneratedWindowCurve() const {
    generateWindowCurve();
    return impl->genWindow.data();
}
const float* ProcessorConfiguration::getGeneratedDispersionCurve() const {
    generateDispersionCurve();
    return impl->genDispersion.data();
}
size_t ProcessorConfiguration::getGeneratedResamplingCurveSize() const { return dataParams.signalLength; }
size_t ProcessorConfiguration::getGeneratedWindowCurveSize() const { return dataParams.signalLength; }
size_t ProcessorConfiguration::getGeneratedDispersionCurveSize() const { return dataParams.signalLength; }

// r[j] = clamp(c0 + j (c1/(N-1) + j (c2/(N-1)^2 + j c3/(N-1)^3)), 7, N-9):
// polynomial normalised to the A-scan length, Horner in float, clamped so the
// widest (Lanczos, 16-tap) kernel stays inside the A-scan
// (processorconfiguration.cpp:438-459).
void ProcessorConfiguration::generateResamplingCurve() const {
    const size_t n = static_cast<size_t>(dataParams.signalLength);
    impl->genResample.resize(n);
    const float* c = resamplingParams.coefficients;
    const float nm1 = static_cast<float>(n) - 1.0f;
    const float k1 = c[1] / nm1;
    const float k2 = c[2] / (nm1 * nm1);
    const float k3 = c[3] / (nm1 * nm1 * nm1);
    const float lo = 7.0f;
    const float hi = static_cast<float>(n - 9);
    for (size_t i = 0; i < n; ++i) {
        const float x = static_cast<float>(i);
        const float r = c[0] + x * (k1 + x * (k2 + x * k3));
        impl->genResample[i] = this->clamp(r, lo, hi);
    }
}

// phi[j] = d0 + j (d1/(N-1) + j (d2/(N-1)^2 + j d3/(N-1)^3)), float Horner;
// `factor` is stored but, as in the reference, not applied
// (processorconfiguration.cpp:461-482).
void ProcessorConfiguration::generateDispersionCurve() const {
    const size_t n = static_cast<size_t>(dataParams.signalLength);
    impl->genDispersion.resize(n);
    const float* d = dispersionParams.coefficients;
    const float den = static_cast<float>(n - 1);
    const float k1 = d[1] / den;
    const float k2 = d[2] / (den * den);
    const float k3 = d[3] / (den * den * den);
    for (size_t i = 0; i < n; ++i) {
        const float x = static_cast<float>(i);
        impl->genDispersion[i] = d[0] + x * (k1 + x * (k2 + x * k3));
    }
}

// Window of `fill * N` samples centred at `center * N`; outside its support
// (normalised position > 0.999 or < 0.0001) the window is 0
// (processorconfiguration.cpp:484-589).
void ProcessorConfiguration::generateWindowCurve() const {
    const unsigned n = static_cast<unsigned>(dataParams.signalLength);
    std::vector<float>& w = impl->genWindow;
    w.assign(n, 0.0f);
    float center = windowingParams.windowCenterPosition;
    const float fill = windowingParams.windowFillFactor;
    if (center > 1.0f) center = 1.0f;
    if (center < 0.0f) center = 0.0f;
    const unsigned width = static_cast<unsigned>(fill * static_cast<float>(n));
    const unsigned centerPos = static_cast<unsigned>(center * static_cast<float>(n));
    const int minPos = static_cast<int>(centerPos - width / 2);  // unsigned wrap, as the reference
    const float widthM1 = static_cast<float>(width) - 1.0f;

    auto norm = [&](unsigned i) { return static_cast<float>(static_cast<int>(i) - minPos) / widthM1; };
    auto outside = [](float u) { return u > 0.999f || u < 0.0001f; };

    switch (windowingParams.windowType) {
        case WindowType::HANN:
            for (unsigned i = 0; i < n; ++i) {
                const float u = norm(i);
                if (!outside(u)) w[i] = static_cast<float>(0.5 * (1.0 - std::cos(2.0 * kPi * static_cast<double>(u))));
            }
            break;
        case WindowType::GAUSS:
            for (unsigned i = 0; i < n; ++i) {
                const int xi = static_cast<int>(i) - static_cast<int>(centerPos);
                const float u = (static_cast<float>(xi) / (static_cast<float>(n) - 1.0f)) / fill;
                w[i] = std::exp(-10.0f * (u * u));
            }
            break;
        case WindowType::SINE:
            for (unsigned i = 0; i < n; ++i) {
                const float u = norm(i);
                if (!outside(u)) w[i] = static_cast<float>(std::sin(kPi * static_cast<double>(u)));
            }
            break;
        case WindowType::LANCZOS:
            for (unsigned i = 0; i < n; ++i) {
                const float u = norm(i);
                if (outside(u)) continue;
                const float arg = 2.0f * u - 1.0f;
                w[i] = (arg == 0.0f) ? 1.0f
                                     : static_cast<float>(std::sin(kPi * static_cast<double>(arg)) /
                                                          (kPi * static_cast<double>(arg)));
            }
            break;
        case WindowType::RECTANGULAR:
            for (unsigned i = 0; i < n; ++i)
                if (!outside(norm(i))) w[i] = 1.0f;
            break;
        case WindowType::FLAT_TOP: {
            const float a0 = 0.215578948f, a1 = 0.416631580f, a2 = 0.277263158f, a3 = 0.083578947f,
                        a4 = 0.006947368f;
            for (unsigned i = 0; i < n; ++i) {
                const float u = norm(i);
                if (outside(u)) continue;
                const double ud = static_cast<double>(u);
                w[i] = a0 - a1 * static_cast<float>(std::cos(2.0 * kPi * ud)) +
                       a2 * static_cast<float>(std::cos(4.0 * kPi * ud)) -
                       a3 * static_cast<float>(std::cos(6.0 * kPi * ud)) +
                       a4 * static_cast<float>(std::cos(8.0 * kPi * ud));
            }
            break;
        }
    }
}

// ---------------------------------------------------------------- file I/O, validation

// The reference does not implement a configuration file format yet
// (processorconfiguration.cpp:595-603 return false); neither does this one.
bool ProcessorConfiguration::saveToFile(const std::string&) const { return false; }
bool ProcessorConfiguration::loadFromFile(const std::string&) { return false; }

bool ProcessorConfiguration::validate() const {
    return dataParams.signalLength > 0 && dataParams.samplesPerBuffer > 0 && dataParams.ascansPerBscan > 0 &&
           dataParams.bscansPerBuffer > 0;
}

}  // namespace ope
