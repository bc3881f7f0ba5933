// Python module `octproengine`: the same names, arguments and exception types
// as OCTproEngine's python/octproengine_bindings.cpp (module :309-834), over
// this library's ope::Processor (Backend.CUDA = the HIP backend).
//
// Buffers are zero-copy numpy views of IOBuffer memory (pinned host memory
// owned by the backend).  The GIL is released around calls that block or
// enqueue GPU work and re-acquired on the callback threads.
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

#include "iobuffer.h"
#include "processor.h"
#include "processorconfiguration.h"
#include "types.h"
#include "version.h"

namespace py = pybind11;

namespace {

struct InitializationError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ConfigurationError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct BufferError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ProcessingError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct BackendError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

py::dtype numpy_dtype(ope::DataType t, size_t* item) {
    switch (t) {
        case ope::DataType::UINT8: *item = 1; return py::dtype::of<uint8_t>();
        case ope::DataType::UINT16: *item = 2; return py::dtype::of<uint16_t>();
        case ope::DataType::UINT32: *item = 4; return py::dtype::of<uint32_t>();
        case ope::DataType::FLOAT32: *item = 4; return py::dtype::of<float>();
        case ope::DataType::FLOAT64: *item = 8; return py::dtype::of<double>();
        default: throw BufferError("Unsupported IOBuffer data type");
    }
}

// numpy view of an IOBuffer's memory; the array's base is the IOBuffer
py::array buffer_view(ope::IOBuffer& b) {
    size_t item = 1;
    py::dtype dt = numpy_dtype(b.getDataType(), &item);
    const size_t n = b.getSizeInBytes() / item;
    return py::array(dt, {n}, {item}, b.getDataPointer(), py::cast(&b, py::return_value_policy::reference));
}

std::vector<float> as_floats(const py::array_t<float, py::array::c_style | py::array::forcecast>& a) {
    py::buffer_info info = a.request();
    const float* p = static_cast<const float*>(info.ptr);
    return std::vector<float>(p, p + info.size);
}

py::array_t<float> to_numpy(const float* d, size_t n) {
    py::array_t<float> out(n);
    if (n) std::memcpy(out.mutable_data(), d, n * sizeof(float));
    return out;
}

class ProcessorWrapper {
public:
    explicit ProcessorWrapper(ope::Backend backend) : processor(backend) {}

    // Stop callback threads and GPU work with the GIL released (a callback
    // thread may be waiting for it) before the Python callables go away.
    ~ProcessorWrapper() {
        py::gil_scoped_release release;
        processor.clearOutputCallbacks();
        processor.cleanup();
    }

    void set_callback(py::function cb, py::object error_cb) {
        callback = cb;
        if (!error_cb.is_none()) error_callback = error_cb.cast<py::function>();
        processor.setOutputCallback([this](const ope::IOBuffer& out) {
            py::gil_scoped_acquire gil;
            try {
                callback(buffer_view(const_cast<ope::IOBuffer&>(out)));
            } catch (const std::exception& e) {
                if (error_callback) {
                    try {
                        error_callback(py::str(e.what()));
                    } catch (...) {
                        py::print("Error in error_callback:", py::str(e.what()));
                    }
                } else {
                    py::print("Error in callback:", py::str(e.what()));
                }
            }
        });
    }

    ope::Processor::CallbackId add_output_callback(py::function cb) {
        // The C++ side copies the std::function freely (without the GIL), so
        // the Python callable is shared through a pointer whose last owner
        // releases it with the GIL held.
        std::shared_ptr<py::function> fn(new py::function(cb), [](py::function* f) {
            py::gil_scoped_acquire gil;
            delete f;
        });
        auto wrapped = [fn](const ope::IOBuffer& out) {
            py::gil_scoped_acquire gil;
            try {
                (*fn)(buffer_view(const_cast<ope::IOBuffer&>(out)));
            } catch (const py::error_already_set& e) {
                py::print("Python error in callback:", e.what());
            } catch (const std::exception& e) {
                py::print("C++ error in callback:", e.what());
            }
        };
        ope::Processor::CallbackId id;
        {
            py::gil_scoped_release release;
            id = processor.addOutputCallback(wrapped);
        }
        std::lock_guard<std::mutex> lk(mu);
        kept[id] = cb;
        return id;
    }

    bool remove_output_callback(ope::Processor::CallbackId id) {
        bool removed;
        {
            py::gil_scoped_release release;
            removed = processor.removeOutputCallback(id);
        }
        if (removed) {
            std::lock_guard<std::mutex> lk(mu);
            kept.erase(id);
        }
        return removed;
    }

    void clear_output_callbacks() {
        {
            py::gil_scoped_release release;
            processor.clearOutputCallbacks();
        }
        std::lock_guard<std::mutex> lk(mu);
        kept.clear();
    }

    void process(py::array buffer) {
        if (!callback && processor.getCallbackCount() == 0)
            throw ProcessingError("Callback must be set before calling process(). Use set_callback() first.");
        py::object base = buffer.attr("base");
        if (base.is_none())
            throw BufferError("Buffer array is not a view of an IOBuffer. Use get_next_available_buffer() first.");
        ope::IOBuffer* b = base.cast<ope::IOBuffer*>();
        try {
            py::gil_scoped_release release;
            processor.process(*b);
        } catch (const std::exception& e) {
            throw ProcessingError(std::string("Processing failed: ") + e.what());
        }
    }

    py::array get_next_available_buffer() {
        ope::IOBuffer* b;
        {
            py::gil_scoped_release release;
            b = &processor.getNextAvailableInputBuffer();
        }
        return buffer_view(*b);
    }

    void initialize() {
        try {
            py::gil_scoped_release release;
            processor.initialize();
        } catch (const std::exception& e) {
            throw InitializationError(std::string("Initialization failed: ") + e.what());
        }
    }

    void stop() {
        py::gil_scoped_release release;
        processor.cleanup();
    }

    ope::Processor processor;
    py::function callback;
    py::function error_callback;
    std::map<ope::Processor::CallbackId, py::function> kept;  // keeps callables alive
    std::mutex mu;
};

}  // namespace

PYBIND11_MODULE(octproengine, m) {
    m.doc() = "OCTproEngine API on AMD Instinct MI355X (HIP backend)";

    py::register_exception<InitializationError>(m, "InitializationError");
    py::register_exception<ConfigurationError>(m, "ConfigurationError");
    py::register_exception<BufferError>(m, "BufferError");
    py::register_exception<ProcessingError>(m, "ProcessingError");
    py::register_exception<BackendError>(m, "BackendError");

    py::class_<ope::IOBuffer>(m, "IOBuffer", py::module_local())
        .def("get_size", &ope::IOBuffer::getSizeInBytes)
        .def("get_data_type", &ope::IOBuffer::getDataType);

    py::enum_<ope::Backend>(m, "Backend")
        .value("CUDA", ope::Backend::CUDA, "GPU backend (HIP on AMD Instinct MI355X)")
        .value("CPU", ope::Backend::CPU, "CPU backend (not part of this library)")
        .export_values();

    py::enum_<ope::DataType>(m, "DataType")
        .value("UINT8", ope::DataType::UINT8)
        .value("UINT16", ope::DataType::UINT16)
        .value("UINT32", ope::DataType::UINT32)
        .value("UINT64", ope::DataType::UINT64)
        .value("INT8", ope::DataType::INT8)
        .value("INT16", ope::DataType::INT16)
        .value("INT32", ope::DataType::INT32)
        .value("INT64", ope::DataType::INT64)
        .value("FLOAT32", ope::DataType::FLOAT32)
        .value("FLOAT64", ope::DataType::FLOAT64)
        .value("COMPLEX_FLOAT32", ope::DataType::COMPLEX_FLOAT32)
        .value("COMPLEX_FLOAT64", ope::DataType::COMPLEX_FLOAT64)
        .export_values();

    py::enum_<ope::InterpolationMethod>(m, "InterpolationMethod")
        .value("LINEAR", ope::InterpolationMethod::LINEAR)
        .value("CUBIC", ope::InterpolationMethod::CUBIC)
        .value("LANCZOS", ope::InterpolationMethod::LANCZOS)
        .export_values();

    py::enum_<ope::WindowType>(m, "WindowType")
        .value("HANN", ope::WindowType::HANN)
        .value("GAUSS", ope::WindowType::GAUSS)
        .value("SINE", ope::WindowType::SINE)
        .value("LANCZOS", ope::WindowType::LANCZOS)
        .value("RECTANGULAR", ope::WindowType::RECTANGULAR)
        .value("FLAT_TOP", ope::WindowType::FLAT_TOP)
        .export_values();

    using PC = ope::ProcessorConfiguration;
    py::class_<PC::DataParameters>(m, "DataParameters")
        .def(py::init<>())
        .def_readwrite("signal_length", &PC::DataParameters::signalLength)
        .def_readwrite("samples_per_buffer", &PC::DataParameters::samplesPerBuffer)
        .def_readwrite("ascans_per_bscan", &PC::DataParameters::ascansPerBscan)
        .def_readwrite("bscans_per_buffer", &PC::DataParameters::bscansPerBuffer)
        .def_readwrite("input_data_type", &PC::DataParameters::inputDataType)
        .def_readwrite("bitshift", &PC::DataParameters::bitshift)
        .def("get_bit_depth", &PC::DataParameters::getBitDepth)
        .def("get_bytes_per_sample", &PC::DataParameters::getBytesPerSample);

    py::class_<PC::ProcessingParameters>(m, "ProcessingParameters")
        .def(py::init<>())
        .def_readwrite("n_streams", &PC::ProcessingParameters::nStreams)
        .def_readwrite("n_buffers", &PC::ProcessingParameters::nBuffers)
        .def_readwrite("grid_size", &PC::ProcessingParameters::gridSize)
        .def_readwrite("block_size", &PC::ProcessingParameters::blockSize);

    py::class_<PC::ResamplingParameters>(m, "ResamplingParameters")
        .def(py::init<>())
        .def_readwrite("enabled", &PC::ResamplingParameters::enabled)
        .def_readwrite("interpolation_method", &PC::ResamplingParameters::interpolationMethod)
        .def_readwrite("use_coefficients", &PC::ResamplingParameters::useCoefficients)
        .def_property(
            "coefficients",
            [](const PC::ResamplingParameters& r) {
                return std::array<float, 4>{r.coefficients[0], r.coefficients[1], r.coefficients[2],
                                            r.coefficients[3]};
            },
            [](PC::ResamplingParameters& r, const std::array<float, 4>& c) {
                for (int i = 0; i < 4; ++i) r.coefficients[i] = c[i];
            })
        .def_readwrite("use_custom_curve", &PC::ResamplingParameters::useCustomCurve);

    py::class_<PC::WindowingParameters>(m, "WindowingParameters")
        .def(py::init<>())
        .def_readwrite("enabled", &PC::WindowingParameters::enabled)
        .def_readwrite("window_type", &PC::WindowingParameters::windowType)
        .def_readwrite("window_center_position", &PC::WindowingParameters::windowCenterPosition)
        .def_readwrite("window_fill_factor", &PC::WindowingParameters::windowFillFactor)
        .def_readwrite("use_custom_curve", &PC::WindowingParameters::useCustomCurve);

    py::class_<PC::DispersionCompensationParameters>(m, "DispersionCompensationParameters")
        .def(py::init<>())
        .def_readwrite("enabled", &PC::DispersionCompensationParameters::enabled)
        .def_readwrite("use_coefficients", &PC::DispersionCompensationParameters::useCoefficients)
        .def_property(
            "coefficients",
            [](const PC::DispersionCompensationParameters& d) {
                return std::array<float, 4>{d.coefficients[0], d.coefficients[1], d.coefficients[2],
                                            d.coefficients[3]};
            },
            [](PC::DispersionCompensationParameters& d, const std::array<float, 4>& c) {
                for (int i = 0; i < 4; ++i) d.coefficients[i] = c[i];
            })
        .def_readwrite("factor", &PC::DispersionCompensationParameters::factor)
        .def_readwrite("use_custom_curve", &PC::DispersionCompensationParameters::useCustomCurve);

    py::class_<PC::BackgroundRemovalParameters>(m, "BackgroundRemovalParameters")
        .def(py::init<>())
        .def_readwrite("enabled", &PC::BackgroundRemovalParameters::enabled)
        .def_readwrite("rolling_average_window_size", &PC::BackgroundRemovalParameters::rollingAverageWindowSize);

    py::class_<PC::PostProcessingParameters>(m, "PostProcessingParameters")
        .def(py::init<>())
        .def_readwrite("background_removal", &PC::PostProcessingParameters::backgroundRemoval)
        .def_readwrite("background_weight", &PC::PostProcessingParameters::backgroundWeight)
        .def_readwrite("background_offset", &PC::PostProcessingParameters::backgroundOffset)
        .def_readwrite("log_scaling", &PC::PostProcessingParameters::logScaling)
        .def_readwrite("grayscale_max", &PC::PostProcessingParameters::grayscaleMax)
        .def_readwrite("grayscale_min", &PC::PostProcessingParameters::grayscaleMin)
        .def_readwrite("addend", &PC::PostProcessingParameters::addend)
        .def_readwrite("multiplicator", &PC::PostProcessingParameters::multiplicator)
        .def_readwrite("bscan_flip", &PC::PostProcessingParameters::bscanFlip)
        .def_readwrite("sinusoidal_scan_correction", &PC::PostProcessingParameters::sinusoidalScanCorrection)
        .def_readwrite("fixed_pattern_noise_removal", &PC::PostProcessingParameters::fixedPatternNoiseRemoval)
        .def_readwrite("fixed_pattern_noise_bscan_count", &PC::PostProcessingParameters::fixedPatternNoiseBscanCount)
        .def_readwrite("continuous_fixed_pattern_noise_determination",
                       &PC::PostProcessingParameters::continuousFixedPatternNoiseDetermination);

    auto curve_getter = [](const float* (PC::*get)() const, size_t (PC::*size)() const) {
        return [get, size](const PC& c) -> py::array_t<float> {
            const float* d = (c.*get)();
            return d ? to_numpy(d, (c.*size)()) : py::array_t<float>(0);
        };
    };

    py::class_<PC>(m, "ProcessorConfiguration")
        .def(py::init<>())
        .def_readwrite("data", &PC::dataParams)
        .def_readwrite("processing", &PC::processingParams)
        .def_readwrite("resampling", &PC::resamplingParams)
        .def_readwrite("windowing", &PC::windowingParams)
        .def_readwrite("dispersion", &PC::dispersionParams)
        .def_readwrite("background_removal", &PC::backgroundRemovalParams)
        .def_readwrite("post_processing", &PC::postProcessingParams)
        .def("set_custom_resampling_curve",
             [](PC& c, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 auto v = as_floats(a);
                 c.setCustomResamplingCurve(v.data(), v.size());
             })
        .def("set_custom_window_curve",
             [](PC& c, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 auto v = as_floats(a);
                 c.setCustomWindowCurve(v.data(), v.size());
             })
        .def("set_custom_dispersion_curve",
             [](PC& c, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 auto v = as_floats(a);
                 c.setCustomDispersionCurve(v.data(), v.size());
             })
        .def("get_custom_resampling_curve", curve_getter(&PC::getCustomResamplingCurve, &PC::getCustomResamplingCurveSize))
        .def("get_custom_window_curve", curve_getter(&PC::getCustomWindowCurve, &PC::getCustomWindowCurveSize))
        .def("get_custom_dispersion_curve", curve_getter(&PC::getCustomDispersionCurve, &PC::getCustomDispersionCurveSize))
        // generated curves (an addition of this library, used by its tests)
        .def("get_generated_resampling_curve",
             curve_getter(&PC::getGeneratedResamplingCurve, &PC::getGeneratedResamplingCurveSize))
        .def("get_generated_window_curve", curve_getter(&PC::getGeneratedWindowCurve, &PC::getGeneratedWindowCurveSize))
        .def("get_generated_dispersion_curve",
             curve_getter(&PC::getGeneratedDispersionCurve, &PC::getGeneratedDispersionCurveSize))
        .def("has_custom_resampling_curve", &PC::hasCustomResamplingCurve)
        .def("has_custom_window_curve", &PC::hasCustomWindowCurve)
        .def("has_custom_dispersion_curve", &PC::hasCustomDispersionCurve)
        .def("adjust_all_custom_curves", &PC::adjustAllCustomCurves)
        .def("save_to_file", &PC::saveToFile)
        .def("load_from_file", &PC::loadFromFile)
        .def("validate", &PC::validate);

    using PW = ProcessorWrapper;
    auto f32 = [](py::array_t<float, py::array::c_style | py::array::forcecast> a) { return as_floats(a); };

    py::class_<PW>(m, "Processor")
        .def(py::init([](ope::Backend b) {
                 try {
                     return new PW(b);
                 } catch (const std::exception& e) {
                     throw BackendError(e.what());
                 }
             }),
             py::arg("backend") = ope::Backend::CUDA)
        .def("initialize", &PW::initialize)
        .def("stop", &PW::stop)
        .def("load_config",
             [](PW& s, const std::string& path) {
                 try {
                     s.processor.loadConfigurationFromFile(path);
                 } catch (const std::exception& e) {
                     throw ConfigurationError(std::string("Failed to load config: ") + e.what());
                 }
             },
             py::arg("filepath"))
        .def("save_config",
             [](PW& s, const std::string& path) {
                 try {
                     s.processor.saveConfigurationToFile(path);
                 } catch (const std::exception& e) {
                     throw ConfigurationError(std::string("Failed to save config: ") + e.what());
                 }
             },
             py::arg("filepath"))
        .def_property_readonly(
            "config", [](PW& s) -> PC& { return const_cast<PC&>(s.processor.getConfig()); },
            py::return_value_policy::reference_internal)
        .def("set_config", [](PW& s, const PC& c) { s.processor.setConfig(c); }, py::arg("config"))
        .def("set_input_parameters",
             [](PW& s, int n, int a, int b, ope::DataType t) {
                 py::gil_scoped_release release;
                 s.processor.setInputParameters(n, a, b, t);
             },
             py::arg("signal_length"), py::arg("ascans_per_bscan"), py::arg("bscans_per_buffer"),
             py::arg("data_type"))
        .def("set_backend",
             [](PW& s, ope::Backend b) {
                 try {
                     py::gil_scoped_release release;
                     s.processor.setBackend(b);
                 } catch (const std::exception& e) {
                     throw BackendError(e.what());
                 }
             },
             py::arg("backend"))
        .def("get_backend", [](const PW& s) { return s.processor.getBackend(); })
        .def("set_callback", &PW::set_callback, py::arg("callback"), py::arg("error_callback") = py::none())
        .def("add_output_callback", &PW::add_output_callback, py::arg("callback"))
        .def("remove_output_callback", &PW::remove_output_callback, py::arg("callback_id"))
        .def("clear_output_callbacks", &PW::clear_output_callbacks)
        .def("get_callback_count", [](const PW& s) { return s.processor.getCallbackCount(); })
        .def("process", &PW::process, py::arg("buffer"))
        .def("get_next_available_buffer", &PW::get_next_available_buffer)
        .def("set_resampling_coefficients",
             [](PW& s, const std::array<float, 4>& c) { s.processor.setResamplingCoefficients(c.data()); },
             py::arg("coefficients"))
        .def("set_custom_resampling_curve",
             [f32](PW& s, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 auto v = f32(a);
                 s.processor.setCustomResamplingCurve(v.data(), v.size());
             },
             py::arg("curve"))
        .def("use_custom_resampling_curve", [](PW& s, bool u) { s.processor.useCustomResamplingCurve(u); },
             py::arg("use_custom"))
        .def("enable_resampling", [](PW& s, bool e) { s.processor.enableResampling(e); }, py::arg("enable"))
        .def("set_interpolation_method", [](PW& s, ope::InterpolationMethod m) { s.processor.setInterpolationMethod(m); },
             py::arg("method"))
        .def("set_dispersion_coefficients",
             [](PW& s, const std::array<float, 4>& c, float factor) { s.processor.setDispersionCoefficients(c.data(), factor); },
             py::arg("coefficients"), py::arg("factor") = 1.0f)
        .def("set_custom_dispersion_curve",
             [f32](PW& s, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 auto v = f32(a);
                 s.processor.setCustomDispersionCurve(v.data(), v.size());
             },
             py::arg("curve"))
        .def("use_custom_dispersion_curve", [](PW& s, bool u) { s.processor.useCustomDispersionCurve(u); },
             py::arg("use_custom"))
        .def("enable_dispersion_compensation", [](PW& s, bool e) { s.processor.enableDispersionCompensation(e); },
             py::arg("enable"))
        .def("set_window_parameters",
             [](PW& s, ope::WindowType t, float c, float f) { s.processor.setWindowParameters(t, c, f); },
             py::arg("window_type"), py::arg("center_position"), py::arg("fill_factor"))
        .def("set_custom_window_curve",
             [f32](PW& s, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 auto v = f32(a);
                 s.processor.setCustomWindowCurve(v.data(), v.size());
             },
             py::arg("curve"))
        .def("use_custom_window_curve", [](PW& s, bool u) { s.processor.useCustomWindowCurve(u); }, py::arg("use_custom"))
        .def("enable_windowing", [](PW& s, bool e) { s.processor.enableWindowing(e); }, py::arg("enable"))
        .def("set_grayscale_range", [](PW& s, float lo, float hi) { s.processor.setGrayscaleRange(lo, hi); },
             py::arg("min"), py::arg("max"))
        .def("set_signal_multiplicator_and_addend",
             [](PW& s, float m_, float a) { s.processor.setSignalMultiplicatorAndAddend(m_, a); },
             py::arg("multiplicator"), py::arg("addend"))
        .def("enable_log_scaling", [](PW& s, bool e) { s.processor.enableLogScaling(e); }, py::arg("enable"))
        .def("enable_background_removal", [](PW& s, bool e) { s.processor.enableBackgroundRemoval(e); }, py::arg("enable"))
        .def("set_background_removal_window_size", [](PW& s, int w) { s.processor.setBackgroundRemovalWindowSize(w); },
             py::arg("window_size"))
        .def("enable_bscan_flip", [](PW& s, bool e) { s.processor.enableBscanFlip(e); }, py::arg("enable"))
        .def("enable_sinusoidal_scan_correction", [](PW& s, bool e) { s.processor.enableSinusoidalScanCorrection(e); },
             py::arg("enable"))
        .def("enable_fixed_pattern_noise_removal", [](PW& s, bool e) { s.processor.enableFixedPatternNoiseRemoval(e); },
             py::arg("enable"))
        .def("enable_post_process_background_subtraction",
             [](PW& s, bool e) { s.processor.enablePostProcessBackgroundSubtraction(e); }, py::arg("enable"))
        .def("request_post_process_background_recording",
             [](PW& s) { s.processor.requestPostProcessBackgroundRecording(); })
        .def("set_post_process_background_weight", [](PW& s, float w) { s.processor.setPostProcessBackgroundWeight(w); },
             py::arg("weight"))
        .def("set_post_process_background_offset", [](PW& s, float o) { s.processor.setPostProcessBackgroundOffset(o); },
             py::arg("offset"))
        .def("has_post_process_background_profile", [](const PW& s) { return s.processor.hasPostProcessBackgroundProfile(); })
        .def("get_post_process_background_profile_size",
             [](const PW& s) { return s.processor.getPostProcessBackgroundProfileSize(); })
        .def("get_post_process_background_profile",
             [](const PW& s) {
                 if (!s.processor.hasPostProcessBackgroundProfile()) throw BufferError("No background profile available");
                 return to_numpy(s.processor.getPostProcessBackgroundProfile(),
                                 s.processor.getPostProcessBackgroundProfileSize());
             })
        .def("set_post_process_background_profile",
             [](PW& s, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 if (a.ndim() != 1) throw BufferError("Background profile must be a 1D array");
                 auto v = as_floats(a);
                 s.processor.setPostProcessBackgroundProfile(v.data(), v.size());
             },
             py::arg("profile"))
        .def("save_post_process_background_profile_to_file",
             [](const PW& s, const std::string& p) {
                 try {
                     s.processor.savePostProcessBackgroundProfileToFile(p);
                 } catch (const std::exception& e) {
                     throw ConfigurationError(std::string("Failed to save background profile: ") + e.what());
                 }
             },
             py::arg("filepath"))
        .def("load_post_process_background_profile_from_file",
             [](PW& s, const std::string& p) {
                 try {
                     s.processor.loadPostProcessBackgroundProfileFromFile(p);
                 } catch (const std::exception& e) {
                     throw ConfigurationError(std::string("Failed to load background profile: ") + e.what());
                 }
             },
             py::arg("filepath"))
        .def("request_fixed_pattern_noise_determination",
             [](PW& s) { s.processor.requestFixedPatternNoiseDetermination(); })
        .def("enable_continuous_fixed_pattern_noise_determination",
             [](PW& s, bool e) { s.processor.enableContinuousFixedPatternNoiseDetermination(e); }, py::arg("enable"))
        .def("set_fixed_pattern_noise_bscan_count", [](PW& s, int n) { s.processor.setFixedPatternNoiseBscanCount(n); },
             py::arg("length"))
        .def("has_fixed_pattern_noise_profile", [](const PW& s) { return s.processor.hasFixedPatternNoiseProfile(); })
        .def("get_fixed_pattern_noise_profile_size",
             [](const PW& s) { return s.processor.getFixedPatternNoiseProfileSize(); })
        .def("get_fixed_pattern_noise_profile",
             [](const PW& s) {
                 if (!s.processor.hasFixedPatternNoiseProfile())
                     throw BufferError("No fixed-pattern noise profile available");
                 return to_numpy(s.processor.getFixedPatternNoiseProfile(),
                                 2 * s.processor.getFixedPatternNoiseProfileSize());
             })
        .def("set_fixed_pattern_noise_profile",
             [](PW& s, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 if (a.ndim() != 1) throw BufferError("fixed-pattern noise profile must be a 1D array");
                 if (a.size() % 2 != 0)
                     throw BufferError("fixed-pattern noise profile must have even length (interleaved real/imag pairs)");
                 auto v = as_floats(a);
                 s.processor.setFixedPatternNoiseProfile(v.data(), v.size() / 2);
             },
             py::arg("profile"))
        .def("save_fixed_pattern_noise_profile_to_file",
             [](const PW& s, const std::string& p) {
                 try {
                     s.processor.saveFixedPatternNoiseProfileToFile(p);
                 } catch (const std::exception& e) {
                     throw ConfigurationError(std::string("Failed to save fixed-pattern noise profile: ") + e.what());
                 }
             },
             py::arg("filepath"))
        .def("load_fixed_pattern_noise_profile_from_file",
             [](PW& s, const std::string& p) {
                 try {
                     s.processor.loadFixedPatternNoiseProfileFromFile(p);
                 } catch (const std::exception& e) {
                     throw ConfigurationError(std::string("Failed to load fixed-pattern noise profile: ") + e.what());
                 }
             },
             py::arg("filepath"))
        .def("get_minimum_variance_mean",
             [](PW& s, py::array_t<float, py::array::c_style | py::array::forcecast> a, int width, int height,
                int segments) {
                 auto v = as_floats(a);
                 if (v.size() < static_cast<size_t>(width) * height * 2) throw BufferError("input too small");
                 std::vector<float> out = s.processor.getMinimumVarianceMean(v.data(), width, height, segments);
                 return to_numpy(out.data(), out.size());
             },
             py::arg("input"), py::arg("width"), py::arg("height"), py::arg("segments") = 8)
        .def("__enter__", [](PW& s) -> PW& { return s; }, py::return_value_policy::reference)
        .def("__exit__", [](PW& s, py::object, py::object, py::object) { s.stop(); })
        .def("__repr__", [](const PW& s) {
            return std::string("<Processor(backend=") +
                   (s.processor.getBackend() == ope::Backend::CUDA ? "CUDA" : "CPU") + ")>";
        });

    m.attr("__version__") = OPE_VERSION_STRING;
}
