// C++ host mirror of lucem-ops' Richardson-Lucy interface, header-only over the C ABI
// (lucem_hip.h). Names, argument meaning and error behaviour follow the reference:
//
//   lucem::PadMode, PadMode parse        crates/lucem-ops/src/core/pad.rs:13-57
//   lucem::OpsError / OpsResult          crates/lucem-ops/src/error.rs:8-59
//   lucem::RichardsonLucy                crates/lucem-ops/src/restore/deconvolve.rs:19-41
//   lucem::richardson_lucy_into          crates/lucem-ops/src/restore/deconvolve.rs:45-149
//   lucem::richardson_lucy               crates/lucem-ops/src/restore/deconvolve.rs:152-162
//   lucem::WienerDeconvolve, wiener_deconvolve_into, wiener_deconvolve
//                                        crates/lucem-ops/src/restore/deconvolve.rs:164-327
//   lucem::UnsupervisedWiener, unsupervised_wiener
//                                        crates/lucem-ops/src/restore/deconvolve.rs:330-509
//   lucem::NoiseModel, add_noise         crates/lucem-ops/src/intensity/noise.rs:17-143
//   lucem::Correlate2d, correlate_2d_into, correlate_2d
//                                        crates/lucem-ops/src/filter/convolve.rs:113-177
//   lucem::PhaseCorrelation, PhaseCorrelationResult, phase_correlation
//                                        crates/lucem-ops/src/register/phase.rs:18-250
//   lucem::fourier_gaussian_into, fourier_box_into, fourier_shift_into, butterworth, gabor_2d
//                                        crates/lucem-ops/src/filter/spectral.rs:20-268
//   lucem::Client, Workspace, DeviceArray crates/lucem-core/src/compute/target.rs:95-115,
//                                        memory/workspace.rs:17-37, memory/array.rs:20-151
//
// Every call enqueues on the client's stream; Client::sync() is the barrier
// (crates/lucem-core/src/compute/bench.rs:10-13).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../lucem_hip.h"

namespace lucem {

enum class PadMode : uint32_t { Reflect = 0, Constant = 1, Nearest = 2, Mirror = 3, Wrap = 4 };

// An error of the C ABI: `status` is the OpsError / CoreError variant, what() its message.
class OpsError : public std::runtime_error {
   public:
    OpsError(lucem_status st, const std::string& msg) : std::runtime_error(msg), status(st) {}
    lucem_status status;
};

// OpsResult<T>: a value or an OpsError (Result<T, OpsError> of the reference).
template <class T>
class OpsResult {
   public:
    OpsResult(T v) : value_(std::move(v)), status_(LUCEM_OK) {}
    OpsResult(lucem_status st, std::string msg) : status_(st), message_(std::move(msg)) {}
    bool is_ok() const { return status_ == LUCEM_OK; }
    lucem_status status() const { return status_; }
    const std::string& message() const { return message_; }
    T& unwrap() {
        if (!is_ok()) throw OpsError(status_, message_);
        return *value_;
    }

   private:
    std::optional<T> value_;
    lucem_status status_;
    std::string message_;
};

struct Unit {};

inline OpsResult<Unit> status_result(lucem_status st) {
    if (st == LUCEM_OK) return Unit{};
    return OpsResult<Unit>(st, lucem_last_error());
}

inline OpsResult<PadMode> parse_pad_mode(const std::string& name) {
    uint32_t id = 0;
    lucem_status st = lucem_pad_mode_parse(name.c_str(), &id);
    if (st != LUCEM_OK) return OpsResult<PadMode>(st, lucem_last_error());
    return static_cast<PadMode>(id);
}

struct RichardsonLucy {
    size_t num_iter = 30;
    bool clip = true;
    std::optional<double> filter_epsilon;
    PadMode mode = PadMode::Reflect;

    lucem_rl_settings to_c() const {
        lucem_rl_settings s;
        lucem_rl_settings_default(&s);
        s.num_iter = num_iter;
        s.clip = clip ? 1 : 0;
        s.has_filter_epsilon = filter_epsilon.has_value() ? 1 : 0;
        s.filter_epsilon = filter_epsilon.value_or(0.0);
        s.mode = static_cast<uint32_t>(mode);
        return s;
    }
};

// One device and one stream (own stream unless `stream` is given).
class Client {
   public:
    explicit Client(int device = 0, hipStream_t stream = nullptr, bool own_stream = true) {
        lucem_status st = lucem_client_create(device, stream, own_stream ? 1 : 0, &h_);
        if (st != LUCEM_OK) throw OpsError(st, lucem_last_error());
    }
    ~Client() {
        if (h_) lucem_client_destroy(h_);
    }
    Client(const Client&) = delete;
    Client& operator=(const Client&) = delete;
    lucem_client* handle() const { return h_; }
    hipStream_t stream() const { return static_cast<hipStream_t>(lucem_client_stream(h_)); }
    void sync() const {
        lucem_status st = lucem_client_sync(h_);
        if (st != LUCEM_OK) throw OpsError(st, lucem_last_error());
    }

   private:
    lucem_client* h_ = nullptr;
};

class Workspace {
   public:
    explicit Workspace(const Client& c) {
        lucem_status st = lucem_workspace_create(c.handle(), &h_);
        if (st != LUCEM_OK) throw OpsError(st, lucem_last_error());
    }
    ~Workspace() {
        if (h_) lucem_workspace_destroy(h_);
    }
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    lucem_workspace* handle() const { return h_; }
    uint64_t capacity_bytes() const { return lucem_workspace_capacity_bytes(h_); }

   private:
    lucem_workspace* h_ = nullptr;
};

// A C-contiguous device array. Copies share the buffer (as DeviceArray clones do).
template <class F>
class DeviceArray {
   public:
    static OpsResult<DeviceArray> empty(const Client&, std::vector<uint64_t> dims) {
        DeviceArray a;
        a.dims_ = std::move(dims);
        const size_t n = a.len();
        if (n) {
            void* p = nullptr;
            hipError_t e = hipMalloc(&p, n * sizeof(F));
            if (e != hipSuccess) return OpsResult<DeviceArray>(LUCEM_OUT_OF_DEVICE_MEMORY, hipGetErrorString(e));
            a.buf_ = std::shared_ptr<F>(static_cast<F*>(p), [](F* q) { (void)hipFree(q); });
        }
        return a;
    }
    static OpsResult<DeviceArray> from_host(const Client& c, std::vector<uint64_t> dims, const std::vector<F>& host) {
        auto r = empty(c, std::move(dims));
        if (!r.is_ok()) return r;
        DeviceArray a = r.unwrap();
        if (host.size() != a.len()) return OpsResult<DeviceArray>(LUCEM_SHAPE_MISMATCH, "host data length != shape");
        if (a.len()) {
            hipError_t e = hipMemcpyAsync(a.data(), host.data(), a.len() * sizeof(F), hipMemcpyHostToDevice, c.stream());
            if (e == hipSuccess) e = hipStreamSynchronize(c.stream());
            if (e != hipSuccess) return OpsResult<DeviceArray>(LUCEM_HIP_ERROR, hipGetErrorString(e));
        }
        return a;
    }
    std::vector<F> to_host(const Client& c) const {
        std::vector<F> out(len());
        if (len()) {
            hipError_t e = hipMemcpyAsync(out.data(), data(), len() * sizeof(F), hipMemcpyDeviceToHost, c.stream());
            if (e == hipSuccess) e = hipStreamSynchronize(c.stream());
            if (e != hipSuccess) throw OpsError(LUCEM_HIP_ERROR, hipGetErrorString(e));
        }
        return out;
    }
    size_t len() const {
        size_t n = 1;
        for (auto d : dims_) n *= d;
        return n;
    }
    bool is_empty() const { return len() == 0; }
    const std::vector<uint64_t>& dims() const { return dims_; }
    F* data() const { return buf_.get(); }
    bool shares_buffer_with(const DeviceArray& o) const { return buf_ && buf_ == o.buf_; }

   private:
    std::vector<uint64_t> dims_;
    std::shared_ptr<F> buf_;
};

// richardson_lucy_into::<f32> (deconvolve.rs:45-149).
inline OpsResult<Unit> richardson_lucy_into(const Client& client, const DeviceArray<float>& input,
                                            const DeviceArray<float>& psf, const DeviceArray<float>& output,
                                            const RichardsonLucy& settings, Workspace& ws) {
    const lucem_rl_settings s = settings.to_c();
    return status_result(lucem_richardson_lucy_into_f32(
        client.handle(), input.data(), input.dims().data(), (uint32_t)input.dims().size(), psf.data(),
        psf.dims().data(), (uint32_t)psf.dims().size(), output.data(), output.dims().data(),
        (uint32_t)output.dims().size(), &s, ws.handle()));
}

// richardson_lucy_into::<f64> (deconvolve.rs:45-149 with F = f64): FP64 direct correlations, 2-D.
inline OpsResult<Unit> richardson_lucy_into(const Client& client, const DeviceArray<double>& input,
                                            const DeviceArray<double>& psf, const DeviceArray<double>& output,
                                            const RichardsonLucy& settings, Workspace& ws) {
    const lucem_rl_settings s = settings.to_c();
    return status_result(lucem_richardson_lucy_into_f64(
        client.handle(), input.data(), input.dims().data(), (uint32_t)input.dims().size(), psf.data(),
        psf.dims().data(), (uint32_t)psf.dims().size(), output.data(), output.dims().data(),
        (uint32_t)output.dims().size(), &s, ws.handle()));
}

// richardson_lucy::<F> (deconvolve.rs:152-162), F = float or double: allocates the output and a
// workspace.
template <class F>
inline OpsResult<DeviceArray<F>> richardson_lucy(const Client& client, const DeviceArray<F>& input,
                                                 const DeviceArray<F>& psf, const RichardsonLucy& settings) {
    auto out = DeviceArray<F>::empty(client, input.dims());
    if (!out.is_ok()) return out;
    Workspace ws(client);
    auto r = richardson_lucy_into(client, input, psf, out.unwrap(), settings, ws);
    if (!r.is_ok()) return OpsResult<DeviceArray<F>>(r.status(), r.message());
    client.sync();  // the workspace is released here; its buffers must not be in use
    return out;
}

// The N-D (2-D or 3-D) form of richardson_lucy_into.
inline OpsResult<Unit> richardson_lucy_nd_into(const Client& client, const DeviceArray<float>& input,
                                               const DeviceArray<float>& psf, const DeviceArray<float>& output,
                                               const RichardsonLucy& settings, Workspace& ws) {
    const lucem_rl_settings s = settings.to_c();
    return status_result(lucem_richardson_lucy_nd_into_f32(
        client.handle(), input.data(), input.dims().data(), (uint32_t)input.dims().size(), psf.data(),
        psf.dims().data(), (uint32_t)psf.dims().size(), output.data(), output.dims().data(),
        (uint32_t)output.dims().size(), &s, ws.handle()));
}

// Settings of wiener_deconvolve (deconvolve.rs:164-181).
struct WienerDeconvolve {
    double balance = 0.1;
    bool clip = true;

    lucem_wiener_settings to_c() const {
        lucem_wiener_settings s;
        lucem_wiener_settings_default(&s);
        s.balance = balance;
        s.clip = clip ? 1 : 0;
        return s;
    }
};

// wiener_deconvolve_into::<f32> (deconvolve.rs:213-315).
inline OpsResult<Unit> wiener_deconvolve_into(const Client& client, const DeviceArray<float>& input,
                                              const DeviceArray<float>& psf, const DeviceArray<float>& output,
                                              const WienerDeconvolve& settings, Workspace& ws) {
    const lucem_wiener_settings s = settings.to_c();
    return status_result(lucem_wiener_deconvolve_into_f32(
        client.handle(), input.data(), input.dims().data(), (uint32_t)input.dims().size(), psf.data(),
        psf.dims().data(), (uint32_t)psf.dims().size(), output.data(), output.dims().data(),
        (uint32_t)output.dims().size(), &s, ws.handle()));
}

// wiener_deconvolve::<f32> (deconvolve.rs:318-327): allocates the output and a workspace.
inline OpsResult<DeviceArray<float>> wiener_deconvolve(const Client& client, const DeviceArray<float>& input,
                                                       const DeviceArray<float>& psf, const WienerDeconvolve& settings) {
    auto out = DeviceArray<float>::empty(client, input.dims());
    if (!out.is_ok()) return out;
    Workspace ws(client);
    auto r = wiener_deconvolve_into(client, input, psf, out.unwrap(), settings, ws);
    if (!r.is_ok()) return OpsResult<DeviceArray<float>>(r.status(), r.message());
    client.sync();
    return out;
}

// Settings of unsupervised_wiener (deconvolve.rs:330-353).
struct UnsupervisedWiener {
    size_t num_iter = 30;
    size_t burnin = 10;
    bool clip = true;
    uint64_t seed = 0;

    lucem_unsupervised_wiener_settings to_c() const {
        lucem_unsupervised_wiener_settings s;
        lucem_unsupervised_wiener_settings_default(&s);
        s.num_iter = num_iter;
        s.burnin = burnin;
        s.clip = clip ? 1 : 0;
        s.seed = seed;
        return s;
    }
};

// unsupervised_wiener::<f32> (deconvolve.rs:357-509): (restored image, balance chain).
inline OpsResult<std::pair<DeviceArray<float>, std::vector<double>>> unsupervised_wiener(
    const Client& client, const DeviceArray<float>& input, const DeviceArray<float>& psf,
    const UnsupervisedWiener& settings) {
    using R = OpsResult<std::pair<DeviceArray<float>, std::vector<double>>>;
    auto out = DeviceArray<float>::empty(client, input.dims());
    if (!out.is_ok()) return R(out.status(), out.message());
    const lucem_unsupervised_wiener_settings s = settings.to_c();
    std::vector<double> chain(std::max<size_t>(settings.num_iter, 1));
    uint64_t len = 0;
    Workspace ws(client);
    const lucem_status st = lucem_unsupervised_wiener_f32(
        client.handle(), input.data(), input.dims().data(), (uint32_t)input.dims().size(), psf.data(),
        psf.dims().data(), (uint32_t)psf.dims().size(), out.unwrap().data(), &s, chain.data(), chain.size(), &len,
        ws.handle());
    if (st != LUCEM_OK) return R(st, lucem_last_error());
    chain.resize(len);
    return std::make_pair(out.unwrap(), std::move(chain));
}

// NoiseModel (intensity/noise.rs:17-32).
struct NoiseModel {
    uint32_t kind;
    double p0 = 0.0, p1 = 0.0;
    static NoiseModel gaussian(double mean, double std) { return {LUCEM_NOISE_GAUSSIAN, mean, std}; }
    static NoiseModel speckle(double mean, double std) { return {LUCEM_NOISE_SPECKLE, mean, std}; }
    static NoiseModel salt(double amount) { return {LUCEM_NOISE_SALT, amount, 0.0}; }
    static NoiseModel pepper(double amount) { return {LUCEM_NOISE_PEPPER, amount, 0.0}; }
    static NoiseModel salt_and_pepper(double amount, double svp) { return {LUCEM_NOISE_SALT_AND_PEPPER, amount, svp}; }
    static NoiseModel poisson(double scale) { return {LUCEM_NOISE_POISSON, scale, 0.0}; }
};

// add_noise::<f32> (intensity/noise.rs:35-143): a new array holding `input` with the noise.
inline OpsResult<DeviceArray<float>> add_noise(const Client& client, const DeviceArray<float>& input,
                                               const NoiseModel& model, uint64_t seed, bool clip) {
    auto out = DeviceArray<float>::empty(client, input.dims());
    if (!out.is_ok()) return out;
    const lucem_noise_model m{model.kind, 0u, model.p0, model.p1};
    const lucem_status st = lucem_add_noise_f32(client.handle(), input.data(), input.dims().data(),
                                                (uint32_t)input.dims().size(), out.unwrap().data(), &m, seed,
                                                clip ? 1 : 0);
    if (st != LUCEM_OK) return OpsResult<DeviceArray<float>>(st, lucem_last_error());
    return out;
}

// Settings of correlate_2d (filter/convolve.rs:113-123): boundary mode, cval, window origin
// per axis (2-D: (y, x)).
struct Correlate2d {
    PadMode mode = PadMode::Reflect;
    double cval = 0.0;
    std::vector<int32_t> origin = {0, 0};

    lucem_correlate_settings to_c() const {
        lucem_correlate_settings s;
        lucem_correlate_settings_default(&s);
        s.mode = (uint32_t)mode;
        s.cval = cval;
        for (size_t i = 0; i < origin.size() && i < 3; i++) s.origin[i] = origin[i];
        return s;
    }
};

// correlate_2d_into (filter/convolve.rs:126-166); f32 through the FFT passes, f64 direct.
inline OpsResult<Unit> correlate_2d_into(const Client& client, const DeviceArray<float>& input,
                                         const DeviceArray<float>& weights, const DeviceArray<float>& output,
                                         const Correlate2d& settings, Workspace& ws) {
    const lucem_correlate_settings s = settings.to_c();
    return status_result(lucem_correlate_2d_into_f32(
        client.handle(), input.data(), input.dims().data(), (uint32_t)input.dims().size(), weights.data(),
        weights.dims().data(), (uint32_t)weights.dims().size(), output.data(), output.dims().data(),
        (uint32_t)output.dims().size(), &s, ws.handle()));
}

inline OpsResult<Unit> correlate_2d_into(const Client& client, const DeviceArray<double>& input,
                                         const DeviceArray<double>& weights, const DeviceArray<double>& output,
                                         const Correlate2d& settings) {
    const lucem_correlate_settings s = settings.to_c();
    return status_result(lucem_correlate_2d_into_f64(
        client.handle(), input.data(), input.dims().data(), (uint32_t)input.dims().size(), weights.data(),
        weights.dims().data(), (uint32_t)weights.dims().size(), output.data(), output.dims().data(),
        (uint32_t)output.dims().size(), &s));
}

// correlate_2d (filter/convolve.rs:169-177): allocates the output.
inline OpsResult<DeviceArray<float>> correlate_2d(const Client& client, const DeviceArray<float>& input,
                                                  const DeviceArray<float>& weights, const Correlate2d& settings) {
    auto out = DeviceArray<float>::empty(client, input.dims());
    if (!out.is_ok()) return out;
    Workspace ws(client);
    auto r = correlate_2d_into(client, input, weights, out.unwrap(), settings, ws);
    if (!r.is_ok()) return OpsResult<DeviceArray<float>>(r.status(), r.message());
    client.sync();
    return out;
}

inline OpsResult<DeviceArray<double>> correlate_2d(const Client& client, const DeviceArray<double>& input,
                                                   const DeviceArray<double>& weights, const Correlate2d& settings) {
    auto out = DeviceArray<double>::empty(client, input.dims());
    if (!out.is_ok()) return out;
    auto r = correlate_2d_into(client, input, weights, out.unwrap(), settings);
    if (!r.is_ok()) return OpsResult<DeviceArray<double>>(r.status(), r.message());
    return out;
}

// Settings and result of phase_correlation (register/phase.rs:18-47).
struct PhaseCorrelation {
    size_t upsample_factor = 1;
    bool normalize_phase = true;
};
struct PhaseCorrelationResult {
    std::vector<double> shift;
    double error = 0.0;
    double phasediff = 0.0;
};

// phase_correlation::<f32> (register/phase.rs:52-250): blocks until the result is on the host.
inline OpsResult<PhaseCorrelationResult> phase_correlation(const Client& client, const DeviceArray<float>& reference,
                                                           const DeviceArray<float>& moving,
                                                           const PhaseCorrelation& settings) {
    lucem_phase_correlation_settings s;
    lucem_phase_correlation_settings_default(&s);
    s.upsample_factor = settings.upsample_factor;
    s.normalize_phase = settings.normalize_phase ? 1 : 0;
    lucem_phase_correlation_result r{};
    const lucem_status st = lucem_phase_correlation_f32(
        client.handle(), reference.data(), reference.dims().data(), (uint32_t)reference.dims().size(), moving.data(),
        moving.dims().data(), (uint32_t)moving.dims().size(), &s, &r, nullptr);
    if (st != LUCEM_OK) return OpsResult<PhaseCorrelationResult>(st, lucem_last_error());
    PhaseCorrelationResult out;
    out.shift.assign(r.shift, r.shift + r.ndim);
    out.error = r.error;
    out.phasediff = r.phasediff;
    return out;
}

// fourier_{gaussian,box,shift}_into (filter/spectral.rs:20-165): in place on a split spectrum,
// one value or one per axis.
inline OpsResult<Unit> fourier_gaussian_into(const Client& client, const DeviceArray<float>& re,
                                             const DeviceArray<float>& im, const std::vector<double>& sigmas) {
    return status_result(lucem_fourier_gaussian_into_f32(client.handle(), re.data(), re.dims().data(),
                                                         (uint32_t)re.dims().size(), im.data(), im.dims().data(),
                                                         (uint32_t)im.dims().size(), sigmas.data(),
                                                         (uint32_t)sigmas.size()));
}
inline OpsResult<Unit> fourier_box_into(const Client& client, const DeviceArray<float>& re,
                                        const DeviceArray<float>& im, const std::vector<double>& sizes) {
    return status_result(lucem_fourier_box_into_f32(client.handle(), re.data(), re.dims().data(),
                                                    (uint32_t)re.dims().size(), im.data(), im.dims().data(),
                                                    (uint32_t)im.dims().size(), sizes.data(), (uint32_t)sizes.size()));
}
inline OpsResult<Unit> fourier_shift_into(const Client& client, const DeviceArray<float>& re,
                                          const DeviceArray<float>& im, const std::vector<double>& shifts) {
    return status_result(lucem_fourier_shift_into_f32(client.handle(), re.data(), re.dims().data(),
                                                      (uint32_t)re.dims().size(), im.data(), im.dims().data(),
                                                      (uint32_t)im.dims().size(), shifts.data(),
                                                      (uint32_t)shifts.size()));
}

// butterworth::<f32> (filter/spectral.rs:166-212): allocates the output.
inline OpsResult<DeviceArray<float>> butterworth(const Client& client, const DeviceArray<float>& input,
                                                 double cutoff_ratio, bool high_pass, double order) {
    auto out = DeviceArray<float>::empty(client, input.dims());
    if (!out.is_ok()) return out;
    auto r = status_result(lucem_butterworth_f32(client.handle(), input.data(), input.dims().data(),
                                                 (uint32_t)input.dims().size(), out.unwrap().data(),
                                                 input.dims().data(), (uint32_t)input.dims().size(), cutoff_ratio,
                                                 high_pass ? 1 : 0, order, nullptr));
    if (!r.is_ok()) return OpsResult<DeviceArray<float>>(r.status(), r.message());
    return out;
}

// gabor_2d::<f32> (filter/spectral.rs:249-268): (real response, imaginary response).
inline OpsResult<std::pair<DeviceArray<float>, DeviceArray<float>>> gabor_2d(
    const Client& client, const DeviceArray<float>& input, double frequency, double theta, double sigma_x,
    double sigma_y, double n_stds, double offset, PadMode mode, double cval) {
    using R = OpsResult<std::pair<DeviceArray<float>, DeviceArray<float>>>;
    auto re = DeviceArray<float>::empty(client, input.dims());
    if (!re.is_ok()) return R(re.status(), re.message());
    auto im = DeviceArray<float>::empty(client, input.dims());
    if (!im.is_ok()) return R(im.status(), im.message());
    lucem_gabor_settings s;
    lucem_gabor_settings_default(&s);
    s.frequency = frequency;
    s.theta = theta;
    s.sigma_x = sigma_x;
    s.sigma_y = sigma_y;
    s.n_stds = n_stds;
    s.offset = offset;
    s.mode = (uint32_t)mode;
    s.cval = cval;
    auto r = status_result(lucem_gabor_2d_f32(client.handle(), input.data(), input.dims().data(),
                                              (uint32_t)input.dims().size(), re.unwrap().data(), im.unwrap().data(),
                                              &s, nullptr));
    if (!r.is_ok()) return R(r.status(), r.message());
    return std::make_pair(re.unwrap(), im.unwrap());
}

}  // namespace lucem
