// scir_gpu.hpp -- C++ host-side mirror of SciR's FIR dispatch API over the
// HIP C ABI (include/scir_hip.h).  Header-only; link with libscir_hip.so.
//
// Mirrors, with the same names and argument meaning:
//   DType, Device, GpuError          crates/scir-gpu/src/lib.rs:15-35, 56-74
//                                    (Device gains the Hip variant)
//   fir1d_batched_f32(x, taps)       CPU FIR, crates/scir-gpu/src/lib.rs:1134-1152
//                                    (the library's CPU arm: same arithmetic order,
//                                    bit-identical)
//   fir1d_batched_f32_hip(x, taps)   replaces cuda::fir1d_batched_f32_cuda,
//                                    crates/scir-gpu/src/lib.rs:1036-1113
//   fir1d_batched_f32_auto(x, taps, device)
//                                    crates/scir-gpu/src/lib.rs:515-531
//   scir_signal::gpu::fir1d_batched_f32(x, taps, device)
//                                    crates/scir-signal/src/lib.rs:364-375
//   scir_signal::sosfilt / filtfilt(sos, x, device)
//                                    crates/scir-signal/src/lib.rs:260-291
//   scir_signal::resample_poly(x, up, down, device)
//                                    crates/scir-signal/src/lib.rs:293-362
//   scir_gpu::fir1d_batched_f32_vjp_hip(x, taps, g)
//                                    no counterpart (the adjoints)
//
// `taps` is in the reference's order (y[i] = sum_t taps[K-1-t] x[i-t]); the C
// ABI takes lfilter-order b = reversed taps, so the reversal happens here.
//
// Failure behaviour: the reference dispatcher silently recomputes on the CPU
// when the CUDA call fails (lib.rs:520-523).  Here the Hip arm throws
// GpuError, so a broken GPU path cannot hide behind a CPU result; set the
// environment variable SCIR_HIP_CPU_FALLBACK=1 for the reference behaviour
// (each fallback then writes one line to stderr).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/scir_hip.h"

namespace scir_gpu {

enum class DType { F32, F64 };

enum class Device { Cpu, Hip };

class GpuError : public std::runtime_error {
  public:
    enum class Kind { BackendUnavailable, ShapeMismatch };
    GpuError(Kind k, const std::string &msg)
        : std::runtime_error(k == Kind::ShapeMismatch ? "shape mismatch: " + msg
                                                      : "backend not available: " + msg),
          kind_(k)
    {
    }
    Kind kind() const { return kind_; }

  private:
    Kind kind_;
};

// Borrowed 2-D view (rows x cols) with element strides, like ndarray's
// ArrayView2 (any layout).
template <class T>
struct ArrayView2 {
    const T *data;
    size_t rows, cols;
    ptrdiff_t row_stride, col_stride;
};

// Owned row-major 2-D array (the reference returns an owned Array2).
template <class T>
struct Array2 {
    size_t rows = 0, cols = 0;
    std::vector<T> data;
    Array2() = default;
    Array2(size_t r, size_t c) : rows(r), cols(c), data(r * c, T(0)) {}
    T &operator()(size_t i, size_t j) { return data[i * cols + j]; }
    const T &operator()(size_t i, size_t j) const { return data[i * cols + j]; }
    ArrayView2<T> view() const
    {
        return {data.data(), rows, cols, (ptrdiff_t)cols, 1};
    }
};

namespace detail {

// Standard (C-contiguous) copy of any view, as as_standard_layout() would.
inline Array2<float> standard(const ArrayView2<float> &x)
{
    Array2<float> out(x.rows, x.cols);
    for (size_t i = 0; i < x.rows; ++i)
        for (size_t j = 0; j < x.cols; ++j)
            out(i, j) = x.data[(ptrdiff_t)i * x.row_stride + (ptrdiff_t)j * x.col_stride];
    return out;
}

inline std::vector<float> reversed(const std::vector<float> &taps)
{
    return std::vector<float>(taps.rbegin(), taps.rend());
}

inline GpuError error_from(int status)
{
    const char *msg = scir_hip_last_error();
    return GpuError(status == SCIR_HIP_ERR_INVALID_ARGUMENT ? GpuError::Kind::ShapeMismatch
                                                            : GpuError::Kind::BackendUnavailable,
                    msg ? msg : "");
}

}  // namespace detail

// CPU FIR (crates/scir-gpu/src/lib.rs:1134-1152), same shape out.
inline Array2<float> fir1d_batched_f32(const ArrayView2<float> &x, const std::vector<float> &taps)
{
    Array2<float> xs = detail::standard(x);
    std::vector<float> b = detail::reversed(taps);
    Array2<float> y(x.rows, x.cols);
    if (x.rows == 0 || x.cols == 0) return y;
    int st = scir_hip_fir_f32_cpu(xs.data.data(), (int64_t)x.cols, b.data(), (int32_t)b.size(),
                                  (int64_t)x.rows, (int64_t)x.cols, y.data.data(),
                                  (int64_t)x.cols, 0);
    if (st != SCIR_HIP_OK) throw detail::error_from(st);
    return y;
}

// HIP FIR: host arrays in and out; throws GpuError on any failure.  AUTO
// (the default) takes the FP32 direct form for short filters and BF16x3 on
// MFMA for long ones, both within the FP32 bound of the reference loop.
inline Array2<float> fir1d_batched_f32_hip(const ArrayView2<float> &x,
                                           const std::vector<float> &taps,
                                           scir_hip_precision precision = SCIR_HIP_PREC_AUTO)
{
    Array2<float> xs = detail::standard(x);
    std::vector<float> b = detail::reversed(taps);
    Array2<float> y(x.rows, x.cols);
    if (x.rows == 0 || x.cols == 0) return y;
    int st = scir_hip_fir_f32_host(xs.data.data(), (int64_t)x.cols, b.data(), (int32_t)b.size(),
                                   (int64_t)x.rows, (int64_t)x.cols, y.data.data(),
                                   (int64_t)x.cols, precision);
    if (st != SCIR_HIP_OK) throw detail::error_from(st);
    return y;
}

// Device dispatch (crates/scir-gpu/src/lib.rs:515-531).
inline Array2<float> fir1d_batched_f32_auto(const ArrayView2<float> &x,
                                            const std::vector<float> &taps, Device device)
{
    if (device == Device::Cpu) return fir1d_batched_f32(x, taps);
    try {
        return fir1d_batched_f32_hip(x, taps);
    } catch (const GpuError &e) {
        const char *fb = std::getenv("SCIR_HIP_CPU_FALLBACK");
        if (fb && std::strcmp(fb, "1") == 0) {
            std::fprintf(stderr, "scir_gpu: HIP FIR failed (%s); recomputing on the CPU\n",
                         e.what());
            return fir1d_batched_f32(x, taps);
        }
        throw;
    }
}

}  // namespace scir_gpu

namespace scir_signal {
namespace gpu {

// Batched FIR for f32 with device selection (crates/scir-signal/src/lib.rs:372).
inline scir_gpu::Array2<float> fir1d_batched_f32(const scir_gpu::ArrayView2<float> &x,
                                                 const std::vector<float> &taps,
                                                 scir_gpu::Device device)
{
    return scir_gpu::fir1d_batched_f32_auto(x, taps, device);
}

}  // namespace gpu

// One second-order section {b0, b1, b2, a0, a1, a2}: the layout of the
// reference's #[repr(C)] SosCoeffs (crates/scir-iir-filters/src/sos.rs);
// a0 is not read, as in the reference.
struct SosCoeffs {
    double b0, b1, b2, a0, a1, a2;
};
using Sos = std::vector<SosCoeffs>;

namespace detail {

inline std::vector<double> sos_filter(const Sos &sos, const std::vector<double> &x,
                                      scir_gpu::Device device, bool zero_phase)
{
    std::vector<double> y(x.size());
    if (x.empty()) return y;
    const double *c = sos.empty() ? nullptr : &sos[0].b0;
    const int64_t n = (int64_t)x.size();
    int st;
    if (device == scir_gpu::Device::Cpu)
        st = zero_phase ? scir_hip_filtfilt_f64_cpu(x.data(), n, 1, n, c, (int32_t)sos.size(),
                                                    y.data(), n, 1)
                        : scir_hip_sosfilt_f64_cpu(x.data(), n, 1, n, c, (int32_t)sos.size(),
                                                   y.data(), n, 1);
    else
        st = zero_phase ? scir_hip_filtfilt_f64_host(x.data(), n, 1, n, c, (int32_t)sos.size(),
                                                     y.data(), n, 0)
                        : scir_hip_sosfilt_f64_host(x.data(), n, 1, n, c, (int32_t)sos.size(),
                                                    y.data(), n, 0);
    if (st != SCIR_HIP_OK) throw scir_gpu::detail::error_from(st);
    return y;
}

}  // namespace detail

// sosfilt / filtfilt (crates/scir-signal/src/lib.rs:260-291) with the device
// choice added; the CPU arm is the reference loop, bit-identical.
inline std::vector<double> sosfilt(const Sos &sos, const std::vector<double> &x,
                                   scir_gpu::Device device = scir_gpu::Device::Cpu)
{
    return detail::sos_filter(sos, x, device, false);
}

inline std::vector<double> filtfilt(const Sos &sos, const std::vector<double> &x,
                                    scir_gpu::Device device = scir_gpu::Device::Cpu)
{
    return detail::sos_filter(sos, x, device, true);
}

// resample_poly (crates/scir-signal/src/lib.rs:293-362) for any up, down and
// filter h, with the device choice added.
inline std::vector<double> resample_poly(const std::vector<double> &x, int up, int down,
                                         const std::vector<double> &h,
                                         scir_gpu::Device device = scir_gpu::Device::Cpu)
{
    int64_t m = 0;
    const int64_t n = (int64_t)x.size();
    int st = scir_hip_resample_len(n, up, down, (int32_t)h.size(), &m);
    if (st != SCIR_HIP_OK) throw scir_gpu::detail::error_from(st);
    std::vector<double> y((size_t)m);
    if (m == 0 || n == 0) return y;
    st = device == scir_gpu::Device::Cpu
             ? scir_hip_resample_poly_f64_cpu(x.data(), n, 1, n, up, down, h.data(),
                                              (int32_t)h.size(), y.data(), m, 1)
             : scir_hip_resample_poly_f64_host(x.data(), n, 1, n, up, down, h.data(),
                                               (int32_t)h.size(), y.data(), m);
    if (st != SCIR_HIP_OK) throw scir_gpu::detail::error_from(st);
    return y;
}

}  // namespace scir_signal

namespace scir_gpu {

// For L = sum(g * fir1d_batched_f32(x, taps)): (dL/dx, dL/dtaps), taps in
// the reference's order, on the HIP device.  No counterpart in the reference.
inline std::pair<Array2<float>, std::vector<float>> fir1d_batched_f32_vjp_hip(
    const ArrayView2<float> &x, const std::vector<float> &taps, const ArrayView2<float> &g,
    scir_hip_precision precision = SCIR_HIP_PREC_F32)
{
    if (g.rows != x.rows || g.cols != x.cols)
        throw GpuError(GpuError::Kind::ShapeMismatch, "g must have the shape of x");
    Array2<float> xs = detail::standard(x), gs = detail::standard(g);
    std::vector<float> b = detail::reversed(taps);
    Array2<float> dx(x.rows, x.cols);
    std::vector<float> db(b.size(), 0.0f);
    const int64_t n = (int64_t)x.cols, ld = n > 0 ? n : 1;
    int st = scir_hip_fir_vjp_f32_host(xs.data.data(), ld, gs.data.data(), ld, b.data(),
                                       (int32_t)b.size(), (int64_t)x.rows, n, dx.data.data(), ld,
                                       db.data(), precision);
    if (st != SCIR_HIP_OK) throw detail::error_from(st);
    return {dx, std::vector<float>(db.rbegin(), db.rend())};
}

}  // namespace scir_gpu
