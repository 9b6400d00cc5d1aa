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

// HIP FIR: host arrays in and out; throws GpuError on any failure.
inline Array2<float> fir1d_batched_f32_hip(const ArrayView2<float> &x,
                                           const std::vector<float> &taps,
                                           scir_hip_precision precision = SCIR_HIP_PREC_F32)
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
}  // namespace scir_signal
