// Device-resident storage and the elementwise operations of DeviceArray on
// the HIP device.
//
// The reference's DeviceArray (crates/scir-gpu/src/lib.rs:76-189) keeps its
// data in a host Vec whatever its device; to_device only tags it, so every
// GPU call copies in and out.  Here an array moved to the HIP device owns a
// device buffer: FIR calls on it (scir_hip_fir_f32_dev with the buffer's
// pointer) and the elementwise operations below run without host
// transfers.  Each elementwise result is one IEEE operation per element, as
// in the reference's CPU loops (lib.rs:206-260), so it is bit-identical.
#include "common.hpp"

#include <memory>

namespace scir {
namespace {

constexpr int kThreads = 256;

template <int OP>  // 0: a + alpha, 1: a * alpha, 2: a + b
__global__ __launch_bounds__(kThreads) void elementwise_kernel(const float *__restrict__ a,
                                                               const float *__restrict__ b,
                                                               float alpha, float *__restrict__ out,
                                                               int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kThreads) {
        if (OP == 0) out[i] = a[i] + alpha;
        if (OP == 1) out[i] = a[i] * alpha;
        if (OP == 2) out[i] = a[i] + b[i];
    }
}

template <int OP>
int launch_elementwise(const float *a, const float *b, float alpha, float *out, int64_t n,
                       void *stream)
{
    clear_error();
    if (n < 0) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative length");
    if (n == 0) return SCIR_HIP_OK;
    if (!a || !out || (OP == 2 && !b))
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
    DeviceState *st = nullptr;
    SCIR_TRY(current_device_state(&st));
    const int64_t want = (n + kThreads - 1) / kThreads;
    const int64_t blocks = std::min<int64_t>(want, (int64_t)st->cu_count * 16);
    hipLaunchKernelGGL((elementwise_kernel<OP>), dim3((unsigned)blocks), dim3(kThreads), 0,
                       as_stream(stream), a, b, alpha, out, n);
    return check_launch("elementwise_kernel");
}

}  // namespace
}  // namespace scir

using namespace scir;

struct scir_hip_buffer_s {
    int device = -1;
    void *ptr = nullptr;
    size_t bytes = 0;
};

extern "C" {

int scir_hip_buffer_create(size_t bytes, scir_hip_buffer *buf)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!buf) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "buf is null");
        *buf = nullptr;
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        host_alloc_point();
        std::unique_ptr<scir_hip_buffer_s> b(new scir_hip_buffer_s);
        b->device = st->device;
        b->bytes = bytes;
        if (bytes > 0) SCIR_TRY(device_alloc(&b->ptr, bytes, "scir_hip_buffer_create"));
        *buf = b.release();
        return SCIR_HIP_OK;
    });
}

int scir_hip_buffer_upload(scir_hip_buffer buf, const void *host, size_t bytes)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!buf || (bytes > 0 && !host)) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null argument");
        if (bytes > buf->bytes)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "upload of %zu bytes into %zu", bytes,
                        buf->bytes);
        if (bytes > 0) SCIR_HIP_TRY(hipMemcpy(buf->ptr, host, bytes, hipMemcpyHostToDevice));
        return SCIR_HIP_OK;
    });
}

int scir_hip_buffer_download(scir_hip_buffer buf, void *host, size_t bytes)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!buf || (bytes > 0 && !host)) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null argument");
        if (bytes > buf->bytes)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "download of %zu bytes from %zu", bytes,
                        buf->bytes);
        // hipMemcpy waits for work on the null stream; callers that used other
        // streams synchronise them first.
        if (bytes > 0) SCIR_HIP_TRY(hipMemcpy(host, buf->ptr, bytes, hipMemcpyDeviceToHost));
        return SCIR_HIP_OK;
    });
}

void *scir_hip_buffer_device_ptr(scir_hip_buffer buf) { return buf ? buf->ptr : nullptr; }

int scir_hip_buffer_destroy(scir_hip_buffer buf)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!buf) return SCIR_HIP_OK;
        if (buf->ptr) {
            // No kernel may still use the memory: wait for the buffer's own
            // device, then restore the caller's current device.
            int cur = -1;
            SCIR_HIP_TRY(hipGetDevice(&cur));
            if (cur != buf->device) SCIR_HIP_TRY(hipSetDevice(buf->device));
            hipError_t e = hipDeviceSynchronize();
            if (e == hipSuccess) e = hipFree(buf->ptr);
            if (cur != buf->device) (void)hipSetDevice(cur);
            SCIR_HIP_TRY(e);
        }
        delete buf;
        return SCIR_HIP_OK;
    });
}

int scir_hip_add_scalar_f32_dev(const float *a, float alpha, float *out, int64_t n, void *stream)
{
    return guarded(__func__, [&]() -> int {
        return launch_elementwise<0>(a, nullptr, alpha, out, n, stream);
    });
}

int scir_hip_mul_scalar_f32_dev(const float *a, float alpha, float *out, int64_t n, void *stream)
{
    return guarded(__func__, [&]() -> int {
        return launch_elementwise<1>(a, nullptr, alpha, out, n, stream);
    });
}

int scir_hip_add_f32_dev(const float *a, const float *b, float *out, int64_t n, void *stream)
{
    return guarded(__func__, [&]() -> int {
        return launch_elementwise<2>(a, b, 0.0f, out, n, stream);
    });
}

}  // extern "C"
