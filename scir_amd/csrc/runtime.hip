// Runtime of libscir_hip.so: error strings, lazily-initialised per-device
// state, library-owned scratch, HIP-graph capture helpers and the synthetic
// input generator.
//
// The reference creates a CUDA context and JIT-loads its PTX module on every
// call (crates/scir-gpu/src/lib.rs:601-622, 814-824) and frees its buffers
// afterwards (:1108-1110).  Here the device state, code objects (loaded by
// the HIP runtime with the library) and scratch live for the process.
#include "common.hpp"

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

namespace scir {

namespace {
std::mutex g_states_mu;
std::vector<std::unique_ptr<DeviceState>> g_states;
}  // namespace

static int init_state(DeviceState *st, int dev)
{
    hipDeviceProp_t prop;
    SCIR_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    st->arch = prop.gcnArchName;
    if (st->arch.compare(0, 6, "gfx950") != 0)
        return fail(SCIR_HIP_ERR_NO_DEVICE,
                    "device %d is %s; libscir_hip.so holds gfx950 (MI355X) code only", dev,
                    prop.gcnArchName);
    st->device = dev;
    st->cu_count = prop.multiProcessorCount;
    st->lds_per_block = prop.sharedMemPerBlock;
    // The allocation-free dL/db scratch: its address on this device (current
    // here), kept in the state; every later call uses this pointer.
    void *scratch = nullptr;
    if (vjp_noalloc_scratch_address(&scratch) == SCIR_HIP_OK) {
        st->vjp_noalloc_scratch = static_cast<float *>(scratch);
    } else {
        st->vjp_noalloc_error = scir_hip_last_error();
        clear_error();
    }
    st->ready = true;
    return SCIR_HIP_OK;
}

int current_device_state(DeviceState **out)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(SCIR_HIP_ERR_NO_DEVICE, "no HIP device is visible");
    int dev = 0;
    SCIR_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(g_states_mu);
    if ((int)g_states.size() < count) g_states.resize(count);
    if (!g_states[dev]) g_states[dev] = std::make_unique<DeviceState>();
    DeviceState *st = g_states[dev].get();
    if (!st->ready) SCIR_TRY(init_state(st, dev));
    *out = st;
    return SCIR_HIP_OK;
}

int device_alloc(void **ptr, size_t bytes, const char *what)
{
    *ptr = nullptr;
    if (alloc_failure_injected())
        return fail(SCIR_HIP_ERR_OUT_OF_MEMORY, "%s: device allocation of %zu bytes failed "
                    "(SCIR_HIP_FAIL_ALLOC is set)", what, bytes);
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(e == hipErrorOutOfMemory ? SCIR_HIP_ERR_OUT_OF_MEMORY : SCIR_HIP_ERR_BACKEND,
                    "%s: hipMalloc(%zu) failed: %s", what, bytes, hipGetErrorString(e));
    }
    return SCIR_HIP_OK;
}

static int refuse_growth_in_capture(size_t bytes, hipStream_t stream)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (stream != nullptr && hipStreamIsCapturing(stream, &cs) == hipSuccess &&
        cs != hipStreamCaptureStatusNone)
        return fail(SCIR_HIP_ERR_CAPTURE,
                    "scratch must grow to %zu bytes during stream capture; run the call "
                    "once outside capture first", bytes);
    return SCIR_HIP_OK;
}

int ensure_buffer(DeviceBuffer &buf, size_t bytes, hipStream_t stream)
{
    if (bytes <= buf.bytes) return SCIR_HIP_OK;
    SCIR_TRY(refuse_growth_in_capture(bytes, stream));
    if (buf.ptr) {
        // Earlier work on any stream may still read the old block.
        SCIR_HIP_TRY(hipDeviceSynchronize());
        SCIR_HIP_TRY(hipFree(buf.ptr));
        buf.ptr = nullptr;
        buf.bytes = 0;
    }
    const size_t want = bytes < 4096 ? 4096 : bytes;
    SCIR_TRY(device_alloc(&buf.ptr, want, "host-tier buffer"));
    buf.bytes = want;
    return SCIR_HIP_OK;
}

int ensure_scratch(DeviceState *st, DeviceBuffer &buf, size_t bytes, hipStream_t stream)
{
    if (bytes <= buf.bytes) return SCIR_HIP_OK;
    SCIR_TRY(refuse_growth_in_capture(bytes, stream));
    host_alloc_point();
    st->retired.reserve(st->retired.size() + 1);
    size_t want = std::max<size_t>(bytes, 2 * buf.bytes);
    if (want < 4096) want = 4096;
    void *p = nullptr;
    SCIR_TRY(device_alloc(&p, want, "library scratch"));
    if (buf.ptr) st->retired.push_back(buf.ptr);  // reserved above: cannot throw
    buf.ptr = p;
    buf.bytes = want;
    return SCIR_HIP_OK;
}

StreamScratch &scratch_for(DeviceState *st, hipStream_t stream)
{
    return st->scratch[reinterpret_cast<void *>(stream)];
}

int check_dense_args(const void *x, int64_t ldx, const void *b, int32_t K, int64_t B,
                     int64_t N, const void *y, int64_t ldy)
{
    if (B < 0 || N < 0 || K < 0)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "negative size (B=%lld N=%lld K=%d)",
                    (long long)B, (long long)N, K);
    if (B == 0 || N == 0) return SCIR_HIP_OK;
    if (!x || !y || (K > 0 && !b))
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (ldx < N || ldy < N)
        return fail(SCIR_HIP_ERR_INVALID_ARGUMENT,
                    "row stride smaller than the row length (ldx=%lld ldy=%lld N=%lld)",
                    (long long)ldx, (long long)ldy, (long long)N);
    return SCIR_HIP_OK;
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(SCIR_HIP_ERR_BACKEND, "launch of %s failed: %s", what, hipGetErrorString(e));
    return SCIR_HIP_OK;
}

int allow_dynamic_lds(const void *fn, int bytes)
{
    // The attribute belongs to the kernel on one device: remember the largest
    // value set per (device, kernel).
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, int> done;
    int dev = -1;
    SCIR_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    int &have = done[{dev, fn}];
    if (have < bytes) {
        SCIR_HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        have = bytes;
    }
    return SCIR_HIP_OK;
}

// ---- small kernels -------------------------------------------------------

__global__ void fill_zero_kernel(float *__restrict__ y, int64_t ldy, int64_t N,
                                 int64_t total)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = i / N, c = i - r * N;
        y[r * ldy + c] = 0.0f;
    }
}

int fill_zero_2d(float *y, int64_t ldy, int64_t B, int64_t N, hipStream_t stream)
{
    int64_t total = B * N;
    if (total == 0) return SCIR_HIP_OK;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, y, ldy,
                       N, total);
    return check_launch("fill_zero_kernel");
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based N(0,1): sample (r, i) hashes key_r + i*C with the SplitMix64
// finaliser; two 24-bit uniforms feed Box-Muller (cosine branch).  A second,
// independent hash decides the outlier set.
__global__ void synth_normal_kernel(float *__restrict__ x, int64_t ldx, int64_t N,
                                    int64_t total, uint64_t seed, float frac, float scale)
{
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = idx / N, i = idx - r * N;
        uint64_t key = seed ^ ((uint64_t)r * 0x9E3779B97F4A7C15ull);
        uint64_t h = mix64(key + (uint64_t)i * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull);
        float u1 = ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
        float u2 = (float)((h >> 16) & 0xFFFFFFull) * (1.0f / 16777216.0f);
        float z = sqrtf(-2.0f * logf(u1)) * cosf(6.283185307179586f * u2);
        if (frac > 0.0f) {
            uint64_t h2 = mix64(h ^ 0x5851F42D4C957F2Dull);
            float u3 = (float)(h2 >> 40) * (1.0f / 16777216.0f);
            if (u3 < frac) z *= scale;
        }
        x[r * ldx + i] = z;
    }
}

}  // namespace scir

using namespace scir;

extern "C" {

int scir_hip_device_count(int *count)
{
    return guarded(__func__, [&]() -> int {
        if (!count) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "count is null");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
        return SCIR_HIP_OK;
    });
}

int scir_hip_init(void)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        DeviceState *st = nullptr;
        return current_device_state(&st);
    });
}

int scir_hip_shutdown(void)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        std::lock_guard<std::mutex> g(g_states_mu);
        int prev = -1;
        (void)hipGetDevice(&prev);
        for (auto &up : g_states) {
            if (!up || !up->ready) continue;
            std::lock_guard<std::mutex> lk(up->mu);
            (void)hipSetDevice(up->device);
            (void)hipDeviceSynchronize();
            auto release = [](DeviceBuffer &b) {
                if (b.ptr) (void)hipFree(b.ptr);
                b.ptr = nullptr;
                b.bytes = 0;
            };
            for (auto &kv : up->scratch) {
                release(kv.second.partials);
                release(kv.second.sos_states);
                release(kv.second.fft_spectrum);
            }
            up->scratch.clear();
            for (void *p : up->retired) (void)hipFree(p);
            up->retired.clear();
            for (DeviceBuffer &t : up->fft_twiddles) release(t);
            release(up->host_x);
            release(up->host_y);
            release(up->host_b);
            release(up->host_z);
            for (void *&p : up->pinned) {
                if (p) (void)hipHostFree(p);
                p = nullptr;
            }
            up->pinned_bytes = 0;
            if (up->host_stream) (void)hipStreamDestroy(up->host_stream);
            up->host_stream = nullptr;
            for (hipStream_t &cs : up->host_copy_stream) {
                if (cs) (void)hipStreamDestroy(cs);
                cs = nullptr;
            }
        }
        if (prev >= 0) (void)hipSetDevice(prev);
        return SCIR_HIP_OK;
    });
}

int scir_hip_synth_normal_f32(float *x, int64_t ldx, int64_t B, int64_t N, uint64_t seed,
                              float outlier_frac, float outlier_scale, void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (B < 0 || N < 0 || ldx < N)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "bad shape for synth (B=%lld N=%lld ldx=%lld)",
                        (long long)B, (long long)N, (long long)ldx);
        if (B == 0 || N == 0) return SCIR_HIP_OK;
        if (!x) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "x is null");
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        int64_t total = B * N;
        int64_t blocks = (total + 255) / 256;
        int64_t cap = (int64_t)st->cu_count * 64;
        if (blocks > cap) blocks = cap;
        hipLaunchKernelGGL(synth_normal_kernel, dim3((unsigned)blocks), dim3(256), 0,
                           as_stream(stream), x, ldx, N, total, seed, outlier_frac, outlier_scale);
        return check_launch("synth_normal_kernel");
    });
}

// ---- HIP graphs -------------------------------------------------------------

struct scir_hip_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int scir_hip_graph_begin(void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!stream)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "graph capture needs a non-null stream");
        DeviceState *st = nullptr;
        SCIR_TRY(current_device_state(&st));
        SCIR_HIP_TRY(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
        return SCIR_HIP_OK;
    });
}

int scir_hip_graph_end(void *stream, scir_hip_graph *graph)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!stream || !graph)
            return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "null stream or graph handle");
        *graph = nullptr;
        hipGraph_t g = nullptr;
        SCIR_HIP_TRY(hipStreamEndCapture(as_stream(stream), &g));
        std::unique_ptr<scir_hip_graph_s> h;
        try {
            host_alloc_point();
            h.reset(new scir_hip_graph_s);
        } catch (...) {
            (void)hipGraphDestroy(g);
            throw;
        }
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(SCIR_HIP_ERR_BACKEND, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
        }
        h->graph = g;
        h->exec = ex;
        *graph = h.release();
        return SCIR_HIP_OK;
    });
}

int scir_hip_graph_launch(scir_hip_graph graph, void *stream)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!graph) return fail(SCIR_HIP_ERR_INVALID_ARGUMENT, "graph is null");
        SCIR_HIP_TRY(hipGraphLaunch(graph->exec, as_stream(stream)));
        return SCIR_HIP_OK;
    });
}

int scir_hip_graph_destroy(scir_hip_graph graph)
{
    return guarded(__func__, [&]() -> int {
        clear_error();
        if (!graph) return SCIR_HIP_OK;
        if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
        if (graph->graph) (void)hipGraphDestroy(graph->graph);
        delete graph;
        return SCIR_HIP_OK;
    });
}

}  // extern "C"
