// Exercises the C++ host mirror (scir_amd/host/scir_gpu.hpp) the way the
// reference's own tests exercise scir_gpu: the known-answer test
// (crates/scir-gpu/src/lib.rs:1249-1261), the auto-dispatch CPU check
// (crates/scir-signal/src/lib.rs:670-696), strided input, and the HIP arm.
// Usage: host_mirror_test [--expect-device | --expect-no-device]
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../scir_amd/host/scir_gpu.hpp"

using scir_gpu::Array2;
using scir_gpu::ArrayView2;
using scir_gpu::Device;

static int failures = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
            ++failures;                                                  \
        }                                                                \
    } while (0)

static Array2<float> kat_x()
{
    Array2<float> x(2, 4);
    const float v[8] = {1, 2, 3, 4, 0.5f, 0, -0.5f, -1};
    std::memcpy(x.data.data(), v, sizeof v);
    return x;
}

int main(int argc, char **argv)
{
    const bool want_dev = argc > 1 && !std::strcmp(argv[1], "--expect-device");
    const std::vector<float> taps = {0.25f, 0.5f, 0.25f};
    const float expect[8] = {0.25f, 1, 2, 3, 0.125f, 0.25f, 0, -0.5f};
    Array2<float> x = kat_x();

    Array2<float> y = scir_gpu::fir1d_batched_f32(x.view(), taps);
    for (int i = 0; i < 8; ++i) CHECK(std::fabs(y.data[i] - expect[i]) <= 1e-7 + 1e-7 * std::fabs(expect[i]));

    Array2<float> ya = scir_signal::gpu::fir1d_batched_f32(x.view(), taps, Device::Cpu);
    CHECK(ya.data == y.data);  // tol 0: same loop

    // column-major view of the same values (non-standard layout)
    std::vector<float> colmajor(8);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 4; ++j) colmajor[j * 2 + i] = x(i, j);
    ArrayView2<float> cv{colmajor.data(), 2, 4, 1, 2};
    CHECK(scir_gpu::fir1d_batched_f32(cv, taps).data == y.data);

    bool threw = false;
    try {
        Array2<float> yh = scir_gpu::fir1d_batched_f32_auto(x.view(), taps, Device::Hip);
        for (int i = 0; i < 8; ++i) CHECK(yh.data[i] == expect[i]);
    } catch (const scir_gpu::GpuError &e) {
        threw = true;
        CHECK(e.kind() == scir_gpu::GpuError::Kind::BackendUnavailable);
        std::printf("hip arm: %s\n", e.what());
    }
    CHECK(threw != want_dev);

    // sosfilt: a one-pole section y[n] = 0.5 x[n] + 0.5 y[n-1] (the CPU arm is
    // the reference loop); the HIP arm gives the same bits (reference order
    // for rows this short) or, without a device, throws.
    const scir_signal::Sos sos = {{0.5, 0.0, 0.0, 1.0, -0.5, 0.0}};
    const std::vector<double> imp = {1.0, 0.0, 0.0, 0.0};
    std::vector<double> ys = scir_signal::sosfilt(sos, imp);
    CHECK(ys.size() == 4 && ys[0] == 0.5 && ys[1] == 0.25 && ys[2] == 0.125 && ys[3] == 0.0625);
    std::vector<double> yf = scir_signal::filtfilt(sos, imp);
    CHECK(yf.size() == 4 && yf[0] > 0.0);
    const std::vector<double> h = {0.25, 0.5, 0.25};
    std::vector<double> yr = scir_signal::resample_poly({1.0, 2.0, 3.0, 4.0, 5.0, 6.0}, 2, 3, h);
    CHECK(!yr.empty());
    bool threw2 = false;
    try {
        CHECK(scir_signal::sosfilt(sos, imp, Device::Hip) == ys);
        CHECK(scir_signal::filtfilt(sos, imp, Device::Hip) == yf);
        CHECK(scir_signal::resample_poly({1.0, 2.0, 3.0, 4.0, 5.0, 6.0}, 2, 3, h, Device::Hip)
                  .size() == yr.size());
        auto gr = scir_gpu::fir1d_batched_f32_vjp_hip(x.view(), taps, x.view());
        // dL/dtaps[j] = sum_i g[i] x[i-(K-1-j)] with g = x
        CHECK(gr.second.size() == 3 && std::fabs(gr.second[2] - 31.5f) < 1e-4f);
    } catch (const scir_gpu::GpuError &e) {
        threw2 = true;
        CHECK(e.kind() == scir_gpu::GpuError::Kind::BackendUnavailable);
    }
    CHECK(threw2 != want_dev);
    if (failures)
        std::printf("host mirror: %d failures\n", failures);
    else
        std::printf("host mirror: ok\n");
    return failures ? 1 : 0;
}
