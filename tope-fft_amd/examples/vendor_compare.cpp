// Vendor comparator (run-only): the same device-resident transforms through
// libtopefft and through rocFFT (ROCm's FFT library, in this image), timed
// with HIP events on one stream, and the two results compared.
//
//   usage: vendor_compare [steps] [workload ...]
//   workloads: 2d4096 3d256 1d65536 batch64 batch32 3d512 (default: all but 3d512)
//
// One JSON object per workload on stdout:
//   {"workload": ..., "topefft_us": ..., "rocfft_us": ..., "rel_l2": ...}
// rel_l2 = ||topefft - rocfft|| / ||rocfft|| of one forward transform of the
// same input.  Timed region: `steps` forward transforms in place, back to
// back, between two events (inputs are scaled so repeated transforms stay
// finite).  Both libraries work in place on their own copy of the input.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "topefft.h"

#define HIPCHECK(x)                                                                  \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(2);                                                            \
        }                                                                            \
    } while (0)
#define RFCHECK(x)                                                                   \
    do {                                                                             \
        rocfft_status s_ = (x);                                                      \
        if (s_ != rocfft_status_success) {                                           \
            std::fprintf(stderr, "%s:%d rocfft status %d\n", __FILE__, __LINE__, (int)s_); \
            std::exit(3);                                                            \
        }                                                                            \
    } while (0)

struct Workload {
    std::string name;
    int dims;             // 1, 2, 3
    size_t len[3];        // x (fastest), y, z
    size_t batch;         // transforms of len[0] (dims == 1)
    bool single;          // complex64
};

template <typename T>
static double rel_l2(const std::vector<std::complex<T>> &a, const std::vector<std::complex<T>> &b)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        const std::complex<double> d = std::complex<double>(a[i]) - std::complex<double>(b[i]);
        num += std::norm(d);
        den += std::norm(std::complex<double>(b[i]));
    }
    return std::sqrt(num / (den > 0 ? den : 1));
}

template <typename T>
static void run(const Workload &w, int steps, hipStream_t st)
{
    const size_t n = w.len[0] * w.len[1] * w.len[2] * w.batch;
    const size_t bytes = n * sizeof(std::complex<T>);
    // uniform [-1, 1) re/im, scaled so `steps` unnormalised forwards stay finite
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const double scale = w.single ? 0x1p-60 : 1.0;
    std::vector<std::complex<T>> host(n);
    for (auto &v : host) v = std::complex<T>((T)(u(rng) * scale), (T)(u(rng) * scale));

    void *d_top = nullptr, *d_roc = nullptr;
    HIPCHECK(hipMalloc(&d_top, bytes));
    HIPCHECK(hipMalloc(&d_roc, bytes));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));

    // ---- libtopefft ----
    struct topeFFT f;
    topeFFTInit(&f);
    if (f.error) std::exit(4);
    topeFFTSetStream(&f, st);
    struct topePlan1D p1;
    struct topePlan2D p2;
    struct topePlan3D p3;
    struct topePlanBatch pb;
    void *impl = nullptr;
    const int prec = w.single ? TOPE_PREC_F32 : TOPE_PREC_F64;
    if (w.dims == 1 && w.batch > 1) {
        topeBatchPlanInit(&f, &pb, (int)w.len[0], (int)w.batch, prec, TOPE_MODE_VALU, host.data());
        impl = pb.impl;
    } else if (w.dims == 1) {
        tope1DPlanInit(&f, &p1, (int)w.len[0], C2C, reinterpret_cast<double *>(host.data()));
        impl = p1.impl;
    } else if (w.dims == 2) {
        tope2DPlanInit(&f, &p2, (int)w.len[0], (int)w.len[1], C2C, reinterpret_cast<double *>(host.data()));
        impl = p2.impl;
    } else {
        tope3DPlanInit(&f, &p3, (int)w.len[0], (int)w.len[1], (int)w.len[2], C2C,
                       reinterpret_cast<double *>(host.data()));
        impl = p3.impl;
    }
    if (f.error) {
        std::fprintf(stderr, "topefft plan %s: %s\n", w.name.c_str(), topeErrorString(f.error));
        std::exit(4);
    }

    // ---- rocFFT ----
    rocfft_plan rp = nullptr;
    const size_t rlen[3] = {w.len[0], w.len[1], w.len[2]};
    RFCHECK(rocfft_plan_create(&rp, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                               w.single ? rocfft_precision_single : rocfft_precision_double, (size_t)w.dims, rlen,
                               w.batch, nullptr));
    size_t wbytes = 0;
    RFCHECK(rocfft_plan_get_work_buffer_size(rp, &wbytes));
    void *work = nullptr;
    rocfft_execution_info info = nullptr;
    RFCHECK(rocfft_execution_info_create(&info));
    RFCHECK(rocfft_execution_info_set_stream(info, st));
    if (wbytes) {
        HIPCHECK(hipMalloc(&work, wbytes));
        RFCHECK(rocfft_execution_info_set_work_buffer(info, work, wbytes));
    }
    void *bufs[1] = {d_roc};

    // ---- one forward each on the same input: agreement ----
    HIPCHECK(hipMemcpy(d_top, host.data(), bytes, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_roc, host.data(), bytes, hipMemcpyHostToDevice));
    topePlanExecPtr(&f, impl, d_top, FORWARD);
    RFCHECK(rocfft_execute(rp, bufs, nullptr, info));
    HIPCHECK(hipStreamSynchronize(st));
    if (f.error) std::exit(4);
    std::vector<std::complex<T>> a(n), b(n);
    HIPCHECK(hipMemcpy(a.data(), d_top, bytes, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(b.data(), d_roc, bytes, hipMemcpyDeviceToHost));
    const double rel = rel_l2(a, b);

    // ---- timing: warm-up, then `steps` forwards between two events ----
    auto time_us = [&](auto &&one) {
        for (int i = 0; i < 3; ++i) one();
        HIPCHECK(hipEventRecord(e0, st));
        for (int i = 0; i < steps; ++i) one();
        HIPCHECK(hipEventRecord(e1, st));
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        return 1e3 * ms / steps;
    };
    HIPCHECK(hipMemcpy(d_top, host.data(), bytes, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_roc, host.data(), bytes, hipMemcpyHostToDevice));
    const double t_top = time_us([&] { topePlanExecPtr(&f, impl, d_top, FORWARD); });
    const double t_roc = time_us([&] { RFCHECK(rocfft_execute(rp, bufs, nullptr, info)); });
    if (f.error) std::exit(4);

    std::printf("{\"workload\": \"%s\", \"precision\": \"%s\", \"topefft_us\": %.2f, \"rocfft_us\": %.2f, "
                "\"speedup_vs_rocfft\": %.3f, \"rel_l2\": %.3g}\n",
                w.name.c_str(), w.single ? "complex64" : "complex128", t_top, t_roc, t_roc / t_top, rel);
    std::fflush(stdout);

    RFCHECK(rocfft_execution_info_destroy(info));
    RFCHECK(rocfft_plan_destroy(rp));
    if (work) HIPCHECK(hipFree(work));
    if (w.dims == 1 && w.batch > 1) topeBatchDestroy(&f, &pb);
    else if (w.dims == 1) tope1DDestroy(&f, &p1);
    else if (w.dims == 2) tope2DDestroy(&f, &p2);
    else tope3DDestroy(&f, &p3);
    HIPCHECK(hipFree(d_top));
    HIPCHECK(hipFree(d_roc));
    HIPCHECK(hipEventDestroy(e0));
    HIPCHECK(hipEventDestroy(e1));
}

int main(int argc, char **argv)
{
    const int steps = argc > 1 ? std::atoi(argv[1]) : 20;
    const std::vector<Workload> all = {
        {"2d4096", 2, {4096, 4096, 1}, 1, false},    {"3d256", 3, {256, 256, 256}, 1, false},
        {"1d65536", 1, {65536, 1, 1}, 1, false},     {"batch64", 1, {4096, 1, 1}, 4096, false},
        {"batch32", 1, {4096, 1, 1}, 4096, true},    {"3d512", 3, {512, 512, 512}, 1, false},
    };
    std::vector<std::string> want;
    for (int i = 2; i < argc; ++i) want.push_back(argv[i]);
    if (want.empty()) want = {"2d4096", "3d256", "1d65536", "batch64", "batch32"};
    hipStream_t st;
    HIPCHECK(hipStreamCreate(&st));
    RFCHECK(rocfft_setup());
    for (const auto &name : want) {
        const Workload *w = nullptr;
        for (const auto &c : all)
            if (c.name == name) w = &c;
        if (!w) {
            std::fprintf(stderr, "unknown workload %s\n", name.c_str());
            return 1;
        }
        if (w->single) run<float>(*w, steps, st);
        else run<double>(*w, steps, st);
    }
    RFCHECK(rocfft_cleanup());
    HIPCHECK(hipStreamDestroy(st));
    return 0;
}
