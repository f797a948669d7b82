// fdcl_device.cpp -- the class layer's bridge to the C ABI: device plans, host <-> device copies,
// and the Euler-angle helpers shared by the SO(3) classes.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>

#include "fdcl_FFTSO3.hpp"
#include "fdcl_sampling.hpp"

namespace fdcl
{

// src/misc_matrix_func.cpp:144-167: beta = acos(R33); at beta = 0 only alpha + gamma is defined
// and is split evenly
std::vector<double> R2Euler323(const Matrix3d &R)
{
    double a, b, g;
    b = std::acos(R(2, 2));
    if (b != 0.) {
        g = std::atan2(R(2, 1), -R(2, 0));
        a = std::atan2(R(1, 2), R(0, 2));
    } else {
        a = std::acos(R(0, 0)) / 2.;
        g = a;
    }
    return {a, b, g};
}

// R = exp(a e3) exp(b e2) exp(g e3) (src/misc_matrix_func.cpp:169-178)
Matrix3d Euler3232R(double a, double b, double g)
{
    const double ca = std::cos(a), sa = std::sin(a), cb = std::cos(b), sb = std::sin(b);
    const double cg = std::cos(g), sg = std::sin(g);
    Matrix3d R;
    R(0, 0) = ca * cb * cg - sa * sg;
    R(0, 1) = -cg * sa - ca * cb * sg;
    R(0, 2) = ca * sb;
    R(1, 0) = ca * sg + cb * cg * sa;
    R(1, 1) = ca * cg - cb * sa * sg;
    R(1, 2) = sa * sb;
    R(2, 0) = -cg * sb;
    R(2, 1) = sb * sg;
    R(2, 2) = cb;
    return R;
}

namespace detail
{

static void check_abi(int st, const char *what)
{
    if (st != FFTSO3_OK)
        throw std::runtime_error(std::string(what) + ": " + fftso3_status_string(st) + ": " + fftso3_last_error());
}

static void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

DevicePlan::~DevicePlan()
{
    if (stream_) (void)hipStreamSynchronize(static_cast<hipStream_t>(stream_));
    if (plan_) fftso3_plan_destroy(plan_);
    if (d_grid_) (void)hipFree(d_grid_);
    if (d_coef_) (void)hipFree(d_coef_);
    if (d_ang_) (void)hipFree(d_ang_);
    if (d_val_) (void)hipFree(d_val_);
    for (int b = 0; b < 2; b++) {
        if (pin_[b]) (void)hipHostFree(pin_[b]);
        if (ev_[b]) (void)hipEventDestroy(static_cast<hipEvent_t>(ev_[b]));
    }
    if (stream_) (void)hipStreamDestroy(static_cast<hipStream_t>(stream_));
}

void DevicePlan::prepare(int domain, int l_max, int n_fn, int kind, int precision)
{
    const int key[5] = {domain, l_max, n_fn, kind, precision};
    if (plan_ && std::memcmp(key, key_, sizeof(key)) == 0) return;
    if (plan_) {
        fftso3_plan_destroy(plan_);
        plan_ = nullptr;
    }
    if (d_grid_) (void)hipFree(d_grid_);
    if (d_coef_) (void)hipFree(d_coef_);
    d_grid_ = d_coef_ = nullptr;
    check_abi(fftso3_plan_create(&plan_, domain, l_max, n_fn, kind, precision, 0), "fftso3_plan_create");
    std::memcpy(key_, key, sizeof(key));
    grid_elems_ = fftso3_plan_grid_elems(plan_);
    coef_elems_ = fftso3_plan_coef_elems(plan_);
    const size_t rsz = precision == FFTSO3_FP64 ? 8 : 4;  // real scalar of device storage
    const size_t gsz = grid_elems_ * rsz * (kind == FFTSO3_COMPLEX ? 2 : 1);
    const size_t csz = coef_elems_ * rsz * (kind == FFTSO3_COMPLEX ? 2 : 1);
    check_hip(hipMalloc(&d_grid_, gsz), "hipMalloc(grid)");
    check_hip(hipMalloc(&d_coef_, csz), "hipMalloc(coefficients)");
}

// copy n doubles from host to device storage of precision `prec` (float storage converts)
static void upload(const double *src, long n, void *dst, int prec, std::vector<char> &stage)
{
    if (prec == FFTSO3_FP64) {
        check_hip(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
        return;
    }
    stage.resize(n * sizeof(float));
    float *f = reinterpret_cast<float *>(stage.data());
    for (long i = 0; i < n; i++) f[i] = (float)src[i];
    check_hip(hipMemcpy(dst, f, n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

static void download(const void *src, long n, double *dst, int prec, std::vector<char> &stage)
{
    if (prec == FFTSO3_FP64) {
        check_hip(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
        return;
    }
    stage.resize(n * sizeof(float));
    float *f = reinterpret_cast<float *>(stage.data());
    check_hip(hipMemcpy(f, src, n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    for (long i = 0; i < n; i++) dst[i] = f[i];
}

void DevicePlan::forward(int domain, int l_max, int n_fn, int kind, int precision, const void *grid, void *coef)
{
    prepare(domain, l_max, n_fn, kind, precision);
    const int w = kind == FFTSO3_COMPLEX ? 2 : 1;
    upload(static_cast<const double *>(grid), grid_elems_ * w, d_grid_, precision, stage_);
    check_abi(fftso3_forward(plan_, d_grid_, d_coef_, nullptr), "fftso3_forward");
    check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    download(d_coef_, coef_elems_ * w, static_cast<double *>(coef), precision, stage_);
}

void DevicePlan::rotate_s2(int l_max, int kind, int precision, const void *F, const double *angles, long n, void *G)
{
    if (n <= 0) return;
    prepare(FFTSO3_DOMAIN_SO3, l_max, 1, kind, precision);
    const size_t w = kind == FFTSO3_COMPLEX ? 2 : 1;
    const size_t nc = (size_t)(l_max + 1) * (l_max + 1) * w;
    void *dF = nullptr, *dA = nullptr, *dG = nullptr;
    auto release = [&]() {
        if (dF) (void)hipFree(dF);
        if (dA) (void)hipFree(dA);
        if (dG) (void)hipFree(dG);
    };
    try {
        check_hip(hipMalloc(&dF, nc * sizeof(double)), "hipMalloc(rotation input)");
        check_hip(hipMalloc(&dA, (size_t)n * 3 * sizeof(double)), "hipMalloc(rotation angles)");
        check_hip(hipMalloc(&dG, (size_t)n * nc * sizeof(double)), "hipMalloc(rotation output)");
        check_hip(hipMemcpy(dF, F, nc * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
        check_hip(hipMemcpy(dA, angles, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
        check_abi(fftso3_rotate_s2(plan_, static_cast<const double *>(dF), static_cast<const double *>(dA), n,
                                   static_cast<double *>(dG), nullptr),
                  "fftso3_rotate_s2");
        check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        check_hip(hipMemcpy(G, dG, (size_t)n * nc * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    } catch (...) {
        release();
        throw;
    }
    release();
}

void DevicePlan::inverse(int domain, int l_max, int n_fn, int kind, int precision, const void *coef, void *grid)
{
    prepare(domain, l_max, n_fn, kind, precision);
    const int w = kind == FFTSO3_COMPLEX ? 2 : 1;
    upload(static_cast<const double *>(coef), coef_elems_ * w, d_coef_, precision, stage_);
    check_abi(fftso3_inverse(plan_, d_coef_, d_grid_, nullptr), "fftso3_inverse");
    check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    download(d_grid_, grid_elems_ * w, static_cast<double *>(grid), precision, stage_);
}

void DevicePlan::forward_fill(int domain, int l_max, int kind, int precision, const Fill &fill, void *coef)
{
    prepare(domain, l_max, 1, kind, precision);
    const int w = kind == FFTSO3_COMPLEX ? 2 : 1;      // doubles per grid value
    const bool fp64 = precision == FFTSO3_FP64;
    const size_t rsz = fp64 ? 8 : 4;                    // device scalar
    const long nv = grid_elems_;                        // grid values of the function
    // chunks of whole 1024-value pieces, about 4 MB of device data each
    const long piece = 1024;
    const long chunk = std::max(piece, ((long)(4 << 20) / (long)(w * rsz)) / piece * piece);
    const size_t bytes = (size_t)std::min(chunk, nv) * w * rsz;
    if (!stream_) {
        hipStream_t st;
        check_hip(hipStreamCreate(&st), "hipStreamCreate");
        stream_ = st;
        for (int b = 0; b < 2; b++) {
            hipEvent_t e;
            check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            ev_[b] = e;
        }
    }
    hipStream_t st = static_cast<hipStream_t>(stream_);
    if (bytes > pin_bytes_) {
        check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
        for (int b = 0; b < 2; b++) {
            if (pin_[b]) (void)hipHostFree(pin_[b]);
            pin_[b] = nullptr;
        }
        pin_bytes_ = 0;
        for (int b = 0; b < 2; b++) check_hip(hipHostMalloc(&pin_[b], bytes, hipHostMallocDefault), "hipHostMalloc");
        pin_bytes_ = bytes;
    }
    char *dg = static_cast<char *>(d_grid_);
    long it = 0;
    for (long c0 = 0; c0 < nv; c0 += chunk, it++) {
        const int b = (int)(it & 1);
        if (it >= 2) check_hip(hipEventSynchronize(static_cast<hipEvent_t>(ev_[b])), "hipEventSynchronize");
        const long n = std::min(chunk, nv - c0);
        void *buf = pin_[b];
        sample_chunk(fill, c0, n, piece, w, fp64, buf);
        check_hip(hipMemcpyAsync(dg + (size_t)c0 * w * rsz, buf, (size_t)n * w * rsz, hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync H2D");
        check_hip(hipEventRecord(static_cast<hipEvent_t>(ev_[b]), st), "hipEventRecord");
    }
    check_abi(fftso3_forward(plan_, d_grid_, d_coef_, st), "fftso3_forward");
    check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    download(d_coef_, coef_elems_ * w, static_cast<double *>(coef), precision, stage_);
}

void DevicePlan::evaluate(int domain, int l_max, int kind, int precision, const void *coef, const double *angles,
                          long n, void *values)
{
    prepare(domain, l_max, 1, kind, precision);
    if (n <= 0) return;
    const int per = domain == FFTSO3_DOMAIN_S2 ? 2 : 3;
    if (n > pts_cap_) {
        if (d_ang_) (void)hipFree(d_ang_);
        if (d_val_) (void)hipFree(d_val_);
        d_ang_ = d_val_ = nullptr;
        check_hip(hipMalloc(&d_ang_, n * per * sizeof(double)), "hipMalloc(angles)");
        check_hip(hipMalloc(&d_val_, n * 2 * sizeof(double)), "hipMalloc(values)");
        pts_cap_ = n;
    }
    const int w = kind == FFTSO3_COMPLEX ? 2 : 1;
    upload(static_cast<const double *>(coef), coef_elems_ * w, d_coef_, precision, stage_);
    check_hip(hipMemcpy(d_ang_, angles, n * per * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
    check_abi(fftso3_evaluate(plan_, 0, d_coef_, static_cast<const double *>(d_ang_), n, d_val_, nullptr),
              "fftso3_evaluate");
    check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    download(d_val_, n * w, static_cast<double *>(values), precision, stage_);
}

}  // namespace detail
}  // namespace fdcl
