// capi_stats.cpp -- the statistics of src/statistics/mod.rs over values
// resident in device memory: mean (mod.rs:321-335), mean_and_variance in the
// two-pass form of the sequential functions (mod.rs:375-399) or the one-pass
// form of the *_par_iter functions (mod.rs:163-181), mean_with_error
// (mod.rs:189-207,401-405) and covariance (mod.rs:427-458, one-pass form
// mod.rs:279-306).  The sums run on the device (k_stats.hip); the final
// arithmetic is the reference's, in double on the host.  Arithmetic on
// complex values is num-complex's, with no conjugation: the complex
// "variance" is the mean of (x - m)(x - m).
#include <cmath>
#include <complex>
#include <cstdint>
#include <mutex>

#include "capi_util.hpp"
#include "internal.hpp"

using namespace lqcd;

// products and quotients rounded one operation at a time, as in the reference
#pragma clang fp contract(off)

constexpr int kStatsSlots = 6;   // doubles per block partial (complex covariance)
// doubles of one call's device scratch: the block partials (at most
// kMaxReduceBlocks blocks, stats_blocks caps the grid there for any n), then
// two vectors of 8 sums
constexpr size_t kScratchDoubles = (size_t)kMaxReduceBlocks * kStatsSlots + 16;
// largest n of the allocation-free entry point: element indices (2 i for
// complex values) stay within int64
constexpr size_t kNoAllocMaxValues = (size_t)1 << 62;

// The fixed scratch of lqcd_stats_mean_and_variance_noalloc: one copy per
// GPU, held until the process ends (192 KiB + 128 B).  In host code this name
// is a host-side stand-in, not a device address: each device's address comes
// from hipGetSymbolAddress (resolve_fixed_scratch, once per device with that
// device current), and only that address reaches kernels and copies.
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

using cd = std::complex<double>;

// num-complex product, component by component as the reference computes it
cd cmul(cd a, cd b)
{
    return {a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()};
}
cd cdiv(cd a, double d) { return {a.real() / d, a.imag() / d}; }

int check_form(int form)
{
    if (form != LQCD_STATS_SEQUENTIAL && form != LQCD_STATS_PARALLEL)
        return fail(LQCD_ERR_INVALID_ARGUMENT, "unknown statistics form");
    return LQCD_OK;
}

// the partials and the two sum vectors of one call, kept on the context
// (allocated on the context's first statistics call)
int ctx_scratch(lqcd_ctx* ctx, double** partials, double** sums)
{
    if (!ctx->stats_scratch) HIP_TRY(hipMalloc((void**)&ctx->stats_scratch, sizeof(double) * kScratchDoubles));
    *partials = ctx->stats_scratch;
    *sums = ctx->stats_scratch + kMaxReduceBlocks * kStatsSlots;
    return LQCD_OK;
}

// per-device addresses of the fixed scratch
constexpr int kMaxDevices = 64;
std::mutex fixed_mu;
double* fixed_ptr[kMaxDevices] = {};

// resolves the address of `scratch` on `device`, which must be current
hipError_t resolve_fixed_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(fixed_mu);
    if (fixed_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    fixed_ptr[device] = (double*)p;
    return hipSuccess;
}

double* fixed_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(fixed_mu);
    return fixed_ptr[device];
}

// element k of a sum vector read back from the device
cd sum_at(const double* h, int k, bool cplx) { return cplx ? cd(h[2 * k], h[2 * k + 1]) : cd(h[k], 0.0); }

void put(cd v, bool cplx, double* out)
{
    out[0] = v.real();
    if (cplx) out[1] = v.imag();
}

// runs op over n values and copies its sums (and, for the two-pass variance,
// the centred sum after them) to h
int device_sums_in(lqcd_ctx* ctx, double* partials, double* sums, int op, bool cplx, const double* x,
                   const double* y, size_t n, bool centred, double* h)
{
    const int nb = stats_blocks((int64_t)n);
    HIP_TRY(launch_stats(ctx, op, cplx, x, y, (int64_t)n, nullptr, partials, nb, sums));
    if (centred)
        HIP_TRY(launch_stats(ctx, STATS_CENTERED, cplx, x, nullptr, (int64_t)n, sums, partials, nb, sums + 8));
    HIP_TRY(hipMemcpyAsync(h, sums, sizeof(double) * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return LQCD_OK;
}

int device_sums(lqcd_ctx* ctx, int op, bool cplx, const double* x, const double* y, size_t n, bool centred,
                double* h)
{
    double *partials, *sums;
    if (int rc = ctx_scratch(ctx, &partials, &sums)) return rc;
    return device_sums_in(ctx, partials, sums, op, cplx, x, y, n, centred, h);
}

// mean and variance of n values; fixed = the partials and sums live in this
// device's copy of `scratch` instead of the context's scratch
int mean_variance(lqcd_ctx* ctx, const double* x, size_t n, bool cplx, int form, cd* mean, cd* var,
                  bool fixed = false)
{
    CHECK_ARG(ctx, "ctx is null");
    if (int rc = check_form(form)) return rc;
    // (len - 1) underflows in the reference for an empty input
    CHECK_ARG(n > 0, "variance of an empty sequence");
    CHECK_ARG(x, "values pointer is null");
    DeviceGuard dg(ctx->device);
    double h[16] = {};
    const bool two_pass = form == LQCD_STATS_SEQUENTIAL;
    if (fixed) {
        CHECK_ARG(n <= kNoAllocMaxValues, "more values than lqcd_stats_mean_and_variance_noalloc accepts (2^62)");
        double* base = fixed_scratch(ctx->device);
        if (!base) {
            const hipError_t e = resolve_fixed_scratch(ctx->device);
            if (e != hipSuccess)
                return fail(LQCD_ERR_DEVICE, std::string("fixed statistics scratch: ") + hipGetErrorString(e));
            base = fixed_scratch(ctx->device);
        }
        // the grid never has more blocks than the scratch holds partials for
        CHECK_ARG((size_t)stats_blocks((int64_t)n) * kStatsSlots + 16 <= kScratchDoubles,
                  "statistics grid larger than the fixed scratch");
        if (int rc = device_sums_in(ctx, base, base + (size_t)kMaxReduceBlocks * kStatsSlots, STATS_MOMENTS, cplx,
                                    x, nullptr, n, two_pass, h))
            return rc;
    } else if (int rc = device_sums(ctx, STATS_MOMENTS, cplx, x, nullptr, n, two_pass, h)) {
        return rc;
    }
    const double len = (double)n;
    const cd s = sum_at(h, 0, cplx);
    *mean = cdiv(s, len);
    if (two_pass) {
        *var = cdiv(sum_at(h + 8, 0, cplx), (double)(n - 1));
    } else {
        const cd s2 = sum_at(h, 1, cplx);
        *var = cdiv(s2 - cdiv(cmul(s, s), len), (double)(n - 1));
    }
    return LQCD_OK;
}

} // namespace

// resolves this device's copy of the fixed statistics scratch at context
// creation (the device is current); a failure here is reported again by the
// first allocation-free call, which retries
hipError_t lqcd::stats_fixed_scratch_init(int device) { return resolve_fixed_scratch(device); }

extern "C" {

int lqcd_stats_mean(lqcd_ctx* ctx, const double* d_values, size_t n, int complex_values, double* mean)
{
    CHECK_ARG(ctx && mean, "null argument");
    const bool cplx = complex_values != 0;
    if (n == 0) {
        // sum of nothing over a length of 0
        put(cdiv(cd(0.0, 0.0), 0.0), cplx, mean);
        return LQCD_OK;
    }
    CHECK_ARG(d_values, "values pointer is null");
    DeviceGuard dg(ctx->device);
    double h[16] = {};
    if (int rc = device_sums(ctx, STATS_MOMENTS, cplx, d_values, nullptr, n, false, h)) return rc;
    put(cdiv(sum_at(h, 0, cplx), (double)n), cplx, mean);
    return LQCD_OK;
}

int lqcd_stats_mean_and_variance(lqcd_ctx* ctx, const double* d_values, size_t n, int complex_values, int form,
                                 double* out)
{
    CHECK_ARG(out, "null argument");
    const bool cplx = complex_values != 0;
    cd m, v;
    if (int rc = mean_variance(ctx, d_values, n, cplx, form, &m, &v)) return rc;
    put(m, cplx, out);
    put(v, cplx, out + (cplx ? 2 : 1));
    return LQCD_OK;
}

int lqcd_stats_mean_and_variance_noalloc(lqcd_ctx* ctx, const double* d_values, size_t n, int complex_values,
                                         int form, double* out)
{
    CHECK_ARG(out, "null argument");
    const bool cplx = complex_values != 0;
    cd m, v;
    if (int rc = mean_variance(ctx, d_values, n, cplx, form, &m, &v, true)) return rc;
    put(m, cplx, out);
    put(v, cplx, out + (cplx ? 2 : 1));
    return LQCD_OK;
}

int lqcd_stats_mean_with_error(lqcd_ctx* ctx, const double* d_values, size_t n, int form, double out[2])
{
    CHECK_ARG(out, "null argument");
    cd m, v;
    if (int rc = mean_variance(ctx, d_values, n, false, form, &m, &v)) return rc;
    out[0] = m.real();
    out[1] = std::sqrt(v.real() / (double)n);
    return LQCD_OK;
}

int lqcd_stats_covariance(lqcd_ctx* ctx, const double* d_x, size_t nx, const double* d_y, size_t ny,
                          int complex_values, int form, double* out, int* has_value)
{
    CHECK_ARG(ctx && out && has_value, "null argument");
    if (int rc = check_form(form)) return rc;
    const bool cplx = complex_values != 0;
    *has_value = 0;
    if (nx != ny) return LQCD_OK;   // the reference's None
    *has_value = 1;
    double h[16] = {};
    if (nx > 0) {
        CHECK_ARG(d_x && d_y, "values pointer is null");
        DeviceGuard dg(ctx->device);
        if (int rc = device_sums(ctx, STATS_COV, cplx, d_x, d_y, nx, false, h)) return rc;
    }
    const double len = (double)nx;
    const cd sx = sum_at(h, 0, cplx), sy = sum_at(h, 1, cplx), sp = sum_at(h, 2, cplx);
    cd c;
    if (form == LQCD_STATS_SEQUENTIAL)
        c = cdiv(sp, len) - cmul(cdiv(sx, len), cdiv(sy, len));
    else
        c = cdiv(sp - cdiv(cmul(sx, sy), len), len);
    put(c, cplx, out);
    return LQCD_OK;
}

} // extern "C"
