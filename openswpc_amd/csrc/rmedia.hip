// rmedia.hip -- random-media synthesis on the device, the generator of
// src/shared/m_rmedia.f90:40-173 (rmedia__3dgen_ani):
//   uniform noise -> forward FFT -> whitening b/|b| -> filter by
//   sqrt(psdf(k) N dV) -> b(0) = 0 -> inverse FFT -> mean removed, rms scaled
//   to epsil.
// Noise comes from a counter-based generator (one 64-bit mix per cell), so a
// realisation depends on the seed and the grid only, not on launch shapes.
// The FFTs are hipFFT double-precision real transforms of any size (the
// reference's radix-2 FFT, m_rfft.f90:46, needs powers of two); sums are
// deterministic two-pass reductions.  The field is written as float.
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "swpc_internal.h"

namespace {

thread_local char g_err[512] = "";

int rfail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

// splitmix64 finaliser: a bijective mix of a 64-bit counter
__device__ static inline unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_noise(double *r, long long n, unsigned long long seed)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long key = mix64(seed);
    for (long long q = t; q < n; q += (long long)gridDim.x * blockDim.x)
        r[q] = (double)(mix64(key ^ (unsigned long long)q) >> 11) * 0x1.0p-53;  // uniform in [0, 1)
}

struct FilterArgs {
    double2 *b;
    int n1h, n2, n3;           // half spectrum n1/2+1 along the fastest axis
    double k1, k2, k3;         // 2 pi / (n d) per axis
    int n1, n2f, n3f;          // full sizes for the negative wavenumbers
    double a1, a2, a3;
    int ptype;
    double coef, expo, scale;  // psdf = coef * shape(q); scale = N dV
};

// wavenumber of index i of an n-point axis (m_fk.f90 x2k3d)
__device__ static inline double wnum(int i, int n, double k0)
{
    return (i <= n / 2 ? i : i - n) * k0;
}

// whitening and filtering of the half spectrum (m_rmedia.f90:107-140)
__global__ void k_filter(FilterArgs a)
{
    const long long total = (long long)a.n1h * a.n2 * a.n3;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += (long long)gridDim.x * blockDim.x) {
        const int i1 = (int)(q % a.n1h);
        const long long t = q / a.n1h;
        const int i2 = (int)(t % a.n2);
        const int i3 = (int)(t / a.n2);
        if (i1 == 0 && i2 == 0 && i3 == 0) {
            a.b[q] = make_double2(0.0, 0.0);  // DC
            continue;
        }
        const double m1 = wnum(i1, a.n1, a.k1) * a.a1;
        const double m2 = wnum(i2, a.n2f, a.k2) * a.a2;
        const double m3 = wnum(i3, a.n3f, a.k3) * a.a3;
        const double q2 = m1 * m1 + m2 * m2 + m3 * m3;
        double psdf;
        if (a.ptype == 1) psdf = a.coef * exp(-q2 * 0.25);
        else if (a.ptype == 2) psdf = a.coef / ((1.0 + q2) * (1.0 + q2));
        else psdf = a.coef * pow(1.0 + q2, -a.expo);
        const double2 v = a.b[q];
        const double mag = sqrt(v.x * v.x + v.y * v.y);
        const double f = mag > 0.0 ? sqrt(psdf * a.scale) / mag : 0.0;
        a.b[q] = make_double2(v.x * f, v.y * f);
    }
}

// per-block partial sums of x (what = 0) or (x - mean)^2 (what = 1)
__global__ void __launch_bounds__(256) k_partial(const double *x, long long n, double mean, int what, double *part)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
        const double v = x[q] - mean;
        acc += what ? v * v : x[q];
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// one block adds the partials in a fixed order
__global__ void __launch_bounds__(256) k_total(const double *part, int np, double *out)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (int q = threadIdx.x; q < np; q += 256) acc += part[q];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

__global__ void k_finish(const double *x, long long n, const double *stats, double epsil, float *out)
{
    const double mean = stats[0], rms = stats[1];
    const double f = rms > 0.0 ? epsil / rms : 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        out[q] = (float)((x[q] - mean) * f);
}

struct Dev {
    void *p = nullptr;
    ~Dev() { if (p) (void)hipFree(p); }
};

}  // namespace

extern "C" const char *swpc_rmedia_last_error(void) { return g_err; }

#define RCHK(x)                                                                           \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) return rfail(SWPC_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)
#define FCHK(x)                                                                 \
    do {                                                                        \
        hipfftResult r_ = (x);                                                  \
        if (r_ != HIPFFT_SUCCESS) return rfail(SWPC_EHIP, "%s: hipFFT error %d", #x, (int)r_); \
    } while (0)

extern "C" int swpc_rmedia_3d(int device, int ptype, int n1, int n2, int n3, float d1, float d2, float d3, float a1,
                              float a2, float a3, float epsil, float kappa, uint64_t seed, float *out)
{
    g_err[0] = 0;
    if (!out || n1 < 2 || n2 < 2 || n3 < 2 || ptype < 1 || ptype > 3 || !(d1 > 0) || !(d2 > 0) || !(d3 > 0) ||
        !(a1 > 0) || !(a2 > 0) || !(a3 > 0) || !(epsil >= 0) || (ptype == 3 && !(kappa > 0)))
        return rfail(SWPC_EINVAL, "swpc_rmedia_3d: bad arguments");
    RCHK(hipSetDevice(device));
    const long long n = (long long)n1 * n2 * n3;
    // out in device memory of this GPU: the field is written there directly
    hipPointerAttribute_t at;
    const bool out_dev = hipPointerGetAttributes(&at, out) == hipSuccess && at.type == hipMemoryTypeDevice &&
                         at.device == device;
    (void)hipGetLastError();
    const int n1h = n1 / 2 + 1;
    const long long nc = (long long)n1h * n2 * n3;
    hipStream_t s;
    RCHK(hipStreamCreate(&s));
    struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{s};
    Dev dr, dc, dpart, dstats, dout;
    const int nblk = 2048;
    RCHK(hipMalloc(&dr.p, (size_t)n * sizeof(double)));
    RCHK(hipMalloc(&dc.p, (size_t)nc * sizeof(double2)));
    RCHK(hipMalloc(&dpart.p, (size_t)nblk * sizeof(double)));
    RCHK(hipMalloc(&dstats.p, 2 * sizeof(double)));
    if (!out_dev) RCHK(hipMalloc(&dout.p, (size_t)n * sizeof(float)));
    double *r = (double *)dr.p;
    double2 *c = (double2 *)dc.p;
    double *part = (double *)dpart.p, *stats = (double *)dstats.p;

    // 1. noise
    hipLaunchKernelGGL(k_noise, dim3(nblk), dim3(256), 0, s, r, n, (unsigned long long)seed);
    RCHK(hipGetLastError());
    // 2. forward transform, axis 1 fastest (hipFFT lists the slowest axis first)
    hipfftHandle fwd, inv;
    FCHK(hipfftPlan3d(&fwd, n3, n2, n1, HIPFFT_D2Z));
    struct PlanGuard { hipfftHandle h; ~PlanGuard() { (void)hipfftDestroy(h); } } pf{fwd};
    FCHK(hipfftSetStream(fwd, s));
    FCHK(hipfftExecD2Z(fwd, (hipfftDoubleReal *)r, (hipfftDoubleComplex *)c));
    // 3-4. whitening and PSDF filter; constants of the PSDFs of
    // m_rmedia.f90:682-732 (they cancel in the final scaling but are kept)
    FilterArgs fa;
    fa.b = c;
    fa.n1h = n1h; fa.n2 = n2; fa.n3 = n3;
    fa.n1 = n1; fa.n2f = n2; fa.n3f = n3;
    const double PI = 3.14159265358979323846;
    fa.k1 = 2 * PI / ((double)n1 * d1);
    fa.k2 = 2 * PI / ((double)n2 * d2);
    fa.k3 = 2 * PI / ((double)n3 * d3);
    fa.a1 = a1; fa.a2 = a2; fa.a3 = a3;
    fa.ptype = ptype;
    const double e2a = (double)epsil * epsil * a1 * a2 * a3;
    if (ptype == 1) fa.coef = e2a * pow(PI, 1.5);
    else if (ptype == 2) fa.coef = 8 * e2a * PI;
    else fa.coef = 8 * pow(PI, 1.5) * e2a * std::tgamma((double)kappa + 1.5) / std::tgamma((double)kappa);
    fa.expo = (double)kappa + 1.5;
    fa.scale = (double)n * d1 * d2 * d3;
    hipLaunchKernelGGL(k_filter, dim3(nblk), dim3(256), 0, s, fa);
    RCHK(hipGetLastError());
    // 5. inverse transform (the output is real: the filtered spectrum keeps
    // the Hermitian symmetry of the noise's)
    FCHK(hipfftPlan3d(&inv, n3, n2, n1, HIPFFT_Z2D));
    struct PlanGuard pi{inv};
    FCHK(hipfftSetStream(inv, s));
    FCHK(hipfftExecZ2D(inv, (hipfftDoubleComplex *)c, (hipfftDoubleReal *)r));
    // 6. mean and rms (m_rmedia.f90:144-159)
    hipLaunchKernelGGL(k_partial, dim3(nblk), dim3(256), 0, s, r, n, 0.0, 0, part);
    hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, s, part, nblk, stats);
    RCHK(hipGetLastError());
    double sum = 0.0;
    RCHK(hipMemcpyAsync(&sum, stats, sizeof(double), hipMemcpyDeviceToHost, s));
    RCHK(hipStreamSynchronize(s));
    const double mean = sum / (double)n;
    hipLaunchKernelGGL(k_partial, dim3(nblk), dim3(256), 0, s, r, n, mean, 1, part);
    hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, s, part, nblk, stats + 1);
    RCHK(hipGetLastError());
    double ss = 0.0;
    RCHK(hipMemcpyAsync(&ss, stats + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    RCHK(hipStreamSynchronize(s));
    const double st[2] = {mean, std::sqrt(ss / (double)n)};
    RCHK(hipMemcpyAsync(stats, st, sizeof st, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_finish, dim3(nblk), dim3(256), 0, s, r, n, stats, (double)epsil,
                       out_dev ? out : (float *)dout.p);
    RCHK(hipGetLastError());
    if (!out_dev) RCHK(hipMemcpyAsync(out, dout.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
    RCHK(hipStreamSynchronize(s));
    return SWPC_OK;
}
