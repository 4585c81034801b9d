// FK dip filter: sh_fkdip replaces the pure-Python fkdip (pyseistr/fk.py:2-68),
// the step after somf2dc in the reference's DAS demo.  The data are padded
// to powers of two (nf samples, nk traces); a real FFT along time keeps the
// nf/2 + 1 non-negative frequencies, a complex FFT along space follows, and
// a cone around zero wavenumber is kept: frequency row i1 (1-based) and
// fftshift-ed wavenumber column i2 pass when i1 > (nn1/nw)*(i2 - nk/2) and
// i1 > (nn1/nw)*(nk/2 - i2), with nn1 = nf/2 + 1 and nw = w*nk, evaluated
// in double as the reference does (fk.py:51-55).  The inverse transforms
// undo both FFTs; the reference's explicit Hermitian extension followed by
// a complex inverse FFT and the real part (fk.py:57-65) is a complex-to-real
// inverse transform, which is what runs here.  The transforms are plain
// library FFTs (hipFFT, float64); the staging, the cone and the crop are
// the kernels below.  Agreement with the reference is to float64 rounding.
#include "common.hpp"

#include <hipfft/hipfft.h>

#include <cstdlib>

namespace {

struct FkPlans {
    int n1 = 0, n2 = 0, two_d = 0;
    hipfftHandle t_fwd = nullptr, x = nullptr, t_inv = nullptr;  // time R2C, space C2C, time C2R
    hipfftHandle f2 = nullptr, i2 = nullptr;                     // both axes at once: R2C, C2R
    ~FkPlans() {
        for (hipfftHandle h : {t_fwd, x, t_inv, f2, i2})
            if (h) hipfftDestroy(h);
    }
};

int pow2_at_least(int n) {
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

// R[c][i] (nf per trace, nc traces) = in[c][i] for i < n1 and c < n2, zero
// beyond
__global__ void fk_stage(const double *__restrict__ in, double *__restrict__ R, int n1, int nf, int n2, int nc) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)nf * nc) return;
    const int i = (int)(k % nf), c = (int)(k / nf);
    R[k] = i < n1 && c < n2 ? in[(long)c * n1 + i] : 0.0;
}

// n1 == 1: the length-1 time transform is the identity, so the spectrum
// columns take the samples directly
__global__ void fk_stage1(const double *__restrict__ in, double2 *__restrict__ C, int n2, int nk) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nk) return;
    C[c] = make_double2(c < n2 ? in[c] : 0.0, 0.0);
}

// zero the padded traces n2 .. nk-1 of the spectrum (nh rows per trace)
__global__ void fk_zero(double2 *__restrict__ C, int nh, int n2, int nk) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)nh * (nk - n2)) return;
    C[(long)n2 * nh + k] = make_double2(0.0, 0.0);
}

// keep the cone; column k of the unshifted spectrum sits at fftshift-ed
// position (k + nk/2) mod nk
__global__ void fk_cone(double2 *__restrict__ C, int nh, int nk, double slope) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)nh * nk) return;
    const int f = (int)(k % nh), col = (int)(k / nh);
    const int nk2 = nk / 2;
    const int i1 = f + 1, i2 = (col + nk2) % nk + 1;
    const bool keep = (double)i1 > slope * (double)(i2 - nk2) && (double)i1 > slope * (double)(nk2 - i2);
    if (!keep) C[k] = make_double2(0.0, 0.0);
}

// out[c][i] = R[c][i] * scale for i < n1 (R holds nf samples per trace)
__global__ void fk_crop(const double *__restrict__ R, double *__restrict__ out, int n1, int nf, int n2, double scale) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)n1 * n2) return;
    const int i = (int)(k % n1), c = (int)(k / n1);
    out[k] = R[(long)c * nf + i] * scale;
}

__global__ void fk_crop1(const double2 *__restrict__ C, double *__restrict__ out, int n2, double scale) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n2) return;
    out[c] = C[c].x * scale;
}

int fft_fail(sh_ctx *ctx, const char *what, hipfftResult r) {
    return sh_fail(ctx, SH_EHIP, "sh_fkdip: %s failed (hipfftResult %d)", what, (int)r);
}

#define FK_FFT(ctx, what, call)                                \
    do {                                                       \
        hipfftResult r_ = (call);                              \
        if (r_ != HIPFFT_SUCCESS) return fft_fail(ctx, what, r_); \
    } while (0)

// plans for (n1, n2), cached on the context
int fk_plans(sh_ctx *ctx, int n1, int n2, FkPlans *&out) {
    auto *p = static_cast<FkPlans *>(ctx->fk.get());
    const int want_2d = pow2_at_least(n1) > 1 && pow2_at_least(n2) > 1 && !getenv("SH_FK_1D");
    if (p && p->n1 == n1 && p->n2 == n2 && p->two_d == want_2d) {
        out = p;
        return SH_OK;
    }
    ctx->fk.reset();
    auto *np_ = new (std::nothrow) FkPlans();
    if (!np_) return sh_fail(ctx, SH_EALLOC, "sh_fkdip: out of host memory");
    ctx->fk = std::shared_ptr<void>(np_, [](void *q) { delete static_cast<FkPlans *>(q); });
    const int nf = pow2_at_least(n1), nk = pow2_at_least(n2), nh = nf / 2 + 1;
    if (nf > 1 && nk > 1 && !getenv("SH_FK_1D")) {
        // the time transform of every padded trace followed by the space
        // transform of every frequency is the 2-D real transform of the
        // nk x nf array (time contiguous); the library then picks its own
        // column kernels for the space axis
        FK_FFT(ctx, "2-D plan", hipfftPlan2d(&np_->f2, nk, nf, HIPFFT_D2Z));
        FK_FFT(ctx, "2-D plan", hipfftPlan2d(&np_->i2, nk, nf, HIPFFT_Z2D));
        np_->two_d = 1;
        np_->n1 = n1;
        np_->n2 = n2;
        out = np_;
        return SH_OK;
    }
    if (nf > 1) {
        int n[1] = {nf}, re[1] = {nf}, ce[1] = {nh};
        FK_FFT(ctx, "time plan", hipfftPlanMany(&np_->t_fwd, 1, n, re, 1, nf, ce, 1, nh, HIPFFT_D2Z, n2));
        FK_FFT(ctx, "time plan", hipfftPlanMany(&np_->t_inv, 1, n, ce, 1, nh, re, 1, nf, HIPFFT_Z2D, n2));
    }
    if (nk > 1) {
        int n[1] = {nk}, e[1] = {nk};
        FK_FFT(ctx, "space plan", hipfftPlanMany(&np_->x, 1, n, e, nh, 1, e, nh, 1, HIPFFT_Z2Z, nh));
    }
    np_->n1 = n1;
    np_->n2 = n2;
    out = np_;
    return SH_OK;
}

}  // namespace

extern "C" int sh_fkdip(sh_ctx *ctx, const double *in, double *out, int n1, int n2, double w) {
    if (!ctx) return SH_ENULL;
    if (!in || !out) return sh_fail(ctx, SH_ENULL, "sh_fkdip: NULL pointer");
    if (n1 < 1 || n2 < 1 || n1 > (1 << 29) || n2 > (1 << 29))
        return sh_fail(ctx, SH_EINVAL, "sh_fkdip: bad shape %d x %d", n1, n2);
    const int nf = pow2_at_least(n1), nk = pow2_at_least(n2), nh = nf / 2 + 1;
    const double nw = w * nk;
    if (nw == 0.0) return sh_fail(ctx, SH_EINVAL, "sh_fkdip: w * nk is zero (the cone slope divides by it)");
    if ((long)nh * nk > (1L << 31) || (long)nf * nk > (1L << 31))
        return sh_fail(ctx, SH_EINVAL, "sh_fkdip: %d x %d spectrum exceeds one transform batch", nh, nk);
    const double slope = (double)nh / nw;
    FkPlans *pl;
    SH_TRY(fk_plans(ctx, n1, n2, pl));
    void *rb = nullptr, *cb = nullptr;
    const int nc = pl->two_d ? nk : n2;  // traces held in the real buffer
    SH_TRY(sh_slot(ctx, SLOT_FK0, sizeof(double) * (size_t)nf * nc, &rb));
    SH_TRY(sh_slot(ctx, SLOT_FK0 + 1, sizeof(double2) * (size_t)nh * nk, &cb));
    double *R = (double *)rb;
    double2 *C = (double2 *)cb;
    hipStream_t st = ctx->stream;
    for (hipfftHandle h : {pl->t_fwd, pl->x, pl->t_inv, pl->f2, pl->i2})
        if (h) FK_FFT(ctx, "set stream", hipfftSetStream(h, st));
    const int T = 256;
    const double scale = 1.0 / ((double)nf * (double)nk);
    if (pl->two_d) {
        hipLaunchKernelGGL(fk_stage, dim3(ceil_div((long)nf * nk, T)), dim3(T), 0, st, in, R, n1, nf, n2, nk);
        SH_TRY(sh_launch_check(ctx, "fk_stage"));
        FK_FFT(ctx, "2-D transform", hipfftExecD2Z(pl->f2, R, (hipfftDoubleComplex *)C));
        hipLaunchKernelGGL(fk_cone, dim3(ceil_div((long)nh * nk, T)), dim3(T), 0, st, C, nh, nk, slope);
        SH_TRY(sh_launch_check(ctx, "fk_cone"));
        FK_FFT(ctx, "2-D transform", hipfftExecZ2D(pl->i2, (hipfftDoubleComplex *)C, R));
        hipLaunchKernelGGL(fk_crop, dim3(ceil_div((long)n1 * n2, T)), dim3(T), 0, st, R, out, n1, nf, n2, scale);
        return sh_launch_check(ctx, "fk_crop");
    }
    if (nf > 1) {
        hipLaunchKernelGGL(fk_stage, dim3(ceil_div((long)nf * n2, T)), dim3(T), 0, st, in, R, n1, nf, n2, n2);
        SH_TRY(sh_launch_check(ctx, "fk_stage"));
        if (nk > n2) {
            hipLaunchKernelGGL(fk_zero, dim3(ceil_div((long)nh * (nk - n2), T)), dim3(T), 0, st, C, nh, n2, nk);
            SH_TRY(sh_launch_check(ctx, "fk_zero"));
        }
        FK_FFT(ctx, "time transform", hipfftExecD2Z(pl->t_fwd, R, (hipfftDoubleComplex *)C));
    } else {
        hipLaunchKernelGGL(fk_stage1, dim3(ceil_div(nk, T)), dim3(T), 0, st, in, C, n2, nk);
        SH_TRY(sh_launch_check(ctx, "fk_stage1"));
    }
    if (nk > 1)
        FK_FFT(ctx, "space transform",
               hipfftExecZ2Z(pl->x, (hipfftDoubleComplex *)C, (hipfftDoubleComplex *)C, HIPFFT_FORWARD));
    hipLaunchKernelGGL(fk_cone, dim3(ceil_div((long)nh * nk, T)), dim3(T), 0, st, C, nh, nk, slope);
    SH_TRY(sh_launch_check(ctx, "fk_cone"));
    if (nk > 1)
        FK_FFT(ctx, "space transform",
               hipfftExecZ2Z(pl->x, (hipfftDoubleComplex *)C, (hipfftDoubleComplex *)C, HIPFFT_BACKWARD));
    if (nf > 1) {
        FK_FFT(ctx, "time transform", hipfftExecZ2D(pl->t_inv, (hipfftDoubleComplex *)C, R));
        hipLaunchKernelGGL(fk_crop, dim3(ceil_div((long)n1 * n2, T)), dim3(T), 0, st, R, out, n1, nf, n2, scale);
        return sh_launch_check(ctx, "fk_crop");
    }
    hipLaunchKernelGGL(fk_crop1, dim3(ceil_div(n2, T)), dim3(T), 0, st, C, out, n2, scale);
    return sh_launch_check(ctx, "fk_crop1");
}
