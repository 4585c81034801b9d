// Butterworth band-pass per trace: sh_bandpass replaces bpcfun.cbp
// (pyseistr/src/bp_cfuns.c:524-645), the C entry behind bandpassc
// (pyseistr/bp.py:128).  The coefficients are computed on the host with the
// reference's expressions (sf_butter_init, bp_cfuns.c:420-462: float
// variables, double literals, libm sinf/cosf); the recursive sections
// (sf_butter_apply, 473-507) run with one thread per trace in the
// reference's order of float operations, and the zero-phase pass over the
// reversed trace (sf_reverse, 509-522) runs as a backward pass, the same
// operations in the same order.  Results are bit-identical to the reference.
#include "common.hpp"

#include <cmath>

namespace {

constexpr int BP_MAXSEC = 64;  // sections per filter: up to 127 poles
constexpr double BP_PI = 3.14159265358979323846264338328;

struct Butter {
    int on, low, nn;
    float den[BP_MAXSEC][2], mid;
};

void butter_init(Butter &bw, bool low, float cutoff, int nn) {
    float arg, ss, sinw, cosw, fact;
    arg = 2. * BP_PI * cutoff;
    sinw = sinf(arg);
    cosw = cosf(arg);
    bw.on = 1;
    bw.nn = nn;
    bw.low = low;
    if (nn % 2) {
        if (low) {
            fact = (1. + cosw) / sinw;
            bw.den[nn / 2][0] = 1. / (1. + fact);
            bw.den[nn / 2][1] = 1. - fact;
        } else {
            fact = sinw / (1. + cosw);
            bw.den[nn / 2][0] = 1. / (fact + 1.);
            bw.den[nn / 2][1] = fact - 1.;
        }
    }
    fact = low ? sinf(0.5 * arg) : cosf(0.5 * arg);
    fact *= fact;
    for (int j = 0; j < nn / 2; j++) {
        ss = sinf(BP_PI * (2 * j + 1) / (2 * nn)) * sinw;
        bw.den[j][0] = fact / (1. + ss);
        bw.den[j][1] = (1. - ss) / fact;
    }
    bw.mid = -2. * cosw / fact;
}

// every section of bw over one trace (x[k * s], k < nx), forward or backward
__device__ void butter_pass(const Butter &bw, float *x, long s, int nx, bool back) {
    float d0, d1 = bw.mid, d2, x0, x1, x2, y0, y1, y2;
    const int nn = bw.nn;
    if (nn % 2) {
        d0 = bw.den[nn / 2][0];
        d2 = bw.den[nn / 2][1];
        x0 = y1 = 0.0f;
        for (int k = 0; k < nx; k++) {
            const long ix = (back ? nx - 1 - k : k) * s;
            x1 = x0;
            x0 = x[ix];
            y0 = bw.low ? (x0 + x1 - d2 * y1) * d0 : (x0 - x1 - d2 * y1) * d0;
            x[ix] = y1 = y0;
        }
    }
    for (int j = 0; j < nn / 2; j++) {
        d0 = bw.den[j][0];
        d2 = bw.den[j][1];
        x1 = x0 = y1 = y2 = 0.0f;
        for (int k = 0; k < nx; k++) {
            const long ix = (back ? nx - 1 - k : k) * s;
            x2 = x1;
            x1 = x0;
            x0 = x[ix];
            y0 = bw.low ? (x0 + 2 * x1 + x2 - d1 * y1 - d2 * y2) * d0 : (x0 - 2 * x1 + x2 - d1 * y1 - d2 * y2) * d0;
            y2 = y1;
            x[ix] = y1 = y0;
        }
    }
}

// one thread per trace (F-order: trace i2 = x[i2 * n1 .. + n1))
__global__ void __launch_bounds__(256)
bandpass_traces(float *__restrict__ x, long ntr, int n1, const Butter blo, const Butter bhi, int phase) {
    const long tr = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tr >= ntr) return;
    float *t = x + tr * n1;
    if (blo.on) {
        butter_pass(blo, t, 1, n1, false);
        if (!phase) butter_pass(blo, t, 1, n1, true);
    }
    if (bhi.on) {
        butter_pass(bhi, t, 1, n1, false);
        if (!phase) butter_pass(bhi, t, 1, n1, true);
    }
}

}  // namespace

extern "C" int sh_bandpass(sh_ctx *ctx, const float *in, float *out, int n1, int n2, int n3, float d1, float flo,
                           float fhi, int nplo, int nphi, int phase) {
    if (!ctx) return SH_ENULL;
    if (!in || !out) return sh_fail(ctx, SH_ENULL, "sh_bandpass: NULL pointer");
    if (n1 < 1 || n2 < 1 || n3 < 1) return sh_fail(ctx, SH_EINVAL, "sh_bandpass: bad shape %d x %d x %d", n1, n2, n3);
    // cbp's parameter handling (bp_cfuns.c:582-606): negative cutoffs are
    // left as they are (the reference only prints), others scaled by d1
    const float eps = 0.0001;
    if (!(flo < 0)) flo = flo * d1;
    if (!(fhi < 0)) fhi = fhi * d1;
    if (nplo < 1) nplo = 1;
    if (nplo > 1 && !phase) nplo /= 2;
    if (nphi < 1) nphi = 1;
    if (nphi > 1 && !phase) nphi /= 2;
    if ((nplo + 1) / 2 > BP_MAXSEC || (nphi + 1) / 2 > BP_MAXSEC)
        return sh_fail(ctx, SH_EUNSUP, "sh_bandpass: at most %d poles", 2 * BP_MAXSEC - 1);
    Butter blo{}, bhi{};
    if (flo > eps) butter_init(blo, false, flo, nplo);
    if (fhi < 0.5 - eps) butter_init(bhi, true, fhi, nphi);
    const long ntr = (long)n2 * n3, n = ntr * n1;
    if (out != in) SH_HIP(ctx, hipMemcpyAsync(out, in, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    if (!blo.on && !bhi.on) return SH_OK;
    hipLaunchKernelGGL(bandpass_traces, dim3((unsigned)ceil_div(ntr, 256)), dim3(256), 0, ctx->stream, out, ntr, n1, blo,
                       bhi, phase);
    return sh_launch_check(ctx, "bandpass_traces");
}
