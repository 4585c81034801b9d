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
#include <cstdlib>

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

// one thread per trace (F-order: trace i2 = x[i2 * n1 .. + n1)); used when
// a pass has more sections than the fused kernel below holds
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

// A cascade of sections applied sample by sample: section s takes the
// output sequence of section s-1 and keeps its own state, so each section
// performs exactly the float operations of its own pass over the whole
// trace (butter_pass), in the same order; only the loop nesting changes.
constexpr int BP_FUSE_MAX = 8;
struct Stage {
    int odd, low;
    float d0, d1, d2;
};
struct Cascade {
    Stage st[BP_FUSE_MAX];
};

int cascade_of(const Butter &bw, Cascade &c, int ns) {
    if (bw.nn % 2) c.st[ns++] = Stage{1, bw.low, bw.den[bw.nn / 2][0], 0.0f, bw.den[bw.nn / 2][1]};
    for (int j = 0; j < bw.nn / 2; j++) c.st[ns++] = Stage{0, bw.low, bw.den[j][0], bw.mid, bw.den[j][1]};
    return ns;
}

// 64 traces per block (one wave), staged through LDS in chunks of 64
// samples: the loads and stores run along the traces (coalesced) and each
// lane then walks its own trace's column of the tile.  tile[k][r] with row
// pitch 65 keeps both access patterns free of bank conflicts.
template <int NS>
__global__ void __launch_bounds__(64)
bandpass_fused(float *__restrict__ x, long ntr, int n1, const Cascade c, int back) {
    __shared__ float tile[64][65];
    const int t = threadIdx.x;
    const long tr0 = (long)blockIdx.x * 64;
    const int nr = (int)min(64L, ntr - tr0);
    float x0[NS], x1[NS], x2[NS], y1[NS], y2[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) x0[s] = x1[s] = x2[s] = y1[s] = y2[s] = 0.0f;
    const int nch = (n1 + 63) / 64;
    for (int ci = 0; ci < nch; ci++) {
        const int k0 = (back ? nch - 1 - ci : ci) * 64;
        const int cnt = min(64, n1 - k0);
        if (t < cnt)
            for (int r = 0; r < nr; r++) tile[t][r] = x[(tr0 + r) * n1 + k0 + t];
        __syncthreads();
        if (t < nr) {
            for (int kk = 0; kk < cnt; kk++) {
                const int k = back ? cnt - 1 - kk : kk;
                float v = tile[k][t];
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const Stage &g = c.st[s];
                    float y0;
                    if (g.odd) {
                        x1[s] = x0[s];
                        x0[s] = v;
                        y0 = g.low ? (x0[s] + x1[s] - g.d2 * y1[s]) * g.d0 : (x0[s] - x1[s] - g.d2 * y1[s]) * g.d0;
                        y1[s] = y0;
                    } else {
                        x2[s] = x1[s];
                        x1[s] = x0[s];
                        x0[s] = v;
                        y0 = g.low ? (x0[s] + 2 * x1[s] + x2[s] - g.d1 * y1[s] - g.d2 * y2[s]) * g.d0
                                   : (x0[s] - 2 * x1[s] + x2[s] - g.d1 * y1[s] - g.d2 * y2[s]) * g.d0;
                        y2[s] = y1[s];
                        y1[s] = y0;
                    }
                    v = y0;
                }
                tile[k][t] = v;
            }
        }
        __syncthreads();
        if (t < cnt)
            for (int r = 0; r < nr; r++) x[(tr0 + r) * n1 + k0 + t] = tile[t][r];
        __syncthreads();
    }
}

int launch_fused(sh_ctx *ctx, float *x, long ntr, int n1, const Cascade &c, int ns, int back) {
    const dim3 grid((unsigned)ceil_div(ntr, 64)), block(64);
    switch (ns) {
#define BP_CASE(N) \
    case N: hipLaunchKernelGGL(bandpass_fused<N>, grid, block, 0, ctx->stream, x, ntr, n1, c, back); break;
        BP_CASE(1) BP_CASE(2) BP_CASE(3) BP_CASE(4) BP_CASE(5) BP_CASE(6) BP_CASE(7) BP_CASE(8)
#undef BP_CASE
        default: return sh_fail(ctx, SH_EINVAL, "bandpass: %d sections in one pass", ns);
    }
    return sh_launch_check(ctx, "bandpass_fused");
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
    const int slo = blo.on ? blo.nn / 2 + blo.nn % 2 : 0, shi = bhi.on ? bhi.nn / 2 + bhi.nn % 2 : 0;
    if (slo > BP_FUSE_MAX || shi > BP_FUSE_MAX || getenv("SH_BANDPASS_PLAIN")) {
        hipLaunchKernelGGL(bandpass_traces, dim3((unsigned)ceil_div(ntr, 256)), dim3(256), 0, ctx->stream, out, ntr, n1,
                           blo, bhi, phase);
        return sh_launch_check(ctx, "bandpass_traces");
    }
    // passes in the reference's order: low cut forward (and backward for
    // zero phase), then high cut; minimum phase runs both filters in one
    // forward pass when they fit
    Cascade lo{}, hi{};
    if (blo.on) cascade_of(blo, lo, 0);
    if (bhi.on) cascade_of(bhi, hi, 0);
    if (phase && slo + shi <= BP_FUSE_MAX) {
        Cascade both = lo;
        if (bhi.on) cascade_of(bhi, both, slo);
        return launch_fused(ctx, out, ntr, n1, both, slo + shi, 0);
    }
    if (blo.on) {
        SH_TRY(launch_fused(ctx, out, ntr, n1, lo, slo, 0));
        if (!phase) SH_TRY(launch_fused(ctx, out, ntr, n1, lo, slo, 1));
    }
    if (bhi.on) {
        SH_TRY(launch_fused(ctx, out, ntr, n1, hi, shi, 0));
        if (!phase) SH_TRY(launch_fused(ctx, out, ntr, n1, hi, shi, 1));
    }
    return SH_OK;
}
