// Patch extraction and reassembly around the batched path: sh_patch and
// sh_patch_inv replace the reference's pure-Python patch2d / patch3d and
// patch2d_inv / patch3d_inv (pyseistr/patch.py:2,86,208,307) for mode 1.
// Each axis is zero-padded so that (n - l) is a multiple of s; patches are
// numbered with the first axis outermost and the last innermost, each
// stored as one row of l1*l2*l3 values in Fortran order.  The inverse sums
// every patch covering a sample in patch order, in double, and divides by
// the number of patches covering it (NaN where none does), as the
// reference's float64 accumulation does; the results are identical.
#include "common.hpp"

namespace {

struct PGeo {
    int n[3], l[3], s[3], P[3];  // sizes, patch sizes, shifts, patches per axis
};

__global__ void patch_cut(const double *__restrict__ in, double *__restrict__ X, PGeo g, long L, long total) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const long pid = idx / L, e = idx % L;
    const int e1 = (int)(e % g.l[0]), e2 = (int)((e / g.l[0]) % g.l[1]), e3 = (int)(e / ((long)g.l[0] * g.l[1]));
    const int p3 = (int)(pid % g.P[2]), p2 = (int)((pid / g.P[2]) % g.P[1]), p1 = (int)(pid / ((long)g.P[1] * g.P[2]));
    const int i1 = p1 * g.s[0] + e1, i2 = p2 * g.s[1] + e2, i3 = p3 * g.s[2] + e3;
    const bool in_ = i1 < g.n[0] && i2 < g.n[1] && i3 < g.n[2];
    X[idx] = in_ ? in[i1 + (long)g.n[0] * (i2 + (long)g.n[1] * i3)] : 0.0;
}

// patches along one axis whose span covers sample i: [lo, hi]
__device__ inline void covering(int i, int l, int s, int P, int &lo, int &hi) {
    const int a = i - l + 1;
    lo = a <= 0 ? 0 : (a + s - 1) / s;
    hi = min(i / s, P - 1);
}

__global__ void patch_paste(const double *__restrict__ X, double *__restrict__ out, PGeo g, long L, long nout) {
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nout) return;
    const int i1 = (int)(v % g.n[0]), i2 = (int)((v / g.n[0]) % g.n[1]), i3 = (int)(v / ((long)g.n[0] * g.n[1]));
    int lo1, hi1, lo2, hi2, lo3, hi3;
    covering(i1, g.l[0], g.s[0], g.P[0], lo1, hi1);
    covering(i2, g.l[1], g.s[1], g.P[1], lo2, hi2);
    covering(i3, g.l[2], g.s[2], g.P[2], lo3, hi3);
    double acc = 0.0, cnt = 0.0;
    for (int p1 = lo1; p1 <= hi1; p1++)
        for (int p2 = lo2; p2 <= hi2; p2++)
            for (int p3 = lo3; p3 <= hi3; p3++) {
                const long id = ((long)p1 * g.P[1] + p2) * g.P[2] + p3;
                const long e = (i1 - p1 * g.s[0]) + (long)g.l[0] * ((i2 - p2 * g.s[1]) + (long)g.l[1] * (i3 - p3 * g.s[2]));
                acc += X[id * L + e];
                cnt += 1.0;
            }
    out[v] = acc / cnt;
}

// padded size N and patch count P per axis (patch.py:66-72, 174-184)
int geometry(sh_ctx *ctx, int n1, int n2, int n3, int l1, int l2, int l3, int s1, int s2, int s3, PGeo &g, long &np,
             long &L) {
    const int n[3] = {n1, n2, n3}, l[3] = {l1, l2, l3}, s[3] = {s1, s2, s3};
    np = 1;
    for (int a = 0; a < 3; a++) {
        if (n[a] < 1 || l[a] < 1 || s[a] < 1)
            return sh_fail(ctx, SH_EINVAL, "patch: sizes, patch sizes and shifts must be >= 1");
        int t = (n[a] - l[a]) % s[a];
        if (t < 0) t += s[a];  // Python's modulo
        const int N = n[a] + (t ? s[a] - t : 0);
        if (N < l[a]) return sh_fail(ctx, SH_EINVAL, "patch: axis %d of %d is shorter than its patch %d", a + 1, n[a], l[a]);
        g.n[a] = n[a];
        g.l[a] = l[a];
        g.s[a] = s[a];
        g.P[a] = (N - l[a]) / s[a] + 1;
        np *= g.P[a];
    }
    L = (long)l1 * l2 * l3;
    if (np * L / 256 >= (1L << 31) || (long)n1 * n2 * n3 / 256 >= (1L << 31))
        return sh_fail(ctx, SH_EINVAL, "patch: %ld patches of %ld samples exceed one launch", np, L);
    return SH_OK;
}

}  // namespace

extern "C" int sh_patch_count(int n1, int n2, int n3, int l1, int l2, int l3, int s1, int s2, int s3, long *npatch) {
    sh_ctx tmp;
    PGeo g;
    long np = 0, L = 0;
    int rc = geometry(&tmp, n1, n2, n3, l1, l2, l3, s1, s2, s3, g, np, L);
    if (rc == SH_OK && npatch) *npatch = np;
    return rc;
}

extern "C" int sh_patch(sh_ctx *ctx, const double *in, double *X, int n1, int n2, int n3, int l1, int l2, int l3,
                        int s1, int s2, int s3) {
    if (!ctx) return SH_ENULL;
    if (!in || !X) return sh_fail(ctx, SH_ENULL, "sh_patch: NULL pointer");
    PGeo g;
    long np, L;
    SH_TRY(geometry(ctx, n1, n2, n3, l1, l2, l3, s1, s2, s3, g, np, L));
    const long total = np * L;
    hipLaunchKernelGGL(patch_cut, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, ctx->stream, in, X, g, L, total);
    return sh_launch_check(ctx, "patch_cut");
}

extern "C" int sh_patch_inv(sh_ctx *ctx, const double *X, double *out, int n1, int n2, int n3, int l1, int l2,
                            int l3, int s1, int s2, int s3) {
    if (!ctx) return SH_ENULL;
    if (!X || !out) return sh_fail(ctx, SH_ENULL, "sh_patch_inv: NULL pointer");
    PGeo g;
    long np, L;
    SH_TRY(geometry(ctx, n1, n2, n3, l1, l2, l3, s1, s2, s3, g, np, L));
    const long nout = (long)n1 * n2 * n3;
    hipLaunchKernelGGL(patch_paste, dim3((unsigned)ceil_div(nout, 256)), dim3(256), 0, ctx->stream, X, out, g, L, nout);
    return sh_launch_check(ctx, "patch_paste");
}
