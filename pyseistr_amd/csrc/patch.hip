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

// One block per patch.  Threads cover the l1 x l2 plane of the patch (q =
// e1 + l1*e2, so runs of l1 samples are contiguous in both the volume and
// the patch row) and walk axis 3 with fixed strides; the only divisions are
// one per plane position and the block's patch-number split.
// Blocks are dealt to the eight XCDs in turn; block b takes patch
// (b % 8) * per + b / 8, so each XCD works through a contiguous run of patch
// numbers and the overlapping neighbours share its L2.
__global__ void patch_cut(const double *__restrict__ in, double *__restrict__ X, PGeo g, long L, int npatch,
                          int per) {
    const int pid = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (pid >= npatch) return;
    const int p3 = pid % g.P[2], p2 = (pid / g.P[2]) % g.P[1], p1 = pid / (g.P[1] * g.P[2]);
    const int b1 = p1 * g.s[0], b2 = p2 * g.s[1], b3 = p3 * g.s[2];
    const int l1 = g.l[0], l12 = g.l[0] * g.l[1], l3 = g.l[2];
    const long n12 = (long)g.n[0] * g.n[1];
    const int m3 = min(l3, g.n[2] - b3);  // axis-3 samples inside the volume (may be <= 0)
    double *row = X + (long)pid * L;
    for (int q = threadIdx.x; q < l12; q += blockDim.x) {
        const int i1 = b1 + q % l1, i2 = b2 + q / l1;
        const bool ok = i1 < g.n[0] && i2 < g.n[1];
        const double *src = in + i1 + (long)g.n[0] * i2 + n12 * b3;
        const int m = ok ? m3 : 0;
        int e3 = 0;
        for (; e3 + 8 <= m; e3 += 8) {  // eight loads in flight before their stores
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = src[(e3 + u) * n12];
#pragma unroll
            for (int u = 0; u < 8; u++) row[q + (long)(e3 + u) * l12] = v[u];
        }
        for (; e3 < m; e3++) row[q + (long)e3 * l12] = src[e3 * n12];
        for (; e3 < l3; e3++) row[q + (long)e3 * l12] = 0.0;
    }
}

// patches along one axis whose span covers sample i: [lo, hi]
__device__ inline void covering(int i, int l, int s, int P, int &lo, int &hi) {
    const int a = i - l + 1;
    lo = a <= 0 ? 0 : (a + s - 1) / s;
    hi = min(i / s, P - 1);
}

// One block per (i2, i3) column of the output, threads along i1.  Every
// sample adds its covering patches in patch order (p1, then p2, then p3),
// the order in which the reference accumulates them.
__global__ void patch_paste(const double *__restrict__ X, double *__restrict__ out, PGeo g, long L) {
    const int c = blockIdx.x;
    const int i2 = c % g.n[1], i3 = c / g.n[1];
    int lo2, hi2, lo3, hi3;
    covering(i2, g.l[1], g.s[1], g.P[1], lo2, hi2);
    covering(i3, g.l[2], g.s[2], g.P[2], lo3, hi3);
    const long l1 = g.l[0], l12 = (long)g.l[0] * g.l[1];
    for (int i1 = threadIdx.x; i1 < g.n[0]; i1 += blockDim.x) {
        int lo1, hi1;
        covering(i1, g.l[0], g.s[0], g.P[0], lo1, hi1);
        double acc = 0.0, cnt = 0.0;
        for (int p1 = lo1; p1 <= hi1; p1++) {
            const long e1 = i1 - p1 * g.s[0];
            for (int p2 = lo2; p2 <= hi2; p2++) {
                const long id0 = ((long)p1 * g.P[1] + p2) * g.P[2];
                const long e12 = e1 + l1 * (i2 - p2 * g.s[1]);
                for (int p3 = lo3; p3 <= hi3; p3++) {
                    acc += X[(id0 + p3) * L + e12 + l12 * (i3 - p3 * g.s[2])];
                    cnt += 1.0;
                }
            }
        }
        out[(long)c * g.n[0] + i1] = acc / cnt;
    }
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
    if (np >= (1L << 31) - 16 || (long)n2 * n3 >= (1L << 31) || (long)l1 * l2 >= (1L << 31))
        return sh_fail(ctx, SH_EINVAL, "patch: %ld patches or %ld columns exceed one launch", np, (long)n2 * n3);
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
    const int tpb = (int)std::min<long>(256, ceil_div((long)l1 * l2, 64) * 64);
    const int per = (int)ceil_div(np, 8);
    hipLaunchKernelGGL(patch_cut, dim3((unsigned)(8L * per)), dim3(tpb), 0, ctx->stream, in, X, g, L, (int)np, per);
    return sh_launch_check(ctx, "patch_cut");
}

extern "C" int sh_patch_inv(sh_ctx *ctx, const double *X, double *out, int n1, int n2, int n3, int l1, int l2,
                            int l3, int s1, int s2, int s3) {
    if (!ctx) return SH_ENULL;
    if (!X || !out) return sh_fail(ctx, SH_ENULL, "sh_patch_inv: NULL pointer");
    PGeo g;
    long np, L;
    SH_TRY(geometry(ctx, n1, n2, n3, l1, l2, l3, s1, s2, s3, g, np, L));
    const int tpb = (int)std::min<long>(256, ceil_div((long)n1, 64) * 64);
    hipLaunchKernelGGL(patch_paste, dim3((unsigned)((long)n2 * n3)), dim3(tpb), 0, ctx->stream, X, out, g, L);
    return sh_launch_check(ctx, "patch_paste");
}
