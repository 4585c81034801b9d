// Batched 2-D dip estimation and structure-oriented median over thousands
// of uneven patches (gathers / DAS windows) stored back to back.
//
// Every stage of the single-patch pipeline (dipc with n3 == 1,
// pyseistr/src/dip_cfuns.c:1694-1989, and csomf2d, pyseistr/src/
// sof_cfuns.c:1534-1672) runs as ONE launch over all patches: the grid is a
// work list of blocks, each block belongs to one patch and looks its patch
// up in a table, and every patch carries its own CG / line-search scalars,
// so a patch that converges or accepts a step turns its blocks into no-ops
// while the others continue.  Per-patch sums are reduced from per-block
// partials in a fixed order (one wave per patch).  The launch count of a
// whole batch equals that of a single patch.
//
// The 2-D triangle smoothing runs as one fused pass (both axes) over tiles
// whose shape is picked per batch by the autotuner (tile index argument).
#include "dipcore.hpp"

#include <algorithm>
#include <vector>

namespace {

using namespace dipcore;

constexpr int TB = 256;

struct BPatch {
    long off;           // first sample of the patch (F-order, n1 fastest)
    int n1, n2;
    int blk0, nblk;     // per-sample blocks of TB
    int tile0, ntile;   // smoothing tiles
    int cblk0, ncblk;   // chain blocks (2*n2 spray chains / factor variants)
    int ttile0, nttile; // 32x32 transpose tiles
};

struct Tile {
    int patch, x0, y0, pad;
};

// Tile shapes offered to the autotuner (i1 extent, i2 extent).
constexpr int NTILE_CFG = 6;
constexpr int TILE_CFG[NTILE_CFG][2] = {{32, 16}, {64, 8}, {64, 16}, {64, 32}, {128, 8}, {128, 16}};
constexpr int DEFAULT_TILE = 2;

// ---------------------------------------------------------------------------
// Host-built tables, uploaded once per call.
// ---------------------------------------------------------------------------
struct BatchTables {
    int np = 0;
    long ntot = 0;
    int n1max = 0;
    int nblk = 0, ntile = 0, ncblk = 0, nttile = 0;
    BPatch *P = nullptr;
    int *blk_patch = nullptr, *cblk_patch = nullptr;
    Tile *tiles = nullptr, *ttiles = nullptr;
};

int build_tables(sh_ctx *ctx, const long *offs, const int *n1s, const int *n2s, int np, int t1, int t2,
                 BatchTables &T) {
    std::vector<BPatch> P(np);
    std::vector<int> blk, cblk;
    std::vector<Tile> tiles, tt;
    T.np = np;
    T.ntot = 0;
    T.n1max = 0;
    for (int p = 0; p < np; p++) {
        int n1 = n1s[p], n2 = n2s[p];
        long n = (long)n1 * n2;
        BPatch &q = P[p];
        q.off = offs[p];
        q.n1 = n1;
        q.n2 = n2;
        T.ntot = std::max(T.ntot, q.off + n);
        T.n1max = std::max(T.n1max, n1);
        q.blk0 = (int)blk.size();
        q.nblk = ceil_div(n, TB);
        for (int b = 0; b < q.nblk; b++) blk.push_back(p);
        q.tile0 = (int)tiles.size();
        for (int y = 0; y < n2; y += t2)
            for (int x = 0; x < n1; x += t1) tiles.push_back({p, x, y, 0});
        q.ntile = (int)tiles.size() - q.tile0;
        q.cblk0 = (int)cblk.size();
        q.ncblk = ceil_div(2L * n2, TB);
        for (int b = 0; b < q.ncblk; b++) cblk.push_back(p);
        q.ttile0 = (int)tt.size();
        for (int k = 0; k < n1; k += 32)
            for (int t = 0; t < n2; t += 32) tt.push_back({p, t, k, 0});
        q.nttile = (int)tt.size() - q.ttile0;
    }
    T.nblk = (int)blk.size();
    T.ntile = (int)tiles.size();
    T.ncblk = (int)cblk.size();
    T.nttile = (int)tt.size();
    size_t bP = sizeof(BPatch) * np, bB = sizeof(int) * blk.size(), bC = sizeof(int) * cblk.size();
    size_t bT = sizeof(Tile) * tiles.size(), bX = sizeof(Tile) * tt.size();
    auto al = [](size_t v) { return (v + 255) / 256 * 256; };
    size_t tot = al(bP) + al(bB) + al(bC) + al(bT) + al(bX);
    std::vector<char> host(tot);
    char *d = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_BATCH0, tot, (void **)&d));
    size_t o = 0;
    auto put = [&](const void *src, size_t b) {
        memcpy(host.data() + o, src, b);
        char *dst = d + o;
        o += al(b);
        return dst;
    };
    T.P = (BPatch *)put(P.data(), bP);
    T.blk_patch = (int *)put(blk.data(), bB);
    T.cblk_patch = (int *)put(cblk.data(), bC);
    T.tiles = (Tile *)put(tiles.data(), bT);
    T.ttiles = (Tile *)put(tt.data(), bX);
    // the host staging vector must outlive the copy: synchronise before
    // it goes out of scope
    SH_HIP(ctx, hipMemcpyAsync(d, host.data(), tot, hipMemcpyHostToDevice, ctx->stream));
    SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SH_OK;
}

int check_patches(sh_ctx *ctx, const long *offs, const int *n1s, const int *n2s, int np, int min_n1,
                  const char *what) {
    if (!offs || !n1s || !n2s) return sh_fail(ctx, SH_ENULL, "%s: NULL patch table", what);
    if (np < 1) return sh_fail(ctx, SH_EINVAL, "%s: no patches", what);
    for (int p = 0; p < np; p++) {
        if (n1s[p] < min_n1 || n2s[p] < 2 || offs[p] < 0)
            return sh_fail(ctx, SH_EINVAL, "%s: patch %d has bad shape %d x %d", what, p, n1s[p], n2s[p]);
        if (p > 0 && offs[p] < offs[p - 1] + (long)n1s[p - 1] * n2s[p - 1])
            return sh_fail(ctx, SH_EINVAL, "%s: patch %d overlaps the previous one", what, p);
    }
    return SH_OK;
}

// ---------------------------------------------------------------------------
// Dip kernels over the block work list.
// ---------------------------------------------------------------------------
template <int NW, int MODE>
__global__ void __launch_bounds__(TB)
bpwd(const float *__restrict__ x, const float *__restrict__ p, const float *__restrict__ dp, float lam, float pmin,
     float pmax, float *__restrict__ pout, float *__restrict__ u1, float *__restrict__ u2,
     const BPatch *__restrict__ P, const int *__restrict__ blk_patch, double *part, const DipScal *sc) {
    const int pi = blk_patch[blockIdx.x];
    if (MODE == 2 && sc[pi].accepted) return;
    __shared__ double sh[16];
    const BPatch q = P[pi];
    const long n = (long)q.n1 * q.n2;
    const long li = (long)(blockIdx.x - q.blk0) * TB + threadIdx.x;
    double acc = 0.0;
    if (li < n) {
        int i1 = (int)(li % q.n1), i2 = (int)(li / q.n1);
        bool valid = i1 >= NW && i1 < q.n1 - NW && i2 < q.n2 - 1;
        acc = pwd_sample<NW, MODE>(x, p, dp, lam, pmin, pmax, pout, u1, u2, false, q.off + li, q.n1, valid);
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(TB)
bcg_start(const float *__restrict__ num, const float *__restrict__ den, float *__restrict__ w, float *__restrict__ r,
          float *__restrict__ x, float *__restrict__ pc, const BPatch *__restrict__ P, const int *__restrict__ blk_patch,
          double *part, const DipScal *sc) {
    __shared__ double sh[16];
    const int pi = blk_patch[blockIdx.x];
    const BPatch q = P[pi];
    const long li = (long)(blockIdx.x - q.blk0) * TB + threadIdx.x;
    double acc = 0.0;
    if (li < (long)q.n1 * q.n2) acc = cg_start_sample(num, den, w, r, x, pc, q.off + li, sc + pi);
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(TB)
bcg_result(const float *__restrict__ xa, const float *__restrict__ xb, const float *__restrict__ sx,
           float *__restrict__ dp, const BPatch *__restrict__ P, const int *__restrict__ blk_patch, const DipScal *sc) {
    const int pi = blk_patch[blockIdx.x];
    const BPatch q = P[pi];
    const long li = (long)(blockIdx.x - q.blk0) * TB + threadIdx.x;
    if (li < (long)q.n1 * q.n2) cg_result_sample(xa, xb, sx, dp, q.off + li, sc + pi);
}

// Both smoothing axes in one pass over a T1 x T2 tile of one patch; edges
// fold by even reflection inside the patch (the direct form of ps_smooth2,
// dip_cfuns.c:603-643, axis 1 then axis 2 as ps_trianglen_lop orders them).
template <int T1, int T2, int PRO, int EPI>
__global__ void __launch_bounds__(TB)
bsmooth2(const float *__restrict__ in, float *__restrict__ out, int r1, int r2, CgArgs A,
         const BPatch *__restrict__ P, const Tile *__restrict__ tiles, const DipScal *sc, double *part, int ntile,
         int first) {
    const Tile t = tiles[blockIdx.x];
    const DipScal *s = sc + t.patch;
    if (s->cg_done) return;
    extern __shared__ float lds[];
    __shared__ float w1[SM_HMAX + 1], w2[SM_HMAX + 1];
    __shared__ double sh[16];
    const BPatch q = P[t.patch];
    const int n1 = q.n1, n2 = q.n2;
    const int h1 = r1 - 1, h2 = r2 - 1, W1 = T1 + 2 * h1, R2 = T2 + 2 * h2;
    float *tile = lds;           // [R2][W1]
    float *s1 = tile + R2 * W1;  // [R2][T1]
    const int tid = threadIdx.x;
    bool pend = (PRO == PRO_CG || EPI == EPI_GP) ? (s->pending != 0) : false;
    float a2 = pend ? (float)s->alpha2 : 0.0f;
    float a1 = (EPI == EPI_S) ? (float)s->alpha : 0.0f;
    int cur = (PRO == PRO_CG) ? s->cur : 0;
    for (int k = tid; k <= SM_HMAX; k += TB) {
        w1[k] = (k < r1) ? (float)((double)(r1 - k) / ((double)r1 * r1)) : 0.0f;
        w2[k] = (k < r2) ? (float)((double)(r2 - k) / ((double)r2 * r2)) : 0.0f;
    }
    for (int e = tid; e < R2 * W1; e += TB) {
        int r = e / W1, c = e - r * W1;
        int yy = t.y0 - h2 + r, xx = t.x0 - h1 + c;
        int iy = fold_idx(yy, n2), ix = fold_idx(xx, n1);
        bool own = xx >= t.x0 && xx < t.x0 + T1 && xx < n1 && yy >= t.y0 && yy < t.y0 + T2 && yy < n2;
        tile[e] = sm_load<PRO>(in, q.off + ix + (long)n1 * iy, own, A, pend, a2, cur);
    }
    __syncthreads();
    for (int e = tid; e < R2 * T1; e += TB) {
        int r = e / T1, c = e - r * T1;
        const float *cp = tile + r * W1 + c + h1;
        float v = w1[0] * cp[0];
        for (int k = 1; k <= h1; k++) v += w1[k] * (cp[-k] + cp[k]);
        s1[e] = v;
    }
    __syncthreads();
    double acc[3] = {0.0, 0.0, 0.0};
    for (int e = tid; e < T2 * T1; e += TB) {
        int r = e / T1, c = e - r * T1;
        int ix = t.x0 + c, iy = t.y0 + r;
        const float *cp = s1 + (r + h2) * T1 + c;
        float v = w2[0] * cp[0];
        for (int k = 1; k <= h2; k++) v += w2[k] * (cp[-k * T1] + cp[k * T1]);
        if (ix < n1 && iy < n2) sm_store<EPI>(v, out, q.off + ix + (long)n1 * iy, A, pend, a2, a1, first, acc);
    }
    sm_partials<EPI>(acc, part, ntile, blockIdx.x, sh);
}

// Per-patch reduction of the partials of its blocks (which = 0) or tiles
// (which = 1), then the patch's scalar step.  One wave per patch.
__global__ void __launch_bounds__(64)
bfinalize(const double *part, int stride, int K, int which, const BPatch *__restrict__ P, DipScal *sc, int mode,
          int it, float tol, int k_ls) {
    const int pi = blockIdx.x, lane = threadIdx.x;
    const BPatch q = P[pi];
    const int b0 = which == 0 ? q.blk0 : q.tile0, nb = which == 0 ? q.nblk : q.ntile;
    double tot[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int b = lane; b < nb; b += 64) s += part[(size_t)k * stride + b0 + b];
        tot[k] = wave_sum(s);
    }
    if (lane == 0) dip_scalar_step(sc + pi, tot, mode, it, (long)q.n1 * q.n2, tol, k_ls);
}

__global__ void bfill(float *__restrict__ y, float v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

__global__ void bzero_scal(DipScal *sc, int np) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) {
        DipScal z{};
        sc[p] = z;
    }
}

__global__ void bcollect_rec(const DipScal *sc, int np, int niter, int *rec) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    for (int k = 0; k < 2 * niter; k++) rec[(size_t)p * 2 * niter + k] = sc[p].rec[k];
}

struct BWork {
    float *p[2], *u1, *u2, *w, *xa, *xb, *ra, *rb, *sp, *sx, *sr, *gp, *pc, *dp;
    double *part;
    DipScal *sc;
};

template <int T1, int T2>
int bsmooth_launch(sh_ctx *ctx, const BatchTables &T, int r1, int r2, const float *in, int pro, int epi,
                   const CgArgs &A, BWork &W, int first) {
    int h1 = r1 - 1, h2 = r2 - 1;
    size_t lds = sizeof(float) * ((size_t)(T2 + 2 * h2) * (T1 + 2 * h1) + (size_t)(T2 + 2 * h2) * T1);
    if (lds > 64 * 1024) return sh_fail(ctx, SH_EUNSUP, "dip2d_batched: smoothing radius too large for the tile");
    dim3 grid((unsigned)T.ntile);
    PROF_A(ctx);
    if (pro == PRO_CG)
        hipLaunchKernelGGL((bsmooth2<T1, T2, PRO_CG, EPI_GP>), grid, dim3(TB), lds, ctx->stream, in, nullptr, r1, r2,
                           A, T.P, T.tiles, W.sc, W.part, T.ntile, first);
    else
        hipLaunchKernelGGL((bsmooth2<T1, T2, PRO_PLAIN, EPI_S>), grid, dim3(TB), lds, ctx->stream, in, nullptr, r1,
                           r2, A, T.P, T.tiles, W.sc, W.part, T.ntile, first);
    // algorithmic bytes: CG prologue + gp epilogue (x, r, w, sx, sr, p, sp
    // read; x, r, p, gp written) or the S epilogue (gp, w, sp, sx, sr read;
    // sp, sx, sr written); first iteration has no pending update
    double bps = (pro == PRO_CG) ? (first ? 4.0 * (3 + 1 + 1) : 4.0 * (5 + 2 + 2 + 2))
                                 : (first ? 4.0 * (2 + 3) : 4.0 * (5 + 3));
    PROF_B(ctx, pro == PRO_CG ? "bsmooth2<cg,gp>" : "bsmooth2<plain,s>", bps * (double)T.ntot);
    return sh_launch_check(ctx, "bsmooth2");
}

int bsmooth(sh_ctx *ctx, int cfg, const BatchTables &T, int r1, int r2, const float *in, int pro, int epi,
            const CgArgs &A, BWork &W, int first) {
    switch (cfg) {
    case 0: return bsmooth_launch<32, 16>(ctx, T, r1, r2, in, pro, epi, A, W, first);
    case 1: return bsmooth_launch<64, 8>(ctx, T, r1, r2, in, pro, epi, A, W, first);
    case 2: return bsmooth_launch<64, 16>(ctx, T, r1, r2, in, pro, epi, A, W, first);
    case 3: return bsmooth_launch<64, 32>(ctx, T, r1, r2, in, pro, epi, A, W, first);
    case 4: return bsmooth_launch<128, 8>(ctx, T, r1, r2, in, pro, epi, A, W, first);
    case 5: return bsmooth_launch<128, 16>(ctx, T, r1, r2, in, pro, epi, A, W, first);
    }
    return sh_fail(ctx, SH_EINVAL, "dip2d_batched: tile config %d out of range", cfg);
}

int bfin(sh_ctx *ctx, const BatchTables &T, BWork &W, int K, int which, int mode, int it, int k_ls = 0) {
    int stride = which == 0 ? T.nblk : T.ntile;
    hipLaunchKernelGGL(bfinalize, dim3(T.np), dim3(64), 0, ctx->stream, W.part, stride, K, which, T.P, W.sc, mode,
                       it, 1.e-6f, k_ls);
    return sh_launch_check(ctx, "bfinalize");
}

template <int NW>
int bpwd_launch(sh_ctx *ctx, const BatchTables &T, int mode, const float *x, const float *p, const float *dp,
                float lam, float *pout, BWork &W) {
    const float pmax = 3.402823466e+38f, pmin = -3.402823466e+38f;
    static const char *names[3] = {"bpwd<apply>", "bpwd<der_divne>", "bpwd<trial>"};
    static const double bps[3] = {12.0, 20.0, 20.0};
    dim3 grid((unsigned)T.nblk);
    PROF_A(ctx);
    if (mode == 0)
        hipLaunchKernelGGL((bpwd<NW, 0>), grid, dim3(TB), 0, ctx->stream, x, p, dp, lam, pmin, pmax, pout, W.u1, W.u2,
                           T.P, T.blk_patch, W.part, W.sc);
    else if (mode == 1)
        hipLaunchKernelGGL((bpwd<NW, 1>), grid, dim3(TB), 0, ctx->stream, x, p, dp, lam, pmin, pmax, pout, W.u1, W.u2,
                           T.P, T.blk_patch, W.part, W.sc);
    else
        hipLaunchKernelGGL((bpwd<NW, 2>), grid, dim3(TB), 0, ctx->stream, x, p, dp, lam, pmin, pmax, pout, W.u1, W.u2,
                           T.P, T.blk_patch, W.part, W.sc);
    PROF_B(ctx, names[mode], bps[mode] * (double)T.ntot);
    return sh_launch_check(ctx, "bpwd");
}

// The Gauss-Newton loop of dip3 (dip_cfuns.c:1619-1691) for every patch at
// once; the result is left in W.p[*fin].
template <int NW>
int bdip(sh_ctx *ctx, const BatchTables &T, int cfg, int r1, int r2, const float *x, int niter, int liter, BWork &W,
         int *fin) {
    int cur = 0;
    CgArgs A{W.w, W.xa, W.xb, W.ra, W.rb, W.sx, W.sr, W.sp, W.pc, W.gp};
    dim3 g((unsigned)T.nblk);
    SH_TRY(bpwd_launch<NW>(ctx, T, 0, x, W.p[cur], nullptr, 0.0f, nullptr, W));
    SH_TRY(bfin(ctx, T, W, 1, 0, F_USUM, 0));
    for (int it = 0; it < niter; it++) {
        SH_TRY(bpwd_launch<NW>(ctx, T, 1, x, W.p[cur], nullptr, 0.0f, nullptr, W));
        SH_TRY(bfin(ctx, T, W, 1, 0, F_DIVN, it));
        hipLaunchKernelGGL(bcg_start, g, dim3(TB), 0, ctx->stream, W.u2, W.u1, W.w, W.ra, W.xa, W.pc, T.P, T.blk_patch,
                           W.part, W.sc);
        SH_TRY(sh_launch_check(ctx, "bcg_start"));
        SH_TRY(bfin(ctx, T, W, 1, 0, F_R0, it));
        for (int k = 0; k < liter; k++) {
            SH_TRY(bsmooth(ctx, cfg, T, r1, r2, nullptr, PRO_CG, EPI_GP, A, W, k == 0));
            SH_TRY(bfin(ctx, T, W, 1, 1, F_GN, k));
            SH_TRY(bsmooth(ctx, cfg, T, r1, r2, W.gp, PRO_PLAIN, EPI_S, A, W, k == 0));
            SH_TRY(bfin(ctx, T, W, 3, 1, F_BETA, k));
        }
        hipLaunchKernelGGL(bcg_result, g, dim3(TB), 0, ctx->stream, W.xa, W.xb, W.sx, W.dp, T.P, T.blk_patch, W.sc);
        SH_TRY(sh_launch_check(ctx, "bcg_result"));
        float lam = 1.0f;
        for (int k = 0; k < 8; k++) {
            SH_TRY(bpwd_launch<NW>(ctx, T, 2, x, W.p[cur], W.dp, lam, W.p[1 - cur], W));
            SH_TRY(bfin(ctx, T, W, 1, 0, F_LS, it, k));
            lam *= 0.5f;
        }
        cur = 1 - cur;
    }
    *fin = cur;
    return SH_OK;
}

// ---------------------------------------------------------------------------
// Batched spray + median (csomf2d per patch).  Layouts: every patch is
// transposed to trace-fast ([k][i2]) at its own offset; slot s of the
// whole batch is the plane U + s * ntot; factor variant v, plane m at
// F + (v * (NB+1) + m) * ntot.
// ---------------------------------------------------------------------------
__global__ void btranspose(const float *__restrict__ src, float *__restrict__ dst, const BPatch *__restrict__ P,
                           const Tile *__restrict__ tt, int to_fast) {
    __shared__ float tile[32][33];
    const Tile t = tt[blockIdx.x];
    const BPatch q = P[t.patch];
    const int n1 = q.n1, ntr = q.n2;
    const int t0 = t.x0, k0 = t.y0;  // trace block, sample block
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    const float *s = src + q.off;
    float *d = dst + q.off;
    if (to_fast) {
        for (int r = ty; r < 32; r += 8) {
            int tr = t0 + r, k = k0 + tx;
            if (tr < ntr && k < n1) tile[r][tx] = s[(long)tr * n1 + k];
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            int k = k0 + r, tr = t0 + tx;
            if (tr < ntr && k < n1) d[(long)k * ntr + tr] = tile[tx][r];
        }
    } else {
        for (int r = ty; r < 32; r += 8) {
            int k = k0 + r, tr = t0 + tx;
            if (tr < ntr && k < n1) tile[r][tx] = s[(long)k * ntr + tr];
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            int tr = t0 + r, k = k0 + tx;
            if (tr < ntr && k < n1) d[(long)tr * n1 + k] = tile[tx][r];
        }
    }
}

#include "predict.hpp"

// Both factor variants of every trace (v = 0: forw = false with the
// target's slope; v = 1: forw = true with the slope of the trace before it).
template <int NW>
__global__ void __launch_bounds__(TB)
bfactor2(const float *__restrict__ dipT, float *__restrict__ F, const BPatch *__restrict__ P,
         const int *__restrict__ cblk_patch, long ntot, float e) {
    const int pi = cblk_patch[blockIdx.x];
    const BPatch q = P[pi];
    const int gid = (blockIdx.x - q.cblk0) * TB + threadIdx.x;
    if (gid >= 2 * q.n2) return;
    const int v = gid / q.n2, i = gid % q.n2;
    if (v == 0 ? i > q.n2 - 2 : i < 1) return;  // never a target in that direction
    constexpr int NB = 2 * NW;
    Acc qa[1] = {Acc{dipT + q.off + (v == 0 ? i : i - 1), q.n2}};
    bool rev[1] = {v == 1};
    const FactorOut Fs = fac_out_at(F + (long)v * fac_plane_floats(NB, ntot), NB, ntot, q.off + i, q.n2);
    factor_sweep_pf<NW, 1, 8>(qa, rev, q.n1, e, Fs);
}

// Forward spray chains of pwspray_lop (sof_cfuns.c:998-1028), one thread per
// (source trace, direction) of every patch.
template <int NW>
__global__ void __launch_bounds__(TB)
bspray2(const float *__restrict__ xT, const float *__restrict__ dipT, const float *__restrict__ F,
        float *__restrict__ U, const BPatch *__restrict__ P, const int *__restrict__ cblk_patch, long ntot, int ns,
        float e, float *__restrict__ scr, long nthr) {
    const int pi = cblk_patch[blockIdx.x];
    const BPatch q = P[pi];
    const int gid = (blockIdx.x - q.cblk0) * TB + threadIdx.x;
    if (gid >= 2 * q.n2) return;
    const int n1 = q.n1, n2 = q.n2;
    const int dir = gid / n2, i = gid % n2;
    const float *xs = xT + q.off;
    const float *ds = dipT + q.off;
    constexpr int NB = 2 * NW;
    const long tid = (long)blockIdx.x * TB + threadIdx.x;
    Sink y{scr + tid, nthr};
    if (dir == 0) {  // centre slot written once
        Sink c{U + (long)ns * ntot + q.off + i, n2};
        for (int k = 0; k < n1; k++) c[k] = xs[(long)k * n2 + i];
    }
    Acc cur{xs + i, n2};
    for (int is = 0; is < ns; is++) {
        int ip = dir == 0 ? i - is - 1 : i + is + 1;
        if (ip < 0 || ip >= n2) break;
        int slot = dir == 0 ? ns - is - 1 : ns + is + 1;
        Acc qa{ds + (dir == 0 ? ip : ip - 1), n2};
        Sink out{U + (long)slot * ntot + q.off + ip, n2};
        Acc in[1] = {cur};
        Acc qs[1] = {qa};
        bool rv[1] = {dir == 1};
        const Factor Fv = fac_at(F + (long)dir * fac_plane_floats(NB, ntot), NB, ntot, q.off + ip, n2);
        solve_factored_pf<NW, 1, 8>(in, qs, rv, Fv, y, n1, e, out);
        cur = Acc{out.p, n2};
    }
}

__global__ void bzero(float *__restrict__ y, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0.0f;
}

int btrans(sh_ctx *ctx, const BatchTables &T, const float *src, float *dst, bool to_fast) {
    hipLaunchKernelGGL(btranspose, dim3((unsigned)T.nttile), dim3(32, 8), 0, ctx->stream, src, dst, T.P, T.ttiles,
                       to_fast ? 1 : 0);
    return sh_launch_check(ctx, "btranspose");
}

}  // namespace

extern "C" {

int sh_batch_tile_count(void) { return NTILE_CFG; }

int sh_batch_tile_shape(int cfg, int *t1, int *t2) {
    if (cfg == -1) cfg = DEFAULT_TILE;
    if (cfg < 0 || cfg >= NTILE_CFG || !t1 || !t2) return SH_EINVAL;
    *t1 = TILE_CFG[cfg][0];
    *t2 = TILE_CFG[cfg][1];
    return SH_OK;
}

int sh_dip2d_batched(sh_ctx *ctx, const float *data, const long *offs, const int *n1s, const int *n2s, int npatch,
                     int niter, int liter, int order, int r1, int r2, int tile, const float *init, float *dip,
                     int *rec) {
    if (!ctx) return SH_ENULL;
    if (!data || !dip) return sh_fail(ctx, SH_ENULL, "dip2d_batched: data and dip must be non-NULL");
    SH_TRY(check_patches(ctx, offs, n1s, n2s, npatch, 1, "dip2d_batched"));
    if (niter < 0 || liter < 0) return sh_fail(ctx, SH_EINVAL, "dip2d_batched: niter and liter must be >= 0");
    if (rec && niter > 64) return sh_fail(ctx, SH_EINVAL, "dip2d_batched: the record holds at most 64 iterations");
    if (r1 < 1 || r2 < 1 || r1 > SM_HMAX + 1 || r2 > SM_HMAX + 1)
        return sh_fail(ctx, SH_EINVAL, "dip2d_batched: bad rect %d %d", r1, r2);
    if (tile < 0) tile = DEFAULT_TILE;
    if (tile >= NTILE_CFG) return sh_fail(ctx, SH_EINVAL, "dip2d_batched: tile config %d out of range", tile);
    BatchTables T;
    SH_TRY(build_tables(ctx, offs, n1s, n2s, npatch, TILE_CFG[tile][0], TILE_CFG[tile][1], T));
    long n = T.ntot;
    const int NF = 15;
    float *fbuf = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_BATCH0 + 1, (size_t)n * NF * sizeof(float), (void **)&fbuf));
    BWork W{};
    float **slots[NF] = {&W.p[0], &W.p[1], &W.u1, &W.u2, &W.w, &W.xa, &W.xb, &W.ra,
                         &W.rb, &W.sp, &W.sx, &W.sr, &W.gp, &W.pc, &W.dp};
    for (int s = 0; s < NF; s++) *slots[s] = fbuf + (size_t)s * n;
    int pmax = std::max(T.nblk, T.ntile);
    SH_TRY(sh_slot(ctx, SLOT_BATCH0 + 2, (size_t)pmax * 3 * sizeof(double), (void **)&W.part));
    SH_TRY(sh_slot(ctx, SLOT_BATCH0 + 3, (size_t)npatch * sizeof(DipScal), (void **)&W.sc));
    // unused gaps between patches stay zero in every field
    hipLaunchKernelGGL(bfill, dim3(ceil_div(n * NF, TB)), dim3(TB), 0, ctx->stream, fbuf, 0.0f, n * NF);
    SH_TRY(sh_launch_check(ctx, "bfill"));
    hipLaunchKernelGGL(bzero_scal, dim3(ceil_div(npatch, 64)), dim3(64), 0, ctx->stream, W.sc, npatch);
    SH_TRY(sh_launch_check(ctx, "bzero_scal"));
    if (init) SH_HIP(ctx, hipMemcpyAsync(W.p[0], init, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    int fin = 0;
    SH_DISPATCH_NW(ctx, order, NW, SH_TRY(bdip<NW>(ctx, T, tile, r1, r2, data, niter, liter, W, &fin)));
    SH_HIP(ctx, hipMemcpyAsync(dip, W.p[fin], (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    if (rec && niter > 0) {
        int *drec = nullptr;
        SH_TRY(sh_slot(ctx, SLOT_BATCH0 + 4, (size_t)npatch * 2 * niter * sizeof(int), (void **)&drec));
        hipLaunchKernelGGL(bcollect_rec, dim3(ceil_div(npatch, 64)), dim3(64), 0, ctx->stream, W.sc, npatch, niter,
                           drec);
        SH_TRY(sh_launch_check(ctx, "bcollect_rec"));
        SH_HIP(ctx, hipMemcpyAsync(rec, drec, (size_t)npatch * 2 * niter * sizeof(int), hipMemcpyDeviceToHost,
                                   ctx->stream));
        SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return SH_OK;
}

}  // extern "C"

namespace {
#include "svmf.hpp"

// space-varying median of every trace of every patch (chain-block grid of
// bspray2: n2 threads of each patch's blocks)
__global__ void bsvmf(const float *__restrict__ U, float *__restrict__ out, const BPatch *__restrict__ P,
                      const int *__restrict__ cblk_patch, long ntot, int np, int nfw) {
    const BPatch q = P[cblk_patch[blockIdx.x]];
    const int gid = (blockIdx.x - q.cblk0) * TB + threadIdx.x;
    if (gid >= q.n2) return;
    svmf_one(U, out, ntot, np, q.n1, nfw, q.off + gid, q.n2);
}
}  // namespace

extern "C" {

int sh_somf2d_batched(sh_ctx *ctx, const float *data, const float *dip, const long *offs, const int *n1s,
                      const int *n2s, int npatch, int ns, int nmf, int option, int order, float eps, float *out) {
    if (!ctx) return SH_ENULL;
    if (!data || !dip || !out) return sh_fail(ctx, SH_ENULL, "somf2d_batched: NULL pointer");
    SH_TRY(check_patches(ctx, offs, n1s, n2s, npatch, 8, "somf2d_batched"));
    if (ns < 0 || ns > 40) return sh_fail(ctx, SH_EINVAL, "somf2d_batched: bad ns %d", ns);
    int np = 2 * ns + 1;
    if (nmf < 1 || nmf % 2 == 0 || nmf > np)
        return sh_fail(ctx, SH_EINVAL, "somf2d_batched: median length %d must be odd and <= %d", nmf, np);
    if (option != 1 && option != 2) return sh_fail(ctx, SH_EINVAL, "somf2d_batched: option must be 1 or 2");
    if (option == 2 && (nmf < 5 || nmf + 2 > 128))
        return sh_fail(ctx, SH_EINVAL, "somf2d_batched: the space-varying median needs 5 <= nmf <= 125");
    if (order < 1 || order > 4) return sh_fail(ctx, SH_EUNSUP, "somf2d_batched: PWD order %d not supported", order);
    BatchTables T;
    SH_TRY(build_tables(ctx, offs, n1s, n2s, npatch, TILE_CFG[DEFAULT_TILE][0], TILE_CFG[DEFAULT_TILE][1], T));
    long n = T.ntot;
    float *b = nullptr, *U = nullptr, *F = nullptr, *scr = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_BATCH0 + 5, (size_t)n * 3 * sizeof(float), (void **)&b));
    float *xT = b, *dT = b + n, *yT = b + 2 * n;
    SH_TRY(sh_slot(ctx, SLOT_BATCH0 + 6, (size_t)n * np * sizeof(float), (void **)&U));
    SH_TRY(sh_slot(ctx, SLOT_BATCH0 + 7, (size_t)2 * fac_plane_floats(2 * order, n) * sizeof(float), (void **)&F));
    long nthr = (long)T.ncblk * TB;
    SH_TRY(sh_slot(ctx, SLOT_BATCH0 + 8, (size_t)nthr * T.n1max * sizeof(float), (void **)&scr));
    hipLaunchKernelGGL(bzero, dim3(ceil_div(n * 3, TB)), dim3(TB), 0, ctx->stream, b, n * 3);
    SH_TRY(sh_launch_check(ctx, "bzero"));
    SH_TRY(btrans(ctx, T, dip, dT, true));
    SH_TRY(btrans(ctx, T, data, xT, true));
    const float e = eps * eps;
    dim3 cg((unsigned)T.ncblk);
    PROF_A(ctx);
    SH_DISPATCH_NW(ctx, order, NW,
                   hipLaunchKernelGGL((bfactor2<NW>), cg, dim3(TB), 0, ctx->stream, dT, F, T.P, T.cblk_patch, n, e));
    PROF_B(ctx, "bfactor2", 4.0 * (double)n);
    SH_TRY(sh_launch_check(ctx, "bfactor2"));
    hipLaunchKernelGGL(bzero, dim3(ceil_div(n * np, TB)), dim3(TB), 0, ctx->stream, U, n * np);
    SH_TRY(sh_launch_check(ctx, "bzero"));
    {
        PROF_A(ctx);
        SH_DISPATCH_NW(ctx, order, NW,
                       hipLaunchKernelGGL((bspray2<NW>), cg, dim3(TB), 0, ctx->stream, xT, dT, F, U, T.P, T.cblk_patch,
                                          n, ns, e, scr, nthr));
        PROF_B(ctx, "bspray2", 12.0 * 2 * ns * (double)n);
        SH_TRY(sh_launch_check(ctx, "bspray2"));
    }
    int m = (nmf - 1) / 2;
    if (option == 2) {
        hipLaunchKernelGGL(bsvmf, cg, dim3(TB), 0, ctx->stream, U, yT, T.P, T.cblk_patch, n, np, nmf);
        SH_TRY(sh_launch_check(ctx, "bsvmf"));
    } else {
        SH_TRY(sh_median_internal(ctx, U, yT, n, ns - m, nmf));
    }
    return btrans(ctx, T, yT, out, false);
}

}  // extern "C"
