// Local-slope estimation on gfx950: plane-wave destruction (PWD) stencils,
// triangle shaping smoothing, the shaping-regularised conjugate-gradient
// smooth division and the Gauss-Newton loop with backtracking line search.
//
// Algorithm: dip3 / ps_divne / ps_conjgrad / ps_trianglen_lop / allpass1-2
// of pyseistr/src/dip_cfuns.c (cited per kernel).  Design:
//  * every kernel streams HBM once, coalesced along n1;
//  * the CG keeps all scalars on the device, so a whole dip estimate is
//    queued without host round trips; convergence and line-search
//    acceptance are device flags that turn later launches into no-ops;
//  * the triangle smoother is the direct (2nb-1)-tap stencil with even
//    reflection, staged through LDS tiles with halos, the vector updates of
//    the CG fused into the first/last smoothing pass.
#include "dipcore.hpp"

#include <algorithm>
#include <cstdlib>

namespace {

using namespace dipcore;


// ---------------------------------------------------------------------------
// Reductions + scalar logic, one block of 1024 threads.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024) dip_finalize(const double *part, int nblk, int K, DipScal *sc,
                                                     int mode, int it, long n, float tol, int k_ls) {
    __shared__ double sh[16];
    __shared__ double tot[4];
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += part[(size_t)k * nblk + i];
        double t = block_sum(s, sh);
        if (threadIdx.x == 0) tot[k] = t;
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    dip_scalar_step(sc, tot, mode, it, n, tol, k_ls);
}

// ---------------------------------------------------------------------------
// PWD stencil kernels (allpass1 / allpass2, dip_cfuns.c:1135-1200, 1399-1465)
// MODE 0: u2 = C(p) x, sum u2^2
// MODE 1: u1 = C'(p) x; divne prologue on (num=u2, den=u1): mask, then
//         scale by 1/hypot(den, 1) (dip_cfuns.c:1656-1663, 802-808), sum den^2
// MODE 2: line-search trial p = clip(p0 + lam dp), u2 = C(p) x, sum u2^2
// ---------------------------------------------------------------------------
template <int NW, int MODE>
__global__ void __launch_bounds__(256)
pwd_kernel(const float *__restrict__ x, const float *__restrict__ p, const float *__restrict__ dp,
           float lam, float pmin, float pmax, float *__restrict__ pout, float *__restrict__ u1,
           float *__restrict__ u2, const unsigned char *__restrict__ mask, int n1, int n2, int n3,
           int axis, double *part, const DipScal *sc, FastDiv f1, FastDiv f2) {
    if (MODE == 2 && sc->accepted) return;
    __shared__ double sh[16];
    const long n = (long)n1 * n2 * n3;
    const long ip = (axis == 2) ? n1 : (long)n1 * n2;
    double acc = 0.0;
    // each block owns a contiguous range of samples (launch-cheap when the
    // device flags make the launch a no-op)
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long i0 = (long)blockIdx.x * chunk, i1e = min(n, i0 + chunk);
    const bool fast = n < (1L << 31);  // multiply-high division (no long division per sample)
    for (long i = i0 + threadIdx.x; i < i1e; i += blockDim.x) {
        int i1, i2, i3;
        if (fast) {
            const unsigned r = f1.div((unsigned)i), q = f2.div(r);
            i1 = (int)((unsigned)i - r * (unsigned)n1);
            i2 = (int)(r - q * (unsigned)n2);
            i3 = (int)q;
        } else {
            i1 = (int)(i % n1);
            const long r = i / n1;
            i2 = (int)(r % n2);
            i3 = (int)(r / n2);
        }
        bool valid = i1 >= NW && i1 < n1 - NW && (axis == 2 ? i2 < n2 - 1 : i3 < n3 - 1);
        acc += pwd_sample<NW, MODE>(x, p, dp, lam, pmin, pmax, pout, u1, u2, mask && mask[i], i, ip, valid);
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// CG start (ps_divne tail + ps_conjgrad head, dip_cfuns.c:819-826, 284-303):
// w = den*norm, r = -(num*norm), x = p = 0; partial sum r^2.
__global__ void __launch_bounds__(256)
cg_start_kernel(const float *__restrict__ num, const float *__restrict__ den, float *__restrict__ w,
                float *__restrict__ r, float *__restrict__ x, float *__restrict__ pc, long n,
                double *part, const DipScal *sc) {
    __shared__ double sh[16];
    double acc = 0.0;
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long i0 = (long)blockIdx.x * chunk, i1e = min(n, i0 + chunk);
    for (long i = i0 + threadIdx.x; i < i1e; i += blockDim.x) acc += cg_start_sample(num, den, w, r, x, pc, i, sc);
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// dp = x (+ alpha2 sx when the last update is still pending)
__global__ void __launch_bounds__(256)
cg_result_kernel(const float *__restrict__ xa, const float *__restrict__ xb, const float *__restrict__ sx,
                 float *__restrict__ dp, long n, const DipScal *sc) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        cg_result_sample(xa, xb, sx, dp, i, sc);
}

// ---------------------------------------------------------------------------
// Triangle smoothing passes with fused CG prologue/epilogue.
// ---------------------------------------------------------------------------



// Axis 1 (contiguous): task = one chunk of one trace; a block runs a
// contiguous range of tasks (ntask = chunks * traces).
template <int PRO, int EPI, bool NT = false>
__global__ void __launch_bounds__(256)
smooth_contig(const float *__restrict__ in, float *__restrict__ out, int n1, long ntask, int nb, CgArgs A,
              const DipScal *sc, double *part, int first) {
    if (sc->cg_done) return;
    __shared__ float tile[SM_CH1 + 2 * SM_HMAX];
    __shared__ float wts[SM_HMAX + 1];
    __shared__ double sh[16];
    int h = nb - 1;
    int nch = (n1 + SM_CH1 - 1) / SM_CH1;
    bool pend = (PRO == PRO_CG || EPI == EPI_GP) ? (sc->pending != 0) : false;
    float a2 = pend ? (float)sc->alpha2 : 0.0f;
    float a1 = (EPI == EPI_S) ? (float)sc->alpha : 0.0f;
    int cur = (PRO == PRO_CG) ? sc->cur : 0;
    sm_weights(wts, nb);
    double acc[3] = {0.0, 0.0, 0.0};
    const long per = (ntask + gridDim.x - 1) / gridDim.x;
    const long t0 = (long)blockIdx.x * per, t1 = min(ntask, t0 + per);
    // the next task's inputs are loaded into registers while the current
    // task is smoothed (up to KS samples per thread: SM_CH1 + 2 h <= 766)
    constexpr int KS = (SM_CH1 + 2 * SM_HMAX + 255) / 256;
    constexpr int NR = (PRO == PRO_PLAIN) ? 1 : 5;
    float raw[KS][NR];
    long gi[KS];
    bool own[KS];
    auto fetch = [&](long task) {
        const long base = (task / nch) * n1;
        const int c0 = (int)(task % nch) * SM_CH1;
        const int len = min(SM_CH1, n1 - c0);
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int t = threadIdx.x + k * 256;
            if (t < len + 2 * h) {
                const int pos = c0 - h + t;
                gi[k] = base + fold_idx(pos, n1);
                own[k] = pos >= c0 && pos < c0 + len;
                sm_fetch<PRO>(in, gi[k], A, pend, cur, raw[k]);
            }
        }
    };
    if (t0 < t1) fetch(t0);
    for (long task = t0; task < t1; task++) {
        long base = (task / nch) * n1;
        int c0 = (int)(task % nch) * SM_CH1;
        int len = min(SM_CH1, n1 - c0);
        __syncthreads();  // the previous task's readers are done with tile (and wts is ready)
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int t = threadIdx.x + k * 256;
            if (t < len + 2 * h) tile[t] = sm_finish<PRO>(raw[k], gi[k], own[k], A, pend, a2, cur);
        }
        __syncthreads();
        if (task + 1 < t1) fetch(task + 1);
        for (int t = threadIdx.x; t < len; t += blockDim.x) {
            const float *c = tile + t + h;
            float s = wts[0] * c[0];
            for (int k = 1; k <= h; k++) s += wts[k] * (c[-k] + c[k]);
            sm_store<EPI, NT>(s, out, base + c0 + t, A, pend, a2, a1, first, acc);
        }
    }
    sm_partials<EPI>(acc, part, gridDim.x, blockIdx.x, sh);
}

// The axis-1 pass without a CG epilogue (plain, or with the CG prologue)
// with one wave per task: each wave stages its chunk (and halo) in its own
// LDS row, so no block-wide barrier is needed, and loads the next task's
// samples into registers while it smooths the current one (up to 12 per
// lane).  Same operations as smooth_contig<PRO, EPI_PLAIN>, including the
// prologue's write-back of the owned samples.
template <int PRO, bool NT>
__global__ void __launch_bounds__(256)
smooth_contig_wave(const float *__restrict__ in, float *__restrict__ out, int n1, long ntask, int nb, CgArgs A,
                   const DipScal *sc) {
    if (sc->cg_done) return;
    constexpr int LW = SM_CH1 + 2 * SM_HMAX, KS = (LW + 63) / 64;
    constexpr int NR = (PRO == PRO_PLAIN) ? 1 : 5;
    __shared__ float tile[4][LW];
    __shared__ float wts[SM_HMAX + 1];
    const int h = nb - 1;
    const int nch = (n1 + SM_CH1 - 1) / SM_CH1;
    const bool pend = (PRO == PRO_CG) ? (sc->pending != 0) : false;
    const float a2 = pend ? (float)sc->alpha2 : 0.0f;
    const int cur = (PRO == PRO_CG) ? sc->cur : 0;
    sm_weights(wts, nb);
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float *tl = tile[wv];
    const long per = (ntask + gridDim.x - 1) / gridDim.x;
    const long t0 = (long)blockIdx.x * per, t1 = min(ntask, t0 + per);
    float raw[KS][NR];
    long gi[KS];
    bool own[KS];
    auto fetch = [&](long task) {
        const long base = (task / nch) * n1;
        const int c0 = (int)(task % nch) * SM_CH1;
        const int len = min(SM_CH1, n1 - c0);
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int t = lane + 64 * k;
            if (t < len + 2 * h) {
                const int pos = c0 - h + t;
                gi[k] = base + fold_idx(pos, n1);
                own[k] = pos >= c0 && pos < c0 + len;
                sm_fetch<PRO>(in, gi[k], A, pend, cur, raw[k]);
            }
        }
    };
    if (t0 + wv < t1) fetch(t0 + wv);
    for (long task = t0 + wv; task < t1; task += 4) {
        const long base = (task / nch) * n1;
        const int c0 = (int)(task % nch) * SM_CH1;
        const int len = min(SM_CH1, n1 - c0);
        // the wave's reads of its previous task are issued before these
        // writes, and LDS instructions of one wave complete in order
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int t = lane + 64 * k;
            if (t < len + 2 * h) tl[t] = sm_finish<PRO>(raw[k], gi[k], own[k], A, pend, a2, cur);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (task + 4 < t1) fetch(task + 4);
        for (int t = lane; t < len; t += 64) {
            const float *c = tl + t + h;
            float v = wts[0] * c[0];
            for (int k = 1; k <= h; k++) v += wts[k] * (c[-k] + c[k]);
            if constexpr (NT) st_nt(&out[base + c0 + t], v);
            else out[base + c0 + t] = v;
        }
    }
}

// The plain strided pass (axis 2 or 3) with one wave per task and each
// lane owning one i1 column of it: the lane keeps the task's CH samples
// along the axis and their halo in registers (no LDS, no barrier).  Same
// stencil operations as smooth_strided<PRO_PLAIN, EPI_PLAIN>.
template <int H, int CH>
__global__ void __launch_bounds__(256)
smooth_strided_col(const float *__restrict__ in, float *__restrict__ out, int n1, int naxis, long stride,
                   long ostride, int nother, const DipScal *sc) {
    if (sc->cg_done) return;
    constexpr int nb = H + 1;
    float wts[H + 1];
#pragma unroll
    for (int k = 0; k <= H; k++) wts[k] = (float)((double)(nb - k) / ((double)nb * nb));
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6), nwave = (long)gridDim.x * 4;
    const int ntx = (n1 + 63) / 64, nty = (naxis + CH - 1) / CH;
    const long ntask = (long)ntx * nty * nother;
    for (long task = wave; task < ntask; task += nwave) {
        const int bx = (int)(task % ntx);
        const long rest = task / ntx;
        const int by = (int)(rest % nty), bz = (int)(rest / nty);
        const int i1 = bx * 64 + lane;
        const int c0 = by * CH, len = min(CH, naxis - c0);
        const long base = (long)bz * ostride + i1;
        if (i1 >= n1) continue;
        float v[CH + 2 * H];
#pragma unroll
        for (int r = 0; r < CH + 2 * H; r++)
            v[r] = (r < len + 2 * H) ? in[base + (long)fold_idx(c0 - H + r, naxis) * stride] : 0.0f;
#pragma unroll
        for (int r = 0; r < CH; r++) {
            if (r < len) {
                float t = wts[0] * v[r + H];
#pragma unroll
                for (int k = 1; k <= H; k++) t += wts[k] * (v[r + H - k] + v[r + H + k]);
                out[base + (long)(c0 + r) * stride] = t;
            }
        }
    }
}

// Axes 2/3 (strided): task = tw = blockDim.x consecutive i1 x a chunk of chs
// samples along the axis x one position of the other strided axis (ntx x
// nty x nother tasks, i1 chunk fastest); index = i1 + o*ostride + q*stride.
template <int PRO, int EPI, bool NT = false>
__global__ void __launch_bounds__(256)
smooth_strided(const float *__restrict__ in, float *__restrict__ out, int n1, int naxis, long stride,
               long ostride, int nother, int nb, int chs, CgArgs A, const DipScal *sc, double *part, int first) {
    if (sc->cg_done) return;
    extern __shared__ float tile[];  // [(chs + 2h) rows][tw]
    __shared__ float wts[SM_HMAX + 1];
    __shared__ double sh[16];
    int h = nb - 1;
    const int tx = threadIdx.x, ty = threadIdx.y, tw = blockDim.x;
    bool pend = (PRO == PRO_CG || EPI == EPI_GP) ? (sc->pending != 0) : false;
    float a2 = pend ? (float)sc->alpha2 : 0.0f;
    float a1 = (EPI == EPI_S) ? (float)sc->alpha : 0.0f;
    int cur = (PRO == PRO_CG) ? sc->cur : 0;
    sm_weights(wts, nb);
    const int ntx = (n1 + tw - 1) / tw, nty = (naxis + chs - 1) / chs;
    const long ntask = (long)ntx * nty * nother;
    const long per = (ntask + gridDim.x - 1) / gridDim.x;
    const long t0 = (long)blockIdx.x * per, t1 = min(ntask, t0 + per);
    double acc[3] = {0.0, 0.0, 0.0};
    // plain-in/plain-out passes: the next task's rows are loaded into
    // registers while the current task is smoothed (up to KR rows per
    // thread).  Not in the passes with a CG epilogue: their extra registers
    // cost more occupancy than the prefetch gains (measured).
    constexpr int KR = 18;
    constexpr bool PIPE = PRO == PRO_PLAIN && EPI == EPI_PLAIN;
    const bool pipe = PIPE && (chs + 2 * h + (int)blockDim.y - 1) / (int)blockDim.y <= KR;
    float nxt[PIPE ? KR : 1];
    auto fetch = [&](long task) {
        const int bx = (int)(task % ntx);
        const long rest = task / ntx;
        const int by = (int)(rest % nty), bz = (int)(rest / nty);
        const int i1 = bx * tw + tx;
        const int c0 = by * chs, len = min(chs, naxis - c0);
        const long base = (long)bz * ostride + i1;
        if (!PIPE || i1 >= n1) return;
#pragma unroll
        for (int k = 0; k < (PIPE ? KR : 1); k++) {
            const int row = ty + k * (int)blockDim.y;
            if (row < len + 2 * h) nxt[k] = in[base + (long)fold_idx(c0 - h + row, naxis) * stride];
        }
    };
    if (pipe && t0 < t1) fetch(t0);
    for (long task = t0; task < t1; task++) {
        int bx = (int)(task % ntx);
        long rest = task / ntx;
        int by = (int)(rest % nty), bz = (int)(rest / nty);
        int i1 = bx * tw + tx;
        int c0 = by * chs;
        int len = min(chs, naxis - c0);
        long base = (long)bz * ostride + i1;
        __syncthreads();
        if (i1 < n1) {
            if (pipe) {
#pragma unroll
                for (int k = 0; k < (PIPE ? KR : 1); k++) {
                    const int row = ty + k * (int)blockDim.y;
                    if (row < len + 2 * h) tile[row * tw + tx] = nxt[k];
                }
            } else {
                for (int row = ty; row < len + 2 * h; row += blockDim.y) {
                    int pos = c0 - h + row;
                    int q = fold_idx(pos, naxis);
                    bool own = pos >= c0 && pos < c0 + len;
                    tile[row * tw + tx] = sm_load<PRO>(in, base + (long)q * stride, own, A, pend, a2, cur);
                }
            }
        }
        __syncthreads();
        if (pipe && task + 1 < t1) fetch(task + 1);
        if (i1 < n1) {
            for (int row = ty; row < len; row += blockDim.y) {
                const float *c = tile + (row + h) * tw + tx;
                float s = wts[0] * c[0];
                for (int k = 1; k <= h; k++) s += wts[k] * (c[-k * tw] + c[k * tw]);
                sm_store<EPI, NT>(s, out, base + (long)(c0 + row) * stride, A, pend, a2, a1, first, acc);
            }
        }
    }
    sm_partials<EPI>(acc, part, gridDim.x, blockIdx.x, sh);
}

// All axes in one pass ("2.5-D"): a block owns a TI1 x TI2 tile of
// (i1, i2) outputs and marches along i3 over a segment of planes.  Each
// input plane (with halos, folded at the volume edges) is smoothed along i1
// and i2 in LDS and pushed into a ring of 2*h3+1 planes; the i3 stencil is
// applied from the ring.  Axes with radius 1 have h = 0 (identity).
constexpr int F_TI1 = 64, F_TI2 = 8;

template <int PRO, int EPI>
__global__ void __launch_bounds__(256)
smooth_fused(const float *__restrict__ in, float *__restrict__ out, int n1, int n2, int n3, int r1, int r2, int r3,
             int seg, CgArgs A, const DipScal *sc, double *part, int first) {
    if (sc->cg_done) return;
    extern __shared__ float lds[];
    __shared__ float w1[SM_HMAX + 1], w2[SM_HMAX + 1], w3[SM_HMAX + 1];
    __shared__ double sh[16];
    const int h1 = r1 - 1, h2 = r2 - 1, h3 = r3 - 1;
    const int W1 = F_TI1 + 2 * h1, R2 = F_TI2 + 2 * h2, RING = 2 * h3 + 1;
    float *tile = lds;                 // [R2][W1]
    float *s1 = tile + R2 * W1;        // [R2][F_TI1]
    float *ring = s1 + R2 * F_TI1;     // [RING][F_TI2][F_TI1]
    const int tx = threadIdx.x, ty = threadIdx.y;  // 64 x 4
    const int ntile1 = (n1 + F_TI1 - 1) / F_TI1;
    const int x0 = (blockIdx.x % ntile1) * F_TI1, y0 = (blockIdx.x / ntile1) * F_TI2;
    const int z0 = blockIdx.y * seg, z1 = min(n3, z0 + seg);
    const long n12 = (long)n1 * n2;
    bool pend = (PRO == PRO_CG || EPI == EPI_GP) ? (sc->pending != 0) : false;
    float a2 = pend ? (float)sc->alpha2 : 0.0f;
    float a1 = (EPI == EPI_S) ? (float)sc->alpha : 0.0f;
    int cur = (PRO == PRO_CG) ? sc->cur : 0;
    {
        int t = tx + 64 * ty;
        for (int k = t; k <= SM_HMAX; k += 256) {
            w1[k] = (k < r1) ? (float)((double)(r1 - k) / ((double)r1 * r1)) : 0.0f;
            w2[k] = (k < r2) ? (float)((double)(r2 - k) / ((double)r2 * r2)) : 0.0f;
            w3[k] = (k < r3) ? (float)((double)(r3 - k) / ((double)r3 * r3)) : 0.0f;
        }
    }
    double acc[3] = {0.0, 0.0, 0.0};
    for (int q = z0 - h3; q < z1 + h3; q++) {
        const int p = fold_idx(q, n3);
        const bool ownz = q >= z0 && q < z1;
        // 1. load the plane tile with halos
        for (int r = ty; r < R2; r += 4) {
            int yy = y0 - h2 + r;
            int iy = fold_idx(yy, n2);
            for (int c = tx; c < W1; c += 64) {
                int xx = x0 - h1 + c;
                int ix = fold_idx(xx, n1);
                bool own = ownz && xx >= x0 && xx < x0 + F_TI1 && xx < n1 && yy >= y0 && yy < y0 + F_TI2 && yy < n2;
                tile[r * W1 + c] = sm_load<PRO>(in, ix + (long)n1 * iy + n12 * p, own, A, pend, a2, cur);
            }
        }
        __syncthreads();
        // 2. along i1
        for (int r = ty; r < R2; r += 4) {
            const float *c = tile + r * W1 + tx + h1;
            float s = w1[0] * c[0];
            for (int k = 1; k <= h1; k++) s += w1[k] * (c[-k] + c[k]);
            s1[r * F_TI1 + tx] = s;
        }
        __syncthreads();
        // 3. along i2, into the ring
        const int slot = (q - (z0 - h3)) % RING;
        for (int r = ty; r < F_TI2; r += 4) {
            const float *c = s1 + (r + h2) * F_TI1 + tx;
            float s = w2[0] * c[0];
            for (int k = 1; k <= h2; k++) s += w2[k] * (c[-k * F_TI1] + c[k * F_TI1]);
            ring[(slot * F_TI2 + r) * F_TI1 + tx] = s;
        }
        __syncthreads();
        // 4. along i3 for plane o = q - h3
        const int o = q - h3;
        if (o >= z0) {
            const int so = (o - (z0 - h3)) % RING;
            for (int r = ty; r < F_TI2; r += 4) {
                int ix = x0 + tx, iy = y0 + r;
                float s = w3[0] * ring[(so * F_TI2 + r) * F_TI1 + tx];
                for (int k = 1; k <= h3; k++) {
                    int sm = (so - k + RING) % RING, spl = (so + k) % RING;
                    s += w3[k] * (ring[(sm * F_TI2 + r) * F_TI1 + tx] + ring[(spl * F_TI2 + r) * F_TI1 + tx]);
                }
                if (ix < n1 && iy < n2) sm_store<EPI>(s, out, ix + (long)n1 * iy + n12 * o, A, pend, a2, a1, first, acc);
            }
        }
        // the next plane's loads may overwrite tile/s1 only after everyone
        // has finished with them; the ring slot reused next is not read here
        __syncthreads();
    }
    sm_partials<EPI>(acc, part, gridDim.x * gridDim.y, blockIdx.x + gridDim.x * blockIdx.y, sh);
}

// Axes 2 and 3 in one pass: a block owns a 64 (i1) x P_T2 (i2) tile and
// marches along i3 over a segment of planes.  Each input plane's rows (with
// the i2 halo, folded at the edges) go through LDS, double-buffered so the
// next plane's loads are in flight while the current one is smoothed along
// i2; every thread keeps the i2-smoothed values of its columns for the last
// 2*H3+1 planes in a register ring and applies the i3 stencil from it.  The
// per-axis sums are those of the separable passes in the same order.
constexpr int P_HMAX2 = 8;

template <int EPI, int H3, int P_T2>
__global__ void __launch_bounds__(256)
smooth_pair23(const float *__restrict__ in, float *__restrict__ out, int n1, int n2, int n3, int r2, int seg,
              CgArgs A, const DipScal *sc, double *part, int first) {
    if (sc->cg_done) return;
    extern __shared__ float lds[];  // two planes of [P_T2 + 2 h2][64]
    __shared__ float w2[P_HMAX2 + 1], w3[H3 + 1];
    __shared__ double sh[16];
    constexpr int NC = P_T2 / 4, RING = 2 * H3 + 1, P_KR = (P_T2 + 2 * P_HMAX2) / 4;
    const int h2 = r2 - 1, rows = P_T2 + 2 * h2;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 64 x 4
    const int ntile1 = (n1 + 63) / 64;
    const int x0 = (blockIdx.x % ntile1) * 64, y0 = (blockIdx.x / ntile1) * P_T2;
    const int z0 = blockIdx.y * seg, z1 = min(n3, z0 + seg);
    const long n12 = (long)n1 * n2;
    const bool pend = (EPI == EPI_GP) ? (sc->pending != 0) : false;
    const float a2 = pend ? (float)sc->alpha2 : 0.0f;
    const float a1 = (EPI == EPI_S) ? (float)sc->alpha : 0.0f;
    {
        const int t = tx + 64 * ty;
        if (t <= P_HMAX2) w2[t] = (t < r2) ? (float)((double)(r2 - t) / ((double)r2 * r2)) : 0.0f;
        if (t <= H3) w3[t] = (float)((double)(H3 + 1 - t) / ((double)(H3 + 1) * (H3 + 1)));
    }
    const int ix = x0 + tx;
    const bool xin = ix < n1;
    // rows this thread loads: r = ty + 4 k (row y0 - h2 + r; rows past
    // n2 - 1 + h2 are never read by an output inside the volume)
    long roff[P_KR];
#pragma unroll
    for (int k = 0; k < P_KR; k++) {
        const int r = ty + 4 * k, yy = y0 - h2 + r;
        roff[k] = (xin && r < rows && yy < n2 + h2) ? ix + (long)n1 * fold_idx(yy, n2) : -1;
    }
    float pre[P_KR];
    auto fetch = [&](int q) {
        const long po = n12 * fold_idx(q, n3);
#pragma unroll
        for (int k = 0; k < P_KR; k++) pre[k] = roff[k] >= 0 ? in[po + roff[k]] : 0.0f;
    };
    float ring[NC][RING];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int m = 0; m < RING; m++) ring[c][m] = 0.0f;
    double acc[3] = {0.0, 0.0, 0.0};
    const int qa = z0 - H3, qb = z1 + H3;
    fetch(qa);
    for (int q = qa; q < qb; q++) {
        float *L = lds + (q & 1) * rows * 64;
#pragma unroll
        for (int k = 0; k < P_KR; k++) {
            const int r = ty + 4 * k;
            if (r < rows) L[r * 64 + tx] = pre[k];
        }
        __syncthreads();
        if (q + 1 < qb) fetch(q + 1);
        const int o = q - H3;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const float *cc = L + (ty + 4 * c + h2) * 64 + tx;
            float s = w2[0] * cc[0];
            for (int k = 1; k <= h2; k++) s += w2[k] * (cc[-k * 64] + cc[k * 64]);
#pragma unroll
            for (int m = 0; m < RING - 1; m++) ring[c][m] = ring[c][m + 1];
            ring[c][RING - 1] = s;
            const int iy = y0 + ty + 4 * c;
            if (o >= z0 && xin && iy < n2) {
                float t = w3[0] * ring[c][H3];
#pragma unroll
                for (int k = 1; k <= H3; k++) t += w3[k] * (ring[c][H3 - k] + ring[c][H3 + k]);
                sm_store<EPI>(t, out, ix + (long)n1 * iy + n12 * o, A, pend, a2, a1, first, acc);
            }
        }
    }
    sm_partials<EPI>(acc, part, gridDim.x * gridDim.y, blockIdx.x + gridDim.x * blockIdx.y, sh);
}

// ---------------------------------------------------------------------------
// Stand-alone smoothing (smoothcf, dip_cfuns.c:2006-2123): every line along
// axis `ax` smoothed `repeat` times by ps_smooth (adj: fold -> doubint ->
// triple) or ps_smooth2 (triple2 -> doubint2 -> fold2), with the box and
// derivative options (dip_cfuns.c:403-625).  One thread per line runs the
// reference's loops in its order of float (and, in triple, double)
// operations, so the result is bit-identical to the reference.  The line's
// work array is tmp[e * nl + line] (consecutive lines, consecutive words).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
smoothc_lines(float *__restrict__ x, float *__restrict__ tmp, int n1, int n2, int n3, int ax, int nb, int box,
              int der, int adj, int repeat) {
    const long n12 = (long)n1 * n2;
    const long nl = ax == 0 ? n12 / n1 * n3 : (ax == 1 ? (long)n1 * n3 : n12);
    const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nl) return;
    long o, d;
    int nx;
    if (ax == 0) { o = line * n1; d = 1; nx = n1; }
    else if (ax == 1) { o = line % n1 + line / n1 * n12; d = n1; nx = n2; }
    else { o = line; d = n12; nx = n3; }
    const int np = nx + 2 * nb;
    const float wt = box ? (float)(1.0 / (2 * nb - 1)) : (float)(1.0 / (nb * nb));
    const bool one = box || der;
    float *t = tmp + line;
    auto T = [&](int e) -> float & { return t[(long)e * nl]; };
    for (int r = 0; r < repeat; r++) {
        if (adj) {
            for (int e = 0; e < np; e++) T(e) = x[o + (long)fold_idx(e - nb, nx) * d];  // fold
            float s = 0.0f;                                                          // doubint
            for (int i = np - 1; i >= 0; i--) { s += T(i); T(i) = s; }
            if (!one) {
                s = 0.0f;
                for (int i = 0; i < np; i++) { s += T(i); T(i) = s; }
            }
            for (int i = 0; i < nx; i++) {                                           // triple
                if (box) x[o + i * d] = (T(i + 1) - T(i + 2 * nb)) * wt;
                else x[o + i * d] = (float)((2. * T(i + nb) - T(i) - T(i + 2 * nb)) * wt);
            }
        } else {
            for (int e = 0; e < np; e++) T(e) = 0.0f;                                // triple2
            if (box) {
                const float mw = -wt;
                for (int i = 0; i < nx; i++) T(i + 1) += wt * x[o + i * d];
                for (int i = 0; i < nx; i++) T(i + 2 * nb) += mw * x[o + i * d];
            } else {
                const float mw = -wt, w2 = (float)(2. * wt);
                for (int i = 0; i < nx; i++) T(i) += mw * x[o + i * d];
                for (int i = 0; i < nx; i++) T(i + nb) += w2 * x[o + i * d];
                for (int i = 0; i < nx; i++) T(i + 2 * nb) += mw * x[o + i * d];
            }
            float s = 0.0f;                                                          // doubint2
            for (int i = 0; i < np; i++) { s += T(i); T(i) = s; }
            if (!one) {
                s = 0.0f;
                for (int i = np - 1; i >= 0; i--) { s += T(i); T(i) = s; }
            }
            for (int i = 0; i < nx; i++) x[o + i * d] = T(i + nb);                   // fold2
            for (int e = nb + nx; e < np; e++) x[o + (long)fold_idx(e - nb, nx) * d] += T(e);
            for (int e = nb - 1; e >= 0; e--) x[o + (long)fold_idx(e - nb, nx) * d] += T(e);
        }
    }
}

// smoothcf with adj != 0 (ps_smooth) in two sweeps per line: the fold
// feeds the backward running sum directly (its sums are stored once in the
// work plane, [sample][line]), and the forward sweep runs the second
// running sum and forms each output from a ring of the last 2*nb integrated
// values in LDS.  Every float operation of fold / doubint / triple happens
// in the reference's order, so results are bit-identical to smoothc_lines;
// the work plane is written and read once instead of five times.
__global__ void __launch_bounds__(64)
smoothc_sweep(float *__restrict__ x, float *__restrict__ tmp, int n1, int n2, int n3, int ax, int nb, int box,
              int der, int repeat) {
    extern __shared__ float ring[];  // [2 * nb][64]
    const long n12 = (long)n1 * n2;
    const long nl = ax == 0 ? n12 / n1 * n3 : (ax == 1 ? (long)n1 * n3 : n12);
    const long line = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nl) return;
    long o, d;
    int nx;
    if (ax == 0) { o = line * n1; d = 1; nx = n1; }
    else if (ax == 1) { o = line % n1 + line / n1 * n12; d = n1; nx = n2; }
    else { o = line; d = n12; nx = n3; }
    const int np = nx + 2 * nb, R = 2 * nb;
    const float wt = box ? (float)(1.0 / (2 * nb - 1)) : (float)(1.0 / (nb * nb));
    const bool one = box || der;
    float *t = tmp + line;
    float *rg = ring + threadIdx.x;
    for (int r = 0; r < repeat; r++) {
        float s = 0.0f;
        for (int e = np - 1; e >= 0; e--) {
            s += x[o + (long)fold_idx(e - nb, nx) * d];
            t[(long)e * nl] = s;
        }
        float s2 = 0.0f;
        int slot = 0;  // e % R
        for (int e = 0; e < np; e++) {
            float v = t[(long)e * nl];
            if (!one) {
                s2 += v;
                v = s2;
            }
            if (e >= R) {  // output i = e - 2 nb; the ring holds T(i) .. T(i + 2 nb - 1)
                const long oi = o + (long)(e - R) * d;
                if (box) {
                    const int s1 = slot + 1 == R ? 0 : slot + 1;
                    x[oi] = (rg[s1 * 64] - v) * wt;
                } else {
                    const int sm = slot + nb >= R ? slot + nb - R : slot + nb;
                    x[oi] = (float)((2. * rg[sm * 64] - rg[slot * 64] - v) * wt);
                }
            }
            rg[slot * 64] = v;
            slot = slot + 1 == R ? 0 : slot + 1;
        }
    }
}

// mask32 (dip_cfuns.c:914-966, both = false)
__global__ void mask_kernel(const float *__restrict__ md, unsigned char *__restrict__ m0,
                            unsigned char *__restrict__ m1, int n1, int n2, int n3, int nw) {
    long n = (long)n1 * n2 * n3;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int i1 = (int)(i % n1);
    long r = i / n1;
    int i2 = (int)(r % n2), i3 = (int)(r / n2);
    unsigned char a = 0, b = 0;
    if (i1 >= nw && i1 < n1 - nw) {
        for (int iw = 0; iw <= 2 * nw; iw++) {
            int is = iw - nw;
            if (i2 < n2 - 1 && (md[i - is] == 0.0f || md[i + n1 + is] == 0.0f)) a = 1;
            if (i3 < n3 - 1 && (md[i - is] == 0.0f || md[i + (long)n1 * n2 + is] == 0.0f)) b = 1;
        }
    }
    m0[i] = a;
    m1[i] = b;
}

__global__ void fill_kernel(float *__restrict__ y, float v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
// Grid cap of the streaming dip kernels: each block walks a contiguous
// range of work, so a launch that the device flags turn into a no-op (CG
// converged, line search accepted) costs a few microseconds instead of a
// full-size dispatch, and the partial sums stay few.
constexpr long DIP_GRID_CAP = 4096;
inline int pw_blocks(long n) { return (int)std::min<long>(ceil_div(n, 256), DIP_GRID_CAP); }

struct DipWork {
    float *p[2], *u1, *u2, *w, *xa, *xb, *ra, *rb, *sp, *sx, *sr, *gp, *pc, *t1, *t2, *dp;
    unsigned char *m0, *m1;
    double *part;
    DipScal *sc;
    int nblk_max;
};

struct SmoothPlan {
    int naxes;
    int axis[3];
    int nb[3];
    int fused;  // all axes in one pass (smooth_fused)
    int r[3];   // radius per axis (1 = inactive)
};

struct DipGeom {
    int n1, n2, n3;
    long n;
};

// Algorithmic bytes per sample of one smoothing pass: each array the pass
// reads or writes counted once (halo re-reads excluded).  After the first
// CG iteration the model update is pending (x, r, p advanced on the fly).
double pass_bytes(int P, int E, bool first, bool single) {
    int rd = 0, wr = 0;
    if (P == PRO_CG) {  // x, r, w (+ sx, sr; writes x', r')
        rd += first ? 3 : 5;
        wr += first ? 0 : 2;
    } else {
        rd += 1;  // the field being smoothed
    }
    if (E == EPI_PLAIN) {
        wr += 1;
    } else if (E == EPI_GP) {  // p (+ sp); writes gp (+ p)
        rd += first ? 1 : 2;
        wr += first ? 1 : 2;
    } else {  // w, gp (unless gp is the input), sp, sx, sr; writes sp, sx, sr
        rd += 1 + (single ? 0 : 1) + (first ? 0 : 3);
        wr += 3;
    }
    return 4.0 * (rd + wr);
}

// One triangle-shaping application in -> out with the given prologue and
// epilogue on the first / last active axis; returns number of blocks of
// the final pass (for its partial sums).
int fused_segments(const DipGeom &g, int h3) {
    long tiles = (long)ceil_div(g.n1, F_TI1) * ceil_div(g.n2, F_TI2);
    int nseg = (int)std::max<long>(1, (2048 + tiles - 1) / tiles);
    int minseg = std::max(8, 4 * h3);  // keep the halo-plane overhead modest
    nseg = std::min(nseg, std::max(1, g.n3 / minseg));
    return ceil_div(g.n3, ceil_div(g.n3, nseg));  // segment length
}

int smooth_apply(sh_ctx *ctx, const DipGeom &g, const SmoothPlan &sp, const float *in, float *out, int pro,
                 int epi, const CgArgs &A, DipWork &W, int first, int *nblk_out) {
    // the fused 2.5-D pass over axes with radii (r1, r[1], r[2])
    auto fused = [&](int P, int E, int r1, const float *src, float *dst) -> int {
        int seg = fused_segments(g, sp.r[2] - 1);
        dim3 grid((unsigned)((long)ceil_div(g.n1, F_TI1) * ceil_div(g.n2, F_TI2)), ceil_div(g.n3, seg));
        int nblk = grid.x * grid.y;
        if (nblk > W.nblk_max) return sh_fail(ctx, SH_EINVAL, "internal: partial buffer too small");
        int h1 = r1 - 1, h2 = sp.r[1] - 1, h3 = sp.r[2] - 1;
        size_t lds = sizeof(float) * ((size_t)(F_TI2 + 2 * h2) * (F_TI1 + 2 * h1) + (size_t)(F_TI2 + 2 * h2) * F_TI1 +
                                      (size_t)(2 * h3 + 1) * F_TI2 * F_TI1);
        double bps = pass_bytes(P, E, first != 0, true);
        static const char *names[2][3] = {{"smooth_fused<0,0>", "smooth_fused<0,1>", "smooth_fused<0,2>"},
                                          {"smooth_fused<1,0>", "smooth_fused<1,1>", "smooth_fused<1,2>"}};
        PROF_A(ctx);
#define SM_F(PP, EE)                                                                                   \
    if (P == PP && E == EE)                                                                            \
        hipLaunchKernelGGL((smooth_fused<PP, EE>), grid, dim3(64, 4), lds, ctx->stream, src, dst, g.n1, g.n2, g.n3, \
                           r1, sp.r[1], sp.r[2], seg, A, W.sc, W.part, first);
        SM_F(0, 0) SM_F(0, 1) SM_F(0, 2) SM_F(1, 0) SM_F(1, 1) SM_F(1, 2)
#undef SM_F
        PROF_B(ctx, names[P][E], bps * (double)g.n);
        SH_TRY(sh_launch_check(ctx, "smooth_fused"));
        *nblk_out = nblk;
        return SH_OK;
    };
    if (sp.fused == 1) return fused(pro, epi, sp.r[0], in, out);
    // axes 2 and 3 in one pass (smooth_pair23) from src into dst
    auto pair = [&](int E, const float *src, float *dst) -> int {
        constexpr int P_T2 = 16;
        const long tiles = (long)ceil_div(g.n1, 64) * ceil_div(g.n2, P_T2);
        const int nseg = (int)std::max<long>(1, std::min<long>(g.n3 / 16, (2048 + tiles - 1) / tiles));
        const int seg = ceil_div(g.n3, nseg);
        dim3 grid((unsigned)tiles, ceil_div(g.n3, seg));
        const int nblk = grid.x * grid.y;
        if (nblk > W.nblk_max) return sh_fail(ctx, SH_EINVAL, "internal: partial buffer too small");
        const size_t lds = sizeof(float) * 2 * (size_t)(P_T2 + 2 * (sp.r[1] - 1)) * 64;
        static const char *names[3] = {"smooth_pair23<0>", "smooth_pair23<1>", "smooth_pair23<2>"};
        PROF_A(ctx);
#define SM_P(EE, HH)                                                                                       \
    if (E == EE && sp.r[2] - 1 == HH)                                                                      \
        hipLaunchKernelGGL((smooth_pair23<EE, HH, P_T2>), grid, dim3(64, 4), lds, ctx->stream, src, dst, g.n1, \
                           g.n2, g.n3, sp.r[1], seg, A, W.sc, W.part, first);
#define SM_PH(EE) SM_P(EE, 1) SM_P(EE, 2) SM_P(EE, 3) SM_P(EE, 4) SM_P(EE, 5) SM_P(EE, 6)
        SM_PH(0) SM_PH(1) SM_PH(2)
#undef SM_PH
#undef SM_P
        PROF_B(ctx, names[E], pass_bytes(PRO_PLAIN, E, first != 0, false) * (double)g.n);
        SH_TRY(sh_launch_check(ctx, "smooth_pair23"));
        *nblk_out = nblk;
        return SH_OK;
    };
    // pair (fused == 3): axis 1 as a separable pass into t1, then axes 2
    // and 3 in one pass (smooth_pair23)
    const int nsep = sp.fused == 3 ? 1 : sp.naxes;
    // plain passes of volumes of 2^24 samples or more write their outputs
    // with non-temporal stores (their arrays do not stay in the L2)
    const bool nt = g.n >= (1L << 24);
    const float *src = in;
    for (int a = 0; a < nsep; a++) {
        bool is_first = a == 0, is_last = a == sp.naxes - 1;
        int P = is_first ? pro : PRO_PLAIN;
        int E = is_last ? epi : EPI_PLAIN;
        float *dst = is_last ? out : ((a % 2 == 0) ? W.t1 : W.t2);
        int ax = sp.axis[a], nb = sp.nb[a];
        int nblk;
        // algorithmic bytes per sample of this pass (pending update assumed
        // after the first CG iteration)
        double bps = pass_bytes(P, E, first != 0, is_first && is_last);
        static const char *names[2][2][3] = {
            {{"smooth_contig<0,0>", "smooth_contig<0,1>", "smooth_contig<0,2>"},
             {"smooth_contig<1,0>", "smooth_contig<1,1>", "smooth_contig<1,2>"}},
            {{"smooth_strided<0,0>", "smooth_strided<0,1>", "smooth_strided<0,2>"},
             {"smooth_strided<1,0>", "smooth_strided<1,1>", "smooth_strided<1,2>"}}};
        PROF_A(ctx);
        if (ax == 0) {
            long ntask = (long)ceil_div(g.n1, SM_CH1) * g.n2 * g.n3;
            dim3 grid((unsigned)std::min<long>(ntask, DIP_GRID_CAP));
            nblk = grid.x;
#define SM_C(PP, EE)                                                                                        \
    if (P == PP && E == EE) {                                                                               \
        if (EE == EPI_PLAIN && nt)                                                                          \
            hipLaunchKernelGGL((smooth_contig<PP, EE, true>), grid, dim3(256), 0, ctx->stream, src, dst, g.n1, ntask, \
                               nb, A, W.sc, W.part, first);                                                 \
        else                                                                                                \
            hipLaunchKernelGGL((smooth_contig<PP, EE, false>), grid, dim3(256), 0, ctx->stream, src, dst, g.n1,      \
                               ntask, nb, A, W.sc, W.part, first);                                          \
    }
            // the plain pass: one wave per task (SH_SMOOTH_WAVE=0: one block
            // per task).  Measured at 512x256x256: warm step 48.2 -> 47.7 ms;
            // with the CG prologue (smooth_contig_wave<1>, 122 VGPRs) that
            // pass went 5.97 -> 6.94 ms per warm step, so it keeps the block
            // form.
            const char *we = getenv("SH_SMOOTH_WAVE");
            if (P == PRO_PLAIN && E == EPI_PLAIN && !(we && we[0] == '0')) {
                if (nt)
                    hipLaunchKernelGGL((smooth_contig_wave<PRO_PLAIN, true>), grid, dim3(256), 0, ctx->stream, src, dst,
                                       g.n1, ntask, nb, A, W.sc);
                else
                    hipLaunchKernelGGL((smooth_contig_wave<PRO_PLAIN, false>), grid, dim3(256), 0, ctx->stream, src, dst,
                                       g.n1, ntask, nb, A, W.sc);
            } else {
                SM_C(0, 0) SM_C(0, 1) SM_C(0, 2) SM_C(1, 0) SM_C(1, 1) SM_C(1, 2)
            }
#undef SM_C
        } else {
            int naxis = ax == 1 ? g.n2 : g.n3;
            long stride = ax == 1 ? g.n1 : (long)g.n1 * g.n2;
            int nother = ax == 1 ? g.n3 : g.n2;
            long ostride = ax == 1 ? (long)g.n1 * g.n2 : g.n1;
            // tile: tw i1-samples x chs samples along the axis (+ halos);
            // SH_SMOOTH_TILE=TWxCHS overrides (tw in {32, 64})
            int tw = SM_TW, chs = SM_CHS;
            const char *te = getenv("SH_SMOOTH_TILE");
            if (te) {
                int a = 0, b = 0;
                if (sscanf(te, "%dx%d", &a, &b) == 2 && (a == 32 || a == 64) && b >= 8 && b <= 512) {
                    tw = a;
                    chs = b;
                }
            }
            long ntask = (long)ceil_div(g.n1, tw) * ceil_div(naxis, chs) * nother;
            // small volumes (a 2-D gather): narrower, then shorter tiles
            // until the launch has a block for every SIMD (the stencils are
            // the same; only the CG sums' block partition changes)
            while (!te && ntask < 1024 && (tw > 32 || chs > 16)) {
                if (tw > 32) tw = 32;
                else chs /= 2;
                ntask = (long)ceil_div(g.n1, tw) * ceil_div(naxis, chs) * nother;
            }
            dim3 grid((unsigned)std::min<long>(ntask, DIP_GRID_CAP));
            nblk = grid.x;
            size_t lds = (size_t)(chs + 2 * (nb - 1)) * tw * sizeof(float);
            // the plain pass with registers per i1 column (SH_SMOOTH_COL=1;
            // radii 2..9).  Measured at 512x256x256: the pass itself 3.82 ->
            // 3.39 ms per warm step, but the warm step (two dip directions
            // side by side) 47.7 -> 48.1 ms, the passes after it slower; so
            // it stays opt-in.
            const char *ce = getenv("SH_SMOOTH_COL");
            const bool col = P == PRO_PLAIN && E == EPI_PLAIN && ce && ce[0] == '1' && nb >= 2 && nb <= 9;
            if (col) {
                const long ct = (long)ceil_div(g.n1, 64) * ceil_div(naxis, 32) * nother;
                dim3 cg((unsigned)std::min<long>(ceil_div(ct, 4), 8192));
#define SM_COL(HH)                                                                                            \
    if (nb - 1 == HH)                                                                                         \
        hipLaunchKernelGGL((smooth_strided_col<HH, 32>), cg, dim3(256), 0, ctx->stream, src, dst, g.n1, naxis, \
                           stride, ostride, nother, W.sc);
                SM_COL(1) SM_COL(2) SM_COL(3) SM_COL(4) SM_COL(5) SM_COL(6) SM_COL(7) SM_COL(8)
#undef SM_COL
            } else {
#define SM_S(PP, EE)                                                                                        \
    if (P == PP && E == EE) {                                                                               \
        if (EE == EPI_PLAIN && nt)                                                                          \
            hipLaunchKernelGGL((smooth_strided<PP, EE, true>), grid, dim3(tw, 256 / tw), lds, ctx->stream, src, dst, \
                               g.n1, naxis, stride, ostride, nother, nb, chs, A, W.sc, W.part, first);      \
        else                                                                                                \
            hipLaunchKernelGGL((smooth_strided<PP, EE, false>), grid, dim3(tw, 256 / tw), lds, ctx->stream, src,    \
                               dst, g.n1, naxis, stride, ostride, nother, nb, chs, A, W.sc, W.part, first); \
    }
            SM_S(0, 0) SM_S(0, 1) SM_S(0, 2) SM_S(1, 0) SM_S(1, 1) SM_S(1, 2)
#undef SM_S
            }
        }
        PROF_B(ctx, names[ax != 0][P][E], bps * (double)g.n);
        SH_TRY(sh_launch_check(ctx, "smooth"));
        if (nblk > W.nblk_max) return sh_fail(ctx, SH_EINVAL, "internal: partial buffer too small");
        if (is_last) *nblk_out = nblk;
        src = dst;
    }
    if (sp.fused == 3) return pair(epi, src, out);
    return SH_OK;
}

int finalize(sh_ctx *ctx, DipWork &W, int nblk, int K, int mode, int it, long n, int k_ls = 0) {
    hipLaunchKernelGGL(dip_finalize, dim3(1), dim3(1024), 0, ctx->stream, W.part, nblk, K, W.sc, mode, it, n,
                       1.e-6f, k_ls);
    return sh_launch_check(ctx, "dip_finalize");
}

template <int NW>
int pwd_launch(sh_ctx *ctx, int mode, const DipGeom &g, int axis, const float *x, const float *p,
               const float *dp, float lam, float *pout, DipWork &W, const unsigned char *mask) {
    int nblk = pw_blocks(g.n);
    const float pmax = 3.402823466e+38f, pmin = -3.402823466e+38f;
    static const char *names[3] = {"pwd_kernel<apply>", "pwd_kernel<der_divne>", "pwd_kernel<trial>"};
    static const double bps[3] = {12.0, 20.0, 20.0};
    const FastDiv fd1((unsigned)g.n1), fd2((unsigned)g.n2);
    PROF_A(ctx);
    if (mode == 0)
        hipLaunchKernelGGL((pwd_kernel<NW, 0>), dim3(nblk), dim3(256), 0, ctx->stream, x, p, dp, lam, pmin, pmax,
                           pout, W.u1, W.u2, mask, g.n1, g.n2, g.n3, axis, W.part, W.sc, fd1, fd2);
    else if (mode == 1)
        hipLaunchKernelGGL((pwd_kernel<NW, 1>), dim3(nblk), dim3(256), 0, ctx->stream, x, p, dp, lam, pmin, pmax,
                           pout, W.u1, W.u2, mask, g.n1, g.n2, g.n3, axis, W.part, W.sc, fd1, fd2);
    else
        hipLaunchKernelGGL((pwd_kernel<NW, 2>), dim3(nblk), dim3(256), 0, ctx->stream, x, p, dp, lam, pmin, pmax,
                           pout, W.u1, W.u2, mask, g.n1, g.n2, g.n3, axis, W.part, W.sc, fd1, fd2);
    PROF_B(ctx, names[mode], bps[mode] * (double)g.n);
    return sh_launch_check(ctx, "pwd_kernel");
}

// One dip direction: dip3 (dip_cfuns.c:1619-1691).  p[0] holds the initial
// dip on entry; the result is left in W.p[*final_idx].
template <int NW>
int dip_direction(sh_ctx *ctx, const DipGeom &g, const SmoothPlan &plan, int axis, const float *x,
                  const unsigned char *mask, int niter, int liter, DipWork &W, int *final_idx) {
    long n = g.n;
    int nb_pw = pw_blocks(n);
    int cur = 0;  // index of the current dip buffer
    CgArgs A{W.w, W.xa, W.xb, W.ra, W.rb, W.sx, W.sr, W.sp, W.pc, W.gp};
    SH_TRY(pwd_launch<NW>(ctx, 0, g, axis, x, W.p[cur], nullptr, 0.0f, nullptr, W, mask));
    SH_TRY(finalize(ctx, W, nb_pw, 1, F_USUM, 0, n));
    for (int it = 0; it < niter; it++) {
        // u1 = C'(p) x, divne prologue
        SH_TRY(pwd_launch<NW>(ctx, 1, g, axis, x, W.p[cur], nullptr, 0.0f, nullptr, W, mask));
        SH_TRY(finalize(ctx, W, nb_pw, 1, F_DIVN, it, n));
        hipLaunchKernelGGL(cg_start_kernel, dim3(nb_pw), dim3(256), 0, ctx->stream, W.u2, W.u1, W.w, W.ra, W.xa,
                           W.pc, n, W.part, W.sc);
        SH_TRY(sh_launch_check(ctx, "cg_start"));
        SH_TRY(finalize(ctx, W, nb_pw, 1, F_R0, it, n));
        for (int k = 0; k < liter; k++) {
            int nblk = 0;
            SH_TRY(smooth_apply(ctx, g, plan, nullptr, nullptr, PRO_CG, EPI_GP, A, W, k == 0, &nblk));
            SH_TRY(finalize(ctx, W, nblk, 1, F_GN, k, n));
            SH_TRY(smooth_apply(ctx, g, plan, W.gp, nullptr, PRO_PLAIN, EPI_S, A, W, k == 0, &nblk));
            SH_TRY(finalize(ctx, W, nblk, 3, F_BETA, k, n));
        }
        hipLaunchKernelGGL(cg_result_kernel, dim3(nb_pw), dim3(256), 0, ctx->stream, W.xa, W.xb, W.sx, W.dp, n,
                           W.sc);
        SH_TRY(sh_launch_check(ctx, "cg_result"));
        float lam = 1.0f;
        for (int k = 0; k < 8; k++) {
            SH_TRY(pwd_launch<NW>(ctx, 2, g, axis, x, W.p[cur], W.dp, lam, W.p[1 - cur], W, mask));
            SH_TRY(finalize(ctx, W, nb_pw, 1, F_LS, it, n, k));
            lam *= 0.5f;
        }
        cur = 1 - cur;
    }
    *final_idx = cur;
    return SH_OK;
}

}  // namespace

extern "C" int shd_dip(sh_ctx *ctx, const float *data, const float *mask, const float *init, int n1, int n2,
                      int n3, int niter, int liter, int order, float eps_dv, int r1, int r2, int r3,
                      float *dip, sh_dip_report *rep) {
    (void)eps_dv;
    if (!ctx) return SH_ENULL;
    if (!data || !dip) return sh_fail(ctx, SH_ENULL, "sh_dip: data and dip must be non-NULL");
    if (n1 < 1 || n2 < 1 || n3 < 1) return sh_fail(ctx, SH_EINVAL, "sh_dip: bad shape %d x %d x %d", n1, n2, n3);
    if (niter < 0 || liter < 0) return sh_fail(ctx, SH_EINVAL, "sh_dip: niter and liter must be >= 0");
    int rect[3] = {r1, r2, r3};
    for (int a = 0; a < 3; a++)
        if (rect[a] > SM_HMAX + 1) return sh_fail(ctx, SH_EUNSUP, "sh_dip: rect %d exceeds %d", rect[a], SM_HMAX + 1);
    DipGeom g{n1, n2, n3, (long)n1 * n2 * n3};
    SmoothPlan plan{};
    for (int a = 0; a < 3; a++)
        if (rect[a] > 1) {
            plan.axis[plan.naxes] = a;
            plan.nb[plan.naxes] = rect[a];
            plan.naxes++;
        }
    for (int a = 0; a < 3; a++) plan.r[a] = rect[a] > 1 ? rect[a] : 1;
    {
        // The separable passes are the default: on 512x256x256 with rect
        // 5,5,5 they run a CG smoothing application in ~0.45 ms against
        // ~0.77 ms for the fused 2.5-D pass, whose CG prologue is recomputed
        // on every halo sample.  SH_SMOOTH_FUSED=1 selects the fused pass
        // where its LDS footprint fits.
        int h1 = plan.r[0] - 1, h2 = plan.r[1] - 1, h3 = plan.r[2] - 1;
        size_t lds = sizeof(float) * ((size_t)(F_TI2 + 2 * h2) * (F_TI1 + 2 * h1) + (size_t)(F_TI2 + 2 * h2) * F_TI1 +
                                      (size_t)(2 * h3 + 1) * F_TI2 * F_TI1);
        const char *env = getenv("SH_SMOOTH_FUSED");
        plan.fused = (lds <= 60 * 1024) && (env && env[0] == '1') ? 1 : 0;
        // SH_SMOOTH_FUSED=3: axis 1 separable, axes 2 and 3 in smooth_pair23
        // (all three axes active, radii r2 <= 9 and 2 <= r3 <= 7).  Measured
        // at 512x256x256 with rect 5,5,5: 8.4 + 9.2 ms per warm step for the
        // two pair passes against 6.7 + 7.8 ms for the four separable ones
        // they replace (tiles of 16 or 32 rows, 4 to 16 segments along i3:
        // no better), so it stays opt-in.
        if (env && env[0] == '3' && plan.naxes == 3 && plan.r[1] - 1 <= P_HMAX2 && plan.r[2] >= 2 && plan.r[2] <= 7)
            plan.fused = 3;
    }
    if (plan.naxes == 0) {  // no smoothing: identity through a 1-tap pass
        plan.naxes = 1;
        plan.axis[0] = 0;
        plan.nb[0] = 1;
    }
    long n = g.n;
    size_t fb = (size_t)n * sizeof(float);
    // workspace
    DipWork W{};
    float *fbuf = nullptr;
    const int NF = 17;
    SH_TRY(sh_slot(ctx, SLOT_DIP0, fb * NF, (void **)&fbuf));
    float **slots[NF] = {&W.p[0], &W.p[1], &W.u1, &W.u2, &W.w, &W.xa, &W.xb, &W.ra, &W.rb,
                         &W.sp, &W.sx, &W.sr, &W.gp, &W.pc, &W.t1, &W.t2, &W.dp};
    for (int s = 0; s < NF; s++) *slots[s] = fbuf + (size_t)s * n;
    long nblk_max = std::max<long>(ceil_div(n, 256), (long)ceil_div(n1, SM_CH1) * n2 * n3);
    nblk_max = std::max<long>(nblk_max, (long)ceil_div(n1, 32) * ceil_div(n2, SM_CHS) * n3);
    nblk_max = std::max<long>(nblk_max, (long)ceil_div(n1, 32) * ceil_div(n3, SM_CHS) * n2);
    nblk_max = std::max<long>(nblk_max, (long)ceil_div(n1, F_TI1) * ceil_div(n2, F_TI2) * n3);
    nblk_max = std::max<long>(nblk_max, DIP_GRID_CAP);  // every capped launch, whatever its tile
    W.nblk_max = (int)nblk_max;
    SH_TRY(sh_slot(ctx, SLOT_RED, (size_t)nblk_max * 3 * sizeof(double), (void **)&W.part));
    SH_TRY(sh_slot(ctx, SLOT_SCAL, sizeof(DipScal), (void **)&W.sc));
    SH_HIP(ctx, hipMemsetAsync(W.sc, 0, sizeof(DipScal), ctx->stream));
    if (mask) {
        SH_TRY(sh_slot(ctx, SLOT_DIP0 + 1, (size_t)n * 2, (void **)&W.m0));
        W.m1 = W.m0 + n;
        hipLaunchKernelGGL(mask_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, mask, W.m0, W.m1, n1, n2,
                           n3, order);
        SH_TRY(sh_launch_check(ctx, "mask_kernel"));
    }
    int ndir = (n3 == 1) ? 1 : 2;
    if (rep) {
        rep->ndir = ndir;
        rep->niter = niter;
    }
    // The two directions of a 3-D estimate are independent Gauss-Newton
    // problems: direction 1 runs on the context's second stream with its
    // own workspace, beside direction 0, and is joined back at the end
    // (SH_DIP_STREAMS=1: both on the main stream, one after the other).
    const char *se = getenv("SH_DIP_STREAMS");
    const bool two = ndir == 2 && !(se && se[0] == '1');
    DipWork W2{};
    if (two) {
        SH_TRY(sh_aux_stream(ctx));
        float *fbuf2 = nullptr;
        SH_TRY(sh_slot(ctx, SLOT_DIP0 + 2, fb * NF, (void **)&fbuf2));
        float **slots2[NF] = {&W2.p[0], &W2.p[1], &W2.u1, &W2.u2, &W2.w, &W2.xa, &W2.xb, &W2.ra, &W2.rb,
                              &W2.sp, &W2.sx, &W2.sr, &W2.gp, &W2.pc, &W2.t1, &W2.t2, &W2.dp};
        for (int q = 0; q < NF; q++) *slots2[q] = fbuf2 + (size_t)q * n;
        W2.nblk_max = W.nblk_max;
        SH_TRY(sh_slot(ctx, SLOT_DIP0 + 3, (size_t)nblk_max * 3 * sizeof(double), (void **)&W2.part));
        SH_TRY(sh_slot(ctx, SLOT_DIP0 + 4, sizeof(DipScal), (void **)&W2.sc));
        SH_HIP(ctx, hipMemsetAsync(W2.sc, 0, sizeof(DipScal), ctx->stream));
        SH_HIP(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
        SH_HIP(ctx, hipStreamWaitEvent(ctx->aux, ctx->ev_fork, 0));
    }
    // restores the main stream on every return; on an early (error) return
    // the second stream is drained so no work of this call is left behind
    struct StreamGuard {
        sh_ctx *c;
        hipStream_t main;
        bool joined = false;
        ~StreamGuard() {
            if (c->stream != main) {
                c->stream = main;
                if (!joined && c->aux) (void)hipStreamSynchronize(c->aux);
            }
        }
    } guard{ctx, ctx->stream};
    for (int d = 0; d < ndir; d++) {
        DipWork &Wd = (two && d == 1) ? W2 : W;
        if (two && d == 1) ctx->stream = ctx->aux;
        if (init)
            SH_HIP(ctx, hipMemcpyAsync(Wd.p[0], init + (size_t)d * n, fb, hipMemcpyDeviceToDevice, ctx->stream));
        else {
            hipLaunchKernelGGL(fill_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, Wd.p[0], 0.0f, n);
            SH_TRY(sh_launch_check(ctx, "fill"));
        }
        int fin = 0;
        const unsigned char *m = mask ? (d == 0 ? W.m0 : W.m1) : nullptr;
        SH_DISPATCH_NW6(ctx, order, NW, SH_TRY(dip_direction<NW>(ctx, g, plan, d == 0 ? 2 : 3, data, m, niter, liter, Wd, &fin)));
        SH_HIP(ctx, hipMemcpyAsync(dip + (size_t)d * n, Wd.p[fin], fb, hipMemcpyDeviceToDevice, ctx->stream));
        if (rep && !two) {
            DipScal hs;
            SH_HIP(ctx, hipMemcpyAsync(&hs, W.sc, sizeof hs, hipMemcpyDeviceToHost, ctx->stream));
            SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
            for (int it = 0; it < niter && it < 64; it++) {
                rep->cg_iters[d][it] = hs.rec[2 * it];
                rep->ls_step[d][it] = hs.rec[2 * it + 1];
            }
        }
    }
    if (two) {
        SH_HIP(ctx, hipEventRecord(ctx->ev_join, ctx->aux));
        ctx->stream = guard.main;
        guard.joined = true;
        SH_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
        if (rep) {
            DipScal hs[2];
            SH_HIP(ctx, hipMemcpyAsync(&hs[0], W.sc, sizeof(DipScal), hipMemcpyDeviceToHost, ctx->stream));
            SH_HIP(ctx, hipMemcpyAsync(&hs[1], W2.sc, sizeof(DipScal), hipMemcpyDeviceToHost, ctx->stream));
            SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
            for (int d = 0; d < 2; d++)
                for (int it = 0; it < niter && it < 64; it++) {
                    rep->cg_iters[d][it] = hs[d].rec[2 * it];
                    rep->ls_step[d][it] = hs[d].rec[2 * it + 1];
                }
        }
    }
    return SH_OK;
}

// Replaces dipcfun.smoothcf (pyseistr/src/dip_cfuns.c:2006-2123), the C
// entry behind smoothc (pyseistr/smooth.py:115).
extern "C" int sh_smooth(sh_ctx *ctx, const float *in, float *out, int n1, int n2, int n3, int repeat, int adj,
                         int r1, int r2, int r3, int diff1, int diff2, int diff3, int box1, int box2, int box3) {
    if (!ctx) return SH_ENULL;
    if (!in || !out) return sh_fail(ctx, SH_ENULL, "sh_smooth: NULL pointer");
    if (n1 < 1 || n2 < 1 || n3 < 1) return sh_fail(ctx, SH_EINVAL, "sh_smooth: bad shape %d x %d x %d", n1, n2, n3);
    if (repeat < 0) return sh_fail(ctx, SH_EINVAL, "sh_smooth: repeat must be >= 0");
    const long n = (long)n1 * n2 * n3;
    if (out != in) SH_HIP(ctx, hipMemcpyAsync(out, in, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    const int nn[3] = {n1, n2, n3}, rect[3] = {r1, r2, r3}, diff[3] = {diff1, diff2, diff3}, box[3] = {box1, box2, box3};
    // SH_SMOOTHC_LINES=1: the one-pass-per-step kernel for adj != 0 as well
    const bool lines_only = getenv("SH_SMOOTHC_LINES") != nullptr;
    for (int ax = 0; ax < 3; ax++) {
        if (rect[ax] <= 1 || repeat == 0) continue;
        const long nl = n / nn[ax];
        float *tmp = nullptr;
        SH_TRY(sh_slot(ctx, SLOT_DIP0 + 5, (size_t)(nn[ax] + 2L * rect[ax]) * nl * sizeof(float), (void **)&tmp));
        const size_t ring = (size_t)2 * rect[ax] * 64 * sizeof(float);
        if (adj && ring <= 32768 && !lines_only) {
            hipLaunchKernelGGL(smoothc_sweep, dim3((unsigned)ceil_div(nl, 64)), dim3(64), ring, ctx->stream, out, tmp, n1,
                               n2, n3, ax, rect[ax], box[ax] != 0, diff[ax] != 0, repeat);
            SH_TRY(sh_launch_check(ctx, "smoothc_sweep"));
            continue;
        }
        hipLaunchKernelGGL(smoothc_lines, dim3((unsigned)ceil_div(nl, 256)), dim3(256), 0, ctx->stream, out, tmp, n1, n2,
                           n3, ax, rect[ax], box[ax] != 0, diff[ax] != 0, adj != 0, repeat);
        SH_TRY(sh_launch_check(ctx, "smoothc_lines"));
    }
    return SH_OK;
}
