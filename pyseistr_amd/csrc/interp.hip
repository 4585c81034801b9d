// Structure-oriented interpolation on gfx950.
//
// soint (PWD-regularised least squares, Claerbout's cgstep solver):
//   csoint3d / allpass3_lop / ps_cgstep / ps_solver
//   (pyseistr/src/soint3d_cfuns.c:625-731, 826-1174, 2405-2508) and the 2-D
//   csoint2d / allpass21_lop (pyseistr/src/soint2d_cfuns.c:296-344, 2260-2419).
// sint (masked shaping CG with structure-oriented smoothing):
//   csint3d / pwsmooth3_lop / ps_conjgrad (soint3d_cfuns.c:2231-2300,
//   1222-1349, 2510-2615) and csint2d (soint2d_cfuns.c:2421-2538).
//
// Design: the PWD operator and its adjoint are gather kernels that form
// each filter tap from the slope where they use it (the adjoint enumerates
// the contributions to each output sample in the reference's scatter
// order, so both are bit-exact); the solver scalars are formed on the
// device from deterministic reductions (device-resident control below).
#include "common.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

void sh_mt_randn(unsigned seed, size_t n, float *out);

namespace {

constexpr int TB = 256;

struct Geo {
    int n1, n2, n3, nw, nj1, nj2, nparts;
    long n;
    int drift = 0;   // shift each stencil by NINT(p) along i1 (allpass "drift")
    int maxid = 0;   // max |NINT(p)| over both slope fields (adjoint search window)
    const int *ids = nullptr;  // NINT(p) per part and sample when drift
    const float *slope[2] = {nullptr, nullptr};  // p (inline part), q (crossline part)
    FastDiv d1, d2;  // n1, n2 (sample index -> i1, i2, i3 when n < 2^31)
};

// (i1, i2, i3) of sample k
__device__ inline void sample_pos(const Geo &g, long k, int &i1, int &i2, int &i3) {
    if (g.n < (1L << 31)) {
        const unsigned r = g.d1.div((unsigned)k);
        i1 = (int)((unsigned)k - r * (unsigned)g.n1);
        const unsigned q = g.d2.div(r);
        i2 = (int)(r - q * (unsigned)g.n2);
        i3 = (int)q;
    } else {
        i1 = (int)(k % g.n1);
        const long r = k / g.n1;
        i2 = (int)(r % g.n2);
        i3 = (int)(r / g.n2);
    }
}

// PS_NINT (soint3d_cfuns.c:11): round half away from zero, in double
__device__ inline int ps_nint(float p) { return (int)(p > 0.0 ? p + 0.5 : p - 0.5); }

// Filter taps a_iw(p) (apfilt, soint3d_cfuns.c:159-250) are formed where
// they are used, from the slope of the sample they belong to; with drift
// they are those of p - NINT(p) (allpass3_lop, soint3d_cfuns.c:646-652),
// and ids[i] = NINT(p) is kept for the shifted stencils.
__global__ void nint_kernel(const float *__restrict__ pp, long n, int *__restrict__ ids) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ids[i] = ps_nint(pp[i]);
}

// the slope whose taps a sample of a part uses
__device__ inline float tap_slope(const Geo &g, int part_id, long i) {
    const float p = g.slope[part_id][i];
    return g.drift ? p - g.ids[(long)part_id * g.n + i] : p;
}

// y[part*n + i] = sum_iw (x[i+ip+is] - x[i-is]) a_iw(i)  (allpass3_lop fwd)
// plus partial dots of y with ss and rr (for cgstep) and y.y.
// The same launch also sums ss.ss, ss.rr and rr.rr (the solver's other
// dots over the data space, over the same blocks), so part holds six
// stacked partial arrays.  Returns at once when the solve has stopped.
struct SolverState;
__device__ inline int solver_done(const SolverState *st);

// One output sample of pwd_fwd; loads are unconditional (indices clamped
// into the arrays) and the result is selected, so the samples a thread
// handles form one straight-line block whose loads issue together.
template <int NW, bool DRIFT>
__device__ inline void pwd_fwd_sample(const float *__restrict__ x, const float *__restrict__ ss,
                                      const float *__restrict__ rr, const Geo &g, const float *base, long i, long ny,
                                      float &vout, double *d) {
    const bool live = i < ny;
    const long ic = live ? i : ny - 1;
    const int part_id = ic >= g.n;  // nparts <= 2
    const long k = ic - part_id * g.n;
    int ix, iy, iz;
    sample_pos(g, k, ix, iy, iz);
    const int nj = part_id ? g.nj2 : g.nj1;
    const long ip = part_id ? (long)g.n1 * g.n2 : g.n1;
    bool valid = ix >= NW * nj && ix < g.n1 - NW * nj && (part_id ? iz < g.n3 - 1 : iy < g.n2 - 1);
    int id = 0;
    if (DRIFT) {
        id = g.ids[(long)part_id * g.n + k];
        valid = valid && ix - NW * nj - id >= 0 && ix + NW * nj - id < g.n1;
    }
    float v = base ? base[ic] : 0.0f;
    const float p = g.slope[part_id][k];
    float a[2 * NW + 1];
    apf_pass<NW>(DRIFT ? p - id : p, a);
#pragma unroll
    for (int iw = 0; iw <= 2 * NW; iw++) {
        const int is = (iw - NW) * nj;
        const float t = (x[valid ? k + ip + is : k] - x[valid ? k - is - id : k]) * a[iw];
        v = valid ? v + t : v;
    }
    const float sv = ss[ic], rv = rr[ic];
    vout = v;
    d[0] = live ? (double)v * v : 0.0;
    d[1] = live ? (double)v * sv : 0.0;
    d[2] = live ? (double)v * rv : 0.0;
    d[3] = live ? (double)sv * sv : 0.0;
    d[4] = live ? (double)sv * rv : 0.0;
    d[5] = live ? (double)rv * rv : 0.0;
}

// Each block handles V consecutive groups of TB samples ("virtual blocks",
// the partition of the partial sums, so the sums are those of V blocks of
// TB threads): all V samples of a thread are loaded before any is stored.
template <int NW, bool DRIFT, int V>
__global__ void __launch_bounds__(TB)
pwd_fwd(const float *__restrict__ x, float *__restrict__ y,
        const float *__restrict__ ss, const float *__restrict__ rr, Geo g, double *part, int nblk,
        const float *base, const SolverState *st) {
    __shared__ double sh[16 * 6];
    if (st && solver_done(st)) return;
    const long ny = g.n * g.nparts;
    float v[V];
    double d[V][6];
#pragma unroll
    for (int u = 0; u < V; u++)
        pwd_fwd_sample<NW, DRIFT>(x, ss, rr, g, base, ((long)blockIdx.x * V + u) * TB + threadIdx.x, ny, v[u], d[u]);
#pragma unroll
    for (int u = 0; u < V; u++) {
        const long i = ((long)blockIdx.x * V + u) * TB + threadIdx.x;
        if (i < ny) y[i] = v[u];
    }
#pragma unroll
    for (int u = 0; u < V; u++) {
        block_sum_multi<6>(d[u], sh);
        const long blk = (long)blockIdx.x * V + u;
        if (threadIdx.x == 0 && blk < nblk)
#pragma unroll
            for (int k = 0; k < 6; k++) part[(long)k * nblk + blk] = d[u][k];
    }
}

// g[k] = (A^T rr)[k], zero where known (allpass3_lop adj + ps_solver
// 1056-1060).  Contributions are summed in the order the reference's
// scatter loop delivers them (increasing source index i, inline part
// first).  One output sample; the non-drift terms are branch-free
// (clamped indices, selected contributions).
template <int NW, bool DRIFT>
__device__ inline float pwd_adj_sample(const float *__restrict__ rr, const unsigned char *__restrict__ known,
                                       const Geo &g, long k) {
    int kx, ky, kz;
    sample_pos(g, k, kx, ky, kz);
    float v = 0.0f;
    for (int part_id = 0; part_id < g.nparts; part_id++) {
        const int nj = part_id ? g.nj2 : g.nj1;
        const int h = NW * nj;
        const long ip = part_id ? (long)g.n1 * g.n2 : g.n1;
        const float *yy = rr + (long)part_id * g.n;
        const int *idp = DRIFT ? g.ids + (long)part_id * g.n : nullptr;
        // with drift a source sample i contributes only when its shifted
        // stencil stays inside the trace
        auto live = [&](long i, int ix) {
            if (!DRIFT) return true;
            const int id = idp[i];
            return ix - h - id >= 0 && ix + h - id < g.n1;
        };
        // "+" terms: i = k - ip - is, source row valid if it has a successor
        const bool prow = part_id ? (kz >= 1) : (ky >= 1);
#pragma unroll
        for (int iw = 2 * NW; iw >= 0; iw--) {
            const int is = (iw - NW) * nj;
            const int ix = kx - is;
            bool on = prow && ix >= h && ix < g.n1 - h;
            const long i = on ? k - ip - is : k;
            if (DRIFT) on = on && live(i, ix);
            const float t = yy[i] * apf_tap<NW>(iw, tap_slope(g, part_id, i));
            v = on ? v + t : v;
        }
        const bool mrow = part_id ? (kz < g.n3 - 1) : (ky < g.n2 - 1);
        if (!DRIFT) {
#pragma unroll
            for (int iw = 0; iw <= 2 * NW; iw++) {
                const int is = (iw - NW) * nj;
                const int ix = kx + is;
                const bool on = mrow && ix >= h && ix < g.n1 - h;
                const long i = on ? k + is : k;
                const float t = yy[i] * apf_tap<NW>(iw, tap_slope(g, part_id, i));
                v = on ? v - t : v;
            }
        } else if (mrow) {
            // "-" terms with drift: sources ix of this row with
            // ix - is - NINT(p(ix)) = kx, in the scatter's order (ix,
            // then iw, ascending); |is + NINT| <= h + maxid
            const int W = h + min(g.maxid, g.n1);
            const int lo = max(h, kx - W), hi = min(g.n1 - h - 1, kx + W);
            for (int ix = lo; ix <= hi; ix++) {
                const long i = k - kx + ix;
                if (!live(i, ix)) continue;
                const int id = idp[i];
#pragma unroll
                for (int iw = 0; iw <= 2 * NW; iw++) {
                    int is = (iw - NW) * nj;
                    if (ix - is - id == kx) v -= yy[i] * apf_tap<NW>(iw, tap_slope(g, part_id, i));
                }
            }
        }
    }
    return known[k] ? 0.0f : v;
}

// g[k] = (A^T rr)[k], zero where known; V virtual blocks of TB samples per
// block as in pwd_fwd.  Partial sums: g.g.
template <int NW, bool DRIFT, int V>
__global__ void __launch_bounds__(TB)
pwd_adj(const float *__restrict__ rr, const unsigned char *__restrict__ known,
        float *__restrict__ gout, Geo g, double *part, int nblk, const SolverState *st) {
    __shared__ double sh[16];
    if (st && solver_done(st)) return;
    float v[V];
#pragma unroll
    for (int u = 0; u < V; u++) {
        const long k = ((long)blockIdx.x * V + u) * TB + threadIdx.x;
        v[u] = pwd_adj_sample<NW, DRIFT>(rr, known, g, k < g.n ? k : g.n - 1);
    }
#pragma unroll
    for (int u = 0; u < V; u++) {
        const long k = ((long)blockIdx.x * V + u) * TB + threadIdx.x;
        if (k < g.n) gout[k] = v[u];
    }
#pragma unroll
    for (int u = 0; u < V; u++) {
        const long k = ((long)blockIdx.x * V + u) * TB + threadIdx.x;
        double d[1] = {k < g.n ? (double)v[u] * v[u] : 0.0};
        block_sum_multi<1>(d, sh);
        const long blk = (long)blockIdx.x * V + u;
        if (threadIdx.x == 0 && blk < nblk) part[blk] = d[0];
    }
}

// partial sums over a vector: sum a*b (stacked into part[k*nblk + b])
__global__ void __launch_bounds__(TB) dot_kernel(const float *__restrict__ a, const float *__restrict__ b, long n,
                                                 double *part, int slot, int nblk) {
    __shared__ double sh[16];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d = (i < n) ? (double)a[i] * b[i] : 0.0;
    double t = block_sum(d, sh);
    if (threadIdx.x == 0) part[(long)slot * nblk + blockIdx.x] = t;
}

// cgstep update (soint3d_cfuns.c:864-877): S = b S + a g, x += S (nx);
// Ss = b Ss + a gg, rr += Ss (ny).  Step sizes and flags come from the
// solver state: nothing once it has stopped; S and Ss cleared on its first
// step (the restart of ps_cgstep), also when no step follows.
__global__ void cgstep_update(float *__restrict__ S, const float *__restrict__ gv, float *__restrict__ x, long nx,
                              float *__restrict__ Ss, const float *__restrict__ gg, float *__restrict__ rr, long ny,
                              const SolverState *st);

__global__ void known_kernel(const float *__restrict__ src, unsigned char *__restrict__ known, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) known[i] = src[i] != 0.0f;
}

__global__ void neg_kernel(const float *__restrict__ a, float *__restrict__ b, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = -a[i];
}

__global__ void zero_f(float *__restrict__ a, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0f;
}

int sum_partials(sh_ctx *ctx, const double *part, int nblk, int K, double *dsum, double *host) {
    SH_TRY(reduce_partials(ctx, part, nblk, K, dsum));
    SH_HIP(ctx, hipMemcpyAsync(host, dsum, sizeof(double) * K, hipMemcpyDeviceToHost, ctx->stream));
    SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SH_OK;
}

// out = max |a[i]| over n ints (block maximum, one atomic per block; out
// must start at 0)
__global__ void __launch_bounds__(TB) maxabs_kernel(const int *__restrict__ a, long n, int *out) {
    __shared__ int sh[TB];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int v = i < n ? a[i] : 0;
    sh[threadIdx.x] = v >= 0 ? v : (v == INT_MIN ? INT_MAX : -v);
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, sh[0]);
}

__global__ void fill_i(int *__restrict__ a, int v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// Device-resident solver control.  The scalars of ps_cgstep / ps_solver and
// of the shaping CG live in device memory: a one-block finalize kernel sums
// the per-block partials in a fixed order (the order of
// reduce_partials_kernel, so every sum is the one the host used to read
// back), runs the scalar logic and publishes the step sizes and the stop
// flag; the vector kernels read them from there.  No iteration waits for
// the host, and every iteration enqueues the same launches, so one
// iteration can be captured into a HIP graph and replayed.
// ---------------------------------------------------------------------------
struct SolverState {
    double r0, g0;        // first residual / gradient norms
    double gnp;           // previous gradient norm (shaping CG)
    double a, b;          // step sizes (cast to float where used)
    double gn;            // current gradient norm (shaping CG)
    int it;               // iterations begun
    int iters;            // iterations the reference's loop counter reaches
    int started, done, step, zero, first;
};

__device__ inline int solver_done(const SolverState *st) { return *(volatile const int *)&st->done; }

__global__ void cgstep_update(float *__restrict__ S, const float *__restrict__ gv, float *__restrict__ x, long nx,
                              float *__restrict__ Ss, const float *__restrict__ gg, float *__restrict__ rr, long ny,
                              const SolverState *st) {
    if (st->done) return;
    const int zero = st->zero, step = st->step;
    const float a = (float)st->a, b = (float)st->b;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nx) {
        const float sv = zero ? 0.0f : S[i];
        if (step) {
            float s = sv * b;
            s += a * gv[i];
            S[i] = s;
            x[i] += s;
        } else if (zero) {
            S[i] = 0.0f;
        }
    }
    if (i < ny) {
        const float sv = zero ? 0.0f : Ss[i];
        if (step) {
            float s = sv * b;
            s += a * gg[i];
            Ss[i] = s;
            rr[i] += s;
        } else if (zero) {
            Ss[i] = 0.0f;
        }
    }
}

// Sum k of the stacked partials part[k*nblk ...] exactly as
// reduce_partials_kernel does (1024 threads, strided, then block_sum);
// the total is valid in thread 0.
__device__ inline double partials_total(const double *part, int nblk, int k, double *sh) {
    double s = 0.0;
#pragma unroll 4
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += part[(size_t)k * nblk + i];
    double t = block_sum(s, sh);
    __syncthreads();
    return t;
}

// K of them in one strided pass (the loads of all K in flight together);
// each sum adds its terms in the order partials_total does, so the totals
// are the same bits.
template <int K>
__device__ inline void partials_totals(const double *part, int nblk, double *v, double *sh) {
    double s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
#pragma unroll 2
    for (int i = threadIdx.x; i < nblk; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < K; k++) s[k] += part[(size_t)k * nblk + i];
#pragma unroll
    for (int k = 0; k < K; k++) {
        v[k] = block_sum(s[k], sh);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(1024) solver_init(const double *part, int nblk, SolverState *st, int stop_if_zero) {
    __shared__ double sh[16];
    const double r0 = partials_total(part, nblk, 0, sh);
    if (threadIdx.x == 0) {
        SolverState z{};
        z.r0 = r0;
        z.done = (stop_if_zero && r0 == 0.) ? 1 : 0;
        *st = z;
    }
}

// ps_cgstep / ps_solver scalars (soint3d_cfuns.c:826-887, 1030-1090):
// pa = |g|^2 partials (nbg blocks); pb = g.g, g.s, g.r, s.s, s.r, r.r
// partials of the data-space vectors (nby blocks).  fnorm != 0: the two
// norms that decide the early stop are rounded to float first
// (ps_solver_prec).
__global__ void __launch_bounds__(1024) cgstep_scalars(const double *pa, int nbg, const double *pb, int nby,
                                                       SolverState *st, int fnorm) {
    __shared__ double sh[16];
    __shared__ int done;
    if (threadIdx.x == 0) done = st->done;
    __syncthreads();
    if (done) return;
    double v[7];
    v[0] = partials_total(pa, nbg, 0, sh);
    partials_totals<6>(pb, nby, v + 1, sh);
    if (threadIdx.x != 0) return;
    const double gnorm = v[0], gdg = v[1], gds = v[2], gdr0 = v[3], sds = v[4], sdr0 = v[5], rnorm = v[6];
    const double TOL = (double)1.e-12f, EPSILON = (double)1.e-12f;
    const int it = st->it;
    st->step = 0;
    st->zero = 0;
    double dpr, dpg;
    if (fnorm) {
        if (it == 0) {
            st->r0 = (float)rnorm;
            st->g0 = (float)gnorm;
            dpr = dpg = 1.0;
        } else {
            dpr = (float)rnorm / st->r0;
            dpg = (float)gnorm / st->g0;
        }
        if (dpr < TOL || dpg < TOL) {
            st->done = 1;
            st->iters = it;
            return;
        }
    } else {
        if (it == 0) {
            st->g0 = gnorm;
            dpr = dpg = 1.0;
        } else {
            dpr = rnorm / st->r0;
            dpg = gnorm / st->g0;
        }
        if ((float)dpr < (float)TOL || (float)dpg < (float)TOL) {
            st->done = 1;
            st->iters = it;
            return;
        }
    }
    st->it = it + 1;
    st->iters = it + 1;
    const bool forget = !st->started;
    st->started = 1;
    if (forget) {
        st->zero = 1;  // S = Ss = 0, also when no step follows
        st->b = 0.0;
        if (gdg <= 0.) return;
        st->a = -gdr0 / gdg;
        st->step = 1;
        return;
    }
    if (gdg == 0. || sds == 0.) return;
    double determ = 1.0 - (gds / gdg) * (gds / sds);
    if (determ > EPSILON) determ *= gdg * sds;
    else determ = gdg * sds * EPSILON;
    const double gdr = -gdr0, sdr = -sdr0;
    st->a = (sds * gdr - gds * sdr) / determ;
    st->b = (-gds * gdr + gdg * sdr) / determ;
    st->step = 1;
}

// Enqueues `body` (one solver iteration, launched on ctx->stream) niter
// times.  By default the iteration is captured once into a HIP graph (on
// the context's second stream, which only records) and the graph is
// replayed niter times on the context's stream; SH_SOLVER_GRAPH=0 launches
// the kernels of every iteration directly.  The launch timers are off
// while capturing; the replays are timed as one entry.
template <class Body>
int run_iterations(sh_ctx *ctx, int niter, const char *name, double bytes_per_iter, Body body) {
    if (niter <= 0) return SH_OK;
    const char *ge = getenv("SH_SOLVER_GRAPH");
    if (ge && ge[0] == '0') {
        for (int it = 0; it < niter; it++) SH_TRY(body());
        return SH_OK;
    }
    SH_TRY(sh_aux_stream(ctx));
    hipStream_t main = ctx->stream;
    const int prof = ctx->prof;
    hipGraph_t graph = nullptr;
    SH_HIP(ctx, hipStreamBeginCapture(ctx->aux, hipStreamCaptureModeThreadLocal));
    ctx->stream = ctx->aux;
    ctx->prof = 0;
    const int rc = body();
    ctx->stream = main;
    ctx->prof = prof;
    const hipError_t ec = hipStreamEndCapture(ctx->aux, &graph);
    if (rc != SH_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return sh_fail(ctx, SH_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return sh_fail(ctx, SH_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    PROF_A(ctx);
    hipError_t el = hipSuccess;
    for (int it = 0; it < niter && el == hipSuccess; it++) el = hipGraphLaunch(exec, main);
    PROF_B(ctx, name, bytes_per_iter * niter);
    // the replays still queued keep their own reference to the executable
    // graph; it is released after the last one has run
    const hipError_t ed = hipStreamSynchronize(main);
    (void)hipGraphExecDestroy(exec);
    if (el != hipSuccess) return sh_fail(ctx, SH_EHIP, "hipGraphLaunch: %s", hipGetErrorString(el));
    if (ed != hipSuccess) return sh_fail(ctx, SH_EHIP, "hipStreamSynchronize: %s", hipGetErrorString(ed));
    return SH_OK;
}

int read_iters(sh_ctx *ctx, const SolverState *st, int *iters) {
    if (!iters) return SH_OK;
    SH_HIP(ctx, hipMemcpyAsync(iters, &st->iters, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SH_OK;
}

// Virtual blocks per block of the soint operator kernels (1, 2 or 4;
// SH_PWD_V overrides): identical results, more loads in flight per wave.
int pwd_vblocks() {
    const char *e = getenv("SH_PWD_V");
    const int v = e ? atoi(e) : 2;
    return (v == 1 || v == 2 || v == 4) ? v : 2;
}

template <int NW>
int pwd_fwd_launch(sh_ctx *ctx, const Geo &g, const float *x, float *y, const float *ss, const float *rr,
                   double *part, int nblk, const float *base, const SolverState *st) {
    const int V = pwd_vblocks();
    const dim3 grid(ceil_div(nblk, V));
#define PF_(D, VV)                                                                                             \
    hipLaunchKernelGGL((pwd_fwd<NW, D, VV>), grid, dim3(TB), 0, ctx->stream, x, y, ss, rr, g, part, nblk, base, st)
    if (g.drift) {
        if (V == 1) PF_(true, 1); else if (V == 2) PF_(true, 2); else PF_(true, 4);
    } else {
        if (V == 1) PF_(false, 1); else if (V == 2) PF_(false, 2); else PF_(false, 4);
    }
#undef PF_
    return sh_launch_check(ctx, "pwd_fwd");
}

template <int NW>
int pwd_adj_launch(sh_ctx *ctx, const Geo &g, const float *rr, const unsigned char *known, float *gout, double *part,
                   int nblk, const SolverState *st) {
    const int V = pwd_vblocks();
    const dim3 grid(ceil_div(nblk, V));
#define PA_(D, VV)                                                                                             \
    hipLaunchKernelGGL((pwd_adj<NW, D, VV>), grid, dim3(TB), 0, ctx->stream, rr, known, gout, g, part, nblk, st)
    if (g.drift) {
        if (V == 1) PA_(true, 1); else if (V == 2) PA_(true, 2); else PA_(true, 4);
    } else {
        if (V == 1) PA_(false, 1); else if (V == 2) PA_(false, 2); else PA_(false, 4);
    }
#undef PA_
    return sh_launch_check(ctx, "pwd_adj");
}

// ps_solver + ps_cgstep with the "known" and "x0" options
template <int NW>
int soint_solve(sh_ctx *ctx, const Geo &g0, const float *pp, const float *qq, const unsigned char *known, float *x,
                const float *dd, int niter, int *iters_out) {
    Geo g = g0;
    long n = g.n, ny = g.n * g.nparts;
    float *buf = nullptr;
    // g, S (n each), rr, gg, Ss (ny each)
    size_t nf = 2 * n + 3 * ny;
    SH_TRY(sh_slot(ctx, SLOT_INTERP0, nf * sizeof(float), (void **)&buf));
    float *gv = buf, *S = gv + n, *rr = S + n, *gg = rr + ny, *Ss = gg + ny;
    g.slope[0] = pp;
    g.slope[1] = qq;
    g.d1 = FastDiv((unsigned)g.n1);
    g.d2 = FastDiv((unsigned)g.n2);
    int nbn = ceil_div(n, TB), nby = ceil_div(ny, TB);
    double *part = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_RED, (size_t)(6 * nby + nbn) * sizeof(double), (void **)&part));
    SolverState *st = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SCAL, sizeof(SolverState), (void **)&st));
    int *ids = nullptr;
    if (g.drift) {
        // NINT(p) per part and sample, then the largest shift (one sync)
        SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 9, ((size_t)g.nparts * n + 64) * sizeof(int), (void **)&ids));
    }
    int *dmax = ids ? ids + (size_t)g.nparts * n : nullptr;
    if (g.drift) {
        hipLaunchKernelGGL(nint_kernel, dim3(nbn), dim3(TB), 0, ctx->stream, pp, n, ids);
        if (g.nparts == 2) hipLaunchKernelGGL(nint_kernel, dim3(nbn), dim3(TB), 0, ctx->stream, qq, n, ids + n);
        SH_TRY(sh_launch_check(ctx, "nint"));
        const long ni = (long)g.nparts * n;
        hipLaunchKernelGGL(fill_i, dim3(1), dim3(TB), 0, ctx->stream, dmax, 0, 1L);
        hipLaunchKernelGGL(maxabs_kernel, dim3(ceil_div(ni, TB)), dim3(TB), 0, ctx->stream, ids, ni, dmax);
        SH_TRY(sh_launch_check(ctx, "maxabs"));
        int h = 0;
        SH_HIP(ctx, hipMemcpyAsync(&h, dmax, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
        g.ids = ids;
        g.maxid = h;
    }
    hipLaunchKernelGGL(zero_f, dim3(nbn), dim3(TB), 0, ctx->stream, S, n);
    hipLaunchKernelGGL(zero_f, dim3(nby), dim3(TB), 0, ctx->stream, Ss, ny);
    // rr = -dd, then rr += A x0 tap by tap (ps_solver 1018-1026); its
    // partial rr.rr starts the solver state
    hipLaunchKernelGGL(neg_kernel, dim3(nby), dim3(TB), 0, ctx->stream, dd, rr, ny);
    SH_TRY(pwd_fwd_launch<NW>(ctx, g, x, rr, Ss, Ss, part, nby, rr, nullptr));
    hipLaunchKernelGGL(solver_init, dim3(1), dim3(1024), 0, ctx->stream, part, nby, st, 0);
    SH_TRY(sh_launch_check(ctx, "soint init"));
    double *pa = part + (size_t)6 * nby;  // |g|^2 partials of pwd_adj
    // one iteration: g = A^T r (zero where known), A g with the six dots,
    // the cgstep scalars, the update (ps_solver 1030-1100, ps_cgstep)
    // (algorithmic bytes: the operator's input, slopes and output; the
    // dots' extra vectors; the update's vectors read and written)
    auto iteration = [&]() -> int {
        {
            PROF_A(ctx);
            SH_TRY(pwd_adj_launch<NW>(ctx, g, rr, known, gv, pa, nbn, st));
            PROF_B(ctx, "pwd_adj", 4.0 * (ny + g.nparts * n + n));
        }
        {
            PROF_A(ctx);
            SH_TRY(pwd_fwd_launch<NW>(ctx, g, gv, gg, Ss, rr, part, nby, nullptr, st));
            PROF_B(ctx, "pwd_fwd", 4.0 * (n + g.nparts * n + 3 * ny));
        }
        hipLaunchKernelGGL(cgstep_scalars, dim3(1), dim3(1024), 0, ctx->stream, pa, nbn, part, nby, st, 0);
        PROF_A(ctx);
        hipLaunchKernelGGL(cgstep_update, dim3(nby), dim3(TB), 0, ctx->stream, S, gv, x, n, Ss, gg, rr, ny, st);
        PROF_B(ctx, "cgstep_update", 4.0 * (5 * n + 5 * ny));
        return sh_launch_check(ctx, "soint iteration");
    };
    // algorithmic bytes per iteration: r, g, A g (2N each side) and the
    // slopes of both applications, the model and residual updates
    SH_TRY(run_iterations(ctx, niter, "soint_iteration", 4.0 * (double)(4 * ny + 4 * n), iteration));
    return read_iters(ctx, st, iters_out);
}

// ---------------------------------------------------------------------------
// sint: masked shaping CG (ps_conjgrad with hasp0) over structure-oriented
// smoothing.  Vectors live in the inline-slice layout [i3][k][i2].
// ---------------------------------------------------------------------------
// dst[i2][k][i3] = src[i3][k][i2]  (inline slices -> crossline slices)
__global__ void permute_xi(const float *__restrict__ src, float *__restrict__ dst, int n1, int n2, int n3, int inv) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)n1 * n2 * n3;
    if (i >= n) return;
    // i indexes the destination; decompose accordingly
    if (!inv) {  // dst [i2][k][i3]
        int i3 = (int)(i % n3);
        long r = i / n3;
        int k = (int)(r % n1), i2 = (int)(r / n1);
        dst[i] = src[((long)i3 * n1 + k) * n2 + i2];
    } else {  // dst [i3][k][i2] from src [i2][k][i3]
        int i2 = (int)(i % n2);
        long r = i / n2;
        int k = (int)(r % n1), i3 = (int)(r / n1);
        dst[i] = src[((long)i2 * n1 + k) * n3 + i3];
    }
}

// The same permutation through 32 x 32 LDS tiles, one k-plane per tile
// row of blocks: src [a][k][b] -> dst [b][k][a] (A, B the extents of a and
// b); reads run along b and writes along a, both coalesced.  inv = 0 is
// a = i3, b = i2; inv = 1 is a = i2, b = i3.
__global__ void __launch_bounds__(256) permute_xi_tiled(const float *__restrict__ src, float *__restrict__ dst,
                                                         int n1, int A, int B, int tilesB) {
    __shared__ float tile[32][33];
    const int tb = blockIdx.x % tilesB;
    const int k = blockIdx.x / tilesB;
    const int b0 = tb * 32, a0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int a = a0 + j, b = b0 + tx;
        if (a < A && b < B) tile[j][tx] = src[((long)a * n1 + k) * B + b];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int b = b0 + j, a = a0 + tx;
        if (a < A && b < B) dst[((long)b * n1 + k) * A + a] = tile[tx][j];
    }
}

int permute(sh_ctx *ctx, const float *src, float *dst, int n1, int n2, int n3, int inv) {
    const char *pe = getenv("SH_PERMUTE_PLAIN");
    if (pe && pe[0] == '1') {
        const long n = (long)n1 * n2 * n3;
        hipLaunchKernelGGL(permute_xi, dim3(ceil_div(n, TB)), dim3(TB), 0, ctx->stream, src, dst, n1, n2, n3, inv);
        return sh_launch_check(ctx, "permute_xi");
    }
    const int A = inv ? n2 : n3, B = inv ? n3 : n2;
    const int tilesB = ceil_div(B, 32);
    if ((long)tilesB * n1 > 0x7fffffffL || ceil_div(A, 32) > 65535)
        return sh_fail(ctx, SH_EINVAL, "permute: %d x %d x %d too large", n1, n2, n3);
    PROF_A(ctx);
    hipLaunchKernelGGL(permute_xi_tiled, dim3((unsigned)(tilesB * n1), ceil_div(A, 32)), dim3(256), 0, ctx->stream, src,
                       dst, n1, A, B, tilesB);
    PROF_B(ctx, "permute_xi_tiled", 8.0 * (double)n1 * n2 * n3);
    return sh_launch_check(ctx, "permute_xi_tiled");
}

struct Shape3 {
    Pwsmooth2 pin, px;  // inline slices (n1 x n2, n3 of them), crossline (n1 x n3, n2 of them)
    bool three;
    float *t1, *t2;     // layout conversion temporaries
    int n1, n2, n3;
    long n;
};

// x = S p  (add = 0)  or  p += S^T x  (adjoint, add = 1); pwsmooth3_lop
int shape_apply(sh_ctx *ctx, Shape3 &Sh, int adj, float *p, float *x) {
    if (!Sh.three) return pws2_apply(ctx, Sh.pin, adj, adj ? 1 : 0, p, x);
    if (!adj) {
        SH_TRY(pws2_apply(ctx, Sh.pin, 0, 0, p, Sh.t1));  // itmp
        SH_TRY(permute(ctx, Sh.t1, Sh.t2, Sh.n1, Sh.n2, Sh.n3, 0));
        SH_TRY(pws2_apply(ctx, Sh.px, 0, 0, Sh.t2, Sh.t1));  // xtmp
        return permute(ctx, Sh.t1, x, Sh.n1, Sh.n2, Sh.n3, 1);
    }
    SH_TRY(permute(ctx, x, Sh.t1, Sh.n1, Sh.n2, Sh.n3, 0));  // xtmp
    SH_TRY(pws2_apply(ctx, Sh.px, 1, 0, Sh.t2, Sh.t1));  // itmp2 = Sx^T xtmp
    SH_TRY(permute(ctx, Sh.t2, Sh.t1, Sh.n1, Sh.n2, Sh.n3, 1));
    return pws2_apply(ctx, Sh.pin, 1, 1, p, Sh.t1);  // trace += Si^T itmp
}

// gp = eps p; gx = -eps x + mask r   (ps_conjgrad 306-316 with ps_mask_lop)
__global__ void sint_pre(const float *__restrict__ p, const float *__restrict__ x, const float *__restrict__ r,
                         const unsigned char *__restrict__ m, float *__restrict__ gp, float *__restrict__ gx, float eps,
                         long n, const SolverState *st) {
    if (solver_done(st)) return;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gp[i] = eps * p[i];
    float v = -eps * x[i];
    if (m[i]) v += r[i];
    gx[i] = v;
}

// gr = mask gx; partial gp.gp
__global__ void __launch_bounds__(TB) sint_post(const float *__restrict__ gp, const float *__restrict__ gx,
                                                const unsigned char *__restrict__ m, float *__restrict__ gr, long n,
                                                double *part, const SolverState *st) {
    __shared__ double sh[16];
    if (solver_done(st)) return;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d = 0;
    if (i < n) {
        gr[i] = m[i] ? (0.0f + gx[i]) : 0.0f;
        d = (double)gp[i] * gp[i];
    }
    double t = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ps_conjgrad's gradient norm and stopping test (soint3d_cfuns.c:1290-1305):
// gn = gp.gp; first iteration g0 = gn, later alpha = gn / gnp, stop when
// alpha or gn / g0 falls below tol.
__global__ void __launch_bounds__(1024) sint_scalars_g(const double *part, int nblk, SolverState *st, float tol) {
    __shared__ double sh[16];
    __shared__ int done;
    if (threadIdx.x == 0) done = st->done;
    __syncthreads();
    if (done) return;
    const double gn = partials_total(part, nblk, 0, sh);
    if (threadIdx.x != 0) return;
    const int it = st->it;
    st->gn = gn;
    st->first = it == 0;
    if (it == 0) {
        st->g0 = gn;
        st->a = 0.0;
    } else {
        const double alpha = gn / st->gnp;
        const double dg = gn / st->g0;
        if (alpha < (double)tol || dg < (double)tol) {
            st->done = 1;
            return;
        }
        st->a = alpha;
    }
}

// s = g (first) or s = g + alpha s; partial sr.sr, sp.sp, sx.sx
__global__ void __launch_bounds__(TB) sint_dir(const float *__restrict__ gp, const float *__restrict__ gx,
                                               const float *__restrict__ gr, float *__restrict__ sp, float *__restrict__ sx,
                                               float *__restrict__ sr, long n, double *part, int nblk,
                                               const SolverState *st) {
    __shared__ double sh[16 * 3];
    if (solver_done(st)) return;
    const int first = st->first;
    const float a = (float)st->a;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d0 = 0, d1 = 0, d2 = 0;
    if (i < n) {
        float vp, vx, vr;
        if (first) {
            vp = gp[i]; vx = gx[i]; vr = gr[i];
        } else {
            vp = gp[i] + a * sp[i];
            vx = gx[i] + a * sx[i];
            vr = gr[i] + a * sr[i];
        }
        sp[i] = vp; sx[i] = vx; sr[i] = vr;
        d0 = (double)vr * vr; d1 = (double)vp * vp; d2 = (double)vx * vx;
    }
    double d[3] = {d0, d1, d2};
    block_sum_multi<3>(d, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = d[0];
        part[nblk + blockIdx.x] = d[1];
        part[2 * nblk + blockIdx.x] = d[2];
    }
}

// beta = sr.sr + eps (sp.sp - sx.sx), step -gn / beta (ps_conjgrad
// 1318-1330)
__global__ void __launch_bounds__(1024) sint_scalars_s(const double *part, int nblk, SolverState *st, float eps) {
    __shared__ double sh[16];
    __shared__ int done;
    if (threadIdx.x == 0) done = st->done;
    __syncthreads();
    if (done) return;
    double h[3];
    partials_totals<3>(part, nblk, h, sh);
    if (threadIdx.x != 0) return;
    const double beta = h[0] + (double)eps * (h[1] - h[2]);
    st->b = -st->gn / beta;
    st->gnp = st->gn;
    st->it = st->it + 1;
    st->iters = st->it;
}

__global__ void sint_update(float *__restrict__ p, float *__restrict__ x, float *__restrict__ r,
                            const float *__restrict__ sp, const float *__restrict__ sx, const float *__restrict__ sr,
                            long n, const SolverState *st) {
    if (solver_done(st)) return;
    const float a = (float)st->b;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] += a * sp[i];
    x[i] += a * sx[i];
    r[i] += a * sr[i];
}

// r = -dat + mask x
__global__ void sint_r0(const float *__restrict__ dat, const float *__restrict__ x, const unsigned char *__restrict__ m,
                        float *__restrict__ r, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = -dat[i];
    if (m[i]) v += x[i];
    r[i] = v;
}

// The shaping CG itself (p = model, x = estimated data, dat = data),
// ps_conjgrad with hasp0 (soint3d_cfuns.c:1222-1349): every scalar stays on
// the device (SolverState), so the iterations are queued without waiting.
int sint_cg(sh_ctx *ctx, Shape3 &Sh, const unsigned char *known, float *p, float *x, const float *dat, float eps1,
            float tol, int niter, int *iters) {
    long n = Sh.n;
    int nb = ceil_div(n, TB);
    float *buf = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 2, (size_t)n * 7 * sizeof(float), (void **)&buf));
    float *r = buf, *sp = r + n, *gp = sp + n, *sx = gp + n, *gx = sx + n, *sr = gx + n, *gr = sr + n;
    double *part = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_RED, (size_t)nb * 4 * sizeof(double), (void **)&part));
    SolverState *st = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SCAL, sizeof(SolverState), (void **)&st));
    float eps = eps1 * eps1;
    SH_TRY(shape_apply(ctx, Sh, 0, p, x));  // x = S p
    hipLaunchKernelGGL(sint_r0, dim3(nb), dim3(TB), 0, ctx->stream, dat, x, known, r, n);
    hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(TB), 0, ctx->stream, r, r, n, part, 0, nb);
    // a zero initial residual ends the solve before its first iteration
    hipLaunchKernelGGL(solver_init, dim3(1), dim3(1024), 0, ctx->stream, part, nb, st, 1);
    SH_TRY(sh_launch_check(ctx, "sint init"));
    // the smoothing sprays of a finished solve return at once
    Sh.pin.stop = &st->done;
    Sh.px.stop = &st->done;
    auto iteration = [&]() -> int {
        {
            PROF_A(ctx);
            hipLaunchKernelGGL(sint_pre, dim3(nb), dim3(TB), 0, ctx->stream, p, x, r, known, gp, gx, eps, n, st);
            PROF_B(ctx, "sint_pre", 4.0 * 6 * (double)n);
        }
        SH_TRY(sh_launch_check(ctx, "sint_pre"));
        SH_TRY(shape_apply(ctx, Sh, 1, gp, gx));  // gp += S^T gx
        SH_TRY(shape_apply(ctx, Sh, 0, gp, gx));  // gx = S gp
        {
            PROF_A(ctx);
            hipLaunchKernelGGL(sint_post, dim3(nb), dim3(TB), 0, ctx->stream, gp, gx, known, gr, n, part, st);
            PROF_B(ctx, "sint_post", 4.0 * 4 * (double)n);
        }
        hipLaunchKernelGGL(sint_scalars_g, dim3(1), dim3(1024), 0, ctx->stream, part, nb, st, tol);
        {
            PROF_A(ctx);
            hipLaunchKernelGGL(sint_dir, dim3(nb), dim3(TB), 0, ctx->stream, gp, gx, gr, sp, sx, sr, n, part, nb, st);
            PROF_B(ctx, "sint_dir", 4.0 * 9 * (double)n);
        }
        hipLaunchKernelGGL(sint_scalars_s, dim3(1), dim3(1024), 0, ctx->stream, part, nb, st, eps);
        PROF_A(ctx);
        hipLaunchKernelGGL(sint_update, dim3(nb), dim3(TB), 0, ctx->stream, p, x, r, sp, sx, sr, n, st);
        PROF_B(ctx, "sint_update", 4.0 * 9 * (double)n);
        return sh_launch_check(ctx, "sint iteration");
    };
    const int rc = run_iterations(ctx, niter, "sint_iteration", 0.0, iteration);
    Sh.pin.stop = nullptr;
    Sh.px.stop = nullptr;
    SH_TRY(rc);
    return read_iters(ctx, st, iters);
}

// ---------------------------------------------------------------------------
// soint2dc with prec: the preconditioned solve of csoint2d
// (soint2d_cfuns.c:2352-2396) with predict_lop / predict2_lop (1537-1616)
// as the preconditioner and ps_solver_prec (1704-1990).  The model is
// x = P p; P runs a trace along i2 and predicts it trace to trace.  Each
// prediction's banded factor depends only on its slope trace, so all
// factors are built in parallel (prec_factor); the trace-to-trace chain is
// inherently sequential (every step needs the whole previous trace, and
// each solve is itself a recurrence along i1), so it runs in one thread
// with the cached factors (prec_chain).  Arithmetic per step is the
// spray's predictor, operation for operation.
// ---------------------------------------------------------------------------
#include "predict.hpp"

// factor of the forward-direction prediction through slope trace i2 (trace
// i2's factor, a Factor with fs = 1, at F + i2 * fac_plane_floats(NB, n1):
// the NB n1 off-diagonal entries row by row, then the n1 diagonal ones)
template <int NW>
__global__ void prec_factor(const float *__restrict__ dip, float *__restrict__ F, int n1, int n2, float e) {
    const int i2 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i2 >= n2) return;
    constexpr int NB = 2 * NW;
    Acc q[1] = {Acc{dip + (long)i2 * n1, 1}};
    bool rev[1] = {true};
    factor_sweep<NW, 1>(q, rev, n1, e, fac_out_at(F + (long)i2 * fac_plane_floats(NB, n1), NB, n1, 0, 1));
}

// predict_lop without add: y = P x (adj = 0) or x = P^T y (adj = 1); tt and
// ys are one trace each of scratch.  One thread.
template <int NW>
__global__ void prec_chain(int adj, float *__restrict__ x, float *__restrict__ y, const float *__restrict__ dip,
                           const float *__restrict__ F, float *tt, float *ys, int n1, int n2, float e,
                           const SolverState *st) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (st && solver_done(st)) return;
    constexpr int NB = 2 * NW;
    for (int i1 = 0; i1 < n1; i1++) tt[i1] = 0.0f;
    bool rev[1] = {true};
    if (adj) {
        for (int i2 = n2 - 1; i2 >= 0; i2--) {
            const long o = (long)i2 * n1;
            predict_adj_factored<NW>(Acc{tt, 1}, Acc{dip + o, 1}, true,
                                     fac_at(F + (long)i2 * fac_plane_floats(NB, n1), NB, n1, 0, 1), Sink{ys, 1}, n1, e,
                                     Sink{tt, 1});
            for (int i1 = 0; i1 < n1; i1++) tt[i1] += y[o + i1];
            for (int i1 = 0; i1 < n1; i1++) x[o + i1] = 0.0f + tt[i1];
        }
    } else {
        for (int i2 = 0; i2 < n2; i2++) {
            const long o = (long)i2 * n1;
            for (int i1 = 0; i1 < n1; i1++) tt[i1] += x[o + i1];
            for (int i1 = 0; i1 < n1; i1++) y[o + i1] = 0.0f + tt[i1];
            Acc in[1] = {Acc{tt, 1}}, q[1] = {Acc{dip + o, 1}};
            solve_factored_pf<NW, 1, 8>(in, q, rev, fac_at(F + (long)i2 * fac_plane_floats(NB, n1), NB, n1, 0, 1),
                                        Sink{ys, 1}, n1, e, Sink{tt, 1});
        }
    }
}

// prec_chain with the current trace's factor and slopes, the running trace
// and the solve scratch in LDS ((NB + 4) * n1 floats): the block stages
// them and does the element-wise steps, thread 0 runs the substitutions
// from LDS.  Same operations in the same order as prec_chain.
template <int NW>
__global__ void __launch_bounds__(256)
prec_chain_lds(int adj, float *__restrict__ x, float *__restrict__ y, const float *__restrict__ dip,
               const float *__restrict__ F, int n1, int n2, float e, const SolverState *st) {
    extern __shared__ __align__(16) float sm[];
    if (st && solver_done(st)) return;
    constexpr int NB = 2 * NW;
    float *Fs = sm, *ds = Fs + (NB + 1) * n1, *tt = ds + n1, *ys = tt + n1;
    const int t = threadIdx.x, T = blockDim.x;
    for (int i = t; i < n1; i += T) tt[i] = 0.0f;
    bool rev[1] = {true};
    for (int s = 0; s < n2; s++) {
        const int i2 = adj ? n2 - 1 - s : s;
        const long o = (long)i2 * n1;
        const float *Fi = F + (long)i2 * fac_plane_floats(NB, n1);
        for (int i = t; i < (NB + 1) * n1; i += T) Fs[i] = Fi[i];
        for (int i = t; i < n1; i += T) ds[i] = dip[o + i];
        if (!adj)
            for (int i = t; i < n1; i += T) {
                const float v = tt[i] + x[o + i];
                tt[i] = v;
                y[o + i] = 0.0f + v;
            }
        __syncthreads();
        if (t == 0) {
            if (adj) {
                predict_adj_factored<NW>(Acc{tt, 1}, Acc{ds, 1}, true, fac_at(Fs, NB, n1, 0, 1), Sink{ys, 1}, n1, e,
                                         Sink{tt, 1});
            } else {
                Acc in[1] = {Acc{tt, 1}}, q[1] = {Acc{ds, 1}};
                solve_factored<NW, 1>(in, q, rev, fac_at(Fs, NB, n1, 0, 1), Sink{ys, 1}, n1, e, Sink{tt, 1});
            }
        }
        __syncthreads();
        if (adj)
            for (int i = t; i < n1; i += T) {
                const float v = tt[i] + y[o + i];
                tt[i] = v;
                x[o + i] = 0.0f + v;
            }
    }
}

// Substitutions of one prediction with a stored factor Fs (a trace's
// Factor with fs = 1: row k's off-diagonal entries at NB k, d[k] at
// NB n1 + k): forward b[k] = src[k] - ...,
// y[k] = b[k] / d[k], then backward into out (which may be y).  The loops of
// solve_factored / predict_adj_factored, operation for operation.  (Loads
// issued ahead into register rings, or the loops unrolled by four, measured
// slower here: the operands are in LDS.)
template <int NW>
__device__ void subst_factored(const float *src, const float *Fs, float *y, int n1, float *out) {
    constexpr int NB = 2 * NW;
    float Lo[NB][NB], Bw[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) {
        Bw[r] = 0.0f;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[r][m] = 0.0f;
    }
    for (int k = 0; k < n1; k++) {
        const int m1 = k < NB ? k : NB;
        float b = src[k];
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) b -= (Lo[m][m]) * Bw[m];
        y[k] = b / Fs[NB * n1 + k];
        float ok[NB];
#pragma unroll
        for (int m = 0; m < NB; m++) ok[m] = Fs[NB * k + m];
#pragma unroll
        for (int r = NB - 1; r > 0; r--) {
            Bw[r] = Bw[r - 1];
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[r][m] = Lo[r - 1][m];
        }
        Bw[0] = b;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[0][m] = ok[m];
    }
    float nx[NB];
#pragma unroll
    for (int m = 0; m < NB; m++) nx[m] = 0.0f;
    for (int k = n1 - 1; k >= 0; k--) {
        float t = y[k];
        const int m1 = (n1 - k - 1) < NB ? (n1 - k - 1) : NB;
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) t -= Fs[NB * k + m] * nx[m];
#pragma unroll
        for (int m = NB - 1; m > 0; m--) nx[m] = nx[m - 1];
        nx[0] = t;
        out[k] = t;
    }
}

// prec_chain with the per-sample work spread over the block: the filter
// taps, the right-hand side A^T A tt + e * boundary (forward) and the
// adjoint product A^T A s + e * boundary(s) are computed for all samples in
// parallel, each sample's sums in the reference's order; only the
// substitutions run in thread 0.  LDS: factor (NB + 1) n1, taps (NB + 1) n1,
// tt, y, tmp, rhs (n1 each).
template <int NW>
__global__ void __launch_bounds__(256)
prec_chain_par(int adj, float *__restrict__ x, float *__restrict__ y, const float *__restrict__ dip,
               const float *__restrict__ F, int n1, int n2, float e, const SolverState *st) {
    extern __shared__ __align__(16) float sm[];
    if (st && solver_done(st)) return;
    constexpr int NB = 2 * NW, NA = 2 * NW + 1;
    float *Fs = sm, *tp = Fs + (NB + 1) * n1, *tt = tp + NA * n1, *ys = tt + n1, *tm = ys + n1, *rh = tm + n1;
    const int t = threadIdx.x, T = blockDim.x;
    for (int i = t; i < n1; i += T) tt[i] = 0.0f;
    for (int s = 0; s < n2; s++) {
        const int i2 = adj ? n2 - 1 - s : s;
        const long o = (long)i2 * n1;
        const float *Fi = F + (long)i2 * fac_plane_floats(NB, n1);
        for (int i = t; i < (NB + 1) * n1; i += T) Fs[i] = Fi[i];
        const Acc q{dip + o, 1};
        for (int i = t; i < n1; i += T) {
            float a[NA];
            taps_at<NW>(q, i, n1, true, a);
#pragma unroll
            for (int j = 0; j < NA; j++) tp[j * n1 + i] = a[j];
        }
        if (!adj) {
            for (int i = t; i < n1; i += T) {
                const float v = tt[i] + x[o + i];
                tt[i] = v;
                y[o + i] = 0.0f + v;
            }
            __syncthreads();
            // tmp = A tt (pwd_set forward, first half)
            for (int i = t; i < n1; i += T) {
                float v = 0.0f;
                if (i >= NW && i < n1 - NW) {
#pragma unroll
                    for (int j = 0; j < NA; j++) v += tp[j * n1 + i] * tt[i + j - NW];
                }
                tm[i] = v;
            }
            __syncthreads();
            // rhs = A^T tmp + e * boundary of tt
            for (int k = t; k < n1; k += T) {
                float v = 0.0f;
#pragma unroll
                for (int j = 0; j < NA; j++) {
                    const int kk = k + j - NW;
                    if (kk >= NW && kk < n1 - NW) v += tp[j * n1 + kk] * tm[kk];
                }
                if (k == 0) v += e * tt[0];
                if (k == 1) v += e * tt[1];
                if (k == n1 - 2) v += e * tt[n1 - 2];
                if (k == n1 - 1) v += e * tt[n1 - 1];
                rh[k] = v;
            }
            __syncthreads();
            if (t == 0) subst_factored<NW>(rh, Fs, ys, n1, tt);
            __syncthreads();
        } else {
            __syncthreads();
            if (t == 0) subst_factored<NW>(tt, Fs, ys, n1, ys);
            __syncthreads();
            // tmp = A s on the solved trace (pwd_set adjoint, gather form)
            for (int i = t; i < n1; i += T) {
                float v = 0.0f;
                if (i >= NW && i < n1 - NW) {
#pragma unroll
                    for (int j = NA - 1; j >= 0; j--) v += tp[j * n1 + i] * ys[i - j + NW];
                }
                tm[i] = v;
            }
            __syncthreads();
            for (int k = t; k < n1; k += T) {
                float v = 0.0f;
#pragma unroll
                for (int w = 0; w < NA; w++) {
                    const int i = k - NW + w;
                    if (i >= NW && i < n1 - NW) v += tp[(2 * NW - w) * n1 + i] * tm[i];
                }
                if (k == 0) v += e * ys[0];
                if (k == 1) v += e * ys[1];
                if (k == n1 - 2) v += e * ys[n1 - 2];
                if (k == n1 - 1) v += e * ys[n1 - 1];
                v += y[o + k];
                tt[k] = v;
                x[o + k] = 0.0f + v;
            }
            __syncthreads();
        }
    }
}

// xt = M^T rr (ps_mask_lop adjoint without add)
__global__ void mask_adj(const unsigned char *__restrict__ known, const float *__restrict__ rr, float *__restrict__ xt,
                         long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xt[i] = known[i] ? 0.0f + rr[i] : 0.0f;
}

__global__ void mask_adj_st(const unsigned char *__restrict__ known, const float *__restrict__ rr,
                            float *__restrict__ xt, long n, const SolverState *st) {
    if (solver_done(st)) return;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xt[i] = known[i] ? 0.0f + rr[i] : 0.0f;
}

// gg = M xt, then gg += eps * (eps * rr) with eps = 0 (the solver's
// regularisation rows, which only touch the sign of zero); partial sums of
// gg.gg, gg.Ss, gg.rr, Ss.Ss, Ss.rr, rr.rr (part[0..5]) and gv.gv
// (part[6]), each over the same blocks as a separate dot would use.
__global__ void __launch_bounds__(TB)
mask_fwd_gg_dots(const unsigned char *__restrict__ known, const float *__restrict__ xt, const float *__restrict__ rr,
                 float *__restrict__ gg, const float *__restrict__ Ss, const float *__restrict__ gv, long n,
                 double *part, int nblk, const SolverState *st) {
    __shared__ double sh[16];
    if (solver_done(st)) return;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        float v = known[i] ? 0.0f + xt[i] : 0.0f;
        const float r = rr[i], sv = Ss[i], g = gv[i];
        const float g2 = 0.0f * r;
        v = v + 0.0f * g2;
        gg[i] = v;
        d[0] = (double)v * v;
        d[1] = (double)v * sv;
        d[2] = (double)v * r;
        d[3] = (double)sv * sv;
        d[4] = (double)sv * r;
        d[5] = (double)r * r;
        d[6] = (double)g * g;
    }
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const double t = block_sum(d[k], sh);
        if (threadIdx.x == 0) part[(long)k * nblk + blockIdx.x] = t;
    }
}

// P (adj = 0: y = P x) or P^T (adj = 1: x = P^T y) for one or two slope
// planes (predict2_lop: P = P(q) P(p))
template <int NW>
int prec_apply(sh_ctx *ctx, int adj, float *x, float *y, const float *pp, const float *qq, const float *Fp,
                      const float *Fq, float *tmp, float *tt, float *ys, int n1, int n2, float e,
                      const SolverState *st) {
    constexpr int NB = 2 * NW;
    const size_t ldp = (size_t)(2 * NB + 6) * n1 * sizeof(float);
    const char *mode = getenv("SH_SOINT_PREC_CHAIN");  // "lds" or "global": the slower kernels
    if (ldp <= 65536 && !mode) {
        auto run = [&](int a, float *xx, float *yy, const float *dp, const float *Fd) {
            hipLaunchKernelGGL((prec_chain_par<NW>), dim3(1), dim3(256), ldp, ctx->stream, a, xx, yy, dp, Fd, n1, n2, e,
                               st);
        };
        if (!qq) run(adj, x, y, pp, Fp);
        else if (adj) {
            run(1, tmp, y, qq, Fq);
            run(1, x, tmp, pp, Fp);
        } else {
            run(0, x, tmp, pp, Fp);
            run(0, tmp, y, qq, Fq);
        }
        return sh_launch_check(ctx, "prec_chain_par");
    }
    const size_t lds = (size_t)(NB + 4) * n1 * sizeof(float);
    if (lds <= 65536 && !(mode && !strcmp(mode, "global"))) {
        auto run = [&](int a, float *xx, float *yy, const float *dp, const float *Fd) {
            hipLaunchKernelGGL((prec_chain_lds<NW>), dim3(1), dim3(256), lds, ctx->stream, a, xx, yy, dp, Fd, n1, n2, e,
                               st);
        };
        if (!qq) run(adj, x, y, pp, Fp);
        else if (adj) {
            run(1, tmp, y, qq, Fq);
            run(1, x, tmp, pp, Fp);
        } else {
            run(0, x, tmp, pp, Fp);
            run(0, tmp, y, qq, Fq);
        }
        return sh_launch_check(ctx, "prec_chain_lds");
    }
    if (!qq) {
        hipLaunchKernelGGL((prec_chain<NW>), dim3(1), dim3(64), 0, ctx->stream, adj, x, y, pp, Fp, tt, ys, n1, n2, e,
                           st);
        return sh_launch_check(ctx, "prec_chain");
    }
    if (adj) {
        hipLaunchKernelGGL((prec_chain<NW>), dim3(1), dim3(64), 0, ctx->stream, 1, tmp, y, qq, Fq, tt, ys, n1, n2, e,
                           st);
        hipLaunchKernelGGL((prec_chain<NW>), dim3(1), dim3(64), 0, ctx->stream, 1, x, tmp, pp, Fp, tt, ys, n1, n2, e,
                           st);
    } else {
        hipLaunchKernelGGL((prec_chain<NW>), dim3(1), dim3(64), 0, ctx->stream, 0, x, tmp, pp, Fp, tt, ys, n1, n2, e,
                           st);
        hipLaunchKernelGGL((prec_chain<NW>), dim3(1), dim3(64), 0, ctx->stream, 0, tmp, y, qq, Fq, tt, ys, n1, n2, e,
                           st);
    }
    return sh_launch_check(ctx, "prec_chain");
}

// ps_solver_prec with the masked data operator, eps = 0, no initial model
// (soint2d_cfuns.c:1704-1990 as csoint2d calls it)
template <int NW>
int soint2d_prec_solve(sh_ctx *ctx, const float *din, const unsigned char *known, const float *pp,
                              const float *qq, int n1, int n2, int niter, float *out, int *iters_out) {
    const long n = (long)n1 * n2;
    constexpr int NB = 2 * NW;
    const float e = 0.0001f;
    const int nb = ceil_div(n, TB);
    // F for pp (and qq), p, g, S, rr, gg, Ss, xt, tmp, tt, ys
    const size_t nF = (size_t)n2 * fac_plane_floats(NB, n1);
    const size_t total = nF * (qq ? 2 : 1) + 8 * (size_t)n + 2 * (size_t)n1;
    float *buf = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 10, total * sizeof(float), (void **)&buf));
    float *Fp = buf, *Fq = qq ? Fp + nF : nullptr;
    float *p = buf + nF * (qq ? 2 : 1), *gv = p + n, *S = gv + n, *rr = S + n, *gg = rr + n, *Ss = gg + n;
    float *xt = Ss + n, *tmp = xt + n, *tt = tmp + n, *ys = tt + n1;
    double *part = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_RED, (size_t)nb * 7 * sizeof(double), (void **)&part));
    SolverState *st = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SCAL, sizeof(SolverState), (void **)&st));
    hipLaunchKernelGGL((prec_factor<NW>), dim3(ceil_div(n2, 64)), dim3(64), 0, ctx->stream, pp, Fp, n1, n2, e);
    if (qq) hipLaunchKernelGGL((prec_factor<NW>), dim3(ceil_div(n2, 64)), dim3(64), 0, ctx->stream, qq, Fq, n1, n2, e);
    hipLaunchKernelGGL(zero_f, dim3(nb), dim3(TB), 0, ctx->stream, p, n);
    hipLaunchKernelGGL(zero_f, dim3(nb), dim3(TB), 0, ctx->stream, S, n);
    hipLaunchKernelGGL(zero_f, dim3(nb), dim3(TB), 0, ctx->stream, Ss, n);
    hipLaunchKernelGGL(neg_kernel, dim3(nb), dim3(TB), 0, ctx->stream, din, rr, n);
    SH_TRY(sh_launch_check(ctx, "soint2d prec setup"));
    hipLaunchKernelGGL(solver_init, dim3(1), dim3(1024), 0, ctx->stream, part, 0, st, 0);
    SH_TRY(sh_launch_check(ctx, "soint2d prec setup"));
    double *pa = part + (size_t)6 * nb;
    // one iteration of ps_solver_prec (soint2d_cfuns.c:1704-1990) with
    // ps_cgstep (493-545); the reference sums the two norms that decide the
    // early stop in float
    auto iteration = [&]() -> int {
        hipLaunchKernelGGL(mask_adj_st, dim3(nb), dim3(TB), 0, ctx->stream, known, rr, xt, n, st);
        SH_TRY(prec_apply<NW>(ctx, 1, gv, xt, pp, qq, Fp, Fq, tmp, tt, ys, n1, n2, e, st));
        SH_TRY(prec_apply<NW>(ctx, 0, gv, xt, pp, qq, Fp, Fq, tmp, tt, ys, n1, n2, e, st));
        hipLaunchKernelGGL(mask_fwd_gg_dots, dim3(nb), dim3(TB), 0, ctx->stream, known, xt, rr, gg, Ss, gv, n, part,
                           nb, st);
        hipLaunchKernelGGL(cgstep_scalars, dim3(1), dim3(1024), 0, ctx->stream, pa, nb, part, nb, st, 1);
        hipLaunchKernelGGL(cgstep_update, dim3(nb), dim3(TB), 0, ctx->stream, S, gv, p, n, Ss, gg, rr, n, st);
        return sh_launch_check(ctx, "soint2d prec iteration");
    };
    SH_TRY(run_iterations(ctx, niter, "soint2d_prec_iteration", 4.0 * 10.0 * (double)n, iteration));
    // the model maps to the output through the preconditioner once the
    // iterations stop (after the last one, or at the early stop)
    if (niter > 0) SH_TRY(prec_apply<NW>(ctx, 0, p, out, pp, qq, Fp, Fq, tmp, tt, ys, n1, n2, e, nullptr));
    return read_iters(ctx, st, iters_out);
}

__global__ void count_kernel(const unsigned char *__restrict__ m, long n, double *part) {
    __shared__ double sh[16];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double d = (i < n && m[i]) ? 1.0 : 0.0;
    double t = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

}  // namespace

extern "C" {

int shd_soint3d(sh_ctx *ctx, const float *din, const float *mask, const float *dipi, const float *dipx, int n1,
               int n2, int n3, int nw, int nj1, int nj2, int niter, int drift, int seed, int hasmask, float var,
               float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    if (!din || !dipi || !dipx || !out || (hasmask == 1 && !mask))
        return sh_fail(ctx, SH_ENULL, "soint3d: NULL pointer");
    if (n1 < 8 || n2 < 2 || n3 < 2 || nj1 < 1 || nj2 < 1) return sh_fail(ctx, SH_EINVAL, "soint3d: bad arguments");
    long n = (long)n1 * n2 * n3;
    Geo g{n1, n2, n3, nw, nj1, nj2, 2, n};
    g.drift = drift != 0;
    unsigned char *known = nullptr;
    float *dd = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 1, (size_t)n + (size_t)2 * n * sizeof(float) + 256, (void **)&dd));
    known = (unsigned char *)(dd + 2 * n);
    int nb = ceil_div(n, TB);
    hipLaunchKernelGGL(known_kernel, dim3(nb), dim3(TB), 0, ctx->stream, hasmask == 1 ? mask : din, known, n);
    // dd = sqrt(var) * N(0,1) (MT19937 + Box-Muller, host-generated; zeros for var = 0)
    float a = sqrtf(var);
    if (a != 0.0f) {
        std::vector<float> h(2 * n);
        sh_mt_randn((unsigned)seed, (size_t)2 * n, h.data());
        for (auto &v : h) v = a * v;
        SH_HIP(ctx, hipMemcpyAsync(dd, h.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, ctx->stream));
        SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        hipLaunchKernelGGL(zero_f, dim3(ceil_div(2 * n, TB)), dim3(TB), 0, ctx->stream, dd, 2 * n);
    }
    SH_HIP(ctx, hipMemcpyAsync(out, din, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->stream));
    int it = 0;
    SH_DISPATCH_NW6(ctx, nw, NW, SH_TRY(soint_solve<NW>(ctx, g, dipi, dipx, known, out, dd, niter, &it)));
    if (iters) *iters = it;
    return SH_OK;
}

int shd_soint2d(sh_ctx *ctx, const float *din, const float *mask, const float *dip, int n1, int n2, int nw, int nj1,
               int niter, int drift, int hasmask, int twoplane, float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    if (!din || !dip || !out || (hasmask == 1 && !mask)) return sh_fail(ctx, SH_ENULL, "soint2d: NULL pointer");
    if (n1 < 8 || n2 < 2 || nj1 < 1) return sh_fail(ctx, SH_EINVAL, "soint2d: bad arguments");
    long n = (long)n1 * n2;
    SH_HIP(ctx, hipMemcpyAsync(out, din, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->stream));
    if (twoplane) {  // solver call commented out in the reference (soint2d_cfuns.c:2391-2392)
        if (iters) *iters = 0;
        return SH_OK;
    }
    Geo g{n1, n2, 1, nw, nj1, nj1, 1, n};
    g.drift = drift != 0;
    float *dd = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 1, (size_t)n + (size_t)n * sizeof(float) + 256, (void **)&dd));
    unsigned char *known = (unsigned char *)(dd + n);
    int nb = ceil_div(n, TB);
    hipLaunchKernelGGL(known_kernel, dim3(nb), dim3(TB), 0, ctx->stream, hasmask == 1 ? mask : din, known, n);
    hipLaunchKernelGGL(zero_f, dim3(nb), dim3(TB), 0, ctx->stream, dd, n);
    int it = 0;
    SH_DISPATCH_NW6(ctx, nw, NW, SH_TRY(soint_solve<NW>(ctx, g, dip, dip, known, out, dd, niter, &it)));
    if (iters) *iters = it;
    return SH_OK;
}

int shd_soint2d_prec(sh_ctx *ctx, const float *din, const float *mask, const float *dip, const float *dip2, int n1,
                    int n2, int nw, int niter, int hasmask, float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    if (!din || !dip || !out || (hasmask == 1 && !mask)) return sh_fail(ctx, SH_ENULL, "soint2d: NULL pointer");
    if (n1 < 8 || n2 < 1 || niter < 0) return sh_fail(ctx, SH_EINVAL, "soint2d: bad arguments");
    const long n = (long)n1 * n2;
    SH_HIP(ctx, hipMemcpyAsync(out, din, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->stream));
    unsigned char *known = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 11, (size_t)n + 256, (void **)&known));
    hipLaunchKernelGGL(known_kernel, dim3(ceil_div(n, TB)), dim3(TB), 0, ctx->stream, hasmask == 1 ? mask : din, known,
                       n);
    SH_TRY(sh_launch_check(ctx, "known_kernel"));
    int it = 0;
    SH_DISPATCH_NW(ctx, nw, NW, SH_TRY(soint2d_prec_solve<NW>(ctx, din, known, dip, dip2, n1, n2, niter, out, &it)));
    if (iters) *iters = it;
    return SH_OK;
}

static int sint_common(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, const float *mask, int n1,
                       int n2, int n3, int niter, int ns1, int ns2, int o1, int o2, float eps, float *out, int *iters) {
    long n = (long)n1 * n2 * n3;
    int nb = ceil_div(n, TB);
    bool three = dipx != nullptr;
    // layout buffers: data, mask, p (xx), x, dipi^T, dipx^X, t1, t2
    float *b = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 3, (size_t)n * 8 * sizeof(float) + n + 256, (void **)&b));
    float *dT = b, *mT = b + n, *pT = b + 2 * n, *xT = b + 3 * n, *diT = b + 4 * n, *dxX = b + 5 * n, *t1 = b + 6 * n,
          *t2 = b + 7 * n;
    unsigned char *known = (unsigned char *)(b + 8 * n);
    // per-slice transposes [i3][n2][n1] -> [i3][n1][n2], all slices in one launch
    SH_TRY(sh_transpose_internal(ctx, din, dT, n1, n2, true, n3));
    SH_TRY(sh_transpose_internal(ctx, mask, mT, n1, n2, true, n3));
    SH_TRY(sh_transpose_internal(ctx, dipi, diT, n1, n2, true, n3));
    if (three) SH_TRY(sh_transpose_internal(ctx, dipx, t1, n1, n2, true, n3));
    if (three) SH_TRY(permute(ctx, t1, dxX, n1, n2, n3, 0));
    hipLaunchKernelGGL(known_kernel, dim3(nb), dim3(TB), 0, ctx->stream, mT, known, n);
    // lam = sqrt(#known / N) (soint3d_cfuns.c:2586-2595)
    double *part = nullptr, *dsum = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_RED, (size_t)nb * 4 * sizeof(double), (void **)&part));
    SH_TRY(sh_slot(ctx, SLOT_SCAL, 16 * sizeof(double), (void **)&dsum));
    hipLaunchKernelGGL(count_kernel, dim3(nb), dim3(TB), 0, ctx->stream, known, n, part);
    double h[2];
    SH_TRY(sum_partials(ctx, part, nb, 1, dsum, h));
    float lam = (float)h[0];
    lam = sqrtf(lam / (float)n);
    Shape3 Sh{};
    Sh.three = three;
    Sh.n1 = n1; Sh.n2 = n2; Sh.n3 = n3; Sh.n = n;
    Sh.t1 = t1; Sh.t2 = t2;
    Sh.pin = Pwsmooth2{n1, n2, n3, ns1, o1, eps * eps, diT};
    SH_TRY(pws2_setup(ctx, Sh.pin, SLOT_INTERP0 + 4));
    if (three) {
        Sh.px = Pwsmooth2{n1, n3, n2, ns2, o2, eps * eps, dxX};
        SH_TRY(pws2_setup(ctx, Sh.px, SLOT_INTERP0 + 7));
    }
    SH_HIP(ctx, hipMemcpyAsync(pT, dT, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->stream));
    SH_TRY(sint_cg(ctx, Sh, known, pT, xT, dT, lam, (float)(10 * 1.19209290E-07F), niter, iters));
    SH_TRY(sh_transpose_internal(ctx, xT, out, n1, n2, false, n3));
    return SH_OK;
}

int shd_sint3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, const float *mask, int n1, int n2,
              int n3, int niter, int ns1, int ns2, int order1, int order2, float eps, float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    if (!din || !dipi || !dipx || !mask || !out) return sh_fail(ctx, SH_ENULL, "sint3d: NULL pointer");
    if (n1 < 8 || n2 < 1 || n3 < 1 || ns1 < 0 || ns2 < 0) return sh_fail(ctx, SH_EINVAL, "sint3d: bad arguments");
    return sint_common(ctx, din, dipi, dipx, mask, n1, n2, n3, niter, ns1, ns2, order1, order2, eps, out, iters);
}

int shd_sint2d(sh_ctx *ctx, const float *din, const float *dip, const float *mask, int n1, int n2, int niter, int ns,
              int order, float eps, float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    if (!din || !dip || !mask || !out) return sh_fail(ctx, SH_ENULL, "sint2d: NULL pointer");
    if (n1 < 8 || n2 < 1 || ns < 0) return sh_fail(ctx, SH_EINVAL, "sint2d: bad arguments");
    return sint_common(ctx, din, dip, nullptr, mask, n1, n2, 1, niter, ns, 0, order, order, eps, out, iters);
}

}  // extern "C"

// MT19937 (Matsumoto-Nishimura reference generator) + polar-free Box-Muller
// pairing of ps_randn_one_bm (soint3d_cfuns.c:2303-2402).
void sh_mt_randn(unsigned seed, size_t n, float *out) {
    static const unsigned mag[2] = {0u, 0x9908b0dfu};
    unsigned mt[624];
    int mti;
    mt[0] = seed;
    for (mti = 1; mti < 624; mti++) mt[mti] = 1812433253u * (mt[mti - 1] ^ (mt[mti - 1] >> 30)) + (unsigned)mti;
    auto next = [&]() {
        if (mti >= 624) {
            int k;
            for (k = 0; k < 624 - 397; k++) {
                unsigned y = (mt[k] & 0x80000000u) | (mt[k + 1] & 0x7fffffffu);
                mt[k] = mt[k + 397] ^ (y >> 1) ^ mag[y & 1u];
            }
            for (; k < 623; k++) {
                unsigned y = (mt[k] & 0x80000000u) | (mt[k + 1] & 0x7fffffffu);
                mt[k] = mt[k + 397 - 624] ^ (y >> 1) ^ mag[y & 1u];
            }
            unsigned y = (mt[623] & 0x80000000u) | (mt[0] & 0x7fffffffu);
            mt[623] = mt[396] ^ (y >> 1) ^ mag[y & 1u];
            mti = 0;
        }
        unsigned y = mt[mti++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= (y >> 18);
        return y;
    };
    bool iset = true;
    float vset = 0.0f;
    for (size_t i = 0; i < n; i++) {
        if (iset) {
            double x1, x2;
            do { x1 = next() * (1.0 / 4294967295.0); } while (x1 == 0.0);
            x2 = next() * (1.0 / 4294967295.0);
            double z1 = std::sqrt(-2.0 * std::log(x1)), z2 = 2.0 * 3.14159265358979323846264338328 * x2;
            iset = false;
            vset = (float)(z1 * std::cos(z2));
            out[i] = (float)(z1 * std::sin(z2));
        } else {
            iset = true;
            out[i] = vset;
        }
    }
}
