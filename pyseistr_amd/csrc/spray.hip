// Plane-wave spray and structure-oriented mean / median on gfx950.
//
// Algorithm: predict_step / predict1_step / predict2_step, pwspray_lop,
// pwsmooth_lop (pyseistr/src/sof_cfuns.c:405-1135), the distance-ordered 3-D
// spray of csomean3d / csomf3d (pyseistr/src/sof3d_cfuns.c:983-1054,
// 1355-1753) and the median of sf_quantile / mf (sof_cfuns.c:1137-1252).
//
// Design:
//  * traces are re-laid out "trace-fast" ([sample][trace]) so that the lanes
//    of a wave, each walking its own trace, touch consecutive addresses;
//  * one thread runs one prediction: the banded LDL^T define, the PWD
//    normal-matrix build and the forward substitution stream through the
//    trace in one pass with register windows; the factor and forward result
//    go to a coalesced per-thread scratch that the backward substitution
//    reads in reverse.  The float operations and their order are those of
//    the reference (contraction off), so spray values are bit-exact;
//  * the 3-D spray runs by dependency level: slot (a,b) of the 9x9 grid
//    depends on its neighbours toward the centre, so all slots with equal
//    |a|+|b| are independent (9 launches for radius 4);
//  * the median is an exact order statistic by rank counting in registers.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

namespace {

constexpr int TB = 256;

// ---------------------------------------------------------------------------
// Layout transposes: F-order [trace][n1] <-> trace-fast [n1][trace], for
// gridDim.z independent slices of n1 * ntr values back to back.
// ---------------------------------------------------------------------------
__global__ void to_trace_fast(const float *__restrict__ src, float *__restrict__ dst, int n1, long ntr) {
    __shared__ float tile[32][33];
    src += (long)blockIdx.z * n1 * ntr;
    dst += (long)blockIdx.z * n1 * ntr;
    long t0 = (long)blockIdx.x * 32;
    int k0 = blockIdx.y * 32;
    int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        long tr = t0 + r;
        int k = k0 + tx;
        if (tr < ntr && k < n1) tile[r][tx] = src[tr * n1 + k];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int k = k0 + r;
        long tr = t0 + tx;
        if (tr < ntr && k < n1) dst[(long)k * ntr + tr] = tile[tx][r];
    }
}

__global__ void from_trace_fast(const float *__restrict__ src, float *__restrict__ dst, int n1, long ntr) {
    __shared__ float tile[32][33];
    src += (long)blockIdx.z * n1 * ntr;
    dst += (long)blockIdx.z * n1 * ntr;
    long t0 = (long)blockIdx.x * 32;
    int k0 = blockIdx.y * 32;
    int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += 8) {
        int k = k0 + r;
        long tr = t0 + tx;
        if (tr < ntr && k < n1) tile[r][tx] = src[(long)k * ntr + tr];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        long tr = t0 + r;
        int k = k0 + tx;
        if (tr < ntr && k < n1) dst[tr * n1 + k] = tile[tx][r];
    }
}

int transpose(sh_ctx *ctx, const float *src, float *dst, int n1, long ntr, bool to_fast, int nslice = 1) {
    if (nslice < 1 || nslice > 65535) return sh_fail(ctx, SH_EINVAL, "transpose: %d slices", nslice);
    dim3 grid((unsigned)((ntr + 31) / 32), (unsigned)((n1 + 31) / 32), (unsigned)nslice);
    if (to_fast)
        hipLaunchKernelGGL(to_trace_fast, grid, dim3(32, 8), 0, ctx->stream, src, dst, n1, ntr);
    else
        hipLaunchKernelGGL(from_trace_fast, grid, dim3(32, 8), 0, ctx->stream, src, dst, n1, ntr);
    return sh_launch_check(ctx, "transpose");
}


#include "predict.hpp"

// ---------------------------------------------------------------------------
// 3-D spray by Manhattan level.  Node table entries: slot, k2-ns2, k3-ns3,
// update, up2, up3.
// ---------------------------------------------------------------------------
struct Node3 {
    int slot, o2, o3, update, up2, up3;
};

// The nodes of one level travel as a kernel argument (no host-to-device copy).
constexpr int MAX_LEVEL_NODES = 64;
struct Level3 {
    Node3 nd[MAX_LEVEL_NODES];
};

// Factor variants per target trace j (the matrix of a prediction depends
// only on the target, the axes used and the directions toward the source):
// v = 0..3: both axes, up2 = v & 1, up3 = v >> 1; v = 4, 5: axis 2 only,
// up2 = v - 4; v = 6, 7: axis 3 only, up3 = v - 6.
__host__ __device__ inline int node_variant(int update, int up2, int up3) {
    return update == 3 ? (up2 | (up3 << 1)) : (update == 1 ? 4 + up2 : 6 + up3);
}

template <int NW>
__global__ void __launch_bounds__(TB)
factor3_kernel(const float *__restrict__ pi, const float *__restrict__ px, float *__restrict__ F, int n1, int n2,
               int n3, float e) {
    long ntr = (long)n2 * n3, n = (long)n1 * ntr;
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntr) return;
    int v = blockIdx.y;
    int j2 = (int)(j % n2), j3 = (int)(j / n2);
    int use2 = (v < 4 || v == 4 || v == 5), use3 = (v < 4 || v >= 6);
    int up2 = v < 4 ? (v & 1) : (v - 4), up3 = v < 4 ? (v >> 1) : (v - 6);
    Acc q[2];
    bool rev[2];
    int nf = 0;
    if (use2) {
        if (up2 ? j2 == 0 : j2 == n2 - 1) return;  // never used at this edge
        q[nf] = Acc{pi + (up2 ? j - 1 : j), ntr};
        rev[nf] = up2 != 0;
        nf++;
    }
    if (use3) {
        if (up3 ? j3 == 0 : j3 == n3 - 1) return;
        q[nf] = Acc{px + (up3 ? j - n2 : j), ntr};
        rev[nf] = up3 != 0;
        nf++;
    }
    constexpr int NB = 2 * NW;
    const FactorOut Fs = fac_out_at(F + (long)v * fac_plane_floats(NB, n), NB, n, j, ntr);
    // slope loads pipelined 8 samples ahead (identical factor)
    // (the factor planes of a 3-D volume are read by the later level
    // launches only: non-temporal stores)
    if (nf == 1) factor_sweep_pf<NW, 1, 8, true>(q, rev, n1, e, Fs);
    else factor_sweep_pf<NW, 2, 8, true>(q, rev, n1, e, Fs);
}

// Block -> (trace block, node) of a level launch over a 1-D grid of
// nodes x gx8 blocks (gx8 = trace blocks rounded up to a multiple of 8).
// Workgroups go to the 8 XCDs round robin (block b to XCD b % 8).  With
// xcd != 0 the nodes of one trace block are consecutive blocks of one XCD,
// so the slopes and factor traces they share (nodes of one variant read
// the same ones) are fetched into that XCD's L2 once, while the nodes walk
// the samples side by side; xcd = 0 is the plain node-major order.
__device__ inline int level_block(int xcd, int cnt, long gx8, long &jb) {
    const long b = blockIdx.x;
    if (xcd) {
        const long l = b >> 3;
        jb = (l / cnt) * 8 + (b & 7);
        return (int)(l % cnt);
    }
    jb = b % gx8;
    return (int)(b / gx8);
}

// Buffer-resource accessors of the level kernels (every plane below 4 GiB):
// the plane's base and the sample row's byte offset (k * row, the
// instruction's scalar offset) are wave-uniform, only the trace's byte
// offset is a per-lane register, so a load or store needs no per-lane
// address arithmetic.  The solve's arithmetic is the same as with pointers.
__device__ inline __amdgpu_buffer_rsrc_t plane_rsrc(const void *p, unsigned long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)(unsigned)bytes, 0x00020000);
}
// Scalar offset of sample row k: k is the same in every active lane at each
// call (the solves walk samples in lockstep); read from the first active
// lane so that the compiler keeps it in a scalar register even where it
// cannot prove it uniform (it would otherwise loop over lanes per load).
__device__ inline int row_off(long k, unsigned row) {
    return (int)__builtin_amdgcn_readfirstlane((unsigned)k * row);
}
struct BTrace {
    __amdgpu_buffer_rsrc_t r;
    unsigned lane, row;
    __device__ float operator[](long k) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)lane, row_off(k, row), 0));
    }
};
// AUX: the cache policy of the stores and loads (2: non-temporal, for
// outputs that only a later launch reads and for the substitution scratch,
// whose rows come back to the same thread after the caches have turned
// over anyway)
template <int AUX>
struct BSinkT {
    __amdgpu_buffer_rsrc_t r;
    unsigned lane, row;
    struct Ref {
        __amdgpu_buffer_rsrc_t r;
        unsigned lane, off;
        __device__ operator float() const {
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)lane, (int)off, AUX));
        }
        __device__ void operator=(float v) const {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)lane, (int)off, AUX);
        }
    };
    __device__ Ref operator[](long k) const { return Ref{r, lane, (unsigned)row_off(k, row)}; }
};
using BSink = BSinkT<0>;
using BSinkNT = BSinkT<2>;
template <int AUX>
__device__ inline BTrace sink_acc(const BSinkT<AUX> &y) { return BTrace{y.r, y.lane, y.row}; }

// factor of one trace (layout of struct Factor): d plane and the rows of
// NB off-diagonal entries side by side
struct BFactor {
    __amdgpu_buffer_rsrc_t rd, ro;
    unsigned lane_d, lane_o, row_d, row_o;
};
__device__ inline float fac_d(const BFactor &F, long k) {
    return __builtin_bit_cast(float,
                              __builtin_amdgcn_raw_buffer_load_b32(F.rd, (int)F.lane_d, row_off(k, F.row_d), 0));
}
// The row's 16- or 8-byte load is bit-cast to a float vector as a whole
// before its lanes are read: read lane by lane from the load builtin's
// integer vector, the load was compiled into one 4-byte load whose value
// every lane received.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
template <int NB>
__device__ inline void fac_row(const BFactor &F, long k, float *ok) {
    const int so = row_off(k, F.row_o);
    if constexpr (NB % 4 == 0) {
#pragma unroll
        for (int c = 0; c < NB / 4; c++) {
            const f32x4 v =
                __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(F.ro, (int)(F.lane_o + 16u * c), so, 0));
#pragma unroll
            for (int i = 0; i < 4; i++) ok[4 * c + i] = v[i];
        }
    } else {
#pragma unroll
        for (int c = 0; c < NB / 2; c++) {
            const f32x2 v =
                __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(F.ro, (int)(F.lane_o + 8u * c), so, 0));
            ok[2 * c] = v[0];
            ok[2 * c + 1] = v[1];
        }
    }
}

// One launch per (level, number of source traces NF): every thread predicts
// one slot at one target trace with the shared factor of the target.  A
// slot whose source lies outside the volume is written as zeros there
// (sof3d_cfuns.c:1478-1503 skips it; the slot stays zero).  BUF: planes,
// factor and scratch through buffer resources (scratch row = the launch's
// thread count), else through pointers (scratch row = nthr); BUF with
// PF == 0 is the pointer path.  Orders 1 and 2 are held to two waves per
// SIMD (256 registers: the two-source solve at prefetch distance 6 fits
// without spills); orders 3 and 4 need more registers than that.
template <bool BUF, int NW, int NF, int PF>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(NW <= 2 ? 2 : 1)))
spray3_level(const float *__restrict__ pi, const float *__restrict__ px, float *__restrict__ U,
             const float *__restrict__ F, const Level3 nodes, int cnt, int xcd, int n1, int n2, int n3, int np2,
             float e, float *__restrict__ scr, long nthr) {
    long ntr = (long)n2 * n3, n = (long)n1 * ntr;
    const long gx8 = (ntr + 8L * TB - 1) / (8L * TB) * 8;
    long jb;
    const int node = level_block(xcd, cnt, gx8, jb);
    long j = jb * TB + threadIdx.x;
    if (j >= ntr) return;
    Node3 nd = nodes.nd[node];
    int j2 = (int)(j % n2), j3 = (int)(j / n2);
    int i2 = j2 - nd.o2, i3 = j3 - nd.o3;
    Sink out{U + (long)nd.slot * n + j, ntr};
    if (i2 < 0 || i2 >= n2 || i3 < 0 || i3 >= n3) {
        for (int k = 0; k < n1; k++) out[k] = 0.0f;
        return;
    }
    long tid = (long)node * gx8 * TB + j;
    constexpr int NB = 2 * NW;
    if constexpr (BUF && PF > 0) {
        // NF == 2: axis 2 then axis 3; NF == 1: the node's one axis
        const unsigned row = (unsigned)ntr * 4u;
        BTrace in[2], q[2];
        bool rev[2];
        const bool ax2 = (nd.update & 1) != 0;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const bool a2 = NF == 2 ? f == 0 : ax2;
            const int up = a2 ? nd.up2 : nd.up3;
            const long d = a2 ? 1 : n2;
            const int ks = nd.slot + (up ? -1 : 1) * (a2 ? 1 : np2);
            in[f] = BTrace{plane_rsrc(U + (long)ks * n, (unsigned long)n * 4), (unsigned)(up ? j - d : j + d) * 4u,
                           row};
            q[f] = BTrace{plane_rsrc(a2 ? pi : px, (unsigned long)n * 4), (unsigned)(up ? j - d : j) * 4u, row};
            rev[f] = up != 0;
        }
        const float *Fp = F + (long)node_variant(nd.update, nd.up2, nd.up3) * fac_plane_floats(NB, n);
        const BFactor Fv{plane_rsrc(Fp + (long)NB * n, (unsigned long)n * 4),
                         plane_rsrc(Fp, (unsigned long)n * NB * 4),
                         (unsigned)j * 4u,
                         (unsigned)j * (4u * NB),
                         row,
                         row * NB};
        const unsigned lt = (unsigned)(cnt * gx8 * TB);  // threads of this launch
        // scratch written and read with non-temporal accesses (keeps the
        // shared factor and slope lines in the L2)
        const BSinkNT y{plane_rsrc(scr, (unsigned long)lt * n1 * 4), (unsigned)tid * 4u, lt * 4u};
        // the slot is read only by later levels and the median: non-temporal
        const BSinkNT o{plane_rsrc(U + (long)nd.slot * n, (unsigned long)n * 4), (unsigned)j * 4u, row};
        solve_factored_pf<NW, NF, PF>(in, q, rev, Fv, y, n1, e, o);
        return;
    }
    Sink y{scr + tid, nthr};
    Acc in[2], q[2];
    bool rev[2];
    int nf = 0;
    if (nd.update & 1) {
        long jj = nd.up2 ? j - 1 : j + 1;
        int ks = nd.up2 ? nd.slot - 1 : nd.slot + 1;
        in[nf] = Acc{U + (long)ks * n + jj, ntr};
        q[nf] = Acc{pi + (nd.up2 ? j - 1 : j), ntr};
        rev[nf] = nd.up2 != 0;
        nf++;
    }
    if (nd.update & 2) {
        long jj = nd.up3 ? j - n2 : j + n2;
        int ks = nd.up3 ? nd.slot - np2 : nd.slot + np2;
        in[nf] = Acc{U + (long)ks * n + jj, ntr};
        q[nf] = Acc{px + (nd.up3 ? j - n2 : j), ntr};
        rev[nf] = nd.up3 != 0;
        nf++;
    }
    const Factor Fv = fac_at(F + (long)node_variant(nd.update, nd.up2, nd.up3) * fac_plane_floats(NB, n), NB, n, j,
                             ntr);
    if constexpr (PF > 0)
        solve_factored_pf<NW, NF, PF>(in, q, rev, Fv, y, n1, e, out);
    else
        solve_factored<NW, NF>(in, q, rev, Fv, y, n1, e, out);
}

// Two-source slots grouped by (level, quadrant): every slot of a group
// arrives at its target through the same slopes, directions and factor, so
// one thread per target trace predicts up to GM of them together
// (solve_factored_multi).  A slot whose source lies outside the volume
// stays zero at that target.
constexpr int GM = 4;
struct Group3 {
    int cnt;
    Node3 nd[GM];
};
constexpr int MAX_GROUPS = 16;
struct GroupLevel {
    Group3 g[MAX_GROUPS];
};

template <int NW, int M>
__global__ void __launch_bounds__(TB)
spray3_group(const float *__restrict__ pi, const float *__restrict__ px, float *__restrict__ U,
             const float *__restrict__ F, const GroupLevel gl, int n1, int n2, int n3, int np2, float e,
             float *__restrict__ scr, long nthr) {
    long ntr = (long)n2 * n3, n = (long)n1 * ntr;
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntr) return;
    const Group3 &G = gl.g[blockIdx.y];
    const int j2 = (int)(j % n2), j3 = (int)(j / n2);
    const int up2 = G.nd[0].up2, up3 = G.nd[0].up3;
    // a slot whose source lies outside the volume is zero at this target
    bool any = false;
    for (int r = 0; r < G.cnt; r++) {
        const Node3 &nd = G.nd[r];
        int i2 = j2 - nd.o2, i3 = j3 - nd.o3;
        if (i2 >= 0 && i2 < n2 && i3 >= 0 && i3 < n3) {
            any = true;
        } else {
            Sink z{U + (long)nd.slot * n + j, ntr};
            for (int k = 0; k < n1; k++) z[k] = 0.0f;
        }
    }
    // (a live source implies both predecessor traces exist)
    if (!any) return;
    Acc in[2 * M];
    bool live[M];
    Sink y[M], out[M];
#pragma unroll
    for (int r = 0; r < M; r++) {
        const Node3 &nd = G.nd[r < G.cnt ? r : 0];
        int i2 = j2 - nd.o2, i3 = j3 - nd.o3;
        live[r] = r < G.cnt && i2 >= 0 && i2 < n2 && i3 >= 0 && i3 < n3;
        in[2 * r] = Acc{U + (long)(up2 ? nd.slot - 1 : nd.slot + 1) * n + (up2 ? j - 1 : j + 1), ntr};
        in[2 * r + 1] = Acc{U + (long)(up3 ? nd.slot - np2 : nd.slot + np2) * n + (up3 ? j - n2 : j + n2), ntr};
        y[r] = Sink{scr + ((long)blockIdx.y * M + r) * gridDim.x * blockDim.x + j, nthr};
        out[r] = Sink{U + (long)nd.slot * n + j, ntr};
    }
    Acc q[2] = {Acc{pi + (up2 ? j - 1 : j), ntr}, Acc{px + (up3 ? j - n2 : j), ntr}};
    bool rev[2] = {up2 != 0, up3 != 0};
    constexpr int NB = 2 * NW;
    const Factor Fv = fac_at(F + (long)node_variant(3, up2, up3) * fac_plane_floats(NB, n), NB, n, j, ntr);
    solve_factored_multi<NW, 2, M>(in, live, q, rev, Fv, y, n1, e, out);
}

// centre slot = data (trace-fast copy)
__global__ void copy_slot(const float *__restrict__ d, float *__restrict__ U, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) U[i] = d[i];
}

// Plain mean over all np slots in slot order (sof3d_cfuns.c:1525-1537).
__global__ void slot_mean(const float *__restrict__ U, float *__restrict__ out, long n, int np) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < np; k++) s = s + U[(long)k * n + i];
    out[i] = s / (float)np;
}

// Median of the nmf slots [c-m, c+m] of each sample (mf, axis 2, centre
// column; sof3d_cfuns.c:1707-1736): a pruned selection network on N
// registers (median_net.hpp).  With nmf < N the window is padded with
// (N-nmf)/2 values of -inf and as many of +inf, which leaves the median of
// the real values in the network's median register.  The order statistic
// is exact; only the sign of a zero median may differ from quickselect.
template <int N> struct MedianNet;
#include "median_net.hpp"

template <int N>
__global__ void __launch_bounds__(TB)
slot_median(const float *__restrict__ U, float *__restrict__ out, long n, int first, int nmf) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int pad = (N - nmf) / 2;
    float v[N];
#pragma unroll
    for (int k = 0; k < N; k++)
        v[k] = (k < nmf) ? U[(long)(first + k) * n + i] : (k - nmf < pad ? -INFINITY : INFINITY);
#pragma unroll
    for (int c = 0; c < MedianNet<N>::NOPS; c++) {
        const int a = MedianNet<N>::OPS[c][0], b = MedianNet<N>::OPS[c][1];
        const float x = v[a], y = v[b];
        v[a] = fminf(x, y);
        v[b] = fmaxf(x, y);
    }
    out[i] = v[MedianNet<N>::MED];
}

int launch_median(sh_ctx *ctx, const float *U, float *out, long n, int first, int nmf) {
    int g = ceil_div(n, TB);
#define MED(NN) \
    if (nmf <= NN) { PROF_A(ctx); hipLaunchKernelGGL((slot_median<NN>), dim3(g), dim3(TB), 0, ctx->stream, U, out, n, first, nmf); \
        PROF_B(ctx, "slot_median", (nmf + 1) * 4.0 * (double)n); return sh_launch_check(ctx, "slot_median"); }
    MED(9) MED(13) MED(17) MED(33) MED(49) MED(65) MED(81)
#undef MED
    return sh_fail(ctx, SH_EUNSUP, "median window %d exceeds 81", nmf);
}

// ---------------------------------------------------------------------------
// 2-D spray chains.  Slice layout trace-fast per slice: [i3][k][i2].
// Thread = (source i2, direction, slice i3).  Forward: the chain walks
// predict_step(forw) to the ns neighbours (pwspray_lop, sof_cfuns.c:998-1028).
// U2 layout: [slot][i3][k][i2].  The system matrix of a step depends only on
// the target trace and the direction, so both factors of every trace are
// built once (factor2_kernel) and shared by the ns chains through it:
// v = 0: forw = false with the target's slope; v = 1: forw = true with the
// slope of the trace before it.  F2 layout [v][NB+1][i3][k][i2].
// ---------------------------------------------------------------------------
template <int NW>
__global__ void __launch_bounds__(TB)
factor2_kernel(const float *__restrict__ dip, float *__restrict__ F, int n1, int n2, int n3, float e) {
    long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i3 = blockIdx.y;
    if (gid >= 2L * n2) return;
    int v = (int)(gid / n2), i = (int)(gid % n2);
    if (v == 0 ? i > n2 - 2 : i < 1) return;  // never a target in that direction
    long slice = (long)n1 * n2, n = slice * n3;
    constexpr int NB = 2 * NW;
    Acc q[1] = {Acc{dip + i3 * slice + (v == 0 ? i : i - 1), n2}};
    bool rev[1] = {v == 1};
    const FactorOut Fs = fac_out_at(F + (long)v * fac_plane_floats(NB, n), NB, n, i3 * slice + i, n2);
    factor_sweep_pf<NW, 1, 8>(q, rev, n1, e, Fs);
}

// Accessors of the 2-D chain kernels over whole buffers (BUF: buffer
// resources with the trace's byte offset per lane and the sample row k * n2
// as the scalar offset; the direction a lane runs differs within a wave,
// so it only enters the lane offset), or plain pointers.
// (buffers of nel floats; element offset off, sample stride in elements)
template <bool BUF>
struct Chain2 {
    int n2;
    __device__ auto rd(const float *b, long nel, long off, long stride) const {
        if constexpr (BUF) return BTrace{plane_rsrc(b, nel * 4), (unsigned)off * 4u, (unsigned)stride * 4u};
        else return Acc{b + off, stride};
    }
    __device__ auto wr(float *b, long nel, long off, long stride) const {
        if constexpr (BUF) return BSink{plane_rsrc(b, nel * 4), (unsigned)off * 4u, (unsigned)stride * 4u};
        else return Sink{b + off, stride};
    }
    // factor (or taps) plane v of F (planes of pf floats, n per sample
    // plane) at element e0 of a sample plane
    template <int NB>
    __device__ auto fac(const float *F, long pf, int nplanes, long n, int v, long e0) const {
        if constexpr (BUF) {
            const __amdgpu_buffer_rsrc_t r = plane_rsrc(F, pf * nplanes * 4);
            return BFactor{r, r, (unsigned)(v * pf + NB * n + e0) * 4u, (unsigned)(v * pf + NB * e0) * 4u,
                           (unsigned)n2 * 4u, (unsigned)n2 * (4u * NB)};
        } else {
            return fac_at(F + (long)v * pf, NB, n, e0, n2);
        }
    }
};

template <bool BUF, int NW, int PFD>
__global__ void __launch_bounds__(TB)
spray2_fwd(const float *__restrict__ x, const float *__restrict__ dip, const float *__restrict__ F,
           float *__restrict__ U, int n1, int n2, int n3, int ns, float e, float *__restrict__ scr, long nthr,
           const int *stop, const float *__restrict__ x2, float *__restrict__ U2) {
    if (stop && *(volatile const int *)stop) return;
    if (blockIdx.z) {  // the second spray of the launch: its own input, output and scratch plane
        x = x2;
        U = U2;
        scr += nthr * (long)n1;
    }
    long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;  // over n2 * 2
    int i3 = blockIdx.y;
    if (gid >= 2L * n2) return;
    int dir = (int)(gid / n2);  // 0: toward lower i2 (forw=false), 1: toward higher (forw=true)
    int i = (int)(gid % n2);
    long slice = (long)n1 * n2;
    const float *xs = x + i3 * slice;
    long plane_slot = slice * n3, n = plane_slot;
    constexpr int NB = 2 * NW;
    const long pf = fac_plane_floats(NB, n);
    const Chain2<BUF> C2{n2};
    long tid = (long)i3 * 2 * n2 + gid;
    const auto y = C2.wr(scr, nthr * n1, tid, nthr);
    if (dir == 0) {  // centre slot written once
        Sink c{U + (long)ns * plane_slot + i3 * slice + i, n2};
        for (int k = 0; k < n1; k++) c[k] = xs[(long)k * n2 + i];
    }
    // slots of this trace that no chain reaches (their source lies outside
    // the slice) are zero: toward lower i2 the source of slot ns-is-1 is
    // i+is+1, toward higher i2 that of slot ns+is+1 is i-is-1
    for (int is = dir == 0 ? max(0, n2 - 1 - i) : min(i, ns); is < ns; is++) {
        const int slot = dir == 0 ? ns - is - 1 : ns + is + 1;
        Sink z{U + (long)slot * plane_slot + i3 * slice + i, n2};
        for (int k = 0; k < n1; k++) z[k] = 0.0f;
    }
    const long nU = (2L * ns + 1) * n;
    auto cur = C2.rd(x, n, i3 * slice + i, n2);
    for (int is = 0; is < ns; is++) {
        int ip = dir == 0 ? i - is - 1 : i + is + 1;
        if (ip < 0 || ip >= n2) break;
        int slot = dir == 0 ? ns - is - 1 : ns + is + 1;
        const long eo = (long)slot * plane_slot + i3 * slice + ip;
        const auto out = C2.wr(U, nU, eo, n2);
        decltype(cur) in[1] = {cur};
        decltype(cur) qs[1] = {C2.rd(dip, n, i3 * slice + (dir == 0 ? ip : ip - 1), n2)};
        bool rv[1] = {dir == 1};
        const auto Fv = C2.template fac<NB>(F, pf, 3, n, dir, i3 * slice + ip);
        const auto Tq = C2.template fac<NB>(F, pf, 3, n, 2, i3 * slice + (dir == 0 ? ip : ip - 1));
        solve_factored_pf<NW, 1, PFD, true, true>(in, qs, rv, Fv, y, n1, e, out, &Tq);
        cur = C2.rd(U, nU, eo, n2);
    }
}

// Adjoint chains (pwspray_lop adj, sof_cfuns.c:958-996): for output trace i,
// chain dir 1 gathers slots ns+is+1 of traces i+is+1 (is = ns-1..0) through
// adjoint predictions; dir 0 likewise from the left.  Results go to
// C[dir][i3][k][i]; the centre slot is added by spray2_adj_combine.
template <bool BUF, int NW, int PFD>
__global__ void __launch_bounds__(TB)
spray2_adj(const float *__restrict__ U, const float *__restrict__ dip, const float *__restrict__ F,
           float *__restrict__ C, int n1, int n2, int n3, int ns, float e, float *__restrict__ scr,
           float *__restrict__ tr, long nthr, const int *stop) {
    if (stop && *(volatile const int *)stop) return;
    long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i3 = blockIdx.y;
    if (gid >= 2L * n2) return;
    int dir = (int)(gid / n2);
    int i = (int)(gid % n2);
    long slice = (long)n1 * n2;
    long plane_slot = slice * n3, n = plane_slot;
    constexpr int NB = 2 * NW;
    const long pf = fac_plane_floats(NB, n);
    const Chain2<BUF> C2{n2};
    long tid = (long)i3 * 2 * n2 + gid;
    const auto y = C2.wr(scr, nthr * n1, tid, nthr);
    const auto t = C2.wr(tr, nthr * n1, tid, nthr);  // running trace
    const auto ta = C2.rd(tr, nthr * n1, tid, nthr);
    {
        Sink z{tr + tid, nthr};
        for (int k = 0; k < n1; k++) z[k] = 0.0f;
    }
    const long nU = (2L * ns + 1) * n;
    for (int is = ns - 1; is >= 0; is--) {
        int ip = dir == 1 ? i + is + 1 : i - is - 1;
        if (ip < 0 || ip >= n2) continue;
        int slot = dir == 1 ? ns + is + 1 : ns - is - 1;
        const auto u = C2.rd(U, nU, (long)slot * plane_slot + i3 * slice + ip, n2);
        const auto q = C2.rd(dip, n, i3 * slice + (dir == 1 ? ip - 1 : ip), n2);
        const auto Fv = C2.template fac<NB>(F, pf, 3, n, dir == 1 ? 1 : 0, i3 * slice + ip);
        const auto Tq = C2.template fac<NB>(F, pf, 3, n, 2, i3 * slice + (dir == 1 ? ip - 1 : ip));
        // t = S^{-1}-adjoint prediction of (t + u), loads issued PFD samples ahead
        predict_adj_factored_pf<NW, PFD, true, true>(ta, &u, q, dir == 1, Fv, y, n1, e, t, &Tq);
    }
    Sink c{C + (long)dir * plane_slot + i3 * slice + i, n2};
    const float *tp = tr + tid;
    for (int k = 0; k < n1; k++) c[k] = tp[(long)k * nthr];
}

// x = ((x or 0) + chain_right) + chain_left + centre: the accumulation
// order of the adjoint spray (sof_cfuns.c:969-995).  C[0] holds the chains
// from the left (dir 0), C[1] those from the right (dir 1).
__global__ void spray2_adj_combine(const float *__restrict__ C, const float *__restrict__ U, float *__restrict__ x,
                                   long n, int ns, int add) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = add ? x[i] : 0.0f;
    v = v + C[n + i];
    v = v + C[i];
    v = v + U[(long)ns * n + i];
    x[i] = v;
}

// pwsmooth forward weighting (sof_cfuns.c:1087-1098):
// out = sum_is (u_is * w_is) * ww, ww = 1/t (t from the spray of ones).
__global__ void pwsmooth_norm(const float *__restrict__ U1, float *__restrict__ ww, long n, int ns) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ns2 = 2 * ns + 1;
    float t = 0.0f;
    for (int is = 0; is < ns2; is++) {
        float w = (float)(ns + 1 - abs(is - ns));
        t += U1[(long)is * n + i] * w * 1.0f;
    }
    ww[i] = (0.0f != t) ? (float)(1.0 / t) : 0.0f;
}

__global__ void pwsmooth_fwd(const float *__restrict__ U, const float *__restrict__ ww, float *__restrict__ out,
                             long n, int ns, int add) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ns2 = 2 * ns + 1;
    float ws = ww[i];
    float s = add ? out[i] : 0.0f;
    for (int is = 0; is < ns2; is++) {
        float w = (float)(ns + 1 - abs(is - ns));
        s += U[(long)is * n + i] * w * ws;
    }
    out[i] = s;
}

// adjoint weighting: u_is = (smooth * w_is) * ww (sof_cfuns.c:1074-1082)
__global__ void pwsmooth_adj_weight(const float *__restrict__ sm, const float *__restrict__ ww, float *__restrict__ U,
                                    long n, int ns) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ns2 = 2 * ns + 1;
    float ws = ww[i], v = sm[i];
    for (int is = 0; is < ns2; is++) {
        float w = (float)(ns + 1 - abs(is - ns));
        U[(long)is * n + i] = v * w * ws;
    }
}

__global__ void fill_one(float *__restrict__ y, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 1.0f;
}

// 3-D visit plan by Manhattan level (dependencies are the axis neighbours
// toward the centre, sof3d_cfuns.c:1022-1054).
int plan3(int ns2, int ns3, std::vector<std::vector<Node3>> &levels) {
    int np2 = 2 * ns2 + 1, np3 = 2 * ns3 + 1;
    levels.assign(ns2 + ns3 + 1, {});
    for (int k3 = 0; k3 < np3; k3++)
        for (int k2 = 0; k2 < np2; k2++) {
            int a = k2 - ns2, b = k3 - ns3;
            int L = std::abs(a) + std::abs(b);
            Node3 nd{};
            nd.slot = k2 + np2 * k3;
            nd.o2 = a;
            nd.o3 = b;
            nd.update = (a != 0 ? 1 : 0) | (b != 0 ? 2 : 0);
            nd.up2 = a > 0;  // predecessor toward the centre along axis 2
            nd.up3 = b > 0;
            if (L > 0) levels[L].push_back(nd);
        }
    return SH_OK;
}

struct SprayBufs {
    float *d, *pi, *px, *U, *scr;
};

// 3-D spray into U (trace-fast, [slot][k][trace]) from trace-fast inputs.
// Slots that a median over the linear slot window [first, first+nmf) reads,
// plus every slot they are predicted from (their predecessors toward the
// centre, recursively).  The other slots never reach the output.
std::vector<char> needed_slots(int ns2, int ns3, int first, int nmf) {
    int np2 = 2 * ns2 + 1, np = np2 * (2 * ns3 + 1);
    std::vector<char> need(np, 0);
    std::vector<int> stack;
    for (int s = first; s < first + nmf; s++) stack.push_back(s);
    while (!stack.empty()) {
        int s = stack.back();
        stack.pop_back();
        if (s < 0 || s >= np || need[s]) continue;
        need[s] = 1;
        int a = s % np2 - ns2, b = s / np2 - ns3;
        if (a != 0) stack.push_back(s + (a > 0 ? -1 : 1));
        if (b != 0) stack.push_back(s + (b > 0 ? -np2 : np2));
    }
    return need;
}

// need (nullable): only these slots are predicted (the others are left
// unwritten and must not be read).
int spray3_run(sh_ctx *ctx, int order, int n1, int n2, int n3, int ns2, int ns3, float e, const float *dT,
               const float *piT, const float *pxT, float *U, float **scr_io, const std::vector<char> *need = nullptr) {
    long ntr = (long)n2 * n3, n = (long)n1 * ntr;
    int np2 = 2 * ns2 + 1, np = np2 * (2 * ns3 + 1);
    std::vector<std::vector<Node3>> levels;
    plan3(ns2, ns3, levels);
    if (need) {
        for (auto &l : levels) {
            std::vector<Node3> kept;
            for (auto &nd : l)
                if ((*need)[nd.slot]) kept.push_back(nd);
            l.swap(kept);
        }
    }
    size_t maxn = 0;
    for (auto &l : levels) maxn = std::max(maxn, l.size());
    if (maxn > (size_t)MAX_LEVEL_NODES) return sh_fail(ctx, SH_EUNSUP, "spray radius too large");
    int g = ceil_div(n, TB);
    // every slot of every target is written below: the prediction, or zeros
    // where the slot's source lies outside the volume
    int ctr = ns3 * np2 + ns2;
    hipLaunchKernelGGL(copy_slot, dim3(g), dim3(TB), 0, ctx->stream, dT, U + (long)ctr * n, n);
    SH_TRY(sh_launch_check(ctx, "copy_slot"));
    int gx = ceil_div(ntr, TB);
    const long gx8 = (long)ceil_div(gx, 8) * 8;  // level launches: gx8 blocks per node
    // per-thread scratch planes: one per node of a level (ungrouped), or
    // MAX_GROUPS x GM per launch (grouped)
    long nthr = gx8 * TB * (long)std::max<size_t>(maxn, (size_t)MAX_GROUPS * GM);
    float *scr = nullptr, *F = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 4, (size_t)nthr * n1 * sizeof(float), (void **)&scr));
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 8, (size_t)8 * fac_plane_floats(2 * order, n) * sizeof(float), (void **)&F));
    {
        PROF_A(ctx);
        SH_DISPATCH_NW(ctx, order, NW,
                       hipLaunchKernelGGL((factor3_kernel<NW>), dim3(gx, 8), dim3(TB), 0, ctx->stream, piT, pxT, F,
                                          n1, n2, n3, e));
        PROF_B(ctx, "factor3_kernel", 2.0 * 4.0 * (double)n);
        SH_TRY(sh_launch_check(ctx, "factor3_kernel"));
    }
    // SH_SPRAY_GROUP=g: at most g two-source slots of a quadrant per thread
    // (1..4); 0 (default) = one thread per target trace and slot.  Measured
    // at 512x256x256 for the 24 two-source slots somf3d needs: g = 0
    // 15.3 ms, g = 1 16.3 ms, g = 2..4 17.5 ms (the register pressure of
    // several right-hand sides costs more than the shared loads save).
    const char *ge = getenv("SH_SPRAY_GROUP");
    const int gmax = ge ? std::max(0, std::min(GM, atoi(ge))) : 0;
    const bool grouped = gmax > 0;
    // SH_SPRAY_XCD=0: node-major block order in the level launches instead
    // of the XCD-grouped one (same results, for timing)
    const char *xe = getenv("SH_SPRAY_XCD");
    const int xcd = (xe && xe[0] == '0') ? 0 : 1;
    // SH_SPRAY_PF=d: prefetch distance (samples) of the level predictions,
    // 0 (no prefetch), 2, 4, 8, or 6 (two-source levels; the one-source
    // ones then use 8); identical results.  Default 8 for one source trace,
    // 6 for two (more registers per sample; at 8 the kernel drops to one
    // wave per SIMD).  Measured at 512x256x256 (somf3d, r = 4), ms per
    // warm step: two sources 9.45 / 8.85 for d = 4 / 6 (12.0 at 8 with the
    // factor in planes), one source 4.63 / 4.42 for d = 4 / 8.
    const char *pe = getenv("SH_SPRAY_PF");
    const char *be = getenv("SH_SPRAY_BUF");
    for (size_t L = 1; L < levels.size(); L++) {
        if (grouped) {  // two-source slots by (quadrant, group size)
            std::vector<Group3> groups;
            for (int quad = 0; quad < 4; quad++) {
                Group3 cur{};
                for (auto &nd : levels[L]) {
                    if (nd.update != 3 || (nd.up2 | (nd.up3 << 1)) != quad) continue;
                    cur.nd[cur.cnt++] = nd;
                    if (cur.cnt == gmax) {
                        groups.push_back(cur);
                        cur = Group3{};
                    }
                }
                if (cur.cnt) groups.push_back(cur);
            }
            for (int m = 1; m <= GM; m++) {
                GroupLevel gl{};
                int cnt = 0;
                double by = 0;
                for (auto &g : groups)
                    if (g.cnt == m) {
                        if (cnt == MAX_GROUPS) return sh_fail(ctx, SH_EUNSUP, "spray radius too large");
                        gl.g[cnt++] = g;
                        by += 20.0 * m * (double)n;
                    }
                if (!cnt) continue;
                dim3 grid(gx, cnt);
                PROF_A(ctx);
#define SG(MM)                                                                                                  \
    if (m == MM)                                                                                                \
        SH_DISPATCH_NW(ctx, order, NW,                                                                          \
                       hipLaunchKernelGGL((spray3_group<NW, MM>), grid, dim3(TB), 0, ctx->stream, piT, pxT, U, F, \
                                          gl, n1, n2, n3, np2, e, scr, nthr));
                SG(1) SG(2) SG(3) SG(4)
#undef SG
                PROF_B(ctx, "spray3_group", by);
                SH_TRY(sh_launch_check(ctx, "spray3_group"));
            }
        }
        for (int nf = 1; nf <= (grouped ? 1 : 2); nf++) {
            Level3 lv{};
            int cnt = 0;
            double by = 0;
            // nodes of one factor variant side by side (they share slopes
            // and factor); the nodes of a level are independent of each other
            for (int v = 0; v < 8; v++)
                for (auto &nd : levels[L])
                    if ((nd.update == 3 ? 2 : 1) == nf && node_variant(nd.update, nd.up2, nd.up3) == v) {
                        lv.nd[cnt++] = nd;
                        by += (nf == 2 ? 20.0 : 12.0) * (double)n;  // per slot-sample: inputs, slopes, output
                    }
            if (!cnt) continue;
            dim3 grid((unsigned)(gx8 * cnt));
            PROF_A(ctx);
            const int pf = pe ? atoi(pe) : (nf == 1 ? 8 : 6);
            // buffer-resource accessors when every plane, the factor's
            // off-diagonal block and this launch's scratch stay below 4 GiB
            // (and a prefetching solve); SH_SPRAY_BUF=0 keeps pointers
            const double lim = 4294967296.0;
            const bool buf = !(be && be[0] == '0') && pf > 0 && (double)n * 4 * (2 * order) < lim &&
                             (double)gx8 * TB * cnt * n1 * 4 < lim;
#define SL(NF_, PF_)                                                                                           \
    if (buf) {                                                                                                 \
        SH_DISPATCH_NW(ctx, order, NW,                                                                         \
                       hipLaunchKernelGGL((spray3_level<true, NW, NF_, PF_>), grid, dim3(TB), 0, ctx->stream,   \
                                          piT, pxT, U, F, lv, cnt, xcd, n1, n2, n3, np2, e, scr, nthr));        \
    } else {                                                                                                   \
        SH_DISPATCH_NW(ctx, order, NW,                                                                         \
                       hipLaunchKernelGGL((spray3_level<false, NW, NF_, PF_>), grid, dim3(TB), 0, ctx->stream,  \
                                          piT, pxT, U, F, lv, cnt, xcd, n1, n2, n3, np2, e, scr, nthr));        \
    }
            // timer names carry the kernel's trailing template arguments
            // (sources, prefetch distance)
            if (nf == 1) {
                if (pf == 8 || pf == 6) { SL(1, 8); PROF_B(ctx, "spray3_level<1,8>", by); }
                else if (pf == 4) { SL(1, 4); PROF_B(ctx, "spray3_level<1,4>", by); }
                else if (pf == 2) { SL(1, 2); PROF_B(ctx, "spray3_level<1,2>", by); }
                else { SL(1, 0); PROF_B(ctx, "spray3_level<1,0>", by); }
            } else {
                if (pf == 8) { SL(2, 8); PROF_B(ctx, "spray3_level<2,8>", by); }
                else if (pf == 6) { SL(2, 6); PROF_B(ctx, "spray3_level<2,6>", by); }
                else if (pf == 4) { SL(2, 4); PROF_B(ctx, "spray3_level<2,4>", by); }
                else if (pf == 2) { SL(2, 2); PROF_B(ctx, "spray3_level<2,2>", by); }
                else { SL(2, 0); PROF_B(ctx, "spray3_level<2,0>", by); }
            }
#undef SL
            SH_TRY(sh_launch_check(ctx, "spray3_level"));
        }
    }
    if (scr_io) *scr_io = scr;
    return SH_OK;
}

int check3(sh_ctx *ctx, const float *a, const float *b, const float *c, int n1, int n2, int n3, int r1, int r2,
           const char *who) {
    if (!a || !b || !c) return sh_fail(ctx, SH_ENULL, "%s: NULL input", who);
    if (n1 < 8 || n2 < 1 || n3 < 1) return sh_fail(ctx, SH_EINVAL, "%s: bad shape %d x %d x %d (n1 >= 8)", who, n1, n2, n3);
    if (r1 < 0 || r2 < 0 || r1 > 16 || r2 > 16) return sh_fail(ctx, SH_EINVAL, "%s: spray radius out of range", who);
    return SH_OK;
}

// inputs F-order -> trace-fast buffers in one workspace slot
int stage3(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, int n1, long ntr, float **dT,
           float **piT, float **pxT) {
    long n = (long)n1 * ntr;
    float *b = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 1, (size_t)n * 3 * sizeof(float), (void **)&b));
    *dT = b; *piT = b + n; *pxT = b + 2 * n;
    SH_TRY(transpose(ctx, din, *dT, n1, ntr, true));
    SH_TRY(transpose(ctx, dipi, *piT, n1, ntr, true));
    if (dipx) SH_TRY(transpose(ctx, dipx, *pxT, n1, ntr, true));
    return SH_OK;
}

}  // namespace

// Internal API (declared in common.hpp): 2-D structure-oriented smoothing.

// Scratch (solver + running trace planes) and the factor cache of P.
// Threads of the 2-D spray chains (two directions per trace and slice).
long pws2_nthr(int n2, int n3) { return (long)ceil_div(2L * n2, TB) * TB * n3; }

// Block size of the 2-D chain kernels (one thread per trace, direction and
// slice); SH_SPRAY2_TB=64 or 256 overrides.
int spray2_tb() {
    const char *e = getenv("SH_SPRAY2_TB");
    if (e && atoi(e) == 64) return 64;
    if (e && atoi(e) == 256) return 256;
    return 256;
}

// Prefetch distance (samples) of the chain kernels' load rings, 8 or 16
// (SH_SPRAY2_PF); identical results.  The chains run at about one wave per
// SIMD and issue-bound: with the taps loaded from the taps plane (five
// values per ring slot) a distance of 16 no longer fits the registers
// (sint3d 300x100x100: chains 17.1 ms at 8, 21.5 ms at 16).
int spray2_pf(bool forward) {
    (void)forward;
    const char *e = getenv("SH_SPRAY2_PF");
    if (e && atoi(e) == 16) return 16;
    return 8;
}

#define SPRAY2_PF_DISPATCH(PF, FWD, ...)         \
    do {                                         \
        if (spray2_pf(FWD) == 16) {              \
            constexpr int PF = 16;               \
            __VA_ARGS__;                         \
        } else {                                 \
            constexpr int PF = 8;                \
            __VA_ARGS__;                         \
        }                                        \
    } while (0)

// Buffer-resource accessors in the chain kernels when every buffer they
// address (input and slopes, slot volume, factor planes, scratch planes)
// stays below 4 GiB; SH_SPRAY2_BUF=0 keeps pointers (identical results).
bool spray2_buf(int n1, long n, int ns, int order, long nthr) {
    const char *e = getenv("SH_SPRAY2_BUF");
    if (e && e[0] == '0') return false;
    const double lim = 4294967295.0;
    return 4.0 * n * (2 * ns + 1) < lim && 4.0 * 3 * fac_plane_floats(2 * order, n) < lim &&
           4.0 * 2 * nthr * n1 < lim;
}

#define SPRAY2_BUF_DISPATCH(BUFV, cond, ...) \
    do {                                     \
        if (cond) {                          \
            constexpr bool BUFV = true;      \
            __VA_ARGS__;                     \
        } else {                             \
            constexpr bool BUFV = false;     \
            __VA_ARGS__;                     \
        }                                    \
    } while (0)

// Factor cache of P into caller-provided buffers: scr holds 2*nthr*n1
// floats (solver scratch, then the running trace planes), F
// 3 * fac_plane_floats(2 * order, n): the two factor variants, then the
// filter taps of the slopes.
int pws2_prepare_bufs(sh_ctx *ctx, Pwsmooth2 &P, float *scr, float *F) {
    long n = (long)P.n1 * P.n2 * P.n3;
    P.nthr = pws2_nthr(P.n2, P.n3);
    P.scr = scr;
    P.tr = P.scr + (long)P.nthr * P.n1;
    P.F = F;
    const int TB2 = spray2_tb();
    dim3 grid(ceil_div(2L * P.n2, TB2), P.n3);
    PROF_A(ctx);
    SH_DISPATCH_NW(ctx, P.order, NW,
                   hipLaunchKernelGGL((factor2_kernel<NW>), grid, dim3(TB2), 0, ctx->stream, P.dip, P.F, P.n1, P.n2,
                                      P.n3, P.e));
    PROF_B(ctx, "factor2_kernel", 4.0 * (double)n);
    SH_TRY(sh_launch_check(ctx, "factor2_kernel"));
    SH_DISPATCH_NW(ctx, P.order, NW,
                   hipLaunchKernelGGL((taps_plane_kernel<NW>), dim3(ceil_div(n, TB)), dim3(TB), 0, ctx->stream, P.dip,
                                      P.F + 2 * fac_plane_floats(2 * NW, n), n));
    return sh_launch_check(ctx, "taps_plane_kernel");
}

int pws2_prepare(sh_ctx *ctx, Pwsmooth2 &P, int slot_scr, int slot_F) {
    long n = (long)P.n1 * P.n2 * P.n3;
    const long nthr = pws2_nthr(P.n2, P.n3);
    float *scr = nullptr, *F = nullptr;
    SH_TRY(sh_slot(ctx, slot_scr, (size_t)nthr * P.n1 * 2 * sizeof(float), (void **)&scr));
    SH_TRY(sh_slot(ctx, slot_F, (size_t)3 * fac_plane_floats(2 * P.order, n) * sizeof(float), (void **)&F));
    return pws2_prepare_bufs(ctx, P, scr, F);
}

// Spray of x into P.U; with x2 also the spray of x2 into U2 in the same
// launch (it uses the scratch plane P.tr, free outside the adjoint), so
// two latency-bound sets of chains run side by side.
int pws2_spray(sh_ctx *ctx, Pwsmooth2 &P, const float *x, const float *x2 = nullptr, float *U2 = nullptr) {
    const int TB2 = spray2_tb();
    dim3 grid(ceil_div(2L * P.n2, TB2), P.n3, x2 ? 2 : 1);
    PROF_A(ctx);
    const bool buf = spray2_buf(P.n1, (long)P.n1 * P.n2 * P.n3, P.ns, P.order, P.nthr);
    SPRAY2_BUF_DISPATCH(BUF, buf, SPRAY2_PF_DISPATCH(PF, true, SH_DISPATCH_NW(ctx, P.order, NW,
                   hipLaunchKernelGGL((spray2_fwd<BUF, NW, PF>), grid, dim3(TB2), 0, ctx->stream, x, P.dip, P.F, P.U,
                                      P.n1, P.n2, P.n3, P.ns, P.e, P.scr, P.nthr, P.stop, x2, U2))));
    PROF_B(ctx, "spray2_fwd", (x2 ? 2 : 1) * 12.0 * 2 * P.ns * (double)P.n1 * P.n2 * P.n3);
    return sh_launch_check(ctx, "spray2_fwd");
}

int pws2_setup(sh_ctx *ctx, Pwsmooth2 &P, int slot_base) {
    long n = (long)P.n1 * P.n2 * P.n3;
    int ns2 = 2 * P.ns + 1;
    float *b = nullptr;
    SH_TRY(sh_slot(ctx, slot_base, (size_t)n * (ns2 + 3) * sizeof(float), (void **)&b));
    P.U = b;
    P.ww = b + (long)ns2 * n;
    P.C = b + (long)(ns2 + 1) * n;
    SH_TRY(pws2_prepare(ctx, P, slot_base + 1, slot_base + 2));
    // normalisation: spray of ones (pwsmooth_set, sof_cfuns.c:1102-1122)
    float *ones = P.C;  // C is free at this point (2n floats)
    int g = ceil_div(n, TB);
    hipLaunchKernelGGL(fill_one, dim3(g), dim3(TB), 0, ctx->stream, ones, n);
    SH_TRY(sh_launch_check(ctx, "fill_one"));
    SH_TRY(pws2_spray(ctx, P, ones));  // writes every slot
    hipLaunchKernelGGL(pwsmooth_norm, dim3(g), dim3(TB), 0, ctx->stream, P.U, P.ww, n, P.ns);
    return sh_launch_check(ctx, "pwsmooth_norm");
}

// pws2_setup + the forward smoothing of x in one pass: the spray of ones
// (normalisation) and the spray of x run in one launch, each into its own
// slot volume; results identical to pws2_setup then pws2_apply(adj = 0).
int pws2_setup_forward(sh_ctx *ctx, Pwsmooth2 &P, int slot_base, const float *x, float *smooth) {
    long n = (long)P.n1 * P.n2 * P.n3;
    int ns2 = 2 * P.ns + 1;
    float *b = nullptr;
    SH_TRY(sh_slot(ctx, slot_base, (size_t)n * (2 * ns2 + 3) * sizeof(float), (void **)&b));
    P.U = b;
    float *U1 = b + (long)ns2 * n;
    P.ww = b + (long)(2 * ns2) * n;
    P.C = b + (long)(2 * ns2 + 1) * n;
    SH_TRY(pws2_prepare(ctx, P, slot_base + 1, slot_base + 2));
    float *ones = P.C;
    int g = ceil_div(n, TB);
    hipLaunchKernelGGL(fill_one, dim3(g), dim3(TB), 0, ctx->stream, ones, n);
    SH_TRY(sh_launch_check(ctx, "fill_one"));
    SH_TRY(pws2_spray(ctx, P, x, ones, U1));
    hipLaunchKernelGGL(pwsmooth_norm, dim3(g), dim3(TB), 0, ctx->stream, U1, P.ww, n, P.ns);
    SH_TRY(sh_launch_check(ctx, "pwsmooth_norm"));
    PROF_A(ctx);
    hipLaunchKernelGGL(pwsmooth_fwd, dim3(g), dim3(TB), 0, ctx->stream, P.U, P.ww, smooth, n, P.ns, 0);
    PROF_B(ctx, "pwsmooth_fwd", 4.0 * (ns2 + 2) * (double)n);
    return sh_launch_check(ctx, "pwsmooth_fwd");
}

// smooth (+)= S trace  or  trace (+)= S^T smooth   (pwsmooth_lop)
int pws2_apply(sh_ctx *ctx, Pwsmooth2 &P, int adj, int add, float *trace, float *smooth) {
    long n = (long)P.n1 * P.n2 * P.n3;
    int ns2 = 2 * P.ns + 1;
    int g = ceil_div(n, TB);
    if (!adj) {
        SH_TRY(pws2_spray(ctx, P, trace));  // writes every slot
        PROF_A(ctx);
        hipLaunchKernelGGL(pwsmooth_fwd, dim3(g), dim3(TB), 0, ctx->stream, P.U, P.ww, smooth, n, P.ns, add);
        PROF_B(ctx, "pwsmooth_fwd", 4.0 * (ns2 + 2 + add) * (double)n);
        return sh_launch_check(ctx, "pwsmooth_fwd");
    }
    {
        PROF_A(ctx);
        hipLaunchKernelGGL(pwsmooth_adj_weight, dim3(g), dim3(TB), 0, ctx->stream, smooth, P.ww, P.U, n, P.ns);
        PROF_B(ctx, "pwsmooth_adj_weight", 4.0 * (ns2 + 2) * (double)n);
    }
    SH_TRY(sh_launch_check(ctx, "pwsmooth_adj_weight"));
    const int TB2 = spray2_tb();
    dim3 grid(ceil_div(2L * P.n2, TB2), P.n3);
    {
        PROF_A(ctx);
        const bool buf = spray2_buf(P.n1, n, P.ns, P.order, P.nthr);
        SPRAY2_BUF_DISPATCH(BUF, buf, SPRAY2_PF_DISPATCH(PF, false, SH_DISPATCH_NW(ctx, P.order, NW,
                       hipLaunchKernelGGL((spray2_adj<BUF, NW, PF>), grid, dim3(TB2), 0, ctx->stream, P.U, P.dip, P.F,
                                          P.C, P.n1, P.n2, P.n3, P.ns, P.e, P.scr, P.tr, P.nthr, P.stop))));
        PROF_B(ctx, "spray2_adj", 12.0 * 2 * P.ns * (double)n);
    }
    SH_TRY(sh_launch_check(ctx, "spray2_adj"));
    PROF_A(ctx);
    hipLaunchKernelGGL(spray2_adj_combine, dim3(g), dim3(TB), 0, ctx->stream, P.C, P.U, trace, n, P.ns, add);
    PROF_B(ctx, "spray2_adj_combine", 4.0 * (4 + add) * (double)n);
    return sh_launch_check(ctx, "spray2_adj_combine");
}

int sh_transpose_internal(sh_ctx *ctx, const float *src, float *dst, int n1, long ntr, bool to_fast, int nslice) {
    return transpose(ctx, src, dst, n1, ntr, to_fast, nslice);
}

int sh_median_internal(sh_ctx *ctx, const float *U, float *out, long n, int first, int nmf) {
    return launch_median(ctx, U, out, n, first, nmf);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int shd_pwspray3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, int n1, int n2, int n3,
                 int ns2, int ns3, int order, float eps, float *u) {
    if (!ctx) return SH_ENULL;
    SH_TRY(check3(ctx, din, dipi, dipx, n1, n2, n3, ns2, ns3, "pwspray3d"));
    if (!u) return sh_fail(ctx, SH_ENULL, "pwspray3d: NULL output");
    long ntr = (long)n2 * n3, n = (long)n1 * ntr;
    int np = (2 * ns2 + 1) * (2 * ns3 + 1);
    float *dT, *piT, *pxT, *U;
    SH_TRY(stage3(ctx, din, dipi, dipx, n1, ntr, &dT, &piT, &pxT));
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 2, (size_t)n * np * sizeof(float), (void **)&U));
    SH_TRY(spray3_run(ctx, order, n1, n2, n3, ns2, ns3, eps * eps, dT, piT, pxT, U, nullptr));
    // [slot][k][trace] -> u[(trace*np + slot)*n1 + k]: per slot transpose into a
    // [trace][k] plane, then interleave slots with a strided copy.
    float *tmp = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 3, (size_t)n * sizeof(float), (void **)&tmp));
    for (int s = 0; s < np; s++) {
        SH_TRY(transpose(ctx, U + (long)s * n, tmp, n1, ntr, false));
        SH_HIP(ctx, hipMemcpy2DAsync(u + (long)s * n1, (size_t)np * n1 * sizeof(float), tmp, (size_t)n1 * sizeof(float),
                                     (size_t)n1 * sizeof(float), (size_t)ntr, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return SH_OK;
}

int shd_somean3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, int n1, int n2, int n3, int r1,
                int r2, float eps, int order, float *out) {
    (void)eps;
    if (!ctx) return SH_ENULL;
    SH_TRY(check3(ctx, din, dipi, dipx, n1, n2, n3, r1, r2, "somean3d"));
    if (!out) return sh_fail(ctx, SH_ENULL, "somean3d: NULL output");
    long ntr = (long)n2 * n3, n = (long)n1 * ntr;
    int np = (2 * r1 + 1) * (2 * r2 + 1);
    float *dT, *piT, *pxT, *U;
    SH_TRY(stage3(ctx, din, dipi, dipx, n1, ntr, &dT, &piT, &pxT));
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 2, (size_t)n * np * sizeof(float), (void **)&U));
    const float e = 0.01f * 0.01f;  // eps forced to 0.01 (sof3d_cfuns.c:1402)
    SH_TRY(spray3_run(ctx, order, n1, n2, n3, r1, r2, e, dT, piT, pxT, U, nullptr));
    hipLaunchKernelGGL(slot_mean, dim3(ceil_div(n, TB)), dim3(TB), 0, ctx->stream, U, dT, n, np);
    SH_TRY(sh_launch_check(ctx, "slot_mean"));
    return transpose(ctx, dT, out, n1, ntr, false);
}

}  // extern "C"

namespace {

#include "svmf.hpp"

// Space-varying median per trace (svmf_one, svmf.hpp); one thread per
// trace, element (slot s, sample k) at U[s * n + base + k * stride].
__global__ void svmf_traces(const float *__restrict__ U, float *__restrict__ out, long n, int np, int n1, int nfw,
                            long ntraces, int n2, int two_d) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntraces) return;
    long base, stride;
    if (two_d) {
        const long i3 = t / n2, i2 = t % n2;
        base = i3 * (long)n1 * n2 + i2;
        stride = n2;
    } else {
        base = t;
        stride = ntraces;
    }
    svmf_one(U, out, n, np, n1, nfw, base, stride);
}

// The space-varying median of the 3-D filter with its work spread over
// the whole grid (svmf, sof3d_cfuns.c:1254-1352), in three launches:
//  1. every first-pass median (slot s, sample k, trace t): the order
//     statistic of the nfw + 2 columns around s (clamped into the panel),
//     by a selection network of N >= nfw + 2 registers padded with -inf and
//     +inf; its magnitude goes to a[s][k][t];
//  2. per trace the float mean of those magnitudes, summed slot by slot and
//     sample by sample as the reference does (one thread per trace, eight
//     loads issued ahead of the additions);
//  3. per (sample, trace) the window length from |x| against the mean and
//     the median of that window, by the same padded network.
// Each step is exact, so the result equals svmf_one's.
template <int N>
__global__ void __launch_bounds__(TB)
svmf_first(const float *__restrict__ U, float *__restrict__ a, long ss, int np, int n1, int nfw, long ntr, long total) {
    const int nf = nfw + 2, m2 = (nfw - 1) / 2, pad = (N - nf) / 2;
    for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (long)gridDim.x * blockDim.x) {
        const long t = g % ntr;
        const long r = g / ntr;
        const int k = (int)(r % n1), s = (int)(r / n1);
        float v[N];
#pragma unroll
        for (int q = 0; q < N; q++) {
            if (q < nf) {
                int col = s + q - m2;
                col = col < 0 ? 0 : (col > np - 1 ? np - 1 : col);
                v[q] = U[(long)col * ss + (long)k * ntr + t];
            } else {
                v[q] = (q - nf < pad) ? -INFINITY : INFINITY;
            }
        }
#pragma unroll
        for (int c = 0; c < MedianNet<N>::NOPS; c++) {
            const int i0 = MedianNet<N>::OPS[c][0], i1 = MedianNet<N>::OPS[c][1];
            const float x = v[i0], y = v[i1];
            v[i0] = fminf(x, y);
            v[i1] = fmaxf(x, y);
        }
        a[g] = fabsf(v[MedianNet<N>::MED]);
    }
}

__global__ void __launch_bounds__(TB)
svmf_mean(const float *__restrict__ a, float *__restrict__ avg, int np, int n1, long ntr) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntr) return;
    const long cnt = (long)np * n1;
    float sum = 0.0f;
    long i = 0;
    for (; i + 8 <= cnt; i += 8) {
        float r[8];
#pragma unroll
        for (int u = 0; u < 8; u++) r[u] = a[(i + u) * ntr + t];
#pragma unroll
        for (int u = 0; u < 8; u++) sum = sum + r[u];
    }
    for (; i < cnt; i++) sum = sum + a[i * ntr + t];
    avg[t] = sum / (float)(n1 * np);
}

template <int N>
__global__ void __launch_bounds__(TB)
svmf_final(const float *__restrict__ U, const float *__restrict__ avg, float *__restrict__ out, long n, long ss,
           int np, int nfw, long ntr) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const long t = g % ntr;
    const long k = g / ntr;
    const int c = (np - 1) / 2;
    const float av = fabsf(U[(long)c * ss + g]), mean = avg[t];
    int len;
    if (av < mean) len = (av < mean / 2) ? nfw + 2 : nfw;
    else len = (av > mean * 2) ? nfw - 4 : nfw - 2;
    const int h = (len - 1) / 2, pad = (N - len) / 2;
    float v[N];
#pragma unroll
    for (int q = 0; q < N; q++) {
        if (q < len) {
            int col = c + q - h;
            col = col < 0 ? 0 : (col > np - 1 ? np - 1 : col);
            v[q] = U[(long)col * ss + k * ntr + t];
        } else {
            v[q] = (q - len < pad) ? -INFINITY : INFINITY;
        }
    }
#pragma unroll
    for (int cc = 0; cc < MedianNet<N>::NOPS; cc++) {
        const int i0 = MedianNet<N>::OPS[cc][0], i1 = MedianNet<N>::OPS[cc][1];
        const float x = v[i0], y = v[i1];
        v[i0] = fminf(x, y);
        v[i1] = fmaxf(x, y);
    }
    out[g] = v[MedianNet<N>::MED];
}

// svmf over the trace-fast slot volume U[np][n1][ntr] into out[n1][ntr];
// the one-thread-per-trace kernel where the window exceeds the networks.
int svmf3(sh_ctx *ctx, const float *U, float *out, int n1, long ntr, int np, int nfw, long ss = 0) {
    const long n = (long)n1 * ntr;
    if (ss == 0) ss = n;  // slot stride of U (the section size unless the slots hold several sections)
    const int nf = nfw + 2;
    if (nf > 81) {
        hipLaunchKernelGGL(svmf_traces, dim3(ceil_div(ntr, 64)), dim3(64), 0, ctx->stream, U, out, n, np, n1, nfw, ntr,
                           1, 0);
        return sh_launch_check(ctx, "svmf_traces");
    }
    float *a = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 10, ((size_t)np * n + ntr) * sizeof(float), (void **)&a));
    float *avg = a + (long)np * n;
    const long total = (long)np * n;
    const int g1 = (int)std::min<long>(ceil_div(total, TB), 1L << 20);
    {
        PROF_A(ctx);
#define SV1(NN) \
    if (nf <= NN) { hipLaunchKernelGGL((svmf_first<NN>), dim3(g1), dim3(TB), 0, ctx->stream, U, a, ss, np, n1, nfw, ntr, total); } else
        SV1(9) SV1(13) SV1(17) SV1(33) SV1(49) SV1(65) SV1(81) {}
#undef SV1
        PROF_B(ctx, "svmf_first", 4.0 * (double)(np + 1) * n);
    }
    SH_TRY(sh_launch_check(ctx, "svmf_first"));
    hipLaunchKernelGGL(svmf_mean, dim3(ceil_div(ntr, TB)), dim3(TB), 0, ctx->stream, a, avg, np, n1, ntr);
    SH_TRY(sh_launch_check(ctx, "svmf_mean"));
    PROF_A(ctx);
#define SV3(NN) \
    if (nf <= NN) { hipLaunchKernelGGL((svmf_final<NN>), dim3(ceil_div(n, TB)), dim3(TB), 0, ctx->stream, U, avg, out, n, ss, np, nfw, ntr); } else
    SV3(9) SV3(13) SV3(17) SV3(33) SV3(49) SV3(65) SV3(81) {}
#undef SV3
    PROF_B(ctx, "svmf_final", 4.0 * (double)(nf + 1) * n);
    return sh_launch_check(ctx, "svmf_final");
}

}  // namespace

extern "C" {

int shd_somf3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, int n1, int n2, int n3, int r1,
              int r2, int nmf, int option, int order, float eps, float *out) {
    (void)eps;
    if (!ctx) return SH_ENULL;
    SH_TRY(check3(ctx, din, dipi, dipx, n1, n2, n3, r1, r2, "somf3d"));
    if (!out) return sh_fail(ctx, SH_ENULL, "somf3d: NULL output");
    int np = (2 * r1 + 1) * (2 * r2 + 1);
    if (nmf < 1 || nmf % 2 == 0 || nmf > np)
        return sh_fail(ctx, SH_EINVAL, "somf3d: median length %d must be odd and <= %d", nmf, np);
    if (option != 1 && option != 2) return sh_fail(ctx, SH_EINVAL, "somf3d: option must be 1 or 2");
    if (option == 2 && (nmf < 5 || nmf + 2 > 128))
        return sh_fail(ctx, SH_EINVAL, "somf3d: the space-varying median needs 5 <= nmf <= 125");
    long ntr = (long)n2 * n3, n = (long)n1 * ntr;
    float *dT, *piT, *pxT, *U;
    SH_TRY(stage3(ctx, din, dipi, dipx, n1, ntr, &dT, &piT, &pxT));
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 2, (size_t)n * np * sizeof(float), (void **)&U));
    const float e = 0.01f * 0.01f;  // eps forced to 0.01 (sof3d_cfuns.c:1586)
    if (option == 2) {  // its first pass reads every slot
        SH_TRY(spray3_run(ctx, order, n1, n2, n3, r1, r2, e, dT, piT, pxT, U, nullptr, nullptr));
        SH_TRY(svmf3(ctx, U, dT, n1, ntr, np, nmf));
        return transpose(ctx, dT, out, n1, ntr, false);
    }
    int c = (np - 1) / 2, m = (nmf - 1) / 2;
    // only the slots the median window reads and their predecessors (37 of
    // 81 for r1 = r2 = 4, nmf = 33); the result is the same
    std::vector<char> need = needed_slots(r1, r2, c - m, nmf);
    SH_TRY(spray3_run(ctx, order, n1, n2, n3, r1, r2, e, dT, piT, pxT, U, nullptr, &need));
    SH_TRY(launch_median(ctx, U, dT, n, c - m, nmf));
    return transpose(ctx, dT, out, n1, ntr, false);
}

int shd_pwspray2d(sh_ctx *ctx, int adj, const float *dip, int n1, int n2, int ns, int order, float eps, float *x,
                 float *u) {
    if (!ctx) return SH_ENULL;
    if (!dip || !x || !u) return sh_fail(ctx, SH_ENULL, "pwspray2d: NULL pointer");
    if (n1 < 8 || n2 < 1 || ns < 0 || ns > 64) return sh_fail(ctx, SH_EINVAL, "pwspray2d: bad arguments");
    long n = (long)n1 * n2;
    int ns2 = 2 * ns + 1;
    float *b = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 1, (size_t)n * 3 * sizeof(float), (void **)&b));
    float *xT = b, *dT = b + n, *tmp = b + 2 * n;
    SH_TRY(transpose(ctx, dip, dT, n1, n2, true));
    Pwsmooth2 P{n1, n2, 1, ns, order, eps * eps, dT};
    float *wb = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 2, (size_t)n * (ns2 + 2) * sizeof(float), (void **)&wb));
    P.U = wb;
    P.C = wb + (long)ns2 * n;
    SH_TRY(pws2_prepare(ctx, P, SLOT_SPRAY0 + 4, SLOT_SPRAY0 + 9));
    if (!adj) {
        SH_TRY(transpose(ctx, x, xT, n1, n2, true));
        SH_TRY(pws2_spray(ctx, P, xT));  // writes every slot
        for (int s = 0; s < ns2; s++) {
            SH_TRY(transpose(ctx, P.U + (long)s * n, tmp, n1, n2, false));
            SH_HIP(ctx, hipMemcpy2DAsync(u + (long)s * n1, (size_t)ns2 * n1 * sizeof(float), tmp,
                                         (size_t)n1 * sizeof(float), (size_t)n1 * sizeof(float), (size_t)n2,
                                         hipMemcpyDeviceToDevice, ctx->stream));
        }
        return SH_OK;
    }
    // adjoint: u (reference layout) -> slots trace-fast -> chains -> x
    for (int s = 0; s < ns2; s++) {
        SH_HIP(ctx, hipMemcpy2DAsync(tmp, (size_t)n1 * sizeof(float), u + (long)s * n1, (size_t)ns2 * n1 * sizeof(float),
                                     (size_t)n1 * sizeof(float), (size_t)n2, hipMemcpyDeviceToDevice, ctx->stream));
        SH_TRY(transpose(ctx, tmp, P.U + (long)s * n, n1, n2, true));
    }
    const int TB2 = spray2_tb();
    dim3 grid(ceil_div(2L * n2, TB2), 1);
    const bool buf = spray2_buf(n1, n, ns, order, P.nthr);
    SPRAY2_BUF_DISPATCH(BUF, buf, SPRAY2_PF_DISPATCH(PF, false, SH_DISPATCH_NW(ctx, order, NW,
                   hipLaunchKernelGGL((spray2_adj<BUF, NW, PF>), grid, dim3(TB2), 0, ctx->stream, P.U, P.dip, P.F, P.C,
                                      n1, n2, 1, ns, P.e, P.scr, P.tr, P.nthr, nullptr))));
    SH_TRY(sh_launch_check(ctx, "spray2_adj"));
    hipLaunchKernelGGL(spray2_adj_combine, dim3(ceil_div(n, TB)), dim3(TB), 0, ctx->stream, P.C, P.U, xT, n, ns, 0);
    SH_TRY(sh_launch_check(ctx, "spray2_adj_combine"));
    return transpose(ctx, xT, x, n1, n2, false);
}

int shd_somean2d(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3, int ns, int order, int adj,
                float eps, float *out) {
    if (!ctx) return SH_ENULL;
    if (!dn || !dip || !out) return sh_fail(ctx, SH_ENULL, "somean2d: NULL pointer");
    if (n1 < 8 || n2 < 1 || n3 < 1 || ns < 0 || ns > 64) return sh_fail(ctx, SH_EINVAL, "somean2d: bad arguments");
    long n = (long)n1 * n2 * n3;
    float *b = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 1, (size_t)n * 3 * sizeof(float), (void **)&b));
    float *xT = b, *dT = b + n, *yT = b + 2 * n;
    // slices [i3][n1][n2] are per-slice transposes of [i3][n2][n1]
    // slices [i3][n1][n2] are per-slice transposes of [i3][n2][n1]
    SH_TRY(transpose(ctx, dip, dT, n1, n2, true, n3));
    SH_TRY(transpose(ctx, dn, xT, n1, n2, true, n3));
    Pwsmooth2 P{n1, n2, n3, ns, order, eps * eps, dT};
    if (!adj) {
        SH_TRY(pws2_setup_forward(ctx, P, SLOT_SPRAY0 + 6, xT, yT));
    } else {
        SH_TRY(pws2_setup(ctx, P, SLOT_SPRAY0 + 6));
        SH_TRY(pws2_apply(ctx, P, 1, 0, yT, xT));
    }
    SH_TRY(transpose(ctx, yT, out, n1, n2, false, n3));
    return SH_OK;
}

}  // extern "C"

namespace {

int somf2d_args(sh_ctx *ctx, const float *dn, const float *dip, const float *out, int n1, int n2, int n3, int ns,
                int nmf, int option) {
    if (!dn || !dip || !out) return sh_fail(ctx, SH_ENULL, "somf2d: NULL pointer");
    if (n1 < 8 || n2 < 1 || n3 < 1 || ns < 0 || ns > 40) return sh_fail(ctx, SH_EINVAL, "somf2d: bad arguments");
    int np = 2 * ns + 1;
    if (nmf < 1 || nmf % 2 == 0 || nmf > np)
        return sh_fail(ctx, SH_EINVAL, "somf2d: median length %d must be odd and <= %d", nmf, np);
    if (option != 1 && option != 2) return sh_fail(ctx, SH_EINVAL, "somf2d: option must be 1 or 2");
    if (option == 2 && (nmf < 5 || nmf + 2 > 128))
        return sh_fail(ctx, SH_EINVAL, "somf2d: the space-varying median needs 5 <= nmf <= 125");
    return SH_OK;
}

// Workspace of one somf2d call, in floats, each part a multiple of 64
// floats (256 bytes): staged data, dip and result (3n), the spray (n*np),
// the chains' solver scratch (2*nthr*n1) and the factor cache
// (2*(2*order+1)*n).
struct Somf2dWs {
    size_t b, U, scr, F, total;
};
Somf2dWs somf2d_ws(int n1, int n2, int n3, int ns, int order) {
    auto up = [](size_t v) { return (v + 63) & ~(size_t)63; };
    const size_t n = (size_t)n1 * n2 * n3;
    Somf2dWs w;
    w.b = up(3 * n);
    w.U = up(n * (2 * ns + 1));
    w.scr = up((size_t)pws2_nthr(n2, n3) * n1 * 2);
    w.F = up((size_t)3 * fac_plane_floats(2 * order, (long)n));
    w.total = w.b + w.U + w.scr + w.F;
    return w;
}

// csomf2d on device buffers; ws (nullable) is the caller's workspace of
// somf2d_ws(...).total floats, otherwise the context's cached slots.
int somf2d_core(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3, int ns, int nmf, int option,
                int order, float eps, float *out, float *ws) {
    int np = 2 * ns + 1;
    long n = (long)n1 * n2 * n3;
    float *b = nullptr;
    Pwsmooth2 P{n1, n2, n3, ns, order, eps * eps, nullptr};
    const Somf2dWs w = somf2d_ws(n1, n2, n3, ns, order);
    if (ws) {
        b = ws;
        P.U = ws + w.b;
    } else {
        SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 1, (size_t)n * 3 * sizeof(float), (void **)&b));
        SH_TRY(sh_slot(ctx, SLOT_SPRAY0 + 6, (size_t)n * np * sizeof(float), (void **)&P.U));
    }
    float *xT = b, *dT = b + n, *yT = b + 2 * n;
    P.dip = dT;
    // slices [i3][n1][n2] are per-slice transposes of [i3][n2][n1]
    SH_TRY(transpose(ctx, dip, dT, n1, n2, true, n3));
    SH_TRY(transpose(ctx, dn, xT, n1, n2, true, n3));
    if (ws)
        SH_TRY(pws2_prepare_bufs(ctx, P, ws + w.b + w.U, ws + w.b + w.U + w.scr));
    else
        SH_TRY(pws2_prepare(ctx, P, SLOT_SPRAY0 + 7, SLOT_SPRAY0 + 9));
    SH_TRY(pws2_spray(ctx, P, xT));  // writes every slot
    int m = (nmf - 1) / 2;
    if (option == 2) {
        const long ntraces = (long)n2 * n3;
        if (!ws && nmf + 2 <= 81) {
            // the slots are U[slot][section][sample][trace]: each section
            // runs the grid-wide kernels of the 3-D filter with the slot
            // stride of the whole volume (exact, the same result)
            const long sec = (long)n1 * n2;
            for (int i3 = 0; i3 < n3; i3++)
                SH_TRY(svmf3(ctx, P.U + i3 * sec, yT + i3 * sec, n1, n2, np, nmf, n));
        } else {
            // the fixed workspace or windows beyond the selection networks:
            // one thread per trace
            hipLaunchKernelGGL(svmf_traces, dim3(ceil_div(ntraces, 64)), dim3(64), 0, ctx->stream, P.U, yT, n, np, n1,
                               nmf, ntraces, n2, 1);
            SH_TRY(sh_launch_check(ctx, "svmf_traces"));
        }
    } else {
        SH_TRY(launch_median(ctx, P.U, yT, n, ns - m, nmf));
    }
    SH_TRY(transpose(ctx, yT, out, n1, n2, false, n3));
    return SH_OK;
}

// Fixed workspace of sh_somf2d_noalloc: every device that runs it holds one
// copy until the process ends.  Host code never uses this name as an
// address (on the host it is only a stand-in): the device address comes
// from hipGetSymbolAddress, once per device.
constexpr size_t NOALLOC_FLOATS = (size_t)SH_SOMF2D_NOALLOC_MAX_BYTES / sizeof(float);
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

constexpr int NOALLOC_MAX_DEV = 64;
std::mutex noalloc_mu;
float *noalloc_ptr[NOALLOC_MAX_DEV] = {};

// Device address of `scratch` on the context's device (looked up on the
// first call there, with that device current).
int noalloc_scratch(sh_ctx *ctx, float **p) {
    const int dev = ctx->device;
    if (dev < 0 || dev >= NOALLOC_MAX_DEV) return sh_fail(ctx, SH_EUNSUP, "somf2d_noalloc: device %d", dev);
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        int cur = 0;
        SH_HIP(ctx, hipGetDevice(&cur));
        if (cur != dev) SH_HIP(ctx, hipSetDevice(dev));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr)
            return sh_fail(ctx, SH_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        noalloc_ptr[dev] = (float *)ptr;
    }
    *p = noalloc_ptr[dev];
    return SH_OK;
}

}  // namespace

extern "C" {

int shd_somf2d(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3, int ns, int nmf, int option,
              int order, float eps, float *out) {
    if (!ctx) return SH_ENULL;
    SH_TRY(somf2d_args(ctx, dn, dip, out, n1, n2, n3, ns, nmf, option));
    return somf2d_core(ctx, dn, dip, n1, n2, n3, ns, nmf, option, order, eps, out, nullptr);
}

size_t sh_somf2d_noalloc_bytes(int n1, int n2, int n3, int ns, int order) {
    if (n1 < 1 || n2 < 1 || n3 < 1 || ns < 0 || order < 1) return 0;
    return somf2d_ws(n1, n2, n3, ns, order).total * sizeof(float);
}

int sh_somf2d_noalloc(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3, int ns, int nmf,
                      int option, int order, float eps, float *out) {
    if (!ctx) return SH_ENULL;
    if (ctx->flags & SH_FLAG_HOST_PTRS)
        return sh_fail(ctx, SH_EINVAL, "somf2d_noalloc: device pointers only (host pointers need staging buffers)");
    SH_TRY(somf2d_args(ctx, dn, dip, out, n1, n2, n3, ns, nmf, option));
    if (order < 1 || order > 4) return sh_fail(ctx, SH_EUNSUP, "PWD order %d not supported (1..4)", order);
    const size_t need = somf2d_ws(n1, n2, n3, ns, order).total;
    if (need > NOALLOC_FLOATS)
        return sh_fail(ctx, SH_EINVAL, "somf2d_noalloc: %d x %d x %d with ns = %d, order = %d needs %zu bytes of "
                       "workspace, above the fixed %zu", n1, n2, n3, ns, order, need * sizeof(float),
                       NOALLOC_FLOATS * sizeof(float));
    float *ws = nullptr;
    SH_TRY(noalloc_scratch(ctx, &ws));
    // launch timers would create events: off for this call
    const int prof = ctx->prof;
    ctx->prof = 0;
    const int rc = somf2d_core(ctx, dn, dip, n1, n2, n3, ns, nmf, option, order, eps, out, ws);
    ctx->prof = prof;
    return rc;
}

}  // extern "C"
