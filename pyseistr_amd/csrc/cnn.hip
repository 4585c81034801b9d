// Learned warm start for dip estimation: a 6-layer 3-D CNN (3x3x3 kernels,
// 32 channels, ReLU) that maps a normalised seismic volume to an initial
// (inline, crossline) dip field.  No counterpart in the reference; the dips
// it produces only seed the reference's Gauss-Newton loop (sh_dip init).
//
// Kernels: BF16 activations in channels-last layout [voxel][32] (voxel =
// F-order index), FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16) for the 32-channel layers, implicit GEMM with
// zero padding at the volume faces.  Backward: input gradients reuse the
// forward kernel with flipped/transposed weights and the ReLU mask of the
// layer below; weight gradients are per-block MFMA partials reduced in a
// fixed order.  Master weights and Adam state are FP32.
#include "common.hpp"

#include <algorithm>
#include <cstdlib>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Raw buffer access (32-bit byte offsets against a resource of known size):
// loads past the size return zeros and stores past it are dropped.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr unsigned BUF_OOB = 0xFFFFE000u;  // offset past every buffer used here
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ inline bf16x8 ld_frag(__amdgpu_buffer_rsrc_t r, unsigned off, bool ok) {
    u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    u32x4 z = {0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf16x8, ok ? v : z);
}

namespace {

constexpr int C = 32, NT = 27, TB = 256;

// Parameter layout (FP32, flat): W1[27][32] b1[32] {Wl[27][32ci][32co] bl[32]} x4  W6[27][32][2] b6[2]
constexpr long OFF_W1 = 0, OFF_B1 = OFF_W1 + NT * C;
constexpr long L_SZ = (long)NT * C * C + C;
constexpr long OFF_L(int l) { return OFF_B1 + C + (long)(l - 2) * L_SZ; }  // l = 2..5
constexpr long OFF_W6 = OFF_B1 + C + 4 * L_SZ, OFF_B6 = OFF_W6 + NT * C * 2;
constexpr long NPARAM = OFF_B6 + 2;

__device__ inline void tap_offset(int t, int &d1, int &d2, int &d3) {
    d1 = t % 3 - 1;
    d2 = (t / 3) % 3 - 1;
    d3 = t / 9 - 1;
}

// ---------------------------------------------------------------------------
// Weight packing: master FP32 -> BF16 [27][32 co][32 ci] for the forward of
// layers 2..6 (layer 6 padded to 32 outputs with zeros) and the flipped,
// transposed [27][32 ci][32 co] for the input-gradient pass.
// ---------------------------------------------------------------------------
__global__ void pack_weights(const float *__restrict__ P, __bf16 *__restrict__ Wf, __bf16 *__restrict__ Wb) {
    // Wf, Wb: [5 layers][27][32][32]
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long per = (long)NT * C * C;
    if (i >= 5 * per) return;
    int l = (int)(i / per);  // 0..4 -> layers 2..6
    long r = i % per;
    int t = (int)(r / (C * C)), a = (int)((r / C) % C), b = (int)(r % C);
    // forward: Wf[l][t][co=a][ci=b] = W[t][ci=b][co=a]
    float wf, wb;
    if (l < 4) {
        const float *W = P + OFF_L(l + 2);
        wf = W[((long)t * C + b) * C + a];
        // backward: Wb[l][t'][co'=a (= ci)][ci'=b (= co)] = W[flip t'][ci=a][co=b]
        wb = W[((long)(NT - 1 - t) * C + a) * C + b];
    } else {
        const float *W = P + OFF_W6;  // [27][32 ci][2 co]
        wf = (a < 2) ? W[((long)t * C + b) * 2 + a] : 0.0f;
        wb = (b < 2) ? W[((long)(NT - 1 - t) * C + a) * 2 + b] : 0.0f;
    }
    Wf[i] = (__bf16)wf;
    Wb[i] = (__bf16)wb;
}

// ---------------------------------------------------------------------------
// Layer 1 (1 -> 32), FP32 input, ReLU, BF16 output.
// ---------------------------------------------------------------------------
// The weights and biases are read with wave-uniform indices from the
// read-only parameter vector, so they arrive through scalar loads and feed
// the FMAs as scalar operands (no LDS traffic per FMA).
__global__ void __launch_bounds__(TB) conv1_fwd(const float *__restrict__ x, float xscale, const float *__restrict__ P,
                                                __bf16 *__restrict__ out, int n1, int n2, int n3, int dil,
                                                FastDiv f1, FastDiv f2, const float *__restrict__ xsp) {
    if (xsp) xscale = xsp[0];  // the scale computed on the device
    const float *__restrict__ w = P + OFF_W1;
    const float *__restrict__ b = P + OFF_B1;
    long n = (long)n1 * n2 * n3;
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int i1, i2, i3;
    if (n < (1L << 31)) {  // multiply-high division (no long division per voxel)
        const unsigned r = f1.div((unsigned)v), q = f2.div(r);
        i1 = (int)((unsigned)v - r * (unsigned)n1);
        i2 = (int)(r - q * (unsigned)n2);
        i3 = (int)q;
    } else {
        i1 = (int)(v % n1);
        const long r = v / n1;
        i2 = (int)(r % n2);
        i3 = (int)(r / n2);
    }
    float xv[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        int d1 = (t % 3 - 1) * dil, d2 = ((t / 3) % 3 - 1) * dil, d3 = (t / 9 - 1) * dil;
        int j1 = i1 + d1, j2 = i2 + d2, j3 = i3 + d3;
        bool in = j1 >= 0 && j1 < n1 && j2 >= 0 && j2 < n2 && j3 >= 0 && j3 < n3;
        xv[t] = in ? x[v + d1 + (long)n1 * (d2 + (long)n2 * d3)] * xscale : 0.0f;
    }
    __bf16 o[C];
#pragma unroll
    for (int co = 0; co < C; co++) {
        float s = b[co];
#pragma unroll
        for (int t = 0; t < NT; t++) s = fmaf(w[t * C + co], xv[t], s);
        o[co] = (__bf16)fmaxf(s, 0.0f);
    }
    bf16x8 *dst = (bf16x8 *)(out + v * C);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        bf16x8 pk;
#pragma unroll
        for (int i = 0; i < 8; i++) pk[i] = o[q * 8 + i];
        dst[q] = pk;
    }
}

// conv1_fwd for dilation 1 with the input staged in LDS: a block owns a
// 64 (i1) x 4 (i2) tile and marches along i3 over a segment of planes,
// keeping the scaled input planes o-1, o, o+1 (with their one-sample halo,
// zero outside the volume) in a ring of three LDS planes, so each input
// sample is read from HBM about once instead of through 27 global loads
// per voxel.  Same operations per output as conv1_fwd.
constexpr int C1_W = 66, C1_H = 6, C1_PL = C1_W * C1_H;

__global__ void __launch_bounds__(TB) conv1_fwd_tile(const float *__restrict__ x, float xscale,
                                                     const float *__restrict__ P, __bf16 *__restrict__ out, int n1,
                                                     int n2, int n3, int seg, const float *__restrict__ xsp) {
    if (xsp) xscale = xsp[0];
    __shared__ float pl[3 * C1_PL];
    const float *__restrict__ w = P + OFF_W1;
    const float *__restrict__ b = P + OFF_B1;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    const int ntile1 = (n1 + 63) / 64;
    const int x0 = (blockIdx.x % ntile1) * 64, y0 = (blockIdx.x / ntile1) * 4;
    const int z0 = blockIdx.y * seg, z1 = min(n3, z0 + seg);
    const long n12 = (long)n1 * n2;
    const int i1 = x0 + tx, i2 = y0 + ty;
    for (int q = z0 - 1; q <= z1; q++) {
        // plane q into ring slot (q + 3) % 3
        float *dst = pl + ((q + 3) % 3) * C1_PL;
        for (int e = threadIdx.x; e < C1_PL; e += TB) {
            const int r = e / C1_W, c = e - r * C1_W;
            const int j1 = x0 - 1 + c, j2 = y0 - 1 + r;
            const bool in = q >= 0 && q < n3 && j1 >= 0 && j1 < n1 && j2 >= 0 && j2 < n2;
            dst[e] = in ? x[j1 + (long)n1 * j2 + n12 * q] * xscale : 0.0f;
        }
        __syncthreads();
        const int o = q - 1;
        if (o >= z0 && i1 < n1 && i2 < n2) {
            float xv[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int d1 = t % 3 - 1, d2 = (t / 3) % 3 - 1, d3 = t / 9 - 1;
                xv[t] = pl[((o + d3 + 3) % 3) * C1_PL + (ty + 1 + d2) * C1_W + tx + 1 + d1];
            }
            __bf16 ov[C];
#pragma unroll
            for (int co = 0; co < C; co++) {
                float s = b[co];
#pragma unroll
                for (int t = 0; t < NT; t++) s = fmaf(w[t * C + co], xv[t], s);
                ov[co] = (__bf16)fmaxf(s, 0.0f);
            }
            bf16x8 *dp = (bf16x8 *)(out + ((long)i1 + (long)n1 * i2 + n12 * o) * C);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                bf16x8 pk;
#pragma unroll
                for (int i = 0; i < 8; i++) pk[i] = ov[k * 8 + i];
                dp[k] = pk;
            }
        }
        __syncthreads();  // the slot of plane o - 1 is overwritten next
    }
}

// ---------------------------------------------------------------------------
// 32-channel 3x3x3 convolution on the matrix cores.
// Block = 4 waves, tile 64 (i1) x 4 (i2) voxels at one i3; wave w owns row
// i2 = y0 + w as two 32-voxel M-tiles.  Per tap and 16-channel half:
// A = input [32 voxels][16 ci] from global (zero outside the volume),
// B = W [16 ci][32 co] from LDS, D += A B in FP32.
// MODE 0: out = relu(D + b) BF16;  MODE 1: out[v][0..1] = D + b FP32;
// MODE 2: out = D * (aux > 0) BF16 (input gradient through the ReLU).
// ---------------------------------------------------------------------------
constexpr int CW = 8;          // waves per conv block (one i2 row of the tile each)
constexpr int CTB = 64 * CW;   // threads per conv block

// B operands in LDS: 16-byte entries [tap][h][hi][co] (h: which 8 of the
// lane's 16 input channels, hi = lane / 32, co = lane % 32), so the 64
// lanes of one fragment read cover 1 KB contiguously and hit distinct
// banks.  (Entries in [tap][co][ci] order put lanes co, co + 4, ... 64 B
// apart on the same banks; at 512x256x256 the layers ran equally fast
// either way, so bank conflicts are not what bounds them.)
__device__ inline void stage_weights(bf16x8 *wl, const __bf16 *__restrict__ Wg) {
    for (int i = threadIdx.x; i < NT * C * C / 8; i += blockDim.x) {
        const int co = i % 32, hi = (i / 32) % 2, h = (i / 64) % 2, t = i / 128;
        wl[i] = ((const bf16x8 *)Wg)[((t * C + co) * C + hi * 16 + h * 8) / 8];
    }
}
__device__ inline int wl_at(int t, int h, int hi, int co) { return ((t * 2 + h) * 2 + hi) * 32 + co; }

// A fragments of one (d2, d3) row of taps: [d1][m][half]
struct RowFrag {
    bf16x8 a[3][2][2];
};

__device__ inline void load_row(RowFrag &f, const __bf16 *__restrict__ in, bool plane, long rowbase, int x0,
                                int lane, int hi, int n1, int dil) {
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int m = 0; m < 2; m++) {
            int j1 = x0 + m * 32 + (lane & 31) + (d - 1) * dil;
            if (plane && j1 >= 0 && j1 < n1) {
                const bf16x8 *src = (const bf16x8 *)(in + (rowbase + j1) * C + hi * 16);
                f.a[d][m][0] = src[0];
                f.a[d][m][1] = src[1];
            } else {
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    f.a[d][m][0][i] = (__bf16)0.0f;
                    f.a[d][m][1][i] = (__bf16)0.0f;
                }
            }
        }
}

template <int MODE>
__global__ void __launch_bounds__(CTB)
conv32(const __bf16 *__restrict__ in, const __bf16 *__restrict__ Wg, const float *__restrict__ bias,
       void *__restrict__ outp, const __bf16 *__restrict__ aux, int n1, int n2, int n3, int dil) {
    __shared__ bf16x8 wl[NT * C * C / 8];  // [27][h][hi][32 co] entries of 8 ci
    stage_weights(wl, Wg);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntx = (n1 + 63) / 64, nty = (n2 + CW - 1) / CW;
    const long ntiles = (long)ntx * nty * n3;
    const long n12 = (long)n1 * n2;
    // tiles ordered column-major in i3 (item = column * n3 + i3); each block
    // walks a contiguous range, so consecutive tiles of a block share two of
    // their three input planes in its XCD's L2
    const long per = (ntiles + gridDim.x - 1) / gridDim.x;
    const long t_begin = (long)blockIdx.x * per, t_end = min(ntiles, t_begin + per);
    // K order of the two MFMAs of a tap: lane half hi holds input channels
    // hi*16 + [0, 8) for the first and hi*16 + [8, 16) for the second, so
    // each lane reads its 16 channels (32 bytes) of a voxel in one go.
    const int co = lane & 31, hi = lane >> 5;
    const float bco = (MODE == 2) ? 0.0f : bias[co];
    // the weights stay in LDS while the block walks over tiles; the A
    // fragments of the next (d2, d3) row of taps are loaded while the
    // current row's 12 MFMAs run
    for (long item = t_begin; item < t_end; item++) {
        const long col = item / n3;
        const int z = (int)(item % n3);
        const int x0 = (int)(col % ntx) * 64;
        const int y = (int)(col / ntx) * CW + w;
        f32x16 acc[2];
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[m][r] = 0.0f;
        const bool rowok = y < n2;
        RowFrag fa, fb;
        auto load = [&](RowFrag &f, int row) {
            int d2 = (row % 3 - 1) * dil, d3 = (row / 3 - 1) * dil;
            int j2 = y + d2, j3 = z + d3;
            bool plane = rowok && j2 >= 0 && j2 < n2 && j3 >= 0 && j3 < n3;
            load_row(f, in, plane, (long)j2 * n1 + (long)j3 * n12, x0, lane, hi, n1, dil);
        };
        auto compute = [&](const RowFrag &f, int row) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int t = row * 3 + d;
                const bf16x8 b0 = wl[wl_at(t, 0, hi, co)];
                const bf16x8 b1 = wl[wl_at(t, 1, hi, co)];
                // independent accumulators back to back (no dependent MFMA
                // pair); each accumulator still sums half 0 then half 1
#pragma unroll
                for (int m = 0; m < 2; m++)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[d][m][0], b0, acc[m], 0, 0, 0);
#pragma unroll
                for (int m = 0; m < 2; m++)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[d][m][1], b1, acc[m], 0, 0, 0);
            }
        };
        load(fa, 0);
#pragma unroll 1
        for (int row = 0; row < 8; row += 2) {
            load(fb, row + 1);
            compute(fa, row);
            load(fa, row + 2);
            compute(fb, row + 1);
        }
        compute(fa, 8);
        if (!rowok) continue;
#pragma unroll
        for (int m = 0; m < 2; m++) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                int rr = (r / 4) * 8 + hi * 4 + (r % 4);
                int i1 = x0 + m * 32 + rr;
                if (i1 >= n1) continue;
                long v = (long)i1 + (long)y * n1 + (long)z * n12;
                float s = acc[m][r] + bco;
                if (MODE == 0) {
                    ((__bf16 *)outp)[v * C + co] = (__bf16)fmaxf(s, 0.0f);
                } else if (MODE == 1) {
                    if (co < 2) ((float *)outp)[v * 2 + co] = s;
                } else {
                    float gate = (float)aux[v * C + co] > 0.0f ? 1.0f : 0.0f;
                    ((__bf16 *)outp)[v * C + co] = (__bf16)(s * gate);
                }
            }
        }
    }
}

// Wide-tile variant for small dilations: lane l of M-tile s holds voxel
// x0 + 4l + s (four interleaved M-tiles, 128 voxels along i1).  The input
// fragment of tile s and tap offset d1 is then the voxel group s + d1*dil,
// so one (d2, d3) row of taps needs only 4 + 2*dil distinct A fragments
// for its 12 MFMA pairs (against 12 in conv32): half the operand loads for
// dil = 1.  The accumulation order per output is that of conv32, so both
// give identical results.
template <int DIL>
struct RowFragW {
    bf16x8 a[4 + 2 * DIL][2];
};

template <int DIL>
__device__ inline void load_row_w(RowFragW<DIL> &f, const __bf16 *__restrict__ in, bool plane, long rowbase, int x0,
                                  int lane, int hi, int n1) {
#pragma unroll
    for (int g = 0; g < 4 + 2 * DIL; g++) {
        int j1 = x0 + 4 * (lane & 31) + g - DIL;
        if (plane && j1 >= 0 && j1 < n1) {
            const bf16x8 *src = (const bf16x8 *)(in + (rowbase + j1) * C + hi * 16);
            f.a[g][0] = src[0];
            f.a[g][1] = src[1];
        } else {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                f.a[g][0][i] = (__bf16)0.0f;
                f.a[g][1][i] = (__bf16)0.0f;
            }
        }
    }
}

template <int MODE, int DIL>
__global__ void __launch_bounds__(CTB)
conv32w(const __bf16 *__restrict__ in, const __bf16 *__restrict__ Wg, const float *__restrict__ bias,
        void *__restrict__ outp, const __bf16 *__restrict__ aux, int n1, int n2, int n3) {
    __shared__ bf16x8 wl[NT * C * C / 8];  // [27][h][hi][32 co] entries of 8 ci
    stage_weights(wl, Wg);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntx = (n1 + 127) / 128, nty = (n2 + CW - 1) / CW;
    const long ntiles = (long)ntx * nty * n3;
    const long n12 = (long)n1 * n2;
    const long per = (ntiles + gridDim.x - 1) / gridDim.x;
    const long t_begin = (long)blockIdx.x * per, t_end = min(ntiles, t_begin + per);
    const int co = lane & 31, hi = lane >> 5;
    const float bco = (MODE == 2) ? 0.0f : bias[co];
    for (long item = t_begin; item < t_end; item++) {
        const long col = item / n3;
        const int z = (int)(item % n3);
        const int x0 = (int)(col % ntx) * 128;
        const int y = (int)(col / ntx) * CW + w;
        f32x16 acc[4];
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[m][r] = 0.0f;
        const bool rowok = y < n2;
        RowFragW<DIL> fa, fb;
        auto load = [&](RowFragW<DIL> &f, int row) {
            int d2 = (row % 3 - 1) * DIL, d3 = (row / 3 - 1) * DIL;
            int j2 = y + d2, j3 = z + d3;
            bool plane = rowok && j2 >= 0 && j2 < n2 && j3 >= 0 && j3 < n3;
            load_row_w<DIL>(f, in, plane, (long)j2 * n1 + (long)j3 * n12, x0, lane, hi, n1);
        };
        auto compute = [&](const RowFragW<DIL> &f, int row) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int t = row * 3 + d;
                const bf16x8 b0 = wl[wl_at(t, 0, hi, co)];
                const bf16x8 b1 = wl[wl_at(t, 1, hi, co)];
#pragma unroll
                for (int m = 0; m < 4; m++)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[m + (d - 1) * DIL + DIL][0], b0, acc[m], 0, 0, 0);
#pragma unroll
                for (int m = 0; m < 4; m++)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[m + (d - 1) * DIL + DIL][1], b1, acc[m], 0, 0, 0);
            }
        };
        load(fa, 0);
#pragma unroll 1
        for (int row = 0; row < 8; row += 2) {
            load(fb, row + 1);
            compute(fa, row);
            load(fa, row + 2);
            compute(fb, row + 1);
        }
        compute(fa, 8);
        if (!rowok) continue;
#pragma unroll
        for (int m = 0; m < 4; m++) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                int rr = (r / 4) * 8 + hi * 4 + (r % 4);
                int i1 = x0 + 4 * rr + m;
                if (i1 >= n1) continue;
                long v = (long)i1 + (long)y * n1 + (long)z * n12;
                float sv = acc[m][r] + bco;
                if (MODE == 0) {
                    ((__bf16 *)outp)[v * C + co] = (__bf16)fmaxf(sv, 0.0f);
                } else if (MODE == 1) {
                    if (co < 2) ((float *)outp)[v * 2 + co] = sv;
                } else {
                    float gate = (float)aux[v * C + co] > 0.0f ? 1.0f : 0.0f;
                    ((__bf16 *)outp)[v * C + co] = (__bf16)(sv * gate);
                }
            }
        }
    }
}

// Row of A fragments through a buffer resource: o1 = byte offset of voxel
// x0 + (lane & 31) of the row (+ the lane's channel half), BUF_OOB for a
// row outside the volume (the hardware returns zeros); xm bit m*3+d clears
// fragment (d, m) whose voxel lies outside [0, n1) along i1.
__device__ inline void load_row_b(RowFrag &f, __amdgpu_buffer_rsrc_t r, unsigned o1, unsigned xm) {
    const unsigned o0 = o1 - 64u;
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const unsigned o = (d == 0 ? o0 : o1) + m * 2048u + (d == 2 ? 64u : 0u);
            const bool ok = (xm >> (m * 3 + d)) & 1u;
            f.a[d][m][0] = ld_frag(r, o, ok);
            f.a[d][m][1] = ld_frag(r, o + 16u, ok);
        }
}

// Plane-scatter variant for dilation 1.  The block walks its z-range one
// INPUT plane at a time: the three (d2) rows of plane p are loaded once
// and feed the taps of the three outputs that read them, z = p + 1
// (d3 = -1), z = p (d3 = 0) and z = p - 1 (d3 = +1), held in three rolling
// accumulators; after plane p, output p - 1 is complete and is stored.
// Each input row is thus read once per block and z-range instead of three
// times, with 36 MFMAs per row load (12 in conv32).  Every output still
// sums its taps in conv32's order (rows (d3, d2) ascending, then d1, then
// the channel half), so both give identical results.  Rows are read and
// outputs written through buffer resources (one lane offset per row or
// plane plus immediates), which keeps the kernel within 256 VGPRs without
// spills; volumes need n * 64 bytes < 4e9.
template <int MODE>
__global__ void __launch_bounds__(CTB)
conv32z(const __bf16 *__restrict__ in, const __bf16 *__restrict__ Wg, const float *__restrict__ bias,
        void *__restrict__ outp, const __bf16 *__restrict__ aux, int n1, int n2, int n3) {
    __shared__ bf16x8 wl[NT * C * C / 8];  // [27][h][hi][32 co] entries of 8 ci
    stage_weights(wl, Wg);
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntx = (n1 + 63) / 64, nty = (n2 + CW - 1) / CW;
    const long ntiles = (long)ntx * nty * n3;
    const long n12 = (long)n1 * n2;
    const long per = (ntiles + gridDim.x - 1) / gridDim.x;
    const long t_begin = (long)blockIdx.x * per, t_end = min(ntiles, t_begin + per);
    const int co = lane & 31, hi = lane >> 5;
    const float bco = (MODE == 2) ? 0.0f : bias[co];
    const unsigned nvox = (unsigned)((long)n12 * n3);
    const __amdgpu_buffer_rsrc_t rin = buf_rsrc(in, nvox * 64u);
    const __amdgpu_buffer_rsrc_t rout = buf_rsrc(outp, nvox * (MODE == 1 ? 8u : 64u));
    const __amdgpu_buffer_rsrc_t raux = buf_rsrc(MODE == 2 ? (const void *)aux : (const void *)in, nvox * 64u);
    long item = t_begin;
    while (item < t_end) {
        // one segment: outputs z in [za, zb) of one column
        const long col = item / n3;
        const int za = (int)(item % n3);
        const int zb = (int)min<long>((long)n3, za + (t_end - item));
        item += zb - za;
        const int x0 = (int)(col % ntx) * 64;
        const int y = (int)(col / ntx) * CW + w;
        const bool rowok = y < n2;
        f32x16 alo[2], amid[2], ahi[2];  // outputs p - 1, p, p + 1
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                alo[m][r] = 0.0f;
                amid[m][r] = 0.0f;
                ahi[m][r] = 0.0f;
            }
        unsigned xm = 0;
#pragma unroll
        for (int m = 0; m < 2; m++)
#pragma unroll
            for (int d = 0; d < 3; d++)
                if ((unsigned)(x0 + m * 32 + (lane & 31) + d - 1) < (unsigned)n1) xm |= 1u << (m * 3 + d);
        const unsigned lb = (unsigned)(x0 + (lane & 31)) * 64u + (unsigned)hi * 32u;
        auto load = [&](RowFrag &f, int p, int r2) {
            int j2 = y + r2 - 1;
            bool plane = rowok && j2 >= 0 && j2 < n2 && p >= 0 && p < n3;
            unsigned o1 = plane ? lb + ((unsigned)j2 * (unsigned)n1 + (unsigned)p * (unsigned)n12) * 64u : BUF_OOB;
            load_row_b(f, rin, o1, xm);
        };
        // taps of row (d3, d2) = (row / 3 - 1, row % 3 - 1) into acc
        auto taps = [&](const RowFrag &f, int row, f32x16 *acc) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int t = row * 3 + d;
                const bf16x8 b0 = wl[wl_at(t, 0, hi, co)];
                const bf16x8 b1 = wl[wl_at(t, 1, hi, co)];
#pragma unroll
                for (int m = 0; m < 2; m++)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[d][m][0], b0, acc[m], 0, 0, 0);
#pragma unroll
                for (int m = 0; m < 2; m++)
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[d][m][1], b1, acc[m], 0, 0, 0);
            }
        };
        auto compute = [&](const RowFrag &f, int r2, bool dlo, bool dmid, bool dhi) {
            if (dhi) taps(f, r2, ahi);        // d3 = -1
            if (dmid) taps(f, 3 + r2, amid);  // d3 = 0
            if (dlo) taps(f, 6 + r2, alo);    // d3 = +1
        };
        // stores through buffer resources: one lane offset per plane plus
        // an immediate per element; the i1 tail tile checks each element
        const bool xfull = x0 + 64 <= n1;
        auto store = [&](const f32x16 *acc, int z) {
            if (!rowok) return;
            const unsigned vb = (unsigned)(x0 + hi * 4) + (unsigned)y * (unsigned)n1 + (unsigned)z * (unsigned)n12;
#pragma unroll
            for (int m = 0; m < 2; m++) {
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const unsigned e = (unsigned)(m * 32 + (r / 4) * 8 + (r % 4));
                    const bool ok = xfull || x0 + m * 32 + (r / 4) * 8 + hi * 4 + (r % 4) < n1;
                    float sv = acc[m][r] + bco;
                    if (MODE == 1) {
                        if (co < 2)
                            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, sv), rout,
                                                                  ok ? (vb + e) * 8u + co * 4u : BUF_OOB, 0, 0);
                    } else {
                        const unsigned o = ok ? (vb + e) * 64u + co * 2u : BUF_OOB;
                        if (MODE == 2) {
                            unsigned short g = __builtin_amdgcn_raw_buffer_load_b16(raux, o, 0, 0);
                            sv = (float)__builtin_bit_cast(__bf16, g) > 0.0f ? sv : sv * 0.0f;
                        } else {
                            sv = fmaxf(sv, 0.0f);
                        }
                        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (__bf16)sv), rout, o, 0,
                                                              0);
                    }
                }
            }
        };
        // one input plane; f0 holds its row 0 on entry, f1 holds the next
        // plane's row 0 on exit
        auto plane = [&](RowFrag &f0, RowFrag &f1, int p) {
            const bool dlo = p - 1 >= za, dmid = p >= za && p < zb, dhi = p + 1 < zb;
            load(f1, p, 1);
            compute(f0, 0, dlo, dmid, dhi);
            load(f0, p, 2);
            compute(f1, 1, dlo, dmid, dhi);
            load(f1, p + 1, 0);
            compute(f0, 2, dlo, dmid, dhi);
            if (dlo) store(alo, p - 1);
#pragma unroll
            for (int m = 0; m < 2; m++) {
                alo[m] = amid[m];
                amid[m] = ahi[m];
#pragma unroll
                for (int r = 0; r < 16; r++) ahi[m][r] = 0.0f;
            }
        };
        RowFrag fa, fb;
        load(fa, za - 1, 0);
#pragma unroll 1
        for (int p = za - 1; p <= zb; p += 2) {
            plane(fa, fb, p);
            if (p + 1 > zb) break;
            plane(fb, fa, p + 1);
        }
    }
}

// Input gradient of layer 1's output when layer 2 is the consumer is given
// by conv32<2>; layer 1 needs no input gradient.

// ---------------------------------------------------------------------------
// Weight gradient of a 32-channel layer: dW[t][ci][co] = sum_v in[v+t][ci]
// g[v][co] over the voxels of each block (64 (i1) x 4 (i2) at one i3), on
// the matrix cores with K = voxels.  The block stages its input tile with
// halo ([ci][voxel] transposed) and the output gradient ([co][voxel]) in
// LDS; wave w accumulates taps t = w, w+4, ... and writes its partial
// [27][32][32] sums to part[block]; db partials go to pb[block][32].
// ---------------------------------------------------------------------------
constexpr int WG_X = 64, WG_Y = 4;
constexpr int MAX_DIL = 8;
constexpr int WG_HXM = WG_X + 2 * MAX_DIL, WG_HYM = WG_Y + 2 * MAX_DIL;

// The input is staged one tap plane (d3) at a time, with an in-plane halo
// of `dil` voxels; the 9 taps of a plane are split over the 4 waves.
__global__ void __launch_bounds__(TB)
conv32_wgrad(const __bf16 *__restrict__ in, const __bf16 *__restrict__ g, float *__restrict__ part,
             float *__restrict__ pb, int n1, int n2, int n3, int dil) {
    // input plane [ci][y HY][x HX], gradient tile [co][y][x]
    __shared__ __bf16 ti[C * WG_HYM * WG_HXM];
    __shared__ __bf16 tg[C * WG_Y * WG_X];
    const int HX = WG_X + 2 * dil, HY = WG_Y + 2 * dil;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntx = (n1 + WG_X - 1) / WG_X, nty = (n2 + WG_Y - 1) / WG_Y;
    long bid = blockIdx.x;
    const int x0 = (int)(bid % ntx) * WG_X;
    bid /= ntx;
    const int y0 = (int)(bid % nty) * WG_Y;
    const int z = (int)(bid / nty);
    const long n12 = (long)n1 * n2;
    for (int i = threadIdx.x; i < WG_Y * WG_X; i += blockDim.x) {
        int xx = i % WG_X, yy = i / WG_X;
        int j1 = x0 + xx, j2 = y0 + yy;
        bool ok = j1 < n1 && j2 < n2;
        const __bf16 *src = g + ((long)j1 + (long)j2 * n1 + (long)z * n12) * C;
        for (int c = 0; c < C; c++) tg[c * (WG_Y * WG_X) + i] = ok ? src[c] : (__bf16)0.0f;
    }
    const int kq = (lane >> 5) * 8;
    for (int zz = 0; zz < 3; zz++) {
        const int j3 = z + (zz - 1) * dil;
        __syncthreads();  // the previous plane's readers are done
        for (int i = threadIdx.x; i < HY * HX; i += blockDim.x) {
            int xx = i % HX, yy = i / HX;
            int j1 = x0 - dil + xx, j2 = y0 - dil + yy;
            bool ok = j1 >= 0 && j1 < n1 && j2 >= 0 && j2 < n2 && j3 >= 0 && j3 < n3;
            const __bf16 *src = in + ((long)j1 + (long)j2 * n1 + (long)j3 * n12) * C;
            for (int c = 0; c < C; c++) ti[c * (HY * HX) + i] = ok ? src[c] : (__bf16)0.0f;
        }
        __syncthreads();
        for (int tt = w; tt < 9; tt += 4) {
            const int t = zz * 9 + tt;
            const int d1 = (tt % 3 - 1) * dil, d2 = (tt / 3 - 1) * dil;
            f32x16 acc;
#pragma unroll
            for (int r = 0; r < 16; r++) acc[r] = 0.0f;
            // K = voxels, 16 per MFMA: rows yy, x chunks of 16
            for (int yy = 0; yy < WG_Y; yy++) {
                for (int xc = 0; xc < WG_X; xc += 16) {
                    // A[row = ci][k = voxel]: lane -> ci = lane & 31, voxels xc + kq .. +7
                    bf16x8 a, b;
                    const __bf16 *ai = ti + (lane & 31) * (HY * HX) + (yy + d2 + dil) * HX + (xc + kq + d1 + dil);
                    const __bf16 *bg = tg + (lane & 31) * (WG_Y * WG_X) + yy * WG_X + xc + kq;
#pragma unroll
                    for (int i = 0; i < 8; i++) {
                        a[i] = ai[i];
                        b[i] = bg[i];
                    }
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
                }
            }
            // D[row = ci][col = co]
            float *dst = part + (size_t)blockIdx.x * NT * C * C + (size_t)t * C * C;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                int ci = (r / 4) * 8 + (lane >> 5) * 4 + (r % 4);
                dst[ci * C + (lane & 31)] = acc[r];
            }
        }
    }
    // bias partials: sum of g over the tile per channel
    if (threadIdx.x < C) {
        float s = 0.0f;
        for (int i = 0; i < WG_Y * WG_X; i++) s += (float)tg[threadIdx.x * (WG_Y * WG_X) + i];
        pb[(size_t)blockIdx.x * C + threadIdx.x] = s;
    }
}

// Layer 1 weight gradient: dW1[t][co] = sum_v x[v+t] g1[v][co]; block
// partials over 256 voxels; thread (t, co) pairs.
__global__ void __launch_bounds__(1024)
conv1_wgrad(const float *__restrict__ x, float xscale, const __bf16 *__restrict__ g, float *__restrict__ part, int n1,
            int n2, int n3, int vox_per_block, int dil) {
    long n = (long)n1 * n2 * n3;
    long v0 = (long)blockIdx.x * vox_per_block;
    int tid = threadIdx.x;
    float s = 0.0f, sb = 0.0f;
    int t = tid / C, co = tid % C;
    if (tid < NT * C) {
        int d1 = (t % 3 - 1) * dil, d2 = ((t / 3) % 3 - 1) * dil, d3 = (t / 9 - 1) * dil;
        for (long v = v0; v < std::min(n, v0 + vox_per_block); v++) {
            int i1 = (int)(v % n1);
            long r = v / n1;
            int i2 = (int)(r % n2), i3 = (int)(r / n2);
            int j1 = i1 + d1, j2 = i2 + d2, j3 = i3 + d3;
            bool ok = j1 >= 0 && j1 < n1 && j2 >= 0 && j2 < n2 && j3 >= 0 && j3 < n3;
            float gv = (float)g[v * C + co];
            if (ok) s += (x[v + d1 + (long)n1 * (d2 + (long)n2 * d3)] * xscale) * gv;
            if (t == 0) sb += gv;
        }
        part[(size_t)blockIdx.x * (NT * C + C) + tid] = s;
        if (t == 0) part[(size_t)blockIdx.x * (NT * C + C) + NT * C + co] = sb;
    }
}

// Sum block partials (fixed order) and scatter into the gradient vector.
// mode 0: 32-channel layer: part [nblk][27][32 ci][32 co] (+ pb [nblk][32])
//         -> G[off + (t*32+ci)*32+co], G[off + 27*1024 + co] (bias)
// mode 1: final layer (co < 2): -> G[OFF_W6 + (t*32+ci)*2+co], G[OFF_B6+co]
// mode 2: layer 1: part [nblk][27*32 + 32] -> G[OFF_W1..], G[OFF_B1..]
__global__ void reduce_wgrad(const float *__restrict__ part, const float *__restrict__ pb, int nblk, int mode, long off,
                             float *__restrict__ G) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 2) {
        if (i >= NT * C + C) return;
        float s = 0.0f;
        for (int b = 0; b < nblk; b++) s += part[(size_t)b * (NT * C + C) + i];
        G[(i < NT * C ? OFF_W1 + i : OFF_B1 + (i - NT * C))] = s;
        return;
    }
    long per = (long)NT * C * C;
    if (i < per) {
        int co = (int)(i % C);
        if (mode == 1 && co >= 2) return;
        float s = 0.0f;
        for (int b = 0; b < nblk; b++) s += part[(size_t)b * per + i];
        if (mode == 0) G[off + i] = s;
        else G[OFF_W6 + (i / C) * 2 + co] = s;
    } else if (i < per + C) {
        int co = (int)(i - per);
        if (mode == 1 && co >= 2) return;
        float s = 0.0f;
        for (int b = 0; b < nblk; b++) s += pb[(size_t)b * C + co];
        if (mode == 0) G[off + per + co] = s;
        else G[OFF_B6 + co] = s;
    }
}

// Loss gradient: gy = 2 (y - t) m / count (BF16, padded to 32 channels),
// partial sums of (y - t)^2 m.
__global__ void __launch_bounds__(TB)
loss_grad(const float *__restrict__ y, const float *__restrict__ tgt, const float *__restrict__ wt,
          __bf16 *__restrict__ gy, long n, float scale, double *part) {
    __shared__ double sh[16];
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (v < n) {
        float m = wt ? wt[v] : 1.0f;
        __bf16 *dst = gy + v * C;
        for (int c = 0; c < 2; c++) {
            float d = y[v * 2 + c] - tgt[c * n + v];
            dst[c] = (__bf16)(2.0f * d * m * scale);
            acc += (double)d * d * m;
        }
        for (int c = 2; c < C; c++) dst[c] = (__bf16)0.0f;
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Adam on the FP32 master parameters.
__global__ void adam_kernel(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M,
                            float *__restrict__ V, long n, float lr, float b1, float b2, float eps, float c1, float c2,
                            float wd) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = G[i] + wd * P[i];
    float m = b1 * M[i] + (1.0f - b1) * g;
    float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// 1 / RMS of x on the device (1 when x is all zero), in two launches: block
// partial sums of x^2 in double, then one block adds them in a fixed order.
__global__ void __launch_bounds__(TB) sumsq_partials(const float *__restrict__ x, long n, double *__restrict__ part) {
    __shared__ double sh[16];
    double s = 0.0;
    for (long i = (long)blockIdx.x * TB + threadIdx.x; i < n; i += (long)gridDim.x * TB) {
        const double v = x[i];
        s += v * v;
    }
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(1024) inv_rms(const double *__restrict__ part, int nblk, long n,
                                                float *__restrict__ xs) {
    __shared__ double sh[16];
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += part[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) {
        const double rms = sqrt(t / (double)n);
        xs[0] = rms > 0.0 ? (float)(1.0 / rms) : 1.0f;
    }
}

// (y interleaved [v][2]) -> dips [2][v]
__global__ void split_dips(const float *__restrict__ y, float *__restrict__ out, long n) {
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    out[v] = y[v * 2];
    out[n + v] = y[v * 2 + 1];
}

struct CnnWork {
    __bf16 *Wf, *Wb;    // packed weights [5][27][32][32]
    __bf16 *act[6];     // layer outputs 1..5 (act[1..5]); act[0] unused
    float *y;           // [V][2]
    __bf16 *gA, *gB;    // gradient ping-pong [V][32]
    float *part, *pb;
    double *red, *dsum;
    long nblk_wg;
};

int cnn_work(sh_ctx *ctx, long n, bool train, CnnWork &W, int n1, int n2, int n3) {
    size_t wsz = (size_t)5 * NT * C * C * sizeof(__bf16);
    char *b = nullptr;
    size_t actsz = (size_t)n * C * sizeof(__bf16);
    size_t ysz = ((size_t)n * 2 * sizeof(float) + 255) / 256 * 256;  // keeps later buffers 16-byte aligned
    int nact = train ? 5 : 2;  // inference ping-pongs two buffers
    size_t need = 2 * wsz + (size_t)nact * actsz + ysz + (train ? 2 * actsz : 0);
    SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 12, need, (void **)&b));
    W.Wf = (__bf16 *)b;
    W.Wb = (__bf16 *)(b + wsz);
    char *p = b + 2 * wsz;
    for (int l = 1; l <= 5; l++) W.act[l] = nullptr;
    if (train) {
        for (int l = 1; l <= 5; l++) { W.act[l] = (__bf16 *)p; p += actsz; }
    } else {
        W.act[1] = (__bf16 *)p; p += actsz;
        W.act[2] = (__bf16 *)p; p += actsz;
    }
    W.y = (float *)p; p += ysz;
    if (train) {
        W.gA = (__bf16 *)p; p += actsz;
        W.gB = (__bf16 *)p; p += actsz;
        long nb = (long)ceil_div(n1, WG_X) * ceil_div(n2, WG_Y) * n3;
        W.nblk_wg = nb;
        SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 13, (size_t)nb * (NT * C * C + C) * sizeof(float), (void **)&W.part));
        W.pb = W.part + (size_t)nb * NT * C * C;
        SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 14, (size_t)ceil_div(n, TB) * sizeof(double) + 64, (void **)&W.red));
        W.dsum = W.red + ceil_div(n, TB);
    }
    return SH_OK;
}

int conv_launch(sh_ctx *ctx, int mode, const __bf16 *in, const __bf16 *Wg, const float *bias, void *out,
                const __bf16 *aux, int n1, int n2, int n3, int dil, const char *name) {
    // persistent blocks: the 55 KB weight tile is staged once per block.
    // SH_CNN_CONV selects the dilation-1 kernel: z (default) plane-scatter,
    // w wide tiles, n the plain one (identical results, for timing)
    const char *cv = getenv("SH_CNN_CONV");
    const char kind = cv && cv[0] ? cv[0] : 'z';
    // (the plane-scatter kernel addresses its buffers with 32-bit offsets)
    if (dil == 1 && kind == 'z' && (double)n1 * n2 * n3 * 64.0 < 4.0e9) {
        long ntiles = (long)ceil_div(n1, 64) * ceil_div(n2, CW) * n3;
        dim3 grid((unsigned)std::min<long>(ntiles, 1024));
        PROF_A(ctx);
        if (mode == 0)
            hipLaunchKernelGGL((conv32z<0>), grid, dim3(CTB), 0, ctx->stream, in, Wg, bias, out, aux, n1, n2, n3);
        else if (mode == 1)
            hipLaunchKernelGGL((conv32z<1>), grid, dim3(CTB), 0, ctx->stream, in, Wg, bias, out, aux, n1, n2, n3);
        else
            hipLaunchKernelGGL((conv32z<2>), grid, dim3(CTB), 0, ctx->stream, in, Wg, bias, out, aux, n1, n2, n3);
        double v = (double)n1 * n2 * n3;
        PROF_B(ctx, name, v * (64.0 + (mode == 1 ? 8.0 : 64.0) + (mode == 2 ? 64.0 : 0.0)));
        return sh_launch_check(ctx, name);
    }
    if (dil == 1 && kind == 'w') {  // wide tiles (half the operand loads)
        long ntiles = (long)ceil_div(n1, 128) * ceil_div(n2, CW) * n3;
        dim3 grid((unsigned)std::min<long>(ntiles, 1024));
        PROF_A(ctx);
#define CW_L(MM) hipLaunchKernelGGL((conv32w<MM, 1>), grid, dim3(CTB), 0, ctx->stream, in, Wg, bias, out, aux, n1, n2, n3)
        if (mode == 0) CW_L(0); else if (mode == 1) CW_L(1); else CW_L(2);
#undef CW_L
        double v = (double)n1 * n2 * n3;
        PROF_B(ctx, name, v * (64.0 + (mode == 1 ? 8.0 : 64.0) + (mode == 2 ? 64.0 : 0.0)));
        return sh_launch_check(ctx, name);
    }
    long ntiles = (long)ceil_div(n1, 64) * ceil_div(n2, CW) * n3;
    dim3 grid((unsigned)std::min<long>(ntiles, 1024));
    PROF_A(ctx);
    if (mode == 0)
        hipLaunchKernelGGL((conv32<0>), grid, dim3(CTB), 0, ctx->stream, in, Wg, bias, out, aux, n1, n2, n3, dil);
    else if (mode == 1)
        hipLaunchKernelGGL((conv32<1>), grid, dim3(CTB), 0, ctx->stream, in, Wg, bias, out, aux, n1, n2, n3, dil);
    else
        hipLaunchKernelGGL((conv32<2>), grid, dim3(CTB), 0, ctx->stream, in, Wg, bias, out, aux, n1, n2, n3, dil);
    double v = (double)n1 * n2 * n3;
    PROF_B(ctx, name, v * (64.0 + (mode == 1 ? 8.0 : 64.0) + (mode == 2 ? 64.0 : 0.0)));
    return sh_launch_check(ctx, name);
}

int cnn_forward(sh_ctx *ctx, const float *P, const float *x, float xscale, const int *D, int n1, int n2, int n3,
                CnnWork &W, bool train, const float *xsp = nullptr) {
    long n = (long)n1 * n2 * n3;
    long per = (long)NT * C * C;
    hipLaunchKernelGGL(pack_weights, dim3(ceil_div(5 * per, TB)), dim3(TB), 0, ctx->stream, P, W.Wf, W.Wb);
    SH_TRY(sh_launch_check(ctx, "pack_weights"));
    PROF_A(ctx);
    // SH_CNN_C1=g: the one-voxel-per-thread kernel with 27 global loads per
    // voxel instead of the LDS-staged one (identical results)
    const char *c1e = getenv("SH_CNN_C1");
    if (D[0] == 1 && !(c1e && c1e[0] == 'g')) {
        const long tiles = (long)ceil_div(n1, 64) * ceil_div(n2, 4);
        const int nseg = (int)std::max<long>(1, std::min<long>(n3 / 8, (2048 + tiles - 1) / tiles));
        const int seg = ceil_div(n3, nseg);
        hipLaunchKernelGGL(conv1_fwd_tile, dim3((unsigned)tiles, ceil_div(n3, seg)), dim3(TB), 0, ctx->stream, x, xscale,
                           P, W.act[1], n1, n2, n3, seg, xsp);
    } else {
        hipLaunchKernelGGL(conv1_fwd, dim3(ceil_div(n, TB)), dim3(TB), 0, ctx->stream, x, xscale, P, W.act[1], n1, n2,
                           n3, D[0], FastDiv((unsigned)n1), FastDiv((unsigned)n2), xsp);
    }
    PROF_B(ctx, "conv1_fwd", (double)n * (4.0 + 64.0));
    SH_TRY(sh_launch_check(ctx, "conv1_fwd"));
    __bf16 *cur = W.act[1];
    for (int l = 2; l <= 5; l++) {
        __bf16 *dst = train ? W.act[l] : (cur == W.act[1] ? W.act[2] : W.act[1]);
        SH_TRY(conv_launch(ctx, 0, cur, W.Wf + (size_t)(l - 2) * per, P + OFF_L(l) + per, dst, nullptr, n1, n2, n3,
                           D[l - 1], "conv32<relu>"));
        cur = dst;
    }
    return conv_launch(ctx, 1, cur, W.Wf + (size_t)4 * per, P + OFF_B6, W.y, nullptr, n1, n2, n3, D[5],
                       "conv32<final>");
}

// Per-layer dilations (NULL = all 1), each in [1, MAX_DIL].
int read_dil(sh_ctx *ctx, const int *dil, int *D) {
    for (int l = 0; l < 6; l++) {
        D[l] = dil ? dil[l] : 1;
        if (D[l] < 1 || D[l] > MAX_DIL) return sh_fail(ctx, SH_EINVAL, "cnn: dilation %d of layer %d not in 1..%d",
                                                       D[l], l + 1, MAX_DIL);
    }
    return SH_OK;
}

}  // namespace

extern "C" {

long sh_cnn_nparams(void) { return NPARAM; }

// Predict (inline, crossline) dips dip[2N] from the normalised volume x[N].
int sh_cnn_forward(sh_ctx *ctx, const float *params, const float *x, float xscale, const int *dil, int n1, int n2,
                   int n3, float *dip) {
    if (!ctx) return SH_ENULL;
    if (!params || !x || !dip) return sh_fail(ctx, SH_ENULL, "cnn_forward: NULL pointer");
    if (n1 < 1 || n2 < 1 || n3 < 1) return sh_fail(ctx, SH_EINVAL, "cnn_forward: bad shape");
    long n = (long)n1 * n2 * n3;
    int D[6];
    SH_TRY(read_dil(ctx, dil, D));
    CnnWork W{};
    SH_TRY(cnn_work(ctx, n, false, W, n1, n2, n3));
    float *xsp = nullptr;
    if (!(xscale > 0.0f)) {  // scale by 1 / RMS of x, computed here (no host round trip)
        const int nb = (int)std::min<long>(ceil_div(n, TB), 4096L);
        double *part = nullptr;
        SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 16, (size_t)nb * sizeof(double) + 256, (void **)&part));
        xsp = (float *)(part + nb);
        hipLaunchKernelGGL(sumsq_partials, dim3(nb), dim3(TB), 0, ctx->stream, x, n, part);
        hipLaunchKernelGGL(inv_rms, dim3(1), dim3(1024), 0, ctx->stream, part, nb, n, xsp);
        SH_TRY(sh_launch_check(ctx, "inv_rms"));
    }
    SH_TRY(cnn_forward(ctx, params, x, xscale, D, n1, n2, n3, W, false, xsp));
    hipLaunchKernelGGL(split_dips, dim3(ceil_div(n, TB)), dim3(TB), 0, ctx->stream, W.y, dip, n);
    return sh_launch_check(ctx, "split_dips");
}

// Loss and gradient: loss = sum wt (y - target)^2 / count over both dip
// components (target [2][N], wt nullable weights [N]); grad[NPARAM] FP32.
int sh_cnn_grad(sh_ctx *ctx, const float *params, const float *x, float xscale, const int *dil, const float *target,
                const float *wt, int n1, int n2, int n3, float *grad, double *loss) {
    if (!ctx) return SH_ENULL;
    if (!params || !x || !target || !grad) return sh_fail(ctx, SH_ENULL, "cnn_grad: NULL pointer");
    if (n1 < 2 || n2 < 1 || n3 < 1) return sh_fail(ctx, SH_EINVAL, "cnn_grad: bad shape");
    long n = (long)n1 * n2 * n3;
    long per = (long)NT * C * C;
    int D[6];
    SH_TRY(read_dil(ctx, dil, D));
    CnnWork W{};
    SH_TRY(cnn_work(ctx, n, true, W, n1, n2, n3));
    SH_TRY(cnn_forward(ctx, params, x, xscale, D, n1, n2, n3, W, true));
    int nb = ceil_div(n, TB);
    float scale = 1.0f / (float)(2 * n);
    hipLaunchKernelGGL(loss_grad, dim3(nb), dim3(TB), 0, ctx->stream, W.y, target, wt, W.gA, n, scale, W.red);
    SH_TRY(sh_launch_check(ctx, "loss_grad"));
    if (loss) {
        double h = 0;
        SH_TRY(reduce_partials(ctx, W.red, nb, 1, W.dsum));
        SH_HIP(ctx, hipMemcpyAsync(&h, W.dsum, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
        *loss = h * (double)scale;
    }
    unsigned gwg = (unsigned)W.nblk_wg;
    int nred = ceil_div(per + C, TB);
    // layer 6 (input act5, output gradient gA)
    __bf16 *g = W.gA, *gn = W.gB;
    for (int l = 6; l >= 2; l--) {
        const __bf16 *ain = W.act[l - 1];
        PROF_A(ctx);
        hipLaunchKernelGGL(conv32_wgrad, dim3(gwg), dim3(TB), 0, ctx->stream, ain, g, W.part, W.pb, n1, n2, n3,
                           D[l - 1]);
        PROF_B(ctx, "conv32_wgrad", (double)n * 128.0);
        SH_TRY(sh_launch_check(ctx, "conv32_wgrad"));
        hipLaunchKernelGGL(reduce_wgrad, dim3(nred), dim3(TB), 0, ctx->stream, W.part, W.pb, (int)gwg, l == 6 ? 1 : 0,
                           l == 6 ? 0L : OFF_L(l), grad);
        SH_TRY(sh_launch_check(ctx, "reduce_wgrad"));
        // input gradient through the ReLU of layer l-1
        SH_TRY(conv_launch(ctx, 2, g, W.Wb + (size_t)(l - 2) * per, nullptr, gn, ain, n1, n2, n3, D[l - 1],
                           "conv32<bwd>"));
        std::swap(g, gn);
    }
    // layer 1 weights
    int vpb = 256;
    int nb1 = ceil_div(n, vpb);
    SH_TRY(sh_slot(ctx, SLOT_INTERP0 + 15, (size_t)nb1 * (NT * C + C) * sizeof(float), (void **)&W.part));
    hipLaunchKernelGGL(conv1_wgrad, dim3(nb1), dim3(1024), 0, ctx->stream, x, xscale, g, W.part, n1, n2, n3, vpb,
                       D[0]);
    SH_TRY(sh_launch_check(ctx, "conv1_wgrad"));
    hipLaunchKernelGGL(reduce_wgrad, dim3(ceil_div(NT * C + C, TB)), dim3(TB), 0, ctx->stream, W.part, nullptr, nb1, 2,
                       0L, grad);
    return sh_launch_check(ctx, "reduce_wgrad");
}

int sh_cnn_adam(sh_ctx *ctx, float *params, const float *grad, float *m, float *v, float lr, float b1, float b2,
                float eps, int step, float wd) {
    if (!ctx) return SH_ENULL;
    if (!params || !grad || !m || !v || step < 1) return sh_fail(ctx, SH_EINVAL, "cnn_adam: bad arguments");
    float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(NPARAM, TB)), dim3(TB), 0, ctx->stream, params, grad, m, v, NPARAM,
                       lr, b1, b2, eps, c1, c2, wd);
    return sh_launch_check(ctx, "adam");
}

}  // extern "C"
