// Shared infrastructure of the gfx950 structural-filter library: context and
// workspace, error reporting, deterministic block reductions, and the
// plane-wave-destruction filter coefficients.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/seistr_hip.h"

#define SH_HD __host__ __device__

// ---------------------------------------------------------------------------
// Context: stream, last error, and a set of cached device buffers ("slots")
// that grow on demand and are reused across calls on the same context.
// ---------------------------------------------------------------------------
struct sh_prof_rec {
    const char *name;
    double bytes;  // algorithmic bytes of the launch
    hipEvent_t a, b;
};

struct sh_ctx {
    hipStream_t stream = nullptr;
    // second stream (default priority) for work that runs beside the main
    // stream inside one call, joined back before the call returns
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // orders the previous stream before the next one in sh_ctx_set_stream
    hipEvent_t ev_switch = nullptr;
    std::string err;
    std::vector<void *> slot_ptr;
    std::vector<size_t> slot_size;
    int device = 0;
    // optional per-launch timing (HIP events on the context stream)
    int prof = 0;
    std::vector<sh_prof_rec> recs;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    // FFT plans of the FK dip filter for the last size it ran on (owned;
    // the deleter destroys them)
    std::shared_ptr<void> fk;
    // SH_FLAG_* of the public entry points (sh_ctx_set_flags)
    int flags = 0;
    // brackets of a dip call's device time (sh_dip_report.elapsed_ms)
    hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
};

hipEvent_t sh_prof_event(sh_ctx *ctx);
// Bracket a launch: PROF_A before, PROF_B(name, bytes) after.
#define PROF_A(ctx) hipEvent_t prof_a_ = (ctx)->prof ? sh_prof_event(ctx) : nullptr; \
    if (prof_a_) (void)hipEventRecord(prof_a_, (ctx)->stream)
#define PROF_B(ctx, name, bytes) do { if (prof_a_) { hipEvent_t b_ = sh_prof_event(ctx); \
    (void)hipEventRecord(b_, (ctx)->stream); (ctx)->recs.push_back({(name), (double)(bytes), prof_a_, b_}); } } while (0)

int sh_fail(sh_ctx *ctx, int code, const char *fmt, ...);
// Create the context's second stream and its fork/join events (once).
int sh_aux_stream(sh_ctx *ctx);

#define SH_HIP(ctx, call)                                                        \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess)                                                    \
            return sh_fail((ctx), SH_EHIP, "%s failed: %s", #call,               \
                           hipGetErrorString(e_));                               \
    } while (0)

#define SH_TRY(expr)                                                             \
    do {                                                                         \
        int rc_ = (expr);                                                        \
        if (rc_ != SH_OK) return rc_;                                            \
    } while (0)

// Device buffer cached in slot `id` of at least `bytes`; contents persist
// between calls only as scratch (callers must initialise what they read).
int sh_slot(sh_ctx *ctx, int id, size_t bytes, void **out);

// Launch-error check after a kernel launch.
int sh_launch_check(sh_ctx *ctx, const char *what);

// Slot ids (each module owns a range so that nested calls do not collide).
enum {
    SLOT_RED = 0,        // reduction partials (double)
    SLOT_SCAL = 1,       // device scalars
    SLOT_DIP0 = 2,       // dip module: 2..19
    SLOT_SPRAY0 = 20,    // spray module: 20..39
    SLOT_INTERP0 = 40,   // interpolation module: 40..51, warm-start network: 52..59
    SLOT_BATCH0 = 64,    // batched 2-D module: 64..79
    SLOT_FK0 = 80,       // FK dip filter: 80..81
    SLOT_IO0 = 84,       // host-pointer staging and the finite check: 84..99
    SLOT_COUNT = 100
};

// ---------------------------------------------------------------------------
// Division by a divisor fixed at launch, for indices below 2^31: a
// multiply-high and a shift instead of the long integer division sequence
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", 1994: m = floor(2^32 (2^s - d) / d) + 1 with
// s = ceil(log2 d), q = (mulhi(n, m) + n) >> s; exact for n, d < 2^31).
// ---------------------------------------------------------------------------
struct FastDiv {
    unsigned d = 1, m = 1;
    int s = 0;
    FastDiv() = default;
    explicit FastDiv(unsigned dv) : d(dv) {
        s = 0;
        while ((1ull << s) < dv) s++;
        m = (unsigned)(((1ull << 32) * ((1ull << s) - dv)) / dv + 1);
    }
    SH_HD unsigned div(unsigned n) const {
        const unsigned t = (unsigned)(((unsigned long long)n * m) >> 32);
        return (t + n) >> s;
    }
};

// ---------------------------------------------------------------------------
// Deterministic reductions.  Each block writes its partial sums (double) to
// part[k*nblk + block]; reduce_partials sums them in a fixed order.
// ---------------------------------------------------------------------------
// Lane 0 of the wave ends with the tree sum v[0] + v[32], then + the
// lane-16 half, ..., then + lane 1 (the order of a shuffle-down loop over
// offsets 32, 16, ..., 1).  The partner values move without LDS: halves of
// the wave by v_permlane32_swap, 16-lane rows by v_permlane16_swap, and
// shifts within a row by DPP row_shl (lane i reads lane i + s of its row;
// lanes 0 .. s-1 of row 0, the only ones lane 0's sum depends on, always
// have their partner).  Every lane of the wave must be active.
__device__ inline double dpp_row_shl(double v, int s) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    int l2, h2;
    switch (s) {
    case 8: l2 = __builtin_amdgcn_update_dpp(0, lo, 0x108, 0xF, 0xF, false);
            h2 = __builtin_amdgcn_update_dpp(0, hi, 0x108, 0xF, 0xF, false); break;
    case 4: l2 = __builtin_amdgcn_update_dpp(0, lo, 0x104, 0xF, 0xF, false);
            h2 = __builtin_amdgcn_update_dpp(0, hi, 0x104, 0xF, 0xF, false); break;
    case 2: l2 = __builtin_amdgcn_update_dpp(0, lo, 0x102, 0xF, 0xF, false);
            h2 = __builtin_amdgcn_update_dpp(0, hi, 0x102, 0xF, 0xF, false); break;
    default: l2 = __builtin_amdgcn_update_dpp(0, lo, 0x101, 0xF, 0xF, false);
             h2 = __builtin_amdgcn_update_dpp(0, hi, 0x101, 0xF, 0xF, false); break;
    }
    return __hiloint2double(h2, l2);
}

__device__ inline double wave_sum(double v) {
    {
        const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
        const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        v += __hiloint2double((int)b[1], (int)a[1]);  // lanes 0..31: + lanes 32..63
    }
    {
        const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
        const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        v += __hiloint2double((int)b[1], (int)a[1]);  // lanes 0..15: + lanes 16..31
    }
    v += dpp_row_shl(v, 8);
    v += dpp_row_shl(v, 4);
    v += dpp_row_shl(v, 2);
    v += dpp_row_shl(v, 1);
    return v;
}

// Sum `v` over the block (blockDim.x a multiple of 64, at most 1024).
// Returns the total in thread 0; `sh` needs 16 doubles.
__device__ inline double block_sum(double v, double *sh) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; w++) t += sh[w];
    return t;
}

// K sums over a 1-D block at once: the shuffle trees of the K values are
// interleaved and share two barriers, and each total is formed exactly as
// block_sum forms it (same tree, waves added in order).  Totals in v[] of
// thread 0; `sh` needs 16 K doubles.
template <int K>
__device__ inline void block_sum_multi(double (&v)[K], double *sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) sh[k * 16 + wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < K; k++) {
            double t = 0.0;
            for (int w = 0; w < nw; w++) t += sh[k * 16 + w];
            v[k] = t;
        }
}

// Linear thread id for 2-D/3-D blocks, used with block_sum.
__device__ inline int flat_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}

__device__ inline double block_sum_flat(double v, double *sh) {
    int t = flat_tid();
    int lane = t & 63, wid = t >> 6;
    int nw = (blockDim.x * blockDim.y * blockDim.z + 63) >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (t == 0)
        for (int w = 0; w < nw; w++) s += sh[w];
    __syncthreads();
    return s;
}

// Sum K arrays of partials (each nblk long, stacked) into out[k] with one
// block of 1024 threads; fixed summation order.
__global__ void reduce_partials_kernel(const double *part, int nblk, int K, double *out);
int reduce_partials(sh_ctx *ctx, const double *part, int nblk, int K, double *out);

// ---------------------------------------------------------------------------
// All-pass PWD filter coefficients (maximally flat, order nw).  The table
// and factor arithmetic follow apfilt_init/passfilter/aderfilter
// (pyseistr/src/dip_cfuns.c:833-910): factors (nf-j-p) and (p+j+1) are
// formed in float, products are accumulated in double, the result rounded
// to float, with contraction disabled so that the coefficients equal the
// reference's bit for bit.
// ---------------------------------------------------------------------------
constexpr double apf_b(int nw, int k) {
    int nf = 2 * nw;
    double v = 1.0;
    for (int j = 0; j < nf; j++)
        v *= (j < nf - k) ? (k + j + 1.0) / (2 * (2 * j + 1) * (j + 1)) : 1.0 / (2 * (2 * j + 1));
    return v;
}

template <int NW>
struct ApfTable {
    double b[2 * NW + 1];
    constexpr ApfTable() : b() {
        for (int k = 0; k <= 2 * NW; k++) b[k] = apf_b(NW, k);
    }
};

#pragma clang fp contract(off)
template <int NW>
SH_HD inline double apf_factor(int k, int j, float p) {
    constexpr int NF = 2 * NW;
    return (j < NF - k) ? (double)((float)(NF - j) - p) : (double)((p + (float)j) + 1.0f);
}

template <int NW>
SH_HD inline void apf_pass(float p, float *a) {
    constexpr ApfTable<NW> T;
    constexpr int NF = 2 * NW;
#pragma unroll
    for (int k = 0; k <= NF; k++) {
        double v = T.b[k];
#pragma unroll
        for (int j = 0; j < NF; j++) v *= apf_factor<NW>(k, j, p);
        a[k] = (float)v;
    }
}

// Tap k alone: the same products as apf_pass's a[k], so the same bits.
template <int NW>
SH_HD inline float apf_tap(int k, float p) {
    constexpr ApfTable<NW> T;
    constexpr int NF = 2 * NW;
    double v = T.b[k];
#pragma unroll
    for (int j = 0; j < NF; j++) v *= apf_factor<NW>(k, j, p);
    return (float)v;
}

template <int NW>
SH_HD inline void apf_der(float p, float *a) {
    constexpr ApfTable<NW> T;
    constexpr int NF = 2 * NW;
#pragma unroll
    for (int k = 0; k <= NF; k++) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < NF; i++) {
            double term = -1.0;
#pragma unroll
            for (int j = 0; j < NF; j++) {
                if (j != i) term *= apf_factor<NW>(k, j, p);
                else if (j < NF - k) term *= -1;
            }
            acc += term;
        }
        a[k] = (float)(acc * T.b[k]);
    }
}

// Nearest-integer shift of the drift variant (PS_NINT, dip_cfuns.c:9).
SH_HD inline int nint_ref(float x) { return (int)((x) > 0.0 ? (x) + 0.5 : (x)-0.5); }

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Dispatch a template on the PWD order (1..6) for the stencil-only paths
// (dip estimation, soint), whose per-sample state is one filter of 2nw+1
// taps; returns SH_EUNSUP otherwise.
#define SH_DISPATCH_NW6(ctx, nw, NAME, ...)                                      \
    switch (nw) {                                                                \
    case 1: { constexpr int NAME = 1; __VA_ARGS__; } break;                      \
    case 2: { constexpr int NAME = 2; __VA_ARGS__; } break;                      \
    case 3: { constexpr int NAME = 3; __VA_ARGS__; } break;                      \
    case 4: { constexpr int NAME = 4; __VA_ARGS__; } break;                      \
    case 5: { constexpr int NAME = 5; __VA_ARGS__; } break;                      \
    case 6: { constexpr int NAME = 6; __VA_ARGS__; } break;                      \
    default: return sh_fail((ctx), SH_EUNSUP, "PWD order %d not supported (1..6)", (nw)); \
    }

// Dispatch a template on the PWD order (1..4) for the predictor paths
// (spray, the preconditioner), whose register windows grow as (2nw+1)^2;
// returns SH_EUNSUP otherwise.
#define SH_DISPATCH_NW(ctx, nw, NAME, ...)                                       \
    switch (nw) {                                                                \
    case 1: { constexpr int NAME = 1; __VA_ARGS__; } break;                      \
    case 2: { constexpr int NAME = 2; __VA_ARGS__; } break;                      \
    case 3: { constexpr int NAME = 3; __VA_ARGS__; } break;                      \
    case 4: { constexpr int NAME = 4; __VA_ARGS__; } break;                      \
    default: return sh_fail((ctx), SH_EUNSUP, "PWD order %d not supported (1..4)", (nw)); \
    }

// ---------------------------------------------------------------------------
// 2-D structure-oriented smoothing of n3 independent n1 x n2 slices held in
// the trace-fast slice layout [slice][sample][trace] (spray.hip); used by
// somean2d and by the sint solvers (interp.hip).
// ---------------------------------------------------------------------------
struct Pwsmooth2 {
    int n1, n2, n3, ns, order;
    float e;            // predictor regularisation (eps^2)
    const float *dip;   // [slice][k][trace]
    float *ww;          // normalisation 1 / smooth(ones)
    float *U;           // slots [slot][slice][k][trace]
    float *C;           // adjoint chains [2][...]
    float *scr, *tr;    // per-thread solver scratch and running trace
    long nthr;
    float *F;           // LDL^T factors, two variants per trace
    // (nullable) device flag: when it reads non-zero the spray kernels
    // return at once (a solver that has stopped)
    const int *stop = nullptr;
};
int pws2_setup(sh_ctx *ctx, Pwsmooth2 &P, int slot_base);
int pws2_prepare(sh_ctx *ctx, Pwsmooth2 &P, int slot_scr, int slot_F);
int pws2_apply(sh_ctx *ctx, Pwsmooth2 &P, int adj, int add, float *trace, float *smooth);
// nslice slices of n1 * ntr values back to back, each transposed alike
int sh_transpose_internal(sh_ctx *ctx, const float *src, float *dst, int n1, long ntr, bool to_fast,
                          int nslice = 1);
// Median of slots [first, first+nmf) of each of n samples (slot-major U).
int sh_median_internal(sh_ctx *ctx, const float *U, float *out, long n, int first, int nmf);
