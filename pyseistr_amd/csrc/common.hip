// Context management, error reporting and the deterministic partial-sum
// reduction shared by all modules.
#include "common.hpp"

int sh_fail(sh_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int sh_slot(sh_ctx *ctx, int id, size_t bytes, void **out) {
    if (id < 0 || id >= SLOT_COUNT) return sh_fail(ctx, SH_EINVAL, "bad workspace slot %d", id);
    if (bytes == 0) bytes = 256;
    if (ctx->slot_size[id] < bytes) {
        if (ctx->slot_ptr[id]) {
            // the stream may still use the old buffer
            hipError_t e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return sh_fail(ctx, SH_EHIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
            (void)hipFree(ctx->slot_ptr[id]);
            ctx->slot_ptr[id] = nullptr;
            ctx->slot_size[id] = 0;
        }
        size_t sz = (bytes + 4095) & ~(size_t)4095;
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess || !p)
            return sh_fail(ctx, SH_EALLOC, "hipMalloc(%zu) for workspace slot %d: %s", sz, id,
                           hipGetErrorString(e));
        ctx->slot_ptr[id] = p;
        ctx->slot_size[id] = sz;
    }
    *out = ctx->slot_ptr[id];
    return SH_OK;
}

int sh_launch_check(sh_ctx *ctx, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return sh_fail(ctx, SH_EHIP, "launch of %s failed: %s", what, hipGetErrorString(e));
    return SH_OK;
}

__global__ void __launch_bounds__(1024) reduce_partials_kernel(const double *part, int nblk, int K, double *out) {
    __shared__ double sh[16];
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nblk; i += blockDim.x) s += part[(size_t)k * nblk + i];
        double t = block_sum(s, sh);
        if (threadIdx.x == 0) out[k] = t;
        __syncthreads();
    }
}

int reduce_partials(sh_ctx *ctx, const double *part, int nblk, int K, double *out) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(1024), 0, ctx->stream, part, nblk, K, out);
    return sh_launch_check(ctx, "reduce_partials");
}

hipEvent_t sh_prof_event(sh_ctx *ctx) {
    if (ctx->pool_used == ctx->pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ctx->pool.push_back(e);
    }
    return ctx->pool[ctx->pool_used++];
}

int sh_aux_stream(sh_ctx *ctx) {
    if (ctx->aux) return SH_OK;
    SH_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    SH_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    SH_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    return SH_OK;
}

// C-order (n1, n2, n3) -> Fortran order: per i2 slab, 64 x 64 tiles of
// the (n1, n3) matrix go through LDS, each wave reading one row of 64
// consecutive i3 samples and writing one run of 64 consecutive i1 samples.
__global__ void __launch_bounds__(256) c_to_f_kernel(const float *__restrict__ in, float *__restrict__ out, int n1,
                                                     int n2, int n3) {
    __shared__ float t[64][65];
    const int i3b = blockIdx.x * 64, i1b = blockIdx.y * 64, i2 = blockIdx.z;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 64; r += 4) {
        const int i1 = i1b + r, i3 = i3b + tx;
        if (i1 < n1 && i3 < n3) t[r][tx] = in[((long)i1 * n2 + i2) * n3 + i3];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int i3 = i3b + r, i1 = i1b + tx;
        if (i1 < n1 && i3 < n3) out[((long)i3 * n2 + i2) * n1 + i1] = t[tx][r];
    }
}

extern "C" {

int sh_c_to_f(sh_ctx *ctx, const float *in, float *out, int n1, int n2, int n3) {
    if (!ctx) return SH_ENULL;
    if (!in || !out) return sh_fail(ctx, SH_ENULL, "sh_c_to_f: NULL pointer");
    if (n1 < 1 || n2 < 1 || n3 < 1) return sh_fail(ctx, SH_EINVAL, "sh_c_to_f: bad shape %d x %d x %d", n1, n2, n3);
    if (n2 > 65535 || (n1 + 63) / 64 > 65535)
        return sh_fail(ctx, SH_EUNSUP, "sh_c_to_f: shape %d x %d x %d exceeds the grid", n1, n2, n3);
    const dim3 grid((unsigned)((n3 + 63) / 64), (unsigned)((n1 + 63) / 64), (unsigned)n2);
    hipLaunchKernelGGL(c_to_f_kernel, grid, dim3(64, 4), 0, ctx->stream, in, out, n1, n2, n3);
    return sh_launch_check(ctx, "c_to_f");
}

int sh_prof_enable(sh_ctx *ctx, int on) {
    if (!ctx) return SH_ENULL;
    SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->prof = on;
    ctx->recs.clear();
    ctx->pool_used = 0;
    return SH_OK;
}

// JSON object {"name": [launches, total_ms, total_algorithmic_bytes], ...}
int sh_prof_report(sh_ctx *ctx, char *buf, int len) {
    if (!ctx || !buf || len < 3) return SH_ENULL;
    SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<std::string> names;
    std::vector<double> cnt, ms, by;
    for (auto &r : ctx->recs) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) t = 0.0f;
        size_t k = 0;
        while (k < names.size() && names[k] != r.name) k++;
        if (k == names.size()) {
            names.push_back(r.name);
            cnt.push_back(0); ms.push_back(0); by.push_back(0);
        }
        cnt[k] += 1; ms[k] += t; by[k] += r.bytes;
    }
    std::string out = "{";
    char tmp[256];
    for (size_t k = 0; k < names.size(); k++) {
        snprintf(tmp, sizeof tmp, "%s\"%s\": [%.0f, %.6f, %.6e]", k ? ", " : "", names[k].c_str(), cnt[k], ms[k], by[k]);
        out += tmp;
    }
    out += "}";
    if ((int)out.size() + 1 > len) return sh_fail(ctx, SH_EINVAL, "profile report needs %zu bytes", out.size() + 1);
    memcpy(buf, out.c_str(), out.size() + 1);
    return SH_OK;
}

sh_ctx *sh_ctx_create(void *hip_stream) {
    sh_ctx *ctx = new (std::nothrow) sh_ctx();
    if (!ctx) return nullptr;
    ctx->stream = (hipStream_t)hip_stream;
    ctx->slot_ptr.assign(SLOT_COUNT, nullptr);
    ctx->slot_size.assign(SLOT_COUNT, 0);
    (void)hipGetDevice(&ctx->device);
    return ctx;
}

int sh_ctx_trim(sh_ctx *ctx) {
    if (!ctx) return SH_ENULL;
    SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->aux) SH_HIP(ctx, hipStreamSynchronize(ctx->aux));
    ctx->fk.reset();
    for (int i = 0; i < SLOT_COUNT; i++) {
        if (ctx->slot_ptr[i]) (void)hipFree(ctx->slot_ptr[i]);
        ctx->slot_ptr[i] = nullptr;
        ctx->slot_size[i] = 0;
    }
    return SH_OK;
}

void sh_ctx_destroy(sh_ctx *ctx) {
    if (!ctx) return;
    sh_ctx_trim(ctx);
    for (auto e : ctx->pool) (void)hipEventDestroy(e);
    if (ctx->ev_t0) (void)hipEventDestroy(ctx->ev_t0);
    if (ctx->ev_t1) (void)hipEventDestroy(ctx->ev_t1);
    if (ctx->ev_switch) (void)hipEventDestroy(ctx->ev_switch);
    if (ctx->aux) {
        (void)hipStreamDestroy(ctx->aux);
        (void)hipEventDestroy(ctx->ev_fork);
        (void)hipEventDestroy(ctx->ev_join);
    }
    delete ctx;
}

int sh_ctx_set_stream(sh_ctx *ctx, void *hip_stream) {
    if (!ctx) return SH_ENULL;
    hipStream_t next = (hipStream_t)hip_stream;
    if (next != ctx->stream) {
        // The workspace slots are shared by every call on this context: work
        // still queued on the old stream may read or write them, so the new
        // stream waits for it before anything is launched there.
        if (!ctx->ev_switch) SH_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_switch, hipEventDisableTiming));
        SH_HIP(ctx, hipEventRecord(ctx->ev_switch, ctx->stream));
        SH_HIP(ctx, hipStreamWaitEvent(next, ctx->ev_switch, 0));
        ctx->stream = next;
    }
    return SH_OK;
}

const char *sh_last_error(const sh_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sh_version(void) { return 100; }

int sh_sync(sh_ctx *ctx) {
    if (!ctx) return SH_ENULL;
    SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SH_OK;
}

}  // extern "C"
