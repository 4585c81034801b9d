// Public entry points of the structural path: the context's flags around
// the device cores (shd_*, defined next to their kernels).
//  * SH_FLAG_CHECK_FINITE: each float input is scanned for NaN / Inf
//    (one block-reduced count per input, read back once) before anything
//    else runs; the call then fails with SH_ENONFINITE naming the input.
//  * SH_FLAG_HOST_PTRS: array arguments are host memory; inputs are copied
//    into device workspace slots, outputs are produced there and copied
//    back before the call returns.
// Without flags a call goes straight to its core.  The reference checks
// neither (its C entries copy NumPy buffers element by element and
// propagate NaN), so both are opt-in.
#include "common.hpp"

#include <cmath>
#include <vector>

extern "C" {
int shd_dip(sh_ctx *, const float *, const float *, const float *, int, int, int, int, int, int, float, int, int,
            int, float *, sh_dip_report *);
int shd_pwspray3d(sh_ctx *, const float *, const float *, const float *, int, int, int, int, int, int, float,
                  float *);
int shd_somean3d(sh_ctx *, const float *, const float *, const float *, int, int, int, int, int, float, int,
                 float *);
int shd_somf3d(sh_ctx *, const float *, const float *, const float *, int, int, int, int, int, int, int, int, float,
               float *);
int shd_pwspray2d(sh_ctx *, int, const float *, int, int, int, int, float, float *, float *);
int shd_somean2d(sh_ctx *, const float *, const float *, int, int, int, int, int, int, float, float *);
int shd_somf2d(sh_ctx *, const float *, const float *, int, int, int, int, int, int, int, float, float *);
int shd_soint3d(sh_ctx *, const float *, const float *, const float *, const float *, int, int, int, int, int, int,
                int, int, int, int, float, float *, int *);
int shd_soint2d(sh_ctx *, const float *, const float *, const float *, int, int, int, int, int, int, int, int,
                float *, int *);
int shd_soint2d_prec(sh_ctx *, const float *, const float *, const float *, const float *, int, int, int, int, int,
                     float *, int *);
int shd_sint3d(sh_ctx *, const float *, const float *, const float *, const float *, int, int, int, int, int, int,
               int, int, float, float *, int *);
int shd_sint2d(sh_ctx *, const float *, const float *, const float *, int, int, int, int, int, float, float *,
               int *);
}

namespace {

constexpr int TB = 256;

__global__ void __launch_bounds__(TB) nonfinite_kernel(const float *__restrict__ a, long n,
                                                        unsigned long long *cnt) {
    __shared__ unsigned int sh[TB / 64];
    unsigned int c = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        c += isfinite(a[i]) ? 0u : 1u;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = 0;
        for (int w = 0; w < TB / 64; w++) t += sh[w];
        if (t) atomicAdd(cnt, (unsigned long long)t);
    }
}

__global__ void clear_u64(unsigned long long *p, int k) {
    if ((int)threadIdx.x < k) p[threadIdx.x] = 0ull;
}

// Counts of non-finite values of k device arrays, read back once.
int count_nonfinite(sh_ctx *ctx, const float *const *a, const long *n, int k, unsigned long long *host) {
    unsigned long long *cnt = nullptr;
    SH_TRY(sh_slot(ctx, SLOT_IO0 + 15, 64 * sizeof(unsigned long long), (void **)&cnt));
    hipLaunchKernelGGL(clear_u64, dim3(1), dim3(64), 0, ctx->stream, cnt, k);
    for (int j = 0; j < k; j++) {
        if (!a[j] || n[j] <= 0) continue;
        const int g = (int)std::min<long>(ceil_div(n[j], TB), 4096);
        hipLaunchKernelGGL(nonfinite_kernel, dim3(g), dim3(TB), 0, ctx->stream, a[j], n[j], cnt + j);
    }
    SH_TRY(sh_launch_check(ctx, "nonfinite_kernel"));
    SH_HIP(ctx, hipMemcpyAsync(host, cnt, k * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    SH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return SH_OK;
}

// Arguments of one call: inputs staged in, outputs staged back, finite
// check over the inputs.
class Boundary {
  public:
    Boundary(sh_ctx *ctx, const char *who) : ctx_(ctx), who_(who) {
        host_ = ctx && (ctx->flags & SH_FLAG_HOST_PTRS);
        check_ = ctx && (ctx->flags & SH_FLAG_CHECK_FINITE);
    }
    // float input of n values (nullable); replaced by its device copy
    int in(const float *&p, long n, const char *name) {
        if (!p || n <= 0) return SH_OK;
        if (host_) {
            float *d = nullptr;
            SH_TRY(stage(n, &d));
            SH_HIP(ctx_, hipMemcpyAsync(d, p, (size_t)n * sizeof(float), hipMemcpyHostToDevice, ctx_->stream));
            p = d;
        }
        ins_.push_back(p);
        ns_.push_back(n);
        names_.push_back(name);
        return SH_OK;
    }
    // float output of n values; with host pointers it is produced in device
    // workspace and copied back by finish()
    int out(float *&p, long n) {
        if (!p || n <= 0 || !host_) return SH_OK;
        float *d = nullptr;
        SH_TRY(stage(n, &d));
        back_.push_back({p, d, n});
        p = d;
        return SH_OK;
    }
    // scan the inputs (after in() so that host arrays are scanned on the
    // device as well)
    int scan() {
        if (!check_ || ins_.empty()) return SH_OK;
        if (ins_.size() > 64) return sh_fail(ctx_, SH_EINVAL, "%s: too many inputs", who_);
        unsigned long long bad[64];
        SH_TRY(count_nonfinite(ctx_, ins_.data(), ns_.data(), (int)ins_.size(), bad));
        for (size_t j = 0; j < ins_.size(); j++)
            if (bad[j])
                return sh_fail(ctx_, SH_ENONFINITE, "%s: %s holds %llu NaN or Inf values", who_, names_[j], bad[j]);
        return SH_OK;
    }
    int finish(int rc) {
        if (rc != SH_OK || back_.empty()) return rc;
        for (auto &b : back_)
            SH_HIP(ctx_, hipMemcpyAsync(b.host, b.dev, (size_t)b.n * sizeof(float), hipMemcpyDeviceToHost,
                                        ctx_->stream));
        SH_HIP(ctx_, hipStreamSynchronize(ctx_->stream));
        return SH_OK;
    }

  private:
    int stage(long n, float **d) {
        if (slot_ >= SLOT_IO0 + 15) return sh_fail(ctx_, SH_EINVAL, "%s: too many staged arrays", who_);
        return sh_slot(ctx_, slot_++, (size_t)n * sizeof(float), (void **)d);
    }
    struct Back {
        float *host;
        float *dev;
        long n;
    };
    sh_ctx *ctx_;
    const char *who_;
    bool host_ = false, check_ = false;
    int slot_ = SLOT_IO0;
    std::vector<const float *> ins_;
    std::vector<long> ns_;
    std::vector<const char *> names_;
    std::vector<Back> back_;
};

}  // namespace

#define SH_IN(B, p, n) SH_TRY((B).in(p, n, #p))

extern "C" {

int sh_ctx_set_flags(sh_ctx *ctx, int flags) {
    if (!ctx) return SH_ENULL;
    if (flags & ~(SH_FLAG_CHECK_FINITE | SH_FLAG_HOST_PTRS)) return sh_fail(ctx, SH_EINVAL, "unknown flags %d", flags);
    ctx->flags = flags;
    return SH_OK;
}

int sh_ctx_get_flags(const sh_ctx *ctx) { return ctx ? ctx->flags : 0; }

int sh_count_nonfinite(sh_ctx *ctx, const float *a, long n, long *nbad) {
    if (!ctx) return SH_ENULL;
    if (!a || !nbad) return sh_fail(ctx, SH_ENULL, "sh_count_nonfinite: NULL pointer");
    unsigned long long c = 0;
    SH_TRY(count_nonfinite(ctx, &a, &n, 1, &c));
    *nbad = (long)c;
    return SH_OK;
}

int sh_dip(sh_ctx *ctx, const float *data, const float *mask, const float *init, int n1, int n2, int n3, int niter,
           int liter, int order, float eps_dv, int r1, int r2, int r3, float *dip, sh_dip_report *rep) {
    if (!ctx) return SH_ENULL;
    if (!data || !dip) return sh_fail(ctx, SH_ENULL, "sh_dip: data and dip must be non-NULL");
    if (n1 < 1 || n2 < 1 || n3 < 1) return sh_fail(ctx, SH_EINVAL, "sh_dip: bad shape %d x %d x %d", n1, n2, n3);
    const long n = (long)n1 * n2 * n3, nd = n3 == 1 ? n : 2 * n;
    Boundary B(ctx, "sh_dip");
    SH_IN(B, data, n);
    SH_IN(B, mask, n);
    SH_IN(B, init, nd);
    SH_TRY(B.out(dip, nd));
    SH_TRY(B.scan());
    if (rep) {
        if (!ctx->ev_t0) SH_HIP(ctx, hipEventCreate(&ctx->ev_t0));
        if (!ctx->ev_t1) SH_HIP(ctx, hipEventCreate(&ctx->ev_t1));
        SH_HIP(ctx, hipEventRecord(ctx->ev_t0, ctx->stream));
    }
    const int rc = shd_dip(ctx, data, mask, init, n1, n2, n3, niter, liter, order, eps_dv, r1, r2, r3, dip, rep);
    if (rc == SH_OK && rep) {
        SH_HIP(ctx, hipEventRecord(ctx->ev_t1, ctx->stream));
        SH_HIP(ctx, hipEventSynchronize(ctx->ev_t1));
        float ms = 0.0f;
        SH_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_t0, ctx->ev_t1));
        rep->elapsed_ms = ms;
    }
    return B.finish(rc);
}

int sh_pwspray3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, int n1, int n2, int n3,
                 int ns2, int ns3, int order, float eps, float *u) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2 * n3;
    const long np = (long)(2 * ns2 + 1) * (2 * ns3 + 1);
    Boundary B(ctx, "sh_pwspray3d");
    SH_IN(B, din, n);
    SH_IN(B, dipi, n);
    SH_IN(B, dipx, n);
    SH_TRY(B.out(u, n * np));
    SH_TRY(B.scan());
    return B.finish(shd_pwspray3d(ctx, din, dipi, dipx, n1, n2, n3, ns2, ns3, order, eps, u));
}

int sh_somean3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, int n1, int n2, int n3, int r1,
                int r2, float eps, int order, float *out) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2 * n3;
    Boundary B(ctx, "sh_somean3d");
    SH_IN(B, din, n);
    SH_IN(B, dipi, n);
    SH_IN(B, dipx, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_somean3d(ctx, din, dipi, dipx, n1, n2, n3, r1, r2, eps, order, out));
}

int sh_somf3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, int n1, int n2, int n3, int r1,
              int r2, int nmf, int option, int order, float eps, float *out) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2 * n3;
    Boundary B(ctx, "sh_somf3d");
    SH_IN(B, din, n);
    SH_IN(B, dipi, n);
    SH_IN(B, dipx, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_somf3d(ctx, din, dipi, dipx, n1, n2, n3, r1, r2, nmf, option, order, eps, out));
}

int sh_pwspray2d(sh_ctx *ctx, int adj, const float *dip, int n1, int n2, int ns, int order, float eps, float *x,
                 float *u) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2, nu = n * (2L * ns + 1);
    Boundary B(ctx, "sh_pwspray2d");
    SH_IN(B, dip, n);
    if (!adj) {
        const float *xi = x;
        SH_TRY(B.in(xi, n, "x"));
        x = const_cast<float *>(xi);
        SH_TRY(B.out(u, nu));
    } else {
        const float *ui = u;
        SH_TRY(B.in(ui, nu, "u"));
        u = const_cast<float *>(ui);
        SH_TRY(B.out(x, n));
    }
    SH_TRY(B.scan());
    return B.finish(shd_pwspray2d(ctx, adj, dip, n1, n2, ns, order, eps, x, u));
}

int sh_somean2d(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3, int ns, int order, int adj,
                float eps, float *out) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2 * n3;
    Boundary B(ctx, "sh_somean2d");
    SH_IN(B, dn, n);
    SH_IN(B, dip, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_somean2d(ctx, dn, dip, n1, n2, n3, ns, order, adj, eps, out));
}

int sh_somf2d(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3, int ns, int nmf, int option,
              int order, float eps, float *out) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2 * n3;
    Boundary B(ctx, "sh_somf2d");
    SH_IN(B, dn, n);
    SH_IN(B, dip, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_somf2d(ctx, dn, dip, n1, n2, n3, ns, nmf, option, order, eps, out));
}

int sh_soint3d(sh_ctx *ctx, const float *din, const float *mask, const float *dipi, const float *dipx, int n1,
               int n2, int n3, int nw, int nj1, int nj2, int niter, int drift, int seed, int hasmask, float var,
               float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2 * n3;
    Boundary B(ctx, "sh_soint3d");
    SH_IN(B, din, n);
    if (hasmask == 1) SH_IN(B, mask, n);
    SH_IN(B, dipi, n);
    SH_IN(B, dipx, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_soint3d(ctx, din, mask, dipi, dipx, n1, n2, n3, nw, nj1, nj2, niter, drift, seed, hasmask,
                                var, out, iters));
}

int sh_soint2d(sh_ctx *ctx, const float *din, const float *mask, const float *dip, int n1, int n2, int nw, int nj1,
               int niter, int drift, int hasmask, int twoplane, float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2;
    Boundary B(ctx, "sh_soint2d");
    SH_IN(B, din, n);
    if (hasmask == 1) SH_IN(B, mask, n);
    SH_IN(B, dip, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_soint2d(ctx, din, mask, dip, n1, n2, nw, nj1, niter, drift, hasmask, twoplane, out, iters));
}

int sh_soint2d_prec(sh_ctx *ctx, const float *din, const float *mask, const float *dip, const float *dip2, int n1,
                    int n2, int nw, int niter, int hasmask, float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2;
    Boundary B(ctx, "sh_soint2d_prec");
    SH_IN(B, din, n);
    if (hasmask == 1) SH_IN(B, mask, n);
    SH_IN(B, dip, n);
    SH_IN(B, dip2, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_soint2d_prec(ctx, din, mask, dip, dip2, n1, n2, nw, niter, hasmask, out, iters));
}

int sh_sint3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx, const float *mask, int n1, int n2,
              int n3, int niter, int ns1, int ns2, int order1, int order2, float eps, float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2 * n3;
    Boundary B(ctx, "sh_sint3d");
    SH_IN(B, din, n);
    SH_IN(B, dipi, n);
    SH_IN(B, dipx, n);
    SH_IN(B, mask, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_sint3d(ctx, din, dipi, dipx, mask, n1, n2, n3, niter, ns1, ns2, order1, order2, eps, out,
                               iters));
}

int sh_sint2d(sh_ctx *ctx, const float *din, const float *dip, const float *mask, int n1, int n2, int niter, int ns,
              int order, float eps, float *out, int *iters) {
    if (!ctx) return SH_ENULL;
    const long n = (long)n1 * n2;
    Boundary B(ctx, "sh_sint2d");
    SH_IN(B, din, n);
    SH_IN(B, dip, n);
    SH_IN(B, mask, n);
    SH_TRY(B.out(out, n));
    SH_TRY(B.scan());
    return B.finish(shd_sint2d(ctx, din, dip, mask, n1, n2, niter, ns, order, eps, out, iters));
}

}  // extern "C"
