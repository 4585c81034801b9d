// Device-side pieces shared by the single-volume and the batched dip
// estimators: scalar CG / line-search state, the finalize logic, edge
// folding, and the CG prologue/epilogue fused into smoothing passes.
#pragma once
#include "common.hpp"

namespace dipcore {

// Output stores of the plain smoothing passes (the intermediate arrays
// between the axis passes, read only by the next pass): non-temporal for
// volumes of 2^24 samples or more (the kernels' NT argument).  Measured at 512x256x256:
// warm step 47.6 -> 47.4 ms; the same for every array the dip kernels
// write made it slower (47.9 ms), and on a 512x256 gather, whose arrays
// stay in the L2, non-temporal outputs cost 5 %.
__device__ inline void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }

constexpr int SM_CH1 = 512;   // samples of a line per block, axis-1 pass
constexpr int SM_CHS = 64;    // samples along the axis per task, strided pass
constexpr int SM_TW = 64;     // consecutive i1 samples per row of a strided pass
constexpr int SM_HMAX = 127;  // largest supported smoothing half-width (rect-1)

// Device-resident scalar state of one dip direction.
struct DipScal {
    double usum;        // sum u2^2 of the current residual
    double divn_norm;   // sqrt(N / sum den^2)
    double gn, gnp, g0; // CG gradient norms
    double alpha;       // gn / gnp (direction update)
    double alpha2;      // -gn / beta (model update, applied lazily)
    int accepted, ls_k;
    int cg_done, pending, cur, cg_iters, divn_zero, pad;
    int rec[2 * 64];    // per GN iteration (first 64): cg iterations, accepted step
};

enum { F_USUM = 0, F_DIVN, F_R0, F_GN, F_BETA, F_LS };

// fold position q (may lie outside [0,n)) back by even reflection
__device__ inline int fold_idx(int q, int n) {
    if (q >= 0 && q < n) return q;
    if (q < 0) q = -1 - q;
    int m = q % (2 * n);
    return (m < n) ? m : 2 * n - 1 - m;
}

// Scalar logic of one finalize step of a dip direction, given the reduced
// sums tot[] of the preceding kernel (thread 0 only).
__device__ inline void dip_scalar_step(DipScal *sc, const double *tot, int mode, int it, long n, float tol,
                                       int k_ls) {
    switch (mode) {
    case F_USUM:
        sc->usum = tot[0];
        break;
    case F_DIVN:  // dip_cfuns.c:810-817: zero denominator -> zero ratio
        sc->accepted = 0;
        sc->cur = 0;
        sc->pending = 0;
        sc->cg_iters = 0;
        if (tot[0] == 0.0) {
            sc->divn_zero = 1;
            sc->cg_done = 1;
            sc->cg_iters = -1;
        } else {
            sc->divn_zero = 0;
            sc->cg_done = 0;
            sc->divn_norm = sqrt((double)n / tot[0]);
        }
        break;
    case F_R0:  // dip_cfuns.c:300-304: zero residual -> zero model
        if (!sc->divn_zero && tot[0] == 0.0) sc->cg_done = 1;
        break;
    case F_GN:  // dip_cfuns.c:337-363
        if (sc->cg_done) break;
        if (sc->pending) sc->cur ^= 1;
        sc->pending = 0;
        sc->gn = tot[0];
        if (it == 0) {
            sc->g0 = sc->gn;
        } else {
            double alpha = sc->gn / sc->gnp, dg = sc->gn / sc->g0;
            if (alpha < (double)tol || dg < (double)tol) {
                sc->cg_done = 1;
                break;
            }
            sc->alpha = alpha;
        }
        break;
    case F_BETA:  // dip_cfuns.c:367-380 (eps = 1)
        if (sc->cg_done) break;
        {
            double beta = tot[0] + 1.0 * (tot[1] - tot[2]);
            sc->alpha2 = -sc->gn / beta;
            sc->gnp = sc->gn;
            sc->pending = 1;
            sc->cg_iters = it + 1;
        }
        break;
    case F_LS:  // dip_cfuns.c:1667-1687: accept on strict decrease
        if (sc->accepted) break;
        if (tot[0] < sc->usum || k_ls == 7) {
            sc->accepted = (tot[0] < sc->usum) ? 1 : 2;
            sc->ls_k = (tot[0] < sc->usum) ? k_ls : 8;
            sc->usum = tot[0];
            if (it < 64) {
                sc->rec[2 * it] = sc->cg_iters;
                sc->rec[2 * it + 1] = sc->ls_k;
            }
        }
        break;
    }

}

enum { PRO_PLAIN = 0, PRO_CG = 1 };
enum { EPI_PLAIN = 0, EPI_GP = 1, EPI_S = 2 };

struct CgArgs {
    const float *w;
    float *xa, *xb, *ra, *rb;  // model and residual, ping-pong by sc->cur
    float *sx, *sr, *sp;       // conjugate directions
    float *pc;                 // shaped model p
    float *gp;                 // shaped gradient
};

// gx = -eps*x + w*r with the pending model update applied first
// (dip_cfuns.c:307-316 and 376-378; eps = 1).
template <int PRO>
__device__ inline float sm_load(const float *in, long gi, bool own, const CgArgs &A, bool pend, float a2,
                                int cur) {
    if (PRO == PRO_PLAIN) return in[gi];
    const float *xi = cur ? A.xb : A.xa;
    const float *ri = cur ? A.rb : A.ra;
    float xv = xi[gi], rv = ri[gi];
    if (pend) {
        xv = xv + a2 * A.sx[gi];
        rv = rv + a2 * A.sr[gi];
        if (own) {
            (cur ? A.xa : A.xb)[gi] = xv;
            (cur ? A.ra : A.rb)[gi] = rv;
        }
    }
    return (-xv) + rv * A.w[gi];
}

// sm_load split in two for software pipelining: sm_fetch issues the loads
// (raw values into registers), sm_finish applies the same arithmetic and
// the own-sample writes later.
template <int PRO>
__device__ inline void sm_fetch(const float *in, long gi, const CgArgs &A, bool pend, int cur, float *raw) {
    if (PRO == PRO_PLAIN) {
        raw[0] = in[gi];
        return;
    }
    raw[0] = (cur ? A.xb : A.xa)[gi];
    raw[1] = (cur ? A.rb : A.ra)[gi];
    raw[2] = A.w[gi];
    if (pend) {
        raw[3] = A.sx[gi];
        raw[4] = A.sr[gi];
    }
}

template <int PRO>
__device__ inline float sm_finish(const float *raw, long gi, bool own, const CgArgs &A, bool pend, float a2,
                                  int cur) {
    if (PRO == PRO_PLAIN) return raw[0];
    float xv = raw[0], rv = raw[1];
    if (pend) {
        xv = xv + a2 * raw[3];
        rv = rv + a2 * raw[4];
        if (own) {
            (cur ? A.xa : A.xb)[gi] = xv;
            (cur ? A.ra : A.rb)[gi] = rv;
        }
    }
    return (-xv) + rv * raw[2];
}

template <int EPI, bool NT = false>
__device__ inline void sm_store(float s, float *out, long gi, const CgArgs &A, bool pend, float a2, float a1,
                                int first, double *acc) {
    if (EPI == EPI_PLAIN) {
        if constexpr (NT) st_nt(&out[gi], s);
        else out[gi] = s;
    } else if (EPI == EPI_GP) {  // gp = eps*p + S(gx), p updated lazily
        float pv = A.pc[gi];
        if (pend) {
            pv = pv + a2 * A.sp[gi];
            A.pc[gi] = pv;
        }
        float g = pv + s;
        A.gp[gi] = g;
        acc[0] += (double)g * g;
    } else {  // gx = S(gp), gr = w gx, conjugate directions (dip_cfuns.c:334-365)
        float gx = s, gr = gx * A.w[gi], gpv = A.gp[gi];
        float sp, sx, sr;
        if (first) {
            sp = gpv; sx = gx; sr = gr;
        } else {
            sp = gpv + a1 * A.sp[gi];
            sx = gx + a1 * A.sx[gi];
            sr = gr + a1 * A.sr[gi];
        }
        A.sp[gi] = sp; A.sx[gi] = sx; A.sr[gi] = sr;
        acc[0] += (double)sr * sr;
        acc[1] += (double)sp * sp;
        acc[2] += (double)sx * sx;
    }
}

__device__ inline void sm_weights(float *wts, int nb) {
    for (int k = threadIdx.x + blockDim.x * threadIdx.y; k < nb; k += blockDim.x * blockDim.y)
        wts[k] = (float)((double)(nb - k) / ((double)nb * nb));
}

// ---------------------------------------------------------------------------
// Per-sample PWD stencil (allpass1 / allpass2, dip_cfuns.c:1135-1200,
// 1399-1465) at flat index i, with ip the offset of the next trace along the
// dip direction.  Returns the sample's contribution to the partial sum.
// MODE 0: u2 = C(p) x, sum u2^2
// MODE 1: u1 = C'(p) x; divne prologue on (num=u2, den=u1): mask, then
//         scale by 1/hypot(den, 1) (dip_cfuns.c:1656-1663, 802-808), sum den^2
// MODE 2: line-search trial p = clip(p0 + lam dp), u2 = C(p) x, sum u2^2
// ---------------------------------------------------------------------------
template <int NW, int MODE>
__device__ inline double pwd_sample(const float *__restrict__ x, const float *__restrict__ p,
                                    const float *__restrict__ dp, float lam, float pmin, float pmax,
                                    float *__restrict__ pout, float *__restrict__ u1, float *__restrict__ u2,
                                    bool masked, long i, long ip, bool valid) {
    float pv;
    if (MODE == 2) {
        pv = p[i] + lam * dp[i];
        if (pv < pmin) pv = pmin;
        if (pv > pmax) pv = pmax;
        pout[i] = pv;
    } else {
        pv = p[i];
    }
    float y = 0.0f;
    if (valid) {
        float a[2 * NW + 1];
        if (MODE == 1) apf_der<NW>(pv, a);
        else apf_pass<NW>(pv, a);
#pragma unroll
        for (int iw = 0; iw <= 2 * NW; iw++) {
            int is = iw - NW;
            y += (x[i + is + ip] - x[i - is]) * a[iw];
        }
    }
    if (MODE == 1) {
        float num = u2[i], den = y;
        if (masked) { num = 0.0f; den = 0.0f; }
        double norm = 1.0 / hypot((double)den, 1.0);
        num = (float)(num * norm);
        den = (float)(den * norm);
        u2[i] = num;
        u1[i] = den;
        return (double)den * den;
    }
    u2[i] = y;
    return (double)y * y;
}

// CG start at sample i: w = den*norm, r = -(num*norm), x = p = 0; returns r^2.
__device__ inline double cg_start_sample(const float *__restrict__ num, const float *__restrict__ den,
                                         float *__restrict__ w, float *__restrict__ r, float *__restrict__ x,
                                         float *__restrict__ pc, long i, const DipScal *sc) {
    float wv = 0.0f, rv = 0.0f;
    if (!sc->divn_zero) {
        double nrm = sc->divn_norm;
        wv = (float)(den[i] * nrm);
        rv = -(float)(num[i] * nrm);
    }
    w[i] = wv;
    r[i] = rv;
    x[i] = 0.0f;
    pc[i] = 0.0f;
    return (double)rv * rv;
}

// dp = x (+ alpha2 sx when the last update is still pending)
__device__ inline void cg_result_sample(const float *__restrict__ xa, const float *__restrict__ xb,
                                        const float *__restrict__ sx, float *__restrict__ dp, long i,
                                        const DipScal *sc) {
    float v = (sc->cur ? xb : xa)[i];
    if (sc->pending) v = v + (float)sc->alpha2 * sx[i];
    dp[i] = v;
}

template <int EPI>
__device__ inline void sm_partials(double *acc, double *part, int nblk, int bid, double *sh) {
    if (EPI == EPI_PLAIN) return;
    constexpr int K = (EPI == EPI_GP) ? 1 : 3;
    for (int k = 0; k < K; k++) {
        double t = block_sum_flat(acc[k], sh);
        if (flat_tid() == 0) part[(size_t)k * nblk + bid] = t;
    }
}

}  // namespace dipcore
