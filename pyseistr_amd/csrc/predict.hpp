// Implicit plane-wave trace prediction, one thread per trace (used by
// spray.hip).  Host+device so the exact arithmetic can be exercised on
// the CPU as well.
#pragma once
#include "common.hpp"

// ---------------------------------------------------------------------------
// The predictor.  Strided accessor: element k of a per-thread array is
// p[k * s].  Forward prediction from NF traces (predict1_step for NF = 1,
// predict2_step for NF = 2, and the forward predict_step with the input as
// its own source):
//   (e R + sum_f A_f^T A_f) out = sum_f A_f^T A_f in_f + e * boundary(in)
// Adjoint predict_step (ADJ): out = A^T A S^{-1} in + e * boundary(S^{-1} in).
// ---------------------------------------------------------------------------
struct Acc {
    const float *p;
    long s;
    SH_HD float operator[](long k) const { return p[k * s]; }
};

struct Sink {
    float *p;
    long s;
    SH_HD float &operator[](long k) const { return p[k * s]; }
};
// read access to what a sink holds
SH_HD inline Acc sink_acc(const Sink &y) { return Acc{y.p, y.s}; }

// LDL^T factor of one trace's banded matrix, rows k = 0 .. n1-1: d[k] at
// d[k * fs]; the NB off-diagonal entries o[0..NB) of row k side by side at
// o[NB * k * fs] (element e of a plane of traces keeps its row at NB * e of
// the off-diagonal block and its d at e of the diagonal plane).  A row is
// one 16-byte load for NB = 4, 8 (8-byte loads for NB = 2, 6), so a step of
// a substitution issues one or two loads for the factor instead of NB + 1.
struct Factor {
    const float *d, *o;
    long fs;
};
struct FactorOut {
    float *d, *o;
    long fs;
};
SH_HD inline float fac_d(const Factor &F, long k) { return F.d[k * F.fs]; }

template <int NB>
SH_HD inline void fac_row(const Factor &F, long k, float *ok) {
    const float *p = F.o + (long)NB * k * F.fs;
    if constexpr (NB % 4 == 0) {
#pragma unroll
        for (int c = 0; c < NB / 4; c++) {
            const float4 v = reinterpret_cast<const float4 *>(p)[c];
            ok[4 * c] = v.x;
            ok[4 * c + 1] = v.y;
            ok[4 * c + 2] = v.z;
            ok[4 * c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NB / 2; c++) {
            const float2 v = reinterpret_cast<const float2 *>(p)[c];
            ok[2 * c] = v.x;
            ok[2 * c + 1] = v.y;
        }
    }
}

// NT: the rows are read only by later launches over a large volume, so
// they are written with non-temporal stores (for small volumes the L2
// keeps them for the readers).
typedef float fac_f4 __attribute__((ext_vector_type(4)));
typedef float fac_f2 __attribute__((ext_vector_type(2)));
template <int NB, bool NT = false>
SH_HD inline void fac_put(const FactorOut &F, long k, float dk, const float *ok) {
    float *p = F.o + (long)NB * k * F.fs;
    if constexpr (NT) {
        __builtin_nontemporal_store(dk, &F.d[k * F.fs]);
        if constexpr (NB % 4 == 0) {
#pragma unroll
            for (int c = 0; c < NB / 4; c++) {
                const fac_f4 v = {ok[4 * c], ok[4 * c + 1], ok[4 * c + 2], ok[4 * c + 3]};
                __builtin_nontemporal_store(v, reinterpret_cast<fac_f4 *>(p) + c);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NB / 2; c++) {
                const fac_f2 v = {ok[2 * c], ok[2 * c + 1]};
                __builtin_nontemporal_store(v, reinterpret_cast<fac_f2 *>(p) + c);
            }
        }
    } else {
        F.d[k * F.fs] = dk;
        if constexpr (NB % 4 == 0) {
#pragma unroll
            for (int c = 0; c < NB / 4; c++)
                reinterpret_cast<float4 *>(p)[c] = make_float4(ok[4 * c], ok[4 * c + 1], ok[4 * c + 2], ok[4 * c + 3]);
        } else {
#pragma unroll
            for (int c = 0; c < NB / 2; c++) reinterpret_cast<float2 *>(p)[c] = make_float2(ok[2 * c], ok[2 * c + 1]);
        }
    }
}

// Filter taps of a slope plane stored like a factor (taps[0..NB) of
// sample k side by side at o[NB * k * fs], taps[NB] at d[k * fs]): a[j],
// reversed when rev, zero outside the trace (taps_at on the stored values).
// a = b, or b reversed when rev.  Written as a branch on rev (uniform in
// every caller): as a select per tap the compiler turns the pair of array
// reads into one read at a variable index, a chain of selects per tap.
template <int NA>
SH_HD inline void tap_order(const float *b, bool rev, float *a) {
    if (rev) {
#pragma unroll
        for (int j = 0; j < NA; j++) a[j] = b[NA - 1 - j];
    } else {
#pragma unroll
        for (int j = 0; j < NA; j++) a[j] = b[j];
    }
}

template <int NW, class FT>
SH_HD inline void tap_row(const FT &T, long k, float *b) {
    constexpr int NB = 2 * NW;
    fac_row<NB>(T, k, b);
    b[NB] = fac_d(T, k);
}
template <int NW>
SH_HD inline void taps_stored(const float *b, int kk, int n1, bool rev, float *a) {
    constexpr int NA = 2 * NW + 1;
    if (kk >= 0 && kk < n1) {
        tap_order<NA>(b, rev, a);
    } else {
#pragma unroll
        for (int j = 0; j < NA; j++) a[j] = 0.0f;
    }
}

// compile-time flag passed to the step lambdas (interior samples: no
// boundary tests)
template <bool V>
struct Mid {
    static constexpr bool value = V;
};

// Floats per factor variant of a plane of n elements (the off-diagonal
// block first, so it starts 16-byte aligned when the buffer does; the
// variant stride is kept a multiple of 4 floats).
SH_HD inline long fac_plane_floats(int nb, long n) { return ((long)(nb + 1) * n + 3) / 4 * 4; }
// Factor of the trace whose sample 0 is element e0 of the variant at Fv.
SH_HD inline Factor fac_at(const float *Fv, int nb, long n, long e0, long fs) {
    return Factor{Fv + (long)nb * n + e0, Fv + (long)nb * e0, fs};
}
SH_HD inline FactorOut fac_out_at(float *Fv, int nb, long n, long e0, long fs) {
    return FactorOut{Fv + (long)nb * n + e0, Fv + (long)nb * e0, fs};
}

// Filter taps of every sample of a slope plane of n elements, stored like a
// factor (chain kernels load them instead of forming them from the slope at
// every step).
template <int NW>
__global__ void taps_plane_kernel(const float *__restrict__ dip, float *__restrict__ T, long n) {
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    constexpr int NB = 2 * NW;
    float b[NB + 1];
    apf_pass<NW>(dip[e], b);
    fac_put<NB>(fac_out_at(T, NB, n, e, 0), 0, b[NB], b);
}

// Taps of the filter at sample kk (zero outside the trace), reversed when rev.
template <int NW, class QA = Acc>
SH_HD inline void taps_at(const QA &q, int kk, int n1, bool rev, float *a) {
    constexpr int NA = 2 * NW + 1;
    if (kk < 0 || kk >= n1) {
#pragma unroll
        for (int j = 0; j < NA; j++) a[j] = 0.0f;
        return;
    }
    float b[NA];
    apf_pass<NW>(q[kk], b);
    tap_order<NA>(b, rev, a);
}

// Regulariser row k (regularization, sof_cfuns.c:548-565).
template <int NB>
SH_HD inline void reg_row(int k, int n1, float e, float &diag, float *offd) {
    diag = (float)(6. * e);
    offd[0] = (float)(-4. * e);
    offd[1] = e;
#pragma unroll
    for (int m = 2; m < NB; m++) offd[m] = 0.0f;
    if (k == 0 || k == n1 - 1) diag = e + e;
    if (k == 1 || k == n1 - 2) diag = (float)(e + 5. * e);
    if (k == 0 || k == n1 - 2) offd[0] = (float)(-2. * e);
}

// Forward sweep: factor (sf_banded_define, sof_cfuns.c:161-185), right-hand
// side, forward substitution (287-297).  Stores y[k] = b[k]/d[k] and o[m][k]
// into scr (NB+1 arrays) for the backward sweep.
template <int NW, int NF, bool ADJ>
SH_HD void pred_forward(const Acc *in, const Acc *q, const bool *rev, Sink scr, long scr_plane, int n1,
                             float e, const float *tb) {
    constexpr int NA = 2 * NW + 1, NB = 2 * NW;
    float A[NF][NA][NA];  // taps at samples k-NW+w
    float I[NF][NA];      // inputs at samples k..k+2NW
    float T[NF][NA];      // A^T-side products at samples k-NW..k+NW
    float Ld[NB], Lo[NB][NB];  // factor rows k-1-r
    float Bw[NB];              // forward-substituted rows k-1-r
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int w = 0; w < NA; w++) taps_at<NW>(q[f], w - NW, n1, rev[f], A[f][w]);
        if (!ADJ) {
#pragma unroll
            for (int w = 0; w < NA; w++) I[f][w] = (w < n1) ? in[f][w] : 0.0f;
            // T at samples -NW..NW: only sample NW.. can be nonzero
#pragma unroll
            for (int w = 0; w < NA; w++) {
                int kk = w - NW;
                float t = 0.0f;
                if (kk >= NW && kk < n1 - NW) {
                    // needs inputs kk-NW..kk+NW; kk <= NW here -> indices 0..2NW
#pragma unroll
                    for (int j = 0; j < NA; j++) t += A[f][w][j] * ((kk + j - NW) < n1 ? in[f][kk + j - NW] : 0.0f);
                }
                T[f][w] = t;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
        Ld[r] = 0.0f;
        Bw[r] = 0.0f;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[r][m] = 0.0f;
    }
    for (int k = 0; k < n1; k++) {
        // normal matrix row k (pwd_define, sof_cfuns.c:481-496)
        float diag, offd[NB];
        reg_row<NB>(k, n1, e, diag, offd);
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int j = 0; j < NA; j++) {
                int kk = k + j - NW;
                if (kk >= NW && kk < n1 - NW) {
                    float aj = A[f][j][j];
                    diag += aj * aj;
                }
            }
#pragma unroll
            for (int m = 0; m < NB; m++) {
#pragma unroll
                for (int j = m + 1; j < NA; j++) {
                    int kk = k + j - NW;
                    if (kk >= NW && kk < n1 - NW) {
                        float aj = A[f][j][j], am = A[f][j][j - m - 1];
                        offd[m] += am * aj;
                    }
                }
            }
        }
        // LDL^T row k
        float t = diag;
        int m1 = k < NB ? k : NB;
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) t -= (Lo[m][m]) * (Lo[m][m]) * (Ld[m]);
        float dk = t;
        float ok[NB];
        int nn = (n1 - k - 1) < NB ? (n1 - k - 1) : NB;
#pragma unroll
        for (int nidx = 0; nidx < NB; nidx++) {
            ok[nidx] = 0.0f;
            if (nidx < nn) {
                float s = offd[nidx];
                int mm = k < (NB - nidx - 1) ? k : (NB - nidx - 1);
#pragma unroll
                for (int m = 0; m < NB; m++)
                    if (m < mm) s -= (Lo[m][m]) * (Lo[m][nidx + m + 1]) * (Ld[m]);
                ok[nidx] = s / dk;
            }
        }
        // right-hand side row k
        float rhs;
        if (ADJ) {
            rhs = in[0][k];
        } else {
            float o0 = 0.0f;
#pragma unroll
            for (int f = 0; f < NF; f++) {
                float of = 0.0f;
#pragma unroll
                for (int j = 0; j < NA; j++) {
                    int kk = k + j - NW;
                    if (kk >= NW && kk < n1 - NW) of += A[f][j][j] * T[f][j];
                }
                o0 = (f == 0) ? of : (o0 + of);
            }
            rhs = o0;
            if (k == 0) rhs += e * tb[0];
            if (k == 1) rhs += e * tb[1];
            if (k == n1 - 2) rhs += e * tb[2];
            if (k == n1 - 1) rhs += e * tb[3];
        }
        // forward substitution
        float b = rhs;
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) b -= (Lo[m][m]) * Bw[m];
        Sink y{scr.p + 0 * scr_plane, scr.s};
        y[k] = b / dk;
#pragma unroll
        for (int m = 0; m < NB; m++) {
            Sink om{scr.p + (long)(1 + m) * scr_plane, scr.s};
            om[k] = ok[m];
        }
        // shift windows
#pragma unroll
        for (int r = NB - 1; r > 0; r--) {
            Ld[r] = Ld[r - 1];
            Bw[r] = Bw[r - 1];
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[r][m] = Lo[r - 1][m];
        }
        Ld[0] = dk;
        Bw[0] = b;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[0][m] = ok[m];
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int w = 0; w < NA - 1; w++)
#pragma unroll
                for (int j = 0; j < NA; j++) A[f][w][j] = A[f][w + 1][j];
            taps_at<NW>(q[f], k + 1 + NW, n1, rev[f], A[f][NA - 1]);
            if (!ADJ) {
#pragma unroll
                for (int w = 0; w < NA - 1; w++) {
                    I[f][w] = I[f][w + 1];
                    T[f][w] = T[f][w + 1];
                }
                int kn = k + 1 + 2 * NW;
                I[f][NA - 1] = (kn < n1) ? in[f][kn] : 0.0f;
                int kk = k + 1 + NW;
                float tn = 0.0f;
                if (kk >= NW && kk < n1 - NW) {
#pragma unroll
                    for (int j = 0; j < NA; j++) tn += A[f][NA - 1][j] * I[f][j];
                }
                T[f][NA - 1] = tn;
            }
        }
    }
}

// Backward substitution (sof_cfuns.c:298-306); writes out[k] (descending)
// and, when keep != nullptr, the solved values into scr plane 0.
template <int NW>
SH_HD void pred_backward(Sink scr, long scr_plane, int n1, Sink out) {
    constexpr int NB = 2 * NW;
    float nx[NB];
#pragma unroll
    for (int m = 0; m < NB; m++) nx[m] = 0.0f;
    Sink y{scr.p, scr.s};
    for (int k = n1 - 1; k >= 0; k--) {
        float t = y[k];
        int m1 = (n1 - k - 1) < NB ? (n1 - k - 1) : NB;
#pragma unroll
        for (int m = 0; m < NB; m++) {
            if (m < m1) {
                Sink om{scr.p + (long)(1 + m) * scr_plane, scr.s};
                t -= om[k] * nx[m];
            }
        }
#pragma unroll
        for (int m = NB - 1; m > 0; m--) nx[m] = nx[m - 1];
        nx[0] = t;
        out[k] = t;
    }
}

// Adjoint PWD product inp = A^T A out (pwd_set adj, sof_cfuns.c:505-522)
// on the solved trace s, plus the boundary terms (predict_step, 636-646).
template <int NW>
SH_HD void pred_adj_apply(const Acc &s, const Acc &q, bool rev, int n1, float e, Sink out) {
    constexpr int NA = 2 * NW + 1;
    float A[NA][NA];   // taps at samples k-NW+w
    float S[NA];       // solved at samples k+NW-NW.. : window for tmp at k+NW: s[k..k+2NW]
    float T[NA];       // tmp at samples k-NW..k+NW
#pragma unroll
    for (int w = 0; w < NA; w++) taps_at<NW>(q, w - NW, n1, rev, A[w]);
#pragma unroll
    for (int w = 0; w < NA; w++) S[w] = (w < n1) ? s[w] : 0.0f;
    // tmp at samples -NW..NW
#pragma unroll
    for (int w = 0; w < NA; w++) {
        int kk = w - NW;
        float t = 0.0f;
        if (kk >= NW && kk < n1 - NW) {
#pragma unroll
            for (int j = NA - 1; j >= 0; j--) t += A[w][j] * s[kk - j + NW];
        }
        T[w] = t;
    }
    float s0 = s[0], s1 = s[1], s2 = s[n1 - 2], s3 = s[n1 - 1];
    for (int k = 0; k < n1; k++) {
        float v = 0.0f;
#pragma unroll
        for (int w = 0; w < NA; w++) {
            int i = k - NW + w;
            if (i >= NW && i < n1 - NW) v += A[w][2 * NW - w] * T[w];
        }
        if (k == 0) v += e * s0;
        if (k == 1) v += e * s1;
        if (k == n1 - 2) v += e * s2;
        if (k == n1 - 1) v += e * s3;
        out[k] = v;
        // advance
#pragma unroll
        for (int w = 0; w < NA - 1; w++) {
#pragma unroll
            for (int j = 0; j < NA; j++) A[w][j] = A[w + 1][j];
            T[w] = T[w + 1];
            S[w] = S[w + 1];
        }
        taps_at<NW>(q, k + 1 + NW, n1, rev, A[NA - 1]);
        int kn = k + 1 + 2 * NW;
        S[NA - 1] = (kn < n1) ? s[kn] : 0.0f;
        int kk = k + 1 + NW;
        float tn = 0.0f;
        if (kk >= NW && kk < n1 - NW) {
            // s[kk - j + NW] = S[2NW - j] (S covers k+1..k+1+2NW)
#pragma unroll
            for (int j = NA - 1; j >= 0; j--) tn += A[NA - 1][j] * S[2 * NW - j];
        }
        T[NA - 1] = tn;
    }
}

// Forward prediction of one trace: out = P(in_0[, in_1]).
template <int NW, int NF>
SH_HD void predict(const Acc *in, const Acc *q, const bool *rev, Sink scr, long plane, int n1, float e,
                        Sink out) {
    float tb[4];
    if (NF == 1) {
        tb[0] = in[0][0]; tb[1] = in[0][1]; tb[2] = in[0][n1 - 2]; tb[3] = in[0][n1 - 1];
    } else {
        tb[0] = (float)(0.5 * (in[0][0] + in[1][0]));
        tb[1] = (float)(0.5 * (in[0][1] + in[1][1]));
        tb[2] = (float)(0.5 * (in[0][n1 - 2] + in[1][n1 - 2]));
        tb[3] = (float)(0.5 * (in[0][n1 - 1] + in[1][n1 - 1]));
    }
    pred_forward<NW, NF, false>(in, q, rev, scr, plane, n1, e, tb);
    pred_backward<NW>(scr, plane, n1, out);
}

// Adjoint prediction step (predict_step with adj, sof_cfuns.c:622-649).
template <int NW>
SH_HD void predict_adj(const Acc &in, const Acc &q, bool rev, Sink scr, long plane, int n1, float e,
                            Sink out) {
    float tb[4] = {0, 0, 0, 0};
    Acc ins[1] = {in};
    Acc qs[1] = {q};
    bool rv[1] = {rev};
    pred_forward<NW, 1, true>(ins, qs, rv, scr, plane, n1, e, tb);
    // solved values overwrite plane 0 (y) in place
    Sink y{scr.p, scr.s};
    pred_backward<NW>(scr, plane, n1, y);
    Acc s{scr.p, scr.s};
    pred_adj_apply<NW>(s, q, rev, n1, e, out);
}


// ---------------------------------------------------------------------------
// Factored form.  The system matrix of a prediction depends only on the
// slope traces and the filter directions, so the LDL^T factor is computed
// once (factor_sweep) and shared by every prediction that uses the same
// matrix; solve_factored then builds the right-hand side, substitutes
// forward and backward with the stored factor.  The arithmetic is the same
// as the fused path, operation for operation.
// Factor layout: struct Factor (d plane, off-diagonal rows side by side).
// ---------------------------------------------------------------------------
template <int NW, int NF>
SH_HD void factor_sweep(const Acc *q, const bool *rev, int n1, float e, FactorOut F) {
    constexpr int NA = 2 * NW + 1, NB = 2 * NW;
    float A[NF][NA][NA];
    float Ld[NB], Lo[NB][NB];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int w = 0; w < NA; w++) taps_at<NW>(q[f], w - NW, n1, rev[f], A[f][w]);
#pragma unroll
    for (int r = 0; r < NB; r++) {
        Ld[r] = 0.0f;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[r][m] = 0.0f;
    }
    for (int k = 0; k < n1; k++) {
        float diag, offd[NB];
        reg_row<NB>(k, n1, e, diag, offd);
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int j = 0; j < NA; j++) {
                int kk = k + j - NW;
                if (kk >= NW && kk < n1 - NW) {
                    float aj = A[f][j][j];
                    diag += aj * aj;
                }
            }
#pragma unroll
            for (int m = 0; m < NB; m++) {
#pragma unroll
                for (int j = m + 1; j < NA; j++) {
                    int kk = k + j - NW;
                    if (kk >= NW && kk < n1 - NW) {
                        float aj = A[f][j][j], am = A[f][j][j - m - 1];
                        offd[m] += am * aj;
                    }
                }
            }
        }
        float t = diag;
        int m1 = k < NB ? k : NB;
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) t -= (Lo[m][m]) * (Lo[m][m]) * (Ld[m]);
        float dk = t;
        float ok[NB];
        int nn = (n1 - k - 1) < NB ? (n1 - k - 1) : NB;
#pragma unroll
        for (int nidx = 0; nidx < NB; nidx++) {
            ok[nidx] = 0.0f;
            if (nidx < nn) {
                float s = offd[nidx];
                int mm = k < (NB - nidx - 1) ? k : (NB - nidx - 1);
#pragma unroll
                for (int m = 0; m < NB; m++)
                    if (m < mm) s -= (Lo[m][m]) * (Lo[m][nidx + m + 1]) * (Ld[m]);
                ok[nidx] = s / dk;
            }
        }
        fac_put<NB>(F, k, dk, ok);
#pragma unroll
        for (int r = NB - 1; r > 0; r--) {
            Ld[r] = Ld[r - 1];
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[r][m] = Lo[r - 1][m];
        }
        Ld[0] = dk;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[0][m] = ok[m];
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int w = 0; w < NA - 1; w++)
#pragma unroll
                for (int j = 0; j < NA; j++) A[f][w][j] = A[f][w + 1][j];
            taps_at<NW>(q[f], k + 1 + NW, n1, rev[f], A[f][NA - 1]);
        }
    }
}

// Taps from a slope value already loaded (taps_at without the load).
template <int NW>
SH_HD inline void taps_from(float qv, int kk, int n1, bool rev, float *a) {
    constexpr int NA = 2 * NW + 1;
    if (kk < 0 || kk >= n1) {
#pragma unroll
        for (int j = 0; j < NA; j++) a[j] = 0.0f;
        return;
    }
    float b[NA];
    apf_pass<NW>(qv, b);
    tap_order<NA>(b, rev, a);
}

// factor_sweep with the slope loads software-pipelined D samples ahead (a
// register ring; loop unrolled by D); same arithmetic, identical factor.
template <int NW, int NF, int D, bool NT = false>
SH_HD void factor_sweep_pf(const Acc *q, const bool *rev, int n1, float e, FactorOut F) {
    constexpr int NA = 2 * NW + 1, NB = 2 * NW;
    float A[NF][NA][NA];
    float Ld[NB], Lo[NB][NB];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int w = 0; w < NA; w++) taps_at<NW>(q[f], w - NW, n1, rev[f], A[f][w]);
#pragma unroll
    for (int r = 0; r < NB; r++) {
        Ld[r] = 0.0f;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[r][m] = 0.0f;
    }
    float rq[D][NF];  // slope at sample k + 1 + NW
    auto fetch = [&](int s, int k) {
        const int kq = k + 1 + NW < n1 ? k + 1 + NW : n1 - 1;
#pragma unroll
        for (int f = 0; f < NF; f++) rq[s][f] = q[f][kq];
    };
    auto step = [&](int s, int k) {
        float diag, offd[NB];
        reg_row<NB>(k, n1, e, diag, offd);
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int j = 0; j < NA; j++) {
                int kk = k + j - NW;
                if (kk >= NW && kk < n1 - NW) {
                    float aj = A[f][j][j];
                    diag += aj * aj;
                }
            }
#pragma unroll
            for (int m = 0; m < NB; m++) {
#pragma unroll
                for (int j = m + 1; j < NA; j++) {
                    int kk = k + j - NW;
                    if (kk >= NW && kk < n1 - NW) {
                        float aj = A[f][j][j], am = A[f][j][j - m - 1];
                        offd[m] += am * aj;
                    }
                }
            }
        }
        float t = diag;
        int m1 = k < NB ? k : NB;
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) t -= (Lo[m][m]) * (Lo[m][m]) * (Ld[m]);
        float dk = t;
        float ok[NB];
        int nn = (n1 - k - 1) < NB ? (n1 - k - 1) : NB;
#pragma unroll
        for (int nidx = 0; nidx < NB; nidx++) {
            ok[nidx] = 0.0f;
            if (nidx < nn) {
                float sm = offd[nidx];
                int mm = k < (NB - nidx - 1) ? k : (NB - nidx - 1);
#pragma unroll
                for (int m = 0; m < NB; m++)
                    if (m < mm) sm -= (Lo[m][m]) * (Lo[m][nidx + m + 1]) * (Ld[m]);
                ok[nidx] = sm / dk;
            }
        }
        fac_put<NB, NT>(F, k, dk, ok);
#pragma unroll
        for (int r = NB - 1; r > 0; r--) {
            Ld[r] = Ld[r - 1];
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[r][m] = Lo[r - 1][m];
        }
        Ld[0] = dk;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[0][m] = ok[m];
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int w = 0; w < NA - 1; w++)
#pragma unroll
                for (int j = 0; j < NA; j++) A[f][w][j] = A[f][w + 1][j];
            taps_from<NW>(rq[s][f], k + 1 + NW, n1, rev[f], A[f][NA - 1]);
        }
        fetch(s, k + D);
    };
#pragma unroll
    for (int s = 0; s < D; s++) fetch(s, s);
    int k0 = 0;
    for (; k0 + D <= n1; k0 += D) {
#pragma unroll
        for (int s = 0; s < D; s++) step(s, k0 + s);
    }
#pragma unroll
    for (int s = 0; s < D; s++)
        if (k0 + s < n1) step(s, k0 + s);
}

// Forward prediction with a precomputed factor F; y is per-thread scratch.
template <int NW, int NF>
SH_HD void solve_factored(const Acc *in, const Acc *q, const bool *rev, const Factor &F, Sink y, int n1, float e,
                          Sink out) {
    constexpr int NA = 2 * NW + 1, NB = 2 * NW;
    float tb[4];
    if (NF == 1) {
        tb[0] = in[0][0]; tb[1] = in[0][1]; tb[2] = in[0][n1 - 2]; tb[3] = in[0][n1 - 1];
    } else {
        tb[0] = (float)(0.5 * (in[0][0] + in[1][0]));
        tb[1] = (float)(0.5 * (in[0][1] + in[1][1]));
        tb[2] = (float)(0.5 * (in[0][n1 - 2] + in[1][n1 - 2]));
        tb[3] = (float)(0.5 * (in[0][n1 - 1] + in[1][n1 - 1]));
    }
    float A[NF][NA][NA], I[NF][NA], T[NF][NA];
    float Lo[NB][NB], Bw[NB];
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int w = 0; w < NA; w++) taps_at<NW>(q[f], w - NW, n1, rev[f], A[f][w]);
#pragma unroll
        for (int w = 0; w < NA; w++) I[f][w] = (w < n1) ? in[f][w] : 0.0f;
#pragma unroll
        for (int w = 0; w < NA; w++) {
            int kk = w - NW;
            float t = 0.0f;
            if (kk >= NW && kk < n1 - NW) {
#pragma unroll
                for (int j = 0; j < NA; j++) t += A[f][w][j] * ((kk + j - NW) < n1 ? in[f][kk + j - NW] : 0.0f);
            }
            T[f][w] = t;
        }
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
        Bw[r] = 0.0f;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[r][m] = 0.0f;
    }
    for (int k = 0; k < n1; k++) {
        float rhs = 0.0f;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            float of = 0.0f;
#pragma unroll
            for (int j = 0; j < NA; j++) {
                int kk = k + j - NW;
                if (kk >= NW && kk < n1 - NW) of += A[f][j][j] * T[f][j];
            }
            rhs = (f == 0) ? of : (rhs + of);
        }
        if (k == 0) rhs += e * tb[0];
        if (k == 1) rhs += e * tb[1];
        if (k == n1 - 2) rhs += e * tb[2];
        if (k == n1 - 1) rhs += e * tb[3];
        int m1 = k < NB ? k : NB;
        float b = rhs;
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) b -= (Lo[m][m]) * Bw[m];
        float dk = F.d[(long)k * F.fs];
        y[k] = b / dk;
        float ok[NB];
        fac_row<NB>(F, k, ok);
#pragma unroll
        for (int r = NB - 1; r > 0; r--) {
            Bw[r] = Bw[r - 1];
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[r][m] = Lo[r - 1][m];
        }
        Bw[0] = b;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[0][m] = ok[m];
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int w = 0; w < NA - 1; w++) {
#pragma unroll
                for (int j = 0; j < NA; j++) A[f][w][j] = A[f][w + 1][j];
                I[f][w] = I[f][w + 1];
                T[f][w] = T[f][w + 1];
            }
            taps_at<NW>(q[f], k + 1 + NW, n1, rev[f], A[f][NA - 1]);
            int kn = k + 1 + 2 * NW;
            I[f][NA - 1] = (kn < n1) ? in[f][kn] : 0.0f;
            int kk = k + 1 + NW;
            float tn = 0.0f;
            if (kk >= NW && kk < n1 - NW) {
#pragma unroll
                for (int j = 0; j < NA; j++) tn += A[f][NA - 1][j] * I[f][j];
            }
            T[f][NA - 1] = tn;
        }
    }
    float nx[NB];
#pragma unroll
    for (int m = 0; m < NB; m++) nx[m] = 0.0f;
    for (int k = n1 - 1; k >= 0; k--) {
        float t = y[k];
        int m1 = (n1 - k - 1) < NB ? (n1 - k - 1) : NB;
        float ok[NB];
        fac_row<NB>(F, k, ok);
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) t -= ok[m] * nx[m];
#pragma unroll
        for (int m = NB - 1; m > 0; m--) nx[m] = nx[m - 1];
        nx[0] = t;
        out[k] = t;
    }
}

// solve_factored with its per-sample loads software-pipelined: the factor
// values, the slope and the input sample that iteration k needs are loaded
// at iteration k - D into a register ring (the sample loops are unrolled
// by D so each ring slot is a fixed register), so a thread waits on memory
// about once per D samples instead of every sample.  Ring loads clamp
// their index into the trace (they are only consumed where the unclamped
// index is valid).  The arithmetic is solve_factored's, operation for
// operation, so the results are identical.
// The input, slope, factor and scratch / output accessors are template
// types (IA, FA, SK): Acc / Factor / Sink, or the buffer-resource ones of
// the 3-D level kernels; the arithmetic does not depend on them.
template <int NW, int NF, int D, bool TP = false, bool MIDR = false, class IA = Acc, class FA = Factor,
          class SK = Sink, class TQ = Factor, class SO = SK>
SH_HD void solve_factored_pf(const IA *in, const IA *q, const bool *rev, const FA &F, SK y, int n1, float e,
                             SO out, const TQ *tq = nullptr) {
    constexpr int NA = 2 * NW + 1, NB = 2 * NW;
    float tb[4];
    if (NF == 1) {
        tb[0] = in[0][0]; tb[1] = in[0][1]; tb[2] = in[0][n1 - 2]; tb[3] = in[0][n1 - 1];
    } else {
        tb[0] = (float)(0.5 * (in[0][0] + in[1][0]));
        tb[1] = (float)(0.5 * (in[0][1] + in[1][1]));
        tb[2] = (float)(0.5 * (in[0][n1 - 2] + in[1][n1 - 2]));
        tb[3] = (float)(0.5 * (in[0][n1 - 1] + in[1][n1 - 1]));
    }
    float A[NF][NA][NA], I[NF][NA], T[NF][NA];
    float Lo[NB][NB], Bw[NB];
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int w = 0; w < NA; w++) {
            if constexpr (TP) {
                float b[NA];
                tap_row<NW>(tq[f], (w - NW) >= 0 && (w - NW) < n1 ? w - NW : 0, b);
                taps_stored<NW>(b, w - NW, n1, rev[f], A[f][w]);
            } else {
                taps_at<NW>(q[f], w - NW, n1, rev[f], A[f][w]);
            }
        }
#pragma unroll
        for (int w = 0; w < NA; w++) I[f][w] = (w < n1) ? in[f][w] : 0.0f;
#pragma unroll
        for (int w = 0; w < NA; w++) {
            int kk = w - NW;
            float t = 0.0f;
            if (kk >= NW && kk < n1 - NW) {
#pragma unroll
                for (int j = 0; j < NA; j++) t += A[f][w][j] * ((kk + j - NW) < n1 ? in[f][kk + j - NW] : 0.0f);
            }
            T[f][w] = t;
        }
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
        Bw[r] = 0.0f;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[r][m] = 0.0f;
    }
    // forward ring: d[k], o[m][k], slope (or its taps) at k + 1 + NW,
    // input at k + 1 + 2 NW
    constexpr int NQ = TP ? NA : 1;
    float rd[D], ro[D][NB], rq[D][NF][NQ], ri[D][NF];
    auto fetch = [&](int s, int k) {
        const long kc = k < n1 ? k : n1 - 1;
        rd[s] = fac_d(F, kc);
        fac_row<NB>(F, kc, ro[s]);
        const int kq = k + 1 + NW < n1 ? k + 1 + NW : n1 - 1;
        const int ki = k + 1 + 2 * NW < n1 ? k + 1 + 2 * NW : n1 - 1;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            if constexpr (TP) tap_row<NW>(tq[f], kq, rq[s][f]);
            else rq[s][f][0] = q[f][kq];
            ri[s][f] = in[f][ki];
        }
    };
    // MID: a sample of [2 NW, n1 - 2 NW - 1), where every boundary test
    // below is known to pass or fail (same operations as the general step)
    auto step = [&](auto mid, int s, int k) {
        constexpr bool MID = decltype(mid)::value;
        float rhs = 0.0f;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            float of = 0.0f;
#pragma unroll
            for (int j = 0; j < NA; j++) {
                int kk = k + j - NW;
                if (MID || (kk >= NW && kk < n1 - NW)) of += A[f][j][j] * T[f][j];
            }
            rhs = (f == 0) ? of : (rhs + of);
        }
        if (!MID) {
            if (k == 0) rhs += e * tb[0];
            if (k == 1) rhs += e * tb[1];
            if (k == n1 - 2) rhs += e * tb[2];
            if (k == n1 - 1) rhs += e * tb[3];
        }
        int m1 = MID ? NB : (k < NB ? k : NB);
        float b = rhs;
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) b -= (Lo[m][m]) * Bw[m];
        float dk = rd[s];
        y[k] = b / dk;
        float ok[NB];
#pragma unroll
        for (int m = 0; m < NB; m++) ok[m] = ro[s][m];
#pragma unroll
        for (int r = NB - 1; r > 0; r--) {
            Bw[r] = Bw[r - 1];
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[r][m] = Lo[r - 1][m];
        }
        Bw[0] = b;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[0][m] = ok[m];
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int w = 0; w < NA - 1; w++) {
#pragma unroll
                for (int j = 0; j < NA; j++) A[f][w][j] = A[f][w + 1][j];
                I[f][w] = I[f][w + 1];
                T[f][w] = T[f][w + 1];
            }
            if constexpr (TP) {
                if (MID) {
                    tap_order<NA>(rq[s][f], rev[f], A[f][NA - 1]);
                } else {
                    taps_stored<NW>(rq[s][f], k + 1 + NW, n1, rev[f], A[f][NA - 1]);
                }
            } else {
                taps_from<NW>(rq[s][f][0], k + 1 + NW, n1, rev[f], A[f][NA - 1]);
            }
            int kn = k + 1 + 2 * NW;
            I[f][NA - 1] = (MID || kn < n1) ? ri[s][f] : 0.0f;
            int kk = k + 1 + NW;
            float tn = 0.0f;
            if (MID || (kk >= NW && kk < n1 - NW)) {
#pragma unroll
                for (int j = 0; j < NA; j++) tn += A[f][NA - 1][j] * I[f][j];
            }
            T[f][NA - 1] = tn;
        }
        fetch(s, k + D);
    };
#pragma unroll
    for (int s = 0; s < D; s++) fetch(s, s);
    int k0 = 0;
    for (; k0 + D <= n1; k0 += D) {
        if (MIDR && k0 >= 2 * NW && k0 + D <= n1 - 2 * NW - 1) {
#pragma unroll
            for (int s = 0; s < D; s++) step(Mid<true>{}, s, k0 + s);
        } else {
#pragma unroll
            for (int s = 0; s < D; s++) step(Mid<false>{}, s, k0 + s);
        }
    }
#pragma unroll
    for (int s = 0; s < D; s++)
        if (k0 + s < n1) step(Mid<false>{}, s, k0 + s);
    // backward ring: y[k], o[m][k]
    float nx[NB];
#pragma unroll
    for (int m = 0; m < NB; m++) nx[m] = 0.0f;
    float ry[D], rb[D][NB];
    auto fetchb = [&](int s, int k) {
        const long kc = k >= 0 ? k : 0;
        ry[s] = y[kc];
        fac_row<NB>(F, kc, rb[s]);
    };
    // MID: k <= n1 - 1 - NB (every band entry below the diagonal is used)
    auto stepb = [&](auto mid, int s, int k) {
        constexpr bool MID = decltype(mid)::value;
        float t = ry[s];
        int m1 = MID ? NB : ((n1 - k - 1) < NB ? (n1 - k - 1) : NB);
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) t -= rb[s][m] * nx[m];
#pragma unroll
        for (int m = NB - 1; m > 0; m--) nx[m] = nx[m - 1];
        nx[0] = t;
        out[k] = t;
        fetchb(s, k - D);
    };
#pragma unroll
    for (int s = 0; s < D; s++) fetchb(s, n1 - 1 - s);
    int kb = n1 - 1;
    for (; kb - D + 1 >= 0; kb -= D) {
        if (MIDR && kb <= n1 - 1 - NB) {
#pragma unroll
            for (int s = 0; s < D; s++) stepb(Mid<true>{}, s, kb - s);
        } else {
#pragma unroll
            for (int s = 0; s < D; s++) stepb(Mid<false>{}, s, kb - s);
        }
    }
#pragma unroll
    for (int s = 0; s < D; s++)
        if (kb - s >= 0) stepb(Mid<false>{}, s, kb - s);
}

// M forward predictions that share the slopes, the directions and the
// factor (the slots of one spray level and quadrant that arrive at the same
// target trace): the filter taps and factor values are loaded once per
// sample for all of them.  Right-hand side r uses in[r*NF + f], scratch
// y[r] and writes out[r] when live[r]; a dead one must still point at
// readable inputs.  Each live result is bit-identical to solve_factored.
template <int NW, int NF, int M>
SH_HD void solve_factored_multi(const Acc *in, const bool *live, const Acc *q, const bool *rev, const Factor &F,
                                const Sink *y, int n1, float e, const Sink *out) {
    constexpr int NA = 2 * NW + 1, NB = 2 * NW;
    float tb[M][4];
#pragma unroll
    for (int r = 0; r < M; r++) {
        const Acc *ir = in + r * NF;
        if (NF == 1) {
            tb[r][0] = ir[0][0]; tb[r][1] = ir[0][1]; tb[r][2] = ir[0][n1 - 2]; tb[r][3] = ir[0][n1 - 1];
        } else {
            tb[r][0] = (float)(0.5 * (ir[0][0] + ir[1][0]));
            tb[r][1] = (float)(0.5 * (ir[0][1] + ir[1][1]));
            tb[r][2] = (float)(0.5 * (ir[0][n1 - 2] + ir[1][n1 - 2]));
            tb[r][3] = (float)(0.5 * (ir[0][n1 - 1] + ir[1][n1 - 1]));
        }
    }
    float A[NF][NA][NA], I[M][NF][NA], T[M][NF][NA];
    float Lo[NB][NB], Bw[M][NB];
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int w = 0; w < NA; w++) taps_at<NW>(q[f], w - NW, n1, rev[f], A[f][w]);
#pragma unroll
        for (int r = 0; r < M; r++) {
            const Acc &ir = in[r * NF + f];
#pragma unroll
            for (int w = 0; w < NA; w++) I[r][f][w] = (w < n1) ? ir[w] : 0.0f;
#pragma unroll
            for (int w = 0; w < NA; w++) {
                int kk = w - NW;
                float t = 0.0f;
                if (kk >= NW && kk < n1 - NW) {
#pragma unroll
                    for (int j = 0; j < NA; j++) t += A[f][w][j] * ((kk + j - NW) < n1 ? ir[kk + j - NW] : 0.0f);
                }
                T[r][f][w] = t;
            }
        }
    }
#pragma unroll
    for (int rr = 0; rr < NB; rr++) {
#pragma unroll
        for (int r = 0; r < M; r++) Bw[r][rr] = 0.0f;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[rr][m] = 0.0f;
    }
    for (int k = 0; k < n1; k++) {
        int m1 = k < NB ? k : NB;
        float dk = F.d[(long)k * F.fs];
        float ok[NB];
        fac_row<NB>(F, k, ok);
        float bnew[M];
#pragma unroll
        for (int r = 0; r < M; r++) {
            float rhs = 0.0f;
#pragma unroll
            for (int f = 0; f < NF; f++) {
                float of = 0.0f;
#pragma unroll
                for (int j = 0; j < NA; j++) {
                    int kk = k + j - NW;
                    if (kk >= NW && kk < n1 - NW) of += A[f][j][j] * T[r][f][j];
                }
                rhs = (f == 0) ? of : (rhs + of);
            }
            if (k == 0) rhs += e * tb[r][0];
            if (k == 1) rhs += e * tb[r][1];
            if (k == n1 - 2) rhs += e * tb[r][2];
            if (k == n1 - 1) rhs += e * tb[r][3];
            float b = rhs;
#pragma unroll
            for (int m = 0; m < NB; m++)
                if (m < m1) b -= (Lo[m][m]) * Bw[r][m];
            bnew[r] = b;
            if (live[r]) y[r][k] = b / dk;
        }
#pragma unroll
        for (int rr = NB - 1; rr > 0; rr--) {
#pragma unroll
            for (int r = 0; r < M; r++) Bw[r][rr] = Bw[r][rr - 1];
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[rr][m] = Lo[rr - 1][m];
        }
#pragma unroll
        for (int r = 0; r < M; r++) Bw[r][0] = bnew[r];
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[0][m] = ok[m];
#pragma unroll
        for (int f = 0; f < NF; f++) {
#pragma unroll
            for (int w = 0; w < NA - 1; w++)
#pragma unroll
                for (int j = 0; j < NA; j++) A[f][w][j] = A[f][w + 1][j];
            taps_at<NW>(q[f], k + 1 + NW, n1, rev[f], A[f][NA - 1]);
            int kn = k + 1 + 2 * NW;
            int kk = k + 1 + NW;
#pragma unroll
            for (int r = 0; r < M; r++) {
#pragma unroll
                for (int w = 0; w < NA - 1; w++) {
                    I[r][f][w] = I[r][f][w + 1];
                    T[r][f][w] = T[r][f][w + 1];
                }
                I[r][f][NA - 1] = (kn < n1) ? in[r * NF + f][kn] : 0.0f;
                float tn = 0.0f;
                if (kk >= NW && kk < n1 - NW) {
#pragma unroll
                    for (int j = 0; j < NA; j++) tn += A[f][NA - 1][j] * I[r][f][j];
                }
                T[r][f][NA - 1] = tn;
            }
        }
    }
    float nx[M][NB];
#pragma unroll
    for (int r = 0; r < M; r++)
#pragma unroll
        for (int m = 0; m < NB; m++) nx[r][m] = 0.0f;
    for (int k = n1 - 1; k >= 0; k--) {
        int m1 = (n1 - k - 1) < NB ? (n1 - k - 1) : NB;
        float ok[NB];
        fac_row<NB>(F, k, ok);
#pragma unroll
        for (int m = 0; m < NB; m++) ok[m] = (m < m1) ? ok[m] : 0.0f;
#pragma unroll
        for (int r = 0; r < M; r++) {
            if (!live[r]) continue;
            float t = y[r][k];
#pragma unroll
            for (int m = 0; m < NB; m++)
                if (m < m1) t -= ok[m] * nx[r][m];
#pragma unroll
            for (int m = NB - 1; m > 0; m--) nx[r][m] = nx[r][m - 1];
            nx[r][0] = t;
            out[r][k] = t;
        }
    }
}

// Adjoint prediction with a precomputed factor: s = S^{-1} in (forward and
// backward substitution), then out = A^T A s + e * boundary(s)
// (predict_step with adj, sof_cfuns.c:622-649).  y holds s.
template <int NW>
SH_HD void predict_adj_factored(const Acc &in, const Acc &q, bool rev, const Factor &F, Sink y, int n1, float e,
                                Sink out) {
    constexpr int NB = 2 * NW;
    float Lo[NB][NB], Bw[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) {
        Bw[r] = 0.0f;
#pragma unroll
        for (int m = 0; m < NB; m++) Lo[r][m] = 0.0f;
    }
    for (int k = 0; k < n1; k++) {
        int m1 = k < NB ? k : NB;
        float b = in[k];
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) b -= (Lo[m][m]) * Bw[m];
        y[k] = b / F.d[(long)k * F.fs];
        float ok[NB];
        fac_row<NB>(F, k, ok);
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
        int m1 = (n1 - k - 1) < NB ? (n1 - k - 1) : NB;
        float ok[NB];
        fac_row<NB>(F, k, ok);
#pragma unroll
        for (int m = 0; m < NB; m++)
            if (m < m1) t -= ok[m] * nx[m];
#pragma unroll
        for (int m = NB - 1; m > 0; m--) nx[m] = nx[m - 1];
        nx[0] = t;
        y[k] = t;
    }
    pred_adj_apply<NW>(Acc{y.p, y.s}, q, rev, n1, e, out);
}

// predict_adj_factored with its loads issued D samples ahead into register
// rings (loops unrolled by D; ring indices clamped into the trace and used
// only where the unclamped index is valid), and with the input formed on
// the fly as in[k] = (a[k] + b[k]) when b is given (the adjoint spray's
// running trace plus the slot it gathers, added in that order).  Same
// float operations in the same order as predict_adj_factored; y holds the
// solved trace, out may alias a.
template <int NW, int D, bool TP = false, bool MIDR = false, class IA = Acc, class FA = Factor, class SK = Sink,
          class TQ = Factor>
SH_HD void predict_adj_factored_pf(const IA &a, const IA *b, const IA &q, bool rev, const FA &F, SK y, int n1,
                                   float e, SK out, const TQ *tq = nullptr) {
    constexpr int NB = 2 * NW, NA = 2 * NW + 1;
    // forward substitution: ring of d[k], o[m][k] and the input at k
    {
        float Lo[NB][NB], Bw[NB];
#pragma unroll
        for (int r = 0; r < NB; r++) {
            Bw[r] = 0.0f;
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[r][m] = 0.0f;
        }
        float rd[D], ro[D][NB], ri[D];
        auto fetch = [&](int s, int k) {
            const long kc = k < n1 ? k : n1 - 1;
            rd[s] = fac_d(F, kc);
            fac_row<NB>(F, kc, ro[s]);
            ri[s] = b ? (a[kc] + (*b)[kc]) : a[kc];
        };
        auto step = [&](auto mid, int s, int k) {
            constexpr bool MID = decltype(mid)::value;  // k >= NB
            const int m1 = MID ? NB : (k < NB ? k : NB);
            float bb = ri[s];
#pragma unroll
            for (int m = 0; m < NB; m++)
                if (m < m1) bb -= (Lo[m][m]) * Bw[m];
            y[k] = bb / rd[s];
#pragma unroll
            for (int r = NB - 1; r > 0; r--) {
                Bw[r] = Bw[r - 1];
#pragma unroll
                for (int m = 0; m < NB; m++) Lo[r][m] = Lo[r - 1][m];
            }
            Bw[0] = bb;
#pragma unroll
            for (int m = 0; m < NB; m++) Lo[0][m] = ro[s][m];
            fetch(s, k + D);
        };
#pragma unroll
        for (int s = 0; s < D; s++) fetch(s, s);
        int k0 = 0;
        for (; k0 + D <= n1; k0 += D) {
            if (MIDR && k0 >= NB) {
#pragma unroll
                for (int s = 0; s < D; s++) step(Mid<true>{}, s, k0 + s);
            } else {
#pragma unroll
                for (int s = 0; s < D; s++) step(Mid<false>{}, s, k0 + s);
            }
        }
#pragma unroll
        for (int s = 0; s < D; s++)
            if (k0 + s < n1) step(Mid<false>{}, s, k0 + s);
    }
    // backward substitution in place: ring of y[k], o[m][k]
    {
        float nx[NB];
#pragma unroll
        for (int m = 0; m < NB; m++) nx[m] = 0.0f;
        float ry[D], rb[D][NB];
        auto fetchb = [&](int s, int k) {
            const long kc = k >= 0 ? k : 0;
            ry[s] = y[kc];
            fac_row<NB>(F, kc, rb[s]);
        };
        auto stepb = [&](auto mid, int s, int k) {
            constexpr bool MID = decltype(mid)::value;  // k <= n1 - 1 - NB
            float t = ry[s];
            const int m1 = MID ? NB : ((n1 - k - 1) < NB ? (n1 - k - 1) : NB);
#pragma unroll
            for (int m = 0; m < NB; m++)
                if (m < m1) t -= rb[s][m] * nx[m];
#pragma unroll
            for (int m = NB - 1; m > 0; m--) nx[m] = nx[m - 1];
            nx[0] = t;
            y[k] = t;
            fetchb(s, k - D);
        };
#pragma unroll
        for (int s = 0; s < D; s++) fetchb(s, n1 - 1 - s);
        int kb = n1 - 1;
        for (; kb - D + 1 >= 0; kb -= D) {
            if (MIDR && kb <= n1 - 1 - NB) {
#pragma unroll
                for (int s = 0; s < D; s++) stepb(Mid<true>{}, s, kb - s);
            } else {
#pragma unroll
                for (int s = 0; s < D; s++) stepb(Mid<false>{}, s, kb - s);
            }
        }
#pragma unroll
        for (int s = 0; s < D; s++)
            if (kb - s >= 0) stepb(Mid<false>{}, s, kb - s);
    }
    // out = A^T A s + e * boundary(s) (pred_adj_apply): ring of the slope
    // at k + 1 + NW and the solved sample at k + 1 + 2 NW
    {
        const auto s = sink_acc(y);
        float A[NA][NA], S[NA], T[NA];
#pragma unroll
        for (int w = 0; w < NA; w++) {
            if constexpr (TP) {
                float bt[NA];
                tap_row<NW>(*tq, (w - NW) >= 0 && (w - NW) < n1 ? w - NW : 0, bt);
                taps_stored<NW>(bt, w - NW, n1, rev, A[w]);
            } else {
                taps_at<NW>(q, w - NW, n1, rev, A[w]);
            }
        }
#pragma unroll
        for (int w = 0; w < NA; w++) S[w] = (w < n1) ? s[w] : 0.0f;
#pragma unroll
        for (int w = 0; w < NA; w++) {
            const int kk = w - NW;
            float t = 0.0f;
            if (kk >= NW && kk < n1 - NW) {
#pragma unroll
                for (int j = NA - 1; j >= 0; j--) t += A[w][j] * s[kk - j + NW];
            }
            T[w] = t;
        }
        const float s0 = s[0], s1 = s[1], s2 = s[n1 - 2], s3 = s[n1 - 1];
        constexpr int NQ = TP ? NA : 1;
        float rq[D][NQ], rs[D];
        auto fetcha = [&](int r, int k) {
            const int kq = k + 1 + NW < n1 ? k + 1 + NW : n1 - 1;
            const int ks = k + 1 + 2 * NW < n1 ? k + 1 + 2 * NW : n1 - 1;
            if constexpr (TP) tap_row<NW>(*tq, kq, rq[r]);
            else rq[r][0] = q[kq];
            rs[r] = s[ks];
        };
        // MID: k in [2 NW, n1 - 2 NW - 1), every boundary test known
        auto stepa = [&](auto mid, int r, int k) {
            constexpr bool MID = decltype(mid)::value;
            float v = 0.0f;
#pragma unroll
            for (int w = 0; w < NA; w++) {
                const int i = k - NW + w;
                if (MID || (i >= NW && i < n1 - NW)) v += A[w][2 * NW - w] * T[w];
            }
            if (!MID) {
                if (k == 0) v += e * s0;
                if (k == 1) v += e * s1;
                if (k == n1 - 2) v += e * s2;
                if (k == n1 - 1) v += e * s3;
            }
            out[k] = v;
#pragma unroll
            for (int w = 0; w < NA - 1; w++) {
#pragma unroll
                for (int j = 0; j < NA; j++) A[w][j] = A[w + 1][j];
                T[w] = T[w + 1];
                S[w] = S[w + 1];
            }
            if constexpr (TP) {
                if (MID) {
                    tap_order<NA>(rq[r], rev, A[NA - 1]);
                } else {
                    taps_stored<NW>(rq[r], k + 1 + NW, n1, rev, A[NA - 1]);
                }
            } else {
                taps_from<NW>(rq[r][0], k + 1 + NW, n1, rev, A[NA - 1]);
            }
            const int kn = k + 1 + 2 * NW;
            S[NA - 1] = (MID || kn < n1) ? rs[r] : 0.0f;
            const int kk = k + 1 + NW;
            float tn = 0.0f;
            if (MID || (kk >= NW && kk < n1 - NW)) {
#pragma unroll
                for (int j = NA - 1; j >= 0; j--) tn += A[NA - 1][j] * S[2 * NW - j];
            }
            T[NA - 1] = tn;
            fetcha(r, k + D);
        };
#pragma unroll
        for (int r = 0; r < D; r++) fetcha(r, r);
        int k0 = 0;
        for (; k0 + D <= n1; k0 += D) {
            if (MIDR && k0 >= 2 * NW && k0 + D <= n1 - 2 * NW - 1) {
#pragma unroll
                for (int r = 0; r < D; r++) stepa(Mid<true>{}, r, k0 + r);
            } else {
#pragma unroll
                for (int r = 0; r < D; r++) stepa(Mid<false>{}, r, k0 + r);
            }
        }
#pragma unroll
        for (int r = 0; r < D; r++)
            if (k0 + r < n1) stepa(Mid<false>{}, r, k0 + r);
    }
}
