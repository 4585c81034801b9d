/*
 * seistr_oracle.c -- CPU restatement of pyseistr's structural hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is
 * checked against; it is compiled twice (REAL=float -> liboracle_f32.so, the
 * FP32 restatement that follows the reference C extension operation by
 * operation; REAL=double -> liboracle_f64.so, the FP64 oracle).  Only tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.
 * The product path (pyseistr_amd) never calls it.
 *
 * Every routine cites the reference file:line whose algorithm it restates
 * (paths relative to the reference repository's pyseistr/src/).  Where the
 * FP32 build must reproduce the reference's rounding, the order of float
 * operations (and which of them run in double) is kept the same; the code
 * layout, data structures and control flow are this project's own
 * (stateless functions, explicit workspaces, no module globals).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

#define EXPORT __attribute__((visibility("default")))

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}

/* ------------------------------------------------------------------ */
/* Maximally-flat all-pass (PWD) filter coefficients.                  */
/* dip_cfuns.c:833-910 (apfilt_init / passfilter / aderfilter).        */
/* ------------------------------------------------------------------ */

/* Normalisation table b[0..2nw] (dip_cfuns.c:836-856). */
static void apf_table(int nw, double *b) {
    int nf = 2 * nw;
    for (int k = 0; k <= nf; k++) {
        double v = 1.0;
        for (int j = 0; j < nf; j++)
            v *= (j < nf - k) ? (k + j + 1.0) / (2 * (2 * j + 1) * (j + 1))
                              : 1.0 / (2 * (2 * j + 1));
        b[k] = v;
    }
}

/* Factor j of coefficient k: (nf-j-p) for j < nf-k, else (p+j+1); both are
 * evaluated in the working precision before promotion (dip_cfuns.c:872-877). */
static inline double apf_factor(int nf, int k, int j, real p) {
    return (j < nf - k) ? (double)((real)(nf - j) - p)
                        : (double)((p + (real)j) + (real)1);
}

static void apf_pass(int nw, const double *b, real p, real *a) {
    int nf = 2 * nw;
    for (int k = 0; k <= nf; k++) {
        double v = b[k];
        for (int j = 0; j < nf; j++) v *= apf_factor(nf, k, j, p);
        a[k] = (real)v;
    }
}

/* Minus the p-derivative of apf_pass (dip_cfuns.c:883-910): the product rule
 * with each factor in turn replaced by its derivative, accumulated in double. */
static void apf_der(int nw, const double *b, real p, real *a) {
    int nf = 2 * nw;
    for (int k = 0; k <= nf; k++) {
        double acc = 0.0;
        for (int i = 0; i < nf; i++) {
            double term = -1.0;
            for (int j = 0; j < nf; j++) {
                if (j != i) term *= apf_factor(nf, k, j, p);
                else if (j < nf - k) term *= -1;
            }
            acc += term;
        }
        a[k] = (real)(acc * b[k]);
    }
}

EXPORT void or_passfilter(int nw, real p, real *a) {
    double b[16];
    apf_table(nw, b);
    apf_pass(nw, b, p, a);
}

EXPORT void or_aderfilter(int nw, real p, real *a) {
    double b[16];
    apf_table(nw, b);
    apf_der(nw, b, p, a);
}

static inline int nint_ref(real x) { /* PS_NINT, dip_cfuns.c:9 */
    return (int)((x) > 0.0 ? (x) + 0.5 : (x) - 0.5);
}

/* ------------------------------------------------------------------ */
/* Plane-wave destruction along axis 2 (stride n1) or axis 3 (stride   */
/* n1*n2).  allpass1 / allpass2 (dip_cfuns.c:1135-1200, 1399-1465).    */
/* y[i] = sum_iw (x[i+is+ip] - x[i-is-id]) * a_iw(p[i]-id), is=(iw-nw)*nj */
/* ------------------------------------------------------------------ */
EXPORT void or_pwd(int n1, int n2, int n3, int axis, int nw, int nj, int drift,
                   int der, const real *x, const real *p, real *y) {
    double b[16];
    real a[16];
    apf_table(nw, b);
    size_t n = (size_t)n1 * n2 * n3;
    for (size_t i = 0; i < n; i++) y[i] = 0;
    int ip = (axis == 2) ? n1 : n1 * n2;
    int i2hi = (axis == 2) ? n2 - 1 : n2;
    int i3hi = (axis == 2) ? n3 : n3 - 1;
    int h = nw * nj;
    for (int i3 = 0; i3 < i3hi; i3++)
        for (int i2 = 0; i2 < i2hi; i2++)
            for (int i1 = h; i1 < n1 - h; i1++) {
                size_t i = i1 + (size_t)n1 * (i2 + (size_t)n2 * i3);
                int id = 0;
                real pv = p[i];
                if (drift) {
                    id = nint_ref(p[i]);
                    if (i1 - h - id < 0 || i1 + h - id >= n1) continue;
                    pv = p[i] - (real)id;
                }
                if (der) apf_der(nw, b, pv, a);
                else apf_pass(nw, b, pv, a);
                real acc = y[i];
                for (int iw = 0; iw <= 2 * nw; iw++) {
                    int is = (iw - nw) * nj;
                    acc += (x[i + is + ip] - x[i - is - id]) * a[iw];
                }
                y[i] = acc;
            }
}

/* Stacked inline+crossline PWD operator [A1; A2] (N -> 2N) and its adjoint.
 * allpass3_lop (soint3d_cfuns.c:625-731).  Adjoint scatters in the same
 * sample order as the reference. */
EXPORT void or_allpass3_lop(int adj, int add, int n1, int n2, int n3, int nw,
                            int nj1, int nj2, int drift, const real *pp,
                            const real *qq, real *x, real *y) {
    double b[16];
    real a[16];
    apf_table(nw, b);
    size_t n = (size_t)n1 * n2 * n3;
    if (!add) {
        if (adj) for (size_t i = 0; i < n; i++) x[i] = 0;
        else for (size_t i = 0; i < 2 * n; i++) y[i] = 0;
    }
    for (int part = 0; part < 2; part++) {
        const real *pr = part ? qq : pp;
        int nj = part ? nj2 : nj1;
        int ip = part ? n1 * n2 : n1;
        int i2hi = part ? n2 : n2 - 1;
        int i3hi = part ? n3 - 1 : n3;
        real *yy = y + (part ? n : 0);
        int h = nw * nj;
        for (int i3 = 0; i3 < i3hi; i3++)
            for (int i2 = 0; i2 < i2hi; i2++)
                for (int i1 = h; i1 < n1 - h; i1++) {
                    size_t i = i1 + (size_t)n1 * (i2 + (size_t)n2 * i3);
                    int id = 0;
                    real pv = pr[i];
                    if (drift) {
                        id = nint_ref(pr[i]);
                        if (i1 - h - id < 0 || i1 + h - id >= n1) continue;
                        pv = pr[i] - (real)id;
                    }
                    apf_pass(nw, b, pv, a);
                    for (int iw = 0; iw <= 2 * nw; iw++) {
                        int is = (iw - nw) * nj;
                        if (adj) {
                            x[i + ip + is] += yy[i] * a[iw];
                            x[i - is - id] -= yy[i] * a[iw];
                        } else {
                            yy[i] += (x[i + ip + is] - x[i - is - id]) * a[iw];
                        }
                    }
                }
    }
}

/* 2-D single-dip PWD operator (N -> N, last trace row zero) and adjoint.
 * allpass21_lop (soint2d_cfuns.c:296-344). */
EXPORT void or_allpass21_lop(int adj, int add, int n1, int n2, int nw, int nj,
                             int drift, const real *pp, real *x, real *y) {
    double b[16];
    real a[16];
    apf_table(nw, b);
    size_t n = (size_t)n1 * n2;
    if (!add) {
        if (adj) for (size_t i = 0; i < n; i++) x[i] = 0;
        else for (size_t i = 0; i < n; i++) y[i] = 0;
    }
    int h = nw * nj;
    for (int i2 = 0; i2 < n2 - 1; i2++)
        for (int i1 = h; i1 < n1 - h; i1++) {
            size_t i = i1 + (size_t)i2 * n1;
            int id = 0;
            real pv = pp[i];
            if (drift) {
                id = nint_ref(pp[i]);
                if (i1 - h - id < 0 || i1 + h - id >= n1) continue;
                pv = pp[i] - (real)id;
            }
            apf_pass(nw, b, pv, a);
            for (int iw = 0; iw <= 2 * nw; iw++) {
                int is = (iw - nw) * nj;
                if (adj) {
                    x[i + is + n1] += y[i] * a[iw];
                    x[i - is - id] -= y[i] * a[iw];
                } else {
                    y[i] += (x[i + is + n1] - x[i - is - id]) * a[iw];
                }
            }
        }
}

/* ------------------------------------------------------------------ */
/* Triangle smoothing (adjoint form, used for both S and S^T).         */
/* ps_smooth2 = triple2 -> doubint2 -> fold2 (dip_cfuns.c:431-625).    */
/* ------------------------------------------------------------------ */

/* Even-reflection fold of extended position q (original index, may be
 * outside [0,nx)) back onto [0,nx); period 2nx (fold2, dip_cfuns.c:458-484). */
static inline int fold_index(int q, int nx) {
    if (q < 0) q = -1 - q;
    int m = q % (2 * nx);
    return (m < nx) ? m : 2 * nx - 1 - m;
}

/* One line: x[o + i*d], i < nx, smoothed in place with radius nb.
 * tmp must hold nx + 2*nb values. */
static void tri_smooth2_line(int nb, int nx, real *x, size_t o, size_t d, real *tmp) {
    int np = nx + 2 * nb;
    real wt = (real)(1.0 / (nb * nb));            /* dip_cfuns.c:421 */
    real mw = -wt, w2 = (real)(2. * wt);
    for (int i = 0; i < np; i++) tmp[i] = 0;
    /* triple2: three scaled copies at offsets 0, nb, 2nb (dip_cfuns.c:569-586) */
    for (int i = 0; i < nx; i++) tmp[i] += mw * x[o + i * d];
    for (int i = 0; i < nx; i++) tmp[i + nb] += w2 * x[o + i * d];
    for (int i = 0; i < nx; i++) tmp[i + 2 * nb] += mw * x[o + i * d];
    /* doubint2: running sum forward, then backward (dip_cfuns.c:509-530) */
    real t = 0;
    for (int i = 0; i < np; i++) { t += tmp[i]; tmp[i] = t; }
    t = 0;
    for (int i = np - 1; i >= 0; i--) { t += tmp[i]; tmp[i] = t; }
    /* fold2: middle, then right tail ascending, then left tail descending */
    for (int i = 0; i < nx; i++) x[o + i * d] = tmp[i + nb];
    for (int e = nb + nx; e < np; e++) x[o + (size_t)fold_index(e - nb, nx) * d] += tmp[e];
    for (int e = nb - 1; e >= 0; e--) x[o + (size_t)fold_index(e - nb, nx) * d] += tmp[e];
}

/* Smoothing along every axis whose radius exceeds 1, axis 1 first
 * (ps_trianglen_lop, dip_cfuns.c:672-712). In place. */
EXPORT void or_trianglen(int n1, int n2, int n3, int r1, int r2, int r3, real *x) {
    int n[3] = {n1, n2, n3}, r[3] = {r1, r2, r3};
    size_t s[3] = {1, (size_t)n1, (size_t)n1 * n2};
    int mx = n1 > n2 ? n1 : n2;
    if (n3 > mx) mx = n3;
    int rmax = r1 > r2 ? r1 : r2;
    if (r3 > rmax) rmax = r3;
    real *tmp = xcalloc(mx + 2 * rmax + 2, sizeof(real));
    for (int ax = 0; ax < 3; ax++) {
        if (r[ax] <= 1) continue;
        /* iterate over the lines orthogonal to axis ax */
        for (int a = 0; a < (ax == 0 ? n2 : n1); a++)
            for (int c = 0; c < (ax == 2 ? n2 : n3); c++) {
                size_t o;
                if (ax == 0) o = (size_t)a * s[1] + (size_t)c * s[2];
                else if (ax == 1) o = (size_t)a + (size_t)c * s[2];
                else o = (size_t)a + (size_t)c * s[1];
                tri_smooth2_line(r[ax], n[ax], x, o, s[ax], tmp);
            }
    }
    free(tmp);
}

/* ------------------------------------------------------------------ */
/* Stand-alone triangle / box smoothing with derivative and repeat:    */
/* smoothcf (dip_cfuns.c:2006-2123) over ps_smooth / ps_smooth2        */
/* (dip_cfuns.c:594-625) with fold, doubint, triple and their adjoints */
/* (433-592); the object's weight from ps_triangle_init (403-427).     */
/* ------------------------------------------------------------------ */

/* adj != 0: ps_smooth = fold -> doubint -> triple; adj == 0: ps_smooth2 =
 * triple2 -> doubint2 -> fold2.  der: derivative; box: box, not triangle. */
static void smooth_line(int nb, int nx, int box, int der, int adj, real *x, size_t o, size_t d, real *tmp) {
    int np = nx + 2 * nb;
    real wt = box ? (real)(1.0 / (2 * nb - 1)) : (real)(1.0 / (nb * nb)); /* 420-424 */
    int one = box || der;                                               /* doubint stops after one sum */
    if (adj) {
        /* fold (433-456): the even-reflection extension, a plain copy */
        for (int e = 0; e < np; e++) tmp[e] = x[o + (size_t)fold_index(e - nb, nx) * d];
        /* doubint (486-507): backward running sum, then forward */
        real t = 0;
        for (int i = np - 1; i >= 0; i--) { t += tmp[i]; tmp[i] = t; }
        if (!one) {
            t = 0;
            for (int i = 0; i < np; i++) { t += tmp[i]; tmp[i] = t; }
        }
        /* triple (532-552): the triangle case is evaluated in double */
        for (int i = 0; i < nx; i++) {
            if (box) x[o + i * d] = (tmp[i + 1] - tmp[i + 2 * nb]) * wt;
            else x[o + i * d] = (real)((2. * tmp[i + nb] - tmp[i] - tmp[i + 2 * nb]) * wt);
        }
    } else {
        /* triple2 (567-586): scaled copies added into zeros, in saxpy order */
        for (int i = 0; i < np; i++) tmp[i] = 0;
        if (box) {
            real mw = -wt;
            for (int i = 0; i < nx; i++) tmp[i + 1] += wt * x[o + i * d];
            for (int i = 0; i < nx; i++) tmp[i + 2 * nb] += mw * x[o + i * d];
        } else {
            real mw = -wt, w2 = (real)(2. * wt);
            for (int i = 0; i < nx; i++) tmp[i] += mw * x[o + i * d];
            for (int i = 0; i < nx; i++) tmp[i + nb] += w2 * x[o + i * d];
            for (int i = 0; i < nx; i++) tmp[i + 2 * nb] += mw * x[o + i * d];
        }
        /* doubint2 (509-530): forward running sum, then backward */
        real t = 0;
        for (int i = 0; i < np; i++) { t += tmp[i]; tmp[i] = t; }
        if (!one) {
            t = 0;
            for (int i = np - 1; i >= 0; i--) { t += tmp[i]; tmp[i] = t; }
        }
        /* fold2 (458-484): middle, right tail ascending, left tail descending */
        for (int i = 0; i < nx; i++) x[o + i * d] = tmp[i + nb];
        for (int e = nb + nx; e < np; e++) x[o + (size_t)fold_index(e - nb, nx) * d] += tmp[e];
        for (int e = nb - 1; e >= 0; e--) x[o + (size_t)fold_index(e - nb, nx) * d] += tmp[e];
    }
}

/* smoothcf's loops (2090-2110): every axis whose radius exceeds 1, axis 1
 * first; each line is smoothed `repeat` times before the next.  (The
 * reference's reshaping of n1/n2 by dim1 only groups the same lines.)
 * In place. */
EXPORT void or_smoothc(int n1, int n2, int n3, int repeat, int adj, const int *rect, const int *diff,
                       const int *box, real *x) {
    int n[3] = {n1, n2, n3};
    size_t s[3] = {1, (size_t)n1, (size_t)n1 * n2};
    int mx = 0, rmax = 0;
    for (int a = 0; a < 3; a++) {
        if (n[a] > mx) mx = n[a];
        if (rect[a] > rmax) rmax = rect[a];
    }
    real *tmp = xcalloc(mx + 2 * rmax + 2, sizeof(real));
    for (int ax = 0; ax < 3; ax++) {
        if (rect[ax] <= 1) continue;
        for (int a = 0; a < (ax == 0 ? n2 : n1); a++)
            for (int c = 0; c < (ax == 2 ? n2 : n3); c++) {
                size_t o;
                if (ax == 0) o = (size_t)a * s[1] + (size_t)c * s[2];
                else if (ax == 1) o = (size_t)a + (size_t)c * s[2];
                else o = (size_t)a + (size_t)c * s[1];
                for (int r = 0; r < repeat; r++) smooth_line(rect[ax], n[ax], box[ax], diff[ax], adj, x, o, s[ax], tmp);
            }
    }
    free(tmp);
}

/* ------------------------------------------------------------------ */
/* Butterworth band-pass per trace (cbp, bp_cfuns.c:524-645):          */
/* coefficients (sf_butter_init, 420-462), recursive first- and        */
/* second-order sections (sf_butter_apply, 473-507); zero phase runs   */
/* each filter again over the reversed trace (sf_reverse, 509-522),    */
/* restated as a backward pass (the same operations in the same order).*/
/* ------------------------------------------------------------------ */
#define BP_PI (3.14159265358979323846264338328)
#define BP_MAXSEC 64
typedef struct {
    int low, nn;
    real den[BP_MAXSEC][2], mid;
} butter_t;

static real bp_sin(real a) { return sizeof(real) == 4 ? (real)sinf((float)a) : (real)sin((double)a); }
static real bp_cos(real a) { return sizeof(real) == 4 ? (real)cosf((float)a) : (real)cos((double)a); }

static void butter_init(butter_t *bw, int low, real cutoff, int nn) {
    real arg, ss, sinw, cosw, fact;
    arg = 2. * BP_PI * cutoff;
    sinw = bp_sin(arg);
    cosw = bp_cos(arg);
    bw->nn = nn;
    bw->low = low;
    if (nn % 2) {
        if (low) {
            fact = (1. + cosw) / sinw;
            bw->den[nn / 2][0] = 1. / (1. + fact);
            bw->den[nn / 2][1] = 1. - fact;
        } else {
            fact = sinw / (1. + cosw);
            bw->den[nn / 2][0] = 1. / (fact + 1.);
            bw->den[nn / 2][1] = fact - 1.;
        }
    }
    fact = low ? bp_sin(0.5 * arg) : bp_cos(0.5 * arg);
    fact *= fact;
    for (int j = 0; j < nn / 2; j++) {
        ss = bp_sin(BP_PI * (2 * j + 1) / (2 * nn)) * sinw;
        bw->den[j][0] = fact / (1. + ss);
        bw->den[j][1] = (1. - ss) / fact;
    }
    bw->mid = -2. * cosw / fact;
}

/* one pass of every section over x[0..nx) (dir = +1) or backwards (-1) */
static void butter_apply(const butter_t *bw, int nx, real *x, int dir) {
    real d0, d1 = bw->mid, d2, x0, x1, x2, y0, y1, y2;
    int nn = bw->nn;
    if (nn % 2) {
        d0 = bw->den[nn / 2][0];
        d2 = bw->den[nn / 2][1];
        x0 = y1 = 0.;
        for (int k = 0; k < nx; k++) {
            int ix = dir > 0 ? k : nx - 1 - k;
            x1 = x0;
            x0 = x[ix];
            y0 = bw->low ? (x0 + x1 - d2 * y1) * d0 : (x0 - x1 - d2 * y1) * d0;
            x[ix] = y1 = y0;
        }
    }
    for (int j = 0; j < nn / 2; j++) {
        d0 = bw->den[j][0];
        d2 = bw->den[j][1];
        x1 = x0 = y1 = y2 = 0.;
        for (int k = 0; k < nx; k++) {
            int ix = dir > 0 ? k : nx - 1 - k;
            x2 = x1;
            x1 = x0;
            x0 = x[ix];
            y0 = bw->low ? (x0 + 2 * x1 + x2 - d1 * y1 - d2 * y2) * d0 : (x0 - 2 * x1 + x2 - d1 * y1 - d2 * y2) * d0;
            y2 = y1;
            x[ix] = y1 = y0;
        }
    }
}

/* cbp's parameter handling (582-606) and trace loop (615-637); in place
 * on ntr traces of n1 samples.  Returns 0, or -1 if a pole count exceeds
 * what this restatement holds. */
EXPORT int or_bandpass(int n1, int ntr, real d1, real flo, real fhi, int nplo, int nphi, int phase, real *x) {
    const float eps = 0.0001;
    butter_t blo, bhi;
    if (!(flo < 0)) flo = flo * d1;
    if (!(fhi < 0)) fhi = fhi * d1;
    if (nplo < 1) nplo = 1;
    if (nplo > 1 && !phase) nplo /= 2;
    if (nphi < 1) nphi = 1;
    if (nphi > 1 && !phase) nphi /= 2;
    if ((nplo + 1) / 2 > BP_MAXSEC || (nphi + 1) / 2 > BP_MAXSEC) return -1;
    int has_lo = flo > eps, has_hi = fhi < 0.5 - eps;
    if (has_lo) butter_init(&blo, 0, flo, nplo);
    if (has_hi) butter_init(&bhi, 1, fhi, nphi);
    for (long i2 = 0; i2 < ntr; i2++) {
        real *tr = x + i2 * (long)n1;
        if (has_lo) {
            butter_apply(&blo, n1, tr, 1);
            if (!phase) butter_apply(&blo, n1, tr, -1);
        }
        if (has_hi) {
            butter_apply(&bhi, n1, tr, 1);
            if (!phase) butter_apply(&bhi, n1, tr, -1);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Reductions with the reference's accumulation types.                 */
/* ps_cblas_dsdot (double), ps_cblas_snrm2 (float sum of squares).     */
/* dip_cfuns.c:122-209.                                                */
/* ------------------------------------------------------------------ */
static double dsdot(size_t n, const real *x, const real *y) {
    double s = 0.;
    for (size_t i = 0; i < n; i++) s += (double)x[i] * y[i];
    return s;
}
static real snrm2(size_t n, const real *x) {
    real s = 0;
    for (size_t i = 0; i < n; i++) s += x[i] * x[i];
    return s;
}

/* ------------------------------------------------------------------ */
/* Shaping-regularised conjugate gradients (Fomel).                    */
/* ps_conjgrad (dip_cfuns.c:257-383; soint3d_cfuns.c:1222-1349).       */
/* The data operator L and shaping operator S are callbacks; prec is   */
/* never used on this path and is omitted.                             */
/* ------------------------------------------------------------------ */
typedef struct {
    /* y (+)= L x  or  x (+)= L^T y */
    void (*oper)(void *ctx, int adj, int add, real *x, real *y);
    /* gx (+)= S gp  or  gp (+)= S^T gx */
    void (*shape)(void *ctx, int adj, int add, real *p, real *x);
    void *octx, *sctx;
} cg_ops;

typedef struct { int iters; int converged; } cg_report;

static void shaping_cg(const cg_ops *ops, size_t n, real eps1, real tol, int hasp0,
                       real *p, real *x, const real *dat, int niter, cg_report *rep) {
    real eps = eps1 * eps1;                           /* dip_cfuns.c:231 */
    real *r = xcalloc(n, sizeof(real)), *sp = xcalloc(n, sizeof(real));
    real *gp = xcalloc(n, sizeof(real)), *sx = xcalloc(n, sizeof(real));
    real *gx = xcalloc(n, sizeof(real)), *sr = xcalloc(n, sizeof(real));
    real *gr = xcalloc(n, sizeof(real));
    double gn, gnp = 0., g0 = 0., alpha, beta;
    if (rep) { rep->iters = 0; rep->converged = 0; }
    for (size_t i = 0; i < n; i++) r[i] = -dat[i];
    if (hasp0) {
        ops->shape(ops->sctx, 0, 0, p, x);
        ops->oper(ops->octx, 0, 1, x, r);
    } else {
        for (size_t i = 0; i < n; i++) { p[i] = 0; x[i] = 0; }
    }
    if (dsdot(n, r, r) == 0.) goto done;              /* dip_cfuns.c:301-304 */
    for (int it = 0; it < niter; it++) {
        for (size_t i = 0; i < n; i++) gp[i] = eps * p[i];
        for (size_t i = 0; i < n; i++) gx[i] = -eps * x[i];
        ops->oper(ops->octx, 1, 1, gx, r);
        ops->shape(ops->sctx, 1, 1, gp, gx);
        ops->shape(ops->sctx, 0, 0, gp, gx);
        ops->oper(ops->octx, 0, 0, gx, gr);
        gn = dsdot(n, gp, gp);
        if (it == 0) {
            g0 = gn;
            memcpy(sp, gp, n * sizeof(real));
            memcpy(sx, gx, n * sizeof(real));
            memcpy(sr, gr, n * sizeof(real));
        } else {
            alpha = gn / gnp;
            double dg = gn / g0;
            if (alpha < tol || dg < tol) {            /* dip_cfuns.c:349-354 */
                if (rep) rep->converged = 1;
                break;
            }
            real a = (real)alpha;
            for (size_t i = 0; i < n; i++) sp[i] = gp[i] + a * sp[i];
            for (size_t i = 0; i < n; i++) sx[i] = gx[i] + a * sx[i];
            for (size_t i = 0; i < n; i++) sr[i] = gr[i] + a * sr[i];
        }
        beta = dsdot(n, sr, sr) + eps * (dsdot(n, sp, sp) - dsdot(n, sx, sx));
        alpha = -gn / beta;
        real a = (real)alpha;
        for (size_t i = 0; i < n; i++) p[i] += a * sp[i];
        for (size_t i = 0; i < n; i++) x[i] += a * sx[i];
        for (size_t i = 0; i < n; i++) r[i] += a * sr[i];
        gnp = gn;
        if (rep) rep->iters = it + 1;
    }
done:
    free(r); free(sp); free(gp); free(sx); free(gx); free(sr); free(gr);
}

/* Diagonal weight operator (ps_weight_lop, dip_cfuns.c:740-758). */
typedef struct { const real *w; size_t n; } weight_ctx;
static void weight_op(void *c, int adj, int add, real *x, real *y) {
    weight_ctx *wc = c;
    if (adj) { for (size_t i = 0; i < wc->n; i++) x[i] = (add ? x[i] : 0) + y[i] * wc->w[i]; }
    else     { for (size_t i = 0; i < wc->n; i++) y[i] = (add ? y[i] : 0) + x[i] * wc->w[i]; }
}

/* N-D triangle shaping operator (ps_trianglen_lop, dip_cfuns.c:672-712). */
typedef struct { int n1, n2, n3, r1, r2, r3; real *tmp; } tri_ctx;
static void tri_op(void *c, int adj, int add, real *x, real *y) {
    tri_ctx *t = c;
    size_t n = (size_t)t->n1 * t->n2 * t->n3;
    const real *src = adj ? y : x;
    real *dst = adj ? x : y;
    memcpy(t->tmp, src, n * sizeof(real));
    or_trianglen(t->n1, t->n2, t->n3, t->r1, t->r2, t->r3, t->tmp);
    for (size_t i = 0; i < n; i++) dst[i] = (add ? dst[i] : 0) + t->tmp[i];
}

/* Smooth division rat = num/den with shaping; num and den are modified
 * in place (ps_divne, dip_cfuns.c:796-827). */
static void divne(int n1, int n2, int n3, int r1, int r2, int r3, int niter,
                  real *num, real *den, real *rat, real eps, cg_report *rep) {
    size_t n = (size_t)n1 * n2 * n3;
    if (eps > 0.0f) {
        for (size_t i = 0; i < n; i++) {
            double norm = 1.0 / hypot(den[i], eps);
            num[i] *= norm;
            den[i] *= norm;
        }
    }
    double norm = dsdot(n, den, den);
    if (norm == 0.0) {
        for (size_t i = 0; i < n; i++) rat[i] = 0;
        if (rep) { rep->iters = -1; rep->converged = 0; }
        return;
    }
    norm = sqrt(n / norm);
    for (size_t i = 0; i < n; i++) { num[i] *= norm; den[i] *= norm; }
    real *pbuf = xcalloc(n, sizeof(real));
    weight_ctx wc = {den, n};
    tri_ctx tc = {n1, n2, n3, r1, r2, r3, xcalloc(n, sizeof(real))};
    cg_ops ops = {weight_op, tri_op, &wc, &tc};
    shaping_cg(&ops, n, (real)1., (real)1.e-6, 0, pbuf, rat, num, niter, rep);
    free(tc.tmp);
    free(pbuf);
}

/* ------------------------------------------------------------------ */
/* Dip estimation: Gauss-Newton with backtracking line search.         */
/* dip3 (dip_cfuns.c:1619-1691) and the dipc driver (1694-1989).       */
/* ------------------------------------------------------------------ */

/* mask32 (dip_cfuns.c:914-966, both=false): a sample is masked if any
 * stencil tap touches a zero of the mask data.  m0 inline, m1 crossline. */
static void dip_masks(int nw, int n1, int n2, int n3, const real *yy,
                      unsigned char *m0, unsigned char *m1) {
    size_t n = (size_t)n1 * n2 * n3;
    for (size_t i = 0; i < n; i++) { m0[i] = 0; m1[i] = 0; }
    for (int i3 = 0; i3 < n3; i3++)
        for (int i2 = 0; i2 < n2 - 1; i2++)
            for (int i1 = nw; i1 < n1 - nw; i1++) {
                size_t i = i1 + (size_t)n1 * (i2 + (size_t)n2 * i3);
                for (int iw = 0; iw <= 2 * nw; iw++) {
                    int is = iw - nw;
                    if (yy[i - is] == 0. || yy[i + n1 + is] == 0.) m0[i] = 1;
                }
            }
    for (int i3 = 0; i3 < n3 - 1; i3++)
        for (int i2 = 0; i2 < n2; i2++)
            for (int i1 = nw; i1 < n1 - nw; i1++) {
                size_t i = i1 + (size_t)n1 * (i2 + (size_t)n2 * i3);
                for (int iw = 0; iw <= 2 * nw; iw++) {
                    int is = iw - nw;
                    if (yy[i - is] == 0. || yy[i + (size_t)n1 * n2 + is] == 0.) m1[i] = 1;
                }
            }
}

/* One direction of dip3.  rep (optional) receives, per GN iteration,
 * [cg iterations, accepted line-search step k (8 = none accepted)]. */
static void dip_gn(int axis, int niter, int liter, int nw, int n1, int n2, int n3,
                   const int *rect, real eps, const real *u, real *p,
                   const unsigned char *mask, real pmin, real pmax, int *rep) {
    size_t n = (size_t)n1 * n2 * n3;
    real *u1 = xcalloc(n, sizeof(real)), *u2 = xcalloc(n, sizeof(real));
    real *dp = xcalloc(n, sizeof(real)), *p0 = xcalloc(n, sizeof(real));
    or_pwd(n1, n2, n3, axis, nw, 1, 0, 0, u, p, u2);
    for (int it = 0; it < niter; it++) {
        or_pwd(n1, n2, n3, axis, nw, 1, 0, 1, u, p, u1);
        real usum = 0;
        for (size_t i = 0; i < n; i++) { p0[i] = p[i]; usum += u2[i] * u2[i]; }
        if (mask)
            for (size_t i = 0; i < n; i++) if (mask[i]) { u1[i] = 0; u2[i] = 0; }
        cg_report cr;
        divne(n1, n2, n3, rect[0], rect[1], rect[2], liter, u2, u1, dp, eps, &cr);
        real lam = 1.;
        int k;
        for (k = 0; k < 8; k++) {
            for (size_t i = 0; i < n; i++) {
                real pi = p0[i] + lam * dp[i];
                if (pi < pmin) pi = pmin;
                if (pi > pmax) pi = pmax;
                p[i] = pi;
            }
            or_pwd(n1, n2, n3, axis, nw, 1, 0, 0, u, p, u2);
            real usum2 = 0;
            for (size_t i = 0; i < n; i++) usum2 += u2[i] * u2[i];
            if (usum2 < usum) break;
            lam *= 0.5;
        }
        if (rep) { rep[2 * it] = cr.iters; rep[2 * it + 1] = k; }
    }
    free(u1); free(u2); free(dp); free(p0);
}

/* dipc: inline dip for n3==1, otherwise inline then crossline (output 2N).
 * eps_dv is accepted and has no effect, as in the reference: dip3_init
 * stores it in the file-scope `eps` (dip_cfuns.c:1582,1596), which is the
 * same object as the conjugate-gradient `eps` (214) that ps_conjgrad_init
 * then sets to 1*1 (231, via ps_divn_init 777); ps_divne receives that
 * shared value (1665).  The smooth division therefore always runs with 1.
 * mask (optional, length N) is the mask data whose zeros disable samples.
 * p_init (optional, N or 2N) replaces the reference's zero initial dip.
 * rep (optional) gets 2*niter ints per direction. */
EXPORT int or_dip(const real *din, const real *mask, int n1, int n2, int n3,
                  int niter, int liter, int order, real eps_dv, int r1, int r2,
                  int r3, const real *p_init, real *dip, int *rep) {
    size_t n = (size_t)n1 * n2 * n3;
    int rect[3] = {r1, r2, r3};
    int ndir = (n3 == 1) ? 1 : 2;
    (void)eps_dv;
    real pmax = (real)3.402823466e+38f, pmin = -pmax;   /* dip_cfuns.c:1783-1786 */
    unsigned char *m0 = NULL, *m1 = NULL;
    if (mask) {
        m0 = xcalloc(n, 1); m1 = xcalloc(n, 1);
        dip_masks(order, n1, n2, n3, mask, m0, m1);
    }
    for (int d = 0; d < ndir; d++) {
        real *p = dip + d * n;
        for (size_t i = 0; i < n; i++) p[i] = p_init ? p_init[d * n + i] : 0;
        dip_gn(d == 0 ? 2 : 3, niter, liter, order, n1, n2, n3, rect, (real)1., din, p,
               d == 0 ? m0 : m1, pmin, pmax, rep ? rep + d * 2 * niter : NULL);
    }
    free(m0); free(m1);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Banded SPD LDL^T solver, band = 2nw (sof_cfuns.c:142-185, 287-307). */
/* ------------------------------------------------------------------ */
typedef struct { int n, band; real *d; real *o[8]; } bands;

static void bands_define(bands *s, const real *diag, real *const *offd) {
    int n = s->n, nb = s->band;
    for (int k = 0; k < n; k++) {
        real t = diag[k];
        int m1 = k < nb ? k : nb;
        for (int m = 0; m < m1; m++)
            t -= (s->o[m][k - m - 1]) * (s->o[m][k - m - 1]) * (s->d[k - m - 1]);
        s->d[k] = t;
        int nn = (n - k - 1) < nb ? (n - k - 1) : nb;
        for (int j = 0; j < nn; j++) {
            t = offd[j][k];
            int mm = k < (nb - j - 1) ? k : (nb - j - 1);
            for (int m = 0; m < mm; m++)
                t -= (s->o[m][k - m - 1]) * (s->o[j + m + 1][k - m - 1]) * (s->d[k - m - 1]);
            s->o[j][k] = t / s->d[k];
        }
    }
}

static void bands_solve(const bands *s, real *b) {
    int n = s->n, nb = s->band;
    for (int k = 1; k < n; k++) {
        real t = b[k];
        int m1 = k < nb ? k : nb;
        for (int m = 0; m < m1; m++) t -= (s->o[m][k - m - 1]) * b[k - m - 1];
        b[k] = t;
    }
    for (int k = n - 1; k >= 0; k--) {
        real t = b[k] / s->d[k];
        int m1 = (n - k - 1) < nb ? (n - k - 1) : nb;
        for (int m = 0; m < m1; m++) t -= s->o[m][k] * b[k + m + 1];
        b[k] = t;
    }
}

/* ------------------------------------------------------------------ */
/* Trace prediction by implicit PWD (sof_cfuns.c:405-716,              */
/* sof3d_cfuns.c:361-661).                                             */
/* ------------------------------------------------------------------ */
typedef struct {
    int n1, nw, na, nb;
    real eps, eps2;
    double btab[16];
    real *a1[8], *a2[8];   /* filter taps per sample, two filters */
    real *diag, *offd[8];
    bands slv;
} predictor;

static void pred_init(predictor *w, int n1, int nw, real e) {
    memset(w, 0, sizeof(*w));
    w->n1 = n1; w->nw = nw; w->na = 2 * nw + 1; w->nb = 2 * nw;
    w->eps = e; w->eps2 = e;
    apf_table(nw, w->btab);
    for (int j = 0; j < w->na; j++) {
        w->a1[j] = xcalloc(n1, sizeof(real));
        w->a2[j] = xcalloc(n1, sizeof(real));
    }
    w->diag = xcalloc(n1, sizeof(real));
    for (int j = 0; j < w->nb; j++) w->offd[j] = xcalloc(n1, sizeof(real));
    w->slv.n = n1; w->slv.band = w->nb;
    w->slv.d = xcalloc(n1, sizeof(real));
    for (int j = 0; j < w->nb; j++) w->slv.o[j] = xcalloc(n1, sizeof(real));
}

static void pred_free(predictor *w) {
    for (int j = 0; j < w->na; j++) { free(w->a1[j]); free(w->a2[j]); }
    free(w->diag);
    for (int j = 0; j < w->nb; j++) { free(w->offd[j]); free(w->slv.o[j]); }
    free(w->slv.d);
}

/* Second-difference-squared regulariser with modified end rows
 * (regularization, sof_cfuns.c:548-565). */
static void pred_regularize(predictor *w) {
    int n1 = w->n1;
    real e = w->eps;
    for (int i = 0; i < n1; i++) {
        w->diag[i] = (real)(6. * e);
        w->offd[0][i] = (real)(-4. * e);
        w->offd[1][i] = e;
        for (int b = 2; b < w->nb; b++) w->offd[b][i] = 0;
    }
    w->diag[0] = w->diag[n1 - 1] = w->eps2 + e;
    w->diag[1] = w->diag[n1 - 2] = (real)(w->eps2 + 5. * e);
    w->offd[0][0] = w->offd[0][n1 - 2] = (real)(-2. * e);
}

/* Filter taps per sample (reversed when rev) and the normal-matrix update
 * diag/offd += A^T A (pwd_define, sof_cfuns.c:453-497). */
static void pred_define(predictor *w, int rev, const real *pp, real **a) {
    int n = w->n1, nw = w->nw, na = w->na;
    real b[16];
    for (int i = 0; i < n; i++) {
        apf_pass(nw, w->btab, pp[i], b);
        for (int j = 0; j < na; j++) a[j][i] = rev ? b[na - 1 - j] : b[j];
    }
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < na; j++) {
            int k = i + j - nw;
            if (k >= nw && k < n - nw) { real aj = a[j][k]; w->diag[i] += aj * aj; }
        }
        for (int m = 0; m < 2 * nw; m++)
            for (int j = m + 1; j < na; j++) {
                int k = i + j - nw;
                if (k >= nw && k < n - nw) {
                    real aj = a[j][k], am = a[j - m - 1][k];
                    w->offd[m][i] += am * aj;
                }
            }
    }
}

/* out = A^T A inp (forward) or its adjoint (pwd_set, sof_cfuns.c:499-550).
 * inp and out may alias. */
static void pred_apply(const predictor *w, int adj, real **a, real *inp, real *out, real *tmp) {
    int n = w->n1, nw = w->nw, na = w->na;
    if (adj) {
        for (int i = 0; i < n; i++) tmp[i] = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < na; j++) {
                int k = i + j - nw;
                if (k >= nw && k < n - nw) tmp[k] += a[j][k] * out[i];
            }
        for (int i = 0; i < n; i++) inp[i] = 0;
        for (int i = nw; i < n - nw; i++)
            for (int j = 0; j < na; j++) inp[i + j - nw] += a[j][i] * tmp[i];
    } else {
        for (int i = 0; i < n; i++) tmp[i] = 0;
        for (int i = nw; i < n - nw; i++)
            for (int j = 0; j < na; j++) tmp[i] += a[j][i] * inp[i + j - nw];
        for (int i = 0; i < n; i++) {
            out[i] = 0;
            for (int j = 0; j < na; j++) {
                int k = i + j - nw;
                if (k >= nw && k < n - nw) out[i] += a[j][k] * tmp[k];
            }
        }
    }
}

static void pred_boundary(const predictor *w, real *trace, real t0, real t1, real t2, real t3) {
    int n1 = w->n1;
    trace[0] += w->eps2 * t0;
    trace[1] += w->eps2 * t1;
    trace[n1 - 2] += w->eps2 * t2;
    trace[n1 - 1] += w->eps2 * t3;
}

/* predict_step (sof_cfuns.c:622-649): in-place one-trace prediction. */
static void pred_step(predictor *w, int adj, int forw, real *trace, const real *pp) {
    int n1 = w->n1;
    pred_regularize(w);
    pred_define(w, forw, pp, w->a1);
    bands_define(&w->slv, w->diag, w->offd);
    if (adj) bands_solve(&w->slv, trace);
    real t0 = trace[0], t1 = trace[1], t2 = trace[n1 - 2], t3 = trace[n1 - 1];
    pred_apply(w, adj, w->a1, trace, trace, w->diag);
    pred_boundary(w, trace, t0, t1, t2, t3);
    if (!adj) bands_solve(&w->slv, trace);
}

/* predict1_step (sof3d_cfuns.c:596-622). */
static void pred1(predictor *w, int forw, const real *trace1, const real *pp, real *trace) {
    int n1 = w->n1;
    pred_regularize(w);
    pred_define(w, forw, pp, w->a1);
    bands_define(&w->slv, w->diag, w->offd);
    real t0 = trace1[0], t1 = trace1[1], t2 = trace1[n1 - 2], t3 = trace1[n1 - 1];
    pred_apply(w, 0, w->a1, (real *)trace1, trace, w->diag);
    pred_boundary(w, trace, t0, t1, t2, t3);
    bands_solve(&w->slv, trace);
}

/* predict2_step (sof3d_cfuns.c:624-661). */
static void pred2(predictor *w, int forw1, int forw2, const real *tr1, const real *tr2,
                  const real *pp1, const real *pp2, real *trace) {
    int n1 = w->n1;
    pred_regularize(w);
    pred_define(w, forw1, pp1, w->a1);
    pred_define(w, forw2, pp2, w->a2);
    bands_define(&w->slv, w->diag, w->offd);
    real t0 = (real)(0.5 * (tr1[0] + tr2[0]));
    real t1 = (real)(0.5 * (tr1[1] + tr2[1]));
    real t2 = (real)(0.5 * (tr1[n1 - 2] + tr2[n1 - 2]));
    real t3 = (real)(0.5 * (tr1[n1 - 1] + tr2[n1 - 1]));
    pred_apply(w, 0, w->a1, (real *)tr1, w->offd[0], trace);
    pred_apply(w, 0, w->a2, (real *)tr2, w->offd[1], trace);
    for (int i = 0; i < n1; i++) trace[i] = w->offd[0][i] + w->offd[1][i];
    pred_boundary(w, trace, t0, t1, t2, t3);
    bands_solve(&w->slv, trace);
}

EXPORT void or_predict_step(int adj, int forw, int n1, int nw, real e, real *trace, const real *pp) {
    predictor w;
    pred_init(&w, n1, nw, e);
    pred_step(&w, adj, forw, trace, pp);
    pred_free(&w);
}

EXPORT void or_predict1(int forw, int n1, int nw, real e, const real *tr1, const real *pp, real *out) {
    predictor w;
    pred_init(&w, n1, nw, e);
    pred1(&w, forw, tr1, pp, out);
    pred_free(&w);
}

EXPORT void or_predict2(int forw1, int forw2, int n1, int nw, real e, const real *tr1,
                        const real *tr2, const real *pp1, const real *pp2, real *out) {
    predictor w;
    pred_init(&w, n1, nw, e);
    pred2(&w, forw1, forw2, tr1, tr2, pp1, pp2, out);
    pred_free(&w);
}

/* ------------------------------------------------------------------ */
/* 2-D plane-wave spray and structure-oriented smoothing.              */
/* pwspray_lop (sof_cfuns.c:948-1031), pwsmooth (1036-1135).           */
/* u layout: u[(i2*ns2 + slot)*n1 + i1], ns2 = 2ns+1.                  */
/* ------------------------------------------------------------------ */
static void spray2_lop(predictor *w, int adj, int add, int n1, int n2, int ns,
                       const real *dip, real *u1, real *u, real *trace) {
    int ns2 = 2 * ns + 1;
    size_t n = (size_t)n1 * n2;
    if (!add) {
        if (adj) for (size_t i = 0; i < n; i++) u1[i] = 0;
        else for (size_t i = 0; i < n * ns2; i++) u[i] = 0;
    }
    for (int i = 0; i < n2; i++) {
        real *ui = u1 + (size_t)i * n1;
        if (adj) {
            for (int k = 0; k < n1; k++) trace[k] = 0;
            for (int is = ns - 1; is >= 0; is--) {
                int ip = i + is + 1;
                if (ip >= n2) continue;
                const real *src = u + ((size_t)ip * ns2 + ns + is + 1) * n1;
                for (int k = 0; k < n1; k++) trace[k] += src[k];
                pred_step(w, 1, 1, trace, dip + (size_t)(ip - 1) * n1);
            }
            for (int k = 0; k < n1; k++) { ui[k] += trace[k]; trace[k] = 0; }
            for (int is = ns - 1; is >= 0; is--) {
                int ip = i - is - 1;
                if (ip < 0) continue;
                const real *src = u + ((size_t)ip * ns2 + ns - is - 1) * n1;
                for (int k = 0; k < n1; k++) trace[k] += src[k];
                pred_step(w, 1, 0, trace, dip + (size_t)ip * n1);
            }
            const real *ctr = u + ((size_t)i * ns2 + ns) * n1;
            for (int k = 0; k < n1; k++) { ui[k] += trace[k]; trace[k] = ctr[k]; ui[k] += trace[k]; }
        } else {
            real *ctr = u + ((size_t)i * ns2 + ns) * n1;
            for (int k = 0; k < n1; k++) { trace[k] = ui[k]; ctr[k] += trace[k]; }
            for (int is = 0; is < ns; is++) {
                int ip = i - is - 1;
                if (ip < 0) break;
                pred_step(w, 0, 0, trace, dip + (size_t)ip * n1);
                real *dst = u + ((size_t)ip * ns2 + ns - is - 1) * n1;
                for (int k = 0; k < n1; k++) dst[k] += trace[k];
            }
            for (int k = 0; k < n1; k++) trace[k] = ui[k];
            for (int is = 0; is < ns; is++) {
                int ip = i + is + 1;
                if (ip >= n2) break;
                pred_step(w, 0, 1, trace, dip + (size_t)(ip - 1) * n1);
                real *dst = u + ((size_t)ip * ns2 + ns + is + 1) * n1;
                for (int k = 0; k < n1; k++) dst[k] += trace[k];
            }
        }
    }
}

/* Caller-facing 2-D spray (pwspray2d drop-in semantics): forward x -> u, or
 * adjoint u -> x; eps is squared for the predictor as in pwspray_init. */
EXPORT void or_pwspray2d(int adj, int n1, int n2, int ns, int order, real eps,
                         const real *dip, real *x, real *u) {
    predictor w;
    pred_init(&w, n1, order, eps * eps);
    real *trace = xcalloc(n1, sizeof(real));
    spray2_lop(&w, adj, 0, n1, n2, ns, dip, x, u, trace);
    free(trace);
    pred_free(&w);
}

/* Structure-oriented smoothing state for one 2-D slice (pwsmooth_*). */
typedef struct {
    int n1, n2, ns, ns2;
    predictor w;
    real *u, *wts, *ww, *trace;
    const real *dip;
} smooth2d;

static void sm2_init(smooth2d *s, int ns, int n1, int n2, int order, real eps) {
    s->n1 = n1; s->n2 = n2; s->ns = ns; s->ns2 = 2 * ns + 1;
    pred_init(&s->w, n1, order, eps * eps);
    s->u = xcalloc((size_t)n1 * n2 * s->ns2, sizeof(real));
    s->wts = xcalloc(s->ns2, sizeof(real));
    s->ww = xcalloc((size_t)n1 * n2, sizeof(real));
    s->trace = xcalloc(n1, sizeof(real));
    for (int is = 0; is < s->ns2; is++) s->wts[is] = (real)(ns + 1 - abs(is - ns));
}

static void sm2_free(smooth2d *s) {
    pred_free(&s->w);
    free(s->u); free(s->wts); free(s->ww); free(s->trace);
}

/* pwsmooth_lop (sof_cfuns.c:1064-1099). */
static void sm2_lop(smooth2d *s, int adj, int add, real *trace, real *smooth) {
    int n1 = s->n1, n2 = s->n2, ns2 = s->ns2;
    size_t n = (size_t)n1 * n2;
    if (!add) {
        if (adj) for (size_t i = 0; i < n; i++) trace[i] = 0;
        else for (size_t i = 0; i < n; i++) smooth[i] = 0;
    }
    if (adj) {
        for (int i2 = 0; i2 < n2; i2++)
            for (int i1 = 0; i1 < n1; i1++) {
                real ws = s->ww[(size_t)i2 * n1 + i1];
                for (int is = 0; is < ns2; is++)
                    s->u[((size_t)i2 * ns2 + is) * n1 + i1] = smooth[(size_t)i2 * n1 + i1] * s->wts[is] * ws;
            }
        spray2_lop(&s->w, 1, 1, n1, n2, s->ns, s->dip, trace, s->u, s->trace);
    } else {
        spray2_lop(&s->w, 0, 0, n1, n2, s->ns, s->dip, trace, s->u, s->trace);
        for (int i2 = 0; i2 < n2; i2++)
            for (int i1 = 0; i1 < n1; i1++) {
                real ws = s->ww[(size_t)i2 * n1 + i1];
                real acc = smooth[(size_t)i2 * n1 + i1];
                for (int is = 0; is < ns2; is++)
                    acc += s->u[((size_t)i2 * ns2 + is) * n1 + i1] * s->wts[is] * ws;
                smooth[(size_t)i2 * n1 + i1] = acc;
            }
    }
}

/* pwsmooth_set (sof_cfuns.c:1102-1122): weights = 1 / smooth(ones). */
static void sm2_set(smooth2d *s, const real *dip) {
    size_t n = (size_t)s->n1 * s->n2;
    s->dip = dip;
    real *t = xcalloc(n, sizeof(real));
    for (size_t i = 0; i < n; i++) s->ww[i] = 1.0f;
    real *ones = xcalloc(n, sizeof(real));
    for (size_t i = 0; i < n; i++) ones[i] = 1.0f;
    sm2_lop(s, 0, 0, ones, t);
    for (size_t i = 0; i < n; i++) s->ww[i] = (0.0f != t[i]) ? (real)(1.0 / t[i]) : 0.0f;
    free(t); free(ones);
}

/* csomean2d (sof_cfuns.c:1433-1532): per n3 slice, normalised triangle-
 * weighted spray average (or its adjoint). */
EXPORT void or_somean2d(const real *dn, const real *dip, int n1, int n2, int n3, int ns,
                        int order, int adj, real eps, real *out) {
    size_t n12 = (size_t)n1 * n2;
    smooth2d s;
    sm2_init(&s, ns, n1, n2, order, eps);
    real *in = xcalloc(n12, sizeof(real));
    for (int i3 = 0; i3 < n3; i3++) {
        memcpy(in, dn + i3 * n12, n12 * sizeof(real));
        sm2_set(&s, dip + i3 * n12);
        if (adj) sm2_lop(&s, 1, 0, out + i3 * n12, in);
        else sm2_lop(&s, 0, 0, in, out + i3 * n12);
    }
    free(in);
    sm2_free(&s);
}

/* Hoare selection of the q-th smallest, in place (sf_quantile,
 * sof_cfuns.c:1137-1164).  Kept step-for-step so that the returned
 * element (including the sign of a zero) is the reference's. */
EXPORT real or_quantile(int q, int n, real *a) {
    real *lo = a, *hi = a + n - 1, *k = a + q;
    while (lo < hi) {
        real pivot = *k;
        real *i = lo, *j = hi;
        do {
            while (*i < pivot) i++;
            while (*j > pivot) j--;
            if (i <= j) { real t = *i; *i++ = *j; *j-- = t; }
        } while (i <= j);
        if (j < k) lo = i;
        if (k < i) hi = j;
    }
    return *k;
}

/* Median over slot windows of one panel tt[np][n1] (slots slow), centre
 * column only, replicate boundary (mf, axis 2, ifbound 1:
 * sof_cfuns.c:1214-1252).  Returns column (np-1)/2 in out[n1]. */
static void panel_mf_centre(const real *tt, int n1, int np, int nfw, real *out, real *win) {
    int m = (nfw - 1) / 2, c = (np - 1) / 2;
    for (int j = 0; j < n1; j++) {
        for (int k = 0; k < nfw; k++) {
            int col = c - m + k;
            if (col < 0) col = 0;
            if (col > np - 1) col = np - 1;
            win[k] = tt[(size_t)col * n1 + j];
        }
        out[j] = or_quantile(m, nfw, win);
    }
}

/* Space-varying median, centre column only (svmf, sof_cfuns.c:1309-1431).
 * The reference's first pass reads one panel row beyond its extended
 * buffer for the last column (index m+i+k-m2 reaches np+2m); here that
 * row is taken as a further replicate of the edge column. */
static void panel_svmf_centre(const real *tt, int n1, int np, int nfw, real *out, real *win) {
    int nf = nfw + 2, m = (nf - 1) / 2, m2 = (nfw - 1) / 2, c = (np - 1) / 2;
    real sum = 0;
    for (int i = 0; i < np; i++)
        for (int j = 0; j < n1; j++) {
            for (int k = 0; k < nf; k++) {
                int col = m + i + k - m2 - m;      /* extended index minus m */
                if (col < 0) col = 0;
                if (col > np - 1) col = np - 1;
                win[k] = tt[(size_t)col * n1 + j];
            }
            sum = sum + (real)fabs(or_quantile(m, nf, win));
        }
    real avg = sum / (real)(n1 * np);
    for (int j = 0; j < n1; j++) {
        real v = tt[(size_t)c * n1 + j];
        real av = (real)fabs(v);
        real wl;
        if (av < avg) wl = (av < avg / 2) ? (real)(nfw + 2) : (real)(nfw + 0);
        else wl = (av > avg * 2) ? (real)(nfw - 4) : (real)(nfw - 2);
        int h = (int)((wl - 1) / 2);
        int len = (int)wl;
        for (int k = 0; k < len; k++) {
            int col = c + k - h;
            if (col < 0) col = 0;
            if (col > np - 1) col = np - 1;
            win[k] = tt[(size_t)col * n1 + j];
        }
        out[j] = or_quantile(h, len, win);
    }
}

/* csomf2d (sof_cfuns.c:1534-1672): spray each slice, then the median (or
 * space-varying median) over nmf slots around the centre slot. */
EXPORT void or_somf2d(const real *dn, const real *dip, int n1, int n2, int n3, int ns,
                      int nmf, int option, int order, real eps, real *out) {
    size_t n12 = (size_t)n1 * n2;
    int np = 2 * ns + 1;
    predictor w;
    pred_init(&w, n1, order, eps * eps);
    real *u = xcalloc(n12 * np, sizeof(real)), *trace = xcalloc(n1, sizeof(real));
    real *x = xcalloc(n12, sizeof(real)), *win = xcalloc(nmf + 8, sizeof(real));
    for (int i3 = 0; i3 < n3; i3++) {
        memcpy(x, dn + i3 * n12, n12 * sizeof(real));
        spray2_lop(&w, 0, 0, n1, n2, ns, dip + i3 * n12, x, u, trace);
        for (int i = 0; i < n2; i++) {
            const real *tt = u + (size_t)i * np * n1;
            real *o = out + i3 * n12 + (size_t)i * n1;
            if (option == 1) panel_mf_centre(tt, n1, np, nmf, o, win);
            else panel_svmf_centre(tt, n1, np, nmf, o, win);
        }
    }
    free(u); free(trace); free(x); free(win);
    pred_free(&w);
}

/* ------------------------------------------------------------------ */
/* 3-D plane-wave spray in order of distance from the source trace.    */
/* update_init / get_update (sof3d_cfuns.c:983-1054) and the spray     */
/* loop of csomean3d / csomf3d (1455-1521, 1639-1705).                 */
/* u layout: u[(trace*np + slot)*n1 + i1], slot = k2 + np2*k3.         */
/* ------------------------------------------------------------------ */
typedef struct { int slot, update, up2, up3; } spray_step;

/* Visit order by increasing hypotf distance; ties are independent because
 * every predecessor is strictly closer. */
static int spray3_plan(int ns2, int ns3, spray_step *plan) {
    int np2 = 2 * ns2 + 1, np3 = 2 * ns3 + 1, np = np2 * np3;
    float *t = xcalloc(np, sizeof(float));
    int *order = xcalloc(np, sizeof(int));
    for (int k3 = 0; k3 < np3; k3++)
        for (int k2 = 0; k2 < np2; k2++) t[k2 + np2 * k3] = hypotf(k2 - ns2, k3 - ns3);
    for (int i = 0; i < np; i++) order[i] = i;
    for (int i = 1; i < np; i++) { /* insertion sort, stable */
        int v = order[i], j = i - 1;
        while (j >= 0 && t[order[j]] > t[v]) { order[j + 1] = order[j]; j--; }
        order[j + 1] = v;
    }
    for (int ip = 0; ip < np; ip++) {
        int jp = order[ip], k2 = jp % np2, k3 = jp / np2, upd = 0, u2 = 0, u3 = 0;
        float tj = t[jp];
        if (np2 > 1) {
            u2 = k2 && (k2 == np2 - 1 || !(t[jp - 1] > t[jp + 1]));
            if (tj > t[u2 ? jp - 1 : jp + 1]) upd |= 1;
        }
        if (np3 > 1) {
            u3 = k3 && (k3 == np3 - 1 || !(t[jp - np2] > t[jp + np2]));
            if (tj > t[u3 ? jp - np2 : jp + np2]) upd |= 2;
        }
        plan[ip].slot = jp; plan[ip].update = upd; plan[ip].up2 = u2; plan[ip].up3 = u3;
    }
    free(t); free(order);
    return np;
}

/* Export of the visit plan for tests: 4 ints per slot. */
EXPORT int or_spray3_plan(int ns2, int ns3, int *out) {
    spray_step plan[4096];
    int np = spray3_plan(ns2, ns3, plan);
    for (int i = 0; i < np; i++) {
        out[4 * i] = plan[i].slot; out[4 * i + 1] = plan[i].update;
        out[4 * i + 2] = plan[i].up2; out[4 * i + 3] = plan[i].up3;
    }
    return np;
}

EXPORT void or_pwspray3d(const real *din, const real *dipi, const real *dipx, int n1,
                         int n2, int n3, int ns2, int ns3, int order, real eps, real *u) {
    int np2 = 2 * ns2 + 1, np3 = 2 * ns3 + 1, np = np2 * np3;
    size_t n23 = (size_t)n2 * n3;
    spray_step *plan = xcalloc(np, sizeof(spray_step));
    spray3_plan(ns2, ns3, plan);
    predictor w;
    pred_init(&w, n1, order, eps * eps);
    memset(u, 0, n23 * np * n1 * sizeof(real));
    int ctr = ns3 * np2 + ns2;
    for (size_t i = 0; i < n23; i++) {
        memcpy(u + (i * np + ctr) * n1, din + i * n1, n1 * sizeof(real));
        int i2 = (int)(i % n2), i3 = (int)(i / n2);
        for (int ip = 0; ip < np; ip++) {
            int jp = plan[ip].slot, upd = plan[ip].update;
            int up2 = plan[ip].up2, up3 = plan[ip].up3;
            int j2 = i2 + jp % np2 - ns2, j3 = i3 + jp / np2 - ns3;
            if (j2 < 0 || j2 >= n2 || j3 < 0 || j3 >= n3) continue;
            size_t j = j2 + (size_t)j3 * n2;
            real *trace = u + (j * np + jp) * n1;
            const real *pa = NULL, *pb = NULL, *qa = NULL, *qb = NULL;
            if (upd & 1) {
                if (up2) { if (j2 == 0) continue; pa = u + ((j - 1) * np + jp - 1) * n1; qa = dipi + (j - 1) * n1; }
                else { if (j2 == n2 - 1) continue; pa = u + ((j + 1) * np + jp + 1) * n1; qa = dipi + j * n1; }
            }
            if (upd & 2) {
                if (up3) { if (j3 == 0) continue; pb = u + ((j - n2) * np + jp - np2) * n1; qb = dipx + (j - n2) * n1; }
                else { if (j3 == n3 - 1) continue; pb = u + ((j + n2) * np + jp + np2) * n1; qb = dipx + j * n1; }
            }
            if (upd == 1) pred1(&w, up2, pa, qa, trace);
            else if (upd == 2) pred1(&w, up3, pb, qb, trace);
            else if (upd == 3) pred2(&w, up2, up3, pa, pb, qa, qb, trace);
        }
    }
    pred_free(&w);
    free(plan);
}

/* csomean3d (sof3d_cfuns.c:1355-1552): eps is forced to 0.01 (1402);
 * plain mean over all np slots, zero slots included (1525-1537). */
EXPORT void or_somean3d(const real *dn, const real *dipi, const real *dipx, int n1, int n2,
                        int n3, int ns2, int ns3, int order, real *out) {
    int np = (2 * ns2 + 1) * (2 * ns3 + 1);
    size_t n23 = (size_t)n2 * n3;
    real *u = xcalloc(n23 * np * n1, sizeof(real));
    or_pwspray3d(dn, dipi, dipx, n1, n2, n3, ns2, ns3, order, (real)0.01f, u);
    for (size_t i = 0; i < n23; i++)
        for (int k = 0; k < n1; k++) {
            real s = 0;
            for (int j = 0; j < np; j++) s = s + u[(i * np + j) * n1 + k];
            out[i * n1 + k] = s / (real)np;
        }
    free(u);
}

/* csomf3d (sof3d_cfuns.c:1554-1753): eps forced to 0.01 (1586); median
 * (option 1) or space-varying median of the nmf linear slot indices
 * centred at (np-1)/2. */
EXPORT void or_somf3d(const real *dn, const real *dipi, const real *dipx, int n1, int n2,
                      int n3, int ns2, int ns3, int nmf, int option, int order, real *out) {
    int np = (2 * ns2 + 1) * (2 * ns3 + 1);
    size_t n23 = (size_t)n2 * n3;
    real *u = xcalloc(n23 * np * n1, sizeof(real)), *win = xcalloc(nmf + 8, sizeof(real));
    or_pwspray3d(dn, dipi, dipx, n1, n2, n3, ns2, ns3, order, (real)0.01f, u);
    for (size_t i = 0; i < n23; i++) {
        if (option == 1) panel_mf_centre(u + i * np * n1, n1, np, nmf, out + i * n1, win);
        else panel_svmf_centre(u + i * np * n1, n1, np, nmf, out + i * n1, win);
    }
    free(u); free(win);
}

/* ------------------------------------------------------------------ */
/* Structure-oriented interpolation.                                   */
/* ------------------------------------------------------------------ */

/* Claerbout's cgstep with its retained step (ps_cgstep,
 * soint3d_cfuns.c:826-877).  state S (nx) and Ss (ny) persist. */
typedef struct { real *S, *Ss; int started; } cgstep_state;

static void cgstep(cgstep_state *st, int forget, size_t nx, size_t ny, real *x,
                   const real *g, real *rr, const real *gg) {
    const real EPSILON = (real)1.e-12f;
    double alfa, beta;
    if (!st->started) { st->started = 1; forget = 1; }
    if (forget) {
        for (size_t i = 0; i < nx; i++) st->S[i] = 0;
        for (size_t i = 0; i < ny; i++) st->Ss[i] = 0;
        beta = 0.0;
        alfa = dsdot(ny, gg, gg);
        if (alfa <= 0.) return;
        alfa = -dsdot(ny, gg, rr) / alfa;
    } else {
        double gdg = dsdot(ny, gg, gg), sds = dsdot(ny, st->Ss, st->Ss);
        double gds = dsdot(ny, gg, st->Ss);
        if (gdg == 0. || sds == 0.) return;
        double determ = 1.0 - (gds / gdg) * (gds / sds);
        if (determ > EPSILON) determ *= gdg * sds;
        else determ = gdg * sds * EPSILON;
        double gdr = -dsdot(ny, gg, rr), sdr = -dsdot(ny, st->Ss, rr);
        alfa = (sds * gdr - gds * sdr) / determ;
        beta = (-gds * gdr + gdg * sdr) / determ;
    }
    real b = (real)beta, a = (real)alfa;
    for (size_t i = 0; i < nx; i++) { st->S[i] *= b; }
    for (size_t i = 0; i < nx; i++) { st->S[i] += a * g[i]; }
    for (size_t i = 0; i < ny; i++) { st->Ss[i] *= b; }
    for (size_t i = 0; i < ny; i++) { st->Ss[i] += a * gg[i]; }
    for (size_t i = 0; i < nx; i++) x[i] += st->S[i];
    for (size_t i = 0; i < ny; i++) rr[i] += st->Ss[i];
}

/* ps_solver restricted to the options soint uses ("known", "x0"):
 * soint3d_cfuns.c:894-1174. */
typedef void (*lin_op)(void *ctx, int adj, int add, real *x, real *y);
static int known_solver(lin_op op, void *ctx, size_t nx, size_t ny, real *x,
                        const real *dat, int niter, const unsigned char *known) {
    const real TOL = (real)1.e-12f;
    real *g = xcalloc(nx, sizeof(real)), *rr = xcalloc(ny, sizeof(real)), *gg = xcalloc(ny, sizeof(real));
    cgstep_state st = {xcalloc(nx, sizeof(real)), xcalloc(ny, sizeof(real)), 0};
    for (size_t i = 0; i < ny; i++) rr[i] = -dat[i];
    op(ctx, 0, 1, x, rr);
    real dpr0 = snrm2(ny, rr), dpg0 = 1.;
    int it;
    for (it = 0; it < niter; it++) {
        op(ctx, 1, 0, g, rr);
        for (size_t i = 0; i < nx; i++) if (known[i]) g[i] = 0.0;
        op(ctx, 0, 0, g, gg);
        real dpr, dpg;
        if (it == 0) { dpg0 = snrm2(nx, g); dpr = 1.; dpg = 1.; }
        else { dpr = snrm2(ny, rr) / dpr0; dpg = snrm2(nx, g) / dpg0; }
        if (dpr < TOL || dpg < TOL) break;
        cgstep(&st, 0, nx, ny, x, g, rr, gg);
    }
    free(g); free(rr); free(gg); free(st.S); free(st.Ss);
    return it;
}

typedef struct { int n1, n2, n3, nw, nj1, nj2, drift; const real *pp, *qq; } ap3_ctx;
static void ap3_op(void *c, int adj, int add, real *x, real *y) {
    ap3_ctx *a = c;
    or_allpass3_lop(adj, add, a->n1, a->n2, a->n3, a->nw, a->nj1, a->nj2, a->drift, a->pp, a->qq, x, y);
}

/* MT19937 + Box-Muller (soint3d_cfuns.c:2303-2402), used for var > 0. */
typedef struct { uint32_t mt[624]; int mti; int iset; float vset; } mt_rng;
static void mt_seed(mt_rng *r, uint32_t s) {
    r->mt[0] = s;
    for (int i = 1; i < 624; i++) r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->mti = 624; r->iset = 1; r->vset = 0;
}
static uint32_t mt_next(mt_rng *r) {
    static const uint32_t mag[2] = {0u, 0x9908b0dfu};
    if (r->mti >= 624) {
        int k;
        for (k = 0; k < 624 - 397; k++) {
            uint32_t y = (r->mt[k] & 0x80000000u) | (r->mt[k + 1] & 0x7fffffffu);
            r->mt[k] = r->mt[k + 397] ^ (y >> 1) ^ mag[y & 1u];
        }
        for (; k < 623; k++) {
            uint32_t y = (r->mt[k] & 0x80000000u) | (r->mt[k + 1] & 0x7fffffffu);
            r->mt[k] = r->mt[k + 397 - 624] ^ (y >> 1) ^ mag[y & 1u];
        }
        uint32_t y = (r->mt[623] & 0x80000000u) | (r->mt[0] & 0x7fffffffu);
        r->mt[623] = r->mt[396] ^ (y >> 1) ^ mag[y & 1u];
        r->mti = 0;
    }
    uint32_t y = r->mt[r->mti++];
    y ^= (y >> 11); y ^= (y << 7) & 0x9d2c5680u; y ^= (y << 15) & 0xefc60000u; y ^= (y >> 18);
    return y;
}
static float mt_randn(mt_rng *r) {
    if (r->iset) {
        double x1, x2;
        do { x1 = mt_next(r) * (1.0 / 4294967295.0); } while (x1 == 0.0);
        x2 = mt_next(r) * (1.0 / 4294967295.0);
        double z1 = sqrt(-2.0 * log(x1)), z2 = 2.0 * 3.14159265358979323846264338328 * x2;
        r->iset = 0;
        r->vset = (float)(z1 * cos(z2));
        return (float)(z1 * sin(z2));
    }
    r->iset = 1;
    return r->vset;
}

EXPORT void or_randn(uint32_t seed, size_t n, float *out) {
    mt_rng r;
    mt_seed(&r, seed);
    for (size_t i = 0; i < n; i++) out[i] = mt_randn(&r);
}

/* csoint3d (soint3d_cfuns.c:2405-2508): known samples are fixed, the
 * rest minimise |[A1;A2] x - dd|, dd = sqrt(var) * N(0,1). */
EXPORT int or_soint3d(const real *din, const real *mask, const real *dipi, const real *dipx,
                      int n1, int n2, int n3, int nw, int nj1, int nj2, int niter, int drift,
                      int seed, int hasmask, real var, real *out) {
    size_t n = (size_t)n1 * n2 * n3;
    real a = (real)sqrtf((float)var);
    unsigned char *known = xcalloc(n, 1);
    real *dd = xcalloc(2 * n, sizeof(real));
    for (size_t i = 0; i < n; i++) {
        out[i] = din[i];
        known[i] = hasmask == 1 ? (mask[i] != 0.) : (din[i] != 0.);
    }
    mt_rng r;
    mt_seed(&r, (uint32_t)seed);
    for (size_t i = 0; i < 2 * n; i++) dd[i] = a * (real)mt_randn(&r);
    ap3_ctx c = {n1, n2, n3, nw, nj1, nj2, drift, dipi, dipx};
    int it = known_solver(ap3_op, &c, n, 2 * n, out, dd, niter, known);
    free(known); free(dd);
    return it;
}

typedef struct { int n1, n2, nw, nj, drift; const real *pp; } ap21_ctx;
static void ap21_op(void *c, int adj, int add, real *x, real *y) {
    ap21_ctx *a = c;
    or_allpass21_lop(adj, add, a->n1, a->n2, a->nw, a->nj, a->drift, a->pp, x, y);
}

/* csoint2d without preconditioning (soint2d_cfuns.c:2260-2419).  With
 * twoplane set (and prec unset) the reference solver call is commented out
 * and the data come back unchanged. */
EXPORT int or_soint2d(const real *din, const real *mask, const real *dip, int n1, int n2,
                      int nw, int nj1, int niter, int drift, int hasmask, int twoplane,
                      real *out) {
    size_t n = (size_t)n1 * n2;
    for (size_t i = 0; i < n; i++) out[i] = din[i];
    if (twoplane) return 0;
    unsigned char *known = xcalloc(n, 1);
    real *dd = xcalloc(n, sizeof(real));
    for (size_t i = 0; i < n; i++) known[i] = hasmask == 1 ? (mask[i] != 0.) : (din[i] != 0.);
    ap21_ctx c = {n1, n2, nw, nj1, drift, dip};
    int it = known_solver(ap21_op, &c, n, n, out, dd, niter, known);
    free(known); free(dd);
    return it;
}

/* predict_lop (soint2d_cfuns.c:1537-1574) without add: yy = P xx, or
 * xx = P^T yy.  A running trace tt walks the traces; each step is the
 * forward-direction prediction (predict_step with forw) through slope
 * trace i2 of dip. */
static void prec_lop(predictor *w, int adj, int n2, const real *dip, real *xx, real *yy, real *tt) {
    int n1 = w->n1;
    size_t n = (size_t)n1 * n2;
    for (size_t i = 0; i < n; i++) { if (adj) xx[i] = 0; else yy[i] = 0; }
    for (int i1 = 0; i1 < n1; i1++) tt[i1] = 0;
    if (adj) {
        for (int i2 = n2 - 1; i2 >= 0; i2--) {
            pred_step(w, 1, 1, tt, dip + (size_t)i2 * n1);
            for (int i1 = 0; i1 < n1; i1++) tt[i1] += yy[i1 + (size_t)i2 * n1];
            for (int i1 = 0; i1 < n1; i1++) xx[i1 + (size_t)i2 * n1] += tt[i1];
        }
    } else {
        for (int i2 = 0; i2 < n2; i2++) {
            for (int i1 = 0; i1 < n1; i1++) tt[i1] += xx[i1 + (size_t)i2 * n1];
            for (int i1 = 0; i1 < n1; i1++) yy[i1 + (size_t)i2 * n1] += tt[i1];
            pred_step(w, 0, 1, tt, dip + (size_t)i2 * n1);
        }
    }
}

/* The preconditioner: predict_lop through dip, or predict2_lop
 * (soint2d_cfuns.c:1597-1616) = P(qq) P(pp) when qq is set. */
typedef struct { predictor w; int n2; const real *pp, *qq; real *tt, *tmp; } prec_ctx;
static void prec_apply(prec_ctx *c, int adj, real *x, real *y) {
    if (!c->qq) { prec_lop(&c->w, adj, c->n2, c->pp, x, y, c->tt); return; }
    if (adj) {
        prec_lop(&c->w, 1, c->n2, c->qq, c->tmp, y, c->tt);
        prec_lop(&c->w, 1, c->n2, c->pp, x, c->tmp, c->tt);
    } else {
        prec_lop(&c->w, 0, c->n2, c->pp, x, c->tmp, c->tt);
        prec_lop(&c->w, 0, c->n2, c->qq, c->tmp, y, c->tt);
    }
}

/* csoint2d with prec (soint2d_cfuns.c:2352-2396): ps_solver_prec
 * (1704-1990) with the masked data operator, this preconditioner, eps = 0
 * and no initial model, so the model is x = P p with p minimising
 * |M P p - d|.  The solver's model temporary is the output itself (the
 * reference aliases x and dat); d is read only at the start.  Vectors p and
 * g have 2n entries (the eps-scaled data part stays zero). */
EXPORT int or_soint2d_prec(const real *din, const real *mask, const real *dip, const real *dip2, int n1,
                           int n2, int nw, int niter, int hasmask, real *out) {
    const real TOLERANCE = (real)1.e-12f;
    const real eps = 0;
    size_t n = (size_t)n1 * n2;
    unsigned char *known = xcalloc(n, 1);
    for (size_t i = 0; i < n; i++) {
        out[i] = din[i];
        known[i] = hasmask == 1 ? (mask[i] != 0.) : (din[i] != 0.);
    }
    prec_ctx c;
    pred_init(&c.w, n1, nw, (real)0.0001f);
    c.n2 = n2; c.pp = dip; c.qq = dip2;
    c.tt = xcalloc(n1, sizeof(real));
    c.tmp = dip2 ? xcalloc(n, sizeof(real)) : NULL;
    real *p = xcalloc(2 * n, sizeof(real)), *g = xcalloc(2 * n, sizeof(real));
    real *rr = xcalloc(n, sizeof(real)), *gg = xcalloc(n, sizeof(real)), *x = xcalloc(n, sizeof(real));
    cgstep_state st = {xcalloc(2 * n, sizeof(real)), xcalloc(n, sizeof(real)), 0};
    for (size_t i = 0; i < n; i++) rr[i] = -din[i];
    double dprr0 = 1., dpgm0 = 1.;
    int iter;
    for (iter = 0; iter < niter; iter++) {
        for (size_t i = 0; i < n; i++) x[i] = known[i] ? (real)0 + rr[i] : (real)0;
        prec_apply(&c, 1, g, x);
        for (size_t i = 0; i < n; i++) g[i + n] = eps * rr[i];
        prec_apply(&c, 0, g, x);
        for (size_t i = 0; i < n; i++) gg[i] = known[i] ? (real)0 + x[i] : (real)0;
        for (size_t i = 0; i < n; i++) gg[i] += eps * g[i + n];
        double dprr, dpgm;
        if (iter == 0) {
            dprr0 = snrm2(n, rr);
            dpgm0 = snrm2(n, g);
            dprr = 1.;
            dpgm = 1.;
        } else {
            dprr = snrm2(n, rr) / dprr0;
            dpgm = snrm2(n, g) / dpgm0;
        }
        if (dprr < TOLERANCE || dpgm < TOLERANCE) {
            prec_apply(&c, 0, p, out);
            break;
        }
        cgstep(&st, 0, 2 * n, n, p, g, rr, gg);
        if (iter == niter - 1) prec_apply(&c, 0, p, out);
    }
    pred_free(&c.w);
    free(c.tt); free(c.tmp); free(known); free(p); free(g); free(rr); free(gg); free(x); free(st.S); free(st.Ss);
    return iter;
}

/* Masked data operator (ps_mask_lop, soint3d_cfuns.c:1359-1372). */
typedef struct { const unsigned char *m; size_t n; } mask_ctx;
static void mask_op(void *c, int adj, int add, real *x, real *y) {
    mask_ctx *mc = c;
    if (!add) {
        if (adj) for (size_t i = 0; i < mc->n; i++) x[i] = 0;
        else for (size_t i = 0; i < mc->n; i++) y[i] = 0;
    }
    for (size_t i = 0; i < mc->n; i++)
        if (mc->m[i]) { if (adj) x[i] += y[i]; else y[i] += x[i]; }
}

typedef struct { smooth2d s; } sm2_ctx;
static void sm2_op(void *c, int adj, int add, real *p, real *x) {
    sm2_ctx *sc = c;
    sm2_lop(&sc->s, adj, add, p, x);
}

/* csint2d (soint2d_cfuns.c:2421-2538): masked shaping CG with 2-D
 * structure-oriented smoothing as the shaping operator. */
EXPORT int or_sint2d(const real *din, const real *dip, const real *mask, int n1, int n2,
                     int niter, int ns, int nw, real eps, real *out) {
    size_t n = (size_t)n1 * n2;
    unsigned char *known = xcalloc(n, 1);
    real *xx = xcalloc(n, sizeof(real));
    real lam = 0.;
    for (size_t i = 0; i < n; i++) {
        out[i] = din[i];
        xx[i] = din[i];
        if (mask[i] != 0.) { known[i] = 1; lam += 1.; }
    }
    lam = sqrtf(lam / n);
    sm2_ctx sc;
    sm2_init(&sc.s, ns, n1, n2, nw, eps);
    sm2_set(&sc.s, dip);
    mask_ctx mc = {known, n};
    cg_ops ops = {mask_op, sm2_op, &mc, &sc};
    real *dat = xcalloc(n, sizeof(real));
    memcpy(dat, out, n * sizeof(real));
    cg_report rep;
    shaping_cg(&ops, n, lam, (real)(10 * 1.19209290E-07F), 1, xx, out, dat, niter, &rep);
    sm2_free(&sc.s);
    free(known); free(xx); free(dat);
    return rep.iters;
}

/* pwsmooth3_lop (soint3d_cfuns.c:2231-2300): inline smoothing of every n3
 * slice, then crossline smoothing of every n2 slice (adjoint: reverse). */
typedef struct {
    int n1, n2, n3, ns1, ns2, o1, o2;
    real eps;
    const real *dipi, *dipx;
    real *xdip;       /* crossline dip regrouped [i2][i3][i1] */
    real *itmp, *itmp2, *xtmp;
} sm3_ctx;

static void sm3_op(void *c, int adj, int add, real *trace, real *smooth) {
    sm3_ctx *s = c;
    int n1 = s->n1, n2 = s->n2, n3 = s->n3;
    size_t n12 = (size_t)n1 * n2, n13 = (size_t)n1 * n3, n = n12 * n3;
    if (!add) {
        if (adj) for (size_t i = 0; i < n; i++) trace[i] = 0;
        else for (size_t i = 0; i < n; i++) smooth[i] = 0;
    }
    smooth2d sm;
    if (adj) {
        for (int i3 = 0; i3 < n3; i3++)
            for (int i2 = 0; i2 < n2; i2++)
                memcpy(s->xtmp + ((size_t)i2 * n3 + i3) * n1, smooth + ((size_t)i3 * n2 + i2) * n1, n1 * sizeof(real));
        sm2_init(&sm, s->ns2, n1, n3, s->o2, s->eps);
        for (int i2 = 0; i2 < n2; i2++) {
            sm2_set(&sm, s->xdip + i2 * n13);
            sm2_lop(&sm, 1, 0, s->itmp2 + i2 * n13, s->xtmp + i2 * n13);
        }
        sm2_free(&sm);
        for (int i3 = 0; i3 < n3; i3++)
            for (int i2 = 0; i2 < n2; i2++)
                memcpy(s->itmp + ((size_t)i3 * n2 + i2) * n1, s->itmp2 + ((size_t)i2 * n3 + i3) * n1, n1 * sizeof(real));
        sm2_init(&sm, s->ns1, n1, n2, s->o1, s->eps);
        for (int i3 = 0; i3 < n3; i3++) {
            sm2_set(&sm, s->dipi + i3 * n12);
            sm2_lop(&sm, 1, 1, trace + i3 * n12, s->itmp + i3 * n12);
        }
        sm2_free(&sm);
    } else {
        sm2_init(&sm, s->ns1, n1, n2, s->o1, s->eps);
        for (int i3 = 0; i3 < n3; i3++) {
            sm2_set(&sm, s->dipi + i3 * n12);
            sm2_lop(&sm, 0, 0, trace + i3 * n12, s->itmp + i3 * n12);
        }
        sm2_free(&sm);
        for (int i3 = 0; i3 < n3; i3++)
            for (int i2 = 0; i2 < n2; i2++)
                memcpy(s->itmp2 + ((size_t)i2 * n3 + i3) * n1, s->itmp + ((size_t)i3 * n2 + i2) * n1, n1 * sizeof(real));
        sm2_init(&sm, s->ns2, n1, n3, s->o2, s->eps);
        for (int i2 = 0; i2 < n2; i2++) {
            sm2_set(&sm, s->xdip + i2 * n13);
            sm2_lop(&sm, 0, 0, s->itmp2 + i2 * n13, s->xtmp + i2 * n13);
        }
        sm2_free(&sm);
        for (int i3 = 0; i3 < n3; i3++)
            for (int i2 = 0; i2 < n2; i2++) {
                real *dst = smooth + ((size_t)i3 * n2 + i2) * n1;
                const real *src = s->xtmp + ((size_t)i2 * n3 + i3) * n1;
                for (int i1 = 0; i1 < n1; i1++) dst[i1] += src[i1];
            }
    }
}

/* csint3d (soint3d_cfuns.c:2510-2615). */
EXPORT int or_sint3d(const real *din, const real *dipi, const real *dipx, const real *mask,
                     int n1, int n2, int n3, int niter, int ns1, int ns2, int o1, int o2,
                     real eps, real *out) {
    size_t n = (size_t)n1 * n2 * n3, n13 = (size_t)n1 * n3;
    unsigned char *known = xcalloc(n, 1);
    real *xx = xcalloc(n, sizeof(real));
    real lam = 0.;
    for (size_t i = 0; i < n; i++) {
        out[i] = din[i];
        xx[i] = din[i];
        if (mask[i] != 0.) { known[i] = 1; lam += 1.; }
    }
    lam = sqrtf(lam / n);
    sm3_ctx s = {n1, n2, n3, ns1, ns2, o1, o2, eps, dipi, dipx,
                 xcalloc(n, sizeof(real)), xcalloc(n, sizeof(real)),
                 xcalloc(n, sizeof(real)), xcalloc(n, sizeof(real))};
    for (int i2 = 0; i2 < n2; i2++)
        for (int i3 = 0; i3 < n3; i3++)
            memcpy(s.xdip + i2 * n13 + (size_t)i3 * n1, dipx + ((size_t)i3 * n2 + i2) * n1, n1 * sizeof(real));
    mask_ctx mc = {known, n};
    cg_ops ops = {mask_op, sm3_op, &mc, &s};
    real *dat = xcalloc(n, sizeof(real));
    memcpy(dat, out, n * sizeof(real));
    cg_report rep;
    shaping_cg(&ops, n, lam, (real)(10 * 1.19209290E-07F), 1, xx, out, dat, niter, &rep);
    free(s.xdip); free(s.itmp); free(s.itmp2); free(s.xtmp);
    free(known); free(xx); free(dat);
    return rep.iters;
}

/* 2-D structure-oriented smoothing of one slice, exported for tests
 * (pwsmooth_set + pwsmooth_lop). */
EXPORT void or_pwsmooth2d(int adj, const real *dip, int n1, int n2, int ns, int order,
                          real eps, real *trace, real *smooth) {
    smooth2d s;
    sm2_init(&s, ns, n1, n2, order, eps);
    sm2_set(&s, dip);
    sm2_lop(&s, adj, 0, trace, smooth);
    sm2_free(&s);
}

EXPORT int or_real_bytes(void) { return (int)sizeof(real); }
