/*
 * seistr_hip.h -- C ABI of the MI355X (gfx950) structural hot path.
 *
 * Every entry point takes a context (HIP stream + cached workspace), DEVICE
 * pointers to Fortran-order float32 arrays (n1 fastest, then n2, then n3)
 * and the reference's own parameters, and writes caller-owned device
 * outputs.  Calls are asynchronous on the context's stream unless stated.
 * Return value: SH_OK (0) or a negative status; sh_last_error() gives the
 * message.  Nothing here allocates host-visible state shared between
 * contexts: one context per thread/stream.
 *
 * Each declaration cites the reference entry point it replaces (paths
 * relative to the reference repository).
 */
#ifndef SEISTR_HIP_H
#define SEISTR_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SH_OK 0
#define SH_EINVAL (-1)   /* bad shape or parameter */
#define SH_ENULL (-2)    /* required pointer is NULL */
#define SH_EALLOC (-3)   /* device allocation failed */
#define SH_EHIP (-4)     /* HIP runtime error */
#define SH_EUNSUP (-5)   /* option the library does not implement */
#define SH_ENONFINITE (-6) /* an input holds NaN or Inf (with SH_FLAG_CHECK_FINITE) */

/* Context flags of the path's entry points (sh_dip, sh_pwspray2d/3d,
 * sh_somean2d/3d, sh_somf2d/3d, sh_soint2d/3d, sh_soint2d_prec,
 * sh_sint2d/3d); the reference checks neither, so both are off by default.
 * SH_FLAG_CHECK_FINITE: every float input is scanned first and the call
 *   returns SH_ENONFINITE, naming the input, if one holds NaN or Inf (the
 *   scan synchronises the stream).
 * SH_FLAG_HOST_PTRS: the array arguments are host memory; inputs are
 *   copied to device workspace, outputs copied back, and the call returns
 *   when they are (synchronous).  Without it every pointer is device
 *   memory. */
#define SH_FLAG_CHECK_FINITE 1
#define SH_FLAG_HOST_PTRS 2

typedef struct sh_ctx sh_ctx;

/* Create a context bound to a HIP stream (NULL = default stream). */
sh_ctx *sh_ctx_create(void *hip_stream);
void sh_ctx_destroy(sh_ctx *ctx);
/* Set / read the context's SH_FLAG_* bits. */
int sh_ctx_set_flags(sh_ctx *ctx, int flags);
int sh_ctx_get_flags(const sh_ctx *ctx);
/* *nbad = number of NaN or Inf values in the device array a[n]
 * (synchronises). */
int sh_count_nonfinite(sh_ctx *ctx, const float *a, long n, long *nbad);
/* Rebind the context to another stream; the new stream first waits for the
 * work already queued on the old one (the workspace is shared). */
int sh_ctx_set_stream(sh_ctx *ctx, void *hip_stream);
/* Release cached device workspace. */
int sh_ctx_trim(sh_ctx *ctx);
/* Layout helper.  The reference's Python wrappers flatten C-order numpy
 * arrays in Fortran order on the host before calling C
 * (pyseistr/dip3d.py:100, pyseistr/somf3d.py:89-91); this does the same
 * reorder on the device.  in: a C-order (n1, n2, n3) array (n3 fastest;
 * a 2-D (n1, n2) array is (n1, 1, n2)); out: the same samples in Fortran
 * order (n1 fastest), the layout every other entry point takes.  Device
 * pointers, not overlapping; n2 <= 65535, n1 <= 4194240. */
int sh_c_to_f(sh_ctx *ctx, const float *in, float *out, int n1, int n2, int n3);
const char *sh_last_error(const sh_ctx *ctx);
int sh_version(void);
/* Block until all work queued on the context's stream has finished. */
int sh_sync(sh_ctx *ctx);

/* Launch timing (no counterpart in the reference; instrumentation for the
 * benchmark): when enabled, every kernel launch of the context is bracketed
 * by HIP events on its stream.  sh_prof_report synchronises and writes
 * {"kernel": [launches, total_ms, total_algorithmic_bytes], ...} to buf. */
int sh_prof_enable(sh_ctx *ctx, int on);
int sh_prof_report(sh_ctx *ctx, char *buf, int len);

/* Per-direction iteration record of a dip call: for GN iteration k < 64,
 * cg_iters[d][k] = shaping-CG iterations run (-1: zero denominator),
 * ls_step[d][k] = accepted line-search halving (8 = none accepted);
 * elapsed_ms = device time of the call on its stream, from its first
 * launch to its last (HIP events). */
typedef struct {
    int ndir, niter;
    int cg_iters[2][64];
    int ls_step[2][64];
    float elapsed_ms;
} sh_dip_report;

/* Local slopes by shaping-regularised plane-wave destruction.
 * Replaces dipcfun.dipc (pyseistr/src/dip_cfuns.c:1694-1989), the C entry
 * behind dip2dc (pyseistr/dip2d.py:115) and dip3dc (pyseistr/dip3d.py:59).
 * n3 == 1: inline dip only, dip[N]; otherwise inline then crossline,
 * dip[2N].  mask (nullable, length N): samples whose PWD stencil touches a
 * zero of mask are excluded (mask32, dip_cfuns.c:914).  init (nullable, N
 * or 2N): initial dip instead of the reference's zeros (the commented
 * idip/xdip inputs, dip_cfuns.c:1861-1875).  eps_dv is accepted and has no
 * effect, as in the reference (its shared `eps` static is reset to 1).
 * rep (nullable, host memory): iteration record; filling it synchronises.
 * With n3 > 1 the two directions are independent: the crossline one runs
 * on a second stream owned by the context, forked from and joined back to
 * the context's stream inside the call, so every result is ordered on the
 * caller's stream as usual (environment SH_DIP_STREAMS=1: one stream). */
int sh_dip(sh_ctx *ctx, const float *data, const float *mask, const float *init,
           int n1, int n2, int n3, int niter, int liter, int order, float eps_dv,
           int r1, int r2, int r3, float *dip, sh_dip_report *rep);

/* Triangle (or box) smoothing with optional derivative, per axis whose
 * radius r exceeds 1, axis 1 first, each line `repeat` times.
 * Replaces dipcfun.smoothcf (pyseistr/src/dip_cfuns.c:2006-2123), the C
 * entry behind smoothc (pyseistr/smooth.py:115); adj != 0 applies
 * ps_smooth, adj == 0 ps_smooth2 (dip_cfuns.c:594-625), as there.
 * Bit-identical to the reference.  out may equal in. */
int sh_smooth(sh_ctx *ctx, const float *in, float *out, int n1, int n2, int n3, int repeat, int adj,
              int r1, int r2, int r3, int diff1, int diff2, int diff3, int box1, int box2, int box3);

/* Butterworth band-pass of every trace (n1 samples, sampling d1): low cut
 * flo and high cut fhi (in the units of 1/d1), nplo / nphi poles, phase
 * != 0 minimum phase, else zero phase (each filter again over the reversed
 * trace, with half the poles).  Replaces bpcfun.cbp
 * (pyseistr/src/bp_cfuns.c:524-645), the C entry behind bandpassc
 * (pyseistr/bp.py:128), including its parameter handling; bit-identical to
 * the reference.  out may equal in.  SH_EUNSUP above 127 poles. */
int sh_bandpass(sh_ctx *ctx, const float *in, float *out, int n1, int n2, int n3, float d1, float flo,
                float fhi, int nplo, int nphi, int phase);

/* Patch extraction and reassembly (mode 1), for 2-D data pass n3 = l3 =
 * s3 = 1.  Each axis is zero-padded so that (n - l) is a multiple of s;
 * patches are numbered with axis 1 outermost and axis 3 innermost, each a
 * row of l1*l2*l3 samples in Fortran order.  Replace the pure-Python
 * patch2d / patch3d (pyseistr/patch.py:2-77, 86-201) and patch2d_inv /
 * patch3d_inv (patch.py:208-300, 307-446); float64 throughout, as the
 * reference computes once it pads.
 * sh_patch_count: *npatch = number of patches; SH_EINVAL if none fits. */
int sh_patch_count(int n1, int n2, int n3, int l1, int l2, int l3, int s1, int s2, int s3, long *npatch);
/* X[npatch][l1*l2*l3] = the patches of in[n1*n2*n3] (Fortran order). */
int sh_patch(sh_ctx *ctx, const double *in, double *X, int n1, int n2, int n3, int l1, int l2, int l3,
             int s1, int s2, int s3);
/* out[n1*n2*n3] = the patches X summed in patch order and divided by the
 * number covering each sample (NaN where none does), identical to the
 * reference's float64 reassembly. */
int sh_patch_inv(sh_ctx *ctx, const double *X, double *out, int n1, int n2, int n3, int l1, int l2, int l3,
                 int s1, int s2, int s3);

/* FK dip filter of a 2-D section in[n1*n2] (Fortran order, float64):
 * padded to powers of two, transformed in time and space, a cone of half
 * width w (fraction of the padded trace count) kept, and transformed back;
 * out[n1*n2].  Replaces the pure-Python fkdip (pyseistr/fk.py:2-68), which
 * computes with numpy's float64 FFTs; agreement is to float64 rounding.
 * SH_EINVAL when w * nk == 0 (the reference divides by it). */
int sh_fkdip(sh_ctx *ctx, const double *in, double *out, int n1, int n2, double w);

/* 3-D plane-wave spray, slots visited by distance from the source trace.
 * Replaces the spray loop of csomean3d/csomf3d (pyseistr/src/
 * sof3d_cfuns.c:1639-1705) and the pure pwspray3d (pyseistr/pwspray3d.py:2).
 * u has the reference's layout u[(trace*np + slot)*n1 + i1],
 * np = (2ns2+1)(2ns3+1); eps is used as given (predictor e = eps*eps). */
int sh_pwspray3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx,
                 int n1, int n2, int n3, int ns2, int ns3, int order, float eps, float *u);

/* Structure-oriented mean: spray, then the plain mean over all np slots.
 * Replaces sof3dcfun.csomean3d (pyseistr/src/sof3d_cfuns.c:1355-1552);
 * eps is ignored and 0.01 used, as in the reference (1402). */
int sh_somean3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx,
                int n1, int n2, int n3, int r1, int r2, float eps, int order, float *out);

/* Structure-oriented median: spray, then the median of the nmf linear slot
 * indices centred at (np-1)/2 (option 1) or the space-varying median over
 * all slots (option 2, svmf, sof3d_cfuns.c:1254-1352; needs nmf >= 5; where
 * the reference reads one row past its extended panel the edge slot is
 * used).  Replaces sof3dcfun.csomf3d (pyseistr/src/sof3d_cfuns.c:1554-1753);
 * eps forced to 0.01 as in the reference (1586). */
int sh_somf3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx,
              int n1, int n2, int n3, int r1, int r2, int nmf, int option, int order,
              float eps, float *out);

/* 2-D plane-wave spray of one n1 x n2 slice, forward (x -> u) or adjoint
 * (u -> x).  Replaces pwspray_lop (pyseistr/src/sof_cfuns.c:948-1031) and
 * the pure pwspray2d (pyseistr/pwspray2d.py:130).  u layout
 * u[(i2*(2ns+1) + slot)*n1 + i1]. */
int sh_pwspray2d(sh_ctx *ctx, int adj, const float *dip, int n1, int n2, int ns, int order,
                 float eps, float *x, float *u);

/* Normalised triangle-weighted structure-oriented smoothing per n3 slice
 * (or its adjoint).  Replaces sofcfun.csomean2d (pyseistr/src/
 * sof_cfuns.c:1433-1532) behind somean2dc (pyseistr/somean2d.py:36). */
int sh_somean2d(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3,
                int ns, int order, int adj, float eps, float *out);

/* Structure-oriented median (option 1) or space-varying median (option 2,
 * svmf, sof_cfuns.c:1336-1431; needs nmf >= 5) per n3 slice.  Replaces
 * sofcfun.csomf2d (pyseistr/src/sof_cfuns.c:1534-1672) behind somf2dc
 * (pyseistr/somf2d.py:60). */
int sh_somf2d(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3,
              int ns, int nmf, int option, int order, float eps, float *out);

/* sh_somf2d without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as sh_somf2d (bit-identical);
 * the workspace is a fixed device array of SH_SOMF2D_NOALLOC_MAX_BYTES
 * inside the library, one per GPU, held until the process ends.  A call
 * whose workspace (sh_somf2d_noalloc_bytes) exceeds it fails with
 * SH_EINVAL before anything is launched; a 512 x 256 gather with ns = 8
 * needs about 18 MB.  Every call on one GPU shares that array, whatever
 * its context: calls that overlap on one GPU (two streams or threads)
 * race and must be ordered by the caller.  Launch timing is off inside
 * the call (it would create events); SH_FLAG_HOST_PTRS is refused
 * (SH_EINVAL) and SH_FLAG_CHECK_FINITE does not apply. */
#define SH_SOMF2D_NOALLOC_MAX_BYTES ((size_t)64 << 20)
size_t sh_somf2d_noalloc_bytes(int n1, int n2, int n3, int ns, int order);
int sh_somf2d_noalloc(sh_ctx *ctx, const float *dn, const float *dip, int n1, int n2, int n3,
                      int ns, int nmf, int option, int order, float eps, float *out);

/* PWD-regularised interpolation of unknown samples.  Replaces
 * soint3dcfun.csoint3d (pyseistr/src/soint3d_cfuns.c:2405-2508) behind
 * soint3dc (pyseistr/soint3d.py:65).  drift != 0 shifts each stencil by
 * NINT(p) along i1 and takes the filter at p - NINT(p)
 * (allpass3_lop, soint3d_cfuns.c:646-664).  iters (nullable, host): solver
 * iterations run; filling it synchronises. */
int sh_soint3d(sh_ctx *ctx, const float *din, const float *mask, const float *dipi,
               const float *dipx, int n1, int n2, int n3, int nw, int nj1, int nj2,
               int niter, int drift, int seed, int hasmask, float var, float *out, int *iters);

/* 2-D analogue without preconditioning.  Replaces soint2dcfun.csoint2d
 * (pyseistr/src/soint2d_cfuns.c:2260-2419) for prec == 0 (sh_soint2d_prec
 * below for prec != 0); twoplane != 0 returns the data unchanged, as the
 * reference does. */
int sh_soint2d(sh_ctx *ctx, const float *din, const float *mask, const float *dip,
               int n1, int n2, int nw, int nj1, int niter, int drift, int hasmask,
               int twoplane, float *out, int *iters);

/* The preconditioned form (prec != 0) of soint2dcfun.csoint2d
 * (pyseistr/src/soint2d_cfuns.c:2352-2396): out = P p with p minimising
 * |M P p - din| by ps_solver_prec (1704-1990; eps = 0), P = predict_lop
 * (1537-1574) through dip, or predict2_lop (1597-1616) = P(dip2) P(dip)
 * when dip2 is non-NULL (twoplane).  M keeps the known samples (mask != 0,
 * or din != 0 when hasmask != 1).  The trace-to-trace prediction chain is
 * sequential and runs in one thread; its banded factors are built in
 * parallel.  *iters (nullable) receives the iterations run. */
int sh_soint2d_prec(sh_ctx *ctx, const float *din, const float *mask, const float *dip,
                    const float *dip2, int n1, int n2, int nw, int niter, int hasmask,
                    float *out, int *iters);

/* Masked shaping-CG interpolation with structure-oriented smoothing as the
 * shaping operator.  Replaces soint3dcfun.csint3d (pyseistr/src/
 * soint3d_cfuns.c:2510-2615) behind sint3dc (pyseistr/sint.py:97). */
int sh_sint3d(sh_ctx *ctx, const float *din, const float *dipi, const float *dipx,
              const float *mask, int n1, int n2, int n3, int niter, int ns1, int ns2,
              int order1, int order2, float eps, float *out, int *iters);

/* 2-D analogue.  Replaces soint2dcfun.csint2d (pyseistr/src/
 * soint2d_cfuns.c:2421-2538) behind sint2dc (pyseistr/sint.py:61). */
int sh_sint2d(sh_ctx *ctx, const float *din, const float *dip, const float *mask,
              int n1, int n2, int niter, int ns, int order, float eps, float *out, int *iters);

/* ---- Batched 2-D path (no counterpart in the reference) ---------------
 * npatch patches stored back to back in one device array: patch p is an
 * n1s[p] x n2s[p] F-order block at element offset offs[p] (offsets
 * ascending, non-overlapping; the three tables are HOST arrays).  Each
 * patch gets exactly the per-patch algorithm; every pipeline stage is one
 * launch over the whole batch. */
int sh_batch_tile_count(void);
/* shape of tile cfg (0 .. count-1; -1: the default tile) */
int sh_batch_tile_shape(int cfg, int *t1, int *t2);
/* dip2dc per patch: dipc with n3 = 1 (pyseistr/src/dip_cfuns.c:1694-1989,
 * behind dip2dc, pyseistr/dip2d.py:115).  tile: smoothing tile shape
 * index (-1 = default; see sh_batch_tile_shape).  init (nullable, same
 * layout as data): initial dips.  rec (nullable, host, npatch*niter*2):
 * per patch and GN iteration the CG iteration count and accepted
 * line-search halving, as sh_dip_report; filling it synchronises. */
int sh_dip2d_batched(sh_ctx *ctx, const float *data, const long *offs, const int *n1s,
                     const int *n2s, int npatch, int niter, int liter, int order, int r1,
                     int r2, int tile, const float *init, float *dip, int *rec);
/* somf2dc per patch: csomf2d (pyseistr/src/sof_cfuns.c:1534-1672, behind
 * somf2dc, pyseistr/somf2d.py:60), option 1 or 2 (n3 = 1 per patch). */
int sh_somf2d_batched(sh_ctx *ctx, const float *data, const float *dip, const long *offs,
                      const int *n1s, const int *n2s, int npatch, int ns, int nmf, int option,
                      int order, float eps, float *out);

/* ---- Learned warm start (no counterpart in the reference) -------------
 * A 3-D CNN (six 3x3x3 convolutions, 1->32->32->32->32->32->2 channels,
 * ReLU, zero padding; BF16 activations/weights on the matrix cores, FP32
 * accumulation, FP32 master parameters) that predicts the initial dip
 * passed to sh_dip's init -- the reference's commented idip/xdip hook
 * (pyseistr/src/dip_cfuns.c:1861-1875).  Parameter vector (FP32):
 *   W1[27][32] b1[32] {W[27][32 in][32 out] b[32]} x4 W6[27][32][2] b6[2],
 * tap t = offset (t%3-1, (t/3)%3-1, t/9-1) along (i1, i2, i3). */
long sh_cnn_nparams(void);
/* dil (nullable: all 1): dilation of each of the six layers, 1..8 (tap t
 * then covers the offset dil * (t%3-1, (t/3)%3-1, t/9-1)). */
/* dip[2N] = inline || crossline prediction for the volume x[N] * xscale;
 * xscale <= 0: the scale is 1 / RMS of x (1 for an all-zero x), computed
 * on the device, so the call needs no host round trip. */
int sh_cnn_forward(sh_ctx *ctx, const float *params, const float *x, float xscale,
                   const int *dil, int n1, int n2, int n3, float *dip);
/* loss = sum wt (y - target)^2 / (2N) over both components (target[2N],
 * wt nullable [N]); grad[nparams] overwritten.  loss (nullable, host):
 * filling it synchronises. */
int sh_cnn_grad(sh_ctx *ctx, const float *params, const float *x, float xscale,
                const int *dil, const float *target, const float *wt, int n1, int n2,
                int n3, float *grad, double *loss);
/* One Adam step (step >= 1; L2 term wd * params added to the gradient). */
int sh_cnn_adam(sh_ctx *ctx, float *params, const float *grad, float *m, float *v,
                float lr, float b1, float b2, float eps, int step, float wd);

#ifdef __cplusplus
}
#endif
#endif /* SEISTR_HIP_H */
