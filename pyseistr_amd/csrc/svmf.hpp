// Space-varying median of one trace's slot panel, centre slot only (svmf,
// sof_cfuns.c:1336-1431, sof3d_cfuns.c:1254-1352): a first pass of medians
// over nfw + 2 slots around every slot and sample (columns clamped into the
// panel; where the reference reads one row past its extended panel the edge
// column is used), their mean absolute value in float summed slot by slot
// and sample by sample as there, then per centre sample a window of
// nfw + 2, nfw, nfw - 2 or nfw - 4 slots chosen by |x| against that mean.
// Element (slot s, sample k) at U[s * n + base + k * stride]; nfw + 2 <= 128.
#pragma once

// q-th smallest of w[0..len) by counting ranks (the value is that of the
// reference's selection; a zero may come back with the other sign)
__device__ inline float select_rank(const float *w, int len, int q) {
    for (int a = 0; a < len; a++) {
        const float v = w[a];
        int less = 0, leq = 0;
        for (int b = 0; b < len; b++) {
            less += w[b] < v;
            leq += w[b] <= v;
        }
        if (less <= q && q < leq) return v;
    }
    return w[0];
}

__device__ inline void svmf_one(const float *__restrict__ U, float *__restrict__ out, long n, int np, int n1, int nfw,
                                long base, long stride) {
    float w[128];
    const int nf = nfw + 2, m = (nf - 1) / 2, m2 = (nfw - 1) / 2, c = (np - 1) / 2;
    float sum = 0.0f;
    for (int i = 0; i < np; i++)
        for (int j = 0; j < n1; j++) {
            for (int k = 0; k < nf; k++) {
                int col = i + k - m2;
                col = col < 0 ? 0 : (col > np - 1 ? np - 1 : col);
                w[k] = U[(long)col * n + base + (long)j * stride];
            }
            sum = sum + fabsf(select_rank(w, nf, m));
        }
    const float avg = sum / (float)(n1 * np);
    for (int j = 0; j < n1; j++) {
        const float v = U[(long)c * n + base + (long)j * stride];
        const float av = fabsf(v);
        int len;
        if (av < avg) len = (av < avg / 2) ? nfw + 2 : nfw;
        else len = (av > avg * 2) ? nfw - 4 : nfw - 2;
        const int h = (len - 1) / 2;
        for (int k = 0; k < len; k++) {
            int col = c + k - h;
            col = col < 0 ? 0 : (col > np - 1 ? np - 1 : col);
            w[k] = U[(long)col * n + base + (long)j * stride];
        }
        out[base + (long)j * stride] = select_rank(w, len, h);
    }
}
