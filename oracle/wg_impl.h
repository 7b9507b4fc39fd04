/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (see wg_common.h).
 *
 * Precision-dependent part of the CPU restatement of rang-core's exact
 * w-gridder (core/src/wgridder.rs): tile layout, stored values, gridding,
 * FFTs, w-screens, corrections and degridding. Included once by wg_f32.c
 * and once by wg_f64.c with REAL, SUF, SINGLE, WMAXT, LOG2TILE defined.
 *
 * The structure follows the reference stage by stage: a counting sort of the
 * active visibilities by tile, a stable sort by first w-plane, one plane at a
 * time (the library's default plane group of one), tiles gridded by parity
 * class into per-task buffers of at most 2048 visibilities merged in task
 * order, v transforms on the touched rows, u transforms on the image columns,
 * projection with the plane's w-screen added in plane order, and the global
 * correction at the end. The FFT is this file's own mixed-radix Stockham
 * transform (the reference uses rustfft 6.4.1): the two agree to rounding.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "wg_common.h"

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SUF)

#if SINGLE
#define RFMA fmaf
#define RSIN sinf
#define RCOS cosf
#else
#define RFMA fma
#define RSIN sin
#define RCOS cos
#endif

typedef struct {
    REAL re, im;
} CX;

static inline CX cx(REAL re, REAL im) {
    CX c = {re, im};
    return c;
}
static inline CX cmul(CX a, CX b) {
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline CX cadd(CX a, CX b) {
    return cx(a.re + b.re, a.im + b.im);
}
static inline CX csub(CX a, CX b) {
    return cx(a.re - b.re, a.im - b.im);
}
static inline CX cscale(CX a, REAL s) {
    return cx(a.re * s, a.im * s);
}

/* ------------------------------------------------------------------ FFT */

typedef struct {
    size_t n;
    int npass;
    int radix[40];
    size_t twoff[40];
    CX *tw;           /* forward twiddles e^{-2 pi i k t / (Ns R)} */
    double cs[12][12]; /* cos / sin (2 pi a / R) per odd radix R */
    double sn[12][12];
} fftplan;

static void fft_plan_init(fftplan *p, size_t n) {
    memset(p, 0, sizeof *p);
    p->n = n;
    size_t m = n;
    while (m % 4 == 0) {
        p->radix[p->npass++] = 4;
        m /= 4;
    }
    while (m % 2 == 0) {
        p->radix[p->npass++] = 2;
        m /= 2;
    }
    static const int odd[] = {3, 5, 7, 11};
    for (int i = 0; i < 4; ++i)
        while (m % odd[i] == 0) {
            p->radix[p->npass++] = odd[i];
            m /= odd[i];
        }
    if (m != 1) abort(); /* sizes are 2^a 3^b 5^c 7^d 11^e by construction */
    size_t total = 0, ns = 1;
    for (int s = 0; s < p->npass; ++s) {
        p->twoff[s] = total;
        total += ns * (p->radix[s] - 1);
        ns *= p->radix[s];
    }
    p->tw = malloc(sizeof(CX) * (total ? total : 1));
    ns = 1;
    for (int s = 0; s < p->npass; ++s) {
        int r = p->radix[s];
        for (size_t k = 0; k < ns; ++k)
            for (int t = 1; t < r; ++t) {
                double ang = -2.0 * M_PI * (double)(k * t) / (double)(ns * r);
                p->tw[p->twoff[s] + k * (r - 1) + (t - 1)] = cx((REAL)cos(ang), (REAL)sin(ang));
            }
        ns *= r;
    }
    for (int r = 3; r <= 11; r += 2)
        for (int a = 0; a < r; ++a) {
            p->cs[r][a] = cos(2.0 * M_PI * a / r);
            p->sn[r][a] = sin(2.0 * M_PI * a / r);
        }
}

static void fft_plan_free(fftplan *p) {
    free(p->tw);
}

/* out[q] = sum_t v[t] e^{sgn 2 pi i t q / R} in place. */
static inline void dft_small(const fftplan *p, CX *v, int r, int sgn) {
    if (r == 2) {
        CX a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
        return;
    }
    if (r == 4) {
        CX t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        CX t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        /* (sgn i) * d */
        CX t3 = sgn > 0 ? cx(-d.im, d.re) : cx(d.im, -d.re);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
        return;
    }
    int m = (r - 1) / 2;
    CX a[6], b[6], out[12];
    CX s0 = v[0];
    for (int t = 1; t <= m; ++t) {
        a[t] = cadd(v[t], v[r - t]);
        b[t] = csub(v[t], v[r - t]);
        s0 = cadd(s0, a[t]);
    }
    out[0] = s0;
    for (int q = 1; q <= m; ++q) {
        CX A = v[0], B = cx(0, 0);
        for (int t = 1; t <= m; ++t) {
            int idx = (t * q) % r;
            A = cadd(A, cscale(a[t], (REAL)p->cs[r][idx]));
            B = cadd(B, cscale(b[t], (REAL)p->sn[r][idx]));
        }
        /* sgn i B */
        CX iB = sgn > 0 ? cx(-B.im, B.re) : cx(B.im, -B.re);
        out[q] = cadd(A, iB);
        out[r - q] = csub(A, iB);
    }
    for (int q = 0; q < r; ++q) v[q] = out[q];
}

/* Unnormalised transform of `data` (n values) with sign sgn (-1 forward,
 * +1 inverse); `work` holds n values. */
static void fft_exec(const fftplan *p, CX *data, CX *work, int sgn) {
    size_t n = p->n;
    CX *x = data, *y = work;
    size_t ns = 1;
    for (int s = 0; s < p->npass; ++s) {
        int r = p->radix[s];
        size_t nb = n / r;
        const CX *tw = p->tw + p->twoff[s];
        CX v[11];
        for (size_t j = 0; j < nb; ++j) {
            size_t k = j % ns;
            for (int t = 0; t < r; ++t) v[t] = x[j + t * nb];
            if (ns > 1)
                for (int t = 1; t < r; ++t) {
                    CX w = tw[k * (r - 1) + (t - 1)];
                    if (sgn > 0) w.im = -w.im;
                    v[t] = cmul(v[t], w);
                }
            dft_small(p, v, r, sgn);
            size_t base = (j / ns) * ns * r + k;
            for (int t = 0; t < r; ++t) y[base + t * ns] = v[t];
        }
        CX *tmp = x;
        x = y;
        y = tmp;
        ns *= r;
    }
    if (x != data) memcpy(data, x, sizeof(CX) * n);
}

/* ------------------------------------------------------------- kernels */

/* wgridder.rs:1485-1507 (coeff_even_odd) */
static REAL *coeff_even_odd(const or_kernel *k, int *ne_out) {
    int w = k->support, d = k->degree, d1 = d + 1;
    int dodd = d | 1;
    int ne = (dodd + 1) / 2;
    REAL *out = calloc((size_t)2 * ne * w, sizeof(REAL));
    for (int i = 0; i < w; ++i)
        for (int kk = 0; kk < ne; ++kk) {
            int m1 = dodd - 1 - 2 * kk, m2 = dodd - 2 * kk;
            out[kk * w + i] = (REAL)(m1 > d ? 0.0 : k->coeff[i * d1 + (d - m1)]);
            out[(ne + kk) * w + i] = (REAL)(m2 > d ? 0.0 : k->coeff[i * d1 + (d - m2)]);
        }
    *ne_out = ne;
    return out;
}

/* wgridder.rs:1514-1543 (eval_taps2) */
static inline __attribute__((always_inline)) void eval_taps2(const REAL *c, int ne, const int S,
                                                             REAL x, REAL y, REAL *ru, REAL *rv) {
    const int h = (S + 1) / 2;
    REAL x2 = x * x, y2 = y * y;
    REAL ex[16], ey[16], ox[16], oy[16];
    for (int i = 0; i < h; ++i) {
        ex[i] = ey[i] = c[i];
        ox[i] = oy[i] = c[ne * S + i];
    }
    for (int k = 1; k < ne; ++k) {
        const REAL *ce = c + k * S, *co = c + (ne + k) * S;
        for (int i = 0; i < h; ++i) {
            ex[i] = RFMA(ex[i], x2, ce[i]);
            ey[i] = RFMA(ey[i], y2, ce[i]);
            ox[i] = RFMA(ox[i], x2, co[i]);
            oy[i] = RFMA(oy[i], y2, co[i]);
        }
    }
    for (int i = 0; i < h; ++i) {
        ru[i] = RFMA(ox[i], x, ex[i]);
        rv[i] = RFMA(oy[i], y, ey[i]);
    }
    for (int i = 0; i < S - h; ++i) {
        ru[S - 1 - i] = RFMA(ox[i], -x, ex[i]);
        rv[S - 1 - i] = RFMA(oy[i], -y, ey[i]);
    }
}

/* wgridder.rs:1548-1560 (eval_tap) */
static inline REAL eval_tap(const REAL *c, int ne, int S, int i, REAL x) {
    int h = (S + 1) / 2;
    if (i >= h) {
        i = S - 1 - i;
        x = -x;
    }
    REAL x2 = x * x;
    REAL e = c[i], o = c[ne * S + i];
    for (int k = 1; k < ne; ++k) {
        e = RFMA(e, x2, c[k * S + i]);
        o = RFMA(o, x2, c[(ne + k) * S + i]);
    }
    return RFMA(o, x, e);
}

/* wgridder.rs:1580-1591 (w_factor) */
static inline REAL w_factor(const or_setup *s, const REAL *c, int ne, int nth, size_t p, double w) {
    double w0 = s->wmin + (double)p * s->dw;
    REAL z = (REAL)(s->xdw * (w0 - w));
    REAL locz = (z - (REAL)nth) * (REAL)2.0 + (REAL)(s->supp - 1.0);
    return eval_tap(c, ne, s->supp, nth, locz);
}

/* ------------------------------------------------------------ inputs */

enum { VIS_UNIT = 0, VIS_C64 = 1, VIS_C128 = 2 };
enum { WGT_NONE = 0, WGT_F32 = 1, WGT_F64 = 2, WGT_F32P = 3 };

typedef struct {
    int kind;
    const void *p;
} visdata;
typedef struct {
    int kind;
    const void *a, *b;
} weights;

/* wgridder.rs:951-957 */
static inline CX vis_get(visdata v, size_t i) {
    if (v.kind == VIS_C64) {
        const float *f = v.p;
        return cx((REAL)(double)f[2 * i], (REAL)(double)f[2 * i + 1]);
    }
    if (v.kind == VIS_C128) {
        const double *f = v.p;
        return cx((REAL)f[2 * i], (REAL)f[2 * i + 1]);
    }
    return cx(1, 0);
}

/* wgridder.rs:998-1005; returns 0 without weights */
static inline int wgt_get(weights w, size_t i, REAL *out) {
    switch (w.kind) {
    case WGT_F32:
        *out = (REAL)(double)((const float *)w.a)[i];
        return 1;
    case WGT_F64:
        *out = (REAL)((const double *)w.a)[i];
        return 1;
    case WGT_F32P: {
        float prod = ((const float *)w.a)[i] * ((const float *)w.b)[i];
        *out = (REAL)(double)prod;
        return 1;
    }
    default:
        return 0;
    }
}

/* wgridder.rs:1033-1041 (Activity::active) */
static inline int active(int use_vis, visdata v, weights w, const uint8_t *mask, size_t idx) {
    double tnorm = 1.0;
    if (use_vis) {
        CX c = vis_get(v, idx);
        REAL ns = c.re * c.re + c.im * c.im;
        tnorm = (double)ns;
    }
    REAL tw;
    double twgt = wgt_get(w, idx, &tw) ? (double)tw : 1.0;
    double tmask = mask ? (double)mask[idx] : 1.0;
    return tnorm * twgt * tmask != 0.0;
}

/* wgridder.rs:1051-1077 (scan) */
static void scan(const or_baselines *bl, int use_vis, visdata v, weights w, const uint8_t *mask,
                 size_t *nvis, double *wlo, double *whi) {
    size_t n = 0;
    double lo = 1e300, hi = -1e300;
    size_t nchan = bl->nchan;
#pragma omp parallel for reduction(+ : n) reduction(min : lo) reduction(max : hi) schedule(static)
    for (size_t row = 0; row < bl->nrow; ++row)
        for (size_t ch = 0; ch < nchan; ++ch)
            if (active(use_vis, v, w, mask, row * nchan + ch)) {
                ++n;
                double ww = fabs(bl->coord[3 * row + 2] * bl->f_over_c[ch]);
                lo = fmin(lo, ww);
                hi = fmax(hi, ww);
            }
    *nvis = n;
    *wlo = lo;
    *whi = hi;
}

/* ------------------------------------------------------------- layout */

typedef struct {
    uint32_t row, ch;
} visref;
typedef struct {
    size_t minplane, start, end;
} block;
typedef struct {
    size_t tu, tv, b0, b1, plo, phi;
    int edge;
} tile;

typedef struct {
    int log2;
    size_t ntu, ntv, su, sv;
    visref *refs;
    size_t nrefs;
    block *blocks;
    size_t nblocks;
    tile *tiles;
    size_t ntiles;
} layout_t;

/* wgridder.rs:1137-1141 */
static inline void uvw_of(const or_baselines *bl, visref r, double *uvw, double *imflip) {
    double c[3];
    or_fix_w(bl->coord + 3 * (size_t)r.row, c, imflip);
    double f = bl->f_over_c[r.ch];
    uvw[0] = c[0] * f;
    uvw[1] = c[1] * f;
    uvw[2] = c[2] * f;
}

static inline size_t shifted(long i0, size_t n) {
    long m = (long)n;
    long v = (i0 + m / 2) % m;
    if (v < 0) v += m;
    return (size_t)v;
}

static inline size_t offset_in(long i0, size_t n, size_t base) {
    long m = (long)n;
    long d = i0 + m / 2 - (long)base;
    if (d < 0)
        d += m;
    else if (d >= m)
        d -= m;
    return (size_t)d;
}

/* wgridder.rs:1244-1255 (Setup::locate) */
static inline void locate(const or_setup *s, const layout_t *L, const or_baselines *bl, size_t row,
                          size_t ch, size_t *t, size_t *mp) {
    visref r = {(uint32_t)row, (uint32_t)ch};
    double uvw[3], imflip, uf, vf;
    long iu0, iv0;
    uvw_of(bl, r, uvw, &imflip);
    or_getpix(s, uvw[0], uvw[1], &uf, &vf, &iu0, &iv0);
    size_t tu = shifted(iu0, s->nu) >> L->log2;
    size_t tv = shifted(iv0, s->nv) >> L->log2;
    size_t nb = s->nplanes > 1 ? s->nplanes : 1;
    size_t m = or_minplane(s, uvw[2]);
    *t = tu * L->ntv + tv;
    *mp = m < nb - 1 ? m : nb - 1;
}

/* wgridder.rs:1261-1376 (layout): counting sort by tile over row chunks,
 * then a stable sort by first plane within each tile. */
static void make_layout(const or_setup *s, const or_baselines *bl, int use_vis, visdata v, weights w,
                        const uint8_t *mask, layout_t *L) {
    L->log2 = LOG2TILE;
    size_t tsz = (size_t)1 << L->log2;
    L->ntu = (s->nu + tsz - 1) / tsz;
    L->ntv = (s->nv + tsz - 1) / tsz;
    L->su = tsz + s->supp - 1;
    L->sv = tsz + s->supp - 1;
    size_t nrow = bl->nrow, nchan = bl->nchan, ntiles = L->ntu * L->ntv;
    size_t rpc = 256;
    if (((nrow * ntiles) >> 22) > rpc) rpc = (nrow * ntiles) >> 22;
    size_t nchunks = (nrow + rpc - 1) / rpc;
    if (nchunks == 0) nchunks = 1;
    size_t *counts = calloc(nchunks * ntiles, sizeof(size_t));
#pragma omp parallel for schedule(dynamic)
    for (size_t c = 0; c < nchunks; ++c) {
        size_t r1 = (c + 1) * rpc < nrow ? (c + 1) * rpc : nrow;
        for (size_t row = c * rpc; row < r1; ++row)
            for (size_t ch = 0; ch < nchan; ++ch)
                if (active(use_vis, v, w, mask, row * nchan + ch)) {
                    size_t t, mp;
                    locate(s, L, bl, row, ch, &t, &mp);
                    counts[c * ntiles + t]++;
                }
    }
    size_t *tile_start = malloc(sizeof(size_t) * (ntiles + 1));
    size_t acc = 0;
    for (size_t t = 0; t < ntiles; ++t) {
        tile_start[t] = acc;
        for (size_t c = 0; c < nchunks; ++c) {
            size_t k = counts[c * ntiles + t];
            counts[c * ntiles + t] = acc;
            acc += k;
        }
    }
    tile_start[ntiles] = acc;
    visref *refs = malloc(sizeof(visref) * (acc ? acc : 1));
    uint32_t *mps = malloc(sizeof(uint32_t) * (acc ? acc : 1));
#pragma omp parallel for schedule(dynamic)
    for (size_t c = 0; c < nchunks; ++c) {
        size_t r1 = (c + 1) * rpc < nrow ? (c + 1) * rpc : nrow;
        size_t *pos = counts + c * ntiles;
        for (size_t row = c * rpc; row < r1; ++row)
            for (size_t ch = 0; ch < nchan; ++ch)
                if (active(use_vis, v, w, mask, row * nchan + ch)) {
                    size_t t, mp;
                    locate(s, L, bl, row, ch, &t, &mp);
                    refs[pos[t]].row = (uint32_t)row;
                    refs[pos[t]].ch = (uint32_t)ch;
                    mps[pos[t]] = (uint32_t)mp;
                    pos[t]++;
                }
    }
    free(counts);
    /* Stable counting sort by first plane within each tile; runs -> blocks. */
    size_t nb = s->nplanes > 1 ? s->nplanes : 1;
    size_t ntile_used = 0;
    for (size_t t = 0; t < ntiles; ++t)
        if (tile_start[t + 1] > tile_start[t]) ++ntile_used;
    size_t *nblk = calloc(ntiles, sizeof(size_t));
#pragma omp parallel
    {
        size_t *cnt = malloc(sizeof(size_t) * (nb + 1));
        visref *tmp = NULL;
        size_t tmpcap = 0;
#pragma omp for schedule(dynamic)
        for (size_t t = 0; t < ntiles; ++t) {
            size_t a = tile_start[t], b = tile_start[t + 1];
            if (a == b) continue;
            if (b - a > tmpcap) {
                free(tmp);
                tmpcap = b - a;
                tmp = malloc(sizeof(visref) * tmpcap);
            }
            memset(cnt, 0, sizeof(size_t) * (nb + 1));
            for (size_t k = a; k < b; ++k) cnt[mps[k] + 1]++;
            size_t nruns = 0;
            for (size_t m = 0; m < nb; ++m) {
                if (cnt[m + 1]) ++nruns;
                cnt[m + 1] += cnt[m];
            }
            nblk[t] = nruns;
            for (size_t k = a; k < b; ++k) tmp[cnt[mps[k]]++] = refs[k];
            memcpy(refs + a, tmp, sizeof(visref) * (b - a));
            /* mps sorted likewise, for the runs below */
            size_t m0 = 0;
            for (size_t m = 0; m < nb; ++m) {
                size_t end = cnt[m];
                for (size_t k = a + m0; k < a + end; ++k) mps[k] = (uint32_t)m;
                m0 = end;
            }
        }
        free(cnt);
        free(tmp);
    }
    size_t nblocks = 0;
    for (size_t t = 0; t < ntiles; ++t) nblocks += nblk[t];
    L->blocks = malloc(sizeof(block) * (nblocks ? nblocks : 1));
    L->tiles = malloc(sizeof(tile) * (ntile_used ? ntile_used : 1));
    size_t bi = 0, ti = 0;
    for (size_t t = 0; t < ntiles; ++t) {
        size_t a = tile_start[t], b = tile_start[t + 1];
        if (a == b) continue;
        size_t b0 = bi;
        size_t k = a;
        while (k < b) {
            size_t e = k;
            while (e < b && mps[e] == mps[k]) ++e;
            L->blocks[bi].minplane = mps[k];
            L->blocks[bi].start = k;
            L->blocks[bi].end = e;
            ++bi;
            k = e;
        }
        tile *T = &L->tiles[ti++];
        T->tu = t / L->ntv;
        T->tv = t % L->ntv;
        T->b0 = b0;
        T->b1 = bi;
        T->plo = L->blocks[b0].minplane;
        T->phi = L->blocks[bi - 1].minplane;
        T->edge = (T->tu << L->log2) + L->su > s->nu || (T->tv << L->log2) + L->sv > s->nv;
    }
    L->nblocks = bi;
    L->ntiles = ti;
    L->refs = refs;
    L->nrefs = acc;
    free(mps);
    free(nblk);
    free(tile_start);
}

static void free_layout(layout_t *L) {
    free(L->refs);
    free(L->blocks);
    free(L->tiles);
}

/* ------------------------------------------------------------- values */

/* wgridder.rs:1395-1408 (compute_phases) */
static inline CX shift_phase(const or_setup *s, const or_baselines *bl, visref r, double sign) {
    double b[3], imflip;
    or_fix_w(bl->coord + 3 * (size_t)r.row, b, &imflip);
    double fct = sign * imflip * (b[0] * s->lshift + b[1] * s->mshift + b[2] * s->nshift);
    double tmp = fct * bl->f_over_c[r.ch];
    REAL ang = SINGLE ? (REAL)(2.0 * M_PI * (tmp - floor(tmp))) : (REAL)(2.0 * M_PI * tmp);
    return cx(RCOS(ang), RSIN(ang));
}

/* wgridder.rs:1414-1424 */
static inline CX unit_value(const or_setup *s, const or_baselines *bl, visref r, int haswgt,
                            REAL wv) {
    CX v = cx(1, 0);
    if (s->shifting) v = cmul(v, shift_phase(s, bl, r, 1.0));
    if (haswgt) v = cscale(v, wv);
    double imflip = bl->coord[3 * (size_t)r.row + 2] < 0.0 ? -1.0 : 1.0;
    return cx(v.re, v.im * (REAL)imflip);
}

/* wgridder.rs:1432-1435 */
static inline CX data_value(CX v, CX p, REAL imflip) {
    CX d = cmul(v, cx(p.re, p.im * imflip));
    return cx(d.re, d.im * imflip);
}

/* --------------------------------------------------------- gridding */

#define CHUNK 2048

/* wgridder.rs:2012-2035 (accumulate) */
static inline __attribute__((always_inline)) void accumulate(REAL *bre, REAL *bim, CX val,
                                                              const REAL *ku, const REAL *kv,
                                                              size_t ou, size_t ov, size_t sv,
                                                              const int S) {
    for (int cu = 0; cu < S; ++cu) {
        REAL tr = val.re * ku[cu], ti = val.im * ku[cu];
        REAL *br = bre + (ou + cu) * sv + ov, *bi = bim + (ou + cu) * sv + ov;
        for (int cv = 0; cv < S; ++cv) br[cv] = RFMA(tr, kv[cv], br[cv]);
        for (int cv = 0; cv < S; ++cv) bi[cv] = RFMA(ti, kv[cv], bi[cv]);
    }
}

typedef struct {
    const or_setup *s;
    const or_baselines *bl;
    const layout_t *L;
    const CX *stored;
    visdata data; /* with both outputs: the data */
    int both;
    const REAL *coeff;
    int ne;
} gridin;

/* wgridder.rs:2102-2156 (grid_task_body) for one plane p: visibilities
 * refs[a..b) (runs of the tile) into the task buffers. */
static inline __attribute__((always_inline)) void grid_range(const gridin *g, size_t mp, size_t a,
                                                             size_t b, size_t p, REAL **bre,
                                                             REAL **bim, size_t bu, size_t bv,
                                                             const int S) {
    const or_setup *s = g->s;
    size_t sv = g->L->sv, nchan = g->bl->nchan;
    if (s->do_wgridding && (p < mp || p >= mp + (size_t)S)) return;
    for (size_t k = a; k < b; ++k) {
        visref r = g->L->refs[k];
        double uvw[3], imflip, uf, vf;
        long iu0, iv0;
        uvw_of(g->bl, r, uvw, &imflip);
        or_getpix(s, uvw[0], uvw[1], &uf, &vf, &iu0, &iv0);
        REAL x0 = (REAL)(-uf * 2.0 + ((double)S - 1.0));
        REAL y0 = (REAL)(-vf * 2.0 + ((double)S - 1.0));
        REAL ku[16], kv[16];
        eval_taps2(g->coeff, g->ne, S, x0, y0, ku, kv);
        size_t ou = offset_in(iu0, s->nu, bu), ov = offset_in(iv0, s->nv, bv);
        CX st = g->stored[k];
        if (s->do_wgridding) {
            REAL kw = w_factor(s, g->coeff, g->ne, (int)(p - mp), p, uvw[2]);
            for (int i = 0; i < S; ++i) ku[i] = ku[i] * kw;
        }
        if (g->both) {
            CX d = data_value(vis_get(g->data, (size_t)r.row * nchan + r.ch), st, (REAL)imflip);
            accumulate(bre[0], bim[0], d, ku, kv, ou, ov, sv, S);
            accumulate(bre[1], bim[1], st, ku, kv, ou, ov, sv, S);
        } else {
            accumulate(bre[0], bim[0], st, ku, kv, ou, ov, sv, S);
        }
    }
}

static void grid_range_s(const gridin *g, size_t mp, size_t a, size_t b, size_t p, REAL **bre,
                         REAL **bim, size_t bu, size_t bv) {
    switch (g->s->supp) {
#define CASE(S)                                                   \
    case S:                                                       \
        grid_range(g, mp, a, b, p, bre, bim, bu, bv, S);          \
        break;
        CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
        CASE(14) CASE(15) CASE(16)
#undef CASE
    default:
        abort();
    }
}

/* wgridder.rs:1871-1875 */
static inline size_t cell(const or_setup *s, size_t bu, size_t bv, size_t i, size_t j) {
    size_t r = (bu + i + s->nu / 2) % s->nu;
    size_t c = (bv + j + s->nv / 2) % s->nv;
    return r * s->nv + c;
}

/* wgridder.rs:1904-1937 (dump) */
static void dump(const or_setup *s, const layout_t *L, const tile *t, REAL *bre, REAL *bim,
                 CX *grid) {
    size_t bu = t->tu << L->log2, bv = t->tv << L->log2;
    for (size_t i = 0; i < L->su; ++i)
        for (size_t j = 0; j < L->sv; ++j) {
            CX *gp = grid + cell(s, bu, bv, i, j);
            gp->re = gp->re + bre[i * L->sv + j];
            gp->im = gp->im + bim[i * L->sv + j];
        }
    memset(bre, 0, sizeof(REAL) * L->su * L->sv);
    memset(bim, 0, sizeof(REAL) * L->su * L->sv);
}

static inline int tile_on_plane(const or_setup *s, const tile *t, size_t p) {
    return !s->do_wgridding || (t->plo <= p && p < t->phi + (size_t)s->supp);
}

/* One tile on plane p: its tasks of at most CHUNK visibilities (runs active
 * on p), each into zeroed buffers then merged into the grids in order
 * (wgridder.rs:1974-2008 plane_tasks, 2163-2221 grid_group). */
static void grid_tile(const gridin *g, const tile *t, size_t p, int nout, REAL **bre, REAL **bim,
                      CX **grids) {
    const or_setup *s = g->s;
    const layout_t *L = g->L;
    size_t bu = t->tu << L->log2, bv = t->tv << L->log2;
    size_t n = 0;
    for (size_t bi = t->b0; bi < t->b1; ++bi) {
        const block *B = &L->blocks[bi];
        if (s->do_wgridding && !(B->minplane <= p && p < B->minplane + (size_t)s->supp)) continue;
        size_t st = B->start;
        while (st < B->end) {
            size_t e = B->end < st + (CHUNK - n) ? B->end : st + (CHUNK - n);
            grid_range_s(g, B->minplane, st, e, p, bre, bim, bu, bv);
            n += e - st;
            st = e;
            if (n == CHUNK) {
                for (int o = 0; o < nout; ++o) dump(s, L, t, bre[o], bim[o], grids[o]);
                n = 0;
            }
        }
    }
    if (n > 0)
        for (int o = 0; o < nout; ++o) dump(s, L, t, bre[o], bim[o], grids[o]);
}

/* Grid plane p (grids zeroed on entry); marks the rows whose v transform
 * is needed (wgridder.rs:1818-1827 tile_rows). Returns whether any tile is
 * active on p. */
static int grid_plane(const gridin *g, size_t p, int nout, CX **grids, unsigned char *rows) {
    const or_setup *s = g->s;
    const layout_t *L = g->L;
    memset(rows, 0, s->nu);
    int any = 0;
    for (size_t ti = 0; ti < L->ntiles; ++ti) {
        const tile *t = &L->tiles[ti];
        if (!tile_on_plane(s, t, p)) continue;
        any = 1;
        size_t base = t->tu << L->log2;
        for (size_t i = 0; i < L->su; ++i) rows[(base + i + s->nu / 2) % s->nu] = 1;
    }
    if (!any) return 0;
    size_t nbuf = L->su * L->sv;
    for (int cls = 0; cls < 5; ++cls) {
        if (cls < 4) {
#pragma omp parallel
            {
                REAL *bre[2], *bim[2];
                for (int o = 0; o < nout; ++o) {
                    bre[o] = calloc(nbuf, sizeof(REAL));
                    bim[o] = calloc(nbuf, sizeof(REAL));
                }
#pragma omp for schedule(dynamic, 4)
                for (size_t ti = 0; ti < L->ntiles; ++ti) {
                    const tile *t = &L->tiles[ti];
                    if (t->edge || (int)((t->tu & 1) * 2 + (t->tv & 1)) != cls) continue;
                    if (!tile_on_plane(s, t, p)) continue;
                    grid_tile(g, t, p, nout, bre, bim, grids);
                }
                for (int o = 0; o < nout; ++o) {
                    free(bre[o]);
                    free(bim[o]);
                }
            }
        } else {
            REAL *bre[2], *bim[2];
            for (int o = 0; o < nout; ++o) {
                bre[o] = calloc(nbuf, sizeof(REAL));
                bim[o] = calloc(nbuf, sizeof(REAL));
            }
            for (size_t ti = 0; ti < L->ntiles; ++ti) {
                const tile *t = &L->tiles[ti];
                if (!t->edge || !tile_on_plane(s, t, p)) continue;
                grid_tile(g, t, p, nout, bre, bim, grids);
            }
            for (int o = 0; o < nout; ++o) {
                free(bre[o]);
                free(bim[o]);
            }
        }
    }
    return 1;
}

/* ------------------------------------------------------ image tables */

static void table_dims(const or_setup *s, size_t *tnx, size_t *tny) {
    if (s->lmshift) {
        *tnx = s->nx;
        *tny = s->ny;
    } else {
        *tnx = s->nx / 2 + 1;
        *tny = s->ny / 2 + 1;
    }
}

static inline size_t table_index(const or_setup *s, size_t i, size_t j, size_t tny) {
    if (s->lmshift) return i * tny + j;
    size_t a = i < s->nx - i ? i : s->nx - i;
    size_t b = j < s->ny - j ? j : s->ny - j;
    return a * tny + b;
}

/* wgridder.rs:806-818 (screen_table) */
static double *screen_table(const or_setup *s) {
    size_t tnx, tny;
    table_dims(s, &tnx, &tny);
    double *out = malloc(sizeof(double) * tnx * tny);
    double x0 = s->lshift - 0.5 * (double)s->nx * s->pixsize_x;
    double y0 = s->mshift - 0.5 * (double)s->ny * s->pixsize_y;
#pragma omp parallel for
    for (size_t i = 0; i < tnx; ++i) {
        double x = x0 + (double)i * s->pixsize_x;
        for (size_t j = 0; j < tny; ++j) {
            double y = y0 + (double)j * s->pixsize_y;
            out[i * tny + j] = or_nm1_of(x, y, s->r_l, s->r_m) + s->nshift;
        }
    }
    return out;
}

/* wgridder.rs:822-861 (global_corrections) */
static double *global_corrections(const or_setup *s) {
    size_t tnx, tny;
    table_dims(s, &tnx, &tny);
    size_t hx = s->nx / 2 + 1, hy = s->ny / 2 + 1;
    double *cfu = malloc(sizeof(double) * hx), *cfv = malloc(sizeof(double) * hy);
    for (size_t i = 0; i < hx; ++i) cfu[i] = or_corfunc(&s->kernel, (double)i * (1.0 / (double)s->nu));
    for (size_t i = 0; i < hy; ++i) cfv[i] = or_corfunc(&s->kernel, (double)i * (1.0 / (double)s->nv));
    double *out = malloc(sizeof(double) * tnx * tny);
    double x0 = s->lshift - 0.5 * (double)s->nx * s->pixsize_x;
    double y0 = s->mshift - 0.5 * (double)s->ny * s->pixsize_y;
#pragma omp parallel for
    for (size_t i = 0; i < tnx; ++i) {
        double x = x0 + (double)i * s->pixsize_x, xsq = x * x;
        size_t iu = s->nx / 2 - (i < s->nx - i ? i : s->nx - i);
        for (size_t j = 0; j < tny; ++j) {
            double y = y0 + (double)j * s->pixsize_y, ysq = y * y;
            double tmp = 1.0 - xsq - ysq, fct;
            if (tmp >= 0.0) {
                double nm1 = (-xsq - ysq) / (sqrt(tmp) + 1.0);
                fct = or_corfunc(&s->kernel, (nm1 + s->r_l * x + s->r_m * y + s->nshift) * s->dw);
                if (s->divide_by_n) fct /= nm1 + 1.0;
            } else if (s->divide_by_n) {
                fct = 0.0;
            } else {
                fct = or_corfunc(&s->kernel, (-sqrt(-tmp) - 1.0 + s->nshift) * s->dw);
            }
            size_t iv = s->ny / 2 - (j < s->ny - j ? j : s->ny - j);
            fct *= cfu[iu] * cfv[iv];
            out[i * tny + j] = fct;
        }
    }
    free(cfu);
    free(cfv);
    return out;
}

/* wgridder.rs:2296-2314 (screen_par) */
static void screen_par(const double *table, size_t n, double w, int adjoint, CX *out) {
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < n; ++t) {
        double phs = w * table[t];
        if (adjoint) phs *= -1.0;
        REAL ang = SINGLE ? (REAL)(2.0 * M_PI * (phs - floor(phs))) : (REAL)(2.0 * M_PI * phs);
        out[t] = cx(RCOS(ang), RSIN(ang));
    }
}

#define COL_BLOCK 8

/* wgridder.rs:2320-2399 (grid_to_image): v transforms of the touched rows,
 * u transforms of the image columns, projection added to the image. */
static void grid_to_image(const or_setup *s, const fftplan *fu, const fftplan *fv, CX *grid,
                          const unsigned char *touched, const CX *screen, REAL *image) {
    size_t nu = s->nu, nv = s->nv, nx = s->nx, ny = s->ny, tnx, tny;
    table_dims(s, &tnx, &tny);
#pragma omp parallel
    {
        CX *work = malloc(sizeof(CX) * (nu > nv ? nu : nv));
#pragma omp for schedule(static)
        for (size_t r = 0; r < nu; ++r)
            if (touched[r]) fft_exec(fv, grid + r * nv, work, +1);
        CX *col = malloc(sizeof(CX) * nu * COL_BLOCK);
        size_t nblk = (ny + COL_BLOCK - 1) / COL_BLOCK;
#pragma omp for schedule(dynamic)
        for (size_t blk = 0; blk < nblk; ++blk) {
            size_t j0 = blk * COL_BLOCK, nc = ny - j0 < COL_BLOCK ? ny - j0 : COL_BLOCK;
            for (size_t r = 0; r < nu; ++r)
                for (size_t c = 0; c < nc; ++c) {
                    size_t jx = (nv - ny / 2 + j0 + c) % nv;
                    col[c * nu + r] = touched[r] ? grid[r * nv + jx] : cx(0, 0);
                }
            for (size_t c = 0; c < nc; ++c) fft_exec(fu, col + c * nu, work, +1);
            for (size_t i = 0; i < nx; ++i) {
                size_t ix = (nu - nx / 2 + i) % nu;
                for (size_t c = 0; c < nc; ++c) {
                    size_t j = j0 + c;
                    CX gg = col[c * nu + ix];
                    REAL term;
                    if (s->do_wgridding) {
                        CX ph = screen[table_index(s, i, j, tny)];
                        term = gg.re * ph.re - gg.im * ph.im;
                    } else {
                        term = gg.re;
                    }
                    image[i * ny + j] = image[i * ny + j] + term;
                }
            }
        }
        free(col);
        free(work);
    }
}

/* wgridder.rs:2403-2428 (correct_image) */
static void correct_image(const or_setup *s, REAL *image, const double *corr, const double *cfu,
                          const double *cfv) {
    size_t nx = s->nx, ny = s->ny, tnx, tny;
    table_dims(s, &tnx, &tny);
#pragma omp parallel for
    for (size_t i = 0; i < nx; ++i) {
        if (s->do_wgridding) {
            for (size_t j = 0; j < ny; ++j) image[i * ny + j] = image[i * ny + j] * (REAL)corr[table_index(s, i, j, tny)];
        } else {
            size_t iu = nx / 2 - (i < nx - i ? i : nx - i);
            for (size_t j = 0; j < ny; ++j) {
                size_t iv = ny / 2 - (j < ny - j ? j : ny - j);
                image[i * ny + j] = (REAL)((double)image[i * ny + j] * cfu[iu] * cfv[iv]);
            }
        }
    }
}

static void uv_tables(const or_setup *s, double **cfu, double **cfv) {
    size_t hx = s->nx / 2 + 1, hy = s->ny / 2 + 1;
    *cfu = malloc(sizeof(double) * hx);
    *cfv = malloc(sizeof(double) * hy);
    for (size_t i = 0; i < hx; ++i) (*cfu)[i] = or_corfunc(&s->kernel, (double)i * (1.0 / (double)s->nu));
    for (size_t i = 0; i < hy; ++i) (*cfv)[i] = or_corfunc(&s->kernel, (double)i * (1.0 / (double)s->nv));
}

/* ------------------------------------------------------------ vis2dirty */

static int lengths_ok(const or_options *o, size_t nrow, size_t nchan) {
    return !or_check_common(o, nrow, nchan);
}

/* wgridder.rs:2436-2496 + 2584-2683 (vis2dirty_multi, grid_images_from) */
int FN(or_vis2dirty)(const or_options *o, size_t nrow, size_t nchan, const double *uvw,
                     const double *freq, const void *vis, int vis_kind, const void *wgt,
                     const void *wgt2, int wgt_kind, const uint8_t *mask, REAL *dirty, REAL *psf,
                     double *info) {
    if (!o || !uvw || !freq || !info || (!dirty && !psf)) {
        or_set_error("null array");
        return 1;
    }
    if (!lengths_ok(o, nrow, nchan)) return 1;
    visdata data = {vis_kind, vis};
    weights w = {wgt_kind, wgt, wgt2};
    if ((vis_kind != VIS_UNIT && !vis) || vis_kind < 0 || vis_kind > 2) {
        or_set_error("invalid visibility array or kind");
        return 1;
    }
    if (wgt_kind < 0 || wgt_kind > 3 || (wgt_kind != WGT_NONE && !wgt) || (wgt_kind == WGT_F32P && !wgt2)) {
        or_set_error("invalid weight array or kind");
        return 1;
    }
    double r_l, r_m;
    or_recentring(o, &r_l, &r_m);
    or_baselines bl;
    if (or_baselines_new(uvw, freq, nrow, nchan, o->flip_u, o->flip_v, !o->flip_w, r_l, r_m, &bl))
        return 1;
    int use_vis = psf == NULL;
    size_t nvis;
    double wmin_d, wmax_d;
    scan(&bl, use_vis, data, w, mask, &nvis, &wmin_d, &wmax_d);
    size_t npix = o->nx * o->ny;
    if (dirty) memset(dirty, 0, sizeof(REAL) * npix);
    if (psf) memset(psf, 0, sizeof(REAL) * npix);
    memset(info, 0, sizeof(double) * OR_N_INFO);
    if (nvis == 0) {
        or_baselines_free(&bl);
        return 0;
    }
    or_setup s;
    if (or_setup_new(o, SINGLE, WMAXT, nvis, wmin_d, wmax_d, r_l, r_m, &s)) {
        or_baselines_free(&bl);
        return 1;
    }
    layout_t L;
    make_layout(&s, &bl, use_vis, data, w, mask, &L);
    int ne;
    REAL *coeff = coeff_even_odd(&s.kernel, &ne);
    /* values (wgridder.rs:1439-1475) */
    CX *stored = malloc(sizeof(CX) * (L.nrefs ? L.nrefs : 1));
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < L.nrefs; ++k) {
        visref r = L.refs[k];
        size_t idx = (size_t)r.row * nchan + r.ch;
        REAL wv = 0;
        int hw = wgt_get(w, idx, &wv);
        CX p = unit_value(&s, &bl, r, hw, wv);
        if (psf) {
            stored[k] = p;
        } else {
            REAL imflip = bl.coord[3 * (size_t)r.row + 2] < 0.0 ? -1 : 1;
            stored[k] = data_value(vis_get(data, idx), p, imflip);
        }
    }
    gridin g = {&s, &bl, &L, stored, data, dirty && psf, coeff, ne};
    REAL *images[2];
    int nout = 0;
    if (dirty) images[nout++] = dirty;
    if (psf) images[nout++] = psf;
    size_t nu = s.nu, nv = s.nv;
    CX *grids[2];
    for (int i = 0; i < nout; ++i) grids[i] = calloc(nu * nv, sizeof(CX));
    unsigned char *rows = malloc(nu);
    fftplan fu, fv;
    fft_plan_init(&fu, nu);
    fft_plan_init(&fv, nv);
    double *table = s.do_wgridding ? screen_table(&s) : NULL;
    size_t tnx, tny;
    table_dims(&s, &tnx, &tny);
    CX *screen = s.do_wgridding ? malloc(sizeof(CX) * tnx * tny) : NULL;
    size_t nplanes = s.nplanes > 1 ? s.nplanes : 1;
    for (size_t p = 0; p < nplanes; ++p) {
        if (!grid_plane(&g, p, nout, grids, rows)) continue;
        if (s.do_wgridding) screen_par(table, tnx * tny, s.wmin + (double)p * s.dw, 1, screen);
        for (int oi = 0; oi < nout; ++oi) {
            grid_to_image(&s, &fu, &fv, grids[oi], rows, screen, images[oi]);
#pragma omp parallel for schedule(static)
            for (size_t r = 0; r < nu; ++r)
                if (rows[r]) memset(grids[oi] + r * nv, 0, sizeof(CX) * nv);
        }
    }
    double *corr = NULL, *cfu = NULL, *cfv = NULL;
    if (s.do_wgridding)
        corr = global_corrections(&s);
    else
        uv_tables(&s, &cfu, &cfv);
    for (int oi = 0; oi < nout; ++oi) correct_image(&s, images[oi], corr, cfu, cfv);
    memcpy(info, s.info, sizeof(double) * OR_N_INFO);
    free(corr);
    free(cfu);
    free(cfv);
    free(table);
    free(screen);
    fft_plan_free(&fu);
    fft_plan_free(&fv);
    free(rows);
    for (int i = 0; i < nout; ++i) free(grids[i]);
    free(stored);
    free(coeff);
    free_layout(&L);
    or_setup_free(&s);
    or_baselines_free(&bl);
    return 0;
}

/* ------------------------------------------------------------ dirty2vis */

/* wgridder.rs:2908-2992 (image_to_grid) */
static void image_to_grid(const or_setup *s, const fftplan *fu, const fftplan *fv, const REAL *image,
                          const CX *screen, const unsigned char *needed, CX *grid) {
    size_t nu = s->nu, nv = s->nv, nx = s->nx, ny = s->ny, tnx, tny;
    table_dims(s, &tnx, &tny);
    long *inv_rows = malloc(sizeof(long) * nu);
    for (size_t r = 0; r < nu; ++r) inv_rows[r] = -1;
    for (size_t i = 0; i < nx; ++i) inv_rows[(nu - nx / 2 + i) % nu] = (long)i;
#pragma omp parallel
    {
#pragma omp for schedule(static)
        for (size_t r = 0; r < nu; ++r)
            if (needed[r]) memset(grid + r * nv + (ny - ny / 2), 0, sizeof(CX) * (nv - ny));
        CX *work = malloc(sizeof(CX) * (nu > nv ? nu : nv));
        CX *col = malloc(sizeof(CX) * nu * COL_BLOCK);
        size_t nblk = (ny + COL_BLOCK - 1) / COL_BLOCK;
#pragma omp for schedule(dynamic)
        for (size_t blk = 0; blk < nblk; ++blk) {
            size_t j0 = blk * COL_BLOCK, nc = ny - j0 < COL_BLOCK ? ny - j0 : COL_BLOCK;
            for (size_t r = 0; r < nu; ++r) {
                long i = inv_rows[r];
                for (size_t c = 0; c < nc; ++c) {
                    if (i < 0) {
                        col[c * nu + r] = cx(0, 0);
                        continue;
                    }
                    size_t j = j0 + c;
                    REAL d = image[(size_t)i * ny + j];
                    if (s->do_wgridding) {
                        CX ph = screen[table_index(s, (size_t)i, j, tny)];
                        col[c * nu + r] = cx(d * ph.re, d * ph.im);
                    } else {
                        col[c * nu + r] = cx(d, 0);
                    }
                }
            }
            for (size_t c = 0; c < nc; ++c) fft_exec(fu, col + c * nu, work, -1);
            for (size_t r = 0; r < nu; ++r)
                if (needed[r])
                    for (size_t c = 0; c < nc; ++c) grid[r * nv + (nv - ny / 2 + j0 + c) % nv] = col[c * nu + r];
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (size_t r = 0; r < nu; ++r)
            if (needed[r]) fft_exec(fv, grid + r * nv, work, -1);
        free(col);
        free(work);
    }
    free(inv_rows);
}

/* wgridder.rs:2793-2901 (degrid_task_body) for one plane: every tile active
 * on p, its visibilities adding their contribution to out. */
static inline __attribute__((always_inline)) void degrid_tile(const or_setup *s, const layout_t *L,
                                                              const or_baselines *bl,
                                                              const REAL *coeff, int ne,
                                                              const tile *t, size_t p, const CX *grid,
                                                              REAL *bre, REAL *bim, CX *out,
                                                              const int S) {
    size_t su = L->su, sv = L->sv, nchan = bl->nchan;
    size_t bu = t->tu << L->log2, bv = t->tv << L->log2;
    for (size_t i = 0; i < su; ++i)
        for (size_t j = 0; j < sv; ++j) {
            CX gg = grid[cell(s, bu, bv, i, j)];
            bre[i * sv + j] = gg.re;
            bim[i * sv + j] = gg.im;
        }
    for (size_t bi = t->b0; bi < t->b1; ++bi) {
        const block *B = &L->blocks[bi];
        size_t mp = B->minplane;
        if (s->do_wgridding && !(mp <= p && p < mp + (size_t)S)) continue;
        for (size_t k = B->start; k < B->end; ++k) {
            visref r = L->refs[k];
            double uvw[3], imflip, uf, vf;
            long iu0, iv0;
            uvw_of(bl, r, uvw, &imflip);
            or_getpix(s, uvw[0], uvw[1], &uf, &vf, &iu0, &iv0);
            REAL x0 = (REAL)(-uf * 2.0 + ((double)S - 1.0));
            REAL y0 = (REAL)(-vf * 2.0 + ((double)S - 1.0));
            REAL ku[16], kv[16];
            eval_taps2(coeff, ne, S, x0, y0, ku, kv);
            size_t ou = offset_in(iu0, s->nu, bu), ov = offset_in(iv0, s->nv, bv);
            if (s->do_wgridding) {
                REAL kw = w_factor(s, coeff, ne, (int)(p - mp), p, uvw[2]);
                for (int i = 0; i < S; ++i) ku[i] = ku[i] * kw;
            }
            REAL rr = 0, ri = 0;
            for (int cu = 0; cu < S; ++cu) {
                const REAL *gr = bre + (ou + cu) * sv + ov, *gi = bim + (ou + cu) * sv + ov;
                REAL tr = 0, tim = 0;
                for (int cv = 0; cv < S; ++cv) {
                    tr = RFMA(kv[cv], gr[cv], tr);
                    tim = RFMA(kv[cv], gi[cv], tim);
                }
                rr = RFMA(ku[cu], tr, rr);
                ri = RFMA(ku[cu], tim, ri);
            }
            size_t idx = (size_t)r.row * nchan + r.ch;
            CX c = cx(rr, ri * (REAL)imflip);
            out[idx] = cadd(c, out[idx]);
        }
    }
}

static void degrid_plane(const or_setup *s, const layout_t *L, const or_baselines *bl,
                         const REAL *coeff, int ne, size_t p, const CX *grid, CX *out) {
#pragma omp parallel
    {
        REAL *bre = malloc(sizeof(REAL) * L->su * L->sv), *bim = malloc(sizeof(REAL) * L->su * L->sv);
#pragma omp for schedule(dynamic, 4)
        for (size_t ti = 0; ti < L->ntiles; ++ti) {
            const tile *t = &L->tiles[ti];
            if (!tile_on_plane(s, t, p)) continue;
            switch (s->supp) {
#define CASE(S)                                                                      \
    case S:                                                                          \
        degrid_tile(s, L, bl, coeff, ne, t, p, grid, bre, bim, out, S);              \
        break;
                CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
                CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
            default:
                abort();
            }
        }
        free(bre);
        free(bim);
    }
}

/* wgridder.rs:2998-3163 (dirty2vis_multi, degrid_images) */
int FN(or_dirty2vis)(const or_options *o, size_t nrow, size_t nchan, const double *uvw,
                     const double *freq, const REAL *dirty, const void *wgt, int wgt_kind,
                     const uint8_t *mask, REAL *vis, double *info) {
    if (!o || !uvw || !freq || !info || !dirty || !vis) {
        or_set_error("null array");
        return 1;
    }
    if (!lengths_ok(o, nrow, nchan)) return 1;
    if (wgt_kind < 0 || wgt_kind > 2 || (wgt_kind != WGT_NONE && !wgt)) {
        or_set_error("invalid weight array or kind");
        return 1;
    }
    weights w = {wgt_kind, wgt, NULL};
    visdata none = {VIS_UNIT, NULL};
    double r_l, r_m;
    or_recentring(o, &r_l, &r_m);
    or_baselines bl;
    if (or_baselines_new(uvw, freq, nrow, nchan, o->flip_u, o->flip_v, !o->flip_w, r_l, r_m, &bl))
        return 1;
    size_t nvis;
    double wmin_d, wmax_d;
    scan(&bl, 0, none, w, mask, &nvis, &wmin_d, &wmax_d);
    size_t n = nrow * nchan;
    memset(vis, 0, sizeof(REAL) * 2 * n);
    memset(info, 0, sizeof(double) * OR_N_INFO);
    if (nvis == 0) {
        or_baselines_free(&bl);
        return 0;
    }
    or_setup s;
    if (or_setup_new(o, SINGLE, WMAXT, nvis, wmin_d, wmax_d, r_l, r_m, &s)) {
        or_baselines_free(&bl);
        return 1;
    }
    layout_t L;
    make_layout(&s, &bl, 0, none, w, mask, &L);
    int ne;
    REAL *coeff = coeff_even_odd(&s.kernel, &ne);
    size_t nx = s.nx, ny = s.ny, nu = s.nu, nv = s.nv, tnx, tny;
    table_dims(&s, &tnx, &tny);
    double *corr = NULL, *cfu = NULL, *cfv = NULL;
    if (s.do_wgridding)
        corr = global_corrections(&s);
    else
        uv_tables(&s, &cfu, &cfv);
    REAL *image = malloc(sizeof(REAL) * nx * ny);
#pragma omp parallel for
    for (size_t i = 0; i < nx; ++i) {
        if (s.do_wgridding) {
            for (size_t j = 0; j < ny; ++j)
                image[i * ny + j] = dirty[i * ny + j] * (REAL)corr[table_index(&s, i, j, tny)];
        } else {
            size_t iu = nx / 2 - (i < nx - i ? i : nx - i);
            for (size_t j = 0; j < ny; ++j) {
                size_t iv = ny / 2 - (j < ny - j ? j : ny - j);
                image[i * ny + j] = dirty[i * ny + j] * (REAL)(cfu[iu] * cfv[iv]);
            }
        }
    }
    double *table = s.do_wgridding ? screen_table(&s) : NULL;
    CX *screen = s.do_wgridding ? malloc(sizeof(CX) * tnx * tny) : NULL;
    CX *grid = calloc(nu * nv, sizeof(CX));
    unsigned char *needed = malloc(nu);
    fftplan fu, fv;
    fft_plan_init(&fu, nu);
    fft_plan_init(&fv, nv);
    CX *out = (CX *)vis;
    size_t nplanes = s.nplanes > 1 ? s.nplanes : 1;
    for (size_t p = 0; p < nplanes; ++p) {
        memset(needed, 0, nu);
        int any = 0;
        for (size_t ti = 0; ti < L.ntiles; ++ti) {
            const tile *t = &L.tiles[ti];
            if (!tile_on_plane(&s, t, p)) continue;
            any = 1;
            size_t base = t->tu << L.log2;
            for (size_t i = 0; i < L.su; ++i) needed[(base + i + nu / 2) % nu] = 1;
        }
        if (!any) continue;
        if (s.do_wgridding) screen_par(table, tnx * tny, s.wmin + (double)p * s.dw, 0, screen);
        image_to_grid(&s, &fu, &fv, image, screen, needed, grid);
        degrid_plane(&s, &L, &bl, coeff, ne, p, grid, out);
    }
    /* weights and the conjugate shift phase (wgridder.rs:3146-3159) */
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < L.nrefs; ++k) {
        visref r = L.refs[k];
        size_t idx = (size_t)r.row * nchan + r.ch;
        REAL tw = 1;
        if (!wgt_get(w, idx, &tw)) tw = 1;
        if (s.shifting) {
            CX ph = shift_phase(&s, &bl, r, -1.0);
            out[idx] = cmul(out[idx], cx(ph.re * tw, ph.im * tw));
        } else {
            out[idx] = cscale(out[idx], tw);
        }
    }
    memcpy(info, s.info, sizeof(double) * OR_N_INFO);
    fft_plan_free(&fu);
    fft_plan_free(&fv);
    free(needed);
    free(grid);
    free(screen);
    free(table);
    free(image);
    free(corr);
    free(cfu);
    free(cfv);
    free(coeff);
    free_layout(&L);
    or_setup_free(&s);
    or_baselines_free(&bl);
    return 0;
}

/* The oracle's FFT alone (tests check it against a direct DFT). */
void FN(or_fft)(size_t n, REAL *data, int sgn) {
    fftplan p;
    fft_plan_init(&p, n);
    CX *work = malloc(sizeof(CX) * n);
    fft_exec(&p, (CX *)data, work, sgn);
    free(work);
    fft_plan_free(&p);
}
