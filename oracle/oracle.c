/*
 * oracle.c -- CPU restatement of the HeddleMD MD-step algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * path in heddlemd_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls
 * into it and has no CPU fallback.
 *
 * Everything here is plain C99 in double precision unless a function name
 * ends in _f32, in which case it reproduces the reference's single-precision
 * arithmetic operation by operation (built with -ffp-contract=off so that
 * no fused multiply-add changes the rounding).
 *
 * Units: Hartree atomic units (k_B = 1, k_C = 1), as in the reference
 * (src/gpu/kernels.rs:176).  Lattice = [lx, ly, lz, xy, xz, yz] with box
 * vectors a = (lx,0,0), b = (xy,ly,0), c = (xz,yz,lz).
 *
 * Each function cites the reference file:line it restates.  How each one
 * is pinned (closed form, brute force, invariants) is listed in the test
 * files under tests/ (test_oracle_*.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Periodic boundaries (kernels/pbc.cuh:18-92)                         */
/* ------------------------------------------------------------------ */

static inline void cart_to_frac(double x, double y, double z, const double *L,
                                double *sa, double *sb, double *sc)
{
    /* pbc.cuh:18-27 -- back substitution */
    *sc = z / L[2];
    *sb = (y - *sc * L[5]) / L[1];
    *sa = (x - *sb * L[3] - *sc * L[4]) / L[0];
}

void or_min_image(double *dx, double *dy, double *dz, const double *L)
{
    /* pbc.cuh:30-43 */
    double sa, sb, sc;
    cart_to_frac(*dx, *dy, *dz, L, &sa, &sb, &sc);
    double ka = floor(sa + 0.5), kb = floor(sb + 0.5), kc = floor(sc + 0.5);
    *dx -= ka * L[0] + kb * L[3] + kc * L[4];
    *dy -= kb * L[1] + kc * L[5];
    *dz -= kc * L[2];
}

void or_wrap_with_image(double *x, double *y, double *z, int *ka_o, int *kb_o,
                        int *kc_o, const double *L)
{
    /* pbc.cuh:74-92 */
    double sa, sb, sc;
    cart_to_frac(*x, *y, *z, L, &sa, &sb, &sc);
    double ka = floor(sa + 0.5), kb = floor(sb + 0.5), kc = floor(sc + 0.5);
    *x -= ka * L[0] + kb * L[3] + kc * L[4];
    *y -= kb * L[1] + kc * L[5];
    *z -= kc * L[2];
    *ka_o = (int)ka;
    *kb_o = (int)kb;
    *kc_o = (int)kc;
}

/* single-precision twins, operation for operation (pbc.cuh with Real=float) */
static inline void cart_to_frac_f(float x, float y, float z, const float *L,
                                  float *sa, float *sb, float *sc)
{
    *sc = z / L[2];
    *sb = (y - *sc * L[5]) / L[1];
    *sa = (x - *sb * L[3] - *sc * L[4]) / L[0];
}

static inline void min_image_f(float *dx, float *dy, float *dz, const float *L)
{
    float sa, sb, sc;
    cart_to_frac_f(*dx, *dy, *dz, L, &sa, &sb, &sc);
    float ka = floorf(sa + 0.5f), kb = floorf(sb + 0.5f), kc = floorf(sc + 0.5f);
    *dx -= ka * L[0] + kb * L[3] + kc * L[4];
    *dy -= kb * L[1] + kc * L[5];
    *dz -= kc * L[2];
}

static inline void wrap_with_image_f(float *x, float *y, float *z, int *ka_o,
                                     int *kb_o, int *kc_o, const float *L)
{
    float sa, sb, sc;
    cart_to_frac_f(*x, *y, *z, L, &sa, &sb, &sc);
    float ka = floorf(sa + 0.5f), kb = floorf(sb + 0.5f), kc = floorf(sc + 0.5f);
    *x -= ka * L[0] + kb * L[3] + kc * L[4];
    *y -= kb * L[1] + kc * L[5];
    *z -= kc * L[2];
    *ka_o = (int)ka;
    *kb_o = (int)kb;
    *kc_o = (int)kc;
}

/* ------------------------------------------------------------------ */
/* Pair enumeration                                                    */
/* ------------------------------------------------------------------ */

/* Visit every unordered pair i<j whose minimum-image distance squared is
 * <= rmax2.  Orthorhombic boxes with >= 3 cells per axis use a cell grid;
 * everything else is the O(N^2) all-pairs sweep (the reference's `Trivial`
 * mode, src/forces/neighbor_list.rs:659).  The visit order does not matter
 * to callers: they accumulate in double and the tests compare at
 * tolerance. */
typedef void (*pair_visit_fn)(int i, int j, double dx, double dy, double dz,
                              double r2, void *ctx);

static void for_each_pair(int n, const double *x, const double *y,
                          const double *z, const double *L, double rmax2,
                          pair_visit_fn fn, void *ctx)
{
    int ortho = (L[3] == 0.0 && L[4] == 0.0 && L[5] == 0.0);
    double rmax = sqrt(rmax2);
    int nc[3] = {0, 0, 0};
    if (ortho) {
        for (int d = 0; d < 3; ++d) nc[d] = (int)floor(L[d] / rmax);
    }
    if (!ortho || nc[0] < 3 || nc[1] < 3 || nc[2] < 3) {
        for (int i = 0; i < n; ++i) {
            for (int j = i + 1; j < n; ++j) {
                double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
                or_min_image(&dx, &dy, &dz, L);
                double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= rmax2) fn(i, j, dx, dy, dz, r2, ctx);
            }
        }
        return;
    }
    int ncell = nc[0] * nc[1] * nc[2];
    int *head = (int *)malloc(sizeof(int) * ncell);
    int *next = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *cidx = (int *)malloc(sizeof(int) * 3 * (n > 0 ? n : 1));
    for (int c = 0; c < ncell; ++c) head[c] = -1;
    for (int i = n - 1; i >= 0; --i) {
        int ci[3];
        double p[3] = {x[i], y[i], z[i]};
        for (int d = 0; d < 3; ++d) {
            double s = p[d] / L[d];
            s -= floor(s + 0.5);
            int k = (int)floor((s + 0.5) * nc[d]);
            if (k < 0) k = 0;
            if (k >= nc[d]) k = nc[d] - 1;
            ci[d] = k;
            cidx[3 * i + d] = k;
        }
        int c = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
        next[i] = head[c];
        head[c] = i;
    }
    for (int i = 0; i < n; ++i) {
        int a = cidx[3 * i], b = cidx[3 * i + 1], cc = cidx[3 * i + 2];
        for (int da = -1; da <= 1; ++da)
            for (int db = -1; db <= 1; ++db)
                for (int dc = -1; dc <= 1; ++dc) {
                    int na = (a + da + nc[0]) % nc[0];
                    int nb = (b + db + nc[1]) % nc[1];
                    int nn = (cc + dc + nc[2]) % nc[2];
                    int c = (na * nc[1] + nb) * nc[2] + nn;
                    for (int j = head[c]; j >= 0; j = next[j]) {
                        if (j <= i) continue;
                        double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
                        or_min_image(&dx, &dy, &dz, L);
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 <= rmax2) fn(i, j, dx, dy, dz, r2, ctx);
                    }
                }
    }
    free(head);
    free(next);
    free(cidx);
}

/* Canonical pair set {(i,j): i<j, r2 <= rs2} (the set P_s of the packed
 * list contract, src/forces/neighbor_list.rs:1366-1506).  Pairs whose r2
 * lies within `band` (relative) of rs2 are written to band_out instead:
 * their membership depends on f32 rounding order.  Returns the number of
 * set pairs (which may exceed cap; only cap are written). */
typedef struct {
    double rs2, band;
    int64_t cap, n_in, bcap, n_band;
    uint32_t *out, *band_out;
} pairset_ctx;

static void pairset_visit(int i, int j, double dx, double dy, double dz,
                          double r2, void *vctx)
{
    (void)dx; (void)dy; (void)dz;
    pairset_ctx *c = (pairset_ctx *)vctx;
    if (fabs(r2 - c->rs2) <= c->band * c->rs2) {
        if (c->n_band < c->bcap) {
            c->band_out[2 * c->n_band] = (uint32_t)i;
            c->band_out[2 * c->n_band + 1] = (uint32_t)j;
        }
        c->n_band++;
        return;
    }
    if (r2 <= c->rs2) {
        if (c->n_in < c->cap) {
            c->out[2 * c->n_in] = (uint32_t)i;
            c->out[2 * c->n_in + 1] = (uint32_t)j;
        }
        c->n_in++;
    }
}

int64_t or_pair_set(int n, const float *posq, const double *L, double rs2,
                    double band, int64_t cap, uint32_t *out, int64_t bcap,
                    uint32_t *band_out, int64_t *n_band)
{
    double *x = (double *)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    double *y = x + n, *z = y + n;
    for (int i = 0; i < n; ++i) {
        x[i] = posq[4 * i];
        y[i] = posq[4 * i + 1];
        z[i] = posq[4 * i + 2];
    }
    pairset_ctx c = {rs2, band, cap, 0, bcap, 0, out, band_out};
    double rmax2 = rs2 * (1.0 + 2.0 * band);
    for_each_pair(n, x, y, z, L, rmax2, pairset_visit, &c);
    free(x);
    *n_band = c.n_band;
    return c.n_in;
}

/* ------------------------------------------------------------------ */
/* Non-bonded pair forces: LJ (src/forces/lj.rs:225-335) plus          */
/* real-space Ewald (src/forces/spme.rs:708-783), exclusion scale       */
/* (src/forces/jit_composed.rs:1445-1458, 1130-1206).                  */
/* ------------------------------------------------------------------ */

/* 0: exact erfc (the reference's f64 path); 1: Hastings polynomial. */
static int g_erfc_hastings = 0;
void or_set_erfc_mode(int hastings) { g_erfc_hastings = hastings; }

typedef struct {
    const uint32_t *type;
    int ntypes;
    const double *sigma, *eps, *cut, *rsw; /* ntypes*ntypes each */
    double kc, alpha, rc_coul;             /* rc_coul <= 0 disables Coulomb */
    const double *q;
    const uint32_t *ex_off, *ex_partner;   /* CSR exclusions (may be NULL) */
    const double *ex_scale;                /* LJ scale per CSR entry */
    const double *ex_scale_coul;           /* real-space Coulomb scale per entry */
    double band;
    double *F, *E, *W;
    uint8_t *band_flag;
} pairforce_ctx;

/* heddle_jit_exclusion_scale (src/forces/jit_composed.rs:1445-1458): a
 * linear scan of i's CSR row; unlisted pairs have scale 1. */
static double excl_scale_of(const pairforce_ctx *c, const double *tab, int i, int j)
{
    if (!c->ex_off) return 1.0;
    for (uint32_t m = c->ex_off[i]; m < c->ex_off[i + 1]; ++m)
        if ((int)c->ex_partner[m] == j) return tab[m];
    return 1.0;
}

static void pairforce_visit(int i, int j, double dx, double dy, double dz,
                            double r2, void *vctx)
{
    pairforce_ctx *c = (pairforce_ctx *)vctx;
    double factor = 0.0, energy = 0.0;
    double inv_r = 1.0 / sqrt(r2);
    double r = r2 * inv_r;
    double sc = excl_scale_of(c, c->ex_scale, i, j);
    double sc_coul = excl_scale_of(c, c->ex_scale_coul, i, j);
    /* LJ fragment, per-type-pair cutoff and optional cubic switch */
    {
        int p = (int)(c->type[i] * (uint32_t)c->ntypes + c->type[j]);
        double cut = c->cut[p];
        double cut2 = cut * cut;
        if (c->band > 0.0 && fabs(r2 - cut2) <= c->band * cut2) {
            c->band_flag[i] = 1;
            c->band_flag[j] = 1;
        }
        if (r2 <= cut2 && c->eps[p] != 0.0) {
            double sigma = c->sigma[p], eps = c->eps[p];
            double inv_r2 = inv_r * inv_r;
            double sr2 = sigma * sigma * inv_r2;
            double sr6 = sr2 * sr2 * sr2;
            double sr12 = sr6 * sr6;
            double f = 24.0 * eps * inv_r2 * (2.0 * sr12 - sr6);
            double e = 4.0 * eps * (sr12 - sr6);
            double rs = c->rsw[p];
            double rs2 = rs * rs, delta = cut2 - rs2;
            if (delta > 0.0 && r2 > rs2) {
                double tau = (r2 - rs2) / delta;
                double omt = 1.0 - tau;
                double s = omt * omt * (1.0 + 2.0 * tau);
                double chain = 12.0 * tau * omt / delta;
                f = s * f + chain * e;
                e = s * e;
            }
            factor += f * sc;
            energy += e * sc;
        }
    }
    /* real-space Coulomb fragment (exact erfc in double) */
    if (c->rc_coul > 0.0) {
        double rc2 = c->rc_coul * c->rc_coul;
        if (c->band > 0.0 && fabs(r2 - rc2) <= c->band * rc2) {
            c->band_flag[i] = 1;
            c->band_flag[j] = 1;
        }
        if (r2 <= rc2) {
            double qq = c->q[i] * c->q[j];
            double ar = c->alpha * r;
            double gauss = exp(-ar * ar);
            double ec;
            if (g_erfc_hastings) {
                /* A&S 7.1.26 (the reference's f32 path, src/forces/spme.rs:744-770) */
                double t = 1.0 / (1.0 + 0.3275911 * ar);
                ec = (0.254829592 + (-0.284496736 + (1.421413741 + (-1.453152027 +
                      1.061405429 * t) * t) * t) * t) * t * gauss;
            } else {
                ec = erfc(ar);
            }
            double e = c->kc * qq * ec * inv_r;
            double f = c->kc * qq * inv_r * inv_r *
                       (ec * inv_r + 2.0 * c->alpha * 0.5641895835477563 * gauss);
            factor += f * sc_coul;
            energy += e * sc_coul;
        }
    }
    double virial = factor * r2;
    c->F[3 * i] += factor * dx;
    c->F[3 * i + 1] += factor * dy;
    c->F[3 * i + 2] += factor * dz;
    c->F[3 * j] -= factor * dx;
    c->F[3 * j + 1] -= factor * dy;
    c->F[3 * j + 2] -= factor * dz;
    c->E[i] += 0.5 * energy;
    c->E[j] += 0.5 * energy;
    c->W[i] += 0.5 * virial;
    c->W[j] += 0.5 * virial;
}

/* Per-atom forces F[3n] (interleaved), energy E[n] and virial W[n]
 * (pair terms split half/half, src/forces/jit_composed.rs:1725-1733).
 * band_flag[n] marks atoms in a pair whose r2 lies within `band`
 * (relative) of a cutoff squared. */
void or_pair_forces(int n, const float *posq, const uint32_t *type, int ntypes,
                    const double *sigma, const double *eps, const double *cut,
                    const double *rsw, double kc, double alpha, double rc_coul,
                    const uint32_t *ex_off, const uint32_t *ex_partner,
                    const double *ex_scale, const double *ex_scale_coul, const double *L,
                    double band, double *F, double *E, double *W, uint8_t *band_flag)
{
    double *x = (double *)malloc(sizeof(double) * 4 * (n > 0 ? n : 1));
    double *y = x + n, *z = y + n, *q = z + n;
    double rmax = rc_coul > 0.0 ? rc_coul : 0.0;
    for (int p = 0; p < ntypes * ntypes; ++p)
        if (cut[p] > rmax) rmax = cut[p];
    for (int i = 0; i < n; ++i) {
        x[i] = posq[4 * i];
        y[i] = posq[4 * i + 1];
        z[i] = posq[4 * i + 2];
        q[i] = posq[4 * i + 3];
    }
    memset(F, 0, sizeof(double) * 3 * n);
    memset(E, 0, sizeof(double) * n);
    memset(W, 0, sizeof(double) * n);
    if (band_flag) memset(band_flag, 0, n);
    pairforce_ctx c = {type, ntypes, sigma, eps, cut, rsw, kc, alpha, rc_coul,
                       q, ex_off, ex_partner, ex_scale, ex_scale_coul ? ex_scale_coul : ex_scale,
                       band_flag ? band : 0.0, F, E, W, band_flag};
    double rmax2 = rmax * rmax * (1.0 + 4.0 * (band > 0.0 ? band : 0.0));
    for_each_pair(n, x, y, z, L, rmax2, pairforce_visit, &c);
    free(x);
}

/* ------------------------------------------------------------------ */
/* SPME reciprocal space                                               */
/* ------------------------------------------------------------------ */

/* Cardinal B-spline M_p(x) by Cox-de Boor (src/forces/spme.rs:589-611,
 * kernels/spme_spread_gather.cuh:36-54). */
double or_bspline(int p, double x)
{
    double vals[16];
    for (int i = 0; i < p; ++i) {
        double xi = x - i;
        vals[i] = (xi >= 0.0 && xi < 1.0) ? 1.0 : 0.0;
    }
    for (int order = 2; order <= p; ++order) {
        double inv = 1.0 / (order - 1.0);
        for (int i = 0; i < p - order + 1; ++i) {
            double xi = x - i;
            vals[i] = xi * inv * vals[i] + (order - xi) * inv * vals[i + 1];
        }
    }
    return vals[0];
}

/* M_p(x) and its derivative M_{p-1}(x) - M_{p-1}(x-1)
 * (kernels/spme_spread_gather.cuh:56-86). */
static void bspline_wd(int p, double x, double *w, double *dw)
{
    *w = or_bspline(p, x);
    *dw = or_bspline(p - 1, x) - or_bspline(p - 1, x - 1.0);
}

/* |b(k)|^2 structure-factor corrections (src/forces/spme.rs:524-585),
 * including the Essmann zero-modulus fix-up for odd orders. */
void or_b_factors(int n, int p, double *out)
{
    double msamp[16];
    for (int j = 1; j < p; ++j) msamp[j - 1] = or_bspline(p, (double)j);
    double *mod = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
        double th = 2.0 * M_PI * k / n, re = 0.0, im = 0.0;
        for (int j = 0; j < p - 1; ++j) {
            re += msamp[j] * cos(th * j);
            im += msamp[j] * sin(th * j);
        }
        mod[k] = re * re + im * im;
    }
    for (int k = 0; k < n; ++k)
        if (mod[k] < 1e-7) mod[k] = 0.5 * (mod[(k + n - 1) % n] + mod[(k + 1) % n]);
    for (int k = 0; k < n; ++k) out[k] = mod[k] > 0.0 ? 1.0 / mod[k] : 0.0;
    free(mod);
}

/* In-place complex FFT of length n with stride (sign -1 forward, +1
 * inverse, unnormalised).  Radix-2 for powers of two, direct DFT
 * otherwise. */
static void fft1d(double *re, double *im, int n, int stride, int sign,
                  double *tre, double *tim)
{
    for (int k = 0; k < n; ++k) {
        tre[k] = re[k * stride];
        tim[k] = im[k * stride];
    }
    if ((n & (n - 1)) == 0) {
        for (int i = 1, j = 0; i < n; ++i) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t = tre[i]; tre[i] = tre[j]; tre[j] = t;
                t = tim[i]; tim[i] = tim[j]; tim[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            double ang = sign * 2.0 * M_PI / len;
            for (int i = 0; i < n; i += len) {
                for (int k = 0; k < len / 2; ++k) {
                    double wr = cos(ang * k), wi = sin(ang * k);
                    int a = i + k, b = i + k + len / 2;
                    double xr = tre[b] * wr - tim[b] * wi;
                    double xi = tre[b] * wi + tim[b] * wr;
                    tre[b] = tre[a] - xr; tim[b] = tim[a] - xi;
                    tre[a] += xr; tim[a] += xi;
                }
            }
        }
        for (int k = 0; k < n; ++k) {
            re[k * stride] = tre[k];
            im[k * stride] = tim[k];
        }
    } else {
        for (int k = 0; k < n; ++k) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; ++j) {
                double ang = sign * 2.0 * M_PI * (double)((long)k * j % n) / n;
                double c = cos(ang), s = sin(ang);
                sr += tre[j] * c - tim[j] * s;
                si += tre[j] * s + tim[j] * c;
            }
            re[k * stride] = sr;
            im[k * stride] = si;
        }
    }
}

/* 3D complex FFT over a row-major (a,b,c) grid, c fastest. */
void or_fft3d(double *re, double *im, int na, int nb, int nc, int sign)
{
    int nmax = na > nb ? na : nb;
    if (nc > nmax) nmax = nc;
    double *tre = (double *)malloc(sizeof(double) * nmax);
    double *tim = (double *)malloc(sizeof(double) * nmax);
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
            size_t o = ((size_t)a * nb + b) * nc;
            fft1d(re + o, im + o, nc, 1, sign, tre, tim);
        }
    for (int a = 0; a < na; ++a)
        for (int c = 0; c < nc; ++c) {
            size_t o = (size_t)a * nb * nc + c;
            fft1d(re + o, im + o, nb, nc, sign, tre, tim);
        }
    for (int b = 0; b < nb; ++b)
        for (int c = 0; c < nc; ++c) {
            size_t o = (size_t)b * nc + c;
            fft1d(re + o, im + o, na, nb * nc, sign, tre, tim);
        }
    free(tre);
    free(tim);
}

static inline int imod(long a, int n)
{
    long r = a % n;
    return (int)(r < 0 ? r + n : r);
}

/* Per-atom spline setup (kernels/spme_spread_gather.cuh:90-126): wrap,
 * fractional coordinates shifted by +1/2, u = s*n, primary bin floor(u),
 * weights M_p(d + t) at grid index (floor(u) - d) mod n. */
static void spline_setup(double px, double py, double pz, const double *L,
                         const int *ng, double *t, long *g0)
{
    int ia, ib, ic;
    or_wrap_with_image(&px, &py, &pz, &ia, &ib, &ic, L);
    double s[3];
    cart_to_frac(px, py, pz, L, &s[0], &s[1], &s[2]);
    for (int d = 0; d < 3; ++d) {
        double u = (s[d] + 0.5) * ng[d];
        double f = floor(u);
        t[d] = u - f;
        g0[d] = (long)f;
    }
}

/* Charge spread in double (the reference accumulates i64 fixed point at
 * 2^32, kernels/spme_spread_gather.cuh:151-221; the oracle keeps the
 * exact sum). rho has na*nb*nc entries. */
void or_spme_spread(int n, const float *posq, const double *L, int na, int nb,
                    int nc, int p, double *rho)
{
    int ng[3] = {na, nb, nc};
    memset(rho, 0, sizeof(double) * (size_t)na * nb * nc);
    for (int i = 0; i < n; ++i) {
        double q = posq[4 * i + 3];
        if (q == 0.0) continue;
        double t[3];
        long g0[3];
        spline_setup(posq[4 * i], posq[4 * i + 1], posq[4 * i + 2], L, ng, t, g0);
        double w[3][16];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < p; ++k) w[d][k] = or_bspline(p, k + t[d]);
        for (int da = 0; da < p; ++da) {
            int ga = imod(g0[0] - da, na);
            for (int db = 0; db < p; ++db) {
                int gb = imod(g0[1] - db, nb);
                for (int dc = 0; dc < p; ++dc) {
                    int gc = imod(g0[2] - dc, nc);
                    rho[((size_t)ga * nb + gb) * nc + gc] += q * w[0][da] * w[1][db] * w[2][dc];
                }
            }
        }
    }
}

/* Reciprocal-lattice rows of H^{-T} (kernels/spme_spread_gather.cuh:128-149) */
static void recip_rows(const double *L, double B[3][3])
{
    double ilx = 1.0 / L[0], ily = 1.0 / L[1], ilz = 1.0 / L[2];
    B[0][0] = ilx;
    B[0][1] = -L[3] * ilx * ily;
    B[0][2] = (L[3] * L[5] - L[4] * L[1]) * ilx * ily * ilz;
    B[1][0] = 0.0; B[1][1] = ily; B[1][2] = -L[5] * ily * ilz;
    B[2][0] = 0.0; B[2][1] = 0.0; B[2][2] = ilz;
}

/* Full SPME reciprocal evaluation in double:
 *   spread -> FFT -> influence multiply (kernels/spme_recip.cu:34-115,
 *   296-342) -> inverse FFT -> gather (kernels/spme_spread_gather.cuh:
 *   223-334), self energy k_C alpha/sqrt(pi) q^2 (src/forces/spme.rs:815-823).
 * Outputs per-atom F[3n], E[n] (1/2 q phi - u_self), W[n] (= W_recip/N),
 * and the totals e_recip (without self term) and w_recip.  Optionally
 * returns the potential grid V (na*nb*nc) when v_out != NULL. */
void or_spme_recip(int n, const float *posq, const double *L, int na, int nb,
                   int nc, int p, double alpha, double kc, double *F,
                   double *E, double *W, double *e_recip, double *w_recip,
                   double *v_out)
{
    size_t M = (size_t)na * nb * nc;
    double *re = (double *)malloc(sizeof(double) * M);
    double *im = (double *)calloc(M, sizeof(double));
    double *ba = (double *)malloc(sizeof(double) * na);
    double *bb = (double *)malloc(sizeof(double) * nb);
    double *bc = (double *)malloc(sizeof(double) * nc);
    or_b_factors(na, p, ba);
    or_b_factors(nb, p, bb);
    or_b_factors(nc, p, bc);
    or_spme_spread(n, posq, L, na, nb, nc, p, re);
    or_fft3d(re, im, na, nb, nc, -1);
    double B[3][3];
    recip_rows(L, B);
    double vbox = L[0] * L[1] * L[2];
    double er = 0.0, wr = 0.0;
    for (int a = 0; a < na; ++a) {
        double ma = (a <= na / 2) ? a : a - (double)na;
        for (int b = 0; b < nb; ++b) {
            double mb = (b <= nb / 2) ? b : b - (double)nb;
            for (int c = 0; c < nc; ++c) {
                double mc = (c <= nc / 2) ? c : c - (double)nc;
                double kx = 2.0 * M_PI * (ma * B[0][0]);
                double ky = 2.0 * M_PI * (ma * B[0][1] + mb * B[1][1]);
                double kz = 2.0 * M_PI * (ma * B[0][2] + mb * B[1][2] + mc * B[2][2]);
                double k2 = kx * kx + ky * ky + kz * kz;
                size_t o = ((size_t)a * nb + b) * nc + c;
                double g = 0.0, vf = 0.0;
                if (k2 != 0.0) {
                    g = kc / vbox * (4.0 * M_PI / k2) * exp(-k2 / (4.0 * alpha * alpha)) *
                        ba[a] * bb[b] * bc[c];
                    vf = g * (1.0 - k2 / (2.0 * alpha * alpha));
                }
                double r2 = re[o] * re[o] + im[o] * im[o];
                er += 0.5 * g * r2;
                wr += 0.5 * vf * r2;
                re[o] *= g;
                im[o] *= g;
            }
        }
    }
    or_fft3d(re, im, na, nb, nc, +1);
    if (v_out) memcpy(v_out, re, sizeof(double) * M);
    int ng[3] = {na, nb, nc};
    double self_pref = kc * alpha / sqrt(M_PI);
    for (int i = 0; i < n; ++i) {
        double q = posq[4 * i + 3];
        double t[3];
        long g0[3];
        spline_setup(posq[4 * i], posq[4 * i + 1], posq[4 * i + 2], L, ng, t, g0);
        double w[3][16], dw[3][16];
        int gi[3][16];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < p; ++k) {
                bspline_wd(p, k + t[d], &w[d][k], &dw[d][k]);
                gi[d][k] = imod(g0[d] - k, ng[d]);
            }
        double phi = 0.0, ga = 0.0, gb = 0.0, gc = 0.0;
        for (int da = 0; da < p; ++da)
            for (int db = 0; db < p; ++db)
                for (int dc = 0; dc < p; ++dc) {
                    double v = re[((size_t)gi[0][da] * nb + gi[1][db]) * nc + gi[2][dc]];
                    phi += v * w[0][da] * w[1][db] * w[2][dc];
                    ga += v * dw[0][da] * w[1][db] * w[2][dc];
                    gb += v * w[0][da] * dw[1][db] * w[2][dc];
                    gc += v * w[0][da] * w[1][db] * dw[2][dc];
                }
        ga *= na; gb *= nb; gc *= nc;
        for (int d = 0; d < 3; ++d)
            F[3 * i + d] = -q * (ga * B[0][d] + gb * B[1][d] + gc * B[2][d]);
        E[i] = 0.5 * q * phi - self_pref * q * q;
        W[i] = n > 0 ? wr / n : 0.0;
    }
    *e_recip = er;
    *w_recip = wr;
    free(re); free(im); free(ba); free(bb); free(bc);
}

/* ------------------------------------------------------------------ */
/* Velocity Verlet, f32 and lossless (kernels/integrate.cu:24-170)     */
/* Reproduces the single-precision arithmetic exactly.                 */
/* ------------------------------------------------------------------ */

void or_vv_kick_drift_f32(int n, float *posq, int *img, float *vx, float *vy,
                          float *vz, const float *fx, const float *fy,
                          const float *fz, const float *mass, const float *L,
                          float dt)
{
    for (int i = 0; i < n; ++i) {
        float m = mass[i];
        float ax = fx[i] / m, ay = fy[i] / m, az = fz[i] / m;
        float half_dt = dt * 0.5f;
        float v0 = vx[i] + ax * half_dt;
        float v1 = vy[i] + ay * half_dt;
        float v2 = vz[i] + az * half_dt;
        vx[i] = v0; vy[i] = v1; vz[i] = v2;
        float px = posq[4 * i] + v0 * dt;
        float py = posq[4 * i + 1] + v1 * dt;
        float pz = posq[4 * i + 2] + v2 * dt;
        int ka, kb, kc;
        wrap_with_image_f(&px, &py, &pz, &ka, &kb, &kc, L);
        posq[4 * i] = px; posq[4 * i + 1] = py; posq[4 * i + 2] = pz;
        img[3 * i] += ka; img[3 * i + 1] += kb; img[3 * i + 2] += kc;
    }
}

void or_vv_kick_f32(int n, float *vx, float *vy, float *vz, const float *fx,
                    const float *fy, const float *fz, const float *mass,
                    float dt)
{
    for (int i = 0; i < n; ++i) {
        float m = mass[i];
        float ax = fx[i] / m, ay = fy[i] / m, az = fz[i] / m;
        float half_dt = dt * 0.5f;
        vx[i] += ax * half_dt;
        vy[i] += ay * half_dt;
        vz[i] += az * half_dt;
    }
}

/* ------------------------------------------------------------------ */
/* Harmonic angles (src/forces/angle.rs:258-313; per-atom thirds of the  */
/* energy and virial, src/forces/jit_composed.rs:2300-2345)              */
/* ------------------------------------------------------------------ */

void or_angle_forces(int n_angles, const uint32_t *ijk, const double *kth, const double *th0,
                     const float *posq, const double *L, double *F, double *E, double *W)
{
    for (int m = 0; m < n_angles; ++m) {
        uint32_t i = ijk[3 * m], j = ijk[3 * m + 1], k = ijk[3 * m + 2];
        double dxi = (double)posq[4 * i] - posq[4 * j];
        double dyi = (double)posq[4 * i + 1] - posq[4 * j + 1];
        double dzi = (double)posq[4 * i + 2] - posq[4 * j + 2];
        double dxk = (double)posq[4 * k] - posq[4 * j];
        double dyk = (double)posq[4 * k + 1] - posq[4 * j + 1];
        double dzk = (double)posq[4 * k + 2] - posq[4 * j + 2];
        or_min_image(&dxi, &dyi, &dzi, L);
        or_min_image(&dxk, &dyk, &dzk, L);
        double dij2 = dxi * dxi + dyi * dyi + dzi * dzi;
        double dkj2 = dxk * dxk + dyk * dyk + dzk * dzk;
        if (dij2 == 0.0 || dkj2 == 0.0) continue;
        double dij = sqrt(dij2), dkj = sqrt(dkj2);
        double inv_dd = 1.0 / (dij * dkj);
        double dot = dxi * dxk + dyi * dyk + dzi * dzk;
        double c = dot * inv_dd;
        if (c > 1.0) c = 1.0;
        if (c < -1.0) c = -1.0;
        double s2 = 1.0 - c * c;
        double sn = sqrt(s2 > 0.0 ? s2 : 0.0);
        if (sn < 1.0e-7) continue;
        double theta = atan2(dij * dkj * sn, dot);
        double dth = theta - th0[m];
        double g = -kth[m] * dth / sn;
        double ii = 1.0 / dij2, kk = 1.0 / dkj2;
        double fi[3] = {g * (c * ii * dxi - inv_dd * dxk), g * (c * ii * dyi - inv_dd * dyk),
                        g * (c * ii * dzi - inv_dd * dzk)};
        double fk[3] = {g * (c * kk * dxk - inv_dd * dxi), g * (c * kk * dyk - inv_dd * dyi),
                        g * (c * kk * dzk - inv_dd * dzi)};
        double u = 0.5 * kth[m] * dth * dth;
        double w = dxi * fi[0] + dyi * fi[1] + dzi * fi[2] + dxk * fk[0] + dyk * fk[1] + dzk * fk[2];
        for (int d = 0; d < 3; ++d) {
            F[3 * i + d] += fi[d];
            F[3 * k + d] += fk[d];
            F[3 * j + d] -= fi[d] + fk[d];
        }
        E[i] += u / 3.0; E[j] += u / 3.0; E[k] += u / 3.0;
        W[i] += w / 3.0; W[j] += w / 3.0; W[k] += w / 3.0;
    }
}

/* ------------------------------------------------------------------ */
/* Two-body bonded terms: harmonic (src/forces/harmonic_bond.rs:288-330) */
/* and Morse (src/forces/morse.rs:290-320).  fmag = -dU/dr / r applied   */
/* along d = r_i - r_j; w = fmag r^2; energy and virial split in halves. */
/* ------------------------------------------------------------------ */
void or_bond_forces(int n_bonds, const uint32_t *ij, const int *kind, const double *prm,
                    const float *posq, const double *L, double *F, double *E, double *W)
{
    for (int m = 0; m < n_bonds; ++m) {
        uint32_t i = ij[2 * m], j = ij[2 * m + 1];
        double d[3];
        for (int a = 0; a < 3; ++a) d[a] = (double)posq[4 * i + a] - posq[4 * j + a];
        or_min_image(&d[0], &d[1], &d[2], L);
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double r = sqrt(r2);
        if (r < 1.0e-7) continue;
        const double *p = prm + 3 * m;
        double fmag, u;
        if (kind[m] == 0) {
            double dr = r - p[1];
            fmag = -p[0] * dr / r;
            u = 0.5 * p[0] * dr * dr;
        } else {
            double e = exp(-p[1] * (r - p[2]));
            fmag = -2.0 * p[0] * p[1] * (1.0 - e) * e / r;
            u = p[0] * (1.0 - e) * (1.0 - e);
        }
        for (int a = 0; a < 3; ++a) {
            F[3 * i + a] += fmag * d[a];
            F[3 * j + a] -= fmag * d[a];
        }
        E[i] += 0.5 * u; E[j] += 0.5 * u;
        W[i] += 0.5 * fmag * r2; W[j] += 0.5 * fmag * r2;
    }
}

/* ------------------------------------------------------------------ */
/* Periodic dihedral U = k (1 + cos(n phi - phi0))                       */
/* (src/forces/dihedral.rs:360-460): b1 = r_i - r_j, b2 = r_k - r_j,      */
/* b3 = r_k - r_l, m = b1 x b2, n = b2 x b3, phi = atan2(|b2| b1.n, m.n); */
/* F_i = f |b2|/|m|^2 m, F_l = -f |b2|/|n|^2 n with f = k n sin(n phi -   */
/* phi0); F_j = (s-1) F_i - t F_l, F_k = -s F_i + (t-1) F_l with          */
/* s = b1.b2/|b2|^2, t = b3.b2/|b2|^2; W = b1.F_i + b2.F_k + (b2-b3).F_l. */
/* ------------------------------------------------------------------ */
void or_dihedral_forces(int n_dih, const uint32_t *ijkl, const double *prm, const float *posq,
                        const double *L, double *F, double *E, double *W)
{
    for (int m = 0; m < n_dih; ++m) {
        uint32_t a[4] = {ijkl[4 * m], ijkl[4 * m + 1], ijkl[4 * m + 2], ijkl[4 * m + 3]};
        double b1[3], b2[3], b3[3];
        for (int d = 0; d < 3; ++d) {
            b1[d] = (double)posq[4 * a[0] + d] - posq[4 * a[1] + d];
            b2[d] = (double)posq[4 * a[2] + d] - posq[4 * a[1] + d];
            b3[d] = (double)posq[4 * a[2] + d] - posq[4 * a[3] + d];
        }
        or_min_image(&b1[0], &b1[1], &b1[2], L);
        or_min_image(&b2[0], &b2[1], &b2[2], L);
        or_min_image(&b3[0], &b3[1], &b3[2], L);
        double mv[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2],
                        b1[0] * b2[1] - b1[1] * b2[0]};
        double nv[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2],
                        b2[0] * b3[1] - b2[1] * b3[0]};
        double m2 = mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
        double n2 = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
        double bb = b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2];
        if (m2 < 1e-14 || n2 < 1e-14 || bb < 1e-14) continue;
        double bl = sqrt(bb);
        double b1n = b1[0] * nv[0] + b1[1] * nv[1] + b1[2] * nv[2];
        double mn = mv[0] * nv[0] + mv[1] * nv[1] + mv[2] * nv[2];
        double phi = atan2(bl * b1n, mn);
        const double *p = prm + 3 * m;
        double delta = p[1] * phi - p[2];
        double f = p[0] * p[1] * sin(delta);
        double u = p[0] * (1.0 + cos(delta));
        double fi[3], fl[3], fj[3], fk[3];
        double s = (b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) / bb;
        double t = (b3[0] * b2[0] + b3[1] * b2[1] + b3[2] * b2[2]) / bb;
        for (int d = 0; d < 3; ++d) {
            fi[d] = f * bl / m2 * mv[d];
            fl[d] = -f * bl / n2 * nv[d];
            fj[d] = (s - 1.0) * fi[d] - t * fl[d];
            fk[d] = -s * fi[d] + (t - 1.0) * fl[d];
        }
        double w = 0.0;
        for (int d = 0; d < 3; ++d)
            w += b1[d] * fi[d] + b2[d] * fk[d] + (b2[d] - b3[d]) * fl[d];
        const double *fs[4] = {fi, fj, fk, fl};
        for (int q = 0; q < 4; ++q) {
            for (int d = 0; d < 3; ++d) F[3 * a[q] + d] += fs[q][d];
            E[a[q]] += 0.25 * u;
            W[a[q]] += 0.25 * w;
        }
    }
}

/* ------------------------------------------------------------------ */
/* SHAKE / RATTLE (kernels/shake.cu:96-490, src/integrator/shake.rs)     */
/* Groups: atoms goff[g]..goff[g+1] of gatoms, constraints coff[g]..     */
/* coff[g+1] with group-local atom indices ci/cj and r2 = (f32 d)^2.     */
/* ------------------------------------------------------------------ */

static void min_image_to_f(float ax, float ay, float az, float *bx, float *by,
                           float *bz, const float *L)
{
    /* shake.cu:27-41 */
    float dx = *bx - ax, dy = *by - ay, dz = *bz - az;
    min_image_f(&dx, &dy, &dz, L);
    *bx = ax + dx; *by = ay + dy; *bz = az + dz;
}

static void weighted_com_f(const float *x, const float *y, const float *z, const float *m,
                           int n, float *cx, float *cy, float *cz)
{
    /* shake.cu:70-91 */
    float total = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int a = 0; a < n; ++a) {
        sx += m[a] * x[a]; sy += m[a] * y[a]; sz += m[a] * z[a];
        total += m[a];
    }
    float inv = total > 0.0f ? 1.0f / total : 0.0f;
    *cx = sx * inv; *cy = sy * inv; *cz = sz * inv;
}

#define OR_SHAKE_MAXA 64
#define OR_SHAKE_REF_ATOMS 8  /* MAX_GROUP_ATOMS, src/integrator/shake.rs:21 */
#define OR_SHAKE_MAXC 64

/* Frame of a group beyond the 8-atom cap (no reference counterpart): the
 * bond vectors d_k = minimg(x_k - x_(k-1)) (d_0 = 0), an inclusive
 * Hillis-Steele scan over 64 lanes (offsets 1, 2, ..., 32; lane k adds lane
 * k - off) giving S_k, and coordinates S_k - S_mid relative to the middle
 * atom mid = na / 2.  The device computes the same sums with wave shuffles. */
static void or_chain_frame(int na, float *x, float *y, float *z, const float *L)
{
    float sx[64], sy[64], sz[64], tx[64], ty[64], tz[64];
    for (int k = 0; k < 64; ++k) {
        sx[k] = sy[k] = sz[k] = 0.0f;
        if (k >= 1 && k < na) {
            float dx = x[k] - x[k - 1], dy = y[k] - y[k - 1], dz = z[k] - z[k - 1];
            min_image_f(&dx, &dy, &dz, L);
            sx[k] = dx; sy[k] = dy; sz[k] = dz;
        }
    }
    for (int off = 1; off < 64; off <<= 1) {
        for (int k = 0; k < 64; ++k) {
            tx[k] = sx[k]; ty[k] = sy[k]; tz[k] = sz[k];
            if (k >= off) {
                tx[k] = sx[k - off] + sx[k];
                ty[k] = sy[k - off] + sy[k];
                tz[k] = sz[k - off] + sz[k];
            }
        }
        for (int k = 0; k < 64; ++k) { sx[k] = tx[k]; sy[k] = ty[k]; sz[k] = tz[k]; }
    }
    int mid = na / 2;
    for (int k = 0; k < na; ++k) {
        x[k] = sx[k] - sx[mid];
        y[k] = sy[k] - sy[mid];
        z[k] = sz[k] - sz[mid];
    }
}

/* Sum of v[0..n-1] as lane 0 of a 64-lane xor butterfly (offsets 32, 16,
 * ..., 1; lane k adds lane k ^ off) over zero-padded lanes. */
static float or_lane0_sum(const float *v, int n)
{
    float a[64], t[64];
    for (int k = 0; k < 64; ++k) a[k] = k < n ? v[k] : 0.0f;
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 64; ++k) t[k] = a[k] + a[k ^ off];
        for (int k = 0; k < 64; ++k) a[k] = t[k];
    }
    return a[0];
}

/* Mass-weighted centre of a long group in the device's butterfly order. */
static void or_long_com(const float *x, const float *y, const float *z, const float *m, int n,
                        float *cx, float *cy, float *cz)
{
    float px[64], py[64], pz[64];
    for (int a = 0; a < n; ++a) { px[a] = m[a] * x[a]; py[a] = m[a] * y[a]; pz[a] = m[a] * z[a]; }
    float total = or_lane0_sum(m, n);
    float inv = total > 0.0f ? 1.0f / total : 0.0f;
    *cx = or_lane0_sum(px, n) * inv;
    *cy = or_lane0_sum(py, n) * inv;
    *cz = or_lane0_sum(pz, n) * inv;
}

/* Form of the group solve.  0: the reference's for groups within its
 * 8-atom cap (one image relative to atom 0, topology-order sweeps,
 * sequential centre of mass) and the device's wave form beyond it.  1: the
 * device's wave form for every group (the device runs all groups, whatever
 * their size, one wave per group). */
static int or_shake_wave_all = 0;

void or_set_shake_form(int wave_all) { or_shake_wave_all = wave_all != 0; }

static int or_wave_form(int na) { return or_shake_wave_all || na > OR_SHAKE_REF_ATOMS; }

/* Sweep order of a group's constraints.  Reference form: topology order.  Longer groups: greedy colouring in topology
 * order (each constraint takes the lowest colour no earlier constraint on
 * either of its atoms holds), then colour by colour, topology order within
 * a colour; constraints of one colour share no atom, which lets the device
 * run each colour's updates at once. */
static void or_sweep_order(int na, int nc, const uint8_t *li, const uint8_t *lj, int *order)
{
    if (!or_wave_form(na)) {
        for (int k = 0; k < nc; ++k) order[k] = k;
        return;
    }
    int color[OR_SHAKE_MAXC];
    uint32_t used[OR_SHAKE_MAXA];
    for (int a = 0; a < na; ++a) used[a] = 0;
    int ncol = 0;
    for (int k = 0; k < nc; ++k) {
        uint32_t busy = used[li[k]] | used[lj[k]];
        int q = 0;
        while (busy & (1u << q)) ++q;
        color[k] = q;
        used[li[k]] |= 1u << q;
        used[lj[k]] |= 1u << q;
        if (q + 1 > ncol) ncol = q + 1;
    }
    int s = 0;
    for (int q = 0; q < ncol; ++q)
        for (int k = 0; k < nc; ++k)
            if (color[k] == q) order[s++] = k;
}

void or_shake_positions_f32(int n_groups, const uint32_t *goff, const uint32_t *gatoms,
                            const uint32_t *coff, const uint8_t *ci, const uint8_t *cj,
                            const float *cr2, float *posq, float *vx, float *vy, float *vz,
                            const float *snap, const float *mass, const float *L, float dt,
                            float tol, int max_iter, float *w_cons, int *iters,
                            const float *nu0, float *nu_out)
{
    /* shake_positions (shake.cu:96-330): Gauss-Seidel sweeps in topology
     * order with gradients from the pre-drift snapshot, in the image of
     * atom 0; delta write-back; position-level constraint virial.  Groups
     * beyond the reference's 8-atom cap: bond-by-bond image, coordinates
     * relative to the middle atom, sweeps in colour order (or_sweep_order) */
    for (int g = 0; g < n_groups; ++g) {
        int na = (int)(goff[g + 1] - goff[g]), nc = (int)(coff[g + 1] - coff[g]);
        const uint32_t *at = gatoms + goff[g];
        float x0[OR_SHAKE_MAXA], y0[OR_SHAKE_MAXA], z0[OR_SHAKE_MAXA];
        float xu[OR_SHAKE_MAXA], yu[OR_SHAKE_MAXA], zu[OR_SHAKE_MAXA];
        float xc[OR_SHAKE_MAXA], yc[OR_SHAKE_MAXA], zc[OR_SHAKE_MAXA];
        float m[OR_SHAKE_MAXA], im[OR_SHAKE_MAXA];
        for (int a = 0; a < na; ++a) {
            uint32_t i = at[a];
            xu[a] = posq[4 * i]; yu[a] = posq[4 * i + 1]; zu[a] = posq[4 * i + 2];
            x0[a] = snap[4 * i]; y0[a] = snap[4 * i + 1]; z0[a] = snap[4 * i + 2];
            m[a] = mass[i];
            im[a] = m[a] > 0.0f ? 1.0f / m[a] : 0.0f;
        }
        /* one image for the whole group: relative to atom 0
         * (shake.cu:213-219); longer groups in the chain frame, bond by bond
         * relative to the middle atom (a chain longer than half the box
         * would fold back onto atom 0's image, and |r|^2 - d^2 resolves to
         * the f32 spacing of the group's extent instead of the box's) */
        const int lg = or_wave_form(na);
        if (lg) {
            or_chain_frame(na, x0, y0, z0, L);
            or_chain_frame(na, xu, yu, zu, L);
        } else {
            for (int a = 1; a < na; ++a) {
                min_image_to_f(x0[0], y0[0], z0[0], &x0[a], &y0[a], &z0[a], L);
                min_image_to_f(xu[0], yu[0], zu[0], &xu[a], &yu[a], &zu[a], L);
            }
        }
        for (int a = 0; a < na; ++a) { xc[a] = xu[a]; yc[a] = yu[a]; zc[a] = zu[a]; }
        float gx[OR_SHAKE_MAXC], gy[OR_SHAKE_MAXC], gz[OR_SHAKE_MAXC], ip[OR_SHAKE_MAXC];
        const uint8_t *li = ci + coff[g], *lj = cj + coff[g];
        const float *r2 = cr2 + coff[g];
        float lam[OR_SHAKE_MAXC];
        for (int k = 0; k < nc; ++k) {
            gx[k] = x0[li[k]] - x0[lj[k]];
            gy[k] = y0[li[k]] - y0[lj[k]];
            gz[k] = z0[li[k]] - z0[lj[k]];
            ip[k] = im[li[k]] + im[lj[k]];
            lam[k] = 0.0f;
        }
        int order[OR_SHAKE_MAXC];
        or_sweep_order(na, nc, li, lj, order);
        if (nu0) {
            /* warm start (no reference counterpart): one move along every
             * constraint gradient by the given nu = lambda (1/m_i + 1/m_j),
             * in sweep order, before the sweeps */
            for (int s = 0; s < nc; ++s) {
                int k = order[s];
                int a = li[k], b = lj[k];
                float lambda = ip[k] > 0.0f ? nu0[coff[g] + k] / ip[k] : 0.0f;
                float ki = lambda * im[a], kj = lambda * im[b];
                xc[a] -= ki * gx[k]; yc[a] -= ki * gy[k]; zc[a] -= ki * gz[k];
                xc[b] += kj * gx[k]; yc[b] += kj * gy[k]; zc[b] += kj * gz[k];
                lam[k] = lambda;
            }
        }
        int it = 0;
        for (; it < max_iter; ++it) {
            int converged = 1;
            for (int s = 0; s < nc; ++s) {
                int k = order[s];
                int a = li[k], b = lj[k];
                float dx = xc[a] - xc[b], dy = yc[a] - yc[b], dz = zc[a] - zc[b];
                float dist2 = dx * dx + dy * dy + dz * dz;
                float sigma = dist2 - r2[k];
                if (fabsf(sigma) > tol) {
                    converged = 0;
                    float ddot = dx * gx[k] + dy * gy[k] + dz * gz[k];
                    float denom = 2.0f * ddot * ip[k];
                    if (denom == 0.0f) continue;
                    float lambda = sigma / denom;
                    float ki = lambda * im[a], kj = lambda * im[b];
                    xc[a] -= ki * gx[k]; yc[a] -= ki * gy[k]; zc[a] -= ki * gz[k];
                    xc[b] += kj * gx[k]; yc[b] += kj * gy[k]; zc[b] += kj * gz[k];
                    lam[k] += lambda;
                }
            }
            if (converged) break;
        }
        if (iters) iters[g] = it;
        if (nu_out)
            for (int k = 0; k < nc; ++k) nu_out[coff[g] + k] = lam[k] * ip[k];
        float cxg, cyg, czg;
        if (lg)
            or_long_com(xc, yc, zc, m, na, &cxg, &cyg, &czg);
        else
            weighted_com_f(xc, yc, zc, m, na, &cxg, &cyg, &czg);
        float inv_dt = dt != 0.0f ? 1.0f / dt : 0.0f;
        float inv_dt2 = dt != 0.0f ? 1.0f / (dt * dt) : 0.0f;
        for (int a = 0; a < na; ++a) {
            uint32_t i = at[a];
            float dxg = xc[a] - xu[a], dyg = yc[a] - yu[a], dzg = zc[a] - zu[a];
            vx[i] += dxg * inv_dt; vy[i] += dyg * inv_dt; vz[i] += dzg * inv_dt;
            /* long groups: every atom moves by its delta */
            if (a == 0 && !lg) {
                posq[4 * i] = xc[0]; posq[4 * i + 1] = yc[0]; posq[4 * i + 2] = zc[0];
            } else {
                posq[4 * i] += dxg; posq[4 * i + 1] += dyg; posq[4 * i + 2] += dzg;
            }
            float rx = xc[a] - cxg, ry = yc[a] - cyg, rz = zc[a] - czg;
            float scale = m[a] * inv_dt2;
            w_cons[i] = scale * (dxg * rx + dyg * ry + dzg * rz);
        }
    }
}

/* Predictor inputs per constraint (no reference counterpart), f32 in the
 * device's operation order: for constraint k with snapshot bond g and
 * unconstrained bond r (each minimum-imaged), nu1 = (|r|^2 - d^2)/(2 g.r);
 * coupling kappa_kl = c_kl (r_k.g_l) / (ip_l g_k.r_k) with c_kl the signed
 * inverse mass of the atom shared by k and l; rho = g.r/d^2 - 1; nuprev =
 * the previous solve's converged nu (nu_prev, or 0).
 * Row: nu1_k, nu1_k-1, nu1_k+1, nuprev_k, nu1_k-2 + nu1_k+2, kappa_k,k-1,
 * kappa_k,k+1, rho_k (zero past the group's ends). */
typedef struct {
    float gx, gy, gz, rx, ry, rz, imi, imj, ip, gr, nu1;
    int li, lj;
} or_bond;

static or_bond or_shake_bond(const uint32_t *at, const uint8_t *ci, const uint8_t *cj,
                             const float *cr2, const float *posq, const float *snap,
                             const float *mass, const float *L, int c)
{
    or_bond b;
    b.li = ci[c];
    b.lj = cj[c];
    uint32_t i = at[b.li], j = at[b.lj];
    b.gx = snap[4 * i] - snap[4 * j];
    b.gy = snap[4 * i + 1] - snap[4 * j + 1];
    b.gz = snap[4 * i + 2] - snap[4 * j + 2];
    min_image_f(&b.gx, &b.gy, &b.gz, L);
    b.rx = posq[4 * i] - posq[4 * j];
    b.ry = posq[4 * i + 1] - posq[4 * j + 1];
    b.rz = posq[4 * i + 2] - posq[4 * j + 2];
    min_image_f(&b.rx, &b.ry, &b.rz, L);
    float mi = mass[i], mj = mass[j];
    b.imi = mi > 0.0f ? 1.0f / mi : 0.0f;
    b.imj = mj > 0.0f ? 1.0f / mj : 0.0f;
    b.ip = b.imi + b.imj;
    b.gr = b.gx * b.rx + b.gy * b.ry + b.gz * b.rz;
    float sigma = (b.rx * b.rx + b.ry * b.ry + b.rz * b.rz) - cr2[c];
    b.nu1 = b.gr != 0.0f ? sigma / (2.0f * b.gr) : 0.0f;
    return b;
}

static float or_coupling(const or_bond *k, const or_bond *l)
{
    float c = 0.0f;
    if (k->li == l->li) c += l->imi;
    if (k->li == l->lj) c -= l->imj;
    if (k->lj == l->li) c -= l->imi;
    if (k->lj == l->lj) c += l->imj;
    float den = l->ip * k->gr;
    if (c == 0.0f || den == 0.0f) return 0.0f;
    float num = c * (k->rx * l->gx + k->ry * l->gy + k->rz * l->gz);
    return num / den;
}

void or_shake_features_f32(int n_groups, const uint32_t *goff, const uint32_t *gatoms,
                           const uint32_t *coff, const uint8_t *ci, const uint8_t *cj,
                           const float *cr2, const float *posq, const float *snap,
                           const float *mass, const float *L, const float *nu_prev, float *feat)
{
    for (int g = 0; g < n_groups; ++g) {
        const uint32_t *at = gatoms + goff[g];
        int cb = (int)coff[g], ce = (int)coff[g + 1];
        for (int c = cb; c < ce; ++c) {
            float *f = feat + 8 * (size_t)c;
            or_bond k = or_shake_bond(at, ci, cj, cr2, posq, snap, mass, L, c);
            for (int q = 0; q < 8; ++q) f[q] = 0.0f;
            f[0] = k.nu1;
            if (c >= cb + 1) {
                or_bond l = or_shake_bond(at, ci, cj, cr2, posq, snap, mass, L, c - 1);
                f[1] = l.nu1;
                f[5] = or_coupling(&k, &l);
            }
            if (c + 1 < ce) {
                or_bond l = or_shake_bond(at, ci, cj, cr2, posq, snap, mass, L, c + 1);
                f[2] = l.nu1;
                f[6] = or_coupling(&k, &l);
            }
            f[3] = nu_prev ? nu_prev[c] : 0.0f;
            float f2l = c >= cb + 2 ? or_shake_bond(at, ci, cj, cr2, posq, snap, mass, L, c - 2).nu1
                                    : 0.0f;
            float f2r = c + 2 < ce ? or_shake_bond(at, ci, cj, cr2, posq, snap, mass, L, c + 2).nu1
                                   : 0.0f;
            f[4] = f2l + f2r;
            f[7] = k.gr / cr2[c] - 1.0f;
        }
    }
}

void or_rattle_velocities_f32(int n_groups, const uint32_t *goff, const uint32_t *gatoms,
                              const uint32_t *coff, const uint8_t *ci, const uint8_t *cj,
                              const float *posq, float *vx, float *vy, float *vz,
                              const float *mass, const float *L, float dt, float tol,
                              int max_iter, float *w_cons, int *iters)
{
    /* rattle_velocities (shake.cu:336-490): Gauss-Seidel on the relative
     * velocity along each bond; with dt > 0 the velocity-level virial is
     * added to the position-level one */
    for (int g = 0; g < n_groups; ++g) {
        int na = (int)(goff[g + 1] - goff[g]), nc = (int)(coff[g + 1] - coff[g]);
        const uint32_t *at = gatoms + goff[g];
        float px[OR_SHAKE_MAXA], py[OR_SHAKE_MAXA], pz[OR_SHAKE_MAXA];
        float v0[OR_SHAKE_MAXA], v1[OR_SHAKE_MAXA], v2[OR_SHAKE_MAXA];
        float dv0[OR_SHAKE_MAXA], dv1[OR_SHAKE_MAXA], dv2[OR_SHAKE_MAXA];
        float m[OR_SHAKE_MAXA], im[OR_SHAKE_MAXA];
        for (int a = 0; a < na; ++a) {
            uint32_t i = at[a];
            px[a] = posq[4 * i]; py[a] = posq[4 * i + 1]; pz[a] = posq[4 * i + 2];
            v0[a] = vx[i]; v1[a] = vy[i]; v2[a] = vz[i];
            dv0[a] = dv1[a] = dv2[a] = 0.0f;
            m[a] = mass[i];
            im[a] = m[a] > 0.0f ? 1.0f / m[a] : 0.0f;
        }
        const int lg = or_wave_form(na);
        if (lg)
            or_chain_frame(na, px, py, pz, L);
        else
            for (int a = 1; a < na; ++a)
                min_image_to_f(px[0], py[0], pz[0], &px[a], &py[a], &pz[a], L);
        const uint8_t *li = ci + coff[g], *lj = cj + coff[g];
        float dx[OR_SHAKE_MAXC], dy[OR_SHAKE_MAXC], dz[OR_SHAKE_MAXC], d2[OR_SHAKE_MAXC],
            ip[OR_SHAKE_MAXC];
        for (int k = 0; k < nc; ++k) {
            dx[k] = px[li[k]] - px[lj[k]];
            dy[k] = py[li[k]] - py[lj[k]];
            dz[k] = pz[li[k]] - pz[lj[k]];
            d2[k] = dx[k] * dx[k] + dy[k] * dy[k] + dz[k] * dz[k];
            ip[k] = im[li[k]] + im[lj[k]];
        }
        int order[OR_SHAKE_MAXC];
        or_sweep_order(na, nc, li, lj, order);
        int it = 0;
        for (; it < max_iter; ++it) {
            int converged = 1;
            for (int s = 0; s < nc; ++s) {
                int k = order[s];
                int a = li[k], b = lj[k];
                float vxr = v0[a] - v0[b], vyr = v1[a] - v1[b], vzr = v2[a] - v2[b];
                float vrel = vxr * dx[k] + vyr * dy[k] + vzr * dz[k];
                if (fabsf(vrel) > tol) {
                    converged = 0;
                    float denom = d2[k] * ip[k];
                    if (denom == 0.0f) continue;
                    float mu = vrel / denom;
                    float ki = mu * im[a], kj = mu * im[b];
                    v0[a] -= ki * dx[k]; v1[a] -= ki * dy[k]; v2[a] -= ki * dz[k];
                    v0[b] += kj * dx[k]; v1[b] += kj * dy[k]; v2[b] += kj * dz[k];
                    dv0[a] -= ki * dx[k]; dv1[a] -= ki * dy[k]; dv2[a] -= ki * dz[k];
                    dv0[b] += kj * dx[k]; dv1[b] += kj * dy[k]; dv2[b] += kj * dz[k];
                }
            }
            if (converged) break;
        }
        if (iters) iters[g] = it;
        if (dt > 0.0f) {
            float inv_dt = 1.0f / dt;
            float cx, cy, cz;
            if (lg)
                or_long_com(px, py, pz, m, na, &cx, &cy, &cz);
            else
                weighted_com_f(px, py, pz, m, na, &cx, &cy, &cz);
            for (int a = 0; a < na; ++a) {
                float rx = px[a] - cx, ry = py[a] - cy, rz = pz[a] - cz;
                float scale = m[a] * inv_dt;
                w_cons[at[a]] += scale * (dv0[a] * rx + dv1[a] * ry + dv2[a] * rz);
            }
        }
        for (int a = 0; a < na; ++a) {
            uint32_t i = at[a];
            vx[i] = v0[a]; vy[i] = v1[a]; vz[i] = v2[a];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Philox-4x32-10 and the Langevin BAOAB sub-steps                     */
/* (kernels/philox.cuh:20-172, kernels/langevin.cu:13-80,              */
/*  src/integrator/langevin_baoab.rs plan: B, A, O, A, F, B)            */
/* ------------------------------------------------------------------ */

void or_philox4x32_10(uint32_t k0, uint32_t k1, const uint32_t *ctr, uint32_t *out)
{
    /* Salmon et al. SC11: ten rounds, multipliers 0xD2511F53/0xCD9E8D57,
     * Weyl key increments 0x9E3779B9/0xBB67AE85 (philox.cuh:20-60) */
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)c0 * 0xD2511F53u;
        uint64_t p2 = (uint64_t)c2 * 0xCD9E8D57u;
        uint32_t n0 = (uint32_t)(p2 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p2;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double or_philox_gaussian(uint32_t seed_lo, uint32_t seed_hi, uint32_t ctr_lo,
                          uint32_t ctr_hi, uint32_t pid, uint32_t axis)
{
    /* philox.cuh philox_gaussian: Box-Muller on the first two lanes, each
     * mapped to the open interval by (x + 0.5) * 2^-32, in double */
    uint32_t ctr[4] = {ctr_lo, ctr_hi, pid, axis}, o[4];
    or_philox4x32_10(seed_lo, seed_hi, ctr, o);
    double u1 = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
    double u2 = ((double)o[1] + 0.5) * (1.0 / 4294967296.0);
    double r = sqrt(-2.0 * log(u1));
    return r * cos(6.283185307179586 * u2);
}

/* CSVR velocity rescale (Bussi, Donadio, Parrinello 2007, eq. A7;        */
/* src/integrator/csvr.rs:188-240, kernels/nose_hoover.cu:145-255):        */
/* K_new = c K + (K_t/g)(1-c)(sum_{i=1}^{g-1} xi_i^2 + r^2)                 */
/*       + 2 r sqrt(c (1-c) K K_t/g), xi_i = normal(seed, counter, i, 0),   */
/* r = normal(seed, counter, 0, 0); a non-finite or non-positive K_new     */
/* keeps K.  Returns sqrt(K_new/K) (1 when K = 0 or K_new = K).            */
double or_csvr_factor(double k, uint32_t g_dof, uint64_t seed, uint64_t counter, double c,
                      double kt_per_dof, double *k_new)
{
    uint32_t slo = (uint32_t)(seed & 0xFFFFFFFFu), shi = (uint32_t)(seed >> 32);
    uint32_t clo = (uint32_t)(counter & 0xFFFFFFFFu), chi = (uint32_t)(counter >> 32);
    double s = 0.0;
    for (uint32_t i = 1; i < g_dof; ++i) {
        double xi = or_philox_gaussian(slo, shi, clo, chi, i, 0u);
        s += xi * xi;
    }
    double r = or_philox_gaussian(slo, shi, clo, chi, 0u, 0u);
    double cross = k > 0.0 ? 2.0 * r * sqrt(c * (1.0 - c) * k * kt_per_dof) : 0.0;
    double kn = c * k + kt_per_dof * (1.0 - c) * (s + r * r) + cross;
    if (!isfinite(kn) || kn <= 0.0) kn = k;
    *k_new = kn;
    return (k > 0.0 && kn != k) ? sqrt(kn / k) : 1.0;
}

/* Berendsen thermostat factor (src/integrator/berendsen.rs:139-170,       */
/* nose_hoover.cu:390-417): sqrt(max(0, 1 + dt/tau (K_t/K - 1))); the      */
/* injected energy K (lambda^2 - 1) goes to *inj.                          */
double or_berendsen_factor(double k, double k_target, double dt_over_tau, double *inj)
{
    *inj = 0.0;
    if (!(k > 0.0)) return 1.0;
    double l2 = 1.0 + dt_over_tau * (k_target / k - 1.0);
    if (!(l2 > 0.0)) l2 = 0.0;
    *inj = k * (l2 - 1.0);
    return sqrt(l2);
}

/* Andersen collisions (kernels/andersen.cu:13-60,                        */
/* src/integrator/andersen.rs:276-300): u = ((x0 + 0.5) 2^-32) of the      */
/* Philox block (seed, counter, i, 3); when u < p the velocity becomes      */
/* sqrt(kT/m) (f32) times the normals of (seed, counter, i, 0..2).          */
void or_andersen(int n, float *vx, float *vy, float *vz, const float *mass, uint64_t seed,
                 uint64_t counter, float p, float kt)
{
    uint32_t slo = (uint32_t)(seed & 0xFFFFFFFFu), shi = (uint32_t)(seed >> 32);
    uint32_t clo = (uint32_t)(counter & 0xFFFFFFFFu), chi = (uint32_t)(counter >> 32);
    for (int i = 0; i < n; ++i) {
        uint32_t ctr[4] = {clo, chi, (uint32_t)i, 3u}, o[4];
        or_philox4x32_10(slo, shi, ctr, o);
        double u = ((double)o[0] + 0.5) * (1.0 / 4294967296.0);
        if (u >= (double)p) continue;
        float sigma = sqrtf(kt / mass[i]);
        vx[i] = sigma * (float)or_philox_gaussian(slo, shi, clo, chi, (uint32_t)i, 0u);
        vy[i] = sigma * (float)or_philox_gaussian(slo, shi, clo, chi, (uint32_t)i, 1u);
        vz[i] = sigma * (float)or_philox_gaussian(slo, shi, clo, chi, (uint32_t)i, 2u);
    }
}

float or_lan_alpha(float friction, float dt)
{
    /* alpha = exp(-gamma dt), f32 on the host (langevin_baoab.rs, "O") */
    return expf(-friction * dt);
}

void or_lan_kick_drift_half_f32(int n, float *posq, int *img, float *vx, float *vy,
                                float *vz, const float *fx, const float *fy,
                                const float *fz, const float *mass, const float *L,
                                float dt)
{
    /* B: vv_kick (integrate.cu) with the current forces, then A:
     * lan_drift_half (langevin.cu:13-50) x += v dt/2 and wrap */
    for (int i = 0; i < n; ++i) {
        float m = mass[i];
        float ax = fx[i] / m, ay = fy[i] / m, az = fz[i] / m;
        float half_dt = dt * 0.5f;
        float v0 = vx[i] + ax * half_dt;
        float v1 = vy[i] + ay * half_dt;
        float v2 = vz[i] + az * half_dt;
        vx[i] = v0; vy[i] = v1; vz[i] = v2;
        float px = posq[4 * i] + v0 * half_dt;
        float py = posq[4 * i + 1] + v1 * half_dt;
        float pz = posq[4 * i + 2] + v2 * half_dt;
        int ka, kb, kc;
        wrap_with_image_f(&px, &py, &pz, &ka, &kb, &kc, L);
        posq[4 * i] = px; posq[4 * i + 1] = py; posq[4 * i + 2] = pz;
        img[3 * i] += ka; img[3 * i + 1] += kb; img[3 * i + 2] += kc;
    }
}

void or_lan_ou_drift_half_f32(int n, float *posq, int *img, float *vx, float *vy,
                              float *vz, const float *mass, const float *L, float dt,
                              float alpha, float kt, uint64_t seed, uint64_t counter)
{
    /* O: lan_ou_step (langevin.cu:54-80) v = alpha v + sigma xi with
     * sigma = sqrt((1 - alpha^2) kT / m) in f32 and xi the double
     * Box-Muller draw rounded to f32; then A: x += v dt/2 and wrap */
    uint32_t slo = (uint32_t)(seed & 0xFFFFFFFFu), shi = (uint32_t)(seed >> 32);
    uint32_t clo = (uint32_t)(counter & 0xFFFFFFFFu), chi = (uint32_t)(counter >> 32);
    float half_dt = dt * 0.5f;
    for (int i = 0; i < n; ++i) {
        float m = mass[i];
        float sigma = sqrtf((1.0f - alpha * alpha) * kt / m);
        float xi0 = (float)or_philox_gaussian(slo, shi, clo, chi, (uint32_t)i, 0u);
        float xi1 = (float)or_philox_gaussian(slo, shi, clo, chi, (uint32_t)i, 1u);
        float xi2 = (float)or_philox_gaussian(slo, shi, clo, chi, (uint32_t)i, 2u);
        float v0 = alpha * vx[i] + sigma * xi0;
        float v1 = alpha * vy[i] + sigma * xi1;
        float v2 = alpha * vz[i] + sigma * xi2;
        vx[i] = v0; vy[i] = v1; vz[i] = v2;
        float px = posq[4 * i] + v0 * half_dt;
        float py = posq[4 * i + 1] + v1 * half_dt;
        float pz = posq[4 * i + 2] + v2 * half_dt;
        int ka, kb, kc;
        wrap_with_image_f(&px, &py, &pz, &ka, &kb, &kc, L);
        posq[4 * i] = px; posq[4 * i + 1] = py; posq[4 * i + 2] = pz;
        img[3 * i] += ka; img[3 * i + 1] += kb; img[3 * i + 2] += kc;
    }
}

static inline void comp_add(float *hi, double *lo, double delta)
{
    double ext = (double)*hi + *lo + delta;
    float nv = (float)ext;
    *lo = ext - (double)nv;
    *hi = nv;
}

void or_vv_kick_drift_lossless(int n, float *posq, int *img, float *vx,
                               float *vy, float *vz, double *xlo, double *ylo,
                               double *zlo, double *vxlo, double *vylo,
                               double *vzlo, const float *fx, const float *fy,
                               const float *fz, const float *mass,
                               const float *L, float dt)
{
    for (int i = 0; i < n; ++i) {
        float m = mass[i];
        float ax = fx[i] / m, ay = fy[i] / m, az = fz[i] / m;
        float half_dt = dt * 0.5f;
        comp_add(&vx[i], &vxlo[i], (double)(ax * half_dt));
        comp_add(&vy[i], &vylo[i], (double)(ay * half_dt));
        comp_add(&vz[i], &vzlo[i], (double)(az * half_dt));
        double dx = ((double)vx[i] + vxlo[i]) * (double)dt;
        double dy = ((double)vy[i] + vylo[i]) * (double)dt;
        double dz = ((double)vz[i] + vzlo[i]) * (double)dt;
        float px = posq[4 * i], py = posq[4 * i + 1], pz = posq[4 * i + 2];
        comp_add(&px, &xlo[i], dx);
        comp_add(&py, &ylo[i], dy);
        comp_add(&pz, &zlo[i], dz);
        int ka, kb, kc;
        wrap_with_image_f(&px, &py, &pz, &ka, &kb, &kc, L);
        posq[4 * i] = px; posq[4 * i + 1] = py; posq[4 * i + 2] = pz;
        img[3 * i] += ka; img[3 * i + 1] += kb; img[3 * i + 2] += kc;
    }
}

void or_vv_kick_lossless(int n, float *vx, float *vy, float *vz, double *vxlo,
                         double *vylo, double *vzlo, const float *fx,
                         const float *fy, const float *fz, const float *mass,
                         float dt)
{
    for (int i = 0; i < n; ++i) {
        float m = mass[i];
        float ax = fx[i] / m, ay = fy[i] / m, az = fz[i] / m;
        float half_dt = dt * 0.5f;
        comp_add(&vx[i], &vxlo[i], (double)(ax * half_dt));
        comp_add(&vy[i], &vylo[i], (double)(ay * half_dt));
        comp_add(&vz[i], &vzlo[i], (double)(az * half_dt));
    }
}

/* ------------------------------------------------------------------ */
/* SETTLE (kernels/settle.cu:145-497), in double.                       */
/* Groups are given as an (n_groups x 3) array of atom indices in       */
/* canonical order O, H1, H2.                                            */
/* ------------------------------------------------------------------ */

static inline double csqrt0(double x) { return sqrt(x > 0.0 ? x : 0.0); }

/* Position reset.  pos_d: current (post-drift) positions as doubles
 * [3n]; snap: pre-drift positions [3n]; vel [3n] updated by corr/dt;
 * virial [3*n_groups] position-level constraint virial per atom slot. */
void or_settle_positions(int n_groups, const uint32_t *groups, double *pos,
                         double *vel, const double *snap, double ra, double rb,
                         double rc, double m_o, double m_h, const double *L,
                         double dt, double *virial)
{
    for (int g = 0; g < n_groups; ++g) {
        const uint32_t *idx = groups + 3 * g;
        double cur[3][3], snp[3][3];
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) {
                cur[a][d] = pos[3 * idx[a] + d];
                snp[a][d] = snap[3 * idx[a] + d];
            }
        double xb0 = snp[1][0] - snp[0][0], yb0 = snp[1][1] - snp[0][1], zb0 = snp[1][2] - snp[0][2];
        double xc0 = snp[2][0] - snp[0][0], yc0 = snp[2][1] - snp[0][1], zc0 = snp[2][2] - snp[0][2];
        or_min_image(&xb0, &yb0, &zb0, L);
        or_min_image(&xc0, &yc0, &zc0, L);
        double xp[3][3];
        for (int a = 0; a < 3; ++a) {
            double dx = cur[a][0] - snp[a][0], dy = cur[a][1] - snp[a][1], dz = cur[a][2] - snp[a][2];
            or_min_image(&dx, &dy, &dz, L);
            xp[a][0] = dx; xp[a][1] = dy; xp[a][2] = dz;
        }
        double m0 = m_o, m1 = m_h, m2 = m_h;
        double itm = 1.0 / (m0 + m1 + m2);
        double d2 = 2.0 * rc, d2sq = d2 * d2;
        double xcom = (xp[0][0] * m0 + (xb0 + xp[1][0]) * m1 + (xc0 + xp[2][0]) * m2) * itm;
        double ycom = (xp[0][1] * m0 + (yb0 + xp[1][1]) * m1 + (yc0 + xp[2][1]) * m2) * itm;
        double zcom = (xp[0][2] * m0 + (zb0 + xp[1][2]) * m1 + (zc0 + xp[2][2]) * m2) * itm;
        double xa1 = xp[0][0] - xcom, ya1 = xp[0][1] - ycom, za1 = xp[0][2] - zcom;
        double xb1 = xb0 + xp[1][0] - xcom, yb1 = yb0 + xp[1][1] - ycom, zb1 = zb0 + xp[1][2] - zcom;
        double xc1 = xc0 + xp[2][0] - xcom, yc1 = yc0 + xp[2][1] - ycom, zc1 = zc0 + xp[2][2] - zcom;
        double ez_x = yb0 * zc0 - zb0 * yc0, ez_y = zb0 * xc0 - xb0 * zc0, ez_z = xb0 * yc0 - yb0 * xc0;
        double ex_x = ya1 * ez_z - za1 * ez_y, ex_y = za1 * ez_x - xa1 * ez_z, ex_z = xa1 * ez_y - ya1 * ez_x;
        double ey_x = ez_y * ex_z - ez_z * ex_y, ey_y = ez_z * ex_x - ez_x * ex_z, ey_z = ez_x * ex_y - ez_y * ex_x;
        double exl = csqrt0(ex_x * ex_x + ex_y * ex_y + ex_z * ex_z);
        double eyl = csqrt0(ey_x * ey_x + ey_y * ey_y + ey_z * ey_z);
        double ezl = csqrt0(ez_x * ez_x + ez_y * ez_y + ez_z * ez_z);
        double iex = exl > 0 ? 1.0 / exl : 0.0, iey = eyl > 0 ? 1.0 / eyl : 0.0, iez = ezl > 0 ? 1.0 / ezl : 0.0;
        double r11 = ex_x * iex, r21 = ex_y * iex, r31 = ex_z * iex;
        double r12 = ey_x * iey, r22 = ey_y * iey, r32 = ey_z * iey;
        double r13 = ez_x * iez, r23 = ez_y * iez, r33 = ez_z * iez;
        double b0x = r11 * xb0 + r21 * yb0 + r31 * zb0, b0y = r12 * xb0 + r22 * yb0 + r32 * zb0;
        double c0x = r11 * xc0 + r21 * yc0 + r31 * zc0, c0y = r12 * xc0 + r22 * yc0 + r32 * zc0;
        double a1z = r13 * xa1 + r23 * ya1 + r33 * za1;
        double b1x = r11 * xb1 + r21 * yb1 + r31 * zb1, b1y = r12 * xb1 + r22 * yb1 + r32 * zb1;
        double b1z = r13 * xb1 + r23 * yb1 + r33 * zb1;
        double c1x = r11 * xc1 + r21 * yc1 + r31 * zc1, c1y = r12 * xc1 + r22 * yc1 + r32 * zc1;
        double c1z = r13 * xc1 + r23 * yc1 + r33 * zc1;
        double sinphi = a1z / ra;
        double cosphi = csqrt0(1.0 - sinphi * sinphi);
        double sinpsi = (b1z - c1z) / (2.0 * rc * cosphi);
        double cospsi = csqrt0(1.0 - sinpsi * sinpsi);
        double a2y = ra * cosphi;
        double b2x = -rc * cospsi;
        double b2y = -rb * cosphi - rc * sinpsi * sinphi;
        double c2y = -rb * cosphi + rc * sinpsi * sinphi;
        double b2xsq = b2x * b2x;
        double hhsq = 4.0 * b2xsq + (b2y - c2y) * (b2y - c2y) + (b1z - c1z) * (b1z - c1z);
        double delta_x = 2.0 * b2x + csqrt0(4.0 * b2xsq - hhsq + d2sq);
        b2x -= delta_x * 0.5;
        double al = b2x * (b0x - c0x) + b0y * b2y + c0y * c2y;
        double be = b2x * (c0y - b0y) + b0x * b2y + c0x * c2y;
        double ga = b0x * b1y - b1x * b0y + c0x * c1y - c1x * c0y;
        double ab2 = al * al + be * be;
        double iab2 = ab2 > 0 ? 1.0 / ab2 : 0.0;
        double sinth = (al * ga - be * csqrt0(ab2 - ga * ga)) * iab2;
        double costh = csqrt0(1.0 - sinth * sinth);
        double a3x = -a2y * sinth, a3y = a2y * costh, a3z = a1z;
        double b3x = b2x * costh - b2y * sinth, b3y = b2x * sinth + b2y * costh, b3z = b1z;
        double c3x = -b2x * costh - c2y * sinth, c3y = -b2x * sinth + c2y * costh, c3z = c1z;
        double rcom[3][3] = {
            {r11 * a3x + r12 * a3y + r13 * a3z, r21 * a3x + r22 * a3y + r23 * a3z, r31 * a3x + r32 * a3y + r33 * a3z},
            {r11 * b3x + r12 * b3y + r13 * b3z, r21 * b3x + r22 * b3y + r23 * b3z, r31 * b3x + r32 * b3y + r33 * b3z},
            {r11 * c3x + r12 * c3y + r13 * c3z, r21 * c3x + r22 * c3y + r23 * c3z, r31 * c3x + r32 * c3y + r33 * c3z}};
        double com[3] = {xcom, ycom, zcom};
        double unc[3][3] = {{xp[0][0], xp[0][1], xp[0][2]},
                            {xb0 + xp[1][0], yb0 + xp[1][1], zb0 + xp[1][2]},
                            {xc0 + xp[2][0], yc0 + xp[2][1], zc0 + xp[2][2]}};
        double mat[3] = {m0, m1, m2};
        double idt = dt != 0 ? 1.0 / dt : 0.0, idt2 = dt != 0 ? 1.0 / (dt * dt) : 0.0;
        for (int a = 0; a < 3; ++a) {
            double corr[3];
            for (int d = 0; d < 3; ++d) corr[d] = (com[d] + rcom[a][d]) - unc[a][d];
            for (int d = 0; d < 3; ++d) {
                vel[3 * idx[a] + d] += corr[d] * idt;
                pos[3 * idx[a] + d] = cur[a][d] + corr[d];
            }
            virial[3 * g + a] = mat[a] * idt2 *
                (corr[0] * rcom[a][0] + corr[1] * rcom[a][1] + corr[2] * rcom[a][2]);
        }
    }
}

/* Velocity reset by the analytic 3x3 solve (kernels/settle.cu:376-494).
 * virial[3*n_groups] is accumulated (+=) when dt > 0. */
void or_settle_velocities(int n_groups, const uint32_t *groups,
                          const double *pos, double *vel, double m_o,
                          double m_h, const double *L, double dt,
                          double *virial)
{
    for (int g = 0; g < n_groups; ++g) {
        const uint32_t *idx = groups + 3 * g;
        double px[3], py[3], pz[3];
        for (int a = 0; a < 3; ++a) {
            px[a] = pos[3 * idx[a]];
            py[a] = pos[3 * idx[a] + 1];
            pz[a] = pos[3 * idx[a] + 2];
        }
        for (int a = 1; a < 3; ++a) {
            double dx = px[a] - px[0], dy = py[a] - py[0], dz = pz[a] - pz[0];
            or_min_image(&dx, &dy, &dz, L);
            px[a] = px[0] + dx; py[a] = py[0] + dy; pz[a] = pz[0] + dz;
        }
        double io = m_o > 0 ? 1.0 / m_o : 0.0, ih = m_h > 0 ? 1.0 / m_h : 0.0;
        double vx[3], vy[3], vz[3];
        for (int a = 0; a < 3; ++a) {
            vx[a] = vel[3 * idx[a]]; vy[a] = vel[3 * idx[a] + 1]; vz[a] = vel[3 * idx[a] + 2];
        }
        double r1x = px[0] - px[1], r1y = py[0] - py[1], r1z = pz[0] - pz[1];
        double r2x = px[0] - px[2], r2y = py[0] - py[2], r2z = pz[0] - pz[2];
        double r3x = px[1] - px[2], r3y = py[1] - py[2], r3z = pz[1] - pz[2];
        double d11 = r1x * r1x + r1y * r1y + r1z * r1z;
        double d22 = r2x * r2x + r2y * r2y + r2z * r2z;
        double d33 = r3x * r3x + r3y * r3y + r3z * r3z;
        double d12 = r1x * r2x + r1y * r2y + r1z * r2z;
        double d13 = r1x * r3x + r1y * r3y + r1z * r3z;
        double d23 = r2x * r3x + r2y * r3y + r2z * r3z;
        double b1 = (vx[0] - vx[1]) * r1x + (vy[0] - vy[1]) * r1y + (vz[0] - vz[1]) * r1z;
        double b2 = (vx[0] - vx[2]) * r2x + (vy[0] - vy[2]) * r2y + (vz[0] - vz[2]) * r2z;
        double b3 = (vx[1] - vx[2]) * r3x + (vy[1] - vy[2]) * r3y + (vz[1] - vz[2]) * r3z;
        double ioh = io + ih;
        double m[3][3] = {{d11 * ioh, d12 * io, -d13 * ih},
                          {d12 * io, d22 * ioh, d23 * ih},
                          {-d13 * ih, d23 * ih, 2.0 * d33 * ih}};
        double rhs[3] = {-b1, -b2, -b3};
        double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                     m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        double gg[3] = {0, 0, 0};
        if (det != 0.0) {
            for (int c = 0; c < 3; ++c) {
                double a[3][3];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) a[i][j] = (j == c) ? rhs[i] : m[i][j];
                double dc = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                            a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                            a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
                gg[c] = dc / det;
            }
        }
        double g1 = gg[0], g2 = gg[1], g3 = gg[2];
        double dv[3][3] = {
            {(g1 * r1x + g2 * r2x) * io, (g1 * r1y + g2 * r2y) * io, (g1 * r1z + g2 * r2z) * io},
            {(-g1 * r1x + g3 * r3x) * ih, (-g1 * r1y + g3 * r3y) * ih, (-g1 * r1z + g3 * r3z) * ih},
            {(-g2 * r2x - g3 * r3x) * ih, (-g2 * r2y - g3 * r3y) * ih, (-g2 * r2z - g3 * r3z) * ih}};
        vel[3 * idx[0]] = vx[0] + dv[0][0]; vel[3 * idx[0] + 1] = vy[0] + dv[0][1]; vel[3 * idx[0] + 2] = vz[0] + dv[0][2];
        vel[3 * idx[1]] = vx[1] + dv[1][0]; vel[3 * idx[1] + 1] = vy[1] + dv[1][1]; vel[3 * idx[1] + 2] = vz[1] + dv[1][2];
        vel[3 * idx[2]] = vx[2] + dv[2][0]; vel[3 * idx[2] + 1] = vy[2] + dv[2][1]; vel[3 * idx[2] + 2] = vz[2] + dv[2][2];
        if (dt > 0.0) {
            double tot = m_o + 2.0 * m_h, iM = tot > 0 ? 1.0 / tot : 0.0;
            double cx = (m_o * px[0] + m_h * px[1] + m_h * px[2]) * iM;
            double cy = (m_o * py[0] + m_h * py[1] + m_h * py[2]) * iM;
            double cz = (m_o * pz[0] + m_h * pz[1] + m_h * pz[2]) * iM;
            double ma[3] = {m_o, m_h, m_h};
            for (int a = 0; a < 3; ++a)
                virial[3 * g + a] += ma[a] / dt *
                    (dv[a][0] * (px[a] - cx) + dv[a][1] * (py[a] - cy) + dv[a][2] * (pz[a] - cz));
        }
    }
}

/* ------------------------------------------------------------------ */
/* Nose-Hoover chain (src/integrator/nose_hoover_chain.rs:109-279)      */
/* ------------------------------------------------------------------ */

double or_nhc_sub_step(int m, double *xi, double *pxi, const double *q,
                       double dt, double k_th, double g_dof, double kt)
{
    if (m == 0) return 1.0;
    double k = k_th;
    for (int j = m - 1; j >= 0; --j) {
        double s = (j == m - 1) ? 1.0 : exp(-dt / 8.0 * pxi[j + 1] / q[j + 1]);
        pxi[j] *= s;
        double gj = (j == 0) ? k - g_dof * kt : pxi[j - 1] * pxi[j - 1] / q[j - 1] - kt;
        pxi[j] += dt / 4.0 * gj;
        pxi[j] *= s;
    }
    double factor = exp(-dt / 2.0 * pxi[0] / q[0]);
    k *= factor * factor;
    for (int j = 0; j < m; ++j) xi[j] += dt / 2.0 * pxi[j] / q[j];
    for (int j = 0; j < m; ++j) {
        double s = (j == m - 1) ? 1.0 : exp(-dt / 8.0 * pxi[j + 1] / q[j + 1]);
        pxi[j] *= s;
        double gj = (j == 0) ? k - g_dof * kt : pxi[j - 1] * pxi[j - 1] / q[j - 1] - kt;
        pxi[j] += dt / 4.0 * gj;
        pxi[j] *= s;
    }
    return factor;
}

/* Yoshida x n_resp cumulative half step; returns the cumulative factor and
 * writes the updated kinetic energy to *k_out. */
double or_nhc_half_step(int m, double *xi, double *pxi, const double *q,
                        int n_yoshida, const double *w, int n_resp, double dt,
                        double k, double g_dof, double kt, double *k_out)
{
    double cum = 1.0;
    for (int y = 0; y < n_yoshida; ++y)
        for (int r = 0; r < n_resp; ++r) {
            double delta = w[y] * dt / (2.0 * n_resp);
            double f = or_nhc_sub_step(m, xi, pxi, q, delta, 2.0 * k, g_dof, kt);
            cum *= f;
            k *= f * f;
        }
    *k_out = k;
    return cum;
}

/* Kinetic energy sum_i 1/2 m v^2 in double (f32 inputs). */
double or_kinetic_energy(int n, const float *vx, const float *vy,
                         const float *vz, const float *mass)
{
    double k = 0.0;
    for (int i = 0; i < n; ++i) {
        double a = vx[i], b = vy[i], c = vz[i];
        k += 0.5 * (double)mass[i] * (a * a + b * b + c * c);
    }
    return k;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
