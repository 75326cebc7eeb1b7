/*
 * cpu_step.c -- the CPU baseline of the benchmark: the reference algorithm
 * for one velocity-Verlet MD step of rigid SPC water with SPME, SETTLE, a
 * Nose-Hoover chain and a C-rescale barostat, restated for the host's cores
 * (f32 arithmetic with FP64 reductions, OpenMP).
 *
 * Test and benchmark infrastructure only, like the rest of oracle/: the
 * benchmark times it beside the GPU path (the reference has no CPU path of
 * its own, book/src/features.md "Hardware") and the tests check it against
 * the FP64 oracle.  Nothing in the product loads it.
 *
 * What it restates, and where from:
 *   - the step order of run_step (src/integrator/mod.rs:344-429): chain
 *     half step, kick + drift, SETTLE on positions, forces, SETTLE on
 *     velocities (here right after the positions: the velocity projection
 *     acts on the kicked velocities, as the fused device step does), chain
 *     half step, closing kick, barostat move;
 *   - neighbour list: cells of at least r_search (neighbor_list.rs:1563),
 *     rebuilt when an atom has moved more than r_skin / 2 since the last
 *     build (neighbor.cu:50-75); a half list per atom (j > i), intramolecular
 *     pairs left out (topology.rs:1225-1245);
 *   - pair terms: LJ 24 eps / r^2 (2 s^12 - s^6) and the real-space Ewald
 *     term with the f32 Hastings erfc (lj.rs:225-300, spme.rs:708-783);
 *   - SPME order 4 (spme_spread_gather.cuh, spme_recip.cu): B-spline spread,
 *     3D FFT, influence function with the b-factors, inverse FFT, gather;
 *   - SETTLE: the three bond conditions solved by Newton's method on the
 *     multipliers (positions) and directly (velocities), in FP64;
 *   - Nose-Hoover chain (nose_hoover_chain.rs:58-279), Suzuki-Yoshida
 *     weights, FP64; C-rescale mu from the virial (barostat.cu:144-232,
 *     the deterministic part).
 * Box: orthorhombic; positions wrapped into [-L/2, L/2).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define CS_ORDER 4
#define CS_MAXNB 640 /* half-list neighbours per atom within r_search */

typedef struct {
    float re, im;
} cs_cplx;

typedef struct {
    int n, n_mol;
    float *x, *y, *z, *q, *m, *vx, *vy, *vz, *fx, *fy, *fz;
    float *x0, *y0, *z0; /* positions at the last list build */
    int *type, *mol;
    double L[3];
    /* LJ: two types, c12 / c6 per type pair (4 eps sigma^n) */
    float c6[4], c12[4];
    float rc2, rs2, alpha, kc, half_skin;
    /* neighbour list */
    int *nbr_count, *nbr;
    int ncell[3];
    int *cell_head, *cell_next;
    /* SPME */
    int ng;
    cs_cplx *grid;
    float *bmod; /* b-factors along one axis (cubic grid) */
    float *infl; /* influence function of the current box */
    double infl_box;
    cs_cplx *tw; /* twiddles e^{-2 pi i k / n} */
    int *slab_atoms, *slab_off;
    /* per-thread j-side force accumulators */
    int nthreads;
    float *tf;
    /* SETTLE */
    double d_oh, d_hh;
    /* integrator */
    float dt;
    double nhc_kt, nhc_tau;
    double xi[3], pxi[3], qm[3];
    int ndof;
    double pressure, tau_p, kappa;
    double virial, e_pot, ke;
    int rebuilds;
} cs_state;

/* ---- neighbour list ---------------------------------------------------- */

static float mi1(float d, float len, float ilen) { return d - len * rintf(d * ilen); }

static void build_list(cs_state *s)
{
    int n = s->n;
    float rs = sqrtf(s->rs2);
    for (int d = 0; d < 3; ++d) {
        s->ncell[d] = (int)floor(s->L[d] / rs);
        if (s->ncell[d] < 3) s->ncell[d] = 3;
    }
    int nc = s->ncell[0] * s->ncell[1] * s->ncell[2];
    s->cell_head = (int *)realloc(s->cell_head, sizeof(int) * nc);
    for (int c = 0; c < nc; ++c) s->cell_head[c] = -1;
    int *cell_of = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        int ca = (int)floorf((s->x[i] / (float)s->L[0] + 0.5f) * s->ncell[0]);
        int cb = (int)floorf((s->y[i] / (float)s->L[1] + 0.5f) * s->ncell[1]);
        int cc = (int)floorf((s->z[i] / (float)s->L[2] + 0.5f) * s->ncell[2]);
        ca = ca < 0 ? 0 : (ca >= s->ncell[0] ? s->ncell[0] - 1 : ca);
        cb = cb < 0 ? 0 : (cb >= s->ncell[1] ? s->ncell[1] - 1 : cb);
        cc = cc < 0 ? 0 : (cc >= s->ncell[2] ? s->ncell[2] - 1 : cc);
        cell_of[i] = (ca * s->ncell[1] + cb) * s->ncell[2] + cc;
    }
    for (int i = n - 1; i >= 0; --i) {
        s->cell_next[i] = s->cell_head[cell_of[i]];
        s->cell_head[cell_of[i]] = i;
    }
    float lx = (float)s->L[0], ly = (float)s->L[1], lz = (float)s->L[2];
    float ilx = 1.0f / lx, ily = 1.0f / ly, ilz = 1.0f / lz;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        int c = cell_of[i];
        int ca = c / (s->ncell[1] * s->ncell[2]), cb = (c / s->ncell[2]) % s->ncell[1],
            cc = c % s->ncell[2];
        int cnt = 0;
        int *out = s->nbr + (size_t)i * CS_MAXNB;
        for (int da = -1; da <= 1; ++da)
            for (int db = -1; db <= 1; ++db)
                for (int dc = -1; dc <= 1; ++dc) {
                    int a = (ca + da + s->ncell[0]) % s->ncell[0];
                    int b = (cb + db + s->ncell[1]) % s->ncell[1];
                    int k = (cc + dc + s->ncell[2]) % s->ncell[2];
                    for (int j = s->cell_head[(a * s->ncell[1] + b) * s->ncell[2] + k]; j >= 0;
                         j = s->cell_next[j]) {
                        if (j <= i || s->mol[j] == s->mol[i]) continue;
                        float dx = mi1(s->x[i] - s->x[j], lx, ilx);
                        float dy = mi1(s->y[i] - s->y[j], ly, ily);
                        float dz = mi1(s->z[i] - s->z[j], lz, ilz);
                        if (dx * dx + dy * dy + dz * dz <= s->rs2 && cnt < CS_MAXNB) out[cnt++] = j;
                    }
                }
        s->nbr_count[i] = cnt;
        s->x0[i] = s->x[i];
        s->y0[i] = s->y[i];
        s->z0[i] = s->z[i];
    }
    free(cell_of);
    s->rebuilds++;
}

static int list_stale(cs_state *s)
{
    float lx = (float)s->L[0], ly = (float)s->L[1], lz = (float)s->L[2];
    float ilx = 1.0f / lx, ily = 1.0f / ly, ilz = 1.0f / lz;
    float h2 = s->half_skin * s->half_skin;
    int stale = 0;
#pragma omp parallel for reduction(| : stale)
    for (int i = 0; i < s->n; ++i) {
        float dx = mi1(s->x[i] - s->x0[i], lx, ilx), dy = mi1(s->y[i] - s->y0[i], ly, ily),
              dz = mi1(s->z[i] - s->z0[i], lz, ilz);
        stale |= dx * dx + dy * dy + dz * dz > h2;
    }
    return stale;
}

/* ---- pair forces ------------------------------------------------------- */

static void pair_forces(cs_state *s, double *e_out, double *w_out)
{
    int n = s->n, T = s->nthreads;
    float lx = (float)s->L[0], ly = (float)s->L[1], lz = (float)s->L[2];
    float ilx = 1.0f / lx, ily = 1.0f / ly, ilz = 1.0f / lz;
    const float two_a_sqrtpi = 2.0f * s->alpha * 0.5641895835477563f;
    double e_tot = 0.0, w_tot = 0.0;
#pragma omp parallel reduction(+ : e_tot, w_tot)
    {
        int t = omp_get_thread_num();
        float *jf = s->tf + (size_t)t * 3 * n;
        memset(jf, 0, sizeof(float) * 3 * n);
#pragma omp for schedule(dynamic, 128)
        for (int i = 0; i < n; ++i) {
            float xi = s->x[i], yi = s->y[i], zi = s->z[i], qi = s->kc * s->q[i];
            int ti = s->type[i];
            float fx = 0.f, fy = 0.f, fz = 0.f, e = 0.f, w = 0.f;
            const int *nb = s->nbr + (size_t)i * CS_MAXNB;
            for (int k = 0; k < s->nbr_count[i]; ++k) {
                int j = nb[k];
                float dx = mi1(xi - s->x[j], lx, ilx), dy = mi1(yi - s->y[j], ly, ily),
                      dz = mi1(zi - s->z[j], lz, ilz);
                float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > s->rc2) continue;
                float ir2 = 1.0f / r2, ir = sqrtf(ir2), r = r2 * ir;
                float ir6 = ir2 * ir2 * ir2;
                int tp = ti * 2 + s->type[j];
                float elj = ir6 * (s->c12[tp] * ir6 - s->c6[tp]);
                float flj = ir2 * ir6 * (12.0f * s->c12[tp] * ir6 - 6.0f * s->c6[tp]);
                float ar = s->alpha * r, gauss = expf(-ar * ar);
                float tt = 1.0f / (1.0f + 0.3275911f * ar);
                float erfc_ar = (0.254829592f +
                                 (-0.284496736f +
                                  (1.421413741f + (-1.453152027f + 1.061405429f * tt) * tt) * tt) *
                                     tt) *
                                tt * gauss;
                float qq = qi * s->q[j];
                float fc = qq * ir2 * (erfc_ar * ir + two_a_sqrtpi * gauss);
                float f = flj + fc;
                fx += f * dx; fy += f * dy; fz += f * dz;
                jf[3 * j] -= f * dx; jf[3 * j + 1] -= f * dy; jf[3 * j + 2] -= f * dz;
                e += elj + qq * erfc_ar * ir;
                w += f * r2;
            }
            jf[3 * i] += fx; jf[3 * i + 1] += fy; jf[3 * i + 2] += fz;
            e_tot += e;
            w_tot += w;
        }
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            float a = 0.f, b = 0.f, c = 0.f;
            for (int u = 0; u < T; ++u) {
                const float *p = s->tf + (size_t)u * 3 * n + 3 * (size_t)i;
                a += p[0]; b += p[1]; c += p[2];
            }
            s->fx[i] = a; s->fy[i] = b; s->fz[i] = c;
        }
    }
    *e_out = e_tot;
    *w_out = w_tot;
}

/* ---- SPME -------------------------------------------------------------- */

static void bspline4(float t, float w[4], float dw[4])
{
    /* M_4 at k + t for k = 0..3 and its derivative (cardinal B-spline) */
    float u = 1.0f - t;
    w[0] = t * t * t / 6.0f;
    w[1] = (1.0f + 3.0f * t + 3.0f * t * t - 3.0f * t * t * t) / 6.0f;
    w[2] = (4.0f - 6.0f * t * t + 3.0f * t * t * t) / 6.0f;
    w[3] = u * u * u / 6.0f;
    dw[0] = 0.5f * t * t;
    dw[1] = 0.5f + t - 1.5f * t * t;
    dw[2] = -2.0f * t + 1.5f * t * t;
    dw[3] = -0.5f * u * u;
}

static void spline_origin(const cs_state *s, int i, int g0[3], float t[3])
{
    float p[3] = {s->x[i], s->y[i], s->z[i]};
    for (int d = 0; d < 3; ++d) {
        float u = (p[d] / (float)s->L[d] + 0.5f) * s->ng;
        float f = floorf(u);
        t[d] = u - f;
        g0[d] = (int)f;
    }
}

static int wrapi(int v, int n) { return ((v % n) + n) % n; }

static void fft_line(cs_cplx *a, int n, const cs_cplx *tw, int inverse)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            cs_cplx t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int step = n / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                cs_cplx w = tw[k * step];
                if (inverse) w.im = -w.im;
                cs_cplx u = a[i + k], v = a[i + k + len / 2];
                float vr = v.re * w.re - v.im * w.im, vi = v.re * w.im + v.im * w.re;
                a[i + k].re = u.re + vr; a[i + k].im = u.im + vi;
                a[i + k + len / 2].re = u.re - vr; a[i + k + len / 2].im = u.im - vi;
            }
    }
}

static void fft3d(cs_state *s, int inverse)
{
    int n = s->ng;
    cs_cplx *g = s->grid;
    /* c axis: contiguous lines */
#pragma omp parallel for schedule(static)
    for (int l = 0; l < n * n; ++l) fft_line(g + (size_t)l * n, n, s->tw, inverse);
    /* b and a axes: gathered lines */
#pragma omp parallel
    {
        cs_cplx *line = (cs_cplx *)malloc(sizeof(cs_cplx) * n);
#pragma omp for schedule(static)
        for (int l = 0; l < n * n; ++l) {
            int a = l / n, c = l % n;
            for (int b = 0; b < n; ++b) line[b] = g[((size_t)a * n + b) * n + c];
            fft_line(line, n, s->tw, inverse);
            for (int b = 0; b < n; ++b) g[((size_t)a * n + b) * n + c] = line[b];
        }
#pragma omp for schedule(static)
        for (int l = 0; l < n * n; ++l) {
            int b = l / n, c = l % n;
            for (int a = 0; a < n; ++a) line[a] = g[((size_t)a * n + b) * n + c];
            fft_line(line, n, s->tw, inverse);
            for (int a = 0; a < n; ++a) g[((size_t)a * n + b) * n + c] = line[a];
        }
        free(line);
    }
}

static void influence(cs_state *s)
{
    int n = s->ng;
    double vol = s->L[0] * s->L[1] * s->L[2];
    double a2 = (double)s->alpha * s->alpha;
#pragma omp parallel for schedule(static)
    for (int a = 0; a < n; ++a) {
        double ma = a <= n / 2 ? a : a - n;
        for (int b = 0; b < n; ++b) {
            double mb = b <= n / 2 ? b : b - n;
            for (int c = 0; c < n; ++c) {
                double mc = c <= n / 2 ? c : c - n;
                double kx = 2.0 * M_PI * ma / s->L[0], ky = 2.0 * M_PI * mb / s->L[1],
                       kz = 2.0 * M_PI * mc / s->L[2];
                double k2 = kx * kx + ky * ky + kz * kz;
                double g = 0.0;
                if (k2 != 0.0)
                    g = s->kc / vol * (4.0 * M_PI / k2) * exp(-k2 / (4.0 * a2)) * s->bmod[a] *
                        s->bmod[b] * s->bmod[c];
                s->infl[((size_t)a * n + b) * n + c] = (float)g;
            }
        }
    }
    s->infl_box = s->L[0];
}

static void spme(cs_state *s, double *e_out, double *w_out)
{
    int n = s->ng, N = s->n;
    size_t M = (size_t)n * n * n;
    /* spread: atoms bucketed by the first stencil plane of the a axis into
     * slabs of 4 planes; even slabs, then odd slabs, are spread in parallel
     * (a stencil spans 4 planes, so slabs of one parity never overlap) */
    int nslab = n / 4;
    memset(s->slab_off, 0, sizeof(int) * (nslab + 1));
    int *sl = (int *)malloc(sizeof(int) * N);
    for (int i = 0; i < N; ++i) {
        int g0[3];
        float t[3];
        spline_origin(s, i, g0, t);
        sl[i] = wrapi(g0[0] - 3, n) / 4;
        s->slab_off[sl[i] + 1]++;
    }
    for (int k = 0; k < nslab; ++k) s->slab_off[k + 1] += s->slab_off[k];
    int *cur = (int *)malloc(sizeof(int) * nslab);
    memcpy(cur, s->slab_off, sizeof(int) * nslab);
    for (int i = 0; i < N; ++i) s->slab_atoms[cur[sl[i]]++] = i;
    free(cur);
    free(sl);
    memset(s->grid, 0, sizeof(cs_cplx) * M);
    for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = parity; k < nslab; k += 2)
            for (int u = s->slab_off[k]; u < s->slab_off[k + 1]; ++u) {
                int i = s->slab_atoms[u];
                int g0[3];
                float t[3], w[3][4], dw[3][4];
                spline_origin(s, i, g0, t);
                for (int d = 0; d < 3; ++d) bspline4(t[d], w[d], dw[d]);
                float q = s->q[i];
                for (int da = 0; da < 4; ++da) {
                    int ga = wrapi(g0[0] - da, n);
                    for (int db = 0; db < 4; ++db) {
                        int gb = wrapi(g0[1] - db, n);
                        float qab = q * w[0][da] * w[1][db];
                        cs_cplx *row = s->grid + ((size_t)ga * n + gb) * n;
                        for (int dc = 0; dc < 4; ++dc) row[wrapi(g0[2] - dc, n)].re += qab * w[2][dc];
                    }
                }
            }
    }
    fft3d(s, 0);
    if (s->infl_box != s->L[0]) influence(s);
    double a2 = (double)s->alpha * s->alpha, er = 0.0, wr = 0.0;
#pragma omp parallel for reduction(+ : er, wr) schedule(static)
    for (int a = 0; a < n; ++a) {
        double ma = a <= n / 2 ? a : a - n;
        for (int b = 0; b < n; ++b) {
            double mb = b <= n / 2 ? b : b - n;
            for (int c = 0; c < n; ++c) {
                double mc = c <= n / 2 ? c : c - n;
                size_t o = ((size_t)a * n + b) * n + c;
                float g = s->infl[o];
                cs_cplx v = s->grid[o];
                double r2 = (double)v.re * v.re + (double)v.im * v.im;
                double kx = ma / s->L[0], ky = mb / s->L[1], kz = mc / s->L[2];
                double k2 = 4.0 * M_PI * M_PI * (kx * kx + ky * ky + kz * kz);
                er += 0.5 * g * r2;
                wr += 0.5 * g * (1.0 - k2 / (2.0 * a2)) * r2;
                s->grid[o].re = v.re * g;
                s->grid[o].im = v.im * g;
            }
        }
    }
    fft3d(s, 1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        int g0[3];
        float t[3], w[3][4], dw[3][4];
        spline_origin(s, i, g0, t);
        for (int d = 0; d < 3; ++d) bspline4(t[d], w[d], dw[d]);
        float ga = 0.f, gb = 0.f, gc = 0.f;
        for (int da = 0; da < 4; ++da) {
            int ia = wrapi(g0[0] - da, n);
            for (int db = 0; db < 4; ++db) {
                int ib = wrapi(g0[1] - db, n);
                const cs_cplx *row = s->grid + ((size_t)ia * n + ib) * n;
                for (int dc = 0; dc < 4; ++dc) {
                    float v = row[wrapi(g0[2] - dc, n)].re;
                    ga += v * dw[0][da] * w[1][db] * w[2][dc];
                    gb += v * w[0][da] * dw[1][db] * w[2][dc];
                    gc += v * w[0][da] * w[1][db] * dw[2][dc];
                }
            }
        }
        float q = s->q[i];
        s->fx[i] -= q * ga * (float)(n / s->L[0]);
        s->fy[i] -= q * gb * (float)(n / s->L[1]);
        s->fz[i] -= q * gc * (float)(n / s->L[2]);
    }
    double self = 0.0;
    for (int i = 0; i < N; ++i) self += (double)s->q[i] * s->q[i];
    *e_out = er - s->kc * s->alpha / sqrt(M_PI) * self;
    *w_out = wr;
}

/* ---- SETTLE ------------------------------------------------------------ */

static const int tri_i[3] = {0, 0, 1}, tri_j[3] = {1, 2, 2};
static double tri_sign(int k, int c) { return k == tri_i[c] ? 1.0 : (k == tri_j[c] ? -1.0 : 0.0); }

static void solve3(double A[3][3], const double b[3], double x[3])
{
    double c[3][3];
    for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) {
            int r1 = (r + 1) % 3, r2 = (r + 2) % 3, q1 = (q + 1) % 3, q2 = (q + 2) % 3;
            c[r][q] = A[r1][q1] * A[r2][q2] - A[r1][q2] * A[r2][q1];
        }
    double det = A[0][0] * c[0][0] + A[0][1] * c[0][1] + A[0][2] * c[0][2];
    double inv = det != 0.0 ? 1.0 / det : 0.0;
    for (int r = 0; r < 3; ++r) x[r] = (c[0][r] * b[0] + c[1][r] * b[1] + c[2][r] * b[2]) * inv;
}

static double mi1d(double d, double len) { return d - len * floor(d / len + 0.5); }

/* one water: u[k] drifted positions, s0[k] snapshot (f32 in, double math) */
static void settle_water(cs_state *s, const int id[3], const float sx[3], const float sy[3],
                         const float sz[3], double *wvir)
{
    double m[3], im[3];
    for (int k = 0; k < 3; ++k) {
        m[k] = s->m[id[k]];
        im[k] = 1.0 / m[k];
    }
    double tgt[3] = {s->d_oh * s->d_oh, s->d_oh * s->d_oh, s->d_hh * s->d_hh};
    double S[3][3], U[3][3];
    for (int k = 0; k < 3; ++k) {
        double snap[3] = {sx[k], sy[k], sz[k]}, o[3] = {sx[0], sy[0], sz[0]};
        double p[3] = {s->x[id[k]], s->y[id[k]], s->z[id[k]]};
        for (int d = 0; d < 3; ++d) {
            S[k][d] = mi1d(snap[d] - o[d], s->L[d]);
            U[k][d] = S[k][d] + mi1d(p[d] - snap[d], s->L[d]);
        }
    }
    double G[3][3], K[3][3], lam[3] = {0, 0, 0}, X[3][3];
    for (int c = 0; c < 3; ++c) {
        for (int d = 0; d < 3; ++d) G[c][d] = S[tri_i[c]][d] - S[tri_j[c]][d];
        for (int q = 0; q < 3; ++q)
            K[c][q] = tri_sign(tri_i[c], q) * im[tri_i[c]] - tri_sign(tri_j[c], q) * im[tri_j[c]];
    }
    memcpy(X, U, sizeof(X));
    for (int it = 0; it < 8; ++it) {
        double B[3][3], res[3], J[3][3], rhs[3], dl[3], worst = 0.0;
        for (int c = 0; c < 3; ++c) {
            for (int d = 0; d < 3; ++d) B[c][d] = X[tri_i[c]][d] - X[tri_j[c]][d];
            res[c] = B[c][0] * B[c][0] + B[c][1] * B[c][1] + B[c][2] * B[c][2] - tgt[c];
            worst = fmax(worst, fabs(res[c]) / tgt[c]);
        }
        if (worst <= 1e-14) break;
        for (int c = 0; c < 3; ++c) {
            for (int q = 0; q < 3; ++q)
                J[c][q] = 2.0 * K[c][q] * (B[c][0] * G[q][0] + B[c][1] * G[q][1] + B[c][2] * G[q][2]);
            rhs[c] = -res[c];
        }
        solve3(J, rhs, dl);
        for (int c = 0; c < 3; ++c) lam[c] += dl[c];
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) {
                double sh = 0.0;
                for (int c = 0; c < 3; ++c) sh += tri_sign(k, c) * lam[c] * G[c][d];
                X[k][d] = U[k][d] + im[k] * sh;
            }
    }
    double idt = 1.0 / s->dt, msum = m[0] + m[1] + m[2], com[3];
    for (int d = 0; d < 3; ++d) com[d] = (m[0] * X[0][d] + m[1] * X[1][d] + m[2] * X[2][d]) / msum;
    double w = 0.0;
    float *pos[3] = {s->x, s->y, s->z}, *vel[3] = {s->vx, s->vy, s->vz};
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) {
            double sh = X[k][d] - U[k][d];
            pos[d][id[k]] = (float)((double)pos[d][id[k]] + sh);
            vel[d][id[k]] = (float)((double)vel[d][id[k]] + sh * idt);
            w += m[k] * idt * idt * sh * (X[k][d] - com[d]);
        }
    /* velocities: zero bond-relative velocities along the new bonds */
    double Y[3][3], R[3][3], A[3][3], b[3], mu[3];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d)
            Y[k][d] = mi1d((double)pos[d][id[k]] - (double)pos[d][id[0]], s->L[d]);
    for (int c = 0; c < 3; ++c) {
        double dv = 0.0;
        for (int d = 0; d < 3; ++d) {
            R[c][d] = Y[tri_i[c]][d] - Y[tri_j[c]][d];
            dv += R[c][d] * ((double)vel[d][id[tri_i[c]]] - (double)vel[d][id[tri_j[c]]]);
        }
        b[c] = -dv;
    }
    for (int c = 0; c < 3; ++c)
        for (int q = 0; q < 3; ++q)
            A[c][q] = K[c][q] * (R[c][0] * R[q][0] + R[c][1] * R[q][1] + R[c][2] * R[q][2]);
    solve3(A, b, mu);
    for (int d = 0; d < 3; ++d) com[d] = (m[0] * Y[0][d] + m[1] * Y[1][d] + m[2] * Y[2][d]) / msum;
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) {
            double sh = 0.0;
            for (int c = 0; c < 3; ++c) sh += tri_sign(k, c) * mu[c] * R[c][d];
            double kick = im[k] * sh;
            vel[d][id[k]] = (float)((double)vel[d][id[k]] + kick);
            w += m[k] * idt * kick * (Y[k][d] - com[d]);
        }
    *wvir += w;
}

/* ---- thermostat / kinetic energy --------------------------------------- */

static double kinetic(cs_state *s)
{
    double ke = 0.0;
#pragma omp parallel for reduction(+ : ke) schedule(static)
    for (int i = 0; i < s->n; ++i)
        ke += 0.5 * s->m[i] * ((double)s->vx[i] * s->vx[i] + (double)s->vy[i] * s->vy[i] +
                               (double)s->vz[i] * s->vz[i]);
    return ke;
}

/* One Nose-Hoover chain half step (chain 3, Suzuki-Yoshida 3 weights),
 * returns the velocity scale factor (nose_hoover_chain.rs:109-279). */
static double nhc_half(cs_state *s, double ke)
{
    static const double yw[3] = {1.3512071919596577, -1.7024143839193155, 1.3512071919596577};
    const int M = 3;
    double scale = 1.0, kt = s->nhc_kt, ndof = s->ndof;
    for (int y = 0; y < 3; ++y) {
        double h = 0.5 * s->dt * yw[y];
        double g[3];
        g[M - 1] = (s->pxi[M - 2] * s->pxi[M - 2] / s->qm[M - 2] - kt);
        s->pxi[M - 1] += 0.5 * h * g[M - 1];
        for (int j = M - 2; j >= 0; --j) {
            double gj = j == 0 ? (2.0 * ke * scale * scale - ndof * kt)
                               : (s->pxi[j - 1] * s->pxi[j - 1] / s->qm[j - 1] - kt);
            double e = exp(-0.25 * h * s->pxi[j + 1] / s->qm[j + 1]);
            s->pxi[j] = s->pxi[j] * e * e + 0.5 * h * gj * e;
        }
        scale *= exp(-h * s->pxi[0] / s->qm[0]);
        for (int j = 0; j < M; ++j) s->xi[j] += h * s->pxi[j] / s->qm[j];
        for (int j = 0; j < M - 1; ++j) {
            double gj = j == 0 ? (2.0 * ke * scale * scale - ndof * kt)
                               : (s->pxi[j - 1] * s->pxi[j - 1] / s->qm[j - 1] - kt);
            double e = exp(-0.25 * h * s->pxi[j + 1] / s->qm[j + 1]);
            s->pxi[j] = s->pxi[j] * e * e + 0.5 * h * gj * e;
        }
        s->pxi[M - 1] += 0.5 * h * (s->pxi[M - 2] * s->pxi[M - 2] / s->qm[M - 2] - kt);
    }
    return scale;
}

/* ---- the step ---------------------------------------------------------- */

static void forces(cs_state *s)
{
    if (list_stale(s)) build_list(s);
    double ep, wp, er, wr;
    pair_forces(s, &ep, &wp);
    spme(s, &er, &wr);
    s->e_pot = ep + er;
    s->virial = wp + wr;
}

static void scale_velocities(cs_state *s, float f)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        s->vx[i] *= f; s->vy[i] *= f; s->vz[i] *= f;
    }
}

void cs_step(cs_state *s, int nsteps)
{
    float dt = s->dt, hdt = 0.5f * dt;
    for (int k = 0; k < nsteps; ++k) {
        scale_velocities(s, (float)nhc_half(s, kinetic(s)));
        float *sx = (float *)malloc(sizeof(float) * 9 * s->n_mol);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < s->n; ++i) {
            float im = 1.0f / s->m[i];
            s->vx[i] += s->fx[i] * im * hdt; s->vy[i] += s->fy[i] * im * hdt;
            s->vz[i] += s->fz[i] * im * hdt;
            if (i % 3 == 0) {
                int w = i / 3;
                for (int a = 0; a < 3; ++a) {
                    sx[9 * w + a] = s->x[i + a];
                    sx[9 * w + 3 + a] = s->y[i + a];
                    sx[9 * w + 6 + a] = s->z[i + a];
                }
            }
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < s->n; ++i) {
            float p[3] = {s->x[i] + s->vx[i] * dt, s->y[i] + s->vy[i] * dt, s->z[i] + s->vz[i] * dt};
            for (int d = 0; d < 3; ++d) p[d] -= (float)s->L[d] * floorf(p[d] / (float)s->L[d] + 0.5f);
            s->x[i] = p[0]; s->y[i] = p[1]; s->z[i] = p[2];
        }
        double wc = 0.0;
#pragma omp parallel for reduction(+ : wc) schedule(static)
        for (int w = 0; w < s->n_mol; ++w) {
            int id[3] = {3 * w, 3 * w + 1, 3 * w + 2};
            settle_water(s, id, sx + 9 * w, sx + 9 * w + 3, sx + 9 * w + 6, &wc);
        }
        free(sx);
        scale_velocities(s, (float)nhc_half(s, kinetic(s)));
        forces(s);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < s->n; ++i) {
            float im = 1.0f / s->m[i];
            s->vx[i] += s->fx[i] * im * hdt; s->vy[i] += s->fy[i] * im * hdt;
            s->vz[i] += s->fz[i] * im * hdt;
        }
        /* C-rescale: isotropic mu from the pressure (deterministic part) */
        if (s->tau_p > 0.0) {
            double vol = s->L[0] * s->L[1] * s->L[2];
            s->ke = kinetic(s);
            double p_inst = (2.0 * s->ke + s->virial + wc) / (3.0 * vol);
            double mu = exp(-s->kappa * s->dt / (3.0 * s->tau_p) * (s->pressure - p_inst));
#pragma omp parallel for schedule(static)
            for (int i = 0; i < s->n; ++i) {
                s->x[i] *= (float)mu; s->y[i] *= (float)mu; s->z[i] *= (float)mu;
            }
            for (int d = 0; d < 3; ++d) s->L[d] = (double)(float)(s->L[d] * mu);
        }
    }
}

/* posq: 4n floats (waters O, H, H consecutive); lj: c6[4], c12[4] over the
 * type pairs (t_i * 2 + t_j); spme grid ng (power of two, cubic) */
cs_state *cs_create(int n, const float *posq, const float *mass, const int *type,
                    const double *L, const float *c6, const float *c12, double rc, double skin,
                    double alpha, double kc, int ng, double d_oh, double d_hh, double dt,
                    double kt, double tau, double pressure, double tau_p, double kappa)
{
    cs_state *s = (cs_state *)calloc(1, sizeof(cs_state));
    s->n = n;
    s->n_mol = n / 3;
    float **arrs[] = {&s->x, &s->y, &s->z, &s->q, &s->m, &s->vx, &s->vy, &s->vz,
                      &s->fx, &s->fy, &s->fz, &s->x0, &s->y0, &s->z0};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a)
        *arrs[a] = (float *)calloc((size_t)n, sizeof(float));
    s->type = (int *)malloc(sizeof(int) * n);
    s->mol = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        s->x[i] = posq[4 * i]; s->y[i] = posq[4 * i + 1]; s->z[i] = posq[4 * i + 2];
        s->q[i] = posq[4 * i + 3];
        s->m[i] = mass[i];
        s->type[i] = type[i];
        s->mol[i] = i / 3;
    }
    for (int d = 0; d < 3; ++d) s->L[d] = L[d];
    for (int k = 0; k < 4; ++k) { s->c6[k] = c6[k]; s->c12[k] = c12[k]; }
    s->rc2 = (float)(rc * rc);
    s->rs2 = (float)((rc + skin) * (rc + skin));
    s->half_skin = (float)(0.5 * skin);
    s->alpha = (float)alpha;
    s->kc = (float)kc;
    s->d_oh = d_oh;
    s->d_hh = d_hh;
    s->dt = (float)dt;
    s->nhc_kt = kt;
    s->nhc_tau = tau;
    s->ndof = 3 * n - 3 * s->n_mol - 3;
    s->qm[0] = s->ndof * kt * tau * tau;
    s->qm[1] = s->qm[2] = kt * tau * tau;
    s->pressure = pressure;
    s->tau_p = tau_p;
    s->kappa = kappa;
    s->nbr_count = (int *)calloc((size_t)n, sizeof(int));
    s->nbr = (int *)malloc(sizeof(int) * (size_t)n * CS_MAXNB);
    s->cell_next = (int *)malloc(sizeof(int) * n);
    s->ng = ng;
    size_t M = (size_t)ng * ng * ng;
    s->grid = (cs_cplx *)malloc(sizeof(cs_cplx) * M);
    s->infl = (float *)malloc(sizeof(float) * M);
    s->bmod = (float *)malloc(sizeof(float) * ng);
    s->tw = (cs_cplx *)malloc(sizeof(cs_cplx) * ng);
    for (int k = 0; k < ng; ++k) {
        s->tw[k].re = (float)cos(-2.0 * M_PI * k / ng);
        s->tw[k].im = (float)sin(-2.0 * M_PI * k / ng);
        /* b-factor 1 / |sum_j M_4(j + 1) e^{2 pi i j k / n}|^2 */
        double mr = 0.0, mi = 0.0, m4[3] = {1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0};
        for (int j = 0; j < 3; ++j) {
            mr += m4[j] * cos(2.0 * M_PI * j * k / ng);
            mi += m4[j] * sin(2.0 * M_PI * j * k / ng);
        }
        s->bmod[k] = (float)(1.0 / (mr * mr + mi * mi));
    }
    s->slab_atoms = (int *)malloc(sizeof(int) * n);
    s->slab_off = (int *)malloc(sizeof(int) * (ng / 4 + 1));
    s->infl_box = -1.0;
    s->nthreads = omp_get_max_threads();
    s->tf = (float *)malloc(sizeof(float) * 3 * (size_t)n * s->nthreads);
    build_list(s);
    s->rebuilds = 0;
    forces(s);
    return s;
}

void cs_set_velocities(cs_state *s, const float *vx, const float *vy, const float *vz)
{
    memcpy(s->vx, vx, sizeof(float) * s->n);
    memcpy(s->vy, vy, sizeof(float) * s->n);
    memcpy(s->vz, vz, sizeof(float) * s->n);
}

/* f: 3n forces (x, y, z per atom); out: [e_pot, virial, ke, rebuilds,
 * threads] */
void cs_get(cs_state *s, float *posq, float *f, double *out)
{
    for (int i = 0; i < s->n; ++i) {
        if (posq) {
            posq[4 * i] = s->x[i]; posq[4 * i + 1] = s->y[i]; posq[4 * i + 2] = s->z[i];
            posq[4 * i + 3] = s->q[i];
        }
        if (f) { f[3 * i] = s->fx[i]; f[3 * i + 1] = s->fy[i]; f[3 * i + 2] = s->fz[i]; }
    }
    if (out) {
        out[0] = s->e_pot;
        out[1] = s->virial;
        out[2] = kinetic(s);
        out[3] = s->rebuilds;
        out[4] = s->nthreads;
    }
}

void cs_free(cs_state *s)
{
    if (!s) return;
    float *arrs[] = {s->x, s->y, s->z, s->q, s->m, s->vx, s->vy, s->vz,
                     s->fx, s->fy, s->fz, s->x0, s->y0, s->z0};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); ++a) free(arrs[a]);
    free(s->type); free(s->mol); free(s->nbr_count); free(s->nbr); free(s->cell_head);
    free(s->cell_next); free(s->grid); free(s->infl); free(s->bmod); free(s->tw);
    free(s->slab_atoms); free(s->slab_off); free(s->tf);
    free(s);
}
