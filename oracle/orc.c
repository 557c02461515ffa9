/*
 * oracle/orc.c -- CPU restatement of the USER-MESO mDPD/DPD per-timestep path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and the
 * cpu_baseline leg of bench.py.  The engine never links or calls it.
 *
 * Compiled with -ffp-contract=off so every FP32 expression that must be
 * bit-exact (neighbor test, wrap, key arithmetic) rounds after each operation.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================
 * TEA, signatures, Morton                      src/USER-MESO/math_meso.h
 * ====================================================================== */

/* author's key and delta: math_meso.h:443-447 */
static const uint32_t TEA_KEY[4] = {0xA341316Cu, 0xC8013EA4u, 0xAD90777Du, 0x7E95761Eu};
#define TEA_DELTA 0x9E3779B9u

void orc_tea_key(uint32_t *v0, uint32_t *v1, int rounds, const uint32_t key[4])
{
    /* one Feistel cycle per round, sum advanced before the two half rounds:
     * math_meso.h:449-455 (same cycle as Wheeler & Needham's TEA encrypt) */
    uint32_t a = *v0, b = *v1, sum = 0;
    for (int r = 0; r < rounds; ++r) {
        sum += TEA_DELTA;
        a += ((b << 4) + key[0]) ^ (b + sum) ^ ((b >> 5) + key[1]);
        b += ((a << 4) + key[2]) ^ (a + sum) ^ ((a >> 5) + key[3]);
    }
    *v0 = a;
    *v1 = b;
}

void orc_tea(uint32_t *v0, uint32_t *v1, int rounds) { orc_tea_key(v0, v1, rounds, TEA_KEY); }

uint32_t orc_premix(uint32_t v0, uint32_t v1, int rounds)
{
    orc_tea(&v0, &v1, rounds);          /* math_meso.h:459-463 */
    return v0 ^ v1;
}

uint32_t orc_brev32(uint32_t x)
{
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}

static uint64_t bit_space3(uint32_t x)
{
    /* bit i of x goes to bit 3i, for i < 21: math_meso.h:178-191 */
    uint64_t r = 0;
    for (int i = 0; i < 21; ++i) r |= (uint64_t)((x >> i) & 1u) << (3 * i);
    return r;
}

uint64_t orc_interleave3(uint32_t i, uint32_t j, uint32_t k)
{
    return bit_space3(i) | (bit_space3(j) << 1) | (bit_space3(k) << 2);   /* :193-201 */
}

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

uint32_t orc_mantissa(float u, float v, float w)
{
    /* math_meso.h:435-441: bits 12..22 of each word, interleaved, truncated to
     * 32 bits by the uint return type */
    uint32_t i = f2u(u) & 0x7FF000u;
    uint32_t j = f2u(v) & 0x7FF000u;
    uint32_t k = f2u(w) & 0x7FF000u;
    return (uint32_t)orc_interleave3(i >> 12, j >> 12, k >> 12);
}

uint32_t orc_seed_now(uint32_t seed, int64_t ntimestep)
{
    return orc_premix(seed, (uint32_t)ntimestep, 64);   /* pair_mdpd_meso.cu:295-297 */
}

uint32_t orc_sig_mdpd(int32_t tag, uint32_t seed_t)
{
    return orc_premix(orc_brev32((uint32_t)tag), seed_t, 32);   /* atom_vec_mdpd_atomic_meso.cu:166 */
}

uint32_t orc_sig_dpd(int32_t tag, uint32_t seed_t, float vx, float vy, float vz)
{
    /* atom_vec_meso.cu:164 */
    return seed_t ^ orc_premix(orc_brev32((uint32_t)tag), orc_mantissa(vx, vy, vz), 16);
}

void orc_pair_bits(uint32_t si, uint32_t sj, uint32_t *v0, uint32_t *v1)
{
    /* gaussian_TEA<4>(si > sj, si, sj): v0 = larger, v1 = smaller (unsigned);
     * pair_mdpd_meso.cu:163, math_meso.h:465-469 */
    int pred = si > sj;
    uint32_t a = pred ? si : sj;
    uint32_t b = pred ? sj : si;
    orc_tea(&a, &b, 4);
    *v0 = a;
    *v1 = b;
}

/* ======================================================================
 * Gaussian from the TEA bits                    math_meso.h:208-237,370-473
 * ====================================================================== */

#define LN_2      6.9314718055994528623E-1
#define SQRT_2    1.4142135623730950488
#define ONE_OVER_SQ2 7.0710678118654757274E-1
#define TWO_M31   4.6566128730773925781E-10

double orc_cospi_poly(double x)
{
    /* author's odd Chebyshev fit of cos(pi x) on x in [0,1] after x -> 2x-1;
     * coefficients from math_meso.h:372-377, evaluation :379-390 */
    static const double C[6] = {-1.57079632662144460E+0, 6.45964092644060746E-1,
                                -7.96925872866600517E-2, 4.68162024021793872E-3,
                                -1.60217135750921262E-4, 3.41817283473266926E-6};
    x = 2.0 * x - 1.0;
    double x2 = x * x;
    double s = C[5];
    for (int k = 4; k >= 0; --k) s = fma(s, x2, C[k]);
    return s * x;
}

double orc_log2u_poly(uint32_t x)
{
    /* log2(x / 2^32) for x >= 1: integer part from the leading bit, fraction
     * from the author's Chebyshev fit on [1/sqrt2, sqrt2]; math_meso.h:392-423 */
    static const double L[5] = {2.88539008179006374E+0, 2.83126505877817866E-2,
                                5.00072802051539862E-4, 1.05013262724846015E-5,
                                2.55854634203511155E-7};
    int I = 31 - __builtin_clz(x);
    double y = (double)x * ldexp(1.0, -I);
    double e = (double)(I - 32);
    int pred = y > SQRT_2;
    y *= pred ? 0.5 : ONE_OVER_SQ2;
    double z = (y - 1.0) * (1.0 / (y + 1.0));
    double t = z * z * 33.9705627484771406;
    double s = L[4];
    for (int k = 3; k >= 0; --k) s = fma(s, t, L[k]);
    return fma(z, s, (pred ? 1.0 : 0.5) + e);
}

static double clamp4(double g) { return g < -4.0 ? -4.0 : (g > 4.0 ? 4.0 : g); }

double orc_gaussian_poly(uint32_t si, uint32_t sj)
{
    uint32_t v0, v1;
    orc_pair_bits(si, sj, &v0, &v1);
    double f = orc_cospi_poly((double)(v0 & 0x7FFFFFFFu) * TWO_M31) * ((v0 & 0x80000000u) ? 1.0 : -1.0);
    double r = sqrt(-2.0 * LN_2 * orc_log2u_poly(v1 > 1u ? v1 : 1u));
    return clamp4(r * f);                                    /* math_meso.h:465-473 */
}

double orc_gaussian_libm(uint32_t si, uint32_t sj)
{
    uint32_t v0, v1;
    orc_pair_bits(si, sj, &v0, &v1);
    double f = cos(M_PI * (double)(v0 & 0x7FFFFFFFu) * TWO_M31) * ((v0 & 0x80000000u) ? 1.0 : -1.0);
    double u = (double)(v1 > 1u ? v1 : 1u) * ldexp(1.0, -32);
    double r = sqrt(-2.0 * log(u));
    return clamp4(r * f);
}

/* ======================================================================
 * Bins, sort keys, stable sort, wrap
 * ====================================================================== */

void orc_setup_bins(const double lo[3], const double hi[3], double cutneigh, orc_bins *b)
{
    /* neighbor_meso.cu:837-853: bin >= cutneighmax, one ghost layer per side */
    double inv = 1.0 / cutneigh;
    int maxm = 0;
    for (int d = 0; d < 3; ++d) {
        double L = hi[d] - lo[d];
        int nb = (int)(L * inv);
        if (nb < 1) nb = 1;
        b->lo[d] = lo[d];
        b->hi[d] = hi[d];
        b->nbin[d] = nb;
        b->mbin[d] = nb + 2;
        b->binsize[d] = L / nb;
        b->bininv[d] = 1.0 / b->binsize[d];
        if (b->mbin[d] > maxm) maxm = b->mbin[d];
    }
    b->cutneigh = cutneigh;
    /* atom_meso.cu:353-360 */
    b->l1_width = 3 * (int)floor(log2(maxm * 2.0));
    b->l2_width = 12;
    b->l1_shift = 12;
}

static int iclamp(int i, int lo, int hi_excl)
{
    int m = i < hi_excl - 1 ? i : hi_excl - 1;    /* clamp at [nmin, nmax): math_meso.h:165-169 */
    return m > lo ? m : lo;
}

void orc_bin_of(const orc_bins *b, float x, float y, float z, int bin[3])
{
    const float c[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        /* (int)((x - lo) * bininv + 1.0) in FP64 from the FP32 value;
         * neighbor_meso.cu:414-416.  Clamped to the interior [1, mbin-2]:
         * after the on-device wrap every particle is local. */
        volatile double t = ((double)c[d] - b->lo[d]) * b->bininv[d];
        double u = t + 1.0;
        bin[d] = iclamp((int)u, 1, b->mbin[d] - 1);
    }
}

uint64_t orc_sort_key(const orc_bins *b, float x, float y, float z)
{
    int bin[3];
    orc_bin_of(b, x, y, z, bin);
    const float c[3] = {x, y, z};
    uint32_t sub[3];
    for (int d = 0; d < 3; ++d) {
        /* sub-bin: (x - (bin-1)*binsize) * 16 * bininv, clamp to [0,16);
         * the box origin is not subtracted, as in atom_meso.cu:298-300 */
        volatile double off = (double)(uint32_t)(bin[d] - 1) * b->binsize[d];
        volatile double rel = (double)c[d] - off;
        double coord_inv = 16.0 * b->bininv[d];
        double s = rel * coord_inv;
        sub[d] = (uint32_t)iclamp((int)s, 0, 16);
    }
    uint64_t z1 = orc_interleave3((uint32_t)bin[0], (uint32_t)bin[1], (uint32_t)bin[2]);
    uint64_t z2 = orc_interleave3(sub[0], sub[1], sub[2]);
    return (z1 << b->l1_shift) | z2;           /* atom_meso.cu:301-302, border bit 0 */
}

typedef struct { uint64_t k; int i; } kv_t;
static int kv_cmp(const void *a, const void *b)
{
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    if (x->k != y->k) return x->k < y->k ? -1 : 1;
    return x->i < y->i ? -1 : (x->i > y->i);
}

void orc_stable_sort_u64(int n, const uint64_t *key, int *perm)
{
    /* stable by key == sort by (key, original index) */
    kv_t *a = (kv_t *)malloc(sizeof(kv_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) { a[i].k = key[i]; a[i].i = i; }
    qsort(a, (size_t)n, sizeof(kv_t), kv_cmp);
    for (int i = 0; i < n; ++i) perm[i] = a[i].i;
    free(a);
}

void orc_wrap_fp32(int n, float *x, const double lo[3], const double hi[3])
{
    /* Domain::pbc (src/domain.cpp): x<lo -> x+=prd; x>=hi -> x-=prd, max(x,lo) */
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            float flo = (float)lo[d], fhi = (float)hi[d], prd = (float)(hi[d] - lo[d]);
            float c = x[3 * i + d];
            if (c < flo) c = c + prd;
            if (c >= fhi) { c = c - prd; if (c < flo) c = flo; }
            x[3 * i + d] = c;
        }
}

/* ======================================================================
 * FP32 neighbor sets (brute force)           neigh_build_meso.cu:86-117
 * ====================================================================== */

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

int orc_neighbors_fp32(int n, const float *x, const double lo[3], const double hi[3],
                       double cut, int maxnbr, int *counts, int *list)
{
    float prd[3];
    for (int d = 0; d < 3; ++d) prd[d] = (float)(hi[d] - lo[d]);
    const float rc2 = (float)(cut * cut);
    int maxc = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxc)
    for (int i = 0; i < n; ++i) {
        int c = 0;
        int *row = list + (size_t)i * maxnbr;
        for (int sz = -1; sz <= 1; ++sz)
            for (int sy = -1; sy <= 1; ++sy)
                for (int sx = -1; sx <= 1; ++sx) {
                    int code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
                    float shx = (float)sx * prd[0], shy = (float)sy * prd[1], shz = (float)sz * prd[2];
                    for (int j = 0; j < n; ++j) {
                        if (j == i && code == 13) continue;
                        float xj = x[3 * j] + shx, yj = x[3 * j + 1] + shy, zj = x[3 * j + 2] + shz;
                        float dx = x[3 * i] - xj, dy = x[3 * i + 1] - yj, dz = x[3 * i + 2] - zj;
                        /* dx*dx + dy*dy + dz*dz as the reference's -O3 nvcc
                         * build contracts it (neigh_build_meso.cu:86-89):
                         * one rounded product, then two fused multiply-adds */
                        float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        if (r2 <= rc2) {
                            if (c < maxnbr) row[c] = j | (code << 27);
                            ++c;
                        }
                    }
                }
        counts[i] = c;
        qsort(row, (size_t)(c < maxnbr ? c : maxnbr), sizeof(int), cmp_int);
        if (c > maxc) maxc = c;
    }
    return maxc;
}

/* ======================================================================
 * FP64 pair enumeration: full list with periodic image codes
 * ====================================================================== */

typedef struct {
    int n, cap;
    int *start;      /* n+1 */
    int *j;          /* entries */
    unsigned char *code;
} nlist_t;

static void nlist_free(nlist_t *nl)
{
    free(nl->start); free(nl->j); free(nl->code);
    memset(nl, 0, sizeof(*nl));
}

static double wrap1(double c, double lo, double L)
{
    double t = c - lo;
    t -= floor(t / L) * L;
    if (t >= L) t -= L;
    if (t < 0) t = 0;
    return lo + t;
}

/* full neighbor list of all (j, image) within rcut of i, built from wrapped
 * copies of x; image code (sx+1)+3(sy+1)+9(sz+1) relative to the wrapped
 * coordinates, converted to a shift of the *given* coordinates */
static void nlist_build(int n, const double *x, const double lo[3], const double hi[3],
                        double rcut, nlist_t *nl, int *wrapk /* 3n: wrap count */)
{
    double L[3];
    int nc[3];
    int brute = 0;
    for (int d = 0; d < 3; ++d) {
        L[d] = hi[d] - lo[d];
        nc[d] = (int)floor(L[d] / rcut);
        if (nc[d] < 3) brute = 1;
    }
    double *xw = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double w = wrap1(x[3 * i + d], lo[d], L[d]);
            xw[3 * i + d] = w;
            wrapk[3 * i + d] = (int)llround((x[3 * i + d] - w) / L[d]);
        }
    const double rc2 = rcut * rcut;
    int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
    int ncell = brute ? 1 : nc[0] * nc[1] * nc[2];
    int *head = (int *)malloc(sizeof(int) * (size_t)ncell);
    int *next = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *cell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int c = 0; c < ncell; ++c) head[c] = -1;
    if (!brute) {
        for (int i = n - 1; i >= 0; --i) {
            int ci[3];
            for (int d = 0; d < 3; ++d) {
                int k = (int)((xw[3 * i + d] - lo[d]) / L[d] * nc[d]);
                ci[d] = k < 0 ? 0 : (k >= nc[d] ? nc[d] - 1 : k);
            }
            int c = ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]);
            cell[i] = c;
            next[i] = head[c];
            head[c] = i;
        }
    }
    /* two passes: count, then fill */
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            nl->start = (int *)malloc(sizeof(int) * ((size_t)n + 1));
            nl->start[0] = 0;
            for (int i = 0; i < n; ++i) nl->start[i + 1] = nl->start[i] + cnt[i];
            size_t tot = (size_t)nl->start[n];
            nl->j = (int *)malloc(sizeof(int) * (tot > 0 ? tot : 1));
            nl->code = (unsigned char *)malloc(tot > 0 ? tot : 1);
            nl->n = n;
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            int k = 0;
            int base = pass ? nl->start[i] : 0;
            const double *xi = xw + 3 * i;
            if (brute) {
                for (int code = 0; code < 27; ++code) {
                    int sx = code % 3 - 1, sy = (code / 3) % 3 - 1, sz = code / 9 - 1;
                    for (int j = 0; j < n; ++j) {
                        if (j == i && code == 13) continue;
                        double dx = xi[0] - (xw[3 * j] + sx * L[0]);
                        double dy = xi[1] - (xw[3 * j + 1] + sy * L[1]);
                        double dz = xi[2] - (xw[3 * j + 2] + sz * L[2]);
                        if (dx * dx + dy * dy + dz * dz < rc2) {
                            if (pass) { nl->j[base + k] = j; nl->code[base + k] = (unsigned char)code; }
                            ++k;
                        }
                    }
                }
            } else {
                int c = cell[i];
                int ci[3] = {c % nc[0], (c / nc[0]) % nc[1], c / (nc[0] * nc[1])};
                for (int oz = -1; oz <= 1; ++oz)
                    for (int oy = -1; oy <= 1; ++oy)
                        for (int ox = -1; ox <= 1; ++ox) {
                            int cc[3] = {ci[0] + ox, ci[1] + oy, ci[2] + oz};
                            int s[3];
                            for (int d = 0; d < 3; ++d) {
                                s[d] = 0;
                                if (cc[d] < 0) { cc[d] += nc[d]; s[d] = -1; }
                                else if (cc[d] >= nc[d]) { cc[d] -= nc[d]; s[d] = 1; }
                            }
                            int code = (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
                            int cid = cc[0] + nc[0] * (cc[1] + nc[1] * cc[2]);
                            for (int j = head[cid]; j >= 0; j = next[j]) {
                                if (j == i && code == 13) continue;
                                double dx = xi[0] - (xw[3 * j] + s[0] * L[0]);
                                double dy = xi[1] - (xw[3 * j + 1] + s[1] * L[1]);
                                double dz = xi[2] - (xw[3 * j + 2] + s[2] * L[2]);
                                if (dx * dx + dy * dy + dz * dz < rc2) {
                                    if (pass) { nl->j[base + k] = j; nl->code[base + k] = (unsigned char)code; }
                                    ++k;
                                }
                            }
                        }
            }
            if (!pass) cnt[i] = k;
        }
    }
    free(cnt); free(head); free(next); free(cell); free(xw);
}

/* displacement of entry (i, j, code) in the caller's coordinates */
static inline void pair_dx(const double *x, const int *wrapk, const double L[3], int i, int j,
                           int code, double d[3])
{
    int s[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
    for (int a = 0; a < 3; ++a) {
        /* wrapped_i - (wrapped_j + s L), expressed with the given coordinates */
        double xi = x[3 * i + a] - wrapk[3 * i + a] * L[a];
        double xj = x[3 * j + a] - wrapk[3 * j + a] * L[a];
        d[a] = xi - (xj + s[a] * L[a]);
    }
}

static double max_cut(const orc_params *p)
{
    double m = 0;
    for (int a = 0; a < p->ntypes; ++a)
        for (int b = 0; b < p->ntypes; ++b)
            if (p->cut[a][b] > m) m = p->cut[a][b];
    return m;
}

/* ======================================================================
 * Density, forces (FP64 arithmetic on the given coordinates)
 * ====================================================================== */

static void density_core(int n, const double *x, const int *wrapk, const nlist_t *nl,
                         const int *type, const int *mask, const orc_params *p, double rd,
                         int lucy, int groupbit, double *rho)
{
    double L[3] = {p->hi[0] - p->lo[0], p->hi[1] - p->lo[1], p->hi[2] - p->lo[2]};
    /* fix_aboundary_rho_meso.cu:88-90 (quadratic) and :89 (Lucy) */
    double fq = 15.0 / (2.0 * M_PI * rd * rd * rd * rd * rd);
    double fl = 105.0 / (16.0 * M_PI * rd * rd * rd * rd * rd * rd * rd);
    const double h2 = (double)((float)rd * (float)rd);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        if (!(mask[i] & groupbit)) { rho[i] = 0.0; continue; }
        double s = 0.0;
        for (int e = nl->start[i]; e < nl->start[i + 1]; ++e) {
            double d[3];
            pair_dx(x, wrapk, L, i, nl->j[e], nl->code[e], d);
            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (rsq < h2 && rsq >= 1e-20) {                    /* :155 */
                double r = sqrt(rsq);
                double w = lucy ? fl * (rd + 3.0 * r) * (rd - r) * (rd - r) * (rd - r)
                                : fq * (rd - r) * (rd - r);    /* :157-158 */
                s += w;
            }
        }
        rho[i] = p->mass[type[i]] * s;                           /* :159 m_i */
    }
}

static void force_core(int style, int n, const double *x, const double *v, const int *wrapk,
                       const nlist_t *nl, const double *rho, const uint32_t *sig,
                       const int *type, const orc_params *p, double *f, double *e, double *vir,
                       const int *mask, const double *phi, const orc_arb *arb)
{
    const int wall_scaling = style == 0 && arb && arb->pair_on && mask && phi;
    double L[3] = {p->hi[0] - p->lo[0], p->hi[1] - p->lo[1], p->hi[2] - p->lo[2]};
    const double dtinvsqrt = 1.0 / sqrt(p->dt);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double fx = 0, fy = 0, fz = 0, en = 0;
        double vxx = 0, vyy = 0, vzz = 0, vxy = 0, vxz = 0, vyz = 0;
        int ti = type[i];
        for (int k = nl->start[i]; k < nl->start[i + 1]; ++k) {
            int j = nl->j[k], tj = type[j];
            double d[3];
            pair_dx(x, wrapk, L, i, j, nl->code[k], d);
            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double cut = p->cut[ti][tj];
            if (!(rsq < cut * cut && rsq >= 1e-20)) continue;
            double rinv = 1.0 / sqrt(rsq);
            double r = rsq * rinv;
            if (style == 0 && r < 1e-10) continue;
            double dvx = v[3 * i] - v[3 * j], dvy = v[3 * i + 1] - v[3 * j + 1], dvz = v[3 * i + 2] - v[3 * j + 2];
            double dot = d[0] * dvx + d[1] * dvy + d[2] * dvz;
            double rn = orc_gaussian_poly(sig[i], sig[j]);
            double fpair;
            if (style == 0) {
                /* pair_mdpd_meso.cu:156-191 */
                double wc = 1.0 - r / cut;
                double wcr = 1.0 - r / p->cut_r[ti][tj];
                if (wcr < 0) wcr = 0;
                double wr = wc;
                double rr = rho[i] + rho[j];
                double gam = p->gamma[ti][tj], sgm = p->sigma[ti][tj];
                if (wall_scaling && (mask[i] & arb->pair_mobile_bit) && (mask[j] & arb->pair_wall_bit)) {
                    /* pair_mdpd_meso.cu:169-185: phi of the mobile particle i */
                    double ratio = orc_wall_ratio(phi[i], cut);
                    sgm *= sqrt(ratio);
                    gam *= ratio;
                }
                fpair = (p->A[ti][tj] * wc + p->B[ti][tj] * rr * wcr)
                        - (gam * wr * wr * dot * rinv)
                        + (sgm * wr * rn * dtinvsqrt);
                fpair *= rinv;
                en += 0.5 * p->A[ti][tj] * cut * wr * wr + 0.5 * p->B[ti][tj] * p->cut_r[ti][tj] * rr * wcr * wcr;
            } else {
                /* pair_dpd_meso.cu:145-175 */
                double wc = 1.0 - r * (1.0 / cut);
                double wr = pow(wc, p->expw[ti][tj]);
                fpair = p->A[ti][tj] * wc - (p->gamma[ti][tj] * wr * wr * dot * rinv)
                        + (p->sigma[ti][tj] * wr * rn * dtinvsqrt);
                fpair *= rinv;
                en += 0.5 * p->A[ti][tj] * cut * wc * wc;
            }
            fx += d[0] * fpair; fy += d[1] * fpair; fz += d[2] * fpair;
            vxx += d[0] * d[0] * fpair; vyy += d[1] * d[1] * fpair; vzz += d[2] * d[2] * fpair;
            vxy += d[0] * d[1] * fpair; vxz += d[0] * d[2] * fpair; vyz += d[1] * d[2] * fpair;
        }
        f[3 * i] = fx; f[3 * i + 1] = fy; f[3 * i + 2] = fz;
        if (e) e[i] = en * 0.5;
        if (vir) {
            vir[6 * i + 0] = 0.5 * vxx; vir[6 * i + 1] = 0.5 * vyy; vir[6 * i + 2] = 0.5 * vzz;
            vir[6 * i + 3] = 0.5 * vxy; vir[6 * i + 4] = 0.5 * vxz; vir[6 * i + 5] = 0.5 * vyz;
        }
    }
}

static double *f2d(int n3, const float *a)
{
    double *d = (double *)malloc(sizeof(double) * (size_t)(n3 > 0 ? n3 : 1));
    for (int i = 0; i < n3; ++i) d[i] = a[i];
    return d;
}

void orc_density(int n, const float *x, const int *type, const int *mask,
                 const orc_params *p, double rd, int lucy, int groupbit, double *rho)
{
    double *xd = f2d(3 * n, x);
    int *wk = (int *)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
    nlist_t nl = {0};
    nlist_build(n, xd, p->lo, p->hi, rd, &nl, wk);
    density_core(n, xd, wk, &nl, type, mask, p, rd, lucy, groupbit, rho);
    nlist_free(&nl);
    free(wk); free(xd);
}

static void force_entry(int style, int n, const float *x, const float *v, const double *rho,
                        const uint32_t *sig, const int *type, const orc_params *p,
                        double *f, double *e, double *vir,
                        const int *mask, const double *phi, const orc_arb *arb)
{
    double *xd = f2d(3 * n, x), *vd = f2d(3 * n, v);
    int *wk = (int *)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
    nlist_t nl = {0};
    nlist_build(n, xd, p->lo, p->hi, max_cut(p), &nl, wk);
    force_core(style, n, xd, vd, wk, &nl, rho, sig, type, p, f, e, vir, mask, phi, arb);
    nlist_free(&nl);
    free(wk); free(xd); free(vd);
}

void orc_force_mdpd(int n, const float *x, const float *v, const double *rho,
                    const uint32_t *sig, const int *type, const orc_params *p,
                    double *f, double *e, double *vir)
{
    force_entry(0, n, x, v, rho, sig, type, p, f, e, vir, NULL, NULL, NULL);
}

void orc_force_mdpd_arb(int n, const float *x, const float *v, const double *rho,
                        const uint32_t *sig, const int *type, const int *mask,
                        const double *phi, const orc_params *p, const orc_arb *arb,
                        double *f, double *e, double *vir)
{
    force_entry(0, n, x, v, rho, sig, type, p, f, e, vir, mask, phi, arb);
}

void orc_force_dpd(int n, const float *x, const float *v, const uint32_t *sig,
                   const int *type, const orc_params *p, double *f, double *e, double *vir)
{
    force_entry(1, n, x, v, NULL, sig, type, p, f, e, vir, NULL, NULL, NULL);
}

/* ======================================================================
 * Arbitrary boundary          fix_aboundary_rho_meso.cu:93-94,220-287,
 *                             pair_mdpd_meso.cu:169-185
 * ====================================================================== */

double orc_wall_ratio(double phi, double rcw)
{
    double ph = phi < 0.5 ? phi : 0.5;                          /* MIN(phi_i, 0.5) */
    double h = 1.0 - pow(2.088 * ph * ph * ph + 1.478 * ph, 0.25);
    if (h < 0.025) h = 0.025;
    h *= rcw;
    double q = 1.0 - h / rcw;
    return 1.0 + 0.187 * (rcw / h - 1.0) - 0.093 * q * q * q;
}

static void aboundary_core(int n, double *x, double *v, const int *wrapk, const nlist_t *nl,
                           const int *type, const int *mask, const orc_params *p, const orc_arb *arb,
                           double *phi, double *nw_out)
{
    double L[3] = {p->hi[0] - p->lo[0], p->hi[1] - p->lo[1], p->hi[2] - p->lo[2]};
    const double c = arb->cut_phi;
    const double c7 = c * c * c * c * c * c * c;
    const double phi_factor = 105.0 / (16.0 * M_PI * arb->rho_wall * c7);   /* :93 */
    const double dw_factor = -315.0 / (4.0 * M_PI * arb->rho_wall * c7);    /* :94 */
    double *nw = (double *)calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        phi[i] = 0.0;
        if (!(mask[i] & arb->fix_mobile_bit)) continue;
        double s = 0.0;
        for (int k = nl->start[i]; k < nl->start[i + 1]; ++k) {
            int j = nl->j[k];
            if (!(mask[j] & arb->fix_wall_bit)) continue;
            double d[3];
            pair_dx(x, wrapk, L, i, j, nl->code[k], d);
            double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(rsq < c * c)) continue;
            double rinv = 1.0 / sqrt(rsq);
            double r = rsq * rinv;
            if (r < 1e-10) r = 1e-10;                            /* MAX(r, EPSILON) */
            double dd = c - r;
            s += (c + 3.0 * r) * dd * dd * dd * phi_factor;      /* :260 */
            for (int a = 0; a < 3; ++a) nw[3 * i + a] -= (r * dd * dd * dw_factor) * d[a] * rinv;   /* :261 */
        }
        phi[i] = s;
    }
    /* bounce-back (:266-283), after every phi is known (walls do not move) */
    for (int i = 0; i < n; ++i) {
        if (!(mask[i] & arb->fix_mobile_bit) || !(phi[i] > arb->phi_c)) continue;
        double *m = nw + 3 * i;
        double len = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        double u[3] = {m[0] / len, m[1] / len, m[2] / len};
        double dtfm = 0.5 * p->dt / p->mass[type[i]];
        double vn = v[3 * i] * u[0] + v[3 * i + 1] * u[1] + v[3 * i + 2] * u[2];
        double tmp = vn > 0.0 ? vn : 0.0;
        for (int a = 0; a < 3; ++a) {
            double orig = x[3 * i + a] - v[3 * i + a] * dtfm;
            v[3 * i + a] = -v[3 * i + a] + 2.0 * tmp * u[a];
            x[3 * i + a] = orig + p->dt * v[3 * i + a];
        }
    }
    if (nw_out) memcpy(nw_out, nw, sizeof(double) * 3 * (size_t)n);
    free(nw);
}

void orc_aboundary(int n, double *x, double *v, const int *type, const int *mask,
                   const orc_params *p, const orc_arb *arb, double *phi, double *nw)
{
    int *wk = (int *)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
    nlist_t nl = {0};
    nlist_build(n, x, p->lo, p->hi, arb->cut_phi, &nl, wk);
    aboundary_core(n, x, v, wk, &nl, type, mask, p, arb, phi, nw);
    nlist_free(&nl);
    free(wk);
}

/* ======================================================================
 * Step loop (FP64 state): ModifiedVerlet::setup/run   mvv_meso.cu:203-572
 * ====================================================================== */

static void wrap_fp64(int n, double *x, const double lo[3], const double hi[3])
{
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double prd = hi[d] - lo[d];
            double c = x[3 * i + d];
            if (c < lo[d]) c += prd;
            if (c >= hi[d]) { c -= prd; if (c < lo[d]) c = lo[d]; }
            x[3 * i + d] = c;
        }
}

static void eval_forces(const orc_run_cfg *cfg, int n, const double *x, const double *v,
                        const int *wk, const nlist_t *nl, double *rho, uint32_t *sig,
                        const int *tag, const int *type, const int *mask, const orc_params *p,
                        double *f)
{
    uint32_t st = orc_seed_now(cfg->seed, cfg->ntimestep);
    if (cfg->style == 0) {
        density_core(n, x, wk, nl, type, mask, p, cfg->cut_rho, cfg->lucy, cfg->rho_groupbit, rho);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) sig[i] = orc_sig_mdpd(tag[i], st);
    } else {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            sig[i] = orc_sig_dpd(tag[i], st, (float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2]);
    }
    force_core(cfg->style, n, x, v, wk, nl, rho, sig, type, p, f, NULL, NULL, mask, cfg->phi, cfg->arb);
    if (cfg->addforce_groupbit) {
        for (int i = 0; i < n; ++i)
            if (mask[i] & cfg->addforce_groupbit)
                for (int d = 0; d < 3; ++d) f[3 * i + d] += cfg->addforce[d];   /* fix_addforce_meso.cu:88-92 */
    }
}

int orc_run(int n, double *x, double *v, double *f, double *rho, const int *tag,
            const int *type, const int *mask, const orc_params *p, orc_run_cfg *cfg, int nsteps)
{
#ifdef _OPENMP
    if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
    size_t n3 = 3 * (size_t)(n > 0 ? n : 1);
    double *xhold = (double *)malloc(sizeof(double) * n3);
    int *wk = (int *)malloc(sizeof(int) * n3);
    uint32_t *sig = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    double rcut = max_cut(p) + cfg->skin;
    nlist_t nl = {0};
    int nbuild = 0;

    /* setup: pbc, build, forces at ntimestep (mvv_meso.cu:203-283) */
    wrap_fp64(n, x, p->lo, p->hi);
    nlist_build(n, x, p->lo, p->hi, rcut, &nl, wk);
    memcpy(xhold, x, sizeof(double) * 3 * (size_t)n);
    ++nbuild;
    eval_forces(cfg, n, x, v, wk, &nl, rho, sig, tag, type, mask, p, f);

    const double dtf = 0.5 * p->dt, trig = 0.25 * cfg->skin * cfg->skin;
    for (int step = 0; step < nsteps; ++step) {
        cfg->ntimestep += 1;
        /* initial_integrate: fix_nve_meso.cu:83-92 */
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            if (!(mask[i] & cfg->nve_groupbit)) continue;
            double dtfm = dtf / p->mass[type[i]];
            for (int d = 0; d < 3; ++d) {
                v[3 * i + d] += dtfm * f[3 * i + d];
                x[3 * i + d] += p->dt * v[3 * i + d];
            }
        }
        /* post_integrate: wall fraction and bounce-back with the list of the
         * last build (fix_aboundary_rho_meso.cu:290-317) */
        if (cfg->arb && cfg->arb->fix_on && cfg->phi)
            aboundary_core(n, x, v, wk, &nl, type, mask, p, cfg->arb, cfg->phi, NULL);
        /* decide(): every 1 delay 0 check yes (src/neighbor.cpp:1216-1290) */
        int flag = 0;
        for (int i = 0; i < n && !flag; ++i) {
            double dx = x[3 * i] - xhold[3 * i], dy = x[3 * i + 1] - xhold[3 * i + 1], dz = x[3 * i + 2] - xhold[3 * i + 2];
            if (dx * dx + dy * dy + dz * dz > trig) flag = 1;
        }
        if (flag) {
            nlist_free(&nl);
            wrap_fp64(n, x, p->lo, p->hi);
            nlist_build(n, x, p->lo, p->hi, rcut, &nl, wk);
            memcpy(xhold, x, sizeof(double) * 3 * (size_t)n);
            ++nbuild;
        }
        eval_forces(cfg, n, x, v, wk, &nl, rho, sig, tag, type, mask, p, f);
        /* final_integrate: fix_nve_meso.cu:168-177 */
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            if (!(mask[i] & cfg->nve_groupbit)) continue;
            double dtfm = dtf / p->mass[type[i]];
            for (int d = 0; d < 3; ++d) v[3 * i + d] += dtfm * f[3 * i + d];
        }
    }
    nlist_free(&nl);
    free(xhold); free(wk); free(sig);
    return nbuild;
}

/* ======================================================================
 * LAMMPS pair_style dpd on the CPU                 src/pair_dpd.cpp:63-160
 * ====================================================================== */

void orc_ranmars_init(orc_ranmars *r, int seed)
{
    /* src/random_mars.cpp:25-62 */
    int ij = (seed - 1) / 30082;
    int kl = (seed - 1) - 30082 * ij;
    int i = (ij / 177) % 177 + 2;
    int j = ij % 177 + 2;
    int k = (kl / 169) % 178 + 1;
    int l = kl % 169;
    for (int ii = 1; ii <= 97; ii++) {
        double s = 0.0, t = 0.5;
        for (int jj = 1; jj <= 24; jj++) {
            int m = ((i * j) % 179) * k % 179;
            i = j; j = k; k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s = s + t;
            t = 0.5 * t;
        }
        r->u[ii] = s;
    }
    r->c = 362436.0 / 16777216.0;
    r->cd = 7654321.0 / 16777216.0;
    r->cm = 16777213.0 / 16777216.0;
    r->i97 = 97;
    r->j97 = 33;
    r->save = 0;
    r->second = 0;
    orc_ranmars_uniform(r);
}

double orc_ranmars_uniform(orc_ranmars *r)
{
    /* src/random_mars.cpp:75-90 */
    double uni = r->u[r->i97] - r->u[r->j97];
    if (uni < 0.0) uni += 1.0;
    r->u[r->i97] = uni;
    if (--r->i97 == 0) r->i97 = 97;
    if (--r->j97 == 0) r->j97 = 97;
    r->c -= r->cd;
    if (r->c < 0.0) r->c += r->cm;
    uni -= r->c;
    if (uni < 0.0) uni += 1.0;
    return uni;
}

double orc_ranmars_gaussian(orc_ranmars *r)
{
    /* polar Box-Muller, pairs cached: src/random_mars.cpp:96-116 */
    double first;
    if (!r->save) {
        double v1, v2, rsq;
        do {
            v1 = 2.0 * orc_ranmars_uniform(r) - 1.0;
            v2 = 2.0 * orc_ranmars_uniform(r) - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while (!(rsq < 1.0 && rsq != 0.0));
        double fac = sqrt(-2.0 * log(rsq) / rsq);
        r->second = v1 * fac;
        first = v2 * fac;
        r->save = 1;
    } else {
        first = r->second;
        r->save = 0;
    }
    return first;
}

/* Pair forces of the half list (newton on: f_j -= F), pair_dpd.cpp:63-160.
 * Thread t of nt accumulates into its own force array with its own Marsaglia
 * stream; the arrays are summed afterwards.  With nt = 1 this is the serial
 * loop with the one stream of PairDPD::settings. */
typedef struct {
    int n, nt, nrdf;
    const double *L;
    double a0, gamma, sigma, rc, rc2, dtinvsqrt;
    orc_ranmars *rng;     /* nt streams */
    double *fthr;         /* nt * 3n (thread 0 is the output f) */
    double *virthr;       /* nt */
    double *rdfthr;       /* nt * nrdf, or NULL */
} dpd_half_ctx;

static void dpd_half_forces(dpd_half_ctx *c, const double *x, const double *v, const int *wk, const nlist_t *nl,
                            int sample)
{
    const size_t n3 = 3 * (size_t)c->n;
#pragma omp parallel num_threads(c->nt)
    {
        const int t = c->nt > 1 ? omp_get_thread_num() : 0;
        double *f = c->fthr + (size_t)t * n3;
        orc_ranmars *rng = &c->rng[t];
        double *rdf = c->rdfthr ? c->rdfthr + (size_t)t * c->nrdf : NULL;
        double virial = 0;
        memset(f, 0, sizeof(double) * n3);
#pragma omp for schedule(static)
        for (int i = 0; i < c->n; ++i) {
            for (int k = nl->start[i]; k < nl->start[i + 1]; ++k) {
                int j = nl->j[k];
                if (j <= i) continue;
                double d[3];
                pair_dx(x, wk, c->L, i, j, nl->code[k], d);
                double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (rsq >= c->rc2) continue;
                double r = sqrt(rsq);
                if (r < 1e-10) continue;
                double rinv = 1.0 / r;
                double dvx = v[3 * i] - v[3 * j];
                double dvy = v[3 * i + 1] - v[3 * j + 1];
                double dvz = v[3 * i + 2] - v[3 * j + 2];
                double dot = d[0] * dvx + d[1] * dvy + d[2] * dvz;
                double wd = 1.0 - r / c->rc;
                double rn = orc_ranmars_gaussian(rng);
                double fp = c->a0 * wd;
                fp -= c->gamma * wd * wd * dot * rinv;
                fp += c->sigma * wd * rn * c->dtinvsqrt;
                fp *= rinv;
                for (int a = 0; a < 3; ++a) {
                    f[3 * i + a] += d[a] * fp;
                    f[3 * j + a] -= d[a] * fp;
                }
                if (sample) virial += rsq * fp;
                if (sample && rdf) {
                    int b = (int)(r / c->rc * c->nrdf);
                    if (b < c->nrdf) rdf[b] += 1.0;
                }
            }
        }
        c->virthr[t] = virial;
        if (c->nt > 1) {
            /* sum the thread arrays into thread 0's (the implicit barrier of
             * the loop above has passed) */
#pragma omp for schedule(static)
            for (size_t a = 0; a < n3; ++a) {
                double s = c->fthr[a];
                for (int q = 1; q < c->nt; ++q) s += c->fthr[(size_t)q * n3 + a];
                c->fthr[a] = s;
            }
        }
    }
}

int orc_lammps_dpd_run(int n, double *x, double *v, const double lo[3], const double hi[3],
                       double a0, double gamma, double temperature, double rc,
                       double skin, double dt, int seed, int nsteps, int equil,
                       int sample_every, double *t_sum, double *p_sum, int *nsamples,
                       int nrdf, double *rdf_hist, int nthreads)
{
    /* PairDPD::settings: RanMars(seed + me) with me = 0 (pair_dpd.cpp:198);
     * with nthreads > 1 thread t draws from RanMars(seed + t), as LAMMPS's
     * OpenMP variant of the pair style seeds its per-thread streams;
     * init_one: sigma = sqrt(2 kT gamma) (pair_dpd.cpp:262) */
    const int nt = nthreads > 1 ? nthreads : 1;
    const int prev_threads = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);   /* the neighbor build too */
    dpd_half_ctx c;
    c.n = n; c.nt = nt; c.nrdf = nrdf;
    c.a0 = a0; c.gamma = gamma; c.sigma = sqrt(2.0 * temperature * gamma); c.rc = rc; c.rc2 = rc * rc;
    c.dtinvsqrt = 1.0 / sqrt(dt);
    double L[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    c.L = L;
    c.rng = (orc_ranmars *)malloc(sizeof(orc_ranmars) * (size_t)nt);
    for (int t = 0; t < nt; ++t) orc_ranmars_init(&c.rng[t], seed + t);
    size_t n3 = 3 * (size_t)n;
    c.fthr = (double *)calloc(n3 * (size_t)nt + 1, sizeof(double));
    c.virthr = (double *)calloc((size_t)nt, sizeof(double));
    c.rdfthr = (rdf_hist && nrdf > 0) ? (double *)calloc((size_t)nt * (size_t)nrdf, sizeof(double)) : NULL;
    double *f = c.fthr;
    const double dtf = 0.5 * dt, trig = 0.25 * skin * skin;
    double V = L[0] * L[1] * L[2];
    double *xhold = (double *)malloc(sizeof(double) * (n3 + 1));
    int *wk = (int *)malloc(sizeof(int) * (n3 + 1));
    nlist_t nl = {0};
    int nbuild = 0;

    /* the full list of nlist_build, used as a half list: entry (i, j) with
     * j > i; its partner (j, i) is skipped */
#define BUILD_HALF()                                                           \
    do {                                                                       \
        nlist_free(&nl);                                                       \
        wrap_fp64(n, x, lo, hi);                                               \
        nlist_build(n, x, lo, hi, rc + skin, &nl, wk);                         \
        memcpy(xhold, x, sizeof(double) * n3);                                 \
        ++nbuild;                                                              \
    } while (0)

    BUILD_HALF();
    dpd_half_forces(&c, x, v, wk, &nl, 0);
    *t_sum = 0; *p_sum = 0; *nsamples = 0;
    for (int step = 1; step <= nsteps; ++step) {
#pragma omp parallel for schedule(static) num_threads(nt)
        for (size_t a = 0; a < n3; ++a) {              /* fix_nve.cpp:64-95 */
            v[a] += dtf * f[a];
            x[a] += dt * v[a];
        }
        int flag = 0;
#pragma omp parallel for schedule(static) num_threads(nt) reduction(| : flag)
        for (int i = 0; i < n; ++i) {
            double dx = x[3 * i] - xhold[3 * i], dy = x[3 * i + 1] - xhold[3 * i + 1], dz = x[3 * i + 2] - xhold[3 * i + 2];
            if (dx * dx + dy * dy + dz * dz > trig) flag |= 1;
        }
        if (flag) BUILD_HALF();
        int sample = (step > equil) && sample_every > 0 && ((step - equil) % sample_every == 0);
        dpd_half_forces(&c, x, v, wk, &nl, sample);
#pragma omp parallel for schedule(static) num_threads(nt)
        for (size_t a = 0; a < n3; ++a) v[a] += dtf * f[a];   /* fix_nve.cpp:100-130 */
        if (sample) {
            double ke = 0, virial = 0;
            for (size_t a = 0; a < n3; ++a) ke += v[a] * v[a];
            for (int t = 0; t < nt; ++t) virial += c.virthr[t];
            double dof = 3.0 * n - 3.0;
            double T = ke / dof;
            *t_sum += T;
            *p_sum += (dof * T + virial) / (3.0 * V);
            *nsamples += 1;
        }
    }
#undef BUILD_HALF
    if (c.rdfthr)
        for (int t = 0; t < nt; ++t)
            for (int b = 0; b < nrdf; ++b) rdf_hist[b] += c.rdfthr[(size_t)t * nrdf + b];
    nlist_free(&nl);
    free(c.fthr); free(c.virthr); free(c.rdfthr); free(c.rng);
    free(xhold); free(wk);
    omp_set_num_threads(prev_threads);
    return nbuild;
}

/* ======================================================================
 * Observables
 * ====================================================================== */

double orc_temperature(int n, const double *v, const int *type, const orc_params *p)
{
    /* compute_temp_meso.cu:48-101: sum m v^2 / (dof kB), dof = 3N - 3 */
    double s = 0;
    for (int i = 0; i < n; ++i)
        s += p->mass[type[i]] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    double dof = 3.0 * n - 3.0;
    return dof > 0 ? s / dof : 0.0;
}

double orc_pressure(int n, const double *v, const double *vir, const int *type, const orc_params *p)
{
    /* compute_pressure_meso.cu:140-152 */
    double t = orc_temperature(n, v, type, p);
    double dof = 3.0 * n - 3.0, vs = 0;
    for (int i = 0; i < n; ++i) vs += vir[6 * i] + vir[6 * i + 1] + vir[6 * i + 2];
    double V = (p->hi[0] - p->lo[0]) * (p->hi[1] - p->lo[1]) * (p->hi[2] - p->lo[2]);
    return (dof * t + vs) / 3.0 / V;
}

/* ======================================================================
 * Array helpers for the Python test harness
 * ====================================================================== */

void orc_premix_n(int n, const uint32_t *a, const uint32_t *b, int rounds, uint32_t *out)
{
    for (int i = 0; i < n; ++i) out[i] = orc_premix(a[i], b[i], rounds);
}

void orc_gaussian_n(int n, const uint32_t *a, const uint32_t *b, int libm, double *out)
{
    for (int i = 0; i < n; ++i) out[i] = libm ? orc_gaussian_libm(a[i], b[i]) : orc_gaussian_poly(a[i], b[i]);
}

void orc_sig_mdpd_n(int n, const int32_t *tag, uint32_t seed_t, uint32_t *out)
{
    for (int i = 0; i < n; ++i) out[i] = orc_sig_mdpd(tag[i], seed_t);
}

void orc_sig_dpd_n(int n, const int32_t *tag, uint32_t seed_t, const float *v, uint32_t *out)
{
    for (int i = 0; i < n; ++i) out[i] = orc_sig_dpd(tag[i], seed_t, v[3 * i], v[3 * i + 1], v[3 * i + 2]);
}

void orc_sort_keys_n(int n, const float *x, const double lo[3], const double hi[3], double cutneigh,
                     uint64_t *key)
{
    orc_bins b;
    orc_setup_bins(lo, hi, cutneigh, &b);
    for (int i = 0; i < n; ++i) key[i] = orc_sort_key(&b, x[3 * i], x[3 * i + 1], x[3 * i + 2]);
}
