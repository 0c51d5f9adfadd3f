/*
 * effm_oracle.c — FP64 CPU restatement of the effmass H|psi> hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle used by tests/, smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product links it.
 *
 * Parity pins (each exercised by tests/test_oracle_*.py):
 *   orc_fft1d/3d   closed-form KATs of test/math/test_fft.cpp:21-126 and
 *                  numpy.fft (pocketfft, the reference's FFT library)
 *   orc_interp     exact at nodes; closed-form beta(q) between nodes;
 *                  q==0 / q<dq / q>qmax rules of ncpp_advance.cppm:656-673
 *   orc_ylm        Y00, Y10, Y1+-1 KATs of test_sph_harmonics.cpp:20-60,
 *                  theta=0 values (:135-139), scipy sph_harm, finite
 *                  differences for both angular gradients (:202-221,:379)
 *   orc_sfac       |S|=1, tau=0 -> 1, direct exp (test_structure_factor.cpp)
 *   orc_hpsi       dense H_GG' brute force (kinetic + V(G-G') convolution +
 *                  sum_ij |p_i> lambda_i <p_i|) built in numpy from the
 *                  mathematics; Hermiticity; V=const shift
 *   orc_dhdk       central finite differences of <psi|H(k+-d)|psi>
 */
#include "effm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static const double PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------- */
/*  1D FFT: recursive mixed-radix decimation in time.                         */
/*  The reference calls pocketfft cfft_forward(plan, buf, 1/n) and            */
/*  cfft_backward(plan, buf, 1) (src/math/fft.cppm:36-41).                    */
/* ------------------------------------------------------------------------- */

typedef struct {
    int n;
    int nf;
    int f[64];
    cplx* w;        /* w[k] = exp(-2 pi i k / n) */
} fft_plan;

static int smallest_factor(int n) {
    if (n % 4 == 0 && n > 4) return 4;
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

static void plan_init(fft_plan* pl, int n) {
    pl->n = n;
    pl->nf = 0;
    int m = n;
    while (m > 1) {
        int p = smallest_factor(m);
        pl->f[pl->nf++] = p;
        m /= p;
    }
    pl->w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double a = -2.0 * PI * (double)k / (double)n;
        pl->w[k] = cos(a) + I * sin(a);
    }
}

static void plan_free(fft_plan* pl) { free(pl->w); pl->w = NULL; }

/* out[0..n) = DFT_n of in[0], in[s], in[2s], ... (sign via conj twiddles). */
static void fft_rec(const cplx* in, int s, cplx* out, int n, int fi,
                    const fft_plan* pl, int sign, cplx* tmp) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = pl->f[fi];
    int m = n / p;
    for (int u = 0; u < p; ++u)
        fft_rec(in + (size_t)u * s, s * p, out + (size_t)u * m, m, fi + 1, pl, sign, tmp);
    int wstep = pl->n / n;      /* W_n = W_N^(N/n) */
    int pstep = pl->n / p;      /* W_p = W_N^(N/p) */
    for (int k1 = 0; k1 < m; ++k1) {
        for (int u = 0; u < p; ++u) {
            cplx w = pl->w[(size_t)u * k1 * wstep];
            if (sign > 0) w = conj(w);
            tmp[u] = out[(size_t)u * m + k1] * w;
        }
        for (int q = 0; q < p; ++q) {
            cplx acc = 0.0;
            for (int u = 0; u < p; ++u) {
                cplx w = pl->w[(size_t)((u * q) % p) * pstep];
                if (sign > 0) w = conj(w);
                acc += tmp[u] * w;
            }
            tmp[p + q] = acc;
        }
        for (int q = 0; q < p; ++q) out[(size_t)q * m + k1] = tmp[p + q];
    }
}

static void fft1d_plan(const fft_plan* pl, cplx* data, int forward, cplx* work) {
    int n = pl->n;
    if (n <= 1) return;
    cplx* out = work;
    cplx* tmp = work + n;
    fft_rec(data, 1, out, n, 0, pl, forward ? -1 : +1, tmp);
    if (forward) {
        double f = 1.0 / (double)n;
        for (int i = 0; i < n; ++i) data[i] = out[i] * f;
    } else {
        memcpy(data, out, sizeof(cplx) * (size_t)n);
    }
}

void orc_fft1d(double* data, int n, int forward) {
    if (n <= 1) return;
    fft_plan pl;
    plan_init(&pl, n);
    cplx* work = (cplx*)malloc(sizeof(cplx) * (size_t)(3 * n + 16));
    fft1d_plan(&pl, (cplx*)data, forward, work);
    free(work);
    plan_free(&pl);
}

/* FFT3D::operator() (src/math/fft.cppm:60-85): k axis (contiguous), then j,
 * then i, each line gathered through a scratch buffer. */
typedef struct {
    fft_plan p1, p2, p3;
    cplx* scratch;
    cplx* work;
} fft3_ctx;

static void fft3_init(fft3_ctx* c, int n1, int n2, int n3) {
    plan_init(&c->p1, n1);
    plan_init(&c->p2, n2);
    plan_init(&c->p3, n3);
    int nm = n1 > n2 ? n1 : n2;
    nm = nm > n3 ? nm : n3;
    c->scratch = (cplx*)malloc(sizeof(cplx) * (size_t)nm);
    c->work = (cplx*)malloc(sizeof(cplx) * (size_t)(3 * nm + 16));
}

static void fft3_free(fft3_ctx* c) {
    plan_free(&c->p1); plan_free(&c->p2); plan_free(&c->p3);
    free(c->scratch); free(c->work);
}

static void fft3_run(fft3_ctx* c, cplx* g, int n1, int n2, int n3, int forward) {
    size_t s23 = (size_t)n2 * n3, s3 = (size_t)n3;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            fft1d_plan(&c->p3, g + ((size_t)i * n2 + j) * n3, forward, c->work);
    for (int i = 0; i < n1; ++i)
        for (int k = 0; k < n3; ++k) {
            for (int j = 0; j < n2; ++j) c->scratch[j] = g[i * s23 + j * s3 + k];
            fft1d_plan(&c->p2, c->scratch, forward, c->work);
            for (int j = 0; j < n2; ++j) g[i * s23 + j * s3 + k] = c->scratch[j];
        }
    for (int j = 0; j < n2; ++j)
        for (int k = 0; k < n3; ++k) {
            for (int i = 0; i < n1; ++i) c->scratch[i] = g[i * s23 + j * s3 + k];
            fft1d_plan(&c->p1, c->scratch, forward, c->work);
            for (int i = 0; i < n1; ++i) g[i * s23 + j * s3 + k] = c->scratch[i];
        }
}

void orc_fft3d(double* grid, int n1, int n2, int n3, int forward) {
    fft3_ctx c;
    fft3_init(&c, n1, n2, n3);
    fft3_run(&c, (cplx*)grid, n1, n2, n3, forward);
    fft3_free(&c);
}

/* ------------------------------------------------------------------------- */
/*  beta(q) interpolation: lagrangeCubic (src/math/numerical.cppm:93-104)    */
/*  with the stencil rule of BetaqTables::interpolate                         */
/*  (src/pseudo/ncpp_advance.cppm:644-674).                                   */
/* ------------------------------------------------------------------------- */

static double lagrange_cubic(double f0, double f1, double f2, double f3, double px) {
    double ux = 1.0 - px;
    double vx = 2.0 - px;
    double wx = 3.0 - px;
    return f0 * ux * vx * wx / 6.0 +
           f1 * px * vx * wx / 2.0 -
           f2 * px * ux * wx / 2.0 +
           f3 * px * ux * vx / 6.0;
}

int orc_interp(const double* t, int nq, double dq, double qmax, double norm,
               double q, double* out) {
    if (q < 0.0) return -2;          /* domain_error: q must be non-negative */
    if (q > qmax) return -3;         /* domain_error: q exceeds q_max */
    if (q == 0.0) { *out = t[0] * norm; return 0; }
    double s = q / dq;
    int i0 = (int)floor(s) - 1;
    double px = s - (double)i0;
    if (i0 < 0) { i0 = 0; px = s; }
    if (i0 + 3 >= nq) return -3;
    *out = lagrange_cubic(t[i0], t[i0 + 1], t[i0 + 2], t[i0 + 3], px) * norm;
    return 0;
}

/* ------------------------------------------------------------------------- */
/*  Real spherical harmonics (src/math/sph_harmonics.cppm).                   */
/*  Q_l^m column recurrences :120-170, angular-gradient recurrences          */
/*  :188-335, Y assembly :702-735, :737-787, :830-875, cos/sin(m phi) by the  */
/*  angle-addition recurrence of setLMax :641-696.                            */
/* ------------------------------------------------------------------------- */

#define ORC_LMAX 6
#define QIDX(l, m) ((l) * ((l) + 1) / 2 + (m))

typedef struct {
    double ct, st, cp, sp;
    double Q[QIDX(ORC_LMAX, ORC_LMAX) + 1];
    double dQ[QIDX(ORC_LMAX, ORC_LMAX) + 1];   /* dQ/dtheta */
    double R[QIDX(ORC_LMAX, ORC_LMAX) + 1];    /* Q / sin(theta) */
    double cm[ORC_LMAX + 1], sm[ORC_LMAX + 1];
} ylm_ws;

static double seed_coeff(int m) {
    double c = 0.5 / sqrt(PI);
    if (m % 2 != 0) c = -c;
    for (int k = 1; k <= m; ++k) c *= sqrt((double)(2 * k + 1) / (2 * k));
    return c;
}

static void adv_coeff(int m, int l, double* r1, double* r2, double* den) {
    *r1 = (2.0 * l - 1.0) * sqrt(((2.0 * l + 1.0) * (l - m)) / ((2.0 * l - 1.0) * (l + m)));
    *r2 = (l + m - 1.0) * sqrt(((2.0 * l + 1.0) * (l - m) * (l - m - 1)) /
                               ((2.0 * l - 3.0) * (l + m) * (l + m - 1)));
    *den = 1.0 / (l - m);
}

static void ylm_fill(ylm_ws* w, double theta, double phi, int lmax, int grads) {
    w->ct = cos(theta); w->st = sin(theta);
    w->cp = cos(phi);   w->sp = sin(phi);
    double ct = w->ct, st = w->st;
    for (int m = 0; m <= lmax; ++m) {
        /* Q column */
        double curr = (m == 0) ? 0.5 / sqrt(PI) : seed_coeff(m) * pow(st, m);
        double prev = 0.0;
        w->Q[QIDX(m, m)] = curr;
        if (m < lmax) {
            double f = sqrt(2.0 * m + 3.0);
            prev = curr; curr = f * ct * curr;
            w->Q[QIDX(m + 1, m)] = curr;
            for (int l = m + 2; l <= lmax; ++l) {
                double r1, r2, den; adv_coeff(m, l, &r1, &r2, &den);
                double nx = (ct * r1 * curr - r2 * prev) * den;
                prev = curr; curr = nx;
                w->Q[QIDX(l, m)] = curr;
            }
        }
    }
    /* cos(m phi), sin(m phi) */
    double c = 1.0, s = 0.0;
    for (int m = 0; m <= lmax; ++m) {
        w->cm[m] = c; w->sm[m] = s;
        if (m < lmax) {
            double oc = c;
            c = c * w->cp - s * w->sp;
            s = s * w->cp + oc * w->sp;
        }
    }
    if (!grads) return;
    /* R = Q/sin column (AngGradPhiRecurrence), m >= 1 */
    for (int m = 1; m <= lmax; ++m) {
        double c0 = seed_coeff(m);
        double curr = (m == 1) ? c0 : c0 * pow(st, m - 1);
        double prev = 0.0;
        w->R[QIDX(m, m)] = curr;
        if (m < lmax) {
            double f = sqrt(2.0 * m + 3.0);
            prev = curr; curr = f * ct * curr;
            w->R[QIDX(m + 1, m)] = curr;
            for (int l = m + 2; l <= lmax; ++l) {
                double r1, r2, den; adv_coeff(m, l, &r1, &r2, &den);
                double nx = (ct * r1 * curr - r2 * prev) * den;
                prev = curr; curr = nx;
                w->R[QIDX(l, m)] = curr;
            }
        }
    }
    /* dQ/dtheta column (AngGradThetaRecurrence) */
    for (int m = 0; m <= lmax; ++m) {
        double curr;
        if (m == 0) curr = 0.0;
        else if (m == 1) curr = seed_coeff(1) * ct;
        else curr = (double)m * seed_coeff(m) * pow(st, m - 1) * ct;
        double prev = 0.0;
        w->dQ[QIDX(m, m)] = curr;
        if (m < lmax) {
            double f = sqrt(2.0 * m + 3.0);
            prev = curr;
            curr = f * (-st * w->Q[QIDX(m, m)] + ct * curr);
            w->dQ[QIDX(m + 1, m)] = curr;
            for (int l = m + 2; l <= lmax; ++l) {
                double r1, r2, den; adv_coeff(m, l, &r1, &r2, &den);
                double nx = (ct * r1 * curr - st * r1 * w->Q[QIDX(l - 1, m)] - r2 * prev) * den;
                prev = curr; curr = nx;
                w->dQ[QIDX(l, m)] = curr;
            }
        }
    }
}

static double ylm_get(const ylm_ws* w, int l, int m) {
    int ma = m < 0 ? -m : m;
    double Q = w->Q[QIDX(l, ma)];
    if (m == 0) return Q;
    double v = sqrt(2.0) * Q;
    return m > 0 ? v * w->cm[ma] : v * w->sm[ma];
}

static double ylm_dtheta(const ylm_ws* w, int l, int m) {
    int ma = m < 0 ? -m : m;
    double d = w->dQ[QIDX(l, ma)];
    if (m == 0) return d;
    return m > 0 ? sqrt(2.0) * d * w->cm[ma] : sqrt(2.0) * d * w->sm[ma];
}

static double ylm_dphi(const ylm_ws* w, int l, int m) {
    if (m == 0) return 0.0;
    int ma = m < 0 ? -m : m;
    double R = w->R[QIDX(l, ma)];
    if (m > 0) return (-(double)m * sqrt(2.0)) * R * w->sm[ma];
    return ((double)ma * sqrt(2.0)) * R * w->cm[ma];
}

void orc_ylm(double theta, double phi, int lmax, double* ylm,
             double* dtheta, double* dphi_over_sin) {
    ylm_ws w;
    if (lmax > ORC_LMAX) lmax = ORC_LMAX;
    ylm_fill(&w, theta, phi, lmax, dtheta != NULL || dphi_over_sin != NULL);
    for (int l = 0; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m) {
            int o = l * l + (m + l);
            ylm[o] = ylm_get(&w, l, m);
            if (dtheta) dtheta[o] = ylm_dtheta(&w, l, m);
            if (dphi_over_sin) dphi_over_sin[o] = ylm_dphi(&w, l, m);
        }
}

/* GKK::computeSpherical (src/io/GKK.cppm:431-449). */
static void spherical(const double* K, double* q, double* th, double* ph) {
    double kx = K[0], ky = K[1], kz = K[2];
    double qq = sqrt(kx * kx + ky * ky + kz * kz);
    *q = qq;
    *th = (qq > 0.0) ? acos(kz / qq) : 0.0;
    *ph = atan2(ky, kx);
}

/* ------------------------------------------------------------------------- */
/*  Structure factor, separable mode (src/H_psi/structure_factor.cppm).       */
/* ------------------------------------------------------------------------- */

typedef struct {
    int n[3];
    int off[3];
    cplx kp;
    cplx* p[3];
} sfac_ws;

#define SF_BUF 4

static void sfac_init(sfac_ws* s, const double tau[3], const double kf[3],
                      int n1, int n2, int n3) {
    s->n[0] = n1; s->n[1] = n2; s->n[2] = n3;
    double kd = -2.0 * PI * (kf[0] * tau[0] + kf[1] * tau[1] + kf[2] * tau[2]);
    s->kp = cexp(I * kd);
    for (int d = 0; d < 3; ++d) {
        int n = s->n[d];
        int sh = n / 2;
        s->off[d] = sh + SF_BUF;
        s->p[d] = (cplx*)malloc(sizeof(cplx) * (size_t)(n + 2 * SF_BUF));
        for (int g = -sh - SF_BUF; g < n - sh + SF_BUF; ++g)
            s->p[d][g + sh + SF_BUF] = cexp(I * (-2.0 * PI * g * tau[d]));
    }
}

static void sfac_free(sfac_ws* s) { for (int d = 0; d < 3; ++d) free(s->p[d]); }

static cplx sfac_eval(const sfac_ws* s, const int32_t* g) {
    return s->kp * s->p[0][g[0] + s->off[0]] * s->p[1][g[1] + s->off[1]] *
           s->p[2][g[2] + s->off[2]];
}

void orc_sfac(const double tau[3], const int g[3], const double kfrac[3],
              int n1, int n2, int n3, double out[2]) {
    sfac_ws s;
    sfac_init(&s, tau, kfrac, n1, n2, n3);
    int32_t gg[3] = {g[0], g[1], g[2]};
    cplx v = sfac_eval(&s, gg);
    out[0] = creal(v); out[1] = cimag(v);
    sfac_free(&s);
}

/* ------------------------------------------------------------------------- */
/*  H|psi> for one band (src/H_psi/hamiltonian_callable.cpp:94-304).          */
/* ------------------------------------------------------------------------- */

typedef struct {
    int npw;
    double* q; double* th; double* ph;
    int lmax;
    double* ylm;     /* [(lmax+1)^2][npw] */
    double* ydt;     /* angular gradients (gradient only) */
    double* ydp;
    double* st; double* ct; double* sp; double* cp;
} kdata;

static int kdata_init(kdata* d, const orc_cell* c, const orc_kpt* kp, int grads) {
    int lmax = -1;
    for (int s = 0; s < c->nspecies; ++s) if (c->sp[s].lmax > lmax) lmax = c->sp[s].lmax;
    if (lmax > ORC_LMAX) return -4;
    d->npw = kp->npw;
    d->lmax = lmax;
    size_t n = (size_t)kp->npw;
    d->q = malloc(sizeof(double) * n);
    d->th = malloc(sizeof(double) * n);
    d->ph = malloc(sizeof(double) * n);
    int nlm = (lmax + 1) * (lmax + 1);
    d->ylm = nlm > 0 ? malloc(sizeof(double) * n * nlm) : NULL;
    d->ydt = (grads && nlm > 0) ? malloc(sizeof(double) * n * nlm) : NULL;
    d->ydp = (grads && nlm > 0) ? malloc(sizeof(double) * n * nlm) : NULL;
    d->st = malloc(sizeof(double) * n); d->ct = malloc(sizeof(double) * n);
    d->sp = malloc(sizeof(double) * n); d->cp = malloc(sizeof(double) * n);
    for (size_t g = 0; g < n; ++g) {
        spherical(kp->K + 3 * g, &d->q[g], &d->th[g], &d->ph[g]);
        if (lmax >= 0) {
            ylm_ws w;
            ylm_fill(&w, d->th[g], d->ph[g], lmax, grads);
            d->st[g] = w.st; d->ct[g] = w.ct; d->sp[g] = w.sp; d->cp[g] = w.cp;
            for (int l = 0; l <= lmax; ++l)
                for (int m = -l; m <= l; ++m) {
                    int o = l * l + m + l;
                    d->ylm[(size_t)o * n + g] = ylm_get(&w, l, m);
                    if (grads) {
                        d->ydt[(size_t)o * n + g] = ylm_dtheta(&w, l, m);
                        d->ydp[(size_t)o * n + g] = ylm_dphi(&w, l, m);
                    }
                }
        }
    }
    return 0;
}

static void kdata_free(kdata* d) {
    free(d->q); free(d->th); free(d->ph); free(d->ylm); free(d->ydt); free(d->ydp);
    free(d->st); free(d->ct); free(d->sp); free(d->cp);
}

/* Nonlocal term for one band, loops as in callable.cpp:211-304. */
static int vnl_band(const orc_cell* c, const orc_kpt* kp, const kdata* kd,
                    const cplx* psi, cplx* hpsi) {
    int ng = kp->npw;
    double norm = 1.0 / sqrt(c->omega);
    double* beta_q = malloc(sizeof(double) * (size_t)ng * 64);
    cplx* sf = malloc(sizeof(cplx) * (size_t)ng);
    int rc = 0;
    for (int s = 0; s < c->nspecies && rc == 0; ++s) {
        const orc_species* sp = &c->sp[s];
        if (sp->lmax < 0) continue;
        int nc = 0;
        int cl[64]; double clam[64];
        for (int ib = 0; ib < sp->nbeta && nc < 64; ++ib) {
            double lam = sp->lambda[ib];
            if (fabs(lam) < 1e-15) continue;
            for (int g = 0; g < ng; ++g) {
                rc = orc_interp(sp->beta_tab + (size_t)ib * sp->nq, sp->nq, sp->dq, sp->qmax,
                                norm, kd->q[g], &beta_q[(size_t)nc * ng + g]);
                if (rc) break;
            }
            if (rc) break;
            cl[nc] = sp->beta_l[ib]; clam[nc] = lam; ++nc;
        }
        if (rc || nc == 0) continue;
        for (int a = 0; a < c->natom; ++a) {
            if (c->species[a] != s) continue;
            sfac_ws sw;
            sfac_init(&sw, c->tau + 3 * a, kp->kfrac, c->n1, c->n2, c->n3);
            for (int g = 0; g < ng; ++g) sf[g] = sfac_eval(&sw, kp->gidx + 3 * g);
            sfac_free(&sw);
            for (int ic = 0; ic < nc; ++ic) {
                int l = cl[ic];
                const double* bq = beta_q + (size_t)ic * ng;
                for (int m = -l; m <= l; ++m) {
                    const double* y = kd->ylm + (size_t)(l * l + m + l) * ng;
                    cplx inner = 0.0;
                    for (int g = 0; g < ng; ++g) {
                        cplx p = (bq[g] * y[g]) * sf[g];
                        inner += conj(p) * psi[g];
                    }
                    for (int g = 0; g < ng; ++g) {
                        cplx p = (bq[g] * y[g]) * sf[g];
                        hpsi[g] += clam[ic] * inner * p;
                    }
                }
            }
        }
    }
    free(beta_q);
    free(sf);
    return rc;
}

static void local_band(const orc_cell* c, const orc_kpt* kp, fft3_ctx* fc, cplx* grid,
                       const cplx* psi, cplx* hpsi) {
    int n1 = c->n1, n2 = c->n2, n3 = c->n3;
    int ng = kp->npw;
    size_t N = (size_t)n1 * n2 * n3;
    /* kinetic (callable.cpp:119-126) */
    for (int g = 0; g < ng; ++g) hpsi[g] += kp->kin[g] * psi[g];
    /* scatter (callable.cpp:138-149) */
    memset(grid, 0, sizeof(cplx) * N);
    for (int g = 0; g < ng; ++g) {
        const int32_t* gi = kp->gidx + 3 * g;
        size_t i = (size_t)(((gi[0] % n1) + n1) % n1);
        size_t j = (size_t)(((gi[1] % n2) + n2) % n2);
        size_t k = (size_t)(((gi[2] % n3) + n3) % n3);
        grid[i * n2 * n3 + j * n3 + k] = psi[g];
    }
    fft3_run(fc, grid, n1, n2, n3, 0);                 /* G2R */
    for (size_t r = 0; r < N; ++r) grid[r] *= c->vloc[r];   /* :154-163 */
    fft3_run(fc, grid, n1, n2, n3, 1);                 /* R2G */
    /* gather (callable.cpp:171-182) */
    for (int g = 0; g < ng; ++g) {
        const int32_t* gi = kp->gidx + 3 * g;
        size_t i = (size_t)(((gi[0] % n1) + n1) % n1);
        size_t j = (size_t)(((gi[1] % n2) + n2) % n2);
        size_t k = (size_t)(((gi[2] % n3) + n3) % n3);
        hpsi[g] += grid[i * n2 * n3 + j * n3 + k];
    }
}

static int hpsi_range(const orc_cell* c, const orc_kpt* kp, const kdata* kd,
                      int b0, int b1, const double* psi, double* hpsi, int nonlocal_only) {
    size_t N = (size_t)c->n1 * c->n2 * c->n3;
    cplx* grid = nonlocal_only ? NULL : malloc(sizeof(cplx) * N);
    fft3_ctx fc;
    if (!nonlocal_only) fft3_init(&fc, c->n1, c->n2, c->n3);
    int rc = 0;
    for (int b = b0; b < b1 && rc == 0; ++b) {
        const cplx* p = (const cplx*)psi + (size_t)b * kp->npw;
        cplx* h = (cplx*)hpsi + (size_t)b * kp->npw;
        if (!nonlocal_only) local_band(c, kp, &fc, grid, p, h);
        rc = vnl_band(c, kp, kd, p, h);
    }
    if (!nonlocal_only) { fft3_free(&fc); free(grid); }
    return rc;
}

int orc_hpsi(const orc_cell* c, const orc_kpt* kp, int nb, const double* psi, double* hpsi) {
    kdata kd;
    int rc = kdata_init(&kd, c, kp, 0);
    if (rc) return rc;
    rc = hpsi_range(c, kp, &kd, 0, nb, psi, hpsi, 0);
    kdata_free(&kd);
    return rc;
}

int orc_vnl(const orc_cell* c, const orc_kpt* kp, int nb, const double* psi, double* hpsi) {
    kdata kd;
    int rc = kdata_init(&kd, c, kp, 0);
    if (rc) return rc;
    rc = hpsi_range(c, kp, &kd, 0, nb, psi, hpsi, 1);
    kdata_free(&kd);
    return rc;
}

int orc_hpsi_parallel(const orc_cell* c, const orc_kpt* kp, int nb, const double* psi,
                      double* hpsi, int nthreads) {
    kdata kd;
    int rc = kdata_init(&kd, c, kp, 0);
    if (rc) return rc;
    int err = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : err)
    {
        int t = omp_get_thread_num(), nt = omp_get_num_threads();
        int b0 = (int)((long)nb * t / nt), b1 = (int)((long)nb * (t + 1) / nt);
        err |= (hpsi_range(c, kp, &kd, b0, b1, psi, hpsi, 0) != 0);
    }
#else
    (void)nthreads;
    err = hpsi_range(c, kp, &kd, 0, nb, psi, hpsi, 0) != 0;
#endif
    kdata_free(&kd);
    return err ? -3 : 0;
}

/* ------------------------------------------------------------------------- */
/*  dH/dk |psi> (src/H_psi/hamiltonian_gradient.cpp:116-294).                */
/* ------------------------------------------------------------------------- */

static int dhdk_band(const orc_cell* c, const orc_kpt* kp, const kdata* kd,
                     const cplx* psi, cplx* ox, cplx* oy, cplx* oz) {
    int ng = kp->npw;
    int nx = ox != NULL, ny = oy != NULL, nz = oz != NULL;
    /* kinetic: out_a += K_a psi (:139-141) */
    if (nx) for (int g = 0; g < ng; ++g) ox[g] += kp->K[3 * g + 0] * psi[g];
    if (ny) for (int g = 0; g < ng; ++g) oy[g] += kp->K[3 * g + 1] * psi[g];
    if (nz) for (int g = 0; g < ng; ++g) oz[g] += kp->K[3 * g + 2] * psi[g];

    double norm = 1.0 / sqrt(c->omega);
    double* bqa = malloc(sizeof(double) * (size_t)ng * 64);
    double* dbqa = malloc(sizeof(double) * (size_t)ng * 64);
    cplx* sf = malloc(sizeof(cplx) * (size_t)ng);
    int rc = 0;
    for (int s = 0; s < c->nspecies && rc == 0; ++s) {
        const orc_species* sp = &c->sp[s];
        if (sp->lmax < 0) continue;
        if (!sp->dbeta_tab) { rc = -5; break; }
        /* beta / beta' cache per (l, ib), shared by all atoms (:189-213) */
        int nc = 0;
        int cl[64]; double clam[64], cdb0[64];
        for (int ib = 0; ib < sp->nbeta && nc < 64 && rc == 0; ++ib) {
            double lam = sp->lambda[ib];
            if (fabs(lam) < 1e-15) continue;
            const double* bt = sp->beta_tab + (size_t)ib * sp->nq;
            const double* dbt = sp->dbeta_tab + (size_t)ib * sp->nq;
            rc = orc_interp(dbt, sp->nq, sp->dq, sp->qmax, norm, 0.0, &cdb0[nc]);
            for (int g = 0; g < ng && rc == 0; ++g) {
                rc = orc_interp(bt, sp->nq, sp->dq, sp->qmax, norm, kd->q[g], &bqa[(size_t)nc * ng + g]);
                if (!rc) rc = orc_interp(dbt, sp->nq, sp->dq, sp->qmax, norm, kd->q[g], &dbqa[(size_t)nc * ng + g]);
            }
            cl[nc] = sp->beta_l[ib]; clam[nc] = lam; ++nc;
        }
        if (rc || nc == 0) continue;
        for (int a = 0; a < c->natom; ++a) {
            if (c->species[a] != s) continue;
            sfac_ws sw;
            sfac_init(&sw, c->tau + 3 * a, kp->kfrac, c->n1, c->n2, c->n3);
            for (int g = 0; g < ng; ++g) sf[g] = sfac_eval(&sw, kp->gidx + 3 * g);
            sfac_free(&sw);
            for (int ic = 0; ic < nc; ++ic) {
                int l = cl[ic];
                double lam = clam[ic], db0 = cdb0[ic];
                const double* bq = bqa + (size_t)ic * ng;
                const double* dbq = dbqa + (size_t)ic * ng;
                for (int m = -l; m <= l; ++m) {
                    int o = l * l + m + l;
                    const double* y = kd->ylm + (size_t)o * ng;
                    const double* yt = kd->ydt + (size_t)o * ng;
                    const double* yp = kd->ydp + (size_t)o * ng;
                    cplx in0 = 0.0, inx = 0.0, iny = 0.0, inz = 0.0;
                    for (int pass = 0; pass < 2; ++pass) {
                        for (int g = 0; g < ng; ++g) {
                            double q = kd->q[g];
                            double by = bq[g] * y[g];
                            double bo = (q > 1e-12) ? bq[g] / q : db0;
                            double gq = dbq[g] * y[g];
                            double gt = bo * yt[g];
                            double gp = bo * yp[g];
                            double sth = kd->st[g], cth = kd->ct[g];
                            double sph = kd->sp[g], cph = kd->cp[g];
                            cplx S = sf[g];
                            cplx p = by * S;
                            cplx dx = (sth * cph * gq + cth * cph * gt - sph * gp) * S;
                            cplx dy = (sth * sph * gq + cth * sph * gt + cph * gp) * S;
                            cplx dz = (cth * gq - sth * gt) * S;
                            if (pass == 0) {
                                in0 += conj(p) * psi[g];
                                if (nx) inx += conj(dx) * psi[g];
                                if (ny) iny += conj(dy) * psi[g];
                                if (nz) inz += conj(dz) * psi[g];
                            } else {
                                if (nx) ox[g] += lam * (inx * p + in0 * dx);
                                if (ny) oy[g] += lam * (iny * p + in0 * dy);
                                if (nz) oz[g] += lam * (inz * p + in0 * dz);
                            }
                        }
                    }
                }
            }
        }
    }
    free(bqa); free(dbqa); free(sf);
    return rc;
}

int orc_dhdk(const orc_cell* c, const orc_kpt* kp, int nb, const double* psi,
             double* ox, double* oy, double* oz) {
    kdata kd;
    int rc = kdata_init(&kd, c, kp, 1);
    if (rc) return rc;
    size_t n = (size_t)kp->npw;
    for (int b = 0; b < nb && rc == 0; ++b) {
        rc = dhdk_band(c, kp, &kd, (const cplx*)psi + b * n,
                       ox ? (cplx*)ox + b * n : NULL,
                       oy ? (cplx*)oy + b * n : NULL,
                       oz ? (cplx*)oz + b * n : NULL);
    }
    kdata_free(&kd);
    return rc;
}
