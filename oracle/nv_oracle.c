/*
 * nv_oracle.c -- TEST INFRASTRUCTURE ONLY.  A plain-C FP64 restatement of the
 * reference's Lanczos path (cavanes1/NadVibS, Fortran), used as the checker
 * of the HIP library by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline.  Nothing in the product links or calls this file.
 *
 * It follows the reference's order of operations where that is cheap:
 *   term numbering           Source/top.f90:150-234
 *   hoelem table             Source/function.f90:186-250, Source/top.f90:501-527
 *   diagonal block list      Source/top.f90:528-553
 *   H*v                      Source/top.f90:672-818 (diagonal rows first, then
 *                            terms -> steps -> rows -> state pairs)
 *   Lanczos                  Source/top.f90:646-670, :820-842
 *   omega recurrence / PRO   Source/top.f90:844-1005
 *   GIVENS (tridiagonal)     Source/yarkony.f90:91-543
 *   start vector (default)   Source/secondary.f90:727-783
 * Differences that only move the last bits: the factor product of a step is
 * formed per row (the reference updates it by division, :799-801), and the
 * noise of the omega estimates comes from a seeded counter-based generator
 * instead of an unseeded random_number stream (Source/main.f90:40).
 *
 * Pinning: tests/test_oracle.py checks every function here against dense
 * Kronecker brute force, closed forms and scipy (see the test docstrings).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/nadvibs.h"

#define MAXF 4
#define MAXM 100  /* modes: basis(100) of the reference's namelist (Source/secondary.f90:830-833) */

typedef struct {
    int cnt;
    int mode[MAXF], pow[MAXF], smax[MAXF]; /* 0-based modes, descending */
    int nblk;
    int *blk;      /* packed block index (0-based) */
    double *raw;   /* raw coefficient per block */
    int npairs;
    int *prow, *pcol; /* state pairs in nzblks order */
    double *pcoef;
    int diag;
    long ndir;
} OTerm;

typedef struct {
    int ns, M, ordr, nstblks, quirks;
    long D, N, nmax;
    long *nfunc, *shft;
    double *omega, *concoef;
    int nterms;
    OTerm *terms;
    long nztrms[MAXF];
    /* homatel and iordr(k, p): offset of the (p, delta = p%2 + 2(k-1)) table */
    double *homatel;
    long iordr[MAXF + 1][MAXF + 2];
    int ndblks;
    int dbl_i[64], dbl_j[64];
    int *bi, *bj;
    double *statew;
    long *ntcom;
    int neworigin;
    double *bomega, *dvec, *tmat;
    long *npirrep, *istate;
    int nirreps;
} OOp;

/* ------------------------------------------------------------------ */
static double o_hoelem(int p, long a, long b, int quirks) {
    /* 1-based a, b as in Source/function.f90:186; only b >= a is used */
    double m = a - 1.0;
    if (p == 1) {
        if (b == a + 1) return sqrt((m + 1) / 2);
        return 0.0;
    }
    if (p == 2) {
        if (b == a + 2) return sqrt((m + 1) * (m + 2) / 4);
        if (b == a) return ((2 * m + 1) / 2);
        return 0.0;
    }
    if (p == 3) {
        if (b == a + 3) return sqrt((m + 1) * (m + 2) * (m + 3) / 8);
        if (b == a + 1) {
            double v = sqrt(((m + 1) * (m + 1) * (m + 1)) / 8);
            return (quirks & NV_QUIRK_Q3ELEM) ? v : 3.0 * v;
        }
        return 0.0;
    }
    if (p == 4) {
        if (b == a + 4) return sqrt((m + 1) * (m + 2) * (m + 3) * (m + 4)) / 4;
        if (b == a + 2) return (2 * m + 3) * sqrt((m + 1) * (m + 2)) / 2;
        if (b == a) return 3 * (2 * m * m + 2 * m + 1) / 4;
        return 0.0;
    }
    return 0.0;
}

static long o_noterms(long M, long p) {
    long r = 1;
    for (long i = 1; i <= p; ++i) r = r * (M - 1 + i) / i;
    return r;
}

/* odometer of Source/top.f90:153-168, 1-based digits */
static void o_next_tuple(int *otab, int p, int M) {
    otab[0] += 1;
    for (int l = 0; l < p - 1; ++l)
        if (otab[l] > M) { otab[l] = 1; otab[l + 1] += 1; }
    for (int l = p - 2; l >= 0; --l)
        if (otab[l] < otab[l + 1]) otab[l] = otab[l + 1];
}

void or_destroy(OOp *op) {
    if (!op) return;
    for (int t = 0; t < op->nterms; ++t) {
        free(op->terms[t].blk); free(op->terms[t].raw);
        free(op->terms[t].prow); free(op->terms[t].pcol); free(op->terms[t].pcoef);
    }
    free(op->terms); free(op->nfunc); free(op->shft); free(op->omega); free(op->concoef);
    free(op->homatel); free(op->bi); free(op->bj); free(op->statew); free(op->ntcom);
    free(op->bomega); free(op->dvec); free(op->tmat); free(op->npirrep); free(op->istate);
    free(op);
}

OOp *or_create(const nv_problem *pr) {
    if (pr->ordr < 1 || pr->ordr > MAXF || pr->nstates < 1 || pr->nmodes < 1 || pr->nmodes > MAXM) return NULL;
    OOp *op = (OOp *)calloc(1, sizeof(OOp));
    op->ns = (int)pr->nstates; op->M = (int)pr->nmodes; op->ordr = (int)pr->ordr; op->quirks = pr->quirks;
    op->nstblks = op->ns * (op->ns + 1) / 2;
    op->nfunc = (long *)malloc(op->M * sizeof(long));
    op->shft = (long *)malloc(op->M * sizeof(long));
    op->omega = (double *)malloc(op->M * sizeof(double));
    op->concoef = (double *)malloc(op->nstblks * sizeof(double));
    op->statew = (double *)malloc(op->ns * sizeof(double));
    op->D = 1; op->nmax = 1;
    for (int k = 0; k < op->M; ++k) {
        op->nfunc[k] = pr->nfunc[k];
        op->omega[k] = pr->aomega[k];
        op->D *= op->nfunc[k];
        if (op->nfunc[k] > op->nmax) op->nmax = op->nfunc[k];
    }
    op->N = op->D * op->ns;
    for (int b = 0; b < op->nstblks; ++b) op->concoef[b] = pr->concoef[b];
    for (int s = 0; s < op->ns; ++s) op->statew[s] = pr->statew ? pr->statew[s] : 1.0;
    op->neworigin = pr->neworigin;
    op->istate = (long *)calloc(op->M, sizeof(long));
    if (pr->istate) for (int k = 0; k < op->M; ++k) op->istate[k] = pr->istate[k];
    op->nirreps = 1;
    op->npirrep = (long *)malloc(sizeof(long) * (pr->nirreps > 1 ? pr->nirreps : 1));
    op->npirrep[0] = op->M;
    if (pr->nirreps > 1 && pr->npirrep) {
        long tot = 0;
        for (long g = 0; g < pr->nirreps; ++g) tot += pr->npirrep[g];
        if (tot == op->M) {
            op->nirreps = (int)pr->nirreps;
            for (long g = 0; g < pr->nirreps; ++g) op->npirrep[g] = pr->npirrep[g];
        }
    }
    if (op->neworigin) {
        op->bomega = (double *)malloc(op->M * sizeof(double));
        op->dvec = (double *)malloc(op->M * sizeof(double));
        op->tmat = (double *)malloc((size_t)op->M * op->M * sizeof(double));
        for (int k = 0; k < op->M; ++k) { op->bomega[k] = pr->bomega[k]; op->dvec[k] = pr->dvec[k]; }
        for (int k = 0; k < op->M * op->M; ++k) op->tmat[k] = pr->tmat[k];
    }
    /* shft(i) = prod_{j>i} nfunc(j), Source/top.f90:379-386 */
    for (int i = 0; i < op->M; ++i) {
        op->shft[i] = 1;
        for (int j = i + 1; j < op->M; ++j) op->shft[i] *= op->nfunc[j];
    }
    op->bi = (int *)malloc(op->nstblks * sizeof(int));
    op->bj = (int *)malloc(op->nstblks * sizeof(int));
    for (int i = 1; i <= op->ns; ++i)
        for (int j = 1; j <= i; ++j) { op->bi[i * (i - 1) / 2 + j - 1] = i - 1; op->bj[i * (i - 1) / 2 + j - 1] = j - 1; }

    /* term numbering: group by number of distinct modes, first-seen order */
    const long ntmax = o_noterms(op->M, op->ordr);
    int cap = 16;
    OTerm *tmp[MAXF + 1];
    int ntmp[MAXF + 1] = {0}, ctmp[MAXF + 1];
    for (int c = 0; c <= MAXF; ++c) { ctmp[c] = cap; tmp[c] = (OTerm *)calloc(cap, sizeof(OTerm)); }
    for (int b = 0; b < op->nstblks; ++b) {
        for (int p = 1; p <= op->ordr; ++p) {
            int otab[MAXF];
            for (int l = 0; l < MAXF; ++l) otab[l] = 1;
            otab[0] = 0;
            long nt = o_noterms(op->M, p);
            for (long k = 0; k < nt; ++k) {
                o_next_tuple(otab, p, op->M);
                double c = pr->potterms[k + (long)(p - 1) * ntmax + (long)b * ntmax * op->ordr];
                int uniq[MAXF], cntv[MAXF], cnt = 0;
                for (int l = 0; l < p; ++l) {
                    int f = -1;
                    for (int q = 0; q < cnt; ++q) if (uniq[q] == otab[l]) f = q;
                    if (f < 0) { uniq[cnt] = otab[l]; cntv[cnt] = 1; cnt++; } else cntv[f]++;
                }
                if (!(fabs(c) >= pr->ztoler)) continue;
                OTerm *t = NULL;
                for (int q = 0; q < ntmp[cnt]; ++q) {
                    int same = 1;
                    for (int l = 0; l < cnt; ++l)
                        if (tmp[cnt][q].mode[l] != uniq[l] - 1 || tmp[cnt][q].pow[l] != cntv[l]) same = 0;
                    if (same) { t = &tmp[cnt][q]; break; }
                }
                if (!t) {
                    if (ntmp[cnt] == ctmp[cnt]) {
                        ctmp[cnt] *= 2;
                        tmp[cnt] = (OTerm *)realloc(tmp[cnt], ctmp[cnt] * sizeof(OTerm));
                        memset(tmp[cnt] + ntmp[cnt], 0, (ctmp[cnt] - ntmp[cnt]) * sizeof(OTerm));
                    }
                    t = &tmp[cnt][ntmp[cnt]++];
                    t->cnt = cnt;
                    for (int l = 0; l < cnt; ++l) { t->mode[l] = uniq[l] - 1; t->pow[l] = cntv[l]; }
                    t->blk = (int *)malloc(op->nstblks * sizeof(int));
                    t->raw = (double *)malloc(op->nstblks * sizeof(double));
                }
                t->blk[t->nblk] = b;
                t->raw[t->nblk] = c;
                t->nblk++;
            }
        }
    }
    op->nterms = 0;
    for (int c = 1; c <= op->ordr; ++c) { op->nztrms[c - 1] = ntmp[c]; op->nterms += ntmp[c]; }
    op->terms = (OTerm *)calloc(op->nterms > 0 ? op->nterms : 1, sizeof(OTerm));
    int at = 0;
    for (int c = 1; c <= op->ordr; ++c)
        for (int q = 0; q < ntmp[c]; ++q) op->terms[at++] = tmp[c][q];
    for (int c = 0; c <= MAXF; ++c) free(tmp[c]);

    /* coefficients: c * prod 1/sqrt(w^p), mirrored pairs (Source/top.f90:202-229) */
    for (int t = 0; t < op->nterms; ++t) {
        OTerm *T = &op->terms[t];
        double dval = 1.0;
        for (int l = 0; l < T->cnt; ++l) {
            double w = op->omega[T->mode[l]], wp = 1.0;
            for (int e = 0; e < T->pow[l]; ++e) wp *= w;
            dval = dval / sqrt(wp);
        }
        T->prow = (int *)malloc(2 * T->nblk * sizeof(int));
        T->pcol = (int *)malloc(2 * T->nblk * sizeof(int));
        T->pcoef = (double *)malloc(2 * T->nblk * sizeof(double));
        T->npairs = 0;
        for (int q = 0; q < T->nblk; ++q) {
            int b = T->blk[q];
            T->prow[T->npairs] = op->bi[b]; T->pcol[T->npairs] = op->bj[b]; T->pcoef[T->npairs] = T->raw[q] * dval;
            T->npairs++;
            if (op->bi[b] != op->bj[b]) {
                T->prow[T->npairs] = op->bj[b]; T->pcol[T->npairs] = op->bi[b]; T->pcoef[T->npairs] = T->raw[q] * dval;
                T->npairs++;
            }
        }
        /* smax, ndir, diagonal-term flag (Source/secondary.f90:29-60) */
        T->ndir = 1;
        T->diag = 1;
        for (int l = 0; l < T->cnt; ++l) {
            int s = T->pow[l];
            while (s > op->nfunc[T->mode[l]] - 1) s -= 2;
            T->smax[l] = s;
            if (s < 0) T->ndir = 0; else T->ndir *= (s + 1);
            if (s % 2 != 0) T->diag = 0;
        }
    }

    /* homatel (Source/top.f90:501-527) */
    long nmatel = 0;
    for (int i = 1; i <= op->ordr; ++i)
        for (int j = i % 2; j <= op->ordr; j += 2)
            if (op->nmax - j > 0) nmatel += op->nmax - j;
    op->homatel = (double *)calloc(nmatel + 1, sizeof(double));
    long off = 0;
    for (int i = 1; i <= op->ordr; ++i) {
        int k = 1;
        for (int j = i % 2; j <= op->ordr; j += 2, ++k) {
            op->iordr[k][i] = off;
            for (long l = 1; l <= op->nmax - j; ++l) op->homatel[off++] = o_hoelem(i, l, l + j, op->quirks);
        }
    }

    /* diagonal block list (Source/top.f90:528-553) */
    int dchk[16][16];
    memset(dchk, 0, sizeof(dchk));
    for (int b = 0; b < op->nstblks; ++b)
        if (fabs(op->concoef[b]) > 1e-16) { dchk[op->bi[b]][op->bj[b]] = 1; dchk[op->bj[b]][op->bi[b]] = 1; }
    for (int t = 0; t < op->nterms; ++t)
        if (op->terms[t].diag && op->terms[t].ndir > 0)
            for (int q = 0; q < op->terms[t].npairs; ++q) dchk[op->terms[t].prow[q]][op->terms[t].pcol[q]] = 1;
    op->ndblks = 0;
    for (int i = 0; i < op->ns; ++i)
        for (int j = 0; j < i; ++j)
            if (dchk[i][j]) { op->dbl_i[op->ndblks] = i; op->dbl_j[op->ndblks] = j; op->ndblks++; }

    /* term counts ntcom(blk, cnt) (Source/top.f90:438-499) */
    op->ntcom = (long *)calloc((size_t)op->ordr * op->nstblks, sizeof(long));
    for (int t = 0; t < op->nterms; ++t) {
        OTerm *T = &op->terms[t];
        if (T->ndir == 0) continue;
        long trm1 = op->D;
        for (int l = 0; l < T->cnt; ++l) trm1 /= op->nfunc[T->mode[l]];
        int d[MAXF];
        for (int l = 0; l < T->cnt; ++l) d[l] = -T->smax[l];
        for (long k = 0; k < T->ndir; ++k) {
            long trm2 = 1;
            for (int l = 0; l < T->cnt; ++l) trm2 *= op->nfunc[T->mode[l]] - abs(d[l]);
            for (int q = 0; q < T->nblk; ++q) op->ntcom[(T->cnt - 1) * op->nstblks + T->blk[q]] += trm1 * trm2;
            int l = T->cnt - 1;
            while (l >= 0 && d[l] >= T->smax[l]) { d[l] = -T->smax[l]; --l; }
            if (l < 0) break;
            d[l] += 2;
        }
    }
    return op;
}

long or_dims(const OOp *op, long *D) { if (D) *D = op->D; return op->N; }
void or_term_counts(const OOp *op, long *out) {
    for (long i = 0; i < (long)op->ordr * op->nstblks; ++i) out[i] = op->ntcom[i];
}
void or_term_summary(const OOp *op, long *out) { for (int c = 0; c < op->ordr; ++c) out[c] = op->nztrms[c]; }

static double o_h(const OOp *op, int p, long nmin, int adelta) {
    /* homatel(iordr(stype, p) + l), stype = adelta/2+1, l = nmin+1 (1-based) */
    return op->homatel[op->iordr[adelta / 2 + 1][p] + nmin];
}

static int g_off32 = 0;
static double g_eps_pro = 0.0;

/* Unit roundoff used by the omega recurrence and its thresholds (0: the
 * reference's 2^-53). */
void or_set_eps_pro(double e) { g_eps_pro = e; }
/* Thresholds of the detection (0 = sqrt(eps), eps^(3/4)): the library's
 * mixed modes use NV_PRO_MIXED_TAU / NV_PRO_MIXED_ETA. */
static double g_tau = 0.0, g_eta = 0.0;
void or_set_pro_thresholds(double tau, double eta) { g_tau = tau; g_eta = eta; }
/* Emulation switch: the off-diagonal (vibrational) part of H*v with the
 * input rounded to float and float accumulation, the diagonal in double --
 * the precision split of the library's mixed mode. */
void or_set_off32(int on) { g_off32 = on; }

/* y = H x, Source/top.f90:672-818 */
void or_hv(const OOp *op, const double *x, double *y) {
    const int ns = op->ns, M = op->M;
    const long D = op->D;
    const int lower = (op->quirks & NV_QUIRK_LOWERDIAG) != 0;
    memset(y, 0, (size_t)op->N * sizeof(double));
    /* diagonal rows: harmonic energy by the odometer of :717-731 */
    long barray[MAXM] = {0};
    for (int j = 0; j < M; ++j) barray[j] = 1;
    barray[M - 1] -= 1;
    double dkin = 0.0;
    for (int j = 0; j < M; ++j) dkin += (barray[j] - 0.5) * op->omega[j];
    double dv[16][16];
    for (long r = 0; r < D; ++r) {
        int k = M - 1;
        while (barray[k] >= op->nfunc[k]) {
            dkin -= (barray[k] - 1) * op->omega[k];
            barray[k] = 1;
            k--;
        }
        barray[k] += 1;
        dkin += op->omega[k];
        for (int b = 0; b < op->nstblks; ++b) {
            dv[op->bi[b]][op->bj[b]] = op->concoef[b];
            dv[op->bj[b]][op->bi[b]] = op->concoef[b];
        }
        for (int t = 0; t < op->nterms; ++t) {
            const OTerm *T = &op->terms[t];
            if (!T->diag || T->ndir == 0) continue;
            double d1 = 1.0;
            for (int l = 0; l < T->cnt; ++l) d1 = d1 * o_h(op, T->pow[l], barray[T->mode[l]] - 1, 0);
            for (int q = 0; q < T->npairs; ++q) dv[T->prow[q]][T->pcol[q]] += T->pcoef[q] * d1;
        }
        for (int s = 0; s < ns; ++s) y[s * D + r] += (dv[s][s] + dkin) * x[s * D + r];
        for (int q = 0; q < op->ndblks; ++q) {
            int i = op->dbl_i[q], j = op->dbl_j[q];
            y[i * D + r] += dv[i][j] * x[j * D + r];
            if (!lower) y[j * D + r] += dv[j][i] * x[i * D + r];
        }
    }
    /* off-diagonal terms: terms -> steps -> rows -> pairs (:758-808) */
    float *yf = NULL, *xf = NULL;
    if (g_off32) {
        yf = (float *)calloc(op->N, sizeof(float));
        xf = (float *)malloc(op->N * sizeof(float));
        for (long i = 0; i < op->N; ++i) xf[i] = (float)x[i];
    }
    for (int t = 0; t < op->nterms; ++t) {
        const OTerm *T = &op->terms[t];
        if (T->ndir == 0) continue;
        int d[MAXF];
        for (int l = 0; l < T->cnt; ++l) d[l] = -T->smax[l];
        for (long k = 0; k < T->ndir; ++k) {
            int zero = 1;
            long stride = 0;
            for (int l = 0; l < T->cnt; ++l) { if (d[l]) zero = 0; stride += d[l] * op->shft[T->mode[l]]; }
            if (!zero) {
                for (long r = 0; r < D; ++r) {
                    double d1 = 1.0;
                    int ok = 1;
                    for (int l = 0; l < T->cnt && ok; ++l) {
                        long n = (r / op->shft[T->mode[l]]) % op->nfunc[T->mode[l]];
                        long n2 = n + d[l];
                        if (n2 < 0 || n2 >= op->nfunc[T->mode[l]]) { ok = 0; break; }
                        d1 = d1 * o_h(op, T->pow[l], n < n2 ? n : n2, abs(d[l]));
                    }
                    if (!ok) continue;
                    if (g_off32) {
                        for (int q = 0; q < T->npairs; ++q)
                            yf[T->prow[q] * D + r] += xf[T->pcol[q] * D + r + stride] * (float)T->pcoef[q] * (float)d1;
                        continue;
                    }
                    for (int q = 0; q < T->npairs; ++q)
                        y[T->prow[q] * D + r] += x[T->pcol[q] * D + r + stride] * T->pcoef[q] * d1;
                }
            }
            int l = T->cnt - 1;
            while (l >= 0 && d[l] >= T->smax[l]) { d[l] = -T->smax[l]; --l; }
            if (l < 0) break;
            d[l] += 2;
        }
    }
    if (g_off32) {
        for (long i = 0; i < op->N; ++i) y[i] += (double)yf[i];
        free(yf);
        free(xf);
    }
}


void or_start_vector(const OOp *op, double *q);

/* One row r of y = H x (all states), the sums of or_hv in the same order for
 * that row: the diagonal part first (harmonic energy from the digits, the
 * dv block of Source/top.f90:733-754), then terms -> steps -> state pairs
 * (:758-808).  With the float emulation the off-diagonal part accumulates
 * in float from a float copy of x, as or_hv does. */
static void o_row(const OOp *op, const double *x, long r, double *yr) {
    const int ns = op->ns, M = op->M;
    const long D = op->D;
    const int lower = (op->quirks & NV_QUIRK_LOWERDIAG) != 0;
    long nd[MAXM];
    double dkin = 0.0;
    for (int j = 0; j < M; ++j) {
        nd[j] = (r / op->shft[j]) % op->nfunc[j];
        dkin += ((double)(nd[j] + 1) - 0.5) * op->omega[j];
    }
    double dv[16][16];
    for (int b = 0; b < op->nstblks; ++b) {
        dv[op->bi[b]][op->bj[b]] = op->concoef[b];
        dv[op->bj[b]][op->bi[b]] = op->concoef[b];
    }
    for (int t = 0; t < op->nterms; ++t) {
        const OTerm *T = &op->terms[t];
        if (!T->diag || T->ndir == 0) continue;
        double d1 = 1.0;
        for (int l = 0; l < T->cnt; ++l) d1 = d1 * o_h(op, T->pow[l], nd[T->mode[l]], 0);
        for (int q = 0; q < T->npairs; ++q) dv[T->prow[q]][T->pcol[q]] += T->pcoef[q] * d1;
    }
    for (int s = 0; s < ns; ++s) yr[s] = (dv[s][s] + dkin) * x[s * D + r];
    for (int q = 0; q < op->ndblks; ++q) {
        int i = op->dbl_i[q], j = op->dbl_j[q];
        yr[i] += dv[i][j] * x[j * D + r];
        if (!lower) yr[j] += dv[j][i] * x[i * D + r];
    }
    float yf[16] = {0};
    for (int t = 0; t < op->nterms; ++t) {
        const OTerm *T = &op->terms[t];
        if (T->ndir == 0) continue;
        int d[MAXF];
        for (int l = 0; l < T->cnt; ++l) d[l] = -T->smax[l];
        for (long k = 0; k < T->ndir; ++k) {
            int zero = 1;
            long stride = 0;
            for (int l = 0; l < T->cnt; ++l) { if (d[l]) zero = 0; stride += d[l] * op->shft[T->mode[l]]; }
            if (!zero) {
                double d1 = 1.0;
                int ok = 1;
                for (int l = 0; l < T->cnt && ok; ++l) {
                    long n = nd[T->mode[l]];
                    long n2 = n + d[l];
                    if (n2 < 0 || n2 >= op->nfunc[T->mode[l]]) { ok = 0; break; }
                    d1 = d1 * o_h(op, T->pow[l], n < n2 ? n : n2, abs(d[l]));
                }
                if (ok) {
                    if (g_off32) {
                        for (int q = 0; q < T->npairs; ++q)
                            yf[T->prow[q]] += (float)x[T->pcol[q] * D + r + stride] * (float)T->pcoef[q] * (float)d1;
                    } else {
                        for (int q = 0; q < T->npairs; ++q)
                            yr[T->prow[q]] += x[T->pcol[q] * D + r + stride] * T->pcoef[q] * d1;
                    }
                }
            }
            int l = T->cnt - 1;
            while (l >= 0 && d[l] >= T->smax[l]) { d[l] = -T->smax[l]; --l; }
            if (l < 0) break;
            d[l] += 2;
        }
    }
    if (g_off32)
        for (int s = 0; s < ns; ++s) yr[s] += (double)yf[s];
}

/* y at the listed rows: y[i*ns + s] = (H x)[s*D + rows[i]].  Agrees with
 * or_hv to the rounding of the harmonic energy (or_hv sums it by the
 * reference's odometer). */
void or_hv_rows(const OOp *op, const double *x, long nrows, const long *rows, double *y) {
    for (long i = 0; i < nrows; ++i) o_row(op, x, rows[i], y + i * op->ns);
}

/* y = H x with nthreads OpenMP threads over contiguous blocks of rows (the
 * reference's Global-Arrays row partition, Source/top.f90:400-418). */
void or_hv_par(const OOp *op, const double *x, double *y, int nthreads) {
    const long D = op->D;
    const int ns = op->ns;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long r = 0; r < D; ++r) {
        double yr[16];
        o_row(op, x, r, yr);
        for (int s = 0; s < ns; ++s) y[s * D + r] = yr[s];
    }
}

/* y = H x for the rows [r0, r1) only (other rows untouched), nthreads
 * threads: a bounded sample of or_hv_par. */
void or_hv_par_range(const OOp *op, const double *x, double *y, long r0, long r1, int nthreads) {
    const long D = op->D;
    const int ns = op->ns;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long r = r0; r < r1; ++r) {
        double yr[16];
        o_row(op, x, r, yr);
        for (int s = 0; s < ns; ++s) y[s * D + r] = yr[s];
    }
}

/* Iterations i0..i1 of the plain recurrence (reorthog = 0) with nthreads
 * threads, in the reference's order (Source/top.f90:646-670, :820-842):
 * state (q1, q2, q3) of 3N doubles; at i0 = 1 the default or FC start
 * vector is built into q3.  alpha/beta as or_lanczos (alpha[i-1], beta[i]). */
void or_lanczos_par(const OOp *op, long i0, long i1, int nthreads, double *state, double *alpha, double *beta) {
    const long N = op->N;
    double *q1 = state, *q2 = state + N, *q3 = state + 2 * N;
    if (i0 == 1) {
        memset(state, 0, (size_t)3 * N * sizeof(double));
        or_start_vector(op, q3);
        double b0 = 0.0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : b0) schedule(static)
        for (long i = 0; i < N; ++i) b0 += q3[i] * q3[i];
        beta[0] = sqrt(b0);
    }
    for (long it = i0; it <= i1; ++it) {
        const double sc = 1.0 / beta[it - 1];
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (long i = 0; i < N; ++i) {
            q1[i] = q2[i];
            q2[i] = q3[i] * sc;
        }
        or_hv_par(op, q2, q3, nthreads);
        double a = 0.0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : a) schedule(static)
        for (long i = 0; i < N; ++i) a += q2[i] * q3[i];
        alpha[it - 1] = a;
        const double b = it > 1 ? beta[it - 1] : 0.0;
        double nb = 0.0;
#pragma omp parallel for num_threads(nthreads) reduction(+ : nb) schedule(static)
        for (long i = 0; i < N; ++i) {
            double v = q3[i] - a * q2[i];
            if (it > 1) v = v - b * q1[i];
            q3[i] = v;
            nb += v * v;
        }
        beta[it] = sqrt(nb);
    }
}

/* ------------------------------------------------------------------ */
/* Franck-Condon start vector (Source/secondary.f90:351-783)            */
/* ------------------------------------------------------------------ */
/* Symmetric eigen-decomposition by cyclic Jacobi (the reference uses GIVENS
 * on the packed A; any accurate decomposition gives the same A^-1, det A). */
static void o_jacobi(int M, double *A, double *V, double *w) {
    for (int i = 0; i < M * M; ++i) V[i] = 0.0;
    for (int i = 0; i < M; ++i) V[i * M + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < M; ++i) for (int j = i + 1; j < M; ++j) off += A[i * M + j] * A[i * M + j];
        if (off < 1e-300) break;
        for (int p = 0; p < M; ++p)
            for (int q = p + 1; q < M; ++q) {
                if (A[p * M + q] == 0.0) continue;
                double th = (A[q * M + q] - A[p * M + p]) / (2 * A[p * M + q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < M; ++k) {
                    double akp = A[k * M + p], akq = A[k * M + q];
                    A[k * M + p] = c * akp - s * akq; A[k * M + q] = s * akp + c * akq;
                }
                for (int k = 0; k < M; ++k) {
                    double apk = A[p * M + k], aqk = A[q * M + k];
                    A[p * M + k] = c * apk - s * aqk; A[q * M + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < M; ++k) {
                    double vkp = V[k * M + p], vkq = V[k * M + q];
                    V[k * M + p] = c * vkp - s * vkq; V[k * M + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < M; ++i) w[i] = A[i * M + i];
}

static void o_fc_start(const OOp *op, double *q) {
    const int M = op->M;
    const long D = op->D;
    const double *a = op->omega, *b = op->bomega, *d = op->dvec;
    double *T = (double *)malloc(M * M * sizeof(double)), *W = (double *)malloc(M * M * sizeof(double));
    double *A = (double *)calloc(M * M, sizeof(double)), *V = (double *)malloc(M * M * sizeof(double));
    double *lam = (double *)malloc(M * sizeof(double)), *Ai = (double *)calloc(M * M, sizeof(double));
    double *A1 = (double *)calloc(M * M, sizeof(double)), *A2 = (double *)calloc(M * M, sizeof(double));
    double *A3 = (double *)calloc(M * M, sizeof(double));
    double *A2d = (double *)calloc(M, sizeof(double)), *A3d = (double *)calloc(M, sizeof(double));
    int *irrep = (int *)malloc(M * sizeof(int));
    for (int i = 0; i < M; ++i) for (int j = 0; j < M; ++j) T[j * M + i] = op->tmat[i * M + j];
    /* |det T| with partial pivoting (gauss_elim, :311-348) */
    memcpy(W, T, M * M * sizeof(double));
    double detT = 1.0;
    for (int i = 0; i < M; ++i) {
        int p = i;
        for (int j = i + 1; j < M; ++j) if (fabs(W[j * M + i]) > fabs(W[p * M + i])) p = j;
        if (p != i) for (int k = 0; k < M; ++k) { double t = W[i * M + k]; W[i * M + k] = W[p * M + k]; W[p * M + k] = t; }
        for (int j = i + 1; j < M; ++j) {
            if (W[j * M + i] == 0.0) continue;
            double v = -W[j * M + i] / W[i * M + i];
            for (int k = 0; k < M; ++k) W[j * M + k] += v * W[i * M + k];
        }
        detT *= fabs(W[i * M + i]);
    }
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            double s = 0.0;
            for (int k = 0; k < M; ++k) s += 0.5 * b[k] * T[k * M + i] * T[k * M + j];
            if (i == j) s += 0.5 * a[i];
            A[i * M + j] = s;
        }
    o_jacobi(M, A, V, lam);
    double detA = 1.0;
    for (int i = 0; i < M; ++i) detA *= fabs(lam[i]);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            double s = 0.0;
            for (int k = 0; k < M; ++k) s += V[i * M + k] * V[j * M + k] / lam[k];
            Ai[i * M + j] = s;
        }
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            A1[i * M + j] = Ai[i * M + j] * sqrt(a[i]) * sqrt(a[j]);
            double s2 = 0.0, s3 = 0.0;
            for (int k = 0; k < M; ++k) {
                s2 += Ai[i * M + k] * T[j * M + k];
                for (int l = 0; l < M; ++l) s3 += T[i * M + k] * Ai[k * M + l] * T[j * M + l];
            }
            A2[i * M + j] = s2 * sqrt(a[i]) * sqrt(b[j]);
            A3[i * M + j] = s3 * sqrt(b[i]) * sqrt(b[j]);
        }
    for (int i = 0; i < M; ++i) {
        A1[i * M + i] -= 1.0;
        for (int j = 0; j < M; ++j) {
            double d1 = A3[j * M + i] - (i == j ? 2.0 : 0.0);
            A2d[i] += A2[i * M + j] * d[j] * sqrt(b[j]);
            A3d[i] += d1 * d[j] * sqrt(b[j]);
        }
    }
    { int m0 = 0; for (int g = 0; g < op->nirreps; ++g) for (long k = 0; k < op->npirrep[g]; ++k) irrep[m0++] = g; }
    /* <0|n>: levels of total quanta, every function of level L pushes to its
     * neighbours n + e_j (Source/secondary.f90:487-570), in index order */
    double *C = (double *)calloc(D, sizeof(double));
    int *lev = (int *)malloc(D * sizeof(int));
    int nlev = 0;
    for (int k = 0; k < M; ++k) nlev += (int)op->nfunc[k] - 1;
    for (long r = 0; r < D; ++r) {
        int L = 0;
        for (int k = 0; k < M; ++k) L += (int)((r / op->shft[k]) % op->nfunc[k]);
        lev[r] = L;
    }
    C[0] = 1.0;
    for (int L = 0; L < nlev; ++L)
        for (long r = 0; r < D; ++r) {
            if (lev[r] != L) continue;
            for (int j = 0; j < M; ++j) {
                long nj = (r / op->shft[j]) % op->nfunc[j];
                if (nj + 1 >= op->nfunc[j]) continue;
                double s = 0.0;
                for (int k = 0; k < M; ++k) {
                    if (irrep[k] != irrep[j]) continue;
                    long nk = (r / op->shft[k]) % op->nfunc[k];
                    if (nk > 0) s += A1[j * M + k] * nk * C[r - op->shft[k]];
                }
                C[r + op->shft[j]] = 2.0 * s - A2d[j] * C[r];
            }
        }
    /* one excited mode: recursion in m (:602-689) */
    int im = 0;
    while (im < M - 1 && op->istate[im] <= 0) ++im;
    int ninit = (int)op->istate[im];
    double *Cp = (double *)calloc(D, sizeof(double)), *Cn = (double *)calloc(D, sizeof(double));
    for (int i = 1; i <= ninit; ++i) {
        for (long r = 0; r < D; ++r) {
            double v = 2.0 * Cp[r] * (i - 1) * (A3[im * M + im] - 1.0) - A3d[im] * C[r];
            for (int k = 0; k < M; ++k) {
                long nk = (r / op->shft[k]) % op->nfunc[k];
                if (nk > 0) v += 2.0 * A2[k * M + im] * nk * C[r - op->shft[k]];
            }
            Cn[r] = v;
        }
        double *t = Cp; Cp = C; C = Cn; Cn = t;
    }
    /* G and the scale factors (:694-725) */
    double e1 = 0.0, pab = 1.0;
    for (int i = 0; i < M; ++i) e1 -= 0.5 * b[i] * d[i] * d[i];
    for (int i = 0; i < M; ++i) for (int j = 0; j < M; ++j) e1 += 0.25 * sqrt(b[i]) * d[i] * A3[i * M + j] * d[j] * sqrt(b[j]);
    for (int i = 0; i < M; ++i) pab *= b[i] * a[i];
    const double G = sqrt(sqrt(pab) * detT / detA) * exp(e1);
    for (long r = 0; r < D; ++r) {
        double dd = 1.0;
        for (int k = 0; k < M; ++k) {
            long n = (r / op->shft[k]) % op->nfunc[k];
            double f = 1.0; for (long t = 2; t <= n; ++t) f *= t;
            double g = 1.0; for (long t = 2; t <= op->istate[k]; ++t) g *= t;
            dd *= pow(2.0, (double)n) * f * pow(2.0, (double)op->istate[k]) * g;
        }
        C[r] = G * C[r] / sqrt(dd);
    }
    for (int s = 0; s < op->ns; ++s) memcpy(q + (size_t)s * D, C, D * sizeof(double));
    free(T); free(W); free(A); free(V); free(lam); free(Ai); free(A1); free(A2); free(A3); free(A2d); free(A3d);
    free(irrep); free(C); free(lev); free(Cp); free(Cn);
}

/* start vector: FC overlaps or |0> in every state, then weights and
 * normalisation (Source/secondary.f90:727-783) */
void or_start_vector(const OOp *op, double *q) {
    memset(q, 0, (size_t)op->N * sizeof(double));
    if (op->neworigin) o_fc_start(op, q);
    else for (int s = 0; s < op->ns; ++s) q[s * op->D] = 1.0;
    double d1 = 0.0;
    for (long i = 0; i < op->N; ++i) d1 += q[i] * q[i];
    d1 = sqrt(d1 / op->ns);
    double d2 = 0.0;
    for (int s = 0; s < op->ns; ++s) d2 += op->statew[s];
    if (d2 == 0.0) d2 = 1.0;
    for (int s = 0; s < op->ns; ++s)
        for (long r = 0; r < op->D; ++r) q[s * op->D + r] *= op->statew[s] / (d2 * d1);
    d2 = 0.0;
    for (int s = 0; s < op->ns; ++s) {
        double t = 0.0;
        for (long r = 0; r < op->D; ++r) t += q[s * op->D + r] * q[s * op->D + r];
        d1 = sqrt(t);
        d2 += d1 * d1;
    }
    double sc = 1.0 / sqrt(d2);
    for (long i = 0; i < op->N; ++i) q[i] *= sc;
}

/* ------------------------------------------------------------------ */
/* Lanczos with optional partial re-orthogonalisation                   */
/* ------------------------------------------------------------------ */
static uint64_t o_splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
/* N(0, sigma) draw number idx of the seeded stream (same generator spec as
 * the library: splitmix64 on seed ^ (2 idx + K), Box-Muller) */
static double o_gauss(uint64_t seed, uint64_t idx, double sigma) {
    uint64_t a = o_splitmix(seed ^ (2 * idx + 0x632BE59BD9B4E019ull));
    uint64_t b = o_splitmix(seed ^ (2 * idx + 1 + 0x632BE59BD9B4E019ull));
    double u1 = ((double)(a >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    return sigma * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static double o_dot(const double *a, const double *b, long n) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

typedef struct {
    long events, touched;
} OrStats;

/*
 * Runs niter iterations.  alpha[0..niter-1] = alpha(1:niter), beta[0..niter]
 * = beta(0:niter), omega (if non-NULL) [2][niter+1] = omega(1:2, 0:niter).
 * start: the normalised start vector (NULL = default).  archive: NULL or
 * (niter+1)*N doubles receiving q_1..q_niter (needed when reorthog > 0;
 * allocated internally when NULL).  Returns 0 on success.
 */
int or_lanczos(const OOp *op, long niter, int reorthog, long chkorthog, uint64_t seed, double noise,
               const double *start, double *alpha, double *beta, double *omega_out, double *archive_in,
               long *stats) {
    const long N = op->N;
    const double eps = g_eps_pro > 0 ? g_eps_pro : 1.1102230246251565e-16;  /* determine_epsilon: 2^-53 */
    const double eta = g_eta > 0 ? g_eta : pow(eps, 0.75);
    const double tau = g_tau > 0 ? g_tau : sqrt(eps);
    const int orthog = reorthog > 0, exact = reorthog > 1;
    double *q1 = (double *)calloc(N, sizeof(double));
    double *q2 = (double *)calloc(N, sizeof(double));
    double *q3 = (double *)calloc(N, sizeof(double));
    double *arch = archive_in;
    int own_arch = 0;
    if (orthog && !arch) {
        arch = (double *)malloc((size_t)(niter + 1) * N * sizeof(double));
        own_arch = 1;
        if (!arch) return 1;
    }
    double *om1 = (double *)calloc(niter + 3, sizeof(double));
    double *om2 = (double *)calloc(niter + 3, sizeof(double));
    const long extent = (niter + 99) / 100;
    const long cap = niter + 2;
    long *loidx = (long *)calloc(cap, sizeof(long));
    long *roidx = (long *)calloc(cap, sizeof(long));
    long northog = 0;
    int firststep = 1;
    uint64_t draws = 0;
    double *dpv = (double *)malloc((size_t)2 * (niter + 2) * sizeof(double));
    long *jl = (long *)malloc((size_t)2 * (niter + 2) * sizeof(long));
    long ev = 0, touched = 0;

    if (start) memcpy(q3, start, N * sizeof(double));
    else or_start_vector(op, q3);
    if (orthog) {
        om1[0] = 0.0; om1[1] = 1.0; om2[0] = 0.0;
        om2[1] = noise * o_gauss(seed, draws++, 0.6) * eps * op->ns * op->D;
        if (niter >= 2) om2[2] = 1.0;
    }
    beta[0] = sqrt(o_dot(q3, q3, N));
    for (long it = 1; it <= niter; ++it) {
        /* update_lanczos */
        memcpy(q1, q2, N * sizeof(double));
        const double sc = 1.0 / beta[it - 1];
        for (long i = 0; i < N; ++i) q3[i] *= sc;
        memcpy(q2, q3, N * sizeof(double));
        if (arch) memcpy(arch + (size_t)it * N, q2, N * sizeof(double));
        /* Hv */
        or_hv(op, q2, q3);
        /* lanczos_step */
        alpha[it - 1] = o_dot(q2, q3, N);
        const double a = alpha[it - 1];
        for (long i = 0; i < N; ++i) q3[i] = q3[i] - a * q2[i];
        if (it > 1) {
            const double b = beta[it - 1];
            for (long i = 0; i < N; ++i) q3[i] = q3[i] - b * q1[i];
        }
        beta[it] = sqrt(o_dot(q3, q3, N));
        if (!orthog || it <= 1) continue;
        /* set_omega */
        if (exact) {
            for (long i = 0; i <= it; ++i) om2[i] = 0.0;
            for (long i = 1; i <= it; ++i) om2[i] = o_dot(q3, arch + (size_t)i * N, N);
        } else {
            if (it < niter) om2[it + 1] = 1.0;
            const double bi = beta[it];
            om1[it] = noise * o_gauss(seed, draws, 0.6) * eps * op->ns * op->D * beta[1] / bi;
            for (long k = 1; k <= it - 1; ++k)
                om1[k] = (1 / bi) * (beta[k] * om2[k + 1] + (alpha[k - 1] - alpha[it - 1]) * om2[k] +
                                     beta[k - 1] * om2[k - 1] - beta[it - 1] * om1[k]) +
                         eps * (beta[k] + bi) * (noise * o_gauss(seed, draws + (uint64_t)k, 0.3));
            draws += (uint64_t)it;
            for (long i = 1; i <= it; ++i) { double t = om2[i]; om2[i] = om1[i]; om1[i] = t; }
            if (it % chkorthog == 0) {
                for (long i = 1; i <= it - 1; ++i) {
                    const double c1 = o_dot(q2, arch + (size_t)i * N, N);
                    const double c2 = o_dot(q3, arch + (size_t)i * N, N) / bi;
                    if (fabs(om1[i]) < fabs(c1)) om1[i] = c1;
                    if (fabs(om2[i]) < fabs(c2)) om2[i] = c2;
                }
            }
        }
        /* partial_orthog */
        if (firststep) {
            northog = 0;
            long plo = 0, pro = 0;
            for (long i = 1; i <= it; ++i) {
                if (!(fabs(om2[i]) >= tau)) continue;
                long l = i;
                while (l >= 1 && !(fabs(om2[l]) < eta)) --l;
                if (l < 1) l = 1;
                long r = i;
                while (r <= it && !(fabs(om2[r]) < eta)) ++r;
                if (r > it) r = it;
                if (northog == 0 || (l != plo && r != pro)) {
                    if (northog < cap) { loidx[northog] = l; roidx[northog] = r; northog++; }
                    plo = l; pro = r;
                }
            }
            if (northog > 0) {
                long span = northog > extent ? northog : extent, mn = 0x7fffffffffffffffL;
                for (long k = 0; k < span && k < cap; ++k)
                    if (loidx[k] > 0 && loidx[k] < mn) mn = loidx[k];
                if (mn > 1 && northog < cap) { loidx[northog] = 1; roidx[northog] = 1; northog++; }
            }
        }
        if (northog > 0) {
            long n = 0;
            for (long k = 0; k < northog; ++k)
                for (long j = loidx[k]; j <= roidx[k]; ++j)
                    if (n < 2 * (niter + 2)) jl[n++] = j;
            for (long c = 0; c < n; ++c) dpv[c] = o_dot(q3, arch + (size_t)jl[c] * N, N);
            for (long c = 0; c < n; ++c) {
                const double *qj = arch + (size_t)jl[c] * N;
                const double d = dpv[c];
                for (long i = 0; i < N; ++i) q3[i] = q3[i] - d * qj[i];
            }
            for (long c = 0; c < n; ++c) om2[jl[c]] = eps * (noise * o_gauss(seed, draws + (uint64_t)c, 2.0));
            draws += (uint64_t)n;
            firststep = !firststep;
            beta[it] = sqrt(o_dot(q3, q3, N));
            ev += 1;
            touched += n;
        }
    }
    if (omega_out && orthog)
        for (long j = 0; j <= niter; ++j) { omega_out[j] = om1[j]; omega_out[niter + 1 + j] = om2[j]; }
    if (stats) { stats[0] = ev; stats[1] = touched; }
    free(q1); free(q2); free(q3); free(om1); free(om2); free(loidx); free(roidx); free(dpv); free(jl);
    if (own_arch) free(arch);
    return 0;
}

/* ------------------------------------------------------------------ */
/* GIVENS for a tridiagonal matrix (Source/yarkony.f90:91-543)          */
/* ------------------------------------------------------------------ */
/*
 * alpha[0..n-1], beta[0..n] (beta[0] unused, beta[n] for bji).  Outputs
 * eigenvalues ascending, inten = v1^2, bji = |beta[n] v_n|, sign of v1,
 * and (if vec != NULL) the eigenvectors, vec[i*n + j] = component j of root i.
 */
void or_givens(long n, const double *alpha, const double *beta, double *root, double *inten, double *bji,
               int *sign, double *vec_out) {
    const double eps = 1.1102230246251565e-16;
    const double THETA = 1.7976931348623157e308;
    const double DEL1 = eps / 100.0, DELTA = eps * eps * 100.0, DELBIG = THETA * DELTA / 1000.0;
    const double THETA1 = 1000.0 / THETA, TOLER = eps * 100.0;
    const double RPOWER = 8388608.0, RPOW1 = RPOWER / 2.0;
    double RAND1 = RPOWER - 3.0;
    double *vec = (double *)malloc((size_t)n * n * sizeof(double));
    double *a = (double *)malloc(n * sizeof(double));   /* diagonal */
    double *b = (double *)calloc(n + 1, sizeof(double)); /* b[j] couples j, j+1 */
    double *b3 = (double *)calloc(n + 1, sizeof(double));
    double *B2 = (double *)malloc(n * sizeof(double)), *B3 = (double *)malloc(n * sizeof(double));
    double *B4 = (double *)malloc(n * sizeof(double)), *B5 = (double *)malloc(n * sizeof(double));
    if (n == 1) {
        root[0] = alpha[0];
        vec[0] = 1.0;
        goto out;
    }
    {
        double factor = 0.0;
        for (long i = 0; i < n; ++i) {
            if (i > 0) factor = fmax(factor, fabs(beta[i]));
            factor = fmax(factor, fabs(alpha[i]));
        }
        if (factor == 0.0) {
            for (long i = 0; i < n; ++i) {
                root[i] = 0.0;
                for (long j = 0; j < n; ++j) vec[i * n + j] = i == j;
            }
            goto out;
        }
        double anorm = 0.0;
        for (long i = 0; i < n; ++i) {
            if (i > 0) anorm += (beta[i] / factor) * (beta[i] / factor);
            anorm += (alpha[i] / factor) * (alpha[i] / factor) / 2.0;
        }
        anorm = sqrt(anorm * 2.0) * factor;
        for (long i = 0; i < n; ++i) a[i] = alpha[i] / anorm;
        for (long i = 0; i + 1 < n; ++i) { b[i] = beta[i + 1] / anorm; b3[i] = b[i] * b[i]; }
        /* bisection, lowest root first (Source/yarkony.f90:331-381) */
        for (long i = 0; i < n; ++i) root[i] = 1.0;
        double rootl = -1.0;
        for (long I = 0; I < n; ++I) {
            double rx = 1.0;
            for (long j = I; j < n; ++j) rx = fmin(rx, root[j]);
            root[I] = rx;
            while (1) {
                double trial = (rootl + root[I]) * 0.5;
                if (trial == rootl || trial == root[I]) break;
                long nomtch = n;
                long j = 0;
                double f0 = a[0] - trial;
                while (1) {
                    if (fabs(f0) < THETA1) {
                        j += 2;
                        nomtch -= 1;
                        if (j < n) { f0 = a[j] - trial; continue; }
                        break;
                    }
                    if (f0 >= 0.0) nomtch -= 1;
                    j += 1;
                    if (j >= n) break;
                    f0 = a[j] - trial - b3[j - 1] / f0;
                }
                if (nomtch >= I + 1) {
                    root[I] = trial;
                    long nom = nomtch < n ? nomtch : n;
                    root[nom - 1] = trial;
                } else {
                    rootl = trial;
                }
            }
        }
        /* inverse iteration (Source/yarkony.f90:382-512) */
        for (long i = 0; i < n * n; ++i) vec[i] = 1.0;
        long ia = 0;
        for (long I = 0; I < n; ++I) {
            double aroot = root[I];
            if (I == 0 || !(fabs(root[I - 1] - aroot) < TOLER)) ia = 0;
            else ia += 1;
            double e1 = a[0] - aroot, e2 = b[0];
            for (long j = 0; j + 1 < n; ++j) {
                double t;
                const double bj1 = (j + 2 < n) ? b[j + 1] : 0.0;
                if (fabs(e1) <= fabs(b[j])) {
                    B2[j] = b[j]; B3[j] = a[j + 1] - aroot; B4[j] = bj1;
                    t = 1.0;
                    if (fabs(b[j]) > THETA1) t = e1 / b[j];
                    e1 = e2 - t * B3[j];
                    e2 = -t * bj1;
                } else {
                    B2[j] = e1; B3[j] = e2; B4[j] = 0.0;
                    t = b[j] / e1;
                    e1 = a[j + 1] - aroot - t * e2;
                    e2 = bj1;
                }
                B5[j] = t;
            }
            B2[n - 1] = e1; B3[n - 1] = 0.0; B4[n - 1] = 0.0; B4[n - 2] = 0.0;
            double *v = vec + I * n;
            if (ia != 0)
                for (long j = 0; j < n; ++j) {
                    RAND1 = fmod(4099.0 * RAND1, RPOWER);
                    v[j] = RAND1 / RPOW1 - 1.0;
                }
            double temp = 1.0;
            for (int pass = 0; pass < 2; ++pass) {
                for (long l = n - 1; l >= 0; --l) {
                    while (1) {
                        double v1 = l + 1 < n ? v[l + 1] : 0.0, v2 = l + 2 < n ? v[l + 2] : 0.0;
                        double el = v[l] - v1 * B3[l] - v2 * B4[l];
                        if (fabs(el) > DELBIG) {
                            for (long k = 0; k < n; ++k) v[k] = v[k] / DELBIG;
                            continue;
                        }
                        temp = B2[l];
                        if (fabs(B2[l]) < DELTA) temp = DELTA;
                        v[l] = el / temp;
                        break;
                    }
                }
                if (pass == 1) break;
                double el = v[0];
                for (long j = 0; j + 1 < n; ++j) {
                    if (B2[j] == b[j]) {
                        v[j] = v[j + 1];
                        el = el - v[j + 1] * temp;
                    } else {
                        v[j] = el;
                        el = v[j + 1] - el * B5[j];
                    }
                }
                v[n - 1] = el;
            }
            for (long j1 = 1; j1 <= ia; ++j1) {
                const double *w = vec + (I - j1) * n;
                double t = 0.0;
                for (long j = 0; j < n; ++j) t += v[j] * w[j];
                for (long j = 0; j < n; ++j) v[j] -= t * w[j];
            }
            double mx = 0.0;
            for (long j = 0; j < n; ++j) mx = fmax(fabs(v[j]), mx);
            double s = 0.0;
            for (long j = 0; j < n; ++j) { double q = v[j] / mx; s += q * q; }
            double tt = 1.0 / (sqrt(s) * mx);
            for (long j = 0; j < n; ++j) {
                v[j] = v[j] * tt;
                if (fabs(v[j]) < DEL1) v[j] = 0.0;
            }
        }
        for (long i = 0; i < n; ++i) root[i] *= anorm;
    }
out:
    for (long i = 0; i < n; ++i) {
        const double v1 = vec[i * n], vn = vec[i * n + n - 1];
        if (inten) inten[i] = v1 * v1;
        if (bji) bji[i] = fabs(beta[n] * vn);
        if (sign) sign[i] = v1 < 0 ? -1 : 1;
    }
    if (vec_out) memcpy(vec_out, vec, (size_t)n * n * sizeof(double));
    free(vec); free(a); free(b); free(b3); free(B2); free(B3); free(B4); free(B5);
}
