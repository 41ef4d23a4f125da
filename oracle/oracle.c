/*
 * oracle.c -- CPU restatement of Photon's gas-phase pressure projection path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never shipped or
 * called by it.  Built with gcc -std=c11 (ISO mode: no FMA contraction, like the
 * reference's -std=c++20 build, CMakeLists.txt:5-6,16-17), so that the FP32
 * path reproduces the reference bit for bit.  Each function cites the
 * reference file:line it restates.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */
typedef struct { int col; float w; } ent;
typedef struct { ent *e; int n, cap; } row_t;

static void row_add(row_t *r, int col, float w) {   /* accumulate or append, pressure_mg.cpp:94-97,171-173 */
    for (int q = 0; q < r->n; ++q)
        if (r->e[q].col == col) { r->e[q].w += w; return; }
    if (r->n == r->cap) {
        r->cap = r->cap ? 2 * r->cap : 8;
        r->e = (ent *)realloc(r->e, sizeof(ent) * (size_t)r->cap);
    }
    r->e[r->n].col = col;
    r->e[r->n].w = w;
    r->n++;
}

/* union-find with path halving, joined under the smaller index (pressure_mg.cpp:20-25) */
static int uf_find(int *p, int a) {
    while (p[a] != a) a = p[a] = p[p[a]];
    return a;
}
static void uf_join(int *p, int a, int b) {
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a != b) { if (a > b) p[a] = b; else p[b] = a; }
}

static void finish_level(orc_level *C, row_t *rows, const double *dg) {
    C->diag = (float *)malloc(sizeof(float) * (size_t)(C->n > 0 ? C->n : 1));
    C->rowStart = (int *)malloc(sizeof(int) * (size_t)(C->n + 1));
    C->rowStart[0] = 0;
    for (int I = 0; I < C->n; ++I) {
        C->diag[I] = (float)(dg[I] > 0.0 ? dg[I] : 0.0);
        C->rowStart[I + 1] = C->rowStart[I] + rows[I].n;
    }
    C->nnz = C->rowStart[C->n];
    C->col = (int *)malloc(sizeof(int) * (size_t)(C->nnz > 0 ? C->nnz : 1));
    C->val = (float *)malloc(sizeof(float) * (size_t)(C->nnz > 0 ? C->nnz : 1));
    int q = 0;
    for (int I = 0; I < C->n; ++I) {
        for (int e = 0; e < rows[I].n; ++e) { C->col[q] = rows[I].e[e].col; C->val[q] = rows[I].e[e].w; ++q; }
        free(rows[I].e);
    }
}

/* ------------------------------------------------------------------------ */
/* buildPressureLevels (pressure_mg.cpp:28-191)                             */
/* ------------------------------------------------------------------------ */
orc_mg *orc_mg_build(int nx, int ny, int nz, const float *ax_, const float *ay_, const float *az_,
                     const uint8_t *blocked, const uint8_t *wind_closed) {
    orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
    const int n0 = nx * ny * nz, sxy = nx * ny;
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz, nfz = (size_t)nx * ny * (nz + 1);
#define XF(i, j, k) ((i) + (nx + 1) * ((j) + ny * (k)))
#define YF(i, j, k) ((i) + nx * ((j) + (ny + 1) * (k)))
#define ZF(i, j, k) ((i) + nx * ((j) + ny * (k)))
    orc_level *L0 = &mg->L[0];
    L0->nx = nx; L0->ny = ny; L0->nz = nz; L0->n = n0;
    L0->ax = (float *)malloc(sizeof(float) * nfx);
    L0->ay = (float *)malloc(sizeof(float) * nfy);
    L0->az = (float *)malloc(sizeof(float) * nfz);
    memcpy(L0->ax, ax_, sizeof(float) * nfx);
    memcpy(L0->ay, ay_, sizeof(float) * nfy);
    memcpy(L0->az, az_, sizeof(float) * nfz);
    /* boundary faces where wind prescribes inflow leave the matrix (:36-46) */
    uint8_t wc[4] = {0, 0, 0, 0};
    if (wind_closed) memcpy(wc, wind_closed, 4);
    for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j) {
            if (wc[0]) L0->ax[XF(0, j, k)] = 0;
            if (wc[1]) L0->ax[XF(nx, j, k)] = 0;
        }
        for (int i = 0; i < nx; ++i) {
            if (wc[2]) L0->ay[YF(i, 0, k)] = 0;
            if (wc[3]) L0->ay[YF(i, ny, k)] = 0;
        }
    }
    L0->diag = (float *)calloc((size_t)n0, sizeof(float));
    L0->leak = (uint8_t *)calloc((size_t)n0, 1);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = i + nx * (j + ny * k);
                if (blocked[c]) continue;
                L0->diag[c] = L0->ax[XF(i, j, k)] + L0->ax[XF(i + 1, j, k)] + L0->ay[YF(i, j, k)] +
                              L0->ay[YF(i, j + 1, k)] + L0->az[ZF(i, j, k)] + L0->az[ZF(i, j, k + 1)];
                /* open (Dirichlet) boundary face of the matrix: the cell's zone is not singular */
                if (L0->diag[c] > 0 &&
                    ((i == 0 && L0->ax[XF(0, j, k)] > 0) || (i == nx - 1 && L0->ax[XF(nx, j, k)] > 0) ||
                     (j == 0 && L0->ay[YF(i, 0, k)] > 0) || (j == ny - 1 && L0->ay[YF(i, ny, k)] > 0) ||
                     (k == 0 && L0->az[ZF(i, j, 0)] > 0) || (k == nz - 1 && L0->az[ZF(i, j, nz)] > 0)))
                    L0->leak[c] = 1;
            }
    /* ---- level 1: connected pieces of 2x2x2 blocks (:58-139) */
    {
        int *p = (int *)malloc(sizeof(int) * (size_t)n0);
        for (int c = 0; c < n0; ++c) p[c] = c;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int c = i + nx * (j + ny * k);
                    if (L0->diag[c] <= 0) continue;
                    if (i + 1 < nx && (i >> 1) == ((i + 1) >> 1) && L0->ax[XF(i + 1, j, k)] > 0 && L0->diag[c + 1] > 0)
                        uf_join(p, c, c + 1);
                    if (j + 1 < ny && (j >> 1) == ((j + 1) >> 1) && L0->ay[YF(i, j + 1, k)] > 0 && L0->diag[c + nx] > 0)
                        uf_join(p, c, c + nx);
                    if (k + 1 < nz && (k >> 1) == ((k + 1) >> 1) && L0->az[ZF(i, j, k + 1)] > 0 && L0->diag[c + sxy] > 0)
                        uf_join(p, c, c + sxy);
                }
        L0->agg = (int *)malloc(sizeof(int) * (size_t)n0);
        int *rootId = (int *)malloc(sizeof(int) * (size_t)n0);
        for (int c = 0; c < n0; ++c) { L0->agg[c] = -1; rootId[c] = -1; }
        orc_level *C = &mg->L[1];
        int cap = 64;
        C->bi = (int *)malloc(sizeof(int) * (size_t)cap);
        C->bj = (int *)malloc(sizeof(int) * (size_t)cap);
        C->bk = (int *)malloc(sizeof(int) * (size_t)cap);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int c = i + nx * (j + ny * k);
                    if (L0->diag[c] <= 0) continue;
                    int r = uf_find(p, c);
                    if (rootId[r] < 0) {
                        if (C->n == cap) {
                            cap *= 2;
                            C->bi = (int *)realloc(C->bi, sizeof(int) * (size_t)cap);
                            C->bj = (int *)realloc(C->bj, sizeof(int) * (size_t)cap);
                            C->bk = (int *)realloc(C->bk, sizeof(int) * (size_t)cap);
                        }
                        rootId[r] = C->n;
                        C->bi[C->n] = i >> 1; C->bj[C->n] = j >> 1; C->bk[C->n] = k >> 1;
                        C->n++;
                    }
                    L0->agg[c] = rootId[r];
                }
        free(rootId);
        free(p);
        row_t *rows = (row_t *)calloc((size_t)(C->n > 0 ? C->n : 1), sizeof(row_t));
        double *dg = (double *)calloc((size_t)(C->n > 0 ? C->n : 1), sizeof(double));
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int c = i + nx * (j + ny * k);
                    if (L0->diag[c] <= 0) continue;
                    int nb[3] = {i + 1 < nx ? c + 1 : -1, j + 1 < ny ? c + nx : -1, k + 1 < nz ? c + sxy : -1};
                    float wv[3] = {i + 1 < nx ? L0->ax[XF(i + 1, j, k)] : 0, j + 1 < ny ? L0->ay[YF(i, j + 1, k)] : 0,
                                   k + 1 < nz ? L0->az[ZF(i, j, k + 1)] : 0};
                    for (int d = 0; d < 3; ++d) {   /* couple() (:90-100) */
                        if (nb[d] < 0) continue;
                        float w = wv[d];
                        if (w <= 0) continue;
                        int A = L0->agg[c], B = L0->agg[nb[d]];
                        if (A < 0 || B < 0 || A == B) continue;
                        row_add(&rows[A], B, w);
                        row_add(&rows[B], A, w);
                    }
                }
        for (int c = 0; c < n0; ++c) if (L0->agg[c] >= 0) dg[L0->agg[c]] += L0->diag[c];
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int c = i + nx * (j + ny * k);
                    if (L0->diag[c] <= 0) continue;
                    /* internal faces were counted twice in the fine diagonal sum (:117-126) */
                    if (i + 1 < nx && L0->diag[c + 1] > 0) {
                        float w = L0->ax[XF(i + 1, j, k)];
                        if (w > 0 && L0->agg[c + 1] == L0->agg[c]) dg[L0->agg[c]] -= 2.0 * w;
                    }
                    if (j + 1 < ny && L0->diag[c + nx] > 0) {
                        float w = L0->ay[YF(i, j + 1, k)];
                        if (w > 0 && L0->agg[c + nx] == L0->agg[c]) dg[L0->agg[c]] -= 2.0 * w;
                    }
                    if (k + 1 < nz && L0->diag[c + sxy] > 0) {
                        float w = L0->az[ZF(i, j, k + 1)];
                        if (w > 0 && L0->agg[c + sxy] == L0->agg[c]) dg[L0->agg[c]] -= 2.0 * w;
                    }
                }
        finish_level(C, rows, dg);
        free(rows);
        free(dg);
        C->leak = (uint8_t *)calloc((size_t)(C->n > 0 ? C->n : 1), 1);
        for (int c = 0; c < n0; ++c) if (L0->agg[c] >= 0 && L0->leak[c]) C->leak[L0->agg[c]] = 1;
        mg->nl = 2;
    }
    /* ---- coarser levels: connected pieces of 2x2x2 groups of blocks (:140-184) */
    while (mg->L[mg->nl - 1].n > 2000 && mg->nl < 12) {
        orc_level *F = &mg->L[mg->nl - 1];
        int *p = (int *)malloc(sizeof(int) * (size_t)F->n);
        for (int I = 0; I < F->n; ++I) p[I] = I;
        for (int I = 0; I < F->n; ++I)
            for (int e = F->rowStart[I]; e < F->rowStart[I + 1]; ++e) {
                int J = F->col[e];
                if (J > I && (F->bi[I] >> 1) == (F->bi[J] >> 1) && (F->bj[I] >> 1) == (F->bj[J] >> 1) &&
                    (F->bk[I] >> 1) == (F->bk[J] >> 1))
                    uf_join(p, I, J);
            }
        int *agg = (int *)malloc(sizeof(int) * (size_t)F->n);
        int *rootId = (int *)malloc(sizeof(int) * (size_t)F->n);
        for (int I = 0; I < F->n; ++I) rootId[I] = -1;
        orc_level C;
        memset(&C, 0, sizeof(C));
        C.bi = (int *)malloc(sizeof(int) * (size_t)F->n);
        C.bj = (int *)malloc(sizeof(int) * (size_t)F->n);
        C.bk = (int *)malloc(sizeof(int) * (size_t)F->n);
        for (int I = 0; I < F->n; ++I) {
            int r = uf_find(p, I);
            if (rootId[r] < 0) {
                rootId[r] = C.n;
                C.bi[C.n] = F->bi[I] >> 1; C.bj[C.n] = F->bj[I] >> 1; C.bk[C.n] = F->bk[I] >> 1;
                C.n++;
            }
            agg[I] = rootId[r];
        }
        free(rootId);
        free(p);
        if (C.n > 0.8 * F->n) {   /* not coarsening any more (:162) */
            free(agg); free(C.bi); free(C.bj); free(C.bk);
            break;
        }
        F->agg = agg;
        row_t *rows = (row_t *)calloc((size_t)(C.n > 0 ? C.n : 1), sizeof(row_t));
        double *dg = (double *)calloc((size_t)(C.n > 0 ? C.n : 1), sizeof(double));
        for (int I = 0; I < F->n; ++I) dg[F->agg[I]] += F->diag[I];
        for (int I = 0; I < F->n; ++I)
            for (int e = F->rowStart[I]; e < F->rowStart[I + 1]; ++e) {
                int J = F->col[e];
                int A = F->agg[I], B = F->agg[J];
                if (A == B) { dg[A] -= F->val[e]; continue; }
                int found = 0;
                for (int q = 0; q < rows[A].n; ++q)
                    if (rows[A].e[q].col == B) { rows[A].e[q].w += F->val[e]; found = 1; break; }
                if (!found) row_add(&rows[A], B, F->val[e]);
            }
        finish_level(&C, rows, dg);
        free(rows);
        free(dg);
        C.leak = (uint8_t *)calloc((size_t)(C.n > 0 ? C.n : 1), 1);
        for (int I = 0; I < F->n; ++I) if (F->leak[I]) C.leak[F->agg[I]] = 1;
        mg->L[mg->nl] = C;
        mg->nl++;
    }
#undef XF
#undef YF
#undef ZF
    return mg;
}

void orc_mg_free(orc_mg *mg) {
    if (!mg) return;
    for (int l = 0; l < mg->nl; ++l) {
        orc_level *L = &mg->L[l];
        free(L->ax); free(L->ay); free(L->az); free(L->rowStart); free(L->col); free(L->val);
        free(L->bi); free(L->bj); free(L->bk); free(L->agg); free(L->diag); free(L->leak);
    }
    free(mg->chol);
    free(mg->cmask);
    free(mg->ccomp);
    free(mg);
}

int orc_mg_nlevels(const orc_mg *mg) { return mg->nl; }
int orc_mg_level_n(const orc_mg *mg, int l) { return mg->L[l].n; }
int orc_mg_level_nnz(const orc_mg *mg, int l) { return mg->L[l].nnz; }
void orc_mg_level_get(const orc_mg *mg, int l, int *rowStart, int *col, float *val, float *diag, int *agg,
                      int *bi, int *bj, int *bk, uint8_t *leak) {
    const orc_level *L = &mg->L[l];
    if (rowStart && L->rowStart) memcpy(rowStart, L->rowStart, sizeof(int) * (size_t)(L->n + 1));
    if (col && L->col) memcpy(col, L->col, sizeof(int) * (size_t)L->nnz);
    if (val && L->val) memcpy(val, L->val, sizeof(float) * (size_t)L->nnz);
    if (diag) memcpy(diag, L->diag, sizeof(float) * (size_t)L->n);
    if (agg) {
        if (L->agg) memcpy(agg, L->agg, sizeof(int) * (size_t)L->n);
        else for (int i = 0; i < L->n; ++i) agg[i] = -1;
    }
    if (bi && L->bi) memcpy(bi, L->bi, sizeof(int) * (size_t)L->n);
    if (bj && L->bj) memcpy(bj, L->bj, sizeof(int) * (size_t)L->n);
    if (bk && L->bk) memcpy(bk, L->bk, sizeof(int) * (size_t)L->n);
    if (leak && L->leak) memcpy(leak, L->leak, (size_t)L->n);
}

/* ------------------------------------------------------------------------ */
/* exact coarsest solve: FP64 Cholesky of the coarsest operator, grounded on */
/* each closed component by adding s*1*1^T (the operator's null space on a   */
/* closed component is the constant vector).  This is the build's GPU        */
/* coarsest solve; the reference uses 20 x forward/backward Gauss-Seidel     */
/* (pressure_mg.cpp:277-281), selected by ORC_COARSE_GS20.                   */
/* ------------------------------------------------------------------------ */
static int coarsest_index(const orc_mg *mg) {
    int l = 0;
    while (l + 1 < mg->nl && mg->L[l].agg != NULL) ++l;
    return l;
}

static void ensure_chol(orc_mg *mg) {
    if (mg->chol) return;
    const orc_level *L = &mg->L[coarsest_index(mg)];
    const int n = L->n;
    double *A = (double *)calloc((size_t)n * (size_t)n + 1, sizeof(double));
    mg->cmask = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    int *p = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int I = 0; I < n; ++I) p[I] = I;
    for (int I = 0; I < n; ++I) {
        if (L->diag[I] <= 0) { mg->cmask[I] = 1; A[(size_t)I * n + I] = 1.0; continue; }
        A[(size_t)I * n + I] += L->diag[I];
        for (int e = L->rowStart[I]; e < L->rowStart[I + 1]; ++e) {
            A[(size_t)I * n + L->col[e]] -= L->val[e];
            if (L->diag[L->col[e]] > 0) uf_join(p, I, L->col[e]);
        }
    }
    /* closed components: no member touches an open boundary face */
    int *open = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    int *cnt = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    double *dsum = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int I = 0; I < n; ++I) {
        if (mg->cmask[I]) continue;
        int r = uf_find(p, I);
        open[r] |= L->leak[I];
        cnt[r]++;
        dsum[r] += L->diag[I];
    }
    for (int I = 0; I < n; ++I) {
        if (mg->cmask[I]) continue;
        int r = uf_find(p, I);
        if (open[r]) continue;
        double s = dsum[r] / ((double)cnt[r] * (double)cnt[r]);
        for (int J = 0; J < n; ++J)
            if (!mg->cmask[J] && uf_find(p, J) == r) A[(size_t)I * n + J] += s;
    }
    /* closed-component ids, for the projection onto zero component sums */
    mg->ccomp = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *cid = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    for (int I = 0; I < n; ++I) cid[I] = -1;
    mg->nccomp = 0;
    for (int I = 0; I < n; ++I) {
        mg->ccomp[I] = -1;
        if (mg->cmask[I]) continue;
        int r = uf_find(p, I);
        if (open[r]) continue;
        if (cid[r] < 0) cid[r] = mg->nccomp++;
        mg->ccomp[I] = cid[r];
    }
    free(cid);
    free(open); free(cnt); free(dsum); free(p);
    /* in-place Cholesky, lower triangle */
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        d = d > 0 ? sqrt(d) : 1e-300;
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = s / d;
        }
    }
    mg->chol = A;
}

/* remove the mean of y on every closed component of the coarsest level */
static void project_closed(const orc_mg *mg, int n, double *y) {
    if (mg->nccomp == 0) return;
    double *sum = (double *)calloc((size_t)mg->nccomp, sizeof(double));
    double *cnt = (double *)calloc((size_t)mg->nccomp, sizeof(double));
    for (int I = 0; I < n; ++I)
        if (mg->ccomp[I] >= 0) { sum[mg->ccomp[I]] += y[I]; cnt[mg->ccomp[I]] += 1; }
    for (int I = 0; I < n; ++I)
        if (mg->ccomp[I] >= 0) y[I] -= sum[mg->ccomp[I]] / cnt[mg->ccomp[I]];
    free(sum);
    free(cnt);
}

static void chol_solve(const double *Lm, int n, double *y) {
    for (int i = 0; i < n; ++i) {
        double s = y[i];
        for (int k = 0; k < i; ++k) s -= Lm[(size_t)i * n + k] * y[k];
        y[i] = s / Lm[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < n; ++k) s -= Lm[(size_t)k * n + i] * y[k];
        y[i] = s / Lm[(size_t)i * n + i];
    }
}

/* ------------------------------------------------------------------------ */
/* V-cycle and CG, FP32 and FP64 instances                                  */
/* ------------------------------------------------------------------------ */
#define REAL float
#define SFX _f32
#include "oracle_mg_impl.h"
#undef REAL
#undef SFX
#define REAL double
#define SFX _f64
#include "oracle_mg_impl.h"
#undef REAL
#undef SFX

double orc_apply_f32(const orc_mg *mg, const float *x, float *y) { return apply_all_f32(mg, x, y); }

/* The exact coarsest solve alone (the ORC_COARSE_EXACT mode of the V-cycle):
 * x = P B^-1 P b on the coarsest level, B the grounded operator, P the
 * removal of the mean on each closed component; rows with diag <= 0 give 0.
 * Exposed so that tests can pin it against a pseudo-inverse directly. */
void orc_coarse_solve_f64(orc_mg *mg, const double *b, double *x) {
    coarse_exact_f64(mg, &mg->L[mg->nl - 1], x, b);
}

int orc_pcg_f32(orc_mg *mg, const float *b, float *x, float inv, float tol, int max_iters, int stop_mode,
                int coarse_mode, double *res_out) {
    return pcg_core_f32(mg, b, x, inv, (double)tol, max_iters, stop_mode, coarse_mode, res_out, NULL, NULL);
}

int orc_pcg_f64(orc_mg *mg, const double *b, double *x, double tol, int max_iters, int coarse_mode,
                double *res_out) {
    return pcg_core_f64(mg, b, x, 1.0f, tol, max_iters, ORC_STOP_REL_L2, coarse_mode, res_out, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* computeZones (gasphase.cpp:111-158): flood fill over faces with a > 0     */
/* ------------------------------------------------------------------------ */
int orc_compute_zones(int nx, int ny, int nz, const float *ax_, const float *ay_, const float *az_,
                      const uint8_t *blocked, int *zone, uint8_t *zone_open, int max_zones) {
    const int N = nx * ny * nz, sxy = nx * ny;
    int *stack = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    int nzones = 0;
    for (int c = 0; c < N; ++c) zone[c] = -1;
    for (int seed = 0; seed < N; ++seed) {
        if (blocked[seed] || zone[seed] >= 0) continue;
        int z = nzones++;
        uint8_t zo = 0;
        int sp = 0;
        stack[sp++] = seed;
        zone[seed] = z;
        while (sp > 0) {
            int c = stack[--sp];
            int i = c % nx, j = (c / nx) % ny, k = c / sxy;
            float a[6] = {ax_[i + (nx + 1) * (j + ny * k)], ax_[i + 1 + (nx + 1) * (j + ny * k)],
                          ay_[i + nx * (j + (ny + 1) * k)], ay_[i + nx * (j + 1 + (ny + 1) * k)],
                          az_[i + nx * (j + ny * k)], az_[i + nx * (j + ny * (k + 1))]};
            int nb[6] = {i > 0 ? c - 1 : -1, i < nx - 1 ? c + 1 : -1, j > 0 ? c - nx : -1,
                         j < ny - 1 ? c + nx : -1, k > 0 ? c - sxy : -1, k < nz - 1 ? c + sxy : -1};
            for (int d = 0; d < 6; ++d) {
                if (a[d] <= 0) continue;
                if (nb[d] < 0) { if (d != 4) zo = 1; continue; }   /* side or top boundary */
                if (blocked[nb[d]] || zone[nb[d]] >= 0) continue;
                zone[nb[d]] = z;
                stack[sp++] = nb[d];
            }
        }
        if (zone_open && z < max_zones) zone_open[z] = zo;
    }
    free(stack);
    return nzones;
}

/* ------------------------------------------------------------------------ */
/* project(dt) pieces (gasphase.cpp:505-677)                                */
/* ------------------------------------------------------------------------ */
static inline float h_out(float vel, int inflow) { return inflow ? -0.5f * vel * vel : 0.0f; }   /* :516 */

/* RHS and diagonal (gasphase.cpp:528-549) */
void orc_rhs(const orc_project_args *a, const float *u, const float *v, const float *w, const float *S, float *b,
             float *diag, float *p) {
    const int nx = a->nx, ny = a->ny, nz = a->nz;
    const float dx = a->dx, inv = dx * dx / a->dt;
    uint8_t wc[4] = {0, 0, 0, 0};
    if (a->wind_closed) memcpy(wc, a->wind_closed, 4);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = i + nx * (j + ny * k);
                if (a->blocked[c]) { b[c] = 0; diag[c] = 0; p[c] = 0; continue; }
                int a0 = i + (nx + 1) * (j + ny * k), a1 = a0 + 1, b0 = i + nx * (j + (ny + 1) * k),
                    b1 = i + nx * (j + 1 + (ny + 1) * k), c0 = i + nx * (j + ny * k), c1 = i + nx * (j + ny * (k + 1));
                float div = (a->fx[a1] * u[a1] - a->fx[a0] * u[a0] + a->fy[b1] * v[b1] - a->fy[b0] * v[b0] +
                             a->fz[c1] * w[c1] - a->fz[c0] * w[c0]) / dx;
                b[c] = (S[c] - div) * inv;
                float co[6] = {a->ax[a0], a->ax[a1], a->ay[b0], a->ay[b1], a->az[c0], a->az[c1]};
                if (i == 0 && wc[0]) co[0] = 0;
                if (i == nx - 1 && wc[1]) co[1] = 0;
                if (j == 0 && wc[2]) co[2] = 0;
                if (j == ny - 1 && wc[3]) co[3] = 0;
                if (i == 0 && co[0] > 0) b[c] += co[0] * h_out(u[a0], u[a0] > 0);
                if (i == nx - 1 && co[1] > 0) b[c] += co[1] * h_out(u[a1], u[a1] < 0);
                if (j == 0 && co[2] > 0) b[c] += co[2] * h_out(v[b0], v[b0] > 0);
                if (j == ny - 1 && co[3] > 0) b[c] += co[3] * h_out(v[b1], v[b1] < 0);
                if (k == nz - 1 && co[5] > 0) b[c] += co[5] * h_out(w[c1], w[c1] < 0);
                diag[c] = co[0] + co[1] + co[2] + co[3] + co[4] + co[5];
            }
}

/* closed zones: zone mean of H to zero (gasphase.cpp:628-639) */
void orc_zone_mean(const orc_project_args *a, int nzones, float *p) {
    const int N = a->nx * a->ny * a->nz;
    double *sum = (double *)calloc((size_t)(nzones > 0 ? nzones : 1), sizeof(double));
    double *cnt = (double *)calloc((size_t)(nzones > 0 ? nzones : 1), sizeof(double));
    for (int c = 0; c < N; ++c)
        if (a->zone[c] >= 0 && !a->zone_open[a->zone[c]]) { sum[a->zone[c]] += p[c]; cnt[a->zone[c]] += 1; }
    int any = 0;
    for (int z = 0; z < nzones; ++z) if (cnt[z] > 0) { sum[z] /= cnt[z]; any = 1; }
    if (any)
        for (int c = 0; c < N; ++c)
            if (a->zone[c] >= 0 && !a->zone_open[a->zone[c]]) p[c] -= (float)sum[a->zone[c]];
    free(sum);
    free(cnt);
}

static inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (hi < x ? hi : x); }
static inline float maxf(float a, float b) { return a < b ? b : a; }
static inline float minf(float a, float b) { return b < a ? b : a; }

/* velocity correction u -= dt grad H on matrix faces (gasphase.cpp:640-676) */
float orc_correct(const orc_project_args *a, const float *p, float *u, float *v, float *w) {
    const int nx = a->nx, ny = a->ny, nz = a->nz;
    const float s = a->dt / a->dx, VMAX = 40.0f;
    uint8_t wc[4] = {0, 0, 0, 0};
    if (a->wind_closed) memcpy(wc, a->wind_closed, 4);
#define P(i, j, k) p[(i) + nx * ((j) + ny * (k))]
    float vmax = 0;
    for (int k = 0; k < nz; ++k) {
        float vm = 0;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                int f = i + (nx + 1) * (j + ny * k);
                if (a->fx[f] <= 0) { u[f] = 0; continue; }
                if (a->ax[f] > 0 && !((i == 0 || i == nx) && wc[i == 0 ? 0 : 1]))
                    u[f] = clampf(u[f] - s * ((i < nx ? P(i, j, k) : h_out(u[f], u[f] < 0)) -
                                              (i > 0 ? P(i - 1, j, k) : h_out(u[f], u[f] > 0))), -VMAX, VMAX);
                vm = maxf(vm, fabsf(u[f]) * a->fx[f]);
            }
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int f = i + nx * (j + (ny + 1) * k);
                if (a->fy[f] <= 0) { v[f] = 0; continue; }
                if (a->ay[f] > 0 && !((j == 0 || j == ny) && wc[j == 0 ? 2 : 3]))
                    v[f] = clampf(v[f] - s * ((j < ny ? P(i, j, k) : h_out(v[f], v[f] < 0)) -
                                              (j > 0 ? P(i, j - 1, k) : h_out(v[f], v[f] > 0))), -VMAX, VMAX);
                vm = maxf(vm, fabsf(v[f]) * a->fy[f]);
            }
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int f = i + nx * (j + ny * (k + 1));
                if (a->fz[f] <= 0) { w[f] = 0; continue; }
                if (a->az[f] > 0)
                    w[f] = clampf(w[f] - s * ((k + 1 < nz ? P(i, j, k + 1) : h_out(w[f], w[f] < 0)) - P(i, j, k)),
                                  -VMAX, VMAX);
                vm = maxf(vm, fabsf(w[f]) * a->fz[f]);
            }
        vmax = maxf(vmax, vm);
    }
#undef P
    return vmax;
}

void orc_project(orc_mg *mg, const orc_project_args *a, float *u, float *v, float *w, float *p, const float *S,
                 float *b_out, float *diag_out, orc_project_out *out) {
    const int N = a->nx * a->ny * a->nz;
    const float inv = a->dx * a->dx / a->dt;
    float *b = (float *)malloc(sizeof(float) * (size_t)N);
    float *diag = (float *)malloc(sizeof(float) * (size_t)N);
    orc_rhs(a, u, v, w, S, b, diag, p);
    double res = 0, maxnorm = 0, rel = 0;
    int it = pcg_core_f32(mg, b, p, inv, (double)a->tol, a->max_iters, a->stop_mode, a->coarse_mode, &res, &maxnorm,
                          &rel);
    out->iters = it;
    out->residual = (float)maxnorm;   /* pressureResidual = maxRes() (gasphase.cpp:627) */
    out->relres = rel;
    int nzones = 0;
    for (int c = 0; c < N; ++c) if (a->zone[c] + 1 > nzones) nzones = a->zone[c] + 1;
    orc_zone_mean(a, nzones, p);
    out->max_vel = orc_correct(a, p, u, v, w);
    if (b_out) memcpy(b_out, b, sizeof(float) * (size_t)N);
    if (diag_out) memcpy(diag_out, diag, sizeof(float) * (size_t)N);
    free(b);
    free(diag);
}

/* ------------------------------------------------------------------------ */
/* energyAndExpansion (gasphase.cpp:251-271)                                */
/* ------------------------------------------------------------------------ */
float orc_energy_expansion(int n, float dx, float dt, const uint8_t *blocked, const float *Pb, const float *Q,
                           const float *massSrc, float *T, float *S) {
    const float RS = 287.0f, CP = 1100.0f, T_MIN = 200.0f, T_MAX = 2300.0f, D_MAX = 50.0f;
    const float V = dx * dx * dx;
    float md = 0;
    for (int c = 0; c < n; ++c) {
        if (blocked[c]) { S[c] = 0; continue; }
        float rho = Pb[c] / (RS * T[c]);
        float m = rho * V;
        float Tn = clampf(T[c] + Q[c] / (m * CP), T_MIN, T_MAX);
        float rhoN = Pb[c] / (RS * Tn);
        float D = Q[c] / (m * CP * T[c] * dt) + massSrc[c] / (rhoN * V * dt);
        T[c] = Tn;
        S[c] = clampf(D, -D_MAX, D_MAX);
        md = maxf(md, fabsf(S[c]));
    }
    return md;
}

/* ------------------------------------------------------------------------ */
/* advectScalars (gasphase.cpp:685-784)                                     */
/* ------------------------------------------------------------------------ */
static inline float superbee(float r) {   /* gasphase.cpp:35 */
    float a = minf(2 * r, 1.0f), b = minf(r, 2.0f);
    float m = maxf(0.0f, a);
    return maxf(m, b);
}

void orc_advect_scalars(int nx, int ny, int nz, float dx, float dt, const float *u, const float *v, const float *w,
                        const float *fx, const float *fy, const float *fz, const uint8_t *blocked, const float *Pb,
                        const float *nu, const float *const *fields, float *const *out) {
    enum { NF = 10 };
    const float T0 = 293.15f, P0 = 101325.0f, RS = 287.0f, T_MIN = 200.0f, SC_T = 0.5f;
    const float amb[NF] = {T0, 0.232f, 0, 0, 0, 0, 0.0006f, 0.0073f, 0, 0};
    const float s = dt / dx;
    const int sxy = nx * ny;
    const float *T = fields[0];
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = i + nx * (j + ny * k);
                if (blocked[c]) {
                    for (int q = 0; q < NF; ++q) out[q][c] = fields[q][c];
                    continue;
                }
                int fa0 = i + (nx + 1) * (j + ny * k), fa1 = fa0 + 1, fb0 = i + nx * (j + (ny + 1) * k),
                    fb1 = i + nx * (j + 1 + (ny + 1) * k), fc0 = i + nx * (j + ny * k), fc1 = i + nx * (j + ny * (k + 1));
                float un[6] = {-fx[fa0] * u[fa0], fx[fa1] * u[fa1], -fy[fb0] * v[fb0], fy[fb1] * v[fb1],
                               -fz[fc0] * w[fc0], fz[fc1] * w[fc1]};
                float fo[6] = {fx[fa0], fx[fa1], fy[fb0], fy[fb1], fz[fc0], fz[fc1]};
                int id[3] = {i, j, k};
                const int dim[3] = {nx, ny, nz}, st[3] = {1, nx, sxy};
                int nb[6], nb2[6], opp[6];
                for (int d = 0; d < 6; ++d) {
                    int ax = d >> 1, sg = (d & 1) ? 1 : -1;
                    int n1 = id[ax] + sg, n2 = id[ax] + 2 * sg, o1 = id[ax] - sg;
                    nb[d] = (n1 >= 0 && n1 < dim[ax]) ? c + sg * st[ax] : -1;
                    nb2[d] = (n2 >= 0 && n2 < dim[ax]) ? c + 2 * sg * st[ax] : -1;
                    opp[d] = (o1 >= 0 && o1 < dim[ax]) ? c - sg * st[ax] : -1;
                    if (nb[d] >= 0 && blocked[nb[d]]) nb[d] = -2;
                    if (nb2[d] >= 0 && blocked[nb2[d]]) nb2[d] = -2;
                    if (opp[d] >= 0 && blocked[opp[d]]) opp[d] = -2;
                }
                float nuc = nu[c];
                float rc = Pb[c] / (RS * T[c]), rUp[6], mden = rc;
                for (int d = 0; d < 6; ++d) {
                    rUp[d] = 0;
                    if (fo[d] <= 0 || nb[d] == -2) continue;
                    float tn = nb[d] >= 0 ? T[nb[d]] : T0, tc = T[c], tf;
                    if (un[d] > 0) {
                        float tu2 = opp[d] >= 0 ? T[opp[d]] : (opp[d] == -1 ? T0 : tc);
                        float den = tn - tc;
                        tf = tc + (fabsf(den) > 1e-6f ? 0.5f * superbee((tc - tu2) / den) * den : 0.0f);
                    } else {
                        float tu2 = nb2[d] >= 0 ? T[nb2[d]] : (nb2[d] == -1 ? T0 : tn);
                        float den = tc - tn;
                        tf = tn + (fabsf(den) > 1e-6f ? 0.5f * superbee((tn - tu2) / den) * den : 0.0f);
                    }
                    float pu = un[d] > 0 ? Pb[c] : (nb[d] >= 0 ? Pb[nb[d]] : P0);
                    rUp[d] = pu / (RS * maxf(tf, T_MIN));
                    mden -= s * rUp[d] * un[d];
                }
                int conservative = mden > 0.25f * rc;
                for (int q = 0; q < NF; ++q) {
                    const float *F = fields[q];
                    float fc = F[c];
                    float lo = fc, hi = fc, acc = 0, macc = 0;
                    for (int d = 0; d < 6; ++d) {
                        if (fo[d] <= 0 || nb[d] == -2) continue;
                        float fn = nb[d] >= 0 ? F[nb[d]] : amb[q];
                        lo = minf(lo, fn);
                        hi = maxf(hi, fn);
                        float flux = un[d];
                        float face = fc;
                        if (flux > 0) {
                            float fu2 = opp[d] >= 0 ? F[opp[d]] : (opp[d] == -1 ? amb[q] : fc);
                            float den = fn - fc;
                            if (fabsf(den) > 1e-12f) face += 0.5f * superbee((fc - fu2) / den) * den;
                        } else if (flux < 0) {
                            float fu2 = nb2[d] >= 0 ? F[nb2[d]] : (nb2[d] == -1 ? amb[q] : fn);
                            float den = fc - fn;
                            face = fn;
                            if (fabsf(den) > 1e-12f) face += 0.5f * superbee((fn - fu2) / den) * den;
                        }
                        acc -= flux * (face - fc);
                        macc -= rUp[d] * flux * face;
                        if (q > 0) {
                            float nun = nb[d] >= 0 ? nu[nb[d]] : nuc;
                            float D = 0.5f * (nuc + nun) / SC_T + 2.0e-5f;
                            float rf = nb[d] >= 0 ? 0.5f * (rc + Pb[nb[d]] / (RS * T[nb[d]])) : rc;
                            acc += fo[d] * D * (fn - fc) / dx;
                            macc += fo[d] * rf * D * (fn - fc) / dx;
                        }
                    }
                    float val = conservative ? (rc * fc + s * macc) / mden : fc + s * acc;
                    out[q][c] = clampf(val, lo, hi);
                }
            }
}
