/*
 * oracle_mg_impl.h -- V-cycle and PCG of the oracle, written once for a
 * generic REAL and included twice by oracle.c (REAL=float, SFX=_f32 and
 * REAL=double, SFX=_f64).  TEST INFRASTRUCTURE ONLY.
 *
 * The float instance follows the reference operation for operation so that it
 * reproduces src/sim/pressure_mg.cpp:195-300 and src/sim/gasphase.cpp:551-627
 * bit for bit (same summation order, FP64 dot products per k slice summed in
 * order as sumSlices does, gasphase.cpp:41-49).  The double instance is the
 * same algorithm in FP64 (the FP64 oracle).
 */
#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* structured level: one coloured Gauss-Seidel half sweep (pressure_mg.cpp:195-216) */
static void FN(gs_color)(const orc_level *L, REAL *x, const REAL *b, int color) {
    const int nx = L->nx, ny = L->ny, nz = L->nz, sx = nx, sxy = nx * ny;
    for (int r = 0; r < ny * nz; ++r) {
        int j = r % ny, k = r / ny;
        int i0 = (j + k + color) & 1;
        for (int i = i0; i < nx; i += 2) {
            int c = i + sx * (j + ny * k);
            REAL dg = L->diag[c];
            if (dg <= 0) { x[c] = 0; continue; }
            REAL s = b[c];
            if (i > 0) s += (REAL)L->ax[i + (nx + 1) * (j + ny * k)] * x[c - 1];
            if (i < nx - 1) s += (REAL)L->ax[i + 1 + (nx + 1) * (j + ny * k)] * x[c + 1];
            if (j > 0) s += (REAL)L->ay[i + nx * (j + (ny + 1) * k)] * x[c - sx];
            if (j < ny - 1) s += (REAL)L->ay[i + nx * (j + 1 + (ny + 1) * k)] * x[c + sx];
            if (k > 0) s += (REAL)L->az[i + nx * (j + ny * k)] * x[c - sxy];
            if (k < nz - 1) s += (REAL)L->az[i + nx * (j + ny * (k + 1))] * x[c + sxy];
            x[c] = s / dg;
        }
    }
}

/* r = b - A x on the structured level (pressure_mg.cpp:217-235) */
static void FN(residual_grid)(const orc_level *L, const REAL *x, const REAL *b, REAL *rr) {
    const int nx = L->nx, ny = L->ny, nz = L->nz, sx = nx, sxy = nx * ny;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = i + sx * (j + ny * k);
                if (L->diag[c] <= 0) { rr[c] = 0; continue; }
                REAL s = b[c] - (REAL)L->diag[c] * x[c];
                if (i > 0) s += (REAL)L->ax[i + (nx + 1) * (j + ny * k)] * x[c - 1];
                if (i < nx - 1) s += (REAL)L->ax[i + 1 + (nx + 1) * (j + ny * k)] * x[c + 1];
                if (j > 0) s += (REAL)L->ay[i + nx * (j + (ny + 1) * k)] * x[c - sx];
                if (j < ny - 1) s += (REAL)L->ay[i + nx * (j + 1 + (ny + 1) * k)] * x[c + sx];
                if (k > 0) s += (REAL)L->az[i + nx * (j + ny * k)] * x[c - sxy];
                if (k < nz - 1) s += (REAL)L->az[i + nx * (j + ny * (k + 1))] * x[c + sxy];
                rr[c] = s;
            }
}

/* damped Jacobi on a sparse level, x <- t (pressure_mg.cpp:237-247) */
static void FN(jacobi)(const orc_level *L, REAL **x, REAL **t, const REAL *b, REAL omega) {
    REAL *X = *x, *T = *t;
    for (int I = 0; I < L->n; ++I) {
        if (L->diag[I] <= 0) { T[I] = 0; continue; }
        REAL s = b[I];
        for (int q = L->rowStart[I]; q < L->rowStart[I + 1]; ++q) s += (REAL)L->val[q] * X[L->col[q]];
        T[I] = (1 - omega) * X[I] + omega * s / (REAL)L->diag[I];
    }
    *x = T;
    *t = X;
}

static void FN(residual_sparse)(const orc_level *L, const REAL *x, const REAL *b, REAL *rr) {
    for (int I = 0; I < L->n; ++I) {
        if (L->diag[I] <= 0) { rr[I] = 0; continue; }
        REAL s = b[I] - (REAL)L->diag[I] * x[I];
        for (int q = L->rowStart[I]; q < L->rowStart[I + 1]; ++q) s += (REAL)L->val[q] * x[L->col[q]];
        rr[I] = s;
    }
}

static void FN(gs_sparse_sweep)(const orc_level *L, REAL *x, const REAL *b, int forward) {
    for (int t = 0; t < L->n; ++t) {
        int I = forward ? t : L->n - 1 - t;
        if (L->diag[I] <= 0) { x[I] = 0; continue; }
        REAL s = b[I];
        for (int q = L->rowStart[I]; q < L->rowStart[I + 1]; ++q) s += (REAL)L->val[q] * x[L->col[q]];
        x[I] = s / (REAL)L->diag[I];
    }
}

static void FN(coarse_exact)(orc_mg *mg, const orc_level *L, REAL *x, const REAL *b) {
    ensure_chol(mg);
    const int n = L->n;
    double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int I = 0; I < n; ++I) y[I] = mg->cmask[I] ? 0.0 : (double)b[I];
    /* P B^-1 P: the pseudo-inverse on closed components (no constant offset) */
    project_closed(mg, n, y);
    chol_solve(mg->chol, n, y);
    project_closed(mg, n, y);
    for (int I = 0; I < n; ++I) x[I] = mg->cmask[I] ? (REAL)0 : (REAL)y[I];
    free(y);
}

/* z = M^-1 r with one symmetric V-cycle (pressure_mg.cpp:270-300) */
void FN(orc_precondition)(orc_mg *mg, const REAL *r, REAL *z, int coarse_mode) {
    const int nl = mg->nl;
    REAL *X[ORC_MAX_LEVELS], *B[ORC_MAX_LEVELS], *R[ORC_MAX_LEVELS], *T[ORC_MAX_LEVELS];
    for (int l = 0; l < nl; ++l) {
        size_t n = (size_t)(mg->L[l].n > 0 ? mg->L[l].n : 1);
        X[l] = (REAL *)calloc(n, sizeof(REAL));
        B[l] = (REAL *)calloc(n, sizeof(REAL));
        R[l] = (REAL *)calloc(n, sizeof(REAL));
        T[l] = (REAL *)calloc(n, sizeof(REAL));
    }
    memcpy(B[0], r, sizeof(REAL) * (size_t)mg->L[0].n);
    for (int l = 0; l < nl; ++l) {
        orc_level *L = &mg->L[l];
        memset(X[l], 0, sizeof(REAL) * (size_t)L->n);
        int coarsest = l == nl - 1 || L->agg == NULL;
        if (coarsest) {
            if (coarse_mode == ORC_COARSE_EXACT) {
                FN(coarse_exact)(mg, L, X[l], B[l]);
            } else {
                for (int it = 0; it < 20; ++it) {
                    FN(gs_sparse_sweep)(L, X[l], B[l], 1);
                    FN(gs_sparse_sweep)(L, X[l], B[l], 0);
                }
            }
            break;
        }
        if (l == 0) {
            FN(gs_color)(L, X[0], B[0], 0);
            FN(gs_color)(L, X[0], B[0], 1);
            FN(residual_grid)(L, X[0], B[0], R[0]);
        } else {
            FN(jacobi)(L, &X[l], &T[l], B[l], (REAL)0.7f);
            FN(jacobi)(L, &X[l], &T[l], B[l], (REAL)0.7f);
            FN(residual_sparse)(L, X[l], B[l], R[l]);
        }
        orc_level *C = &mg->L[l + 1];
        memset(B[l + 1], 0, sizeof(REAL) * (size_t)C->n);
        for (int i = 0; i < L->n; ++i)
            if (L->agg[i] >= 0) B[l + 1][L->agg[i]] += R[l][i];
    }
    int last = 0;
    while (last + 1 < nl && mg->L[last].agg != NULL) ++last;
    for (int l = last - 1; l >= 0; --l) {
        orc_level *L = &mg->L[l];
        for (int i = 0; i < L->n; ++i)
            if (L->agg[i] >= 0) X[l][i] += X[l + 1][L->agg[i]];
        if (l == 0) {
            FN(gs_color)(L, X[0], B[0], 1);
            FN(gs_color)(L, X[0], B[0], 0);
        } else {
            FN(jacobi)(L, &X[l], &T[l], B[l], (REAL)0.7f);
            FN(jacobi)(L, &X[l], &T[l], B[l], (REAL)0.7f);
        }
    }
    memcpy(z, X[0], sizeof(REAL) * (size_t)mg->L[0].n);
    for (int l = 0; l < nl; ++l) { free(X[l]); free(B[l]); free(R[l]); free(T[l]); }
}

/* y = A x for one k slice, returns sum y*x (gasphase.cpp:551-570).  The
 * coefficients are the masked level-0 faces, whose values equal coefs(). */
static double FN(apply_slice)(const orc_level *L, const REAL *x, REAL *y, int k) {
    const int nx = L->nx, ny = L->ny, nz = L->nz, sxy = nx * ny;
    double dot = 0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int c = i + nx * (j + ny * k);
            REAL dg = L->diag[c];
            if (dg <= 0) { y[c] = 0; continue; }
            REAL a0 = L->ax[i + (nx + 1) * (j + ny * k)], a1 = L->ax[i + 1 + (nx + 1) * (j + ny * k)];
            REAL a2 = L->ay[i + nx * (j + (ny + 1) * k)], a3 = L->ay[i + nx * (j + 1 + (ny + 1) * k)];
            REAL a4 = L->az[i + nx * (j + ny * k)], a5 = L->az[i + nx * (j + ny * (k + 1))];
            /* the diagonal is the sum of the faces: in FP32 this is bitwise the
             * stored diag (same order of float adds); in FP64 it is the exact sum,
             * so A 1 = 0 holds exactly on closed zones with fractional faces */
            REAL s = (a0 + a1 + a2 + a3 + a4 + a5) * x[c];
            if (i > 0) s -= a0 * x[c - 1];
            if (i < nx - 1) s -= a1 * x[c + 1];
            if (j > 0) s -= a2 * x[c - nx];
            if (j < ny - 1) s -= a3 * x[c + nx];
            if (k > 0) s -= a4 * x[c - sxy];
            if (k < nz - 1) s -= a5 * x[c + sxy];
            y[c] = s;
            dot += (double)s * x[c];
        }
    return dot;
}

static double FN(apply_all)(const orc_mg *mg, const REAL *x, REAL *y) {
    double s = 0;
    for (int k = 0; k < mg->L[0].nz; ++k) s += FN(apply_slice)(&mg->L[0], x, y, k);
    return s;
}

static double FN(dot_slices)(const orc_level *L, const REAL *a, const REAL *b) {
    const int sxy = L->nx * L->ny;
    double s = 0;
    for (int k = 0; k < L->nz; ++k) {
        double part = 0;
        for (int c = k * sxy; c < (k + 1) * sxy; ++c) part += (double)a[c] * b[c];
        s += part;
    }
    return s;
}

/* largest |r| back in 1/s (gasphase.cpp:592-602) */
static double FN(maxres)(const REAL *r, int n, int sxy, int nz, float inv) {
    (void)n;
    float m = 0;
    for (int k = 0; k < nz; ++k) {
        float part = 0;
        for (int c = k * sxy; c < (k + 1) * sxy; ++c) {
            float a = (float)fabs((double)r[c]);
            part = part > a ? part : a;
        }
        m = m > part ? m : part;
    }
    return (double)(m / inv);
}

/* the CG loop of project() (gasphase.cpp:571-627) on a given b, warm start x */
static int FN(pcg_core)(orc_mg *mg, const REAL *b, REAL *x, float inv, double tol, int max_iters, int stop_mode,
                        int coarse_mode, double *res_out, double *maxnorm_out, double *rel_out) {
    const orc_level *L0 = &mg->L[0];
    const int n = L0->n, sxy = L0->nx * L0->ny, nz = L0->nz;
    REAL *r = (REAL *)calloc((size_t)n, sizeof(REAL));
    REAL *z = (REAL *)calloc((size_t)n, sizeof(REAL));
    REAL *d = (REAL *)calloc((size_t)n, sizeof(REAL));
    REAL *q = (REAL *)calloc((size_t)n, sizeof(REAL));
    for (int k = 0; k < nz; ++k) {
        FN(apply_slice)(L0, x, q, k);
        for (int c = k * sxy; c < (k + 1) * sxy; ++c) r[c] = L0->diag[c] > 0 ? b[c] - q[c] : (REAL)0;
    }
    double bnorm = sqrt(FN(dot_slices)(L0, b, b));
    FN(orc_precondition)(mg, r, z, coarse_mode);
    memcpy(d, z, sizeof(REAL) * (size_t)n);
    double rz = FN(dot_slices)(L0, r, z);
    /* stop quantity: max|r|/inv in 1/s (gasphase.cpp:592-602) or relative L2 */
#define STOPQ() (stop_mode == ORC_STOP_MAXNORM ? FN(maxres)(r, n, sxy, nz, inv) \
                                               : (bnorm > 0 ? sqrt(FN(dot_slices)(L0, r, r)) / bnorm : 0.0))
    int it = 0;
    double res = STOPQ();
    while (res > tol && it < max_iters && rz > 0) {
        double dq = FN(apply_all)(mg, d, q);
        if (dq <= 0) break;
        double alpha = rz / dq;
        for (int c = 0; c < n; ++c) {
            if (L0->diag[c] <= 0) continue;
            x[c] += (REAL)(alpha * d[c]);
            r[c] -= (REAL)(alpha * q[c]);
        }
        FN(orc_precondition)(mg, r, z, coarse_mode);
        double rzNew = FN(dot_slices)(L0, r, z);
        double beta = rzNew / rz;
        rz = rzNew;
        for (int c = 0; c < n; ++c) d[c] = z[c] + (REAL)(beta * d[c]);
        ++it;
        res = STOPQ();
    }
#undef STOPQ
    if (res_out) *res_out = res;
    if (maxnorm_out) *maxnorm_out = FN(maxres)(r, n, sxy, nz, inv);
    if (rel_out) *rel_out = bnorm > 0 ? sqrt(FN(dot_slices)(L0, r, r)) / bnorm : 0.0;
    free(r); free(z); free(d); free(q);
    return it;
}

double FN(orc_true_relres)(const orc_mg *mg, const REAL *b, const REAL *x) {
    const orc_level *L = &mg->L[0];
    const int nx = L->nx, ny = L->ny, nz = L->nz, sxy = nx * ny;
    double rr = 0, bb = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = i + nx * (j + ny * k);
                bb += (double)b[c] * (double)b[c];
                if (L->diag[c] <= 0) continue;
                const double dsum = (double)L->ax[i + (nx + 1) * (j + ny * k)] + (double)L->ax[i + 1 + (nx + 1) * (j + ny * k)] +
                                    (double)L->ay[i + nx * (j + (ny + 1) * k)] + (double)L->ay[i + nx * (j + 1 + (ny + 1) * k)] +
                                    (double)L->az[i + nx * (j + ny * k)] + (double)L->az[i + nx * (j + ny * (k + 1))];
                double s = (sizeof(REAL) == sizeof(double) ? dsum : (double)L->diag[c]) * (double)x[c];
                if (i > 0) s -= (double)L->ax[i + (nx + 1) * (j + ny * k)] * (double)x[c - 1];
                if (i < nx - 1) s -= (double)L->ax[i + 1 + (nx + 1) * (j + ny * k)] * (double)x[c + 1];
                if (j > 0) s -= (double)L->ay[i + nx * (j + (ny + 1) * k)] * (double)x[c - nx];
                if (j < ny - 1) s -= (double)L->ay[i + nx * (j + 1 + (ny + 1) * k)] * (double)x[c + nx];
                if (k > 0) s -= (double)L->az[i + nx * (j + ny * k)] * (double)x[c - sxy];
                if (k < nz - 1) s -= (double)L->az[i + nx * (j + ny * (k + 1))] * (double)x[c + sxy];
                double e = (double)b[c] - s;
                rr += e * e;
            }
    return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

#undef FN
#undef CAT
#undef CAT_
