// hierarchy.cpp -- host-side multigrid set-up for the HIP pressure solver.
// Algorithm: src/sim/pressure_mg.cpp:28-191 (connected aggregation, Galerkin
// coarse operators); zones: src/sim/gasphase.cpp:111-158.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <thread>

#include "hierarchy.hpp"

namespace phg {
namespace {

// Host threads for the dense coarsest inverse (O(n^3) with n up to 2048).
// Every element is computed by one thread with the same operation order as a
// serial loop, so the result does not depend on the thread count.
int host_threads(int work) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    return std::max(1, std::min({hw, 16, work}));
}

// f(i) for i in [0, n), rows handed out one at a time (unbalanced triangles)
template <class F>
void parallel_rows(int n, F f) {
    const int T = host_threads(n / 8);
    if (T <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

// spinning barrier for a fixed team (one crossing per Cholesky column)
struct SpinBarrier {
    const int n;
    std::atomic<int> count{0};
    std::atomic<int> phase{0};
    explicit SpinBarrier(int n_) : n(n_) {}
    void wait() {
        const int ph = phase.load(std::memory_order_acquire);
        if (count.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
            count.store(0, std::memory_order_relaxed);
            phase.store(ph + 1, std::memory_order_release);
        } else {
            while (phase.load(std::memory_order_acquire) == ph) std::this_thread::yield();
        }
    }
};

// Disjoint sets whose representative is always the smallest member, so the
// partition and the representative do not depend on the order of joins.
struct MinSets {
    std::vector<int> up;
    explicit MinSets(int n) : up(n) {
        for (int i = 0; i < n; ++i) up[i] = i;
    }
    int root(int a) {
        int r = a;
        while (up[r] != r) r = up[r];
        while (up[a] != r) {
            int nxt = up[a];
            up[a] = r;
            a = nxt;
        }
        return r;
    }
    void merge(int a, int b) {
        a = root(a);
        b = root(b);
        if (a == b) return;
        if (a < b) up[b] = a;
        else up[a] = b;
    }
};

// Row builder keeping couplings in first-insertion order (the order fixes the
// FP32 summation of repeated couplings, as in the reference).
struct Rows {
    std::vector<std::vector<std::pair<int, float>>> r;
    explicit Rows(int n) : r(n) {}
    void add(int row, int col, float w) {
        for (auto& e : r[row])
            if (e.first == col) {
                e.second += w;
                return;
            }
        r[row].push_back({col, w});
    }
};

void to_csr(HostLevel& C, Rows& rows, const std::vector<double>& dg) {
    C.diag.resize(C.n);
    C.rowStart.assign(C.n + 1, 0);
    for (int I = 0; I < C.n; ++I) {
        C.diag[I] = (float)std::max(dg[I], 0.0);
        C.rowStart[I + 1] = C.rowStart[I] + (int)rows.r[I].size();
    }
    C.col.clear();
    C.val.clear();
    for (int I = 0; I < C.n; ++I)
        for (auto& e : rows.r[I]) {
            C.col.push_back(e.first);
            C.val.push_back(e.second);
        }
}

// Partition level F into connected pieces of 2x2x2 groups of its block
// coordinates; returns the coarse count, fills agg and the coarse block coords.
int group_pieces(const HostLevel& F, std::vector<int>& agg, HostLevel& C) {
    MinSets s(F.n);
    for (int I = 0; I < F.n; ++I)
        for (int e = F.rowStart[I]; e < F.rowStart[I + 1]; ++e) {
            int J = F.col[e];
            if (J > I && (F.bi[I] >> 1) == (F.bi[J] >> 1) && (F.bj[I] >> 1) == (F.bj[J] >> 1) &&
                (F.bk[I] >> 1) == (F.bk[J] >> 1))
                s.merge(I, J);
        }
    agg.assign(F.n, -1);
    std::vector<int> id(F.n, -1);
    C = HostLevel();
    for (int I = 0; I < F.n; ++I) {
        int r = s.root(I);
        if (id[r] < 0) {
            id[r] = C.n++;
            C.bi.push_back(F.bi[I] >> 1);
            C.bj.push_back(F.bj[I] >> 1);
            C.bk.push_back(F.bk[I] >> 1);
        }
        agg[I] = id[r];
    }
    return C.n;
}

void galerkin_sparse(const HostLevel& F, HostLevel& C) {
    Rows rows(C.n);
    std::vector<double> dg(C.n, 0.0);
    for (int I = 0; I < F.n; ++I) dg[F.agg[I]] += F.diag[I];
    for (int I = 0; I < F.n; ++I)
        for (int e = F.rowStart[I]; e < F.rowStart[I + 1]; ++e) {
            int A = F.agg[I], B = F.agg[F.col[e]];
            if (A == B) dg[A] -= F.val[e];   // internal coupling: seen from both ends
            else rows.add(A, B, F.val[e]);
        }
    to_csr(C, rows, dg);
    C.leak.assign(C.n, 0);
    for (int I = 0; I < F.n; ++I)
        if (F.leak[I]) C.leak[F.agg[I]] = 1;
}

void make_ell(HostLevel& L) {
    int k = 0;
    for (int I = 0; I < L.n; ++I) k = std::max(k, L.rowStart[I + 1] - L.rowStart[I]);
    k = std::max(k, 1);   // at least one (padding) entry: the kernels take the first eight without a test
    L.ellk = k;
    L.ell_col.assign((size_t)k * L.n, 0);
    L.ell_val.assign((size_t)k * L.n, 0.0f);
    for (int I = 0; I < L.n; ++I) {
        int q = 0;
        for (int e = L.rowStart[I]; e < L.rowStart[I + 1]; ++e, ++q) {
            L.ell_col[(size_t)q * L.n + I] = L.col[e];
            L.ell_val[(size_t)q * L.n + I] = L.val[e];
        }
        for (; q < k; ++q) L.ell_col[(size_t)q * L.n + I] = I;
    }
}

void make_children(HostLevel& F, int ncoarse) {
    F.cptr.assign(ncoarse + 1, 0);
    for (int i = 0; i < F.n; ++i)
        if (F.agg[i] >= 0) F.cptr[F.agg[i] + 1]++;
    for (int I = 0; I < ncoarse; ++I) F.cptr[I + 1] += F.cptr[I];
    F.cidx.assign(F.cptr[ncoarse], 0);
    std::vector<int> fill(F.cptr.begin(), F.cptr.end() - 1);
    for (int i = 0; i < F.n; ++i)   // ascending: the gather sums in the reference's serial order
        if (F.agg[i] >= 0) F.cidx[fill[F.agg[i]]++] = i;
}

// Dense inverse of the coarsest operator.  A closed component (no member on
// an open boundary) has the constant vector as null space; adding s*1*1^T on
// it gives an SPD matrix whose inverse acts as the pseudo-inverse on vectors
// of zero component sum.  Rows with diag <= 0 are isolated and held at zero.
// forward sweep: rows in ascending order, a row after its lower neighbours;
// backward sweep: descending, after its higher neighbours
void gs_schedule(HostHierarchy& H) {
    const HostLevel& C = H.L.back();
    const int n = C.n;
    H.gs_ord.assign(2 * (size_t)n, 0);
    H.gs_lptr.clear();
    std::vector<int> lev(n, 0);
    for (int dir = 0; dir < 2; ++dir) {
        int nlev = 0;
        for (int t = 0; t < n; ++t) {
            const int I = dir == 0 ? t : n - 1 - t;
            int v = 0;
            for (int e = C.rowStart[I]; e < C.rowStart[I + 1]; ++e) {
                const int J = C.col[e];
                if (dir == 0 ? J < I : J > I) v = std::max(v, lev[J] + 1);
            }
            lev[I] = v;
            nlev = std::max(nlev, v + 1);
        }
        std::vector<int> start(nlev + 1, 0);
        for (int I = 0; I < n; ++I) start[lev[I] + 1]++;
        for (int l = 0; l < nlev; ++l) start[l + 1] += start[l];
        std::vector<int> fill(start.begin(), start.end() - 1);
        int* ord = H.gs_ord.data() + (size_t)dir * n;
        for (int I = 0; I < n; ++I) ord[fill[lev[I]]++] = I;
        H.gs_lptr.insert(H.gs_lptr.end(), start.begin(), start.end());
        (dir == 0 ? H.gs_nf : H.gs_nb) = nlev;
    }
}

void dense_inverse(HostHierarchy& H) {
    const HostLevel& L = H.L.back();
    const int n = L.n;
    H.nc = n;
    H.cmask.assign(n, 0);
    std::vector<double> A((size_t)n * n, 0.0);
    MinSets comp(std::max(n, 1));
    for (int I = 0; I < n; ++I) {
        if (L.diag[I] <= 0) {
            H.cmask[I] = 1;
            A[(size_t)I * n + I] = 1.0;
            continue;
        }
        A[(size_t)I * n + I] += L.diag[I];
        for (int e = L.rowStart[I]; e < L.rowStart[I + 1]; ++e) {
            A[(size_t)I * n + L.col[e]] -= L.val[e];
            if (L.diag[L.col[e]] > 0) comp.merge(I, L.col[e]);
        }
    }
    std::vector<int> open(n, 0), cnt(n, 0);
    std::vector<double> dsum(n, 0.0);
    for (int I = 0; I < n; ++I) {
        if (H.cmask[I]) continue;
        int r = comp.root(I);
        open[r] |= L.leak[I];
        cnt[r]++;
        dsum[r] += L.diag[I];
    }
    std::vector<std::vector<int>> members(n);
    for (int I = 0; I < n; ++I)
        if (!H.cmask[I]) members[comp.root(I)].push_back(I);
    for (int r = 0; r < n; ++r) {
        if (members[r].empty() || open[r]) continue;
        double s = dsum[r] / ((double)cnt[r] * cnt[r]);
        for (int I : members[r])
            for (int J : members[r]) A[(size_t)I * n + J] += s;
    }
    // Cholesky A = G G^T (lower G stored in A), column by column; the rows
    // below the diagonal of a column are shared out among the team, and every
    // thread forms the column's diagonal itself from the finished row j.  A
    // thread runs the sums of four of its rows side by side, each in the
    // order of k, so they overlap without changing a bit.
    {
        const int T = host_threads(n / 64);
        SpinBarrier bar(T);
        auto team = [&](int t) {
            for (int j = 0; j < n; ++j) {
                const double* aj = &A[(size_t)j * n];
                double d = aj[j];
                for (int k = 0; k < j; ++k) d -= aj[k] * aj[k];
                d = d > 0 ? std::sqrt(d) : 1e-300;
                int i = j + 1 + t;
                for (; i + 3 * T < n; i += 4 * T) {
                    double* a0 = &A[(size_t)i * n];
                    double* a1 = &A[(size_t)(i + T) * n];
                    double* a2 = &A[(size_t)(i + 2 * T) * n];
                    double* a3 = &A[(size_t)(i + 3 * T) * n];
                    double s0 = a0[j], s1 = a1[j], s2 = a2[j], s3 = a3[j];
                    for (int k = 0; k < j; ++k) {
                        const double b = aj[k];
                        s0 -= a0[k] * b;
                        s1 -= a1[k] * b;
                        s2 -= a2[k] * b;
                        s3 -= a3[k] * b;
                    }
                    a0[j] = s0 / d;
                    a1[j] = s1 / d;
                    a2[j] = s2 / d;
                    a3[j] = s3 / d;
                }
                for (; i < n; i += T) {
                    double* ai = &A[(size_t)i * n];
                    double s = ai[j];
                    for (int k = 0; k < j; ++k) s -= ai[k] * aj[k];
                    ai[j] = s / d;
                }
                bar.wait();
                if (t == 0) A[(size_t)j * n + j] = d;
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < T; ++t) pool.emplace_back(team, t);
        team(0);
        for (auto& th : pool) th.join();
    }
    // G^-1 (lower, columns independent), then A^-1 = G^-T G^-1 (rows
    // independent).  G^-1 is kept transposed (Gt[j][i] = G^-1[i][j]) so both
    // loops run along contiguous rows; the sums keep the order of k.
    std::vector<double> Gt((size_t)n * n, 0.0);
    // four rows at a time: their sums over the finished k < i side by side,
    // then the terms of the rows just finished, in the order of k
    parallel_rows(n, [&](int j) {
        double* gj = &Gt[(size_t)j * n];
        gj[j] = 1.0 / A[(size_t)j * n + j];
        int i = j + 1;
        for (; i + 3 < n; i += 4) {
            const double* a0 = &A[(size_t)i * n];
            const double* a1 = &A[(size_t)(i + 1) * n];
            const double* a2 = &A[(size_t)(i + 2) * n];
            const double* a3 = &A[(size_t)(i + 3) * n];
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            for (int k = j; k < i; ++k) {
                const double g = gj[k];
                s0 -= a0[k] * g;
                s1 -= a1[k] * g;
                s2 -= a2[k] * g;
                s3 -= a3[k] * g;
            }
            gj[i] = s0 / a0[i];
            s1 -= a1[i] * gj[i];
            gj[i + 1] = s1 / a1[i + 1];
            s2 -= a2[i] * gj[i];
            s2 -= a2[i + 1] * gj[i + 1];
            gj[i + 2] = s2 / a2[i + 2];
            s3 -= a3[i] * gj[i];
            s3 -= a3[i + 1] * gj[i + 1];
            s3 -= a3[i + 2] * gj[i + 2];
            gj[i + 3] = s3 / a3[i + 3];
        }
        for (; i < n; ++i) {
            const double* ai = &A[(size_t)i * n];
            double s = 0;
            for (int k = j; k < i; ++k) s -= ai[k] * gj[k];
            gj[i] = s / ai[i];
        }
    });
    H.cinv.assign((size_t)n * n, 0.0);
    parallel_rows(n, [&](int i) {
        const double* gi = &Gt[(size_t)i * n];
        double* ci = &H.cinv[(size_t)i * n];
        int j = 0;
        for (; j + 3 <= i; j += 4) {
            const double* g0 = &Gt[(size_t)j * n];
            const double* g1 = g0 + n;
            const double* g2 = g1 + n;
            const double* g3 = g2 + n;
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            for (int k = i; k < n; ++k) {
                const double g = gi[k];
                s0 += g * g0[k];
                s1 += g * g1[k];
                s2 += g * g2[k];
                s3 += g * g3[k];
            }
            ci[j] = s0;
            ci[j + 1] = s1;
            ci[j + 2] = s2;
            ci[j + 3] = s3;
        }
        for (; j <= i; ++j) {
            const double* gj = &Gt[(size_t)j * n];
            double s = 0;
            for (int k = i; k < n; ++k) s += gi[k] * gj[k];
            ci[j] = s;
        }
    });
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) H.cinv[(size_t)j * n + i] = H.cinv[(size_t)i * n + j];
    for (int I = 0; I < n; ++I)
        if (H.cmask[I])
            for (int J = 0; J < n; ++J) H.cinv[(size_t)I * n + J] = H.cinv[(size_t)J * n + I] = 0.0;
    // On a closed component project both sides onto vectors of zero component
    // sum: P B^-1 P is the pseudo-inverse of A there, so the coarse correction
    // carries no constant offset (A's null space) whatever rounding left in
    // the restricted residual.
    for (int r = 0; r < n; ++r) {
        const std::vector<int>& m = members[r];
        if (m.empty() || open[r]) continue;
        const double inv_c = 1.0 / (double)m.size();
        std::vector<double> rowmean(m.size(), 0.0);
        double total = 0.0;
        for (size_t a = 0; a < m.size(); ++a) {
            double acc = 0.0;
            for (int J : m) acc += H.cinv[(size_t)m[a] * n + J];
            rowmean[a] = acc * inv_c;
            total += acc;
        }
        total *= inv_c * inv_c;
        for (size_t a = 0; a < m.size(); ++a)
            for (size_t b2 = 0; b2 < m.size(); ++b2)
                H.cinv[(size_t)m[a] * n + m[b2]] += total - rowmean[a] - rowmean[b2];   // symmetric: col mean = row mean
    }
    // exactly symmetric (the update above rounds the two halves differently):
    // the device reads column i as row i
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double v = 0.5 * (H.cinv[(size_t)i * n + j] + H.cinv[(size_t)j * n + i]);
            H.cinv[(size_t)i * n + j] = H.cinv[(size_t)j * n + i] = v;
        }
}

}  // namespace

bool build_hierarchy(int nx, int ny, int nz, const float* ax, const float* ay, const float* az,
                     const uint8_t* blocked, const uint8_t* wind_closed, int dense_max, HostHierarchy& H,
                     std::string* err) {
    if (nx < 2 || ny < 2 || nz < 2) {
        if (err) *err = "grid must be at least 2 cells in each direction";
        return false;
    }
    H = HostHierarchy();
    const int n0 = nx * ny * nz, sxy = nx * ny;
    auto xf = [&](int i, int j, int k) { return i + (nx + 1) * (j + ny * k); };
    auto yf = [&](int i, int j, int k) { return i + nx * (j + (ny + 1) * k); };
    auto zf = [&](int i, int j, int k) { return i + nx * (j + ny * k); };
    H.L.emplace_back();
    HostLevel& L0 = H.L[0];
    L0.n = n0;
    L0.nx = nx;
    L0.ny = ny;
    L0.nz = nz;
    L0.ax.assign(ax, ax + (size_t)(nx + 1) * ny * nz);
    L0.ay.assign(ay, ay + (size_t)nx * (ny + 1) * nz);
    L0.az.assign(az, az + (size_t)nx * ny * (nz + 1));
    bool wc[4] = {false, false, false, false};
    if (wind_closed)
        for (int s = 0; s < 4; ++s) wc[s] = wind_closed[s] != 0;
    for (int k = 0; k < nz; ++k) {
        for (int j = 0; j < ny; ++j) {
            if (wc[0]) L0.ax[xf(0, j, k)] = 0;
            if (wc[1]) L0.ax[xf(nx, j, k)] = 0;
        }
        for (int i = 0; i < nx; ++i) {
            if (wc[2]) L0.ay[yf(i, 0, k)] = 0;
            if (wc[3]) L0.ay[yf(i, ny, k)] = 0;
        }
    }
    L0.diag.assign(n0, 0.0f);
    L0.leak.assign(n0, 0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int c = i + nx * (j + ny * k);
                if (blocked[c]) continue;
                float a0 = L0.ax[xf(i, j, k)], a1 = L0.ax[xf(i + 1, j, k)], a2 = L0.ay[yf(i, j, k)],
                      a3 = L0.ay[yf(i, j + 1, k)], a4 = L0.az[zf(i, j, k)], a5 = L0.az[zf(i, j, k + 1)];
                L0.diag[c] = a0 + a1 + a2 + a3 + a4 + a5;
                bool bnd = (i == 0 && a0 > 0) || (i == nx - 1 && a1 > 0) || (j == 0 && a2 > 0) ||
                           (j == ny - 1 && a3 > 0) || (k == 0 && a4 > 0) || (k == nz - 1 && a5 > 0);
                L0.leak[c] = (L0.diag[c] > 0 && bnd) ? 1 : 0;
            }
    // ---- level 1: connected pieces of 2x2x2 cell blocks
    {
        MinSets s(n0);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int c = i + nx * (j + ny * k);
                    if (L0.diag[c] <= 0) continue;
                    if ((i & 1) == 0 && i + 1 < nx && L0.ax[xf(i + 1, j, k)] > 0 && L0.diag[c + 1] > 0) s.merge(c, c + 1);
                    if ((j & 1) == 0 && j + 1 < ny && L0.ay[yf(i, j + 1, k)] > 0 && L0.diag[c + nx] > 0) s.merge(c, c + nx);
                    if ((k & 1) == 0 && k + 1 < nz && L0.az[zf(i, j, k + 1)] > 0 && L0.diag[c + sxy] > 0)
                        s.merge(c, c + sxy);
                }
        L0.agg.assign(n0, -1);
        std::vector<int> id(n0, -1);
        HostLevel C;
        for (int c = 0; c < n0; ++c) {
            if (L0.diag[c] <= 0) continue;
            int r = s.root(c);
            if (id[r] < 0) {
                id[r] = C.n++;
                int i = c % nx, j = (c / nx) % ny, k = c / sxy;
                C.bi.push_back(i >> 1);
                C.bj.push_back(j >> 1);
                C.bk.push_back(k >> 1);
            }
            L0.agg[c] = id[r];
        }
        Rows rows(C.n);
        std::vector<double> dg(C.n, 0.0);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int c = i + nx * (j + ny * k);
                    if (L0.diag[c] <= 0) continue;
                    const int nb[3] = {i + 1 < nx ? c + 1 : -1, j + 1 < ny ? c + nx : -1, k + 1 < nz ? c + sxy : -1};
                    const float w[3] = {i + 1 < nx ? L0.ax[xf(i + 1, j, k)] : 0.0f,
                                        j + 1 < ny ? L0.ay[yf(i, j + 1, k)] : 0.0f,
                                        k + 1 < nz ? L0.az[zf(i, j, k + 1)] : 0.0f};
                    for (int d = 0; d < 3; ++d) {
                        if (nb[d] < 0 || w[d] <= 0) continue;
                        int A = L0.agg[c], B = L0.agg[nb[d]];
                        if (A < 0 || B < 0 || A == B) continue;
                        rows.add(A, B, w[d]);
                        rows.add(B, A, w[d]);
                    }
                }
        for (int c = 0; c < n0; ++c)
            if (L0.agg[c] >= 0) dg[L0.agg[c]] += L0.diag[c];
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    int c = i + nx * (j + ny * k);
                    if (L0.diag[c] <= 0) continue;
                    const int nb[3] = {i + 1 < nx ? c + 1 : -1, j + 1 < ny ? c + nx : -1, k + 1 < nz ? c + sxy : -1};
                    const float w[3] = {i + 1 < nx ? L0.ax[xf(i + 1, j, k)] : 0.0f,
                                        j + 1 < ny ? L0.ay[yf(i, j + 1, k)] : 0.0f,
                                        k + 1 < nz ? L0.az[zf(i, j, k + 1)] : 0.0f};
                    for (int d = 0; d < 3; ++d)
                        if (nb[d] >= 0 && L0.diag[nb[d]] > 0 && w[d] > 0 && L0.agg[nb[d]] == L0.agg[c])
                            dg[L0.agg[c]] -= 2.0 * w[d];   // internal face counted from both cells
                }
        to_csr(C, rows, dg);
        C.leak.assign(C.n, 0);
        for (int c = 0; c < n0; ++c)
            if (L0.agg[c] >= 0 && L0.leak[c]) C.leak[L0.agg[c]] = 1;
        H.L.push_back(std::move(C));
    }
    return finish_hierarchy(H, dense_max, err);
}

bool finish_hierarchy(HostHierarchy& H, int dense_max, std::string* err) {
    // ---- coarser levels (reference rule), then optional extension for the dense solve
    bool reference_phase = true;
    while (true) {
        const int n = H.L.back().n;
        if ((int)H.L.size() >= 12) break;
        if (reference_phase && !(n > 2000)) {
            reference_phase = false;
            H.ref_levels = (int)H.L.size();
        }
        if (!reference_phase && !(n > dense_max)) break;
        HostLevel C;
        std::vector<int> agg;
        int nc = group_pieces(H.L.back(), agg, C);
        if (nc > 0.8 * n) break;
        H.L.back().agg = std::move(agg);
        galerkin_sparse(H.L.back(), C);
        H.L.push_back(std::move(C));
    }
    if (H.ref_levels == 0) H.ref_levels = (int)H.L.size();
    for (size_t l = 1; l < H.L.size(); ++l) make_ell(H.L[l]);
    for (size_t l = 0; l + 1 < H.L.size(); ++l)   // level 0's lists may come with a device-built level 1
        if (!(l == 0 && H.L[0].cptr.size() == (size_t)H.L[1].n + 1)) make_children(H.L[l], H.L[l + 1].n);
    // coarsening can stall (many small sealed pockets, each its own unknown):
    // such a coarsest level is solved by the reference's Gauss-Seidel sweeps
    // only, without the O(n^2) memory and O(n^3) work of the dense inverse
    H.nc = H.L.back().n;
    if (H.nc > kCoarseGsMax) {
        if (err)
            *err = "coarsest level has " + std::to_string(H.nc) + " unknowns (limit " + std::to_string(kCoarseGsMax) +
                   "): coarsening stalled";
        return false;
    }
    if (H.L.size() >= 2) gs_schedule(H);
    if (H.nc <= dense_max) dense_inverse(H);
    return true;
}

int compute_zones(int nx, int ny, int nz, const float* ax, const float* ay, const float* az, const uint8_t* blocked,
                  std::vector<int>& zone, std::vector<uint8_t>& zone_open) {
    const int N = nx * ny * nz, sxy = nx * ny;
    zone.assign(N, -1);
    zone_open.clear();
    std::vector<int> stack;
    for (int seed = 0; seed < N; ++seed) {
        if (blocked[seed] || zone[seed] >= 0) continue;
        const int z = (int)zone_open.size();
        uint8_t open = 0;
        zone[seed] = z;
        stack.assign(1, seed);
        while (!stack.empty()) {
            int c = stack.back();
            stack.pop_back();
            int i = c % nx, j = (c / nx) % ny, k = c / sxy;
            const float a[6] = {ax[i + (nx + 1) * (j + ny * k)], ax[i + 1 + (nx + 1) * (j + ny * k)],
                                ay[i + nx * (j + (ny + 1) * k)], ay[i + nx * (j + 1 + (ny + 1) * k)],
                                az[i + nx * (j + ny * k)], az[i + nx * (j + ny * (k + 1))]};
            const int nb[6] = {i > 0 ? c - 1 : -1, i < nx - 1 ? c + 1 : -1, j > 0 ? c - nx : -1,
                               j < ny - 1 ? c + nx : -1, k > 0 ? c - sxy : -1, k < nz - 1 ? c + sxy : -1};
            for (int d = 0; d < 6; ++d) {
                if (a[d] <= 0) continue;
                if (nb[d] < 0) {
                    if (d != 4) open = 1;   // the ground never opens a zone
                    continue;
                }
                if (blocked[nb[d]] || zone[nb[d]] >= 0) continue;
                zone[nb[d]] = z;
                stack.push_back(nb[d]);
            }
        }
        zone_open.push_back(open);
    }
    return (int)zone_open.size();
}

}  // namespace phg
