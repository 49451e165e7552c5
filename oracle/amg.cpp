// Smoothed-aggregation AMG + SPAI-0 (see amg.hpp for the sources of the method).
// Test infrastructure only (see oracle.h).
#include "amg.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {
namespace {

constexpr int kCoarseEnough = 500;  // dense LU below this many unknowns
constexpr double kEpsStrong = 0.08; // strength threshold, halved per level
constexpr double kRelax = 1.0;      // prolongation-smoother damping factor

void spmv(const Csr &A, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < A.nrows; r++) {
        double s = 0;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) s += A.val[k] * x[A.col[k]];
        y[r] = s;
    }
}

Csr transpose(const Csr &A) {
    Csr T;
    T.nrows = A.ncols;
    T.ncols = A.nrows;
    T.ptr.assign(T.nrows + 1, 0);
    for (int k = 0; k < A.ptr[A.nrows]; k++) T.ptr[A.col[k] + 1]++;
    for (int r = 0; r < T.nrows; r++) T.ptr[r + 1] += T.ptr[r];
    T.col.resize(A.ptr[A.nrows]);
    T.val.resize(A.ptr[A.nrows]);
    std::vector<int> pos(T.ptr.begin(), T.ptr.end() - 1);
    for (int r = 0; r < A.nrows; r++)
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
            int p = pos[A.col[k]]++;
            T.col[p] = r;
            T.val[p] = A.val[k];
        }
    return T;
}

// Row-wise (Gustavson) sparse product C = A B, two passes, OpenMP over rows.
Csr spgemm(const Csr &A, const Csr &B) {
    Csr C;
    C.nrows = A.nrows;
    C.ncols = B.ncols;
    C.ptr.assign(C.nrows + 1, 0);
#pragma omp parallel
    {
        std::vector<int> mark(B.ncols, -1);
#pragma omp for schedule(dynamic, 1024)
        for (int r = 0; r < A.nrows; r++) {
            int cnt = 0;
            for (int ka = A.ptr[r]; ka < A.ptr[r + 1]; ka++) {
                int j = A.col[ka];
                for (int kb = B.ptr[j]; kb < B.ptr[j + 1]; kb++) {
                    int c = B.col[kb];
                    if (mark[c] != r) {
                        mark[c] = r;
                        cnt++;
                    }
                }
            }
            C.ptr[r + 1] = cnt;
        }
    }
    for (int r = 0; r < C.nrows; r++) C.ptr[r + 1] += C.ptr[r];
    C.col.resize(C.ptr[C.nrows]);
    C.val.resize(C.ptr[C.nrows]);
#pragma omp parallel
    {
        std::vector<int> where(B.ncols, -1);
#pragma omp for schedule(dynamic, 1024)
        for (int r = 0; r < A.nrows; r++) {
            int beg = C.ptr[r], end = beg;
            for (int ka = A.ptr[r]; ka < A.ptr[r + 1]; ka++) {
                int j = A.col[ka];
                double a = A.val[ka];
                for (int kb = B.ptr[j]; kb < B.ptr[j + 1]; kb++) {
                    int c = B.col[kb];
                    if (where[c] < beg) {
                        where[c] = end;
                        C.col[end] = c;
                        C.val[end] = a * B.val[kb];
                        end++;
                    } else {
                        C.val[where[c]] += a * B.val[kb];
                    }
                }
            }
            for (int k = beg; k < end; k++) where[C.col[k]] = -1;
        }
    }
    return C;
}

// Plain (Vanek) aggregation over the strength graph. agg[i] = aggregate id or
// -1 for nodes without strong connections (they are left to the smoother).
int aggregate(const Csr &A, double eps, std::vector<int> &agg, std::vector<char> &strong) {
    int n = A.nrows;
    std::vector<double> dia(n, 0.0);
    for (int r = 0; r < n; r++)
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
            if (A.col[k] == r) dia[r] = A.val[k];
    strong.assign(A.ptr[n], 0);
    double eps2 = eps * eps;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; r++)
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
            int c = A.col[k];
            if (c != r && A.val[k] * A.val[k] > eps2 * std::fabs(dia[r] * dia[c])) strong[k] = 1;
        }
    agg.assign(n, -2);  // -2: not yet visited
    for (int r = 0; r < n; r++) {
        bool any = false;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) any |= strong[k] != 0;
        if (!any) agg[r] = -1;
    }
    int na = 0;
    // pass 1: roots whose strong neighbourhood is still free
    for (int r = 0; r < n; r++) {
        if (agg[r] != -2) continue;
        bool free = true;
        for (int k = A.ptr[r]; k < A.ptr[r + 1] && free; k++)
            if (strong[k] && agg[A.col[k]] >= 0) free = false;
        if (!free) continue;
        agg[r] = na;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
            if (strong[k]) agg[A.col[k]] = na;
        na++;
    }
    // pass 2: attach leftovers to the strongest aggregated neighbour
    std::vector<int> snapshot(agg);
    for (int r = 0; r < n; r++) {
        if (snapshot[r] != -2) continue;
        int best = -1;
        double bv = -1;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
            if (strong[k] && snapshot[A.col[k]] >= 0 && std::fabs(A.val[k]) > bv) {
                bv = std::fabs(A.val[k]);
                best = snapshot[A.col[k]];
            }
        agg[r] = best;
    }
    // pass 3: anything still unvisited starts its own aggregate with its free neighbours
    for (int r = 0; r < n; r++) {
        if (agg[r] != -2) continue;
        agg[r] = na;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
            if (strong[k] && agg[A.col[k]] == -2) agg[A.col[k]] = na;
        na++;
    }
    return na;
}

// P = (I - omega D_F^{-1} A_F) P_tent with the filtered matrix A_F (weak
// connections lumped into the diagonal) and omega = relax * 4/3 / rho(D_F^{-1} A_F),
// rho bounded by Gershgorin.
Csr smoothed_prolongation(const Csr &A, const std::vector<int> &agg, int na,
                          const std::vector<char> &strong) {
    int n = A.nrows;
    std::vector<double> dF(n, 0.0);
    double rho = 0;
#pragma omp parallel for reduction(max : rho) schedule(static)
    for (int r = 0; r < n; r++) {
        double d = 0, weak = 0, abs_sum = 0;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
            if (A.col[k] == r)
                d += A.val[k];
            else if (!strong[k])
                weak += A.val[k];
        }
        dF[r] = d + weak;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++)
            if (A.col[k] == r)
                abs_sum += std::fabs(dF[r]);
            else if (strong[k])
                abs_sum += std::fabs(A.val[k]);
        double q = dF[r] != 0 ? abs_sum / std::fabs(dF[r]) : 0;
        rho = std::max(rho, q);
    }
    double omega = rho > 0 ? kRelax * (4.0 / 3.0) / rho : 0.0;
    Csr P;
    P.nrows = n;
    P.ncols = na;
    P.ptr.assign(n + 1, 0);
    std::vector<std::vector<std::pair<int, double>>> rows(n);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int r = 0; r < n; r++) {
        auto &row = rows[r];
        auto add = [&](int c, double v) {
            for (auto &e : row)
                if (e.first == c) {
                    e.second += v;
                    return;
                }
            row.emplace_back(c, v);
        };
        if (agg[r] >= 0) add(agg[r], 1.0);
        if (dF[r] == 0) continue;
        double s = -omega / dF[r];
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
            int c = A.col[k];
            double a;
            if (c == r)
                a = dF[r];
            else if (strong[k])
                a = A.val[k];
            else
                continue;
            if (agg[c] >= 0) add(agg[c], s * a);
        }
        std::sort(row.begin(), row.end());
    }
    for (int r = 0; r < n; r++) P.ptr[r + 1] = P.ptr[r] + (int)rows[r].size();
    P.col.resize(P.ptr[n]);
    P.val.resize(P.ptr[n]);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; r++) {
        int p = P.ptr[r];
        for (auto const &e : rows[r]) {
            P.col[p] = e.first;
            P.val[p] = e.second;
            p++;
        }
    }
    return P;
}

std::vector<double> spai0(const Csr &A) {
    std::vector<double> m(A.nrows, 0.0);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < A.nrows; r++) {
        double d = 0, s = 0;
        for (int k = A.ptr[r]; k < A.ptr[r + 1]; k++) {
            if (A.col[k] == r) d = A.val[k];
            s += A.val[k] * A.val[k];
        }
        m[r] = s > 0 ? d / s : 0.0;
    }
    return m;
}

}  // namespace

Amg::Amg(int n, const int *rowptr, const int *col, const double *val) {
    Level L0;
    L0.A.nrows = L0.A.ncols = n;
    L0.A.ptr.assign(rowptr, rowptr + n + 1);
    L0.A.col.assign(col, col + rowptr[n]);
    L0.A.val.assign(val, val + rowptr[n]);
    levels_.push_back(std::move(L0));
    double eps = kEpsStrong;
    while (levels_.back().A.nrows > kCoarseEnough) {
        Level &L = levels_.back();
        std::vector<int> agg;
        std::vector<char> strong;
        int na = aggregate(L.A, eps, agg, strong);
        if (na == 0 || na >= L.A.nrows) break;
        L.P = smoothed_prolongation(L.A, agg, na, strong);
        L.R = transpose(L.P);
        Csr AP = spgemm(L.A, L.P);
        Level C;
        C.A = spgemm(L.R, AP);
        eps *= 0.5;
        levels_.push_back(std::move(C));
    }
    for (auto &L : levels_) {
        L.spai0 = spai0(L.A);
        L.x.assign(L.A.nrows, 0.0);
        L.b.assign(L.A.nrows, 0.0);
        L.t.assign(L.A.nrows, 0.0);
    }
    // dense LU with partial pivoting of the coarsest matrix
    Csr const &C = levels_.back().A;
    coarse_n_ = C.nrows;
    int m = coarse_n_;
    coarse_lu_.assign((size_t)m * m, 0.0);
    coarse_piv_.resize(m);
    for (int r = 0; r < m; r++)
        for (int k = C.ptr[r]; k < C.ptr[r + 1]; k++) coarse_lu_[(size_t)r * m + C.col[k]] += C.val[k];
    double scale = 0;
    for (int i = 0; i < m; i++) scale = std::max(scale, std::fabs(coarse_lu_[(size_t)i * m + i]));
    for (int k = 0; k < m; k++) {
        int p = k;
        for (int i = k + 1; i < m; i++)
            if (std::fabs(coarse_lu_[(size_t)i * m + k]) > std::fabs(coarse_lu_[(size_t)p * m + k])) p = i;
        coarse_piv_[k] = p;
        if (p != k)
            for (int j = 0; j < m; j++) std::swap(coarse_lu_[(size_t)k * m + j], coarse_lu_[(size_t)p * m + j]);
        double piv = coarse_lu_[(size_t)k * m + k];
        if (std::fabs(piv) <= 1e-13 * scale) {  // singular (pure-Neumann) block
            coarse_lu_[(size_t)k * m + k] = 0.0;
            continue;
        }
        for (int i = k + 1; i < m; i++) {
            double f = coarse_lu_[(size_t)i * m + k] / piv;
            coarse_lu_[(size_t)i * m + k] = f;
            for (int j = k + 1; j < m; j++) coarse_lu_[(size_t)i * m + j] -= f * coarse_lu_[(size_t)k * m + j];
        }
    }
}

void Amg::cycle(int l) {
    Level &L = levels_[l];
    int n = L.A.nrows;
    if (l + 1 == (int)levels_.size()) {
        int m = coarse_n_;
        std::vector<double> y(L.b);
        for (int k = 0; k < m; k++)
            if (coarse_piv_[k] != k) std::swap(y[k], y[coarse_piv_[k]]);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < i; j++) y[i] -= coarse_lu_[(size_t)i * m + j] * y[j];
        for (int i = m - 1; i >= 0; i--) {
            double d = coarse_lu_[(size_t)i * m + i];
            for (int j = i + 1; j < m; j++) y[i] -= coarse_lu_[(size_t)i * m + j] * y[j];
            y[i] = d != 0 ? y[i] / d : 0.0;
        }
        L.x = y;
        return;
    }
    Level &C = levels_[l + 1];
    // pre-smoothing from x = 0: x = M b
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) L.x[i] = L.spai0[i] * L.b[i];
    spmv(L.A, L.x.data(), L.t.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) L.t[i] = L.b[i] - L.t[i];
    spmv(L.R, L.t.data(), C.b.data());
    cycle(l + 1);
    spmv(L.P, C.x.data(), L.t.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) L.x[i] += L.t[i];
    // post-smoothing
    spmv(L.A, L.x.data(), L.t.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) L.x[i] += L.spai0[i] * (L.b[i] - L.t[i]);
}

void Amg::apply(const double *r, double *z) {
    Level &L = levels_[0];
    std::memcpy(L.b.data(), r, sizeof(double) * L.A.nrows);
    cycle(0);
    std::memcpy(z, L.x.data(), sizeof(double) * L.A.nrows);
}

}  // namespace orc
