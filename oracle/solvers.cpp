// FP64 Krylov solvers of the oracle: CG (the magnetic solve, core/MagneticField.h:17-21)
// and BiCGStab (the pressure solve, core/Pressure.cpp:126-131), with no, Jacobi,
// or smoothed-aggregation AMG + SPAI-0 preconditioning (amg.cpp).
// Test infrastructure only (see oracle.h).
#include "oracle.h"
#include "amg.hpp"

#include <chrono>
#include <cmath>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

double dot(int n, const double *a, const double *b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

struct Precond {
    int kind = 0;
    std::vector<double> inv_diag;
    std::unique_ptr<orc::Amg> amg;

    void setup(int n, const int *rowptr, const int *col, const double *val) {
        if (kind == 1) {
            inv_diag.assign(n, 1.0);
            for (int r = 0; r < n; r++)
                for (int k = rowptr[r]; k < rowptr[r + 1]; k++)
                    if (col[k] == r && val[k] != 0) inv_diag[r] = 1.0 / val[k];
        } else if (kind == 2) {
            amg = std::make_unique<orc::Amg>(n, rowptr, col, val);
        }
    }
    void apply(int n, const double *r, double *z) const {
        if (kind == 0) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++) z[i] = r[i];
        } else if (kind == 1) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++) z[i] = inv_diag[i] * r[i];
        } else {
            amg->apply(r, z);
        }
    }
};

}  // namespace

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_spmv(int n, const int *rowptr, const int *col, const double *val, const double *x,
              double *y) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; r++) {
        double s = 0;
        for (int k = rowptr[r]; k < rowptr[r + 1]; k++) s += val[k] * x[col[k]];
        y[r] = s;
    }
}

double orc_rel_residual(int n, const int *rowptr, const int *col, const double *val,
                        const double *b, const double *x) {
    std::vector<double> ax(n);
    orc_spmv(n, rowptr, col, val, x, ax.data());
    double rr = 0, bb = 0;
#pragma omp parallel for reduction(+ : rr, bb) schedule(static)
    for (int i = 0; i < n; i++) {
        double r = b[i] - ax[i];
        rr += r * r;
        bb += b[i] * b[i];
    }
    if (bb == 0) return std::sqrt(rr);
    return std::sqrt(rr / bb);
}

int orc_cg(int n, const int *rowptr, const int *col, const double *val, const double *b,
           double *x, double tol, int maxit, int precond, double *rel, double *setup_s,
           double *solve_s) {
    double t0 = now();
    Precond M;
    M.kind = precond;
    M.setup(n, rowptr, col, val);
    double t1 = now();
    std::vector<double> r(n), z(n), p(n), q(n);
    double bb = dot(n, b, b);
    int it = 0;
    if (bb == 0) {
        for (int i = 0; i < n; i++) x[i] = 0;
        *rel = 0;
    } else {
        double nb = std::sqrt(bb);
        orc_spmv(n, rowptr, col, val, x, q.data());
        for (int i = 0; i < n; i++) r[i] = b[i] - q[i];
        double res = std::sqrt(dot(n, r.data(), r.data())) / nb;
        double rz_old = 0;
        while (res > tol && it < maxit) {
            M.apply(n, r.data(), z.data());
            double rz = dot(n, r.data(), z.data());
            double beta = it == 0 ? 0.0 : rz / rz_old;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
            orc_spmv(n, rowptr, col, val, p.data(), q.data());
            double alpha = rz / dot(n, p.data(), q.data());
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            rz_old = rz;
            res = std::sqrt(dot(n, r.data(), r.data())) / nb;
            it++;
        }
        *rel = orc_rel_residual(n, rowptr, col, val, b, x);
    }
    double t2 = now();
    *setup_s = t1 - t0;
    *solve_s = t2 - t1;
    return it;
}

int orc_bicgstab(int n, const int *rowptr, const int *col, const double *val, const double *b,
                 double *x, double tol, int maxit, int precond, double *rel, double *setup_s,
                 double *solve_s) {
    // Right-preconditioned BiCGStab (van der Vorst 1992).
    double t0 = now();
    Precond M;
    M.kind = precond;
    M.setup(n, rowptr, col, val);
    double t1 = now();
    std::vector<double> r(n), rh(n), p(n), v(n), s(n), t(n), ph(n), sh(n);
    double bb = dot(n, b, b);
    int it = 0;
    if (bb == 0) {
        for (int i = 0; i < n; i++) x[i] = 0;
        *rel = 0;
    } else {
        double nb = std::sqrt(bb);
        orc_spmv(n, rowptr, col, val, x, t.data());
        for (int i = 0; i < n; i++) r[i] = rh[i] = b[i] - t[i];
        double res = std::sqrt(dot(n, r.data(), r.data())) / nb;
        double rho_old = 1, alpha = 1, omega = 1;
        while (res > tol && it < maxit) {
            double rho = dot(n, rh.data(), r.data());
            if (rho == 0) break;
            if (it == 0) {
                for (int i = 0; i < n; i++) p[i] = r[i];
            } else {
                double beta = (rho / rho_old) * (alpha / omega);
#pragma omp parallel for schedule(static)
                for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
            }
            M.apply(n, p.data(), ph.data());
            orc_spmv(n, rowptr, col, val, ph.data(), v.data());
            alpha = rho / dot(n, rh.data(), v.data());
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
            double sres = std::sqrt(dot(n, s.data(), s.data())) / nb;
            if (sres <= tol) {
#pragma omp parallel for schedule(static)
                for (int i = 0; i < n; i++) x[i] += alpha * ph[i];
                it++;
                res = sres;
                break;
            }
            M.apply(n, s.data(), sh.data());
            orc_spmv(n, rowptr, col, val, sh.data(), t.data());
            double tt = dot(n, t.data(), t.data());
            omega = tt > 0 ? dot(n, t.data(), s.data()) / tt : 0.0;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++) {
                x[i] += alpha * ph[i] + omega * sh[i];
                r[i] = s[i] - omega * t[i];
            }
            rho_old = rho;
            res = std::sqrt(dot(n, r.data(), r.data())) / nb;
            it++;
            if (omega == 0) break;
        }
        *rel = orc_rel_residual(n, rowptr, col, val, b, x);
    }
    double t2 = now();
    *setup_s = t1 - t0;
    *solve_s = t2 - t1;
    return it;
}

}  // extern "C"
