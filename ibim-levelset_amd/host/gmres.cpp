#include "gmres.h"

#include <cmath>
#include <iomanip>
#include <iostream>
#include <vector>

using bbfmm::Vector;

namespace {

double dot(const Vector &a, const Vector &b) {
    double s = 0.0;
    for (int i = 0; i < a.row(); ++i) s += a(i) * b(i);
    return s;
}

void gen_rot(double dx, double dy, double &cs, double &sn) {
    if (dy == 0.0) {
        cs = 1.0;
        sn = 0.0;
    } else if (std::fabs(dy) > std::fabs(dx)) {
        const double t = dx / dy;
        sn = 1.0 / std::sqrt(1.0 + t * t);
        cs = t * sn;
    } else {
        const double t = dy / dx;
        cs = 1.0 / std::sqrt(1.0 + t * t);
        sn = t * cs;
    }
}

void apply_rot(double &dx, double &dy, double cs, double sn) {
    const double t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

void update(Vector &x, int k, const std::vector<std::vector<double>> &H, const std::vector<double> &s,
            const std::vector<Vector> &v) {
    std::vector<double> y(s.begin(), s.begin() + k + 1);
    for (int i = k; i >= 0; i--) {
        y[i] /= H[i][i];
        for (int j = i - 1; j >= 0; j--) y[j] -= H[j][i] * y[i];
    }
    for (int j = 0; j <= k; j++)
        for (int q = 0; q < x.row(); ++q) x(q) += y[j] * v[j](q);
}

}  // namespace

int GMRES(const std::function<Vector(Vector &)> A, Vector &x, Vector &b, int m, int max_iter, double tol) {
    double normb = std::sqrt(dot(b, b));
    Vector r = b, p = A(x);
    for (int q = 0; q < r.row(); ++q) r(q) -= p(q);
    double beta = std::sqrt(dot(r, r));
    if (normb == 0.0) normb = 1;
    double resid = beta / normb;
    std::cout << "=============== GMRES =================" << std::endl;
    if (resid <= tol) return 0;
    std::vector<Vector> v(m + 1);
    std::vector<std::vector<double>> H(m + 1, std::vector<double>(m + 1, 0.0));
    std::vector<double> s(m + 1), cs(m + 1), sn(m + 1);
    int i = 0, j = 1;
    while (j <= max_iter) {
        v[0] = r;
        for (int q = 0; q < r.row(); ++q) v[0](q) *= 1.0 / beta;
        std::fill(s.begin(), s.end(), 0.0);
        s[0] = beta;
        for (i = 0; i < m && j <= max_iter; i++, j++) {
            std::cout << std::setw(6) << j << std::setw(20) << std::scientific << resid << std::fixed << std::endl;
            Vector w = A(v[i]);
            for (int k = 0; k <= i; k++) {
                H[k][i] = dot(w, v[k]);
                for (int q = 0; q < w.row(); ++q) w(q) -= H[k][i] * v[k](q);
            }
            H[i + 1][i] = std::sqrt(dot(w, w));
            v[i + 1] = w;
            for (int q = 0; q < w.row(); ++q) v[i + 1](q) *= 1.0 / H[i + 1][i];
            for (int k = 0; k < i; k++) apply_rot(H[k][i], H[k + 1][i], cs[k], sn[k]);
            gen_rot(H[i][i], H[i + 1][i], cs[i], sn[i]);
            apply_rot(H[i][i], H[i + 1][i], cs[i], sn[i]);
            apply_rot(s[i], s[i + 1], cs[i], sn[i]);
            if ((resid = std::fabs(s[i + 1]) / normb) < tol) {
                update(x, i, H, s, v);
                std::cout << "=============== CONVERGED =============" << std::endl;
                return 0;
            }
        }
        update(x, i - 1, H, s, v);
        r = b;
        p = A(x);
        for (int q = 0; q < r.row(); ++q) r(q) -= p(q);
        beta = std::sqrt(dot(r, r));
        if ((resid = beta / normb) < tol) {
            std::cout << "=============== CONVERGED =============" << std::endl;
            return 0;
        }
    }
    std::cout << "============= NOT CONVERGED ===========" << std::endl;
    return 1;
}

int GMRES(ibim_plan *plan, Vector &x, Vector &b, int m, int max_iter, double tol, ibim_solve_info *info) {
    ibim_solve_info local{};
    ibim_solve_info *inf = info ? info : &local;
    std::cout << "=============== GMRES =================" << std::endl;
    const int rc = ibim::check(ibim_gmres(plan, b.data(), x.data(), m, max_iter, tol, inf), "ibim_gmres");
    std::cout << std::setw(6) << inf->iterations << std::setw(20) << std::scientific << inf->rel_residual
              << std::fixed << "  (" << inf->matvecs << " matvecs)" << std::endl;
    std::cout << (rc == 0 ? "=============== CONVERGED =============" : "============= NOT CONVERGED ===========")
              << std::endl;
    return rc;
}
