// Host FP64 K x K algebra for the estimation core (see dense.hpp).
#include "dense.hpp"

#include <cmath>

namespace fehip {
namespace dense {

int chol_solve_invert(const std::vector<double>& A, const std::vector<double>& b, int n, std::vector<double>& x,
                      std::vector<double>& Ainv) {
    // U (upper) with A = U^T U, stored column-major in U[j*n + i], i <= j
    std::vector<double> U(A);
    for (int j = 0; j < n; ++j) {
        double s = U[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= U[(size_t)j * n + k] * U[(size_t)j * n + k];
        if (!(s > 0.0)) return j + 1;
        const double ujj = std::sqrt(s);
        U[(size_t)j * n + j] = ujj;
        for (int i = j + 1; i < n; ++i) {
            double t = U[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) t -= U[(size_t)j * n + k] * U[(size_t)i * n + k];
            U[(size_t)i * n + j] = t / ujj;
        }
    }
    x.assign(b.begin(), b.end());
    for (int i = 0; i < n; ++i) {  // U^T y = b
        double t = x[i];
        for (int k = 0; k < i; ++k) t -= U[(size_t)i * n + k] * x[k];
        x[i] = t / U[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {  // U x = y
        double t = x[i];
        for (int k = i + 1; k < n; ++k) t -= U[(size_t)k * n + i] * x[k];
        x[i] = t / U[(size_t)i * n + i];
    }
    // A^{-1} = U^{-1} U^{-T}
    std::vector<double> Ui((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        Ui[(size_t)j * n + j] = 1.0 / U[(size_t)j * n + j];
        for (int i = j - 1; i >= 0; --i) {
            double t = 0.0;
            for (int k = i + 1; k <= j; ++k) t += U[(size_t)k * n + i] * Ui[(size_t)j * n + k];
            Ui[(size_t)j * n + i] = -t / U[(size_t)i * n + i];
        }
    }
    Ainv.assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            double t = 0.0;
            for (int k = j; k < n; ++k) t += Ui[(size_t)k * n + i] * Ui[(size_t)k * n + j];
            Ainv[(size_t)j * n + i] = t;
            Ainv[(size_t)i * n + j] = t;
        }
    return 0;
}

void sym_eig(const std::vector<double>& A, int n, std::vector<double>& w, std::vector<double>& V) {
    std::vector<double> a(A);
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (size_t e = 0; e < a.size(); ++e) tot += a[e] * a[e];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                if (i != j) off += a[(size_t)j * n + i] * a[(size_t)j * n + i];
        if (off == 0.0 || off <= 1e-30 * tot) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[(size_t)q * n + p];
                if (apq == 0.0) continue;
                const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[(size_t)p * n + k], akq = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * akp - s * akq;
                    a[(size_t)q * n + k] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[(size_t)k * n + p], aqk = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * apk - s * aqk;
                    a[(size_t)k * n + q] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)p * n + k], vkq = V[(size_t)q * n + k];
                    V[(size_t)p * n + k] = c * vkp - s * vkq;
                    V[(size_t)q * n + k] = s * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
}

// True when A (symmetric) factors as U^T U with every pivot above 1e-12 times the largest diagonal
// entry: A is then positive definite with a clear margin, so clipping its negative eigenvalues would
// return A itself (its eigen-reconstruction V diag(w) V^T equals A up to rounding).
static bool clearly_positive_definite(const std::vector<double>& A, int n) {
    double dmax = 0.0;
    for (int i = 0; i < n; ++i) dmax = std::fmax(dmax, A[(size_t)i * n + i]);
    if (!(dmax > 0.0)) return false;
    const double floor_pivot = 1e-12 * dmax;
    std::vector<double> U(A);
    for (int j = 0; j < n; ++j) {
        double s = U[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= U[(size_t)j * n + k] * U[(size_t)j * n + k];
        if (!(s > floor_pivot)) return false;
        const double ujj = std::sqrt(s);
        U[(size_t)j * n + j] = ujj;
        for (int i = j + 1; i < n; ++i) {
            double t = U[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) t -= U[(size_t)j * n + k] * U[(size_t)i * n + k];
            U[(size_t)i * n + j] = t / ujj;
        }
    }
    return true;
}

int psd_clip(std::vector<double>& A, int n) {
    // the reference's dsyev + clip of negative eigenvalues + rebuild (src/core/fe_pipeline.f90:1797-1894),
    // taken only when A is not clearly positive definite (otherwise it would give back A)
    if (clearly_positive_definite(A, n)) return 0;
    std::vector<double> w, V;
    sym_eig(A, n, w, V);
    int clipped = 0;
    for (int k = 0; k < n; ++k) {
        double lam = w[k];
        if (lam < 0.0) {
            lam = 0.0;
            clipped = 1;
        }
        const double sl = lam > 0.0 ? std::sqrt(lam) : 0.0;
        for (int i = 0; i < n; ++i) V[(size_t)k * n + i] *= sl;
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double t = 0.0;
            for (int k = 0; k < n; ++k) t += V[(size_t)k * n + i] * V[(size_t)k * n + j];
            A[(size_t)j * n + i] = t;
        }
    return clipped;
}

}  // namespace dense
}  // namespace fehip
