// Host FP64 model arithmetic and the MlModel native stubs of the C ABI.
#include "model_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/phylo_hip.h"

namespace phy {

static const double kComposeEps = 1e-10;  // lib/mlmodel.h EPSILON

// Regularized lower incomplete gamma P(a, x) = gamma(a, x) / Gamma(a), with
// F = x^a e^-x / Gamma(a):
//  * x < a + 1: the power series P = F * sum_{n>=0} x^n Gamma(a) / Gamma(a+n+1)
//    (Abramowitz & Stegun 6.5.29), whose terms t_0 = 1/a,
//    t_n = t_{n-1} x / (a + n) fall off fast there;
//  * otherwise Q = 1 - P = F / f with the continued fraction
//    f = b_0 + a_1 / (b_1 + a_2 / (b_2 + ...)),  b_k = x + 2k + 1 - a,
//    a_k = -k (k - a)  (the even form of A&S 6.5.31), evaluated forward by the
//    modified Lentz method (W. J. Lentz, Appl. Opt. 15 (1976) 668;
//    I. J. Thompson and A. R. Barnett, J. Comput. Phys. 64 (1986) 490).
double gammp(double a, double x) {
  if (x <= 0.0) return 0.0;
  if (std::isinf(x)) return 1.0;
  const double log_f = a * std::log(x) - x - std::lgamma(a);
  if (x < a + 1.0) {
    double term = 1.0 / a, total = term;
    for (int n = 1; n < 100000; ++n) {
      term *= x / (a + n);
      total += term;
      if (std::fabs(term) < std::fabs(total) * 1e-17) break;
    }
    return total * std::exp(log_f);
  }
  const double tiny = 1e-300;  // stands in for a zero denominator
  double f = x + 1.0 - a;
  if (std::fabs(f) < tiny) f = tiny;
  double C = f, D = 0.0;
  for (int k = 1; k < 100000; ++k) {
    const double ak = -k * (k - a), bk = x + 2.0 * k + 1.0 - a;
    D = bk + ak * D;
    if (std::fabs(D) < tiny) D = tiny;
    D = 1.0 / D;
    C = bk + ak / C;
    if (std::fabs(C) < tiny) C = tiny;
    const double step = C * D;
    f *= step;
    if (std::fabs(step - 1.0) < 1e-17) break;
  }
  return 1.0 - std::exp(log_f) / f;
}

double gammp_inv(double a, double p) {
  if (p <= 0.0) return 0.0;
  if (p >= 1.0) return INFINITY;
  // bracket
  double lo = 0.0, hi = std::max(1.0, a);
  while (gammp(a, hi) < p) {
    lo = hi;
    hi *= 2.0;
  }
  double x = 0.5 * (lo + hi);
  const double gln = std::lgamma(a);
  for (int it = 0; it < 200; ++it) {
    const double f = gammp(a, x) - p;
    if (f < 0.0) lo = x; else hi = x;
    // density of Gamma(a,1) at x
    const double dens = std::exp(-x + (a - 1.0) * std::log(x) - gln);
    double xn = (dens > 0.0) ? x - f / dens : 0.5 * (lo + hi);
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    if (std::fabs(xn - x) <= 1e-16 * std::max(1.0, std::fabs(x)) || hi - lo <= 1e-300) {
      x = xn;
      break;
    }
    x = xn;
  }
  return x;
}

bool stationary(int n, const double* Q, double* pi, std::string* err) {
  // Solve Q^T pi = 0 with the last equation replaced by sum(pi) = 1.
  std::vector<double> A((size_t)n * (n + 1));
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) A[(size_t)i * (n + 1) + j] = (i == n - 1) ? 1.0 : Q[(size_t)j * n + i];
    A[(size_t)i * (n + 1) + n] = (i == n - 1) ? 1.0 : 0.0;
  }
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[(size_t)r * (n + 1) + c]) > std::fabs(A[(size_t)piv * (n + 1) + c])) piv = r;
    if (std::fabs(A[(size_t)piv * (n + 1) + c]) < 1e-300) {
      if (err) *err = "rate matrix has no unique stationary distribution";
      return false;
    }
    if (piv != c)
      for (int j = 0; j <= n; ++j) std::swap(A[(size_t)piv * (n + 1) + j], A[(size_t)c * (n + 1) + j]);
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = A[(size_t)r * (n + 1) + c] / A[(size_t)c * (n + 1) + c];
      if (f == 0.0) continue;
      for (int j = c; j <= n; ++j) A[(size_t)r * (n + 1) + j] -= f * A[(size_t)c * (n + 1) + j];
    }
  }
  for (int i = 0; i < n; ++i) {
    pi[i] = A[(size_t)i * (n + 1) + n] / A[(size_t)i * (n + 1) + i];
    if (!(pi[i] > 0.0)) {
      if (err) *err = "rate matrix stationary distribution is not positive";
      return false;
    }
  }
  return true;
}

void jacobi_sym(int n, const double* Ain, double* V, double* w) {
  std::vector<double> A(Ain, Ain + (size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[(size_t)i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < n; ++p) {
      diag += A[(size_t)p * n + p] * A[(size_t)p * n + p];
      for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
    }
    if (off <= 1e-34 * diag || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        // A <- J^T A J with J the rotation in the (p,q) plane
        for (int k = 0; k < n; ++k) {
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        A[(size_t)p * n + q] = A[(size_t)q * n + p] = 0.0;
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < n; ++i) w[i] = A[(size_t)i * n + i];
}

bool eigen_reversible(int n, const double* Q, double* U, double* lambda, double* Ui,
                      std::string* err) {
  for (int i = 0; i < n * n; ++i)
    if (!std::isfinite(Q[i])) {
      if (err) *err = "rate matrix contains NaN or infinity";
      return false;
    }
  std::vector<double> pi(n);
  if (!stationary(n, Q, pi.data(), err)) return false;
  std::vector<double> sq(n), S((size_t)n * n), V((size_t)n * n);
  for (int i = 0; i < n; ++i) sq[i] = std::sqrt(pi[i]);
  double scale = 0.0;
  for (int i = 0; i < n * n; ++i) scale = std::max(scale, std::fabs(Q[i]));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double sij = sq[i] / sq[j] * Q[(size_t)i * n + j];
      const double sji = sq[j] / sq[i] * Q[(size_t)j * n + i];
      if (std::fabs(sij - sji) > 1e-9 * scale) {
        if (err) *err = "Imaginary eigenvalues: rate matrix is not reversible";
        return false;
      }
      S[(size_t)i * n + j] = 0.5 * (sij + sji);
    }
  jacobi_sym(n, S.data(), V.data(), lambda);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      U[(size_t)i * n + k] = V[(size_t)i * n + k] / sq[i];
      Ui[(size_t)k * n + i] = V[(size_t)i * n + k] * sq[i];
    }
  return true;
}

void compose(int n, const double* U, const double* lambda, const double* Ui, double t,
             double* P) {
  if (t == -1.0) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += U[(size_t)i * n + k] * lambda[k] * Ui[(size_t)k * n + j];
        P[(size_t)i * n + j] = s;
      }
    return;
  }
  if (!(t >= kComposeEps)) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) P[(size_t)i * n + j] = (i == j) ? 1.0 : 0.0;
    return;
  }
  double e[64];
  for (int k = 0; k < n; ++k) e[k] = std::exp(lambda[k] * t);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += U[(size_t)i * n + k] * e[k] * Ui[(size_t)k * n + j];
      P[(size_t)i * n + j] = s;
    }
}

bool gtr_q(int n, const double* pi, const double* co, double* Q, std::string* err) {
  if (n < 2 || n > 64) {
    if (err) *err = "alphabet size out of range";
    return false;
  }
  const int m = n * (n - 1) / 2;
  std::vector<double> c(m);
  for (int k = 0; k < m - 1; ++k) c[k] = co[k];
  c[m - 1] = 1.0;  // last element of GTR = 1.0 (lib/mlModel.ml:396-400)
  std::memset(Q, 0, sizeof(double) * n * n);
  int k = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j, ++k) {
      Q[(size_t)i * n + j] = c[k] * pi[j];
      Q[(size_t)j * n + i] = c[k] * pi[i];
    }
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int j = 0; j < n; ++j)
      if (j != i) d += Q[(size_t)i * n + j];
    Q[(size_t)i * n + i] = -d;
  }
  double mr = 0.0;
  for (int i = 0; i < n; ++i) mr += -Q[(size_t)i * n + i] * pi[i];
  for (int i = 0; i < n * n; ++i) Q[i] /= mr;
  return true;
}

bool gamma_rates(int mode, double a, double b, int n, double* rates, std::string* err) {
  if (n < 1 || !(a > 0.0)) {
    if (err) *err = "gamma_rates: need n >= 1 and shape > 0";
    return false;
  }
  if (mode == PHY_GAMMA_REF) {
    if (!(b > 0.0)) {
      if (err) *err = "gamma_rates: scale must be > 0";
      return false;
    }
    for (int i = 0; i < n; ++i) rates[i] = b * gammp_inv(a, (double)i / (double)n);
    return true;
  }
  if (mode == PHY_GAMMA_MEAN) {
    if (n == 1) {
      rates[0] = 1.0;
      return true;
    }
    std::vector<double> x(n + 1);
    x[0] = 0.0;
    x[n] = INFINITY;
    for (int i = 1; i < n; ++i) x[i] = gammp_inv(a, (double)i / (double)n);
    for (int i = 0; i < n; ++i) rates[i] = n * (gammp(a + 1.0, x[i + 1]) - gammp(a + 1.0, x[i]));
    return true;
  }
  if (err) *err = "gamma_rates: unknown mode";
  return false;
}

}  // namespace phy

// --------------------------------------------------------------------------
// C ABI: MlModel native stubs.
// --------------------------------------------------------------------------
extern "C" {

int phy_diagonalize_gtr(int32_t n, const double* Q, double* U, double* D, double* Ui) {
  if (n < 1 || n > 64 || !Q || !U || !D || !Ui) return PHY_EINVAL;
  std::vector<double> lam(n), u((size_t)n * n), ui((size_t)n * n);
  std::string err;
  if (!phy::eigen_reversible(n, Q, u.data(), lam.data(), ui.data(), &err)) return PHY_ENUMERIC;
  std::memcpy(U, u.data(), sizeof(double) * n * n);
  std::memcpy(Ui, ui.data(), sizeof(double) * n * n);
  std::memset(D, 0, sizeof(double) * n * n);
  for (int i = 0; i < n; ++i) D[(size_t)i * n + i] = lam[i];
  return PHY_OK;
}

int phy_diagonalize_sym(int32_t n, double* A, double* D) {
  if (n < 1 || n > 64 || !A || !D) return PHY_EINVAL;
  for (int i = 0; i < n * n; ++i)
    if (!std::isfinite(A[i])) return PHY_ENUMERIC;
  std::vector<double> V((size_t)n * n), w(n);
  phy::jacobi_sym(n, A, V.data(), w.data());
  // rows of A = eigenvectors
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < n; ++i) A[(size_t)k * n + i] = V[(size_t)i * n + k];
  std::memset(D, 0, sizeof(double) * n * n);
  for (int i = 0; i < n; ++i) D[(size_t)i * n + i] = w[i];
  return PHY_OK;
}

int phy_compose_gtr(int32_t n, const double* U, const double* D, const double* Ui, double t,
                    double* P) {
  if (n < 1 || n > 64 || !U || !D || !Ui || !P) return PHY_EINVAL;
  double lam[64];
  for (int i = 0; i < n; ++i) lam[i] = D[(size_t)i * n + i];
  phy::compose(n, U, lam, Ui, t, P);
  return PHY_OK;
}

int phy_compose_sym(int32_t n, const double* U, const double* D, float t, double* P) {
  if (n < 1 || n > 64 || !U || !D || !P) return PHY_EINVAL;
  // U rows are eigenvectors: P = U^T diag(e^{lambda t}) U
  std::vector<double> Ut((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) Ut[(size_t)i * n + k] = U[(size_t)k * n + i];
  double lam[64];
  for (int i = 0; i < n; ++i) lam[i] = D[(size_t)i * n + i];
  // the float branch length is compared and used in single precision, then
  // widened, as in compose_sym's `const float t` (lib/mlmodel.c:280-301)
  const double tt = (double)t;
  if (t == -1.0f) {
    phy::compose(n, Ut.data(), lam, U, -1.0, P);
  } else if (tt >= 1e-10) {
    phy::compose(n, Ut.data(), lam, U, tt, P);
  } else {
    phy::compose(n, Ut.data(), lam, U, 0.0, P);
  }
  return PHY_OK;
}

int phy_gamma_rates(int32_t mode, double a, double b, int32_t n, double* rates) {
  if (!rates) return PHY_EINVAL;
  std::string err;
  return phy::gamma_rates(mode, a, b, n, rates, &err) ? PHY_OK : PHY_EINVAL;
}

int phy_gtr_q(int32_t n, const double* pi, const double* co, double* Q) {
  if (!pi || !co || !Q) return PHY_EINVAL;
  std::string err;
  return phy::gtr_q(n, pi, co, Q, &err) ? PHY_OK : PHY_EINVAL;
}

}  // extern "C"
