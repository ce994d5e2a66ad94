// Host-side FP64 substitution-model arithmetic: GTR rate matrix, reversible
// eigen-decomposition by cyclic Jacobi, P(t) composition and Gamma rate
// categories.  No LAPACK: the reference's LAPACK calls (lib/mlmodel.c:163-365)
// are replaced by a symmetric Jacobi solve of Pi^1/2 Q Pi^-1/2 (Golub & Van
// Loan, Matrix Computations, section 8.5), which has a real spectrum by
// construction.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace phy {

// Regularized lower incomplete gamma P(a, x) (series / continued fraction,
// Press et al., Numerical Recipes 3rd ed., section 6.2).
double gammp(double a, double x);
// x with P(a, x) = p, by safeguarded Newton iteration.
double gammp_inv(double a, double p);

// Stationary distribution of a rate matrix (pi Q = 0, sum pi = 1).
bool stationary(int n, const double* Q, double* pi, std::string* err);

// Q = U diag(lambda) Ui for reversible Q.  Returns false with *err set for
// NaN or a non-reversible Q.
bool eigen_reversible(int n, const double* Q, double* U, double* lambda, double* Ui,
                      std::string* err);

// Symmetric Jacobi: A = V diag(w) V^T; V columns are eigenvectors (row-major V).
void jacobi_sym(int n, const double* A, double* V, double* w);

// P = U diag(exp(lambda t)) Ui with the reference's special cases
// (t == -1 -> Q, t < 1e-10 -> I).
void compose(int n, const double* U, const double* lambda, const double* Ui, double t,
             double* P);

bool gtr_q(int n, const double* pi, const double* co, double* Q, std::string* err);

bool gamma_rates(int mode, double a, double b, int n, double* rates, std::string* err);

}  // namespace phy
