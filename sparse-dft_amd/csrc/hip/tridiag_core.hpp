// Arithmetic of the symmetric tridiagonal eigensolver, shared by the device
// kernels (tridiag_eig.hip) and host-side checks: the Sturm count of T - x I
// and the row-pivoted LU factorisation of T - x I with its solve.
//
// T has diagonal d[0..n) and off-diagonal e[0..n-1) (e[i] couples i and i+1).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define SDFT_HOST_DEVICE __host__ __device__
#else
#define SDFT_HOST_DEVICE
#endif

namespace sdft_hip {

// Number of eigenvalues of T not above x: the signs of the pivots of the LDL^T
// factorisation of T - x I (LAPACK dlaebz), e2[i] = e[i]^2; a pivot smaller
// than pivmin in magnitude is replaced by -pivmin.
SDFT_HOST_DEVICE inline int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q <= 0.0;
  for (int j = 1; j < n; ++j) {
    q = d[j] - e2[j - 1] / q - x;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q <= 0.0;
  }
  return c;
}

// One elimination step of T - x I with partial pivoting: (p, q) are the
// entries at columns i and i + 1 of the row not yet used as pivot, s, a, b the
// entries of row i + 1 at columns i, i + 1, i + 2. Writes row i of U (1 / u0,
// u1, u2), the multiplier and the exchange flag; pivots of magnitude below
// pert become +-pert (the perturbation of LAPACK dlagts, job = -1), so that
// the solve stays finite at an eigenvalue.
SDFT_HOST_DEVICE inline void lu_step(double& p, double& q, double s, double a, double b, double pert, double* inv0,
                                     double* u1, double* u2, double* l, unsigned char* piv) {
  double u;
  if (fabs(p) >= fabs(s)) {
    const double m = p != 0.0 ? s / p : 0.0;
    u = p;
    *u1 = q;
    *u2 = 0.0;
    *l = m;
    *piv = 0;
    p = a - m * q;
    q = b;
  } else {
    const double m = p / s;
    u = s;
    *u1 = a;
    *u2 = b;
    *l = m;
    *piv = 1;
    p = q - m * a;
    q = -m * b;
  }
  if (fabs(u) < pert) u = u < 0.0 ? -pert : pert;
  *inv0 = 1.0 / u;
}

// T - x I = P L U by Gaussian elimination with partial pivoting between rows i
// and i + 1: l[i] is the multiplier of step i, piv[i] = 1 when rows i and
// i + 1 were exchanged; U has the diagonals u0 (stored as 1 / u0), u1 and u2.
// The operands of C steps are read ahead of the dependent chain, so that one
// memory latency serves C steps (the arithmetic is that of the plain loop).
template <int C = 8>
SDFT_HOST_DEVICE inline void shifted_lu(const double* d, const double* e, int n, double x, double pert,
                                        double* inv0, double* u1, double* u2, double* l, unsigned char* piv) {
  double p = d[0] - x;
  double q = n > 1 ? e[0] : 0.0;
  int i = 0;
  for (; i + C + 1 < n; i += C) {
    double sv[C], av[C], bv[C];
#pragma unroll
    for (int t = 0; t < C; ++t) {
      sv[t] = e[i + t];
      av[t] = d[i + 1 + t] - x;
      bv[t] = e[i + 1 + t];
    }
#pragma unroll
    for (int t = 0; t < C; ++t)
      lu_step(p, q, sv[t], av[t], bv[t], pert, inv0 + i + t, u1 + i + t, u2 + i + t, l + i + t, piv + i + t);
  }
  for (; i + 1 < n; ++i)
    lu_step(p, q, e[i], d[i + 1] - x, i + 2 < n ? e[i + 1] : 0.0, pert, inv0 + i, u1 + i, u2 + i, l + i, piv + i);
  if (fabs(p) < pert) p = p < 0.0 ? -pert : pert;
  inv0[n - 1] = 1.0 / p;
  u1[n - 1] = 0.0;
  u2[n - 1] = 0.0;
  l[n - 1] = 0.0;
  piv[n - 1] = 0;
}

// y <- (T - x I)^(-1) y with the factors of shifted_lu; operands read C steps
// ahead as in shifted_lu.
template <int C = 8>
SDFT_HOST_DEVICE inline void shifted_solve(const double* inv0, const double* u1, const double* u2, const double* l,
                                           const unsigned char* piv, int n, double* y) {
  double cur = y[0];
  int i = 0;
  for (; i + C < n; i += C) {
    double yn[C], lv[C];
    unsigned char pv[C];
#pragma unroll
    for (int t = 0; t < C; ++t) {
      yn[t] = y[i + 1 + t];
      lv[t] = l[i + t];
      pv[t] = piv[i + t];
    }
#pragma unroll
    for (int t = 0; t < C; ++t) {
      double nxt = yn[t];
      if (pv[t]) {
        const double tmp = cur;
        cur = nxt;
        nxt = tmp;
      }
      y[i + t] = cur;
      cur = nxt - lv[t] * cur;
    }
  }
  for (; i + 1 < n; ++i) {
    double nxt = y[i + 1];
    if (piv[i]) {
      const double tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
    y[i] = cur;
    cur = nxt - l[i] * cur;
  }
  y[n - 1] = cur;
  double y1 = cur * inv0[n - 1], y2 = 0.0;
  y[n - 1] = y1;
  i = n - 2;
  for (; i - C + 1 >= 0; i -= C) {
    double yv[C], av[C], bv[C], iv[C];
#pragma unroll
    for (int t = 0; t < C; ++t) {
      yv[t] = y[i - t];
      av[t] = u1[i - t];
      bv[t] = u2[i - t];
      iv[t] = inv0[i - t];
    }
#pragma unroll
    for (int t = 0; t < C; ++t) {
      const double v = (yv[t] - av[t] * y1 - bv[t] * y2) * iv[t];
      y[i - t] = v;
      y2 = y1;
      y1 = v;
    }
  }
  for (; i >= 0; --i) {
    const double v = (y[i] - u1[i] * y1 - u2[i] * y2) * inv0[i];
    y[i] = v;
    y2 = y1;
    y1 = v;
  }
}

// Deterministic start vector entry in [-1, 1) of inverse iteration for
// eigenvalue index j (splitmix64 of (j, i)).
SDFT_HOST_DEVICE inline double start_entry(int j, int i) {
  unsigned long long z = (static_cast<unsigned long long>(j) << 32 | static_cast<unsigned int>(i)) +
                         0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return static_cast<double>(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

}  // namespace sdft_hip
