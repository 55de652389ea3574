/*
 * The reference CPU path of helios's r-ccsd, restated in C with the
 * reference's own loop nests and OpenMP pattern (one parallel reduction per
 * term, inside the loop over excitation pairs).  TEST AND MEASUREMENT
 * INFRASTRUCTURE ONLY: the tests pin it against the numpy oracle, and
 * bench.py times it on the host cores as the CPU baseline.  Nothing in the
 * product links it.
 *
 * Reference files (paths relative to the helios source tree):
 *   R build        src/methods/coupledCluster/mod_cc.f90:547-558 (loop),
 *                  2258-2277 (spat_hf_int)
 *   guess          mod_cc.f90:634-650
 *   singles        src/methods/coupledCluster/ccsd/proj-ccsd-spat-hf-1.f90:14-119
 *   doubles        src/methods/coupledCluster/ccsd/proj-ccsd-spat-hf-2.f90:14-256
 *   step           mod_cc.f90:987-1017
 *   DIIS           mod_cc.f90:1390-1423 (push), 1514-1555 (extrapolation)
 *   energy         mod_cc.f90:1916-1952
 *   iterator       src/calculation/gIterator.f90:165-206
 *
 * Storage follows the reference (Fortran order, 0-based here): R(p,q,r,s) at
 * p + n(q + n(r + n s)); F(p,q) at p + n q; t1(i,a) at i + o A and
 * t2(i,j,a,b) at i + o(j + o(A + v B)) with A = a - o, B = b - o for the
 * absolute virtual indices a, b (o .. n-1); excitation mm = i v + A.
 *
 * One liberty, stated: the DIIS bordered system is solved by Gaussian
 * elimination with partial pivoting on B scaled by its largest diagonal
 * instead of the reference's eigen-decomposition (gltred4); the coefficients
 * are the same in exact arithmetic and the solve is a negligible part of an
 * iteration.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define RX(p, q, r, s) R[(p) + n * ((q) + n * ((r) + (size_t)n * (s)))]
#define FX(p, q) F[(p) + n * (q)]
#define T1(i, a) t1[(i) + o * ((a)-o)]
#define T2(i, j, a, b) t2[(i) + o * ((j) + o * (((a)-o) + (size_t)v * ((b)-o)))]
#define D2(i, j, a, b) d2[(i) + o * ((j) + o * (((a)-o) + (size_t)v * ((b)-o)))]
#define D1(i, a) d1[(i) + o * ((a)-o)]

/* R(a,b,c,d) = 2(ab|cd) - (ad|cb) with ZDO integrals (spat_hf_int), for all
 * n^4 index sets, parallel over the outer index as the reference. */
void ref_build_R(int n, const double* C, const double* G, double* R) {
#pragma omp parallel for
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
          /* spat_hf_int(i, j, a, b): sum_mu,nu G(mu,nu) (2 C_mi C_mj C_nb - C_mi C_mb C_nj) C_na */
          double sum = 0.0;
          for (int mu = 0; mu < n; ++mu) {
            const double ir1 = 2.0 * C[mu + n * i] * C[mu + n * j];
            const double ir2 = C[mu + n * i] * C[mu + n * b];
            for (int nu = 0; nu < n; ++nu) {
              const double irt = (ir1 * C[nu + n * b] - ir2 * C[nu + n * j]) * C[nu + n * a];
              sum += irt * G[mu + n * nu];
            }
          }
          R[i + n * (j + n * (a + (size_t)n * b))] = sum;
        }
}

/* guessCC r-ccsd: t1 = F(i,a)/(F(i,i)-F(a,a)); t2 = R(i,a,j,b)/D, mirrored */
void ref_guess(int n, int o, const double* F, const double* R, double* t1, double* t2) {
  const int v = n - o, ne = o * v;
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    T1(i, a) = FX(i, a) / (FX(i, i) - FX(a, a));
    for (int nn = mm; nn < ne; ++nn) {
      const int j = nn / v, b = o + nn % v;
      T2(i, j, a, b) = RX(i, a, j, b) / (FX(i, i) + FX(j, j) - FX(a, a) - FX(b, b));
      T2(j, i, b, a) = T2(i, j, a, b);
    }
  }
}

/* projection_ccsd_singles_spatial_hf */
void ref_singles(int n, int o, const double* F, const double* R, const double* t1, const double* t2, double* d1) {
  const int v = n - o, ne = o * v;
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    double rez = FX(i, a), sum;
    sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int j = 0; j < o; ++j)
      for (int b = o; b < n; ++b) sum += RX(a, i, j, b) * T1(j, b);
    rez += sum;
    sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int b = o; b < n; ++b) sum += FX(a, b) * T1(i, b);
    rez += sum;
    sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int j = 0; j < o; ++j) sum += FX(i, j) * T1(j, a);
    rez -= sum;
    sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int j = 0; j < o; ++j)
      for (int b = o; b < n; ++b) sum += FX(j, b) * (2 * T2(i, j, a, b) - T2(j, i, a, b));
    rez += sum;
    sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int j = 0; j < o; ++j)
      for (int b = o; b < n; ++b)
        for (int c = o; c < n; ++c) sum += RX(a, b, j, c) * (T2(i, j, b, c) + T1(i, b) * T1(j, c));
    rez += sum;
    sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int j = 0; j < o; ++j)
      for (int k = 0; k < o; ++k)
        for (int b = o; b < n; ++b) sum += RX(k, b, j, i) * (T2(j, k, a, b) + T1(j, a) * T1(k, b));
    rez -= sum;
    sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int j = 0; j < o; ++j)
      for (int b = o; b < n; ++b) sum += FX(j, b) * T1(i, b) * T1(j, a);
    rez -= sum;
    sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (int j = 0; j < o; ++j)
      for (int k = 0; k < o; ++k)
        for (int b = o; b < n; ++b)
          for (int c = o; c < n; ++c)
            sum += RX(k, c, j, b) * (T1(k, c) * (2 * T2(i, j, a, b) - T2(j, i, a, b)) - T1(i, b) * T2(j, k, a, c) -
                                     T1(k, a) * (T2(j, i, b, c) + T1(j, b) * T1(i, c)));
    rez += sum;
    D1(i, a) = rez;
  }
}

/* projection_ccsd_doubles_spatial_hf: every pair mm <= nn, then the mirror */
void ref_doubles(int n, int o, const double* F, const double* R, const double* t1, const double* t2, double* d2) {
  const int v = n - o, ne = o * v;
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    for (int nn = mm; nn < ne; ++nn) {
      const int j = nn / v, b = o + nn % v;
      double rez = RX(a, i, b, j), sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int c = o; c < n; ++c) sum += RX(b, j, a, c) * T1(i, c) + RX(a, i, b, c) * T1(j, c);
      rez += sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int k = 0; k < o; ++k) sum += RX(k, i, b, j) * T1(k, a) + RX(k, j, a, i) * T1(k, b);
      rez -= sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int k = 0; k < o; ++k)
        for (int l = 0; l < o; ++l) sum += RX(k, i, l, j) * (T2(k, l, a, b) + T1(k, a) * T1(l, b));
      rez += sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int c = o; c < n; ++c)
        for (int d = o; d < n; ++d) sum += RX(b, d, a, c) * (T2(i, j, c, d) + T1(i, c) * T1(j, d));
      rez += sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int k = 0; k < o; ++k)
        for (int c = o; c < n; ++c) {
          sum += 2 * (T2(i, k, a, c) * RX(b, j, k, c) + T2(j, k, b, c) * RX(a, i, k, c) +
                      (-T2(i, k, c, a) * RX(k, c, b, j) - T2(j, k, c, b) * RX(k, c, a, i) -
                       T2(j, k, a, c) * RX(k, c, b, i) - T2(i, k, b, c) * RX(k, c, a, j) -
                       T2(k, j, a, c) * RX(k, i, b, c) - T2(k, i, b, c) * RX(k, j, a, c)) / 2);
          sum += -RX(a, c, k, j) * T1(i, c) * T1(k, b) - RX(b, c, k, i) * T1(j, c) * T1(k, a) -
                 RX(k, c, b, j) * T1(i, c) * T1(k, a) - RX(k, c, a, i) * T1(j, c) * T1(k, b);
        }
      rez += sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int k = 0; k < o; ++k)
        for (int c = o; c < n; ++c)
          for (int d = o; d < n; ++d) {
            sum += RX(k, c, b, d) * (T1(j, d) * (2 * T2(i, k, a, c) - T2(k, i, a, c)) - T2(j, k, a, c) * T1(i, d) -
                                     T2(k, j, a, d) * T1(i, c)) +
                   RX(k, c, a, d) * (T1(i, d) * (2 * T2(j, k, b, c) - T2(k, j, b, c)) - T2(i, k, b, c) * T1(j, d) -
                                     T2(k, i, b, d) * T1(j, c));
            sum += T1(k, d) * (RX(k, d, b, c) * (2 * T2(i, j, a, c) - T2(i, j, c, a)) +
                               RX(k, d, a, c) * (2 * T2(j, i, b, c) - T2(j, i, c, b)));
            sum += -RX(a, c, k, d) * T1(k, b) * (T2(i, j, c, d) + T1(i, c) * T1(j, d)) -
                   RX(b, c, k, d) * T1(k, a) * (T2(j, i, c, d) + T1(j, c) * T1(i, d));
          }
      rez += sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int k = 0; k < o; ++k)
        for (int c = o; c < n; ++c)
          for (int l = 0; l < o; ++l) {
            sum += RX(k, c, l, j) * (T1(l, b) * (-2 * T2(i, k, a, c) + T2(k, i, a, c)) + T2(i, k, b, c) * T1(l, a) +
                                     T2(i, l, c, b) * T1(k, a)) +
                   RX(k, c, l, i) * (T1(l, a) * (-2 * T2(j, k, b, c) + T2(k, j, b, c)) + T2(j, k, a, c) * T1(l, b) +
                                     T2(j, l, c, a) * T1(k, b));
            sum += T1(l, c) * (RX(k, j, l, c) * (-2 * T2(i, k, a, b) + T2(k, i, a, b)) +
                               RX(k, i, l, c) * (-2 * T2(j, k, b, a) + T2(k, j, b, a)));
            sum += RX(k, c, l, j) * T1(i, c) * (T2(k, l, a, b) + T1(k, a) * T1(l, b)) +
                   RX(k, c, l, i) * T1(j, c) * (T2(k, l, b, a) + T1(k, b) * T1(l, a));
          }
      rez += sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int k = 0; k < o; ++k)
        for (int c = o; c < n; ++c)
          for (int l = 0; l < o; ++l)
            for (int d = o; d < n; ++d) {
              const double inner =
                  -2 * (T2(l, j, d, b) * T1(i, c) * T1(k, a) + T2(l, i, d, a) * T1(j, c) * T1(k, b) +
                        T2(l, j, a, b) * T1(k, c) * T1(i, d) + T2(l, i, b, a) * T1(k, c) * T1(j, d) +
                        T2(i, j, a, c) * T1(k, b) * T1(l, d) + T2(j, i, b, c) * T1(k, a) * T1(l, d) +
                        T2(k, i, c, d) * T2(l, j, a, b) + T2(k, j, c, d) * T2(l, i, b, a) +
                        T2(i, j, a, c) * T2(k, l, b, d) + T2(j, i, b, c) * T2(k, l, a, d) +
                        T2(i, k, c, a) * T2(l, j, d, b) + T2(j, k, c, b) * T2(l, i, d, a) -
                        T2(i, k, a, c) * T2(l, j, d, b) - T2(j, k, b, c) * T2(l, i, d, a)) +
                  T2(j, l, d, b) * T1(i, c) * T1(k, a) + T2(i, l, d, a) * T1(j, c) * T1(k, b) +
                  T2(l, i, d, b) * T1(j, c) * T1(k, a) + T2(l, j, d, a) * T1(i, c) * T1(k, b) +
                  T2(i, l, c, b) * T1(j, d) * T1(k, a) + T2(j, l, c, a) * T1(i, d) * T1(k, b) +
                  T2(l, i, a, b) * T1(k, c) * T1(j, d) + T2(l, j, b, a) * T1(k, c) * T1(i, d) +
                  T2(i, j, c, a) * T1(k, b) * T1(l, d) + T2(j, i, c, b) * T1(k, a) * T1(l, d) +
                  T2(k, l, a, b) * T1(i, c) * T1(j, d) + T2(i, j, c, d) * (T2(k, l, a, b) + T1(k, a) * T1(l, b)) +
                  T1(i, c) * T1(k, a) * T1(j, d) * T1(l, b) + T2(k, j, c, d) * T2(l, i, a, b) +
                  T2(k, i, c, d) * T2(l, j, b, a) + T2(i, j, b, c) * T2(k, l, a, d) +
                  T2(j, i, a, c) * T2(k, l, b, d) - T2(l, i, d, b) * (T2(j, k, a, c) - T2(j, k, c, a)) -
                  T2(l, j, d, a) * (T2(i, k, b, c) - T2(i, k, c, b)) +
                  (T2(i, k, c, a) * T2(j, l, d, b) + T2(j, k, c, b) * T2(i, l, d, a) +
                   T2(j, k, d, a) * T2(i, l, c, b) + T2(i, k, d, b) * T2(j, l, c, a)) /
                      2;
              sum += RX(k, c, l, d) * inner;
            }
      rez += sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int c = o; c < n; ++c)
        sum += FX(c, a) * (2 * T2(i, j, c, b) - T2(j, i, c, b)) + FX(c, b) * (2 * T2(j, i, c, a) - T2(i, j, c, a));
      rez += sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int k = 0; k < o; ++k)
        sum += FX(k, i) * (2 * T2(k, j, a, b) - T2(k, j, b, a)) + FX(k, j) * (2 * T2(k, i, b, a) - T2(k, i, a, b));
      rez -= sum;
      sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
      for (int k = 0; k < o; ++k)
        for (int c = o; c < n; ++c)
          sum += FX(k, c) * (-2 * (T2(k, j, a, b) * T1(i, c) + T2(k, i, b, a) * T1(j, c) + T2(i, j, a, c) * T1(k, b) +
                                   T2(j, i, b, c) * T1(k, a)) +
                             T2(k, i, a, b) * T1(j, c) + T2(k, j, b, a) * T1(i, c) + T2(j, i, a, c) * T1(k, b) +
                             T2(i, j, b, c) * T1(k, a));
      rez += sum;
      D2(i, j, a, b) = rez;
      D2(j, i, b, a) = rez;
    }
  }
}

/* energyCC r-ccsd: energy(1..4) */
void ref_energy(int n, int o, const double* F, const double* R, const double* t1, const double* t2, double eref,
                double* energy) {
  const int v = n - o, ne = o * v;
  double sum1 = 0.0, sum2 = 0.0, sum3 = 0.0;
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    sum1 += FX(i, a) * T1(i, a);
    for (int nn = mm + 1; nn < ne; ++nn) {
      const int j = nn / v, b = o + nn % v;
      sum2 += RX(a, i, b, j) * (T2(i, j, a, b) + T1(i, a) * T1(j, b));
    }
    sum3 += RX(a, i, a, i) * (T2(i, i, a, a) + T1(i, a) * T1(i, a));
  }
  energy[0] = 2 * sum1 + (2 * sum2 + sum3) + eref;
  energy[1] = eref;
  energy[2] = 2 * sum1;
  energy[3] = 2 * sum2 + sum3;
  energy[4] = 0.0;
}

/* stepCC r-ccsd; returns max(max|d1|, max|d2|) in acc[0..1] */
static void ref_step(int n, int o, const double* F, double* t1, double* t2, const double* d1, const double* d2,
                     double s1, double s2, int projtype, double* acc) {
  const int v = n - o, ne = o * v;
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    T1(i, a) += s1 * D1(i, a) / (FX(i, i) - FX(a, a));
  }
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    for (int nn = mm; nn < ne; ++nn) {
      const int j = nn / v, b = o + nn % v;
      const double num = projtype == 1 ? D2(i, j, a, b) : 2 * D2(i, j, a, b) + D2(i, j, b, a);
      T2(i, j, a, b) += s2 * num / (FX(i, i) - FX(a, a) + FX(j, j) - FX(b, b));
      T2(j, i, b, a) = T2(i, j, a, b);
    }
  }
  double m1 = 0.0, m2 = 0.0;
  for (long e = 0; e < (long)ne; ++e) m1 = fmax(m1, fabs(d1[e]));
  for (long e = 0; e < (long)ne * ne; ++e) m2 = fmax(m2, fabs(d2[e]));
  acc[0] = m1;
  acc[1] = m2;
}

/* DIIS history (pushCCVectors / newCCVectors): vectors packed over mm <= nn */
typedef struct {
  int nd, len, count;
  double *st, *sd, *B;   /* st, sd: nd slots of len (newest first); B: (nd+1)^2 */
} Diis;

static void diis_push(Diis* h, int n, int o, const double* t1, const double* t2, const double* d1, const double* d2,
                      int projtype) {
  const int v = n - o, ne = o * v;
  memmove(h->st + h->len, h->st, sizeof(double) * (size_t)h->len * (h->nd - 1));
  memmove(h->sd + h->len, h->sd, sizeof(double) * (size_t)h->len * (h->nd - 1));
  long q = 0;
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    h->sd[q] = D1(i, a);
    h->st[q++] = T1(i, a);
  }
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    for (int nn = mm; nn < ne; ++nn) {
      const int j = nn / v, b = o + nn % v;
      h->st[q] = T2(i, j, a, b);
      h->sd[q++] = projtype == 1 ? D2(i, j, a, b) : 2 * D2(i, j, a, b) + D2(i, j, b, a);
    }
  }
  if (h->count < h->nd) ++h->count;
}

/* newCCVectors(it): B shift, new dot products, bordered solve, t = sum c_k st_k.
 * Returns 0 when no extrapolation happened (it == 1 or a singular system). */
static int diis_extrapolate(Diis* h, int it, int n, int o, double* t1, double* t2) {
  const int m1 = h->nd + 1, v = n - o, ne = o * v;
  double* B = h->B;
  for (int k = it - 1; k >= 1; --k)
    for (int l = it - 1; l >= 1; --l) B[k + m1 * l] = B[(k - 1) + m1 * (l - 1)];
  for (int k = 0; k < it; ++k) {
    double s = 0.0;
    for (long e = 0; e < h->len; ++e) s += h->sd[e + (long)h->len * k] * h->sd[e];
    B[k] = B[m1 * k] = s;
  }
  if (it == 1) return 0;
  const int ub = it + 1;
  double A[33 * 34], x[33];
  double scale = 0.0;
  for (int k = 0; k < it; ++k) scale = fmax(scale, B[k + m1 * k]);
  if (!(scale > 0.0)) scale = 1.0;
  for (int r = 0; r < ub; ++r) {
    for (int c = 0; c < ub; ++c)
      A[r * (ub + 1) + c] = (r < it && c < it) ? B[r + m1 * c] / scale : (r == it && c == it ? 0.0 : 1.0);
    A[r * (ub + 1) + ub] = r == it ? 1.0 : 0.0;
  }
  for (int c = 0; c < ub; ++c) {
    int piv = c;
    for (int r = c + 1; r < ub; ++r)
      if (fabs(A[r * (ub + 1) + c]) > fabs(A[piv * (ub + 1) + c])) piv = r;
    if (!(fabs(A[piv * (ub + 1) + c]) > 1e-300)) return 0;
    for (int q = 0; q <= ub; ++q) {
      const double tmp = A[c * (ub + 1) + q];
      A[c * (ub + 1) + q] = A[piv * (ub + 1) + q];
      A[piv * (ub + 1) + q] = tmp;
    }
    for (int r = c + 1; r < ub; ++r) {
      const double f = A[r * (ub + 1) + c] / A[c * (ub + 1) + c];
      for (int q = c; q <= ub; ++q) A[r * (ub + 1) + q] -= f * A[c * (ub + 1) + q];
    }
  }
  for (int r = ub - 1; r >= 0; --r) {
    double s = A[r * (ub + 1) + ub];
    for (int q = r + 1; q < ub; ++q) s -= A[r * (ub + 1) + q] * x[q];
    x[r] = s / A[r * (ub + 1) + r];
  }
  long q = 0;
  for (int mm = 0; mm < ne; ++mm, ++q) {
    const int i = mm / v, a = o + mm % v;
    double s = 0.0;
    for (int k = 0; k < it; ++k) s += x[k] * h->st[q + (long)h->len * k];
    T1(i, a) = s;
  }
  for (int mm = 0; mm < ne; ++mm) {
    const int i = mm / v, a = o + mm % v;
    for (int nn = mm; nn < ne; ++nn, ++q) {
      const int j = nn / v, b = o + nn % v;
      double s = 0.0;
      for (int k = 0; k < it; ++k) s += x[k] * h->st[q + (long)h->len * k];
      T2(i, j, a, b) = s;
      T2(j, i, b, a) = s;
    }
  }
  return 1;
}

/* Whole r-ccsd solve after the SCF (initCC's R build and guess, then the
 * iterator around iterationCC and energyCC) with `threads` OpenMP threads.
 * Returns the status (1, -1, -2, -3) and fills energy[5], *iterations, t1, t2
 * (any output may be null). */
int ref_solve(int n, int o, const double* C, const double* F, const double* G, double eref, double eps, int maxiter,
              int diis_steps, int threads, double* energy, int* iterations, double* t1_out, double* t2_out) {
  if (threads > 0) omp_set_num_threads(threads);
  const int v = n - o, ne = o * v;
  const size_t n4 = (size_t)n * n * n * n, nt2 = (size_t)ne * ne;
  double* R = malloc(sizeof(double) * n4);
  double* t1 = calloc(ne, sizeof(double));
  double* t2 = calloc(nt2, sizeof(double));
  double* d1 = calloc(ne, sizeof(double));
  double* d2 = calloc(nt2, sizeof(double));
  Diis h;
  h.nd = diis_steps;
  h.len = ne + ne * (ne + 1) / 2;
  h.count = 0;
  h.st = calloc((size_t)h.len * diis_steps, sizeof(double));
  h.sd = calloc((size_t)h.len * diis_steps, sizeof(double));
  h.B = calloc((size_t)(diis_steps + 1) * (diis_steps + 1), sizeof(double));
  int status = 0, it = 0;
  if (!R || !t1 || !t2 || !d1 || !d2 || !h.st || !h.sd || !h.B) {
    status = -100;
    goto out;
  }
  ref_build_R(n, C, G, R);
  ref_guess(n, o, F, R, t1, t2);
  double cur = 0.0, prev = 10.0;
  int successful = 1;
  for (;;) {
    ++it;
    double acc[2];
    ref_singles(n, o, F, R, t1, t2, d1);
    ref_doubles(n, o, F, R, t1, t2, d2);
    ref_step(n, o, F, t1, t2, d1, d2, 1.0, 1.0, 2, acc);
    const double a = fmax(acc[0], acc[1]);
    if (!(a < eps) && diis_steps > 1) {
      diis_push(&h, n, o, t1, t2, d1, d2, 2);
      diis_extrapolate(&h, it < diis_steps ? it : diis_steps, n, o, t1, t2);
    }
    prev = cur;
    cur = a;
    if (cur < prev) ++successful;
    if (cur < eps) { status = 1; break; }
    if (!(cur <= 1e3)) { status = -2; break; }
    if (it + 1 > maxiter) { status = -1; break; }
    if (it > 200 && 100.0 * successful / it <= 60.0) { status = -3; break; }
  }
  if (energy) ref_energy(n, o, F, R, t1, t2, eref, energy);
  if (t1_out) memcpy(t1_out, t1, sizeof(double) * ne);
  if (t2_out) memcpy(t2_out, t2, sizeof(double) * nt2);
out:
  if (iterations) *iterations = it;
  free(R); free(t1); free(t2); free(d1); free(d2); free(h.st); free(h.sd); free(h.B);
  return status;
}

/* Wall time (omp_get_wtime) of the R build and of `iters` iterations
 * (residuals, step, DIIS push and extrapolation) from the guess, with
 * `threads` OpenMP threads: the parts of a solve, for timing samples whose
 * full solve would take too long. */
int ref_time_parts(int n, int o, const double* C, const double* F, const double* G, int iters, int diis_steps,
                   int threads, double* t_build, double* t_iters) {
  if (threads > 0) omp_set_num_threads(threads);
  const int v = n - o, ne = o * v;
  const size_t n4 = (size_t)n * n * n * n, nt2 = (size_t)ne * ne;
  double* R = malloc(sizeof(double) * n4);
  double* t1 = calloc(ne, sizeof(double));
  double* t2 = calloc(nt2, sizeof(double));
  double* d1 = calloc(ne, sizeof(double));
  double* d2 = calloc(nt2, sizeof(double));
  Diis h;
  h.nd = diis_steps;
  h.len = ne + ne * (ne + 1) / 2;
  h.count = 0;
  h.st = calloc((size_t)h.len * diis_steps, sizeof(double));
  h.sd = calloc((size_t)h.len * diis_steps, sizeof(double));
  h.B = calloc((size_t)(diis_steps + 1) * (diis_steps + 1), sizeof(double));
  int rc = 0;
  if (!R || !t1 || !t2 || !d1 || !d2 || !h.st || !h.sd || !h.B) {
    rc = -100;
  } else {
    double w0 = omp_get_wtime();
    ref_build_R(n, C, G, R);
    *t_build = omp_get_wtime() - w0;
    ref_guess(n, o, F, R, t1, t2);
    w0 = omp_get_wtime();
    for (int it = 1; it <= iters; ++it) {
      double acc[2];
      ref_singles(n, o, F, R, t1, t2, d1);
      ref_doubles(n, o, F, R, t1, t2, d2);
      ref_step(n, o, F, t1, t2, d1, d2, 1.0, 1.0, 2, acc);
      diis_push(&h, n, o, t1, t2, d1, d2, 2);
      diis_extrapolate(&h, it < diis_steps ? it : diis_steps, n, o, t1, t2);
    }
    *t_iters = omp_get_wtime() - w0;
  }
  free(R); free(t1); free(t2); free(d1); free(d2); free(h.st); free(h.sd); free(h.B);
  return rc;
}
