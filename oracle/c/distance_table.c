/* Oracle (test infrastructure only): FP64 C restatement of bilby's distance
 * marginalisation lookup table, bilby/gw/likelihood/base.py:994-1018.
 *
 *   table[ii][jj] = logsumexp_k( f(d_jj * s_k) - h_ii * s_k^2 / 2, b = prior_k * delta )
 *                   - logsumexp_k( 0, b = prior_k * delta )
 *   s_k = dref / dist_k,  f = ln I0(|x|) (phase marginalisation) or identity.
 *
 * logsumexp as scipy.special.logsumexp (max over the entries with b != 0,
 * then log of the weighted sum of exp(a - max)).  ln I0 by its power series
 * (x < 30) and the large-argument expansion (Abramowitz & Stegun 9.7.1).
 * Built with: gcc -O2 -fopenmp -shared -fPIC (oracle/build_oracle.py). */
#include <math.h>

static double ln_i0(double x) {
  x = fabs(x);
  if (x < 30.0) {
    double q = 0.25 * x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 120; ++k) {
      term *= q / ((double)k * (double)k);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    return log(sum);
  }
  double y = 1.0 / (8.0 * x), term = 1.0, sum = 1.0;
  for (int k = 1; k < 40; ++k) {
    double a = 2.0 * k - 1.0, next = term * a * a * y / (double)k;
    if (next >= term) break;
    term = next;
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return x - 0.5 * log(2.0 * M_PI * x) + log(sum);
}

void oracle_distance_table(int nh, int nd, const double* haxis, const double* daxis, int nk,
                           const double* dist, const double* prior, double dref, int phase,
                           double* out) {
  const double delta = dist[1] - dist[0];
  double norm_max = -INFINITY, norm_sum = 0.0;
  for (int k = 0; k < nk; ++k) {
    double b = prior[k] * delta;
    if (b != 0.0) norm_max = 0.0;
  }
  for (int k = 0; k < nk; ++k) norm_sum += prior[k] * delta * exp(0.0 - norm_max);
  const double log_norm = log(norm_sum) + norm_max;
#pragma omp parallel for schedule(dynamic, 4)
  for (int jj = 0; jj < nd; ++jj) {
    double fterm[nk > 0 ? nk : 1];
    double s2[nk > 0 ? nk : 1];
    for (int k = 0; k < nk; ++k) {
      double s = dref / dist[k];
      double v = daxis[jj] * s;
      fterm[k] = phase ? ln_i0(fabs(v)) : v;
      s2[k] = s * s;
    }
    for (int ii = 0; ii < nh; ++ii) {
      const double h = haxis[ii];
      double m = -INFINITY;
      for (int k = 0; k < nk; ++k) {
        if (prior[k] * delta == 0.0) continue;
        double a = fterm[k] - h * s2[k] / 2;
        if (a > m) m = a;
      }
      double s = 0.0;
      for (int k = 0; k < nk; ++k) s += prior[k] * delta * exp(fterm[k] - h * s2[k] / 2 - m);
      out[(long)ii * nd + jj] = (log(s) + m) - log_norm;
    }
  }
}
