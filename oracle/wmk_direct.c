/* Direct wavemaker sums of the reference, restated in C.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline of the
 * additional-potential sums.  Only the test suite and bench.py's cpu_baseline
 * leg load it (through oracle/wmk_direct.py); the product path never does.
 *
 * phiadd_sl_rows    : phiadd_SL (sources/main/resol_wmkr.f90:1124-1187), 5 fields
 *                     at z = eta, the reference's loop order i2, i1, i3, j2 with
 *                     cos/sin of kx_add (eta + 1) and cos/sin(ky cy) per iteration;
 * phiadd_1st_rows   : the 12 z = 0 sums of phiadd_1st_SWEET
 *                     (resol_wmkr.f90:485-532), same loop order;
 * tables            : csh_add_x, k_add_sh_add_x, kycsh_add_x, kx_add_csh_add_x of
 *                     HOS-NWT.f90:364-405 (exponential underflow guards at 700),
 *                     built for the rows [i1_lo, i1_hi) only (table row r = i1 - i1_lo).
 *
 * Arrays are column-major with the first index fastest: a(i3, j2) is a[i3 + n3a*j2],
 * fields (i1, i2) are f[i1 + n1*i2], tables (r, j2, i3) are t[r + nr*(j2 + n2*i3)].
 * Rows are independent; OpenMP splits them over threads.
 */
#include <math.h>
#include <stdlib.h>

static void build_tables(int n1, int n2, int n3a, int i1_lo, int i1_hi, double xlen, double ylen, double* csh,
                         double* ksh, double* kycsh, double* kxcsh) {
  const int nr = i1_hi - i1_lo;
  const double xmax = 700.0;
  for (int i3 = 0; i3 < n3a; ++i3) {
    const double kx = (2.0 * i3 + 1.0) * M_PI / 4.0; /* kx_add(i3) = (2 i3 - 1) pi / xlen_add, xlen_add = 4 */
    for (int j2 = 0; j2 < n2; ++j2) {
      const double ky = n2 > 1 ? (double)j2 * M_PI / ylen : 0.0;
      const double kk = kx * kx + ky * ky;
      const double k = sqrt(kk);
      const double k2 = 2.0 * k; /* k_add_2 */
      const double e3 = (k2 * xlen <= xmax) ? exp(-k2 * xlen) : 0.0;
      const double e3p1 = e3 + 1.0;
      for (int r = 0; r < nr; ++r) {
        const double cx = (double)(i1_lo + r) * xlen / (double)(n1 - 1);
        const double x1 = k * cx;
        const double e1 = (x1 <= xmax) ? exp(-x1) / e3p1 : 0.0;
        const double x2 = k2 * (xlen - cx);
        const double e2 = (x2 <= xmax) ? exp(-x2) : 0.0;
        const double e12 = (x1 + x2 <= xmax) ? e1 * e2 : 0.0;
        const size_t o = (size_t)r + (size_t)nr * ((size_t)j2 + (size_t)n2 * i3);
        csh[o] = e1 + e12;
        ksh[o] = (e1 - e12) * k;
        kycsh[o] = (e1 + e12) * ky;
        kxcsh[o] = (e1 + e12) * kx;
      }
    }
  }
}

/* phi, phit, phix, phiy, phiz (n1 x n2 each, rows [i1_lo, i1_hi) written) */
int phiadd_sl_rows(int n1, int n2, int n3a, int i1_lo, int i1_hi, double xlen, double ylen, const double* a,
                   const double* da, const double* eta, double* phi, double* phit, double* phix, double* phiy,
                   double* phiz) {
  const int nr = i1_hi - i1_lo;
  const size_t nt = (size_t)nr * n2 * n3a;
  double* csh = malloc(4 * nt * sizeof(double));
  if (!csh) return 1;
  double *ksh = csh + nt, *kycsh = ksh + nt, *kxcsh = kycsh + nt;
  build_tables(n1, n2, n3a, i1_lo, i1_hi, xlen, ylen, csh, ksh, kycsh, kxcsh);
  const double dy = n2 > 1 ? ylen / (double)(n2 - 1) : 0.0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i2 = 0; i2 < n2; ++i2) {
    for (int r = 0; r < nr; ++r) {
      const int i1 = i1_lo + r;
      const size_t o = (size_t)i1 + (size_t)n1 * i2;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
      const double cy = (double)i2 * dy;
      for (int i3 = 0; i3 < n3a; ++i3) {
        for (int j2 = 0; j2 < n2; ++j2) {
          const double kx = (2.0 * i3 + 1.0) * M_PI / 4.0;
          const double ky = n2 > 1 ? (double)j2 * M_PI / ylen : 0.0;
          const double arg = kx * (eta[o] + 1.0);
          const double c = cos(arg), s = sin(arg);
          const size_t t = (size_t)r + (size_t)nr * ((size_t)j2 + (size_t)n2 * i3);
          const double ay = a[i3 + (size_t)n3a * j2], day = da[i3 + (size_t)n3a * j2];
          s0 += ay * c * csh[t] * cos(ky * cy);
          s1 += day * c * csh[t] * cos(ky * cy);
          s2 -= ay * c * ksh[t] * cos(ky * cy);
          s3 -= ay * c * kycsh[t] * sin(ky * cy);
          s4 -= ay * s * kxcsh[t] * cos(ky * cy);
        }
      }
      phi[o] = s0;
      phit[o] = s1;
      phix[o] = s2;
      phiy[o] = s3;
      phiz[o] = s4;
    }
  }
  free(csh);
  return 0;
}

/* the 12 z = 0 fields of phiadd_1st_SWEET into out[f * n1 * n2 + (i1 + n1 i2)], f in the
 * reference's order phi, phit, phix, phiy, phiz, phizz, phixt, phiyt, phizt, phixx, phixz, phiyz */
int phiadd_1st_rows(int n1, int n2, int n3a, int i1_lo, int i1_hi, double xlen, double ylen, const double* a,
                    const double* da, double* out) {
  const int nr = i1_hi - i1_lo;
  const size_t nt = (size_t)nr * n2 * n3a, ns = (size_t)n1 * n2;
  double* csh = malloc(4 * nt * sizeof(double));
  if (!csh) return 1;
  double *ksh = csh + nt, *kycsh = ksh + nt, *kxcsh = kycsh + nt;
  build_tables(n1, n2, n3a, i1_lo, i1_hi, xlen, ylen, csh, ksh, kycsh, kxcsh);
  const double dy = n2 > 1 ? ylen / (double)(n2 - 1) : 0.0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i2 = 0; i2 < n2; ++i2) {
    for (int r = 0; r < nr; ++r) {
      const int i1 = i1_lo + r;
      const size_t o = (size_t)i1 + (size_t)n1 * i2;
      double s[12] = {0};
      const double cy = (double)i2 * dy;
      for (int i3 = 0; i3 < n3a; ++i3) {
        for (int j2 = 0; j2 < n2; ++j2) {
          const double kx = (2.0 * i3 + 1.0) * M_PI / 4.0;
          const double ky = n2 > 1 ? (double)j2 * M_PI / ylen : 0.0;
          const double k = sqrt(kx * kx + ky * ky);
          const double c = cos(kx), sn = sin(kx);
          const size_t t = (size_t)r + (size_t)nr * ((size_t)j2 + (size_t)n2 * i3);
          const double ay = a[i3 + (size_t)n3a * j2], day = da[i3 + (size_t)n3a * j2];
          s[0] += ay * c * csh[t] * cos(ky * cy);
          s[1] += day * c * csh[t] * cos(ky * cy);
          s[2] -= ay * c * ksh[t] * cos(ky * cy);
          s[3] -= ay * c * kycsh[t] * sin(ky * cy);
          s[4] -= ay * sn * kxcsh[t] * cos(ky * cy);
          s[5] -= ay * c * kxcsh[t] * kx * cos(ky * cy);
          s[6] -= day * c * ksh[t] * cos(ky * cy);
          s[7] -= day * c * kycsh[t] * sin(ky * cy);
          s[8] -= day * sn * kxcsh[t] * cos(ky * cy);
          s[9] += ay * c * csh[t] * k * k * cos(ky * cy);
          s[10] += ay * sn * ksh[t] * kx * cos(ky * cy);
          s[11] += ay * sn * ky * csh[t] * sin(ky * cy);
        }
      }
      for (int f = 0; f < 12; ++f) out[(size_t)f * ns + o] = s[f];
    }
  }
  free(csh);
  return 0;
}
