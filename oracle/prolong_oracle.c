/*
 * prolong_oracle.c -- CPU oracle for EnzoProlong's conservative
 * interpolation of coarse ghost data onto a finer block.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of EnzoProlong::apply_
 * (src/Enzo/mesh/EnzoProlong.cpp:112-190) and the Fortran routines it
 * calls, for rank 3 and refinement 2:
 *   interpolate  (src/Enzo/mesh/interpolate.F:1-299, the dispatch),
 *   tscint3d     (src/Enzo/hydro-mhd/ppm_fortran/tscint3d.F, method "3A"),
 *   interp3d     (src/Enzo/hydro-mhd/ppm_fortran/interp3d.F, "2A"),
 *   int_lin3d    (src/Enzo/hydro-mhd/ppm_fortran/int_lin3d.F, "2B"),
 * loop by loop, with the routines' weights, summation order and branch
 * tests (Fortran indices are kept 1-based through the macros P and G).
 * Only tests/ load it.
 *
 * Pins (tests/test_oracle_prolong.py): closed forms -- a constant coarse
 * field gives the constant, "2A"/"2B" reproduce a coarse field linear along
 * a cell diagonal exactly inside the block and are limited to flat children
 * at an extremum, "3A" conserves the coarse value of every coarse cell, all
 * three conserve to roundoff, "2B" with positivity keeps a positive field
 * positive -- and a brute-force evaluation of the same formulas in Python
 * on small random data.
 */
#include "vlct_oracle.h"

#include <math.h>
#include <stdlib.h>

#define TINY 1.0e-20 /* tiny of fortran_types.h for double precision */

/* Fortran SIGN(a, b): |a| with the sign of b (signed zeros honoured) */
static double fsign(double a, double b) { return copysign(fabs(a), b); }
/* Fortran MIN / MAX of two values */
static double fmin2(double a, double b) { return (b < a) ? b : a; }
static double fmax2(double a, double b) { return (b > a) ? b : a; }

/* parent (coarse) value at Fortran indices (ip, jp, kp), 1 = first cell of
 * the coarse region; grid (fine) value at (ig, jg, kg), 1 = first fine
 * cell of the region */
#define P(ip, jp, kp) pc[((ip) - 1) + (long)mc3[0] * (((jp) - 1) + (long)mc3[1] * ((kp) - 1))]
#define G(ig, jg, kg) gf[((ig) - 1) + (long)mf3[0] * (((jg) - 1) + (long)mf3[1] * ((kg) - 1))]

/* tscint3d with start = 3, end = nf + 2, refine = 2, gstart = 1 */
static void tscint3d(const double *pc, const int mc3[3], double *gf, const int mf3[3],
                     const int nf3[3]) {
  const int r = 2;
  double cm[3], cp[3], c0[3];  /* coef{m,p,0}(1..2) at [1..2] */
  for (int i = 0; i < r; i++) {
    cm[i + 1] = (double)((r - i) * (r - i) * (r - i) - (r - i - 1) * (r - i - 1) * (r - i - 1)) /
                (double)(6 * r * r * r);
    cp[i + 1] = (double)(3 * i * i + 3 * i + 1) / (double)(6 * r * r * r);
    c0[i + 1] = 1.0 / r - cm[i + 1] - cp[i + 1];
  }
  const int s1 = 3, s2 = 3, s3 = 3;
  const int e1 = nf3[0] + 2, e2 = nf3[1] + 2, e3 = nf3[2] + 2;
  for (int k = s3; k <= e3; k++) {
    const int kp = (k - 1) / r + 1, kc = k - (kp - 1) * r;
    const double wm3 = cm[kc], w03 = c0[kc], wp3 = cp[kc];
    for (int j = s2; j <= e2; j++) {
      const int jp = (j - 1) / r + 1, jc = j - (jp - 1) * r;
      const double wm2 = cm[jc], w02 = c0[jc], wp2 = cp[jc];
      for (int i = s1; i <= e1; i++) {
        const int ip = (i - 1) / r + 1, ic = i - (ip - 1) * r;
        const double wm1 = cm[ic], w01 = c0[ic], wp1 = cp[ic];
        G(i - s1 + 1, j - s2 + 1, k - s3 + 1) =
            wm1 * wm2 * wm3 * P(ip - 1, jp - 1, kp - 1) +
            w01 * wm2 * wm3 * P(ip, jp - 1, kp - 1) +
            wp1 * wm2 * wm3 * P(ip + 1, jp - 1, kp - 1) +
            wm1 * w02 * wm3 * P(ip - 1, jp, kp - 1) +
            w01 * w02 * wm3 * P(ip, jp, kp - 1) +
            wp1 * w02 * wm3 * P(ip + 1, jp, kp - 1) +
            wm1 * wp2 * wm3 * P(ip - 1, jp + 1, kp - 1) +
            w01 * wp2 * wm3 * P(ip, jp + 1, kp - 1) +
            wp1 * wp2 * wm3 * P(ip + 1, jp + 1, kp - 1) +
            wm1 * wm2 * w03 * P(ip - 1, jp - 1, kp) +
            w01 * wm2 * w03 * P(ip, jp - 1, kp) +
            wp1 * wm2 * w03 * P(ip + 1, jp - 1, kp) +
            wm1 * w02 * w03 * P(ip - 1, jp, kp) +
            w01 * w02 * w03 * P(ip, jp, kp) +
            wp1 * w02 * w03 * P(ip + 1, jp, kp) +
            wm1 * wp2 * w03 * P(ip - 1, jp + 1, kp) +
            w01 * wp2 * w03 * P(ip, jp + 1, kp) +
            wp1 * wp2 * w03 * P(ip + 1, jp + 1, kp) +
            wm1 * wm2 * wp3 * P(ip - 1, jp - 1, kp + 1) +
            w01 * wm2 * wp3 * P(ip, jp - 1, kp + 1) +
            wp1 * wm2 * wp3 * P(ip + 1, jp - 1, kp + 1) +
            wm1 * w02 * wp3 * P(ip - 1, jp, kp + 1) +
            w01 * w02 * wp3 * P(ip, jp, kp + 1) +
            wp1 * w02 * wp3 * P(ip + 1, jp, kp + 1) +
            wm1 * wp2 * wp3 * P(ip - 1, jp + 1, kp + 1) +
            w01 * wp2 * wp3 * P(ip, jp + 1, kp + 1) +
            wp1 * wp2 * wp3 * P(ip + 1, jp + 1, kp + 1);
      }
    }
  }
  /* conservation: each coarse cell's children sum to its value */
  const int wd1 = (e1 - s1 + 1) / r, wd2 = (e2 - s2 + 1) / r, wd3 = (e3 - s3 + 1) / r;
  double *temp = (double *)malloc(sizeof(double) * (size_t)(wd1 + 1));
  for (int k = 1; k <= wd3; k++)
    for (int j = 1; j <= wd2; j++) {
      for (int i = 1; i <= wd1; i++) temp[i] = 0.0;
      for (int ks = 1; ks <= r; ks++)
        for (int js = 1; js <= r; js++)
          for (int is = 1; is <= r; is++)
            for (int i = 1; i <= wd1; i++)
              temp[i] = temp[i] + G((i - 1) * r + is, (j - 1) * r + js, (k - 1) * r + ks);
      for (int i = 1; i <= wd1; i++)
        temp[i] = (P((s1 - 1) / r + i, (s2 - 1) / r + j, (s3 - 1) / r + k) - temp[i]) /
                  (double)(r * r * r);
      for (int ks = 1; ks <= r; ks++)
        for (int js = 1; js <= r; js++)
          for (int is = 1; is <= r; is++)
            for (int i = 1; i <= wd1; i++)
              G((i - 1) * r + is, (j - 1) * r + js, (k - 1) * r + ks) =
                  G((i - 1) * r + is, (j - 1) * r + js, (k - 1) * r + ks) + temp[i];
    }
  /* the two monotonicity iterations along x, y, z.  The routine's second
   * bound, max(max(min(d1right/d2right, 0), 1), fraction), is at least 1
   * and the first leaves fraction <= 1, so for finite data every fraction is
   * 1 and each pass sets the children of a coarse cell to their mean along
   * the axis (the del* values it computes do not reach the result) */
  const double frac = 1.0;
  for (int iter = 1; iter <= 2; iter++) {
    for (int k = s3; k <= e3; k++) {
      const int kg = k - s3 + 1;
      for (int j = s2; j <= e2; j++) {
        const int jg = j - s2 + 1;
        for (int i = 1; i <= wd1; i++) temp[i] = 0.0;
        for (int i1 = 1; i1 <= r; i1++)
          for (int i2 = 1; i2 <= wd1; i2++)
            temp[i2] = temp[i2] + G((i2 - 1) * r + i1, jg, kg) / (double)r;
        for (int i1 = 1; i1 <= r; i1++)
          for (int i2 = 1; i2 <= wd1; i2++)
            G((i2 - 1) * r + i1, jg, kg) =
                frac * temp[i2] + (1.0 - frac) * G((i2 - 1) * r + i1, jg, kg);
      }
    }
    double *tj = (double *)malloc(sizeof(double) * (size_t)(wd2 + 1));
    for (int k = s3; k <= e3; k++) {
      const int kg = k - s3 + 1;
      for (int i = s1; i <= e1; i++) {
        const int ig = i - s1 + 1;
        for (int j = 1; j <= wd2; j++) tj[j] = 0.0;
        for (int j1 = 1; j1 <= r; j1++)
          for (int j2 = 1; j2 <= wd2; j2++)
            tj[j2] = tj[j2] + G(ig, (j2 - 1) * r + j1, kg) / (double)r;
        for (int j1 = 1; j1 <= r; j1++)
          for (int j2 = 1; j2 <= wd2; j2++)
            G(ig, (j2 - 1) * r + j1, kg) =
                frac * tj[j2] + (1.0 - frac) * G(ig, (j2 - 1) * r + j1, kg);
      }
    }
    free(tj);
    double *tk = (double *)malloc(sizeof(double) * (size_t)(wd3 + 1));
    for (int i = s1; i <= e1; i++) {
      const int ig = i - s1 + 1;
      for (int j = s2; j <= e2; j++) {
        const int jg = j - s2 + 1;
        for (int k = 1; k <= wd3; k++) tk[k] = 0.0;
        for (int k1 = 1; k1 <= r; k1++)
          for (int k2 = 1; k2 <= wd3; k2++)
            tk[k2] = tk[k2] + G(ig, jg, (k2 - 1) * r + k1) / (double)r;
        for (int k1 = 1; k1 <= r; k1++)
          for (int k2 = 1; k2 <= wd3; k2++)
            G(ig, jg, (k2 - 1) * r + k1) =
                frac * tk[k2] + (1.0 - frac) * G(ig, jg, (k2 - 1) * r + k1);
      }
    }
    free(tk);
  }
  free(temp);
  for (int k = 1; k <= e3 - s3 + 1; k++)
    for (int j = 1; j <= e2 - s2 + 1; j++)
      for (int i = 1; i <= e1 - s1 + 1; i++)
        G(i, j, k) = G(i, j, k) * (double)r * (double)r * (double)r;
}

/* the corner averages work(i, j, k), 1 <= i <= pdim+1, of interp3d /
 * int_lin3d (pstart = 1): the mean of the 8 coarse cells around the corner */
static double corner(const double *pc, const int mc3[3], int i, int j, int k) {
  const int ps = 1;
  return 0.125 * (P(ps + i - 1, ps + j - 1, ps + k - 1) + P(ps + i - 1, ps + j - 1, ps + k) +
                  P(ps + i - 1, ps + j, ps + k - 1) + P(ps + i - 1, ps + j, ps + k) +
                  P(ps + i, ps + j - 1, ps + k - 1) + P(ps + i, ps + j - 1, ps + k) +
                  P(ps + i, ps + j, ps + k - 1) + P(ps + i, ps + j, ps + k));
}

/* the 8 children of coarse cell (i, j, k): fbar + (i1 - 1/2)/2 fx + ... */
static void children(double *gf, const int mf3[3], int i, int j, int k, double fbar,
                     double fx, double fy, double fz) {
  const int r = 2;
  for (int k1 = 0; k1 < r; k1++)
    for (int j1 = 0; j1 < r; j1++)
      for (int i1 = 0; i1 < r; i1++)
        G(1 + (i - 1) * r + i1, 1 + (j - 1) * r + j1, 1 + (k - 1) * r + k1) =
            (fbar + ((double)i1 + 0.5 - 0.5 * (double)r) / (double)r * fx +
             ((double)j1 + 0.5 - 0.5 * (double)r) / (double)r * fy +
             ((double)k1 + 0.5 - 0.5 * (double)r) / (double)r * fz);
}

/* interp3d ("2A"): returns 0, or -2 where the routine returns with its
 * error flag (a NaN coarse value; a non-positive child of a field whose
 * corner and cell values are all positive; a NaN slope): the children of
 * the rows before are filled, the rest untouched, as there */
static int interp3d(const double *pc, const int mc3[3], double *gf, const int mf3[3],
                    const int nf3[3]) {
  const int pd1 = nf3[0] / 2, pd2 = nf3[1] / 2, pd3 = nf3[2] / 2;
  const int ps = 1;
  const long nw = (long)(pd1 + 2) * (pd2 + 2) * (pd3 + 2);
  double *w = (double *)malloc(sizeof(double) * (size_t)nw);
#define W(i, j, k) w[(i) + (long)(pd1 + 2) * ((j) + (long)(pd2 + 2) * (k))]
  int nneg = 0;
  for (int k = 1; k <= pd3 + 1; k++)
    for (int j = 1; j <= pd2 + 1; j++)
      for (int i = 1; i <= pd1 + 1; i++) {
        W(i, j, k) = corner(pc, mc3, i, j, k);
        const double pv = P(ps + i, ps + j, ps + k);
        if (pv != pv) {
          free(w);
          return -2;
        }
        if (W(i, j, k) <= 0.0 || pv <= 0.0) nneg = nneg + 1;
      }
  double *fx = (double *)malloc(sizeof(double) * (size_t)(pd1 + 1));
  double *fy = (double *)malloc(sizeof(double) * (size_t)(pd1 + 1));
  double *fz = (double *)malloc(sizeof(double) * (size_t)(pd1 + 1));
  double *fb = (double *)malloc(sizeof(double) * (size_t)(pd1 + 1));
  int rc = 0;
  for (int k = 1; k <= pd3 && !rc; k++)
    for (int j = 1; j <= pd2 && !rc; j++) {
      for (int i = 1; i <= pd1; i++) {
        const double fbar = P(ps + i, ps + j, ps + k);
        double d0 = fmin2(fabs(fbar - W(i, j, k)), fabs(fbar - W(i + 1, j + 1, k + 1))) *
                    fsign(1.0, fbar - W(i, j, k));
        if ((W(i + 1, j + 1, k + 1) - fbar) * (fbar - W(i, j, k)) <= 0.0) d0 = 0.0;
        double d1 = fmin2(fabs(fbar - W(i + 1, j, k)), fabs(fbar - W(i, j + 1, k + 1))) *
                    fsign(1.0, fbar - W(i + 1, j, k));
        if ((W(i, j + 1, k + 1) - fbar) * (fbar - W(i + 1, j, k)) <= 0.0) d1 = 0.0;
        double d2 = fmin2(fabs(fbar - W(i, j + 1, k)), fabs(fbar - W(i + 1, j, k + 1))) *
                    fsign(1.0, fbar - W(i, j + 1, k));
        if ((W(i + 1, j, k + 1) - fbar) * (fbar - W(i, j + 1, k)) <= 0.0) d2 = 0.0;
        double d3 = fmin2(fabs(fbar - W(i, j, k + 1)), fabs(fbar - W(i + 1, j + 1, k))) *
                    fsign(1.0, fbar - W(i, j, k + 1));
        if ((W(i + 1, j + 1, k) - fbar) * (fbar - W(i, j, k + 1)) <= 0.0) d3 = 0.0;
        double s, sp;
        if (d0 == 0.0) {
          d0 = 1.0e-5 * fsign(TINY, (d1 + d2 + d3));
          sp = 1.0;
        } else {
          s = (d1 + d2 + d3) / d0;
          sp = fmin2(fmax2(1.0 * s, 0.0), 1.0);
        }
        double c1 = 1.0, c2 = 1.0, c3 = 1.0;
        if (d1 / d0 > 0.0) c1 = 0.0;
        if (d2 / d0 > 0.0) c2 = 0.0;
        if (d3 / d0 > 0.0) c3 = 0.0;
        c1 = (1.0 - sp) * c1 + sp * (1.0 - c1);
        c2 = (1.0 - sp) * c2 + sp * (1.0 - c2);
        c3 = (1.0 - sp) * c3 + sp * (1.0 - c3);
        if (sp != 0.0 && sp != 1.0) {
          c1 = 0.0;
          c2 = 0.0;
          c3 = 0.0;
        }
        double frac = -((-d0 * sp) + (1.0 - c1) * d1 + (1.0 - c2) * d2 + (1.0 - c3) * d3) /
                      (c1 * d1 + c2 * d2 + c3 * d3 + 1.0e-35);
        frac = fmin2(frac, 1.0);
        if (c1 + c2 + c3 == 0) frac = 0.0;
        frac = fmax2(frac, 0.0);
        const double d1n = frac * c1 * d1 + (1.0 - c1) * d1;
        const double d2n = frac * c2 * d2 + (1.0 - c2) * d2;
        const double d3n = frac * c3 * d3 + (1.0 - c3) * d3;
        s = d1n + d2n + d3n;
        if (nneg == 0 && (fbar + s <= 0.0 || fbar - s <= 0.0)) {
          rc = -2;
          break;
        }
        if (s != s) {
          rc = -2;
          break;
        }
        fx[i] = d2n + d3n;
        fy[i] = d1n + d3n;
        fz[i] = d1n + d2n;
        fb[i] = fbar;
      }
      if (rc) break;
      for (int i = 1; i <= pd1; i++) children(gf, mf3, i, j, k, fb[i], fx[i], fy[i], fz[i]);
    }
#undef W
  free(w);
  free(fx);
  free(fy);
  free(fz);
  free(fb);
  return rc;
}

/* int_lin3d ("2B"), iposflag = positive */
static void int_lin3d(const double *pc, const int mc3[3], double *gf, const int mf3[3],
                      const int nf3[3], int positive) {
  const int pd1 = nf3[0] / 2, pd2 = nf3[1] / 2, pd3 = nf3[2] / 2;
  const int ps = 1;
  const long nw = (long)(pd1 + 2) * (pd2 + 2) * (pd3 + 2);
  double *w = (double *)malloc(sizeof(double) * (size_t)nw);
#define W(i, j, k) w[(i) + (long)(pd1 + 2) * ((j) + (long)(pd2 + 2) * (k))]
  for (int k = 1; k <= pd3 + 1; k++)
    for (int j = 1; j <= pd2 + 1; j++)
      for (int i = 1; i <= pd1 + 1; i++) W(i, j, k) = corner(pc, mc3, i, j, k);
  for (int k = 1; k <= pd3; k++)
    for (int j = 1; j <= pd2; j++)
      for (int i = 1; i <= pd1; i++) {
        const double fbar = P(ps + i, ps + j, ps + k);
        double d0 = 0.5 * (W(i + 1, j + 1, k + 1) - W(i, j, k));
        double d1 = 0.5 * (W(i, j + 1, k + 1) - W(i + 1, j, k));
        double d2 = 0.5 * (W(i + 1, j, k + 1) - W(i, j + 1, k));
        double d3 = 0.5 * (W(i + 1, j + 1, k) - W(i, j, k + 1));
        if (positive == 1) {
          if (fbar + d0 < 0.0) d0 = -0.8 * fbar;
          if (fbar - d0 < 0.0) d0 = 0.8 * fbar;
          if (fbar + d1 < 0.0) d1 = -0.8 * fbar;
          if (fbar - d1 < 0.0) d1 = 0.8 * fbar;
          if (fbar + d2 < 0.0) d2 = -0.8 * fbar;
          if (fbar - d2 < 0.0) d2 = 0.8 * fbar;
          if (fbar + d3 < 0.0) d3 = -0.8 * fbar;
          if (fbar - d3 < 0.0) d3 = 0.8 * fbar;
        }
        if (d0 == 0.0) d0 = fsign(TINY, d0 + d1 + d2);
        const double s = (d1 + d2 + d3) / d0;
        const double sp = fmin2(fmax2(s, 0.0), 1.0);
        double c1 = 1.0, c2 = 1.0, c3 = 1.0;
        if (d1 / d0 >= 0.0) c1 = 0.0;
        if (d2 / d0 >= 0.0) c2 = 0.0;
        if (d3 / d0 >= 0.0) c3 = 0.0;
        c1 = (1.0 - sp) * c1 + sp * (1.0 - c1);
        c2 = (1.0 - sp) * c2 + sp * (1.0 - c2);
        c3 = (1.0 - sp) * c3 + sp * (1.0 - c3);
        if (sp != 0.0 && sp != 1.0) {
          c1 = 0.0;
          c2 = 0.0;
          c3 = 0.0;
        }
        double frac = -((-d0 * sp) + (1.0 - c1) * d1 + (1.0 - c2) * d2 + (1.0 - c3) * d3) /
                      (c1 * d1 + c2 * d2 + c3 * d3 + TINY);
        frac = fmin2(frac, 1.0);
        const double d1n = frac * c1 * d1 + (1.0 - c1) * d1;
        const double d2n = frac * c2 * d2 + (1.0 - c2) * d2;
        const double d3n = frac * c3 * d3 + (1.0 - c3) * d3;
        children(gf, mf3, i, j, k, fbar, d2n + d3n, d1n + d3n, d1n + d2n);
      }
#undef W
  free(w);
}

int o_enzo_prolong(int method, int positive, double *values_f, const int mf3[3],
                   const int of3[3], const int nf3[3], const double *values_c,
                   const int mc3[3], const int oc3[3], const int nc3[3]) {
  if (method < 0 || method > 2) return -1;
  for (int a = 0; a < 3; a++) {
    /* the coarse region holds the children's parents and one cell of
     * stencil on each side; the reference's own check is weaker
     * (interpolate.F: pend < (pdims-1)*refine) and goes on after it */
    if (nf3[a] < 2 || nf3[a] % 2 != 0 || nc3[a] != nf3[a] / 2 + 2) return -1;
    if (oc3[a] < 0 || oc3[a] + nc3[a] > mc3[a]) return -1;
    if (of3[a] < 0 || of3[a] + nf3[a] > mf3[a]) return -1;
  }
  const double *pc = values_c + oc3[0] + (long)mc3[0] * (oc3[1] + (long)mc3[1] * oc3[2]);
  double *gf = values_f + of3[0] + (long)mf3[0] * (of3[1] + (long)mf3[1] * of3[2]);
  if (method == 0) {
    tscint3d(pc, mc3, gf, mf3, nf3);
    return 0;
  }
  if (method == 1) return interp3d(pc, mc3, gf, mf3, nf3);
  int_lin3d(pc, mc3, gf, mf3, nf3, positive);
  return 0;
}
