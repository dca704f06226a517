/*
 * mesh_oracle.c -- CPU oracle for the coarse/fine transfer operators that
 * fill ghost zones across refinement levels.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of
 *   ProlongLinear::apply_ (src/Cello/problem_ProlongLinear.cpp:85-370) and
 *   RestrictLinear::apply_ (src/Cello/problem_RestrictLinear.cpp:60-198),
 * loop by loop, with the reference's weights and summation order.  Only
 * tests/, the smoke entry point and the CPU-baseline leg of bench.py load it.
 *
 * Pins (tests/test_oracle_mesh.py): linear prolongation reproduces constant
 * and linear fields exactly (interior weights 3/4, 1/4 and the one-sided
 * 5/4, -1/4 extrapolation where no coarse ghost cells are available), and
 * restriction returns the mean of the 2^rank children (closed forms on
 * dyadic data, so every sum is exact).
 */
#include "vlct_oracle.h"

#include <stdlib.h>

/* per-axis source index and weights of fine cell i_f (of n_f) */
static void prolong_axis(int i_f, int n_f, int gc, int *ic, double *w0,
                         double *w1) {
  int c = ((i_f + 1) >> 1) - gc;
  int w[2] = {1, 3};
  if (i_f == 0) c += gc;
  if (i_f == n_f - 1) c -= gc;
  if (i_f == 0 || i_f == n_f - 1) {
    w[0] += 4 * gc;
    w[1] -= 4 * gc;
  }
  *ic = c;
  *w0 = 0.25 * w[i_f & 1];
  *w1 = 0.25 * w[~i_f & 1];
}

int o_prolong_linear(double *values_f, const int mf3[3], const int of3[3],
                     const int nf3[3], const double *values_c,
                     const int mc3[3], const int oc3[3], const int nc3[3],
                     int accumulate) {
  const int dcx = 1, dcy = mc3[0], dcz = mc3[0] * mc3[1];
  const int rank = (mf3[1] == 1) ? 1 : ((mf3[2] == 1) ? 2 : 3);
  for (int i = 0; i < rank; i++)
    if (!(nf3[i] == 2 * nc3[i] || nf3[i] == 2 * (nc3[i] - 2))) return -1;
  const int gcx = (nf3[0] == 2 * nc3[0]) ? 1 : 0;
  const int gcy = (nf3[1] == 2 * nc3[1]) ? 1 : 0;
  const int gcz = (nf3[2] == 2 * nc3[2]) ? 1 : 0;
  const long mf = (long)mf3[0] * mf3[1] * mf3[2];
  double *temp_f = accumulate ? (double *)malloc((size_t)mf * sizeof(double))
                              : values_f;
  if (rank == 1) {
    for (int ifx = 0; ifx < nf3[0]; ifx++) {
      int icx;
      double wx0, wx1;
      prolong_axis(ifx, nf3[0], gcx, &icx, &wx0, &wx1);
      const long i_c = oc3[0] + icx, i_f = of3[0] + ifx;
      temp_f[i_f] = wx0 * values_c[i_c] + wx1 * values_c[i_c + dcx];
    }
  } else if (rank == 2) {
    for (int ify = 0; ify < nf3[1]; ify++) {
      int icy;
      double wy0, wy1;
      prolong_axis(ify, nf3[1], gcy, &icy, &wy0, &wy1);
      for (int ifx = 0; ifx < nf3[0]; ifx++) {
        int icx;
        double wx0, wx1;
        prolong_axis(ifx, nf3[0], gcx, &icx, &wx0, &wx1);
        const long i_c = (oc3[0] + icx) + (long)mc3[0] * (oc3[1] + icy);
        const long i_f = (of3[0] + ifx) + (long)mf3[0] * (of3[1] + ify);
        temp_f[i_f] = wx0 * wy0 * values_c[i_c] +
                      wx1 * wy0 * values_c[i_c + dcx] +
                      wx0 * wy1 * values_c[i_c + dcy] +
                      wx1 * wy1 * values_c[i_c + dcx + dcy];
      }
    }
  } else {
    for (int ifz = 0; ifz < nf3[2]; ifz++) {
      int icz;
      double wz0, wz1;
      prolong_axis(ifz, nf3[2], gcz, &icz, &wz0, &wz1);
      for (int ify = 0; ify < nf3[1]; ify++) {
        int icy;
        double wy0, wy1;
        prolong_axis(ify, nf3[1], gcy, &icy, &wy0, &wy1);
        for (int ifx = 0; ifx < nf3[0]; ifx++) {
          int icx;
          double wx0, wx1;
          prolong_axis(ifx, nf3[0], gcx, &icx, &wx0, &wx1);
          const long i_c = (oc3[0] + icx) +
                           (long)mc3[0] * ((oc3[1] + icy) + (long)mc3[1] * (oc3[2] + icz));
          const long i_f = (of3[0] + ifx) +
                           (long)mf3[0] * ((of3[1] + ify) + (long)mf3[1] * (of3[2] + ifz));
          temp_f[i_f] = wx0 * wy0 * wz0 * values_c[i_c] +
                        wx1 * wy0 * wz0 * values_c[i_c + dcx] +
                        wx0 * wy1 * wz0 * values_c[i_c + dcy] +
                        wx1 * wy1 * wz0 * values_c[i_c + dcx + dcy] +
                        wx0 * wy0 * wz1 * values_c[i_c + dcz] +
                        wx1 * wy0 * wz1 * values_c[i_c + dcx + dcz] +
                        wx0 * wy1 * wz1 * values_c[i_c + dcy + dcz] +
                        wx1 * wy1 * wz1 * values_c[i_c + dcx + dcy + dcz];
        }
      }
    }
  }
  if (accumulate) {
    const long i0 = of3[0] + (long)mf3[0] * (of3[1] + (long)mf3[1] * of3[2]);
    for (int iz = 0; iz < nf3[2]; iz++)
      for (int iy = 0; iy < nf3[1]; iy++)
        for (int ix = 0; ix < nf3[0]; ix++) {
          const long i = i0 + ix + (long)mf3[0] * (iy + (long)mf3[1] * iz);
          values_f[i] += temp_f[i];
        }
    free(temp_f);
  }
  return 0;
}

int o_restrict_linear(double *values_c, const int nd3_c[3], const int im3_c[3],
                      const int n3_c[3], const double *values_f,
                      const int nd3_f[3], const int im3_f[3],
                      const int n3_f[3], int accumulate) {
  (void)n3_f;
  const int rank = (nd3_f[1] == 1) ? 1 : ((nd3_f[2] == 1) ? 2 : 3);
  const long dx = 1, dy = nd3_f[0], dz = (long)nd3_f[0] * nd3_f[1];
  if (rank == 1) {
    for (int ix_c = 0; ix_c < n3_c[0]; ix_c++) {
      const long i_c = im3_c[0] + ix_c, i_f = im3_f[0] + 2 * ix_c;
      const double v = 0.5 * (values_f[i_f] + values_f[i_f + dx]);
      if (accumulate) values_c[i_c] += v; else values_c[i_c] = v;
    }
  } else if (rank == 2) {
    for (int ix_c = 0; ix_c < n3_c[0]; ix_c++)
      for (int iy_c = 0; iy_c < n3_c[1]; iy_c++) {
        const long i_c = (im3_c[0] + ix_c) + (long)nd3_c[0] * (im3_c[1] + iy_c);
        const long i_f = (im3_f[0] + 2 * ix_c) + (long)nd3_f[0] * (im3_f[1] + 2 * iy_c);
        const double v = 0.25 * (values_f[i_f] + values_f[i_f + dx] +
                                 values_f[i_f + dy] + values_f[i_f + dx + dy]);
        if (accumulate) values_c[i_c] += v; else values_c[i_c] = v;
      }
  } else {
    for (int ix_c = 0; ix_c < n3_c[0]; ix_c++)
      for (int iy_c = 0; iy_c < n3_c[1]; iy_c++)
        for (int iz_c = 0; iz_c < n3_c[2]; iz_c++) {
          const long i_c = (im3_c[0] + ix_c) +
                           (long)nd3_c[0] * ((im3_c[1] + iy_c) + (long)nd3_c[1] * (im3_c[2] + iz_c));
          const long i_f = (im3_f[0] + 2 * ix_c) +
                           (long)nd3_f[0] * ((im3_f[1] + 2 * iy_c) +
                                             (long)nd3_f[1] * (im3_f[2] + 2 * iz_c));
          const double v = 0.125 * (values_f[i_f] + values_f[i_f + dx] +
                                    values_f[i_f + dy] + values_f[i_f + dx + dy] +
                                    values_f[i_f + dz] + values_f[i_f + dx + dz] +
                                    values_f[i_f + dy + dz] +
                                    values_f[i_f + dx + dy + dz]);
          if (accumulate) values_c[i_c] += v; else values_c[i_c] = v;
        }
  }
  return 0;
}
