/*
 * fluxcorr_oracle.c -- CPU oracle for the flux corrections at refinement
 * boundaries that accompany the VL+CT step in hydro mode.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of
 *   EnzoMethodMHDVlct::save_fluxes_for_corrections_
 *       (src/Enzo/hydro-mhd/EnzoMethodMHDVlct.cpp:250-330),
 *   FluxData single-array storage (src/Cello/data_FluxData.cpp:14-110),
 *   FaceFluxes::coarsen (src/Cello/data_FaceFluxes.cpp:128-250),
 *   FluxData::sum_neighbor_fluxes (src/Cello/data_FluxData.hpp:136-145),
 *   MethodFluxCorrect::flux_correct_ with flux_correct_helper_ and
 *   divide_flux_corrected_cells_by_density_
 *       (src/Cello/problem_MethodFluxCorrect.cpp:275-560),
 * loop by loop, so that every sum is formed in the reference's order.  Only
 * tests/, the smoke entry point and the CPU-baseline leg of bench.py load it.
 *
 * Pins (tests/test_oracle_fluxcorr.py): closed forms of the coarsening
 * (constant and linear fine fluxes), the saved fluxes of a uniform flow
 * (the exact Euler flux), the density correction's closed form, and the
 * composite-mesh identity that the corrected coarse + fine blocks change
 * their conserved totals by exactly the fluxes through the outer faces.
 */
#include "vlct_oracle.h"

#include <stdlib.h>
#include <string.h>

/* face size of axis for an (nx, ny, nz) block: FaceFluxes sets the extent
 * normal to the face to 1 (data_FaceFluxes.cpp:21-46) */
static void face_dims(int axis, int nx, int ny, int nz, int *fx, int *fy,
                      int *fz) {
  *fx = (axis == 0) ? 1 : nx;
  *fy = (axis == 1) ? 1 : ny;
  *fz = (axis == 2) ? 1 : nz;
}

static long face_size(int axis, int nx, int ny, int nz) {
  int fx, fy, fz;
  face_dims(axis, nx, ny, nz, &fx, &fy, &fz);
  return (long)fx * fy * fz;
}

/* storage order of FluxData::allocate with single_array = true
 * (data_FluxData.cpp:57-87): for i_f, axis, face: block then neighbour */
long o_flux_data_offset(int i_f, int axis, int face, int neighbor, int nx,
                        int ny, int nz) {
  long per_field = 0;
  for (int a = 0; a < 3; a++) per_field += 4 * face_size(a, nx, ny, nz);
  long off = (long)i_f * per_field;
  for (int a = 0; a < axis; a++) off += 4 * face_size(a, nx, ny, nz);
  const long m = face_size(axis, nx, ny, nz);
  off += (long)face * 2 * m;
  if (neighbor) off += m;
  return off;
}

long o_flux_data_size(int nf, int nx, int ny, int nz) {
  return o_flux_data_offset(nf, 0, 0, 0, nx, ny, nz);
}

void o_save_fluxes_for_corrections(double *const *flux, int dim,
                                   double cell_width, double dt, int mx,
                                   int my, int mz, int g, int nf,
                                   const int *fields, double *flux_data) {
  const int nx = mx - 2 * g, ny = my - 2 * g, nz = mz - 2 * g;
  /* dense face-trimmed flux arrays */
  const int fx = mx - (dim == 0), fy = my - (dim == 1);
  const double dt_dxi = dt / cell_width;
  for (int i_f = 0; i_f < nf; i_f++) {
    const double *F = flux[fields[i_f]];
    double *left = flux_data + o_flux_data_offset(i_f, dim, 0, 0, nx, ny, nz);
    double *right = flux_data + o_flux_data_offset(i_f, dim, 1, 0, nx, ny, nz);
    int ax, ay, az;
    face_dims(dim, nx, ny, nz, &ax, &ay, &az);
    /* FaceFluxes::flux_array strides (data_FaceFluxes.cpp:114-126) */
    const int dx = 1, dy = ax, dz = ax * ay;
    if (dim == 0) {
      const int left_ix = g - 1, right_ix = mx - g - 1;
      for (int iz = 0; iz < az; iz++)
        for (int iy = 0; iy < ay; iy++) {
          left[dz * iz + dy * iy] =
              dt_dxi * F[((long)(g + iz) * fy + (g + iy)) * fx + left_ix];
          right[dz * iz + dy * iy] =
              dt_dxi * F[((long)(g + iz) * fy + (g + iy)) * fx + right_ix];
        }
    } else if (dim == 1) {
      const int left_iy = g - 1, right_iy = my - g - 1;
      for (int iz = 0; iz < az; iz++)
        for (int ix = 0; ix < ax; ix++) {
          left[dz * iz + dx * ix] =
              dt_dxi * F[((long)(g + iz) * fy + left_iy) * fx + (g + ix)];
          right[dz * iz + dx * ix] =
              dt_dxi * F[((long)(g + iz) * fy + right_iy) * fx + (g + ix)];
        }
    } else {
      const int left_iz = g - 1, right_iz = mz - g - 1;
      for (int iy = 0; iy < ay; iy++)
        for (int ix = 0; ix < ax; ix++) {
          left[dy * iy + dx * ix] =
              dt_dxi * F[((long)left_iz * fy + (g + iy)) * fx + (g + ix)];
          right[dy * iy + dx * ix] =
              dt_dxi * F[((long)right_iz * fy + (g + iy)) * fx + (g + ix)];
        }
    }
  }
}

void o_face_fluxes_coarsen(const double *fine, double *out, int axis, int nx,
                           int ny, int nz, int cx, int cy, int cz) {
  int mxf, myf, mzf;
  face_dims(axis, nx, ny, nz, &mxf, &myf, &mzf);
  const long m = (long)mxf * myf * mzf;
  for (long i = 0; i < m; i++) out[i] = 0.0;
  /* nx_ etc. are the face dimensions; centering 0 */
  const int nx_ = mxf, ny_ = myf, nz_ = mzf;
  const int mxc = (nx_ > 1) ? nx_ / 2 : 1;
  const int myc = (ny_ > 1) ? ny_ / 2 : 1;
  const int mzc = (nz_ > 1) ? nz_ / 2 : 1;
  const int dxc = 1, dyc = nx_, dzc = nx_ * ny_;
  const int ixc0 = cx * mxc * dxc, iyc0 = cy * myc * dyc, izc0 = cz * mzc * dzc;
  const double dArea = 0.125;
  if (mxf == 1) {
    const int ic0 = iyc0 + izc0;
    for (int izf = 0; izf < mzf; izf++) {
      const int izcm = dzc * (izf >> 1), izcp = dzc * (izf >> 1);
      for (int iyf = 0; iyf < myf; iyf++) {
        const int iycm = dyc * (iyf >> 1), iycp = dyc * (iyf >> 1);
        const int i_f = mxf * (iyf + myf * izf);
        out[ic0 + izcm + iycm] += 0.25 * dArea * fine[i_f];
        out[ic0 + izcm + iycp] += 0.25 * dArea * fine[i_f];
        out[ic0 + izcp + iycm] += 0.25 * dArea * fine[i_f];
        out[ic0 + izcp + iycp] += 0.25 * dArea * fine[i_f];
      }
    }
  } else if (myf == 1) {
    const int ic0 = ixc0 + izc0;
    for (int izf = 0; izf < mzf; izf++) {
      const int izcm = dzc * (izf >> 1), izcp = dzc * (izf >> 1);
      for (int ixf = 0; ixf < mxf; ixf++) {
        const int ixcm = dxc * (ixf >> 1), ixcp = dxc * (ixf >> 1);
        const int i_f = ixf + mxf * (myf * izf);
        out[ic0 + izcm + ixcm] += 0.25 * dArea * fine[i_f];
        out[ic0 + izcm + ixcp] += 0.25 * dArea * fine[i_f];
        out[ic0 + izcp + ixcm] += 0.25 * dArea * fine[i_f];
        out[ic0 + izcp + ixcp] += 0.25 * dArea * fine[i_f];
      }
    }
  } else {
    const int ic0 = ixc0 + iyc0;
    for (int iyf = 0; iyf < myf; iyf++) {
      const int iycm = dyc * (iyf >> 1), iycp = dyc * (iyf >> 1);
      for (int ixf = 0; ixf < mxf; ixf++) {
        const int ixcm = dxc * (ixf >> 1), ixcp = dxc * (ixf >> 1);
        const int i_f = ixf + mxf * iyf;
        out[ic0 + iycm + ixcm] += 0.25 * dArea * fine[i_f];
        out[ic0 + iycm + ixcp] += 0.25 * dArea * fine[i_f];
        out[ic0 + iycp + ixcm] += 0.25 * dArea * fine[i_f];
        out[ic0 + iycp + ixcp] += 0.25 * dArea * fine[i_f];
      }
    }
  }
}

void o_flux_send_to_coarse(const double *fine_fd, double *coarse_fd, int nf,
                           int nx, int ny, int nz, int axis, int face,
                           const int child[3]) {
  const long m = face_size(axis, nx, ny, nz);
  double *tmp = (double *)malloc((size_t)m * sizeof(double));
  for (int i_f = 0; i_f < nf; i_f++) {
    const double *src =
        fine_fd + o_flux_data_offset(i_f, axis, face, 0, nx, ny, nz);
    o_face_fluxes_coarsen(src, tmp, axis, nx, ny, nz, child[0], child[1],
                          child[2]);
    /* the message lands on the neighbour's opposite face
     * (data_DataMsg.cpp:383-396) */
    double *dst =
        coarse_fd + o_flux_data_offset(i_f, axis, 1 - face, 1, nx, ny, nz);
    for (long i = 0; i < m; i++) dst[i] += tmp[i];
  }
  free(tmp);
}

/* flux_correct_helper_ (problem_MethodFluxCorrect.cpp:275-370), rank 3;
 * a points at the first active cell */
static void flux_correct_helper(double *a, int mx, int my, int nx, int ny,
                                int nz, int i_f, const double *fd,
                                const int correct[6]) {
  for (int axis = 0; axis < 3; axis++) {
    int fx, fy, fz;
    face_dims(axis, nx, ny, nz, &fx, &fy, &fz);
    const int dbx = 1, dby = fx, dbz = fx * fy;
    for (int face = 0; face < 2; face++) {
      if (!correct[2 * axis + face]) continue;
      const double *bf = fd + o_flux_data_offset(i_f, axis, face, 0, nx, ny, nz);
      const double *nb = fd + o_flux_data_offset(i_f, axis, face, 1, nx, ny, nz);
      if (axis == 0) {
        const int ix = (face == 0) ? 0 : nx - 1;
        for (int iz = 0; iz < nz; iz++)
          for (int iy = 0; iy < ny; iy++) {
            long i = ix + (long)mx * (iy + (long)my * iz);
            int ib = iy * dby + iz * dbz;
            a[i] += (2 * face - 1) * (bf[ib] - nb[ib]);
          }
      } else if (axis == 1) {
        const int iy = (face == 0) ? 0 : ny - 1;
        for (int iz = 0; iz < nz; iz++)
          for (int ix = 0; ix < nx; ix++) {
            long i = ix + (long)mx * (iy + (long)my * iz);
            int ib = ix * dbx + iz * dbz;
            a[i] += (2 * face - 1) * (bf[ib] - nb[ib]);
          }
      } else {
        const int iz = (face == 0) ? 0 : nz - 1;
        for (int iy = 0; iy < ny; iy++)
          for (int ix = 0; ix < nx; ix++) {
            long i = ix + (long)mx * (iy + (long)my * iz);
            int ib = ix * dbx + iy * dby;
            a[i] += (2 * face - 1) * (bf[ib] - nb[ib]);
          }
      }
    }
  }
}

/* divide_flux_corrected_cells_by_density_
 * (problem_MethodFluxCorrect.cpp:372-423) */
static void divide_by_density(const double *in, const double *rho, double *out,
                              int mx, int my, int nx, int ny, int nz,
                              const int correct[6]) {
  for (int axis = 0; axis < 3; axis++)
    for (int face = 0; face < 2; face++) {
      if (!correct[2 * axis + face]) continue;
      if (axis == 0) {
        const int ix = (face == 0) ? 0 : nx - 1;
        for (int iz = 0; iz < nz; iz++)
          for (int iy = 0; iy < ny; iy++) {
            long i = ix + (long)mx * (iy + (long)my * iz);
            out[i] = in[i] / rho[i];
          }
      } else if (axis == 1) {
        const int iy = (face == 0) ? 0 : ny - 1;
        for (int iz = 0; iz < nz; iz++)
          for (int ix = 0; ix < nx; ix++) {
            long i = ix + (long)mx * (iy + (long)my * iz);
            out[i] = in[i] / rho[i];
          }
      } else {
        const int iz = (face == 0) ? 0 : nz - 1;
        for (int iy = 0; iy < ny; iy++)
          for (int ix = 0; ix < nx; ix++) {
            long i = ix + (long)mx * (iy + (long)my * iz);
            out[i] = in[i] / rho[i];
          }
      }
    }
}

int o_flux_correct(double *const *fields, int nf, int i_density,
                   const int *conservative, const double *flux_data, int nx,
                   int ny, int nz, int g, const int correct[6]) {
  if (nf == 0) return 0;
  const int mx = nx + 2 * g, my = ny + 2 * g, mz = nz + 2 * g;
  const long n = (long)mx * my * mz;
  const long off = g + (long)mx * (g + (long)my * g);
  for (int i_f = 0; i_f < nf; i_f++)
    if (conservative[i_f] && i_f != i_density && i_density < 0) return -1;
  double *scratch = (double *)calloc((size_t)(2 * n), sizeof(double));
  double *old_rho = scratch + off, *temp = scratch + n + off;
  double *rho = (i_density >= 0) ? fields[i_density] + off : NULL;
  if (rho) {
    for (int iz = 0; iz < nz; iz++)
      for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
          long i = ix + (long)mx * (iy + (long)my * iz);
          old_rho[i] = rho[i];
        }
    flux_correct_helper(rho, mx, my, nx, ny, nz, i_density, flux_data,
                        correct);
  }
  for (int i_f = 0; i_f < nf; i_f++) {
    if (i_f == i_density) continue;
    double *a = fields[i_f] + off;
    if (conservative[i_f]) {
      for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
          for (int ix = 0; ix < nx; ix++) {
            long i = ix + (long)mx * (iy + (long)my * iz);
            temp[i] = old_rho[i] * a[i];
          }
      flux_correct_helper(temp, mx, my, nx, ny, nz, i_f, flux_data, correct);
      divide_by_density(temp, rho, a, mx, my, nx, ny, nz, correct);
    } else {
      flux_correct_helper(a, mx, my, nx, ny, nz, i_f, flux_data, correct);
    }
  }
  free(scratch);
  return 0;
}
