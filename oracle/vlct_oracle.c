/*
 * vlct_oracle.c -- CPU oracle (test infrastructure, not product code) for the
 * Enzo-E VL+CT MHD step.  Plain C99, built with -ffp-contract=off so that no
 * multiply-add is fused and every expression rounds exactly as written.
 *
 * Each function restates one reference routine; the comment above it cites
 * the reference file:line it follows.  Expressions keep the reference's
 * parenthesisation and C's left-to-right association, because the HIP
 * library is checked against this file cell by cell.
 *
 * Region conventions (from the reference's stale-depth bookkeeping,
 * src/Enzo/hydro-mhd/toolkit/EnzoReconstructor.hpp:66-94):
 *   - a cell-centred array has extent m along each axis;
 *   - a face array "along dim" has extent m-1 along dim, index f is the face
 *     between cells f and f+1 (EnzoRiemannImpl.hpp:71-82);
 *   - bfieldi arrays include the exterior faces: extent m+1 along their axis,
 *     index X is the face between cells X-1 and X;
 *   - "stale depth s" trims s entries from both ends of every axis
 *     (src/Cello/view_ViewMap.hpp:310-327).
 */
#include "vlct_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* element-wise loops: threads split the outer (z) loop; results do not
 * depend on the thread count. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static) if((long)mx * my * mz > 300000L)")
#else
#define OMP_FOR
#endif

#define IDX(k, j, i, ny, nx) \
  ((((long)(k)) * (long)(ny) + (long)(j)) * (long)(nx) + (long)(i))

/* enzo_utils::min (src/Enzo/utils/utils.hpp:82-89) */
static inline double min3(double a, double b, double c) {
  if (a < b) {
    return (c < a) ? c : a;
  } else {
    return (c < b) ? c : b;
  }
}
/* std::min / std::max semantics */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }
/* enzo_utils::apply_floor (utils.hpp:102-114) = std::max(value, floor) */
static inline double apply_floor(double v, double f) { return smax(v, f); }
/* sign() of EnzoReconstructorPLM.hpp:253 */
static inline double sgn(double v) {
  return (double)((0.0 < v) - (v < 0.0));
}

/* ------------------------------------------------------------------------ */
/* Limiters                                                                 */
/* ------------------------------------------------------------------------ */

/* PLM_EnzoRKLimiter::operator() (EnzoReconstructorPLM.hpp:287-302) */
double o_limiter_enzo(double vm1, double v, double vp1, double theta) {
  double dv_c = 0.5 * (vp1 - vm1);
  double dv_l = (v - vm1) * theta;
  double dv_r = (vp1 - v) * theta;
  return (0.5 * (sgn(dv_l) + sgn(dv_r))) *
         min3(fabs(dv_l), fabs(dv_r), fabs(dv_c));
}

/* PLM_AthenaLimiter::operator() (EnzoReconstructorPLM.hpp:337-345) */
double o_limiter_athena(double vm1, double v, double vp1) {
  double dv_l = (v - vm1);
  double dv_r = (vp1 - v);
  double temp = dv_l * dv_r;
  if (temp <= 0.) return 0.;
  return 2. * temp / (dv_l + dv_r);
}

/* ------------------------------------------------------------------------ */
/* Reconstruction                                                           */
/* ------------------------------------------------------------------------ */

static void face_shape(int mz, int my, int mx, int dim, int *fz, int *fy,
                       int *fx) {
  *fz = mz - (dim == 2);
  *fy = my - (dim == 1);
  *fx = mx - (dim == 0);
}

/* kind O_NN: EnzoReconstructorNN::reconstruct_interface
 *   (EnzoReconstructorNN.cpp:14-49): faces f in [s, m_d-s-1) along dim,
 *   transverse [s, m-s): wl(f) = w(f), wr(f) = w(f+1).
 * kind O_PLM / O_PLM_ATHENA: EnzoReconstructorPLM<Limiter>::
 *   reconstruct_interface (EnzoReconstructorPLM.hpp:166-248): centre cells c
 *   in [s+1, m_d-s-1) along dim, transverse [s, m-s):
 *   wr(c-1) = floor(w_c - dv/2), wl(c) = floor(w_c + dv/2).            */
void o_reconstruct(int kind, double theta, const double *w, double *wl,
                   double *wr, int mz, int my, int mx, int dim, int s,
                   int use_floor, double floor_val) {
  int fz, fy, fx;
  face_shape(mz, my, mx, dim, &fz, &fy, &fx);
  const int di = (dim == 0), dj = (dim == 1), dk = (dim == 2);
  const long cstr = (dim == 0) ? 1 : ((dim == 1) ? (long)mx : (long)mx * my);
  const int m_d = (dim == 0) ? mx : ((dim == 1) ? my : mz);
  (void)di; (void)dj; (void)dk;

  if (kind == O_NN) {
    OMP_FOR
    for (int k = s; k < mz - s - dk; k++)
      for (int j = s; j < my - s - dj; j++)
        for (int i = s; i < mx - s - di; i++) {
          long c = IDX(k, j, i, my, mx);
          long f = IDX(k, j, i, fy, fx);
          wl[f] = w[c];
          wr[f] = w[c + cstr];
        }
    return;
  }
  (void)m_d;
  OMP_FOR
  for (int k = s + dk; k < mz - s - dk; k++)
    for (int j = s + dj; j < my - s - dj; j++)
      for (int i = s + di; i < mx - s - di; i++) {
        long c = IDX(k, j, i, my, mx);
        double val = w[c];
        double dv = (kind == O_PLM)
                        ? o_limiter_enzo(w[c - cstr], val, w[c + cstr], theta)
                        : o_limiter_athena(w[c - cstr], val, w[c + cstr]);
        double half_dv = dv * 0.5;
        double right_val, left_val;
        if (use_floor) {
          right_val = apply_floor(val - half_dv, floor_val);
          left_val = apply_floor(val + half_dv, floor_val);
        } else {
          right_val = val - half_dv;
          left_val = val + half_dv;
        }
        /* face index of the face on the left of cell c is c-1 */
        long f_right_of_c = IDX(k, j, i, fy, fx);
        long f_left_of_c = IDX(k - dk, j - dj, i - di, fy, fx);
        wr[f_left_of_c] = right_val;
        wl[f_right_of_c] = left_val;
      }
}

/* ------------------------------------------------------------------------ */
/* Riemann solvers (one face, LUT order)                                    */
/* ------------------------------------------------------------------------ */

enum { L_D = 0, L_BI, L_BJ, L_BK, L_VI, L_VJ, L_VK, L_E };
enum { H_D = 0, H_VI, H_VJ, H_VK, H_E };

/* EnzoEOSIdeal::fast_magnetosonic_speed<-1> (EnzoEOSIdeal.hpp:111-140);
 * std::pow(x,2) is exact and equal to x*x. */
static inline double fast_ms_speed(double d, double p, double bi, double bj,
                                   double bk, double gamma) {
  const double B2 = ((bi * bi) + ((bj * bj) + (bk * bk)));
  const double cs2 = gamma * p / d;
  const double inv_density = 1.0 / d;
  const double va2 = B2 * inv_density;
  const double va2_cos2 = (bi * bi) * inv_density;
  const double sum = cs2 + va2;
  return sqrt(0.5 * (va2 + cs2 + sqrt(sum * sum - 4. * cs2 * va2_cos2)));
}

typedef struct { double d, mx, my, mz, e, by, bz; } cons1d;

/* enzo_riemann_utils::passive_eint_flux (EnzoRiemannUtils.hpp:224-249):
 * the specific internal energy eint = p / ((gamma - 1) rho) of the side the
 * density flux comes from (branch-free selection as the reference writes
 * it), times the density flux. */
static double passive_eint_flux(double density_l, double pressure_l,
                                double density_r, double pressure_r,
                                double gamma, double density_flux) {
  double eint_l = pressure_l / ((gamma - 1.0) * density_l);
  double eint_r = pressure_r / ((gamma - 1.0) * density_r);
  double upwind = (double)(density_flux > 0) * eint_l +
                  (double)(density_flux <= 0) * eint_r;
  return upwind * density_flux;
}

/* HLLDKernel::operator() (EnzoRiemannHLLD.hpp:40-448); with de, the dual
 * energy outputs of :404-447: the internal-energy flux and vi_bar from the
 * outer and contact speeds. */
static void hlld_impl(const double wli[8], const double wri[8], double gamma,
                      double flxo[8], double *de) {
  const double SMALL_NUMBER = 1.0e-8;
  const double igm1 = 1.0 / (gamma - 1.0);
  double spd[5];
  cons1d ul, ur, ulst, uldst, urdst, urst, fl, fr;
  double flxi[8];

  const double pressure_l = wli[L_E];
  const double pressure_r = wri[L_E];
  const double bxi = wli[L_BI];

  double bxsq = bxi * bxi;
  double pbl = 0.5 * (bxsq + (wli[L_BJ] * wli[L_BJ] + wli[L_BK] * wli[L_BK]));
  double pbr = 0.5 * (bxsq + (wri[L_BJ] * wri[L_BJ] + wri[L_BK] * wri[L_BK]));
  double kel = 0.5 * wli[L_D] *
               (wli[L_VI] * wli[L_VI] +
                (wli[L_VJ] * wli[L_VJ] + wli[L_VK] * wli[L_VK]));
  double ker = 0.5 * wri[L_D] *
               (wri[L_VI] * wri[L_VI] +
                (wri[L_VJ] * wri[L_VJ] + wri[L_VK] * wri[L_VK]));

  ul.d = wli[L_D];
  ul.mx = wli[L_VI] * ul.d;
  ul.my = wli[L_VJ] * ul.d;
  ul.mz = wli[L_VK] * ul.d;
  ul.e = pressure_l * igm1 + kel + pbl;
  ul.by = wli[L_BJ];
  ul.bz = wli[L_BK];

  ur.d = wri[L_D];
  ur.mx = wri[L_VI] * ur.d;
  ur.my = wri[L_VJ] * ur.d;
  ur.mz = wri[L_VK] * ur.d;
  ur.e = pressure_r * igm1 + ker + pbr;
  ur.by = wri[L_BJ];
  ur.bz = wri[L_BK];

  double cfl = fast_ms_speed(wli[L_D], pressure_l, wli[L_BI], wli[L_BJ],
                             wli[L_BK], gamma);
  double cfr = fast_ms_speed(wri[L_D], pressure_r, wri[L_BI], wri[L_BJ],
                             wri[L_BK], gamma);

  spd[0] = smin(wli[L_VI] - cfl, wri[L_VI] - cfr);
  spd[4] = smax(wli[L_VI] + cfl, wri[L_VI] + cfr);

  double ptl = pressure_l + pbl;
  double ptr = pressure_r + pbr;

  fl.d = ul.mx;
  fl.mx = ul.mx * wli[L_VI] + ptl - bxsq;
  fl.my = ul.my * wli[L_VI] - bxi * ul.by;
  fl.mz = ul.mz * wli[L_VI] - bxi * ul.bz;
  fl.e = wli[L_VI] * (ul.e + ptl - bxsq) -
         bxi * (wli[L_VJ] * ul.by + wli[L_VK] * ul.bz);
  fl.by = ul.by * wli[L_VI] - bxi * wli[L_VJ];
  fl.bz = ul.bz * wli[L_VI] - bxi * wli[L_VK];

  fr.d = ur.mx;
  fr.mx = ur.mx * wri[L_VI] + ptr - bxsq;
  fr.my = ur.my * wri[L_VI] - bxi * ur.by;
  fr.mz = ur.mz * wri[L_VI] - bxi * ur.bz;
  fr.e = wri[L_VI] * (ur.e + ptr - bxsq) -
         bxi * (wri[L_VJ] * ur.by + wri[L_VK] * ur.bz);
  fr.by = ur.by * wri[L_VI] - bxi * wri[L_VJ];
  fr.bz = ur.bz * wri[L_VI] - bxi * wri[L_VK];

  double sdl = spd[0] - wli[L_VI];
  double sdr = spd[4] - wri[L_VI];

  spd[2] = (sdr * ur.mx - sdl * ul.mx + (ptl - ptr)) /
           (sdr * ur.d - sdl * ul.d);

  double sdml = spd[0] - spd[2];
  double sdmr = spd[4] - spd[2];
  double sdml_inv = 1.0 / sdml;
  double sdmr_inv = 1.0 / sdmr;
  ulst.d = ul.d * sdl * sdml_inv;
  urst.d = ur.d * sdr * sdmr_inv;
  double ulst_d_inv = 1.0 / ulst.d;
  double urst_d_inv = 1.0 / urst.d;
  double sqrtdl = sqrt(ulst.d);
  double sqrtdr = sqrt(urst.d);

  spd[1] = spd[2] - fabs(bxi) / sqrtdl;
  spd[3] = spd[2] + fabs(bxi) / sqrtdr;

  double ptstl = ptl + ul.d * sdl * (spd[2] - wli[L_VI]);
  double ptstr = ptr + ur.d * sdr * (spd[2] - wri[L_VI]);
  double ptst = 0.5 * (ptstr + ptstl);

  ulst.mx = ulst.d * spd[2];
  if (fabs(ul.d * sdl * sdml - bxsq) < (SMALL_NUMBER)*ptst) {
    ulst.my = ulst.d * wli[L_VJ];
    ulst.mz = ulst.d * wli[L_VK];
    ulst.by = ul.by;
    ulst.bz = ul.bz;
  } else {
    double tmp = bxi * (sdl - sdml) / (ul.d * sdl * sdml - bxsq);
    ulst.my = ulst.d * (wli[L_VJ] - ul.by * tmp);
    ulst.mz = ulst.d * (wli[L_VK] - ul.bz * tmp);
    tmp = (ul.d * (sdl * sdl) - bxsq) / (ul.d * sdl * sdml - bxsq);
    ulst.by = ul.by * tmp;
    ulst.bz = ul.bz * tmp;
  }
  double vbstl = (ulst.mx * bxi + (ulst.my * ulst.by + ulst.mz * ulst.bz)) *
                 ulst_d_inv;
  ulst.e = (sdl * ul.e - ptl * wli[L_VI] + ptst * spd[2] +
            bxi * (wli[L_VI] * bxi +
                   (wli[L_VJ] * ul.by + wli[L_VK] * ul.bz) - vbstl)) *
           sdml_inv;

  urst.mx = urst.d * spd[2];
  if (fabs(ur.d * sdr * sdmr - bxsq) < (SMALL_NUMBER)*ptst) {
    urst.my = urst.d * wri[L_VJ];
    urst.mz = urst.d * wri[L_VK];
    urst.by = ur.by;
    urst.bz = ur.bz;
  } else {
    double tmp = bxi * (sdr - sdmr) / (ur.d * sdr * sdmr - bxsq);
    urst.my = urst.d * (wri[L_VJ] - ur.by * tmp);
    urst.mz = urst.d * (wri[L_VK] - ur.bz * tmp);
    tmp = (ur.d * (sdr * sdr) - bxsq) / (ur.d * sdr * sdmr - bxsq);
    urst.by = ur.by * tmp;
    urst.bz = ur.bz * tmp;
  }
  double vbstr = (urst.mx * bxi + (urst.my * urst.by + urst.mz * urst.bz)) *
                 urst_d_inv;
  urst.e = (sdr * ur.e - ptr * wri[L_VI] + ptst * spd[2] +
            bxi * (wri[L_VI] * bxi +
                   (wri[L_VJ] * ur.by + wri[L_VK] * ur.bz) - vbstr)) *
           sdmr_inv;

  if (0.5 * bxsq < (SMALL_NUMBER)*ptst) {
    uldst = ulst;
    urdst = urst;
  } else {
    double invsumd = 1.0 / (sqrtdl + sqrtdr);
    double bxsig = (bxi > 0.0 ? 1.0 : -1.0);

    uldst.d = ulst.d;
    urdst.d = urst.d;
    uldst.mx = ulst.mx;
    urdst.mx = urst.mx;

    double tmp = invsumd * (sqrtdl * (ulst.my * ulst_d_inv) +
                            sqrtdr * (urst.my * urst_d_inv) +
                            bxsig * (urst.by - ulst.by));
    uldst.my = uldst.d * tmp;
    urdst.my = urdst.d * tmp;

    tmp = invsumd * (sqrtdl * (ulst.mz * ulst_d_inv) +
                     sqrtdr * (urst.mz * urst_d_inv) +
                     bxsig * (urst.bz - ulst.bz));
    uldst.mz = uldst.d * tmp;
    urdst.mz = urdst.d * tmp;

    tmp = invsumd * (sqrtdl * urst.by + sqrtdr * ulst.by +
                     bxsig * sqrtdl * sqrtdr *
                         ((urst.my * urst_d_inv) - (ulst.my * ulst_d_inv)));
    uldst.by = urdst.by = tmp;

    tmp = invsumd * (sqrtdl * urst.bz + sqrtdr * ulst.bz +
                     bxsig * sqrtdl * sqrtdr *
                         ((urst.mz * urst_d_inv) - (ulst.mz * ulst_d_inv)));
    uldst.bz = urdst.bz = tmp;

    tmp = spd[2] * bxi + (uldst.my * uldst.by + uldst.mz * uldst.bz) / uldst.d;
    uldst.e = ulst.e - sqrtdl * bxsig * (vbstl - tmp);
    urdst.e = urst.e + sqrtdr * bxsig * (vbstr - tmp);
  }

  uldst.d = spd[1] * (uldst.d - ulst.d);
  uldst.mx = spd[1] * (uldst.mx - ulst.mx);
  uldst.my = spd[1] * (uldst.my - ulst.my);
  uldst.mz = spd[1] * (uldst.mz - ulst.mz);
  uldst.e = spd[1] * (uldst.e - ulst.e);
  uldst.by = spd[1] * (uldst.by - ulst.by);
  uldst.bz = spd[1] * (uldst.bz - ulst.bz);

  ulst.d = spd[0] * (ulst.d - ul.d);
  ulst.mx = spd[0] * (ulst.mx - ul.mx);
  ulst.my = spd[0] * (ulst.my - ul.my);
  ulst.mz = spd[0] * (ulst.mz - ul.mz);
  ulst.e = spd[0] * (ulst.e - ul.e);
  ulst.by = spd[0] * (ulst.by - ul.by);
  ulst.bz = spd[0] * (ulst.bz - ul.bz);

  urdst.d = spd[3] * (urdst.d - urst.d);
  urdst.mx = spd[3] * (urdst.mx - urst.mx);
  urdst.my = spd[3] * (urdst.my - urst.my);
  urdst.mz = spd[3] * (urdst.mz - urst.mz);
  urdst.e = spd[3] * (urdst.e - urst.e);
  urdst.by = spd[3] * (urdst.by - urst.by);
  urdst.bz = spd[3] * (urdst.bz - urst.bz);

  urst.d = spd[4] * (urst.d - ur.d);
  urst.mx = spd[4] * (urst.mx - ur.mx);
  urst.my = spd[4] * (urst.my - ur.my);
  urst.mz = spd[4] * (urst.mz - ur.mz);
  urst.e = spd[4] * (urst.e - ur.e);
  urst.by = spd[4] * (urst.by - ur.by);
  urst.bz = spd[4] * (urst.bz - ur.bz);

  if (spd[0] >= 0.0) {
    flxi[L_D] = fl.d; flxi[L_VI] = fl.mx; flxi[L_VJ] = fl.my;
    flxi[L_VK] = fl.mz; flxi[L_E] = fl.e; flxi[L_BJ] = fl.by;
    flxi[L_BK] = fl.bz;
  } else if (spd[4] <= 0.0) {
    flxi[L_D] = fr.d; flxi[L_VI] = fr.mx; flxi[L_VJ] = fr.my;
    flxi[L_VK] = fr.mz; flxi[L_E] = fr.e; flxi[L_BJ] = fr.by;
    flxi[L_BK] = fr.bz;
  } else if (spd[1] >= 0.0) {
    flxi[L_D] = fl.d + ulst.d;
    flxi[L_VI] = fl.mx + ulst.mx;
    flxi[L_VJ] = fl.my + ulst.my;
    flxi[L_VK] = fl.mz + ulst.mz;
    flxi[L_E] = fl.e + ulst.e;
    flxi[L_BJ] = fl.by + ulst.by;
    flxi[L_BK] = fl.bz + ulst.bz;
  } else if (spd[2] >= 0.0) {
    flxi[L_D] = fl.d + ulst.d + uldst.d;
    flxi[L_VI] = fl.mx + ulst.mx + uldst.mx;
    flxi[L_VJ] = fl.my + ulst.my + uldst.my;
    flxi[L_VK] = fl.mz + ulst.mz + uldst.mz;
    flxi[L_E] = fl.e + ulst.e + uldst.e;
    flxi[L_BJ] = fl.by + ulst.by + uldst.by;
    flxi[L_BK] = fl.bz + ulst.bz + uldst.bz;
  } else if (spd[3] > 0.0) {
    flxi[L_D] = fr.d + urst.d + urdst.d;
    flxi[L_VI] = fr.mx + urst.mx + urdst.mx;
    flxi[L_VJ] = fr.my + urst.my + urdst.my;
    flxi[L_VK] = fr.mz + urst.mz + urdst.mz;
    flxi[L_E] = fr.e + urst.e + urdst.e;
    flxi[L_BJ] = fr.by + urst.by + urdst.by;
    flxi[L_BK] = fr.bz + urst.bz + urdst.bz;
  } else {
    flxi[L_D] = fr.d + urst.d;
    flxi[L_VI] = fr.mx + urst.mx;
    flxi[L_VJ] = fr.my + urst.my;
    flxi[L_VK] = fr.mz + urst.mz;
    flxi[L_E] = fr.e + urst.e;
    flxi[L_BJ] = fr.by + urst.by;
    flxi[L_BK] = fr.bz + urst.bz;
  }
  flxo[L_D] = flxi[L_D];
  flxo[L_BI] = 0.0;
  flxo[L_BJ] = flxi[L_BJ];
  flxo[L_BK] = flxi[L_BK];
  flxo[L_VI] = flxi[L_VI];
  flxo[L_VJ] = flxi[L_VJ];
  flxo[L_VK] = flxi[L_VK];
  flxo[L_E] = flxi[L_E];
  if (de) {
    de[0] = passive_eint_flux(wli[L_D], pressure_l, wri[L_D], pressure_r,
                              gamma, flxo[L_D]);
    const double S_M = spd[2];
    const double S_l = spd[0];
    const double S_r = spd[4];
    const double l_coef = (S_l - wli[L_VI]) / (S_l - S_M);
    const double r_coef = (S_r - wri[L_VI]) / (S_r - S_M);
    if (S_l > 0) {
      de[1] = wli[L_VI];
    } else if (S_r < 0) {
      de[1] = wri[L_VI];
    } else if (S_M >= 0) {
      de[1] = S_M * l_coef;
    } else {
      de[1] = S_M * r_coef;
    }
  }
}

void o_hlld_face(const double wli[8], const double wri[8], double gamma,
                 double flxo[8]) {
  hlld_impl(wli, wri, gamma, flxo, NULL);
}

void o_hlld_face_de(const double wli[8], const double wri[8], double gamma,
                    double flxo[8], double de[2]) {
  hlld_impl(wli, wri, gamma, flxo, de);
}

/* enzo_riemann_utils::compute_conserved (EnzoRiemannUtils.hpp:48-82) for
 * the MHD LUT, and active_fluxes (:111-149). */
static void mhd_cons_and_flux(const double *w, double gamma, double *cons,
                              double *flux) {
  for (int i = 0; i < 4; i++) cons[i] = w[i];
  for (int i = 4; i < 8; i++) cons[i] = w[i] * w[L_D];
  {
    double density = w[L_D];
    double pressure = w[L_E];
    double internal_edens = pressure / (gamma - 1.0);
    const double vi = w[L_VI], vj = w[L_VJ], vk = w[L_VK];
    const double kinetic = 0.5 * density * ((vi * vi) + ((vj * vj) + (vk * vk)));
    double magnetic = 0.5 * ((w[L_BI] * w[L_BI]) +
                             ((w[L_BJ] * w[L_BJ]) + (w[L_BK] * w[L_BK])));
    cons[L_E] = internal_edens + kinetic + magnetic;
  }
  double vi = w[L_VI], vj = w[L_VJ], vk = w[L_VK];
  double Bi = w[L_BI], Bj = w[L_BJ], Bk = w[L_BK];
  double etot = cons[L_E];
  double ptot = w[L_E] + 0.5 * ((Bi * Bi) + ((Bj * Bj) + (Bk * Bk)));
  double mom_i = cons[L_VI];
  flux[L_D] = mom_i;
  flux[L_VI] = mom_i * vi - Bi * Bi + ptot;
  flux[L_VJ] = mom_i * vj - Bj * Bi;
  flux[L_VK] = mom_i * vk - Bk * Bi;
  flux[L_E] = ((etot + ptot) * vi - (Bi * vi + (Bj * vj + Bk * vk)) * Bi);
  flux[L_BI] = 0;
  flux[L_BJ] = Bj * vi - Bi * vj;
  flux[L_BK] = Bk * vi - Bi * vk;
}

/* EinfeldtWavespeed<MHDLUT>::operator() + roe_cfast_
 * (EnzoRiemannHLL.hpp:44-115, 134-172) */
static void einfeldt_mhd(const double *wl, const double *wr, const double *Ul,
                         const double *Ur, double gamma, double *bp,
                         double *bm) {
  double c_l = fast_ms_speed(wl[L_D], wl[L_E], wl[L_BI], wl[L_BJ], wl[L_BK],
                             gamma);
  double c_r = fast_ms_speed(wr[L_D], wr[L_E], wr[L_BI], wr[L_BJ], wr[L_BK],
                             gamma);
  double left_speed = (wl[L_VI] - c_l);
  double right_speed = (wr[L_VI] + c_r);
  double sqrtrho_l = sqrt(wl[L_D]);
  double sqrtrho_r = sqrt(wr[L_D]);
  double inv_sqrtrho_tot = 1.0 / (sqrtrho_l + sqrtrho_r);
  double vi_roe = (sqrtrho_l * wl[L_VI] + sqrtrho_r * wr[L_VI]) * inv_sqrtrho_tot;
  double vj_roe = (sqrtrho_l * wl[L_VJ] + sqrtrho_r * wr[L_VJ]) * inv_sqrtrho_tot;
  double vk_roe = (sqrtrho_l * wl[L_VK] + sqrtrho_r * wr[L_VK]) * inv_sqrtrho_tot;
  double v_roe2 = vi_roe * vi_roe + vj_roe * vj_roe + vk_roe * vk_roe;
  double ptot_l = wl[L_E];
  double ptot_r = wr[L_E];
  ptot_l += 0.5 * ((wl[L_BI] * wl[L_BI]) +
                   ((wl[L_BJ] * wl[L_BJ]) + (wl[L_BK] * wl[L_BK])));
  ptot_r += 0.5 * ((wr[L_BI] * wr[L_BI]) +
                   ((wr[L_BJ] * wr[L_BJ]) + (wr[L_BK] * wr[L_BK])));
  double h_l = (Ul[L_E] + ptot_l) / wl[L_D];
  double h_r = (Ur[L_E] + ptot_r) / wr[L_D];
  double h_roe = (sqrtrho_l * h_l + sqrtrho_r * h_r) * inv_sqrtrho_tot;

  /* roe_cfast_ */
  double rho_roe = sqrtrho_l * sqrtrho_r;
  double bi_roe = wl[L_BI];
  double bj_roe = (sqrtrho_l * wr[L_BJ] + sqrtrho_r * wl[L_BJ]) * inv_sqrtrho_tot;
  double bk_roe = (sqrtrho_l * wr[L_BK] + sqrtrho_r * wl[L_BK]) * inv_sqrtrho_tot;
  double b_roe2 = bi_roe * bi_roe + bj_roe * bj_roe + bk_roe * bk_roe;
  double gamma_prime = gamma - 1.;
  double dbj = wl[L_BJ] - wr[L_BJ];
  double dbk = wl[L_BK] - wr[L_BK];
  double x_prime = ((dbj * dbj + dbk * dbk) * 0.5 * (gamma_prime - 1) *
                    inv_sqrtrho_tot);
  double y_prime = ((gamma_prime - 1) * (wl[L_D] + wr[L_D]) * 0.5 / rho_roe);
  double tilde_a2 = (gamma_prime * (h_roe - 0.5 * v_roe2 - b_roe2 / rho_roe) -
                     x_prime);
  double tilde_vai2 = bi_roe * bi_roe / rho_roe;
  double tilde_va2 = (tilde_vai2 + (gamma_prime - y_prime) *
                                       (bj_roe * bj_roe + bk_roe * bk_roe) /
                                       rho_roe);
  double s = tilde_a2 + tilde_va2;
  double c_roe = sqrt(0.5 * (tilde_a2 + tilde_va2 +
                             sqrt(s * s - 4 * tilde_a2 * tilde_vai2)));
  *bp = fmax(vi_roe + c_roe, right_speed);
  *bm = fmin(vi_roe - c_roe, left_speed);
}

/* HLLKernel<EinfeldtWavespeed<MHDLUT>>::operator() (EnzoRiemannHLL.hpp:
 * 238-345) -- the "hlle" MHD solver; with de, the internal-energy flux and
 * vi_bar interpolated between the outer speeds (:322-344). */
static void hlle_mhd_impl(const double wl[8], const double wr[8],
                          double gamma, double flux[8], double *de) {
  double cl[8], cr[8], fl[8], fr[8];
  mhd_cons_and_flux(wl, gamma, cl, fl);
  mhd_cons_and_flux(wr, gamma, cr, fr);
  double bp, bm;
  einfeldt_mhd(wl, wr, cl, cr, gamma, &bp, &bm);
  bp = fmax(bp, 0.0);
  bm = fmin(bm, 0.0);
  double inv_speed_diff = 1. / (bp - bm);
  for (int f = 0; f < 8; f++) {
    flux[f] = ((bp * fl[f] - bm * fr[f] + (cr[f] - cl[f]) * bp * bm) *
               inv_speed_diff);
  }
  if (de) {
    de[0] = passive_eint_flux(wl[L_D], wl[L_E], wr[L_D], wr[L_E], gamma,
                              flux[L_D]);
    de[1] = (bp * wl[L_VI] - bm * wr[L_VI]) * inv_speed_diff;
  }
}

void o_hlle_mhd_face(const double wl[8], const double wr[8], double gamma,
                     double flux[8]) {
  hlle_mhd_impl(wl, wr, gamma, flux, NULL);
}

void o_hlle_mhd_face_de(const double wl[8], const double wr[8], double gamma,
                        double flux[8], double de[2]) {
  hlle_mhd_impl(wl, wr, gamma, flux, de);
}

/* compute_conserved for the hydro LUT (EnzoRiemannUtils.hpp:48-82) */
static void hydro_cons(const double *w, double gamma, double *cons) {
  cons[H_D] = w[H_D];
  for (int i = 1; i < 5; i++) cons[i] = w[i] * w[H_D];
  double density = w[H_D];
  double internal_edens = w[H_E] / (gamma - 1.0);
  const double vi = w[H_VI], vj = w[H_VJ], vk = w[H_VK];
  const double kinetic = 0.5 * density * ((vi * vi) + ((vj * vj) + (vk * vk)));
  double magnetic = 0.5 * ((0.0 * 0.0) + ((0.0 * 0.0) + (0.0 * 0.0)));
  cons[H_E] = internal_edens + kinetic + magnetic;
}

/* HLLCKernel::operator() (EnzoRiemannHLLC.hpp:34-172) with
 * EinfeldtWavespeed<HydroLUT> and roe_cs_ (EnzoRiemannHLL.hpp:44-130). */
static void hllc_impl(const double wl[5], const double wr[5], double gamma,
                      double flux[5], double *de) {
  const double pressure_l = wl[H_E];
  const double pressure_r = wr[H_E];
  double cons_l[5], cons_r[5];
  hydro_cons(wl, gamma, cons_l);
  hydro_cons(wr, gamma, cons_r);

  /* EinfeldtWavespeed: (bp -> cs_r, bm -> cs_l) */
  double cs_l, cs_r;
  {
    double c_l = sqrt(gamma * pressure_l / wl[H_D]);
    double c_r = sqrt(gamma * pressure_r / wr[H_D]);
    double left_speed = (wl[H_VI] - c_l);
    double right_speed = (wr[H_VI] + c_r);
    double sqrtrho_l = sqrt(wl[H_D]);
    double sqrtrho_r = sqrt(wr[H_D]);
    double inv_sqrtrho_tot = 1.0 / (sqrtrho_l + sqrtrho_r);
    double vi_roe = (sqrtrho_l * wl[H_VI] + sqrtrho_r * wr[H_VI]) * inv_sqrtrho_tot;
    double vj_roe = (sqrtrho_l * wl[H_VJ] + sqrtrho_r * wr[H_VJ]) * inv_sqrtrho_tot;
    double vk_roe = (sqrtrho_l * wl[H_VK] + sqrtrho_r * wr[H_VK]) * inv_sqrtrho_tot;
    double v_roe2 = vi_roe * vi_roe + vj_roe * vj_roe + vk_roe * vk_roe;
    double ptot_l = pressure_l;
    double ptot_r = pressure_r;
    double h_l = (cons_l[H_E] + ptot_l) / wl[H_D];
    double h_r = (cons_r[H_E] + ptot_r) / wr[H_D];
    double h_roe = (sqrtrho_l * h_l + sqrtrho_r * h_r) * inv_sqrtrho_tot;
    double tiny2 = 0.;
    double temp = h_roe - 0.5 * v_roe2;
    double c_roe = sqrt((gamma - 1) * smax(temp, tiny2));
    cs_r = fmax(vi_roe + c_roe, right_speed);
    cs_l = fmin(vi_roe - c_roe, left_speed);
  }

  double bm = fmin(cs_l, 0.0);
  double bp = fmax(cs_r, 0.0);

  double tl = (pressure_l - (cs_l - wl[H_VI]) * wl[H_D] * wl[H_VI]);
  double tr = (pressure_r - (cs_r - wr[H_VI]) * wr[H_D] * wr[H_VI]);
  double dl = wl[H_D] * (cs_l - wl[H_VI]);
  double dr = -wr[H_D] * (cs_r - wr[H_VI]);
  double q1 = 1.0 / (dl + dr);
  double cw = (tr - tl) * q1;
  double cp = (dl * tr + dr * tl) * q1;

  double sl, sr, sm;
  if (cw >= 0.) {
    sl = cw / (cw - bm);
    sr = 0.;
    sm = -bm / (cw - bm);
  } else {
    sl = 0.;
    sr = -cw / (bp - cw);
    sm = bp / (bp - cw);
  }
  cp = smax(cp, 0.0);

  double momentumi_l = cons_l[H_VI];
  double momentumi_r = cons_r[H_VI];
  double dfl = momentumi_l - bm * wl[H_D];
  double dfr = momentumi_r - bp * wr[H_D];
  double ufl = momentumi_l * (wl[H_VI] - bm) + pressure_l;
  double ufr = momentumi_r * (wr[H_VI] - bp) + pressure_r;
  double vfl = (wl[H_D] * wl[H_VJ] * (wl[H_VI] - bm));
  double vfr = (wr[H_D] * wr[H_VJ] * (wr[H_VI] - bp));
  double wfl = (wl[H_D] * wl[H_VK] * (wl[H_VI] - bm));
  double wfr = (wr[H_D] * wr[H_VK] * (wr[H_VI] - bp));
  double efl = (cons_l[H_E] * (wl[H_VI] - bm) + pressure_l * wl[H_VI]);
  double efr = (cons_r[H_E] * (wr[H_VI] - bp) + pressure_r * wr[H_VI]);

  flux[H_D] = sl * dfl + sr * dfr;
  flux[H_VI] = sl * ufl + sr * ufr;
  flux[H_VJ] = sl * vfl + sr * vfr;
  flux[H_VK] = sl * wfl + sr * wfr;
  flux[H_E] = sl * efl + sr * efr;
  flux[H_VI] += (sm * cp);
  flux[H_E] += (sm * cp * cw);
  if (de) {
    /* EnzoRiemannHLLC.hpp:159-170 */
    de[0] = passive_eint_flux(wl[H_D], pressure_l, wr[H_D], pressure_r, gamma,
                              flux[H_D]);
    de[1] = (sl * (wl[H_VI] - bm) + sr * (wr[H_VI] - bp));
  }
}

void o_hllc_face(const double wl[5], const double wr[5], double gamma,
                 double flux[5]) {
  hllc_impl(wl, wr, gamma, flux, NULL);
}

void o_hllc_face_de(const double wl[5], const double wr[5], double gamma,
                    double flux[5], double de[2]) {
  hllc_impl(wl, wr, gamma, flux, de);
}

/* EnzoRiemannImpl<K>::solve / solve_ (EnzoRiemannImpl.hpp:265-338):
 * faces in [s, F-s) of the face array along every axis.  pl/pr/flux are
 * indexed by cell-centred field slot (O_RHO.. with O_ETOT holding pressure
 * for pl/pr). */
static void riemann_solve_impl(const o_config *cfg, double *const *pl,
                               double *const *pr, double **flux,
                               double *eflux, double *vbar, int mz, int my,
                               int mx, int dim, int s) {
  int fz, fy, fx;
  face_shape(mz, my, mx, dim, &fz, &fy, &fx);
  const int vi = O_VX + dim, vj = O_VX + (dim + 1) % 3,
            vk = O_VX + (dim + 2) % 3;
  const int bi = O_BX + dim, bj = O_BX + (dim + 1) % 3,
            bk = O_BX + (dim + 2) % 3;
  OMP_FOR
  for (int k = s; k < fz - s; k++)
    for (int j = s; j < fy - s; j++)
      for (int i = s; i < fx - s; i++) {
        long f = IDX(k, j, i, fy, fx);
        if (cfg->mhd) {
          double wl[8], wr[8], fo[8];
          wl[L_D] = pl[O_RHO][f]; wr[L_D] = pr[O_RHO][f];
          wl[L_VI] = pl[vi][f];   wr[L_VI] = pr[vi][f];
          wl[L_VJ] = pl[vj][f];   wr[L_VJ] = pr[vj][f];
          wl[L_VK] = pl[vk][f];   wr[L_VK] = pr[vk][f];
          wl[L_BI] = pl[bi][f];   wr[L_BI] = pr[bi][f];
          wl[L_BJ] = pl[bj][f];   wr[L_BJ] = pr[bj][f];
          wl[L_BK] = pl[bk][f];   wr[L_BK] = pr[bk][f];
          wl[L_E] = pl[O_ETOT][f]; wr[L_E] = pr[O_ETOT][f];
          double de[2];
          double *dep = eflux ? de : NULL;
          if (cfg->riemann == O_HLLD) {
            hlld_impl(wl, wr, cfg->gamma, fo, dep);
          } else {
            hlle_mhd_impl(wl, wr, cfg->gamma, fo, dep);
          }
          if (eflux) {
            eflux[f] = de[0];
            vbar[f] = de[1];
          }
          flux[O_RHO][f] = fo[L_D];
          flux[vi][f] = fo[L_VI];
          flux[vj][f] = fo[L_VJ];
          flux[vk][f] = fo[L_VK];
          flux[bi][f] = fo[L_BI];
          flux[bj][f] = fo[L_BJ];
          flux[bk][f] = fo[L_BK];
          flux[O_ETOT][f] = fo[L_E];
        } else {
          double wl[5], wr[5], fo[5];
          wl[H_D] = pl[O_RHO][f]; wr[H_D] = pr[O_RHO][f];
          wl[H_VI] = pl[vi][f];   wr[H_VI] = pr[vi][f];
          wl[H_VJ] = pl[vj][f];   wr[H_VJ] = pr[vj][f];
          wl[H_VK] = pl[vk][f];   wr[H_VK] = pr[vk][f];
          wl[H_E] = pl[O_ETOT][f]; wr[H_E] = pr[O_ETOT][f];
          double de[2];
          hllc_impl(wl, wr, cfg->gamma, fo, eflux ? de : NULL);
          if (eflux) {
            eflux[f] = de[0];
            vbar[f] = de[1];
          }
          flux[O_RHO][f] = fo[H_D];
          flux[vi][f] = fo[H_VI];
          flux[vj][f] = fo[H_VJ];
          flux[vk][f] = fo[H_VK];
          flux[O_ETOT][f] = fo[H_E];
        }
      }
}

void o_riemann_solve(const o_config *cfg, double *const *pl,
                     double *const *pr, double **flux, int mz, int my,
                     int mx, int dim, int s) {
  riemann_solve_impl(cfg, pl, pr, flux, NULL, NULL, mz, my, mx, dim, s);
}

/* The solve with the dual-energy outputs (EnzoRiemannImpl.hpp:270-301):
 * the internal-energy flux and vi_bar of every face, face-trimmed arrays. */
void o_riemann_solve_de(const o_config *cfg, double *const *pl,
                        double *const *pr, double **flux, double *eflux,
                        double *vbar, int mz, int my, int mx, int dim,
                        int s) {
  riemann_solve_impl(cfg, pl, pr, flux, eflux, vbar, mz, my, mx, dim, s);
}

/* ------------------------------------------------------------------------ */
/* Pressure, primitives, dt                                                 */
/* ------------------------------------------------------------------------ */

/* EnzoComputePressure::compute_pressure, rank 3 (EnzoComputePressure.cpp:
 * 127-189): gm1 * rho * eint with dual energy, otherwise from the total
 * energy; region from enzo_utils::exec_loop (utils.hpp:122-148). */
static void pressure_region(const o_config *cfg, double *const *U,
                            const double *eint, double *p, int mz, int my,
                            int mx, int s) {
  const double gm1 = cfg->gamma - 1.0;
  OMP_FOR
  for (int k = s; k < mz - s; k++)
    for (int j = s; j < my - s; j++)
      for (int i = s; i < mx - s; i++) {
        long c = IDX(k, j, i, my, mx);
        if (cfg->dual_energy) {
          p[c] = gm1 * U[O_RHO][c] * eint[c];
          continue;
        }
        double vx = U[O_VX][c], vy = U[O_VY][c], vz = U[O_VZ][c];
        double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
        double me_den = 0.;
        if (cfg->mhd) {
          double bx = U[O_BX][c], by = U[O_BY][c], bz = U[O_BZ][c];
          me_den = 0.5 * (bx * bx + by * by + bz * bz);
        }
        p[c] = gm1 * (U[O_RHO][c] * (U[O_ETOT][c] - ke) - me_den);
      }
}

void o_compute_pressure(const o_config *cfg, const o_block *b, double *p,
                        int s) {
  pressure_region(cfg, b->cc, b->eint, p, b->mz, b->my, b->mx, s);
}

/* EnzoPhysicsFluidProps::primitive_from_integration
 * (EnzoPhysicsFluidProps.cpp:64-138): copy density, velocity and bfield over
 * [s, m-s), then pressure into the O_ETOT slot. */
static void primitive_from_integration(const o_config *cfg, double *const *U,
                                       const double *eint, double **W,
                                       int mz, int my, int mx, int s) {
  for (int q = 0; q < O_NCC; q++) {
    if (q == O_ETOT) continue;
    if (!cfg->mhd && q >= O_BX) continue;
    OMP_FOR
    for (int k = s; k < mz - s; k++)
      for (int j = s; j < my - s; j++)
        for (int i = s; i < mx - s; i++) {
          long c = IDX(k, j, i, my, mx);
          W[q][c] = U[q][c];
        }
  }
  pressure_region(cfg, U, eint, W[O_ETOT], mz, my, mx, s);
}

/* EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync
 * (EnzoPhysicsFluidProps.cpp:158-290).  Without dual energy the total
 * energy gets the floor eint_floor + non-thermal energy; with the "modern"
 * formulation the internal energy is replaced by etot - non-thermal where
 * the sound speed that implies exceeds eta times the kinetic and magnetic
 * scales (and eint_1 > eint / 2 unless eta = 0), floored, and the total
 * energy rebuilt from it.  ggm1 is a float in the reference (:236). */
void o_floor_energy_and_sync(const o_config *cfg, double *const *U,
                             double *eint, int mz, int my, int mx, int s) {
  const int idual = cfg->dual_energy;
  const double eta = idual ? cfg->eta : 0.0;
  const float ggm1 = (float)(cfg->gamma * (cfg->gamma - 1.));
  const double pressure_floor = cfg->pressure_floor;
  const double inv_gm1 = 1. / (cfg->gamma - 1.);
  const double half_factor = (eta != 0.) ? 0.5 : 0.;
  OMP_FOR
  for (int k = s; k < mz - s; k++)
    for (int j = s; j < my - s; j++)
      for (int i = s; i < mx - s; i++) {
        long c = IDX(k, j, i, my, mx);
        double inv_rho = 1. / U[O_RHO][c];
        double eint_floor = pressure_floor * inv_gm1 * inv_rho;
        double vx = U[O_VX][c], vy = U[O_VY][c], vz = U[O_VZ][c];
        double v2 = (vx * vx + vy * vy + vz * vz);
        double non_thermal_e = 0.5 * v2;
        double b2 = 0;
        if (cfg->mhd) {
          double bx = U[O_BX][c], by = U[O_BY][c], bz = U[O_BZ][c];
          b2 = (bx * bx + by * by + bz * bz);
          non_thermal_e += (0.5 * b2 * inv_rho);
        }
        if (idual) {
          double eint_1 = U[O_ETOT][c] - non_thermal_e;
          double cur_eint = eint[c];
          double cs2_1 = fmax(0., ggm1 * eint_1);
          if ((cs2_1 > fmax(eta * v2, eta * b2 * inv_rho)) &&
              (eint_1 > half_factor * cur_eint)) {
            cur_eint = eint_1;
          }
          cur_eint = apply_floor(cur_eint, eint_floor);
          eint[c] = cur_eint;
          U[O_ETOT][c] = cur_eint + non_thermal_e;
        } else {
          double etot_floor = eint_floor + non_thermal_e;
          U[O_ETOT][c] = apply_floor(U[O_ETOT][c], etot_floor);
        }
      }
}

/* EnzoMethodMHDVlct::timestep (EnzoMethodMHDVlct.cpp:551-588) ->
 * EnzoMHDIntegratorStageCommands::timestep
 * (EnzoMHDIntegratorStageCommands.cpp:299-366): min over every cell,
 * ghosts included.  With dual energy the block's energies are synchronised
 * first, in place (:563-568). */
double o_vlct_timestep(const o_config *cfg, o_block *b) {
  const int mz = b->mz, my = b->my, mx = b->mx;
  if (cfg->dual_energy) o_floor_energy_and_sync(cfg, b->cc, b->eint, mz, my, mx, 0);
  pressure_region(cfg, b->cc, b->eint, b->pressure, mz, my, mx, 0);
  const double dx = b->cw[0], dy = b->cw[1], dz = b->cw[2];
  const double gamma = cfg->gamma;
  double dt = 1.7976931348623157e308; /* numeric_limits<double>::max() */
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        long c = IDX(k, j, i, my, mx);
        double d = b->cc[O_RHO][c], p = b->pressure[c];
        double sig;
        if (cfg->mhd) {
          double bx = b->cc[O_BX][c], by = b->cc[O_BY][c],
                 bz = b->cc[O_BZ][c];
          double B2 = ((bx * bx) + ((by * by) + (bz * bz)));
          double cs2 = gamma * p / d;
          double va2 = B2 / d;
          sig = sqrt(va2 + cs2);
        } else {
          sig = sqrt(gamma * p / d);
        }
        double local_dt = min3(dx / (fabs(b->cc[O_VX][c]) + sig),
                               dy / (fabs(b->cc[O_VY][c]) + sig),
                               dz / (fabs(b->cc[O_VZ][c]) + sig));
        dt = smin(dt, local_dt);
      }
  return dt * cfg->courant;
}

/* ------------------------------------------------------------------------ */
/* Flux divergence and update                                               */
/* ------------------------------------------------------------------------ */

/* integration keys updated through dU (EnzoIntegrationQuanUpdate.cpp:61-79,
 * skip_B_update=true): density; velocity_x, _y, _z, total_energy. */
static const int dU_fields[5] = {O_RHO, O_VX, O_VY, O_VZ, O_ETOT};

/* EnzoIntegrationQuanUpdate::accumulate_flux_component
 * (EnzoIntegrationQuanUpdate.cpp:105-147): cells c in [s+1, m_d-s-1) along
 * dim, [s, m-s) transverse: dU(c) -= dtdx * (F(c) - F(c-1)).  With dual
 * energy the "internal_energy" key (eflux into dUe) follows. */
static void accumulate_one(int dim, double dtdx_i, const double *F, double *D,
                           int mz, int my, int mx, int s) {
  int fz, fy, fx;
  face_shape(mz, my, mx, dim, &fz, &fy, &fx);
  const int di = (dim == 0), dj = (dim == 1), dk = (dim == 2);
  OMP_FOR
  for (int k = s + dk; k < mz - s - dk; k++)
    for (int j = s + dj; j < my - s - dj; j++)
      for (int i = s + di; i < mx - s - di; i++) {
        long c = IDX(k, j, i, my, mx);
        double fr = F[IDX(k, j, i, fy, fx)];
        double fl = F[IDX(k - dk, j - dj, i - di, fy, fx)];
        D[c] -= dtdx_i * (fr - fl);
      }
}

static void accumulate_flux(int dim, double dt, double cw, double **flux,
                            double **dU, const double *eflux, double *dUe,
                            int mz, int my, int mx, int s) {
  double dtdx_i = dt / cw;
  for (int q = 0; q < 5; q++)
    accumulate_one(dim, dtdx_i, flux[dU_fields[q]], dU[q], mz, my, mx, s);
  if (eflux) accumulate_one(dim, dtdx_i, eflux, dUe, mz, my, mx, s);
}

void o_internal_energy_source(const o_config *cfg, int dim, double dt,
                              double cw, const double *p, double *deint,
                              const double *vbar, int mz, int my, int mx,
                              int s) {
  const double dtdx = dt / cw;
  const double p_floor = cfg->pressure_floor;
  int fz, fy, fx;
  face_shape(mz, my, mx, dim, &fz, &fy, &fx);
  const int di = (dim == 0), dj = (dim == 1), dk = (dim == 2);
  OMP_FOR
  for (int k = s + dk; k < mz - s - dk; k++)
    for (int j = s + dj; j < my - s - dj; j++)
      for (int i = s + di; i < mx - s - di; i++) {
        long c = IDX(k, j, i, my, mx);
        double vr = vbar[IDX(k, j, i, fy, fx)];
        double vl = vbar[IDX(k - dk, j - dj, i - di, fy, fx)];
        double pc = apply_floor(p[c], p_floor);
        deint[c] -= dtdx * pc * (vr - vl);
      }
}

/* EnzoIntegrationQuanUpdate::update_quantities
 * (EnzoIntegrationQuanUpdate.cpp:183-238) over [s+1, m-s-1) -- the specific
 * quantities include internal_energy with dual energy -- followed by
 * EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync at stale s+1. */
static void update_quantities_de(const o_config *cfg, double *const *U0,
                                 const double *eint0, double *const *dU,
                                 const double *dUe, double **Uo,
                                 double *eint_o, int mz, int my, int mx,
                                 int s) {
  const double density_floor = cfg->density_floor;
  OMP_FOR
  for (int k = s + 1; k < mz - s - 1; k++)
    for (int j = s + 1; j < my - s - 1; j++)
      for (int i = s + 1; i < mx - s - 1; i++) {
        long c = IDX(k, j, i, my, mx);
        double old_rho = U0[O_RHO][c];
        double new_rho = U0[O_RHO][c] + dU[0][c];
        new_rho = apply_floor(new_rho, density_floor);
        Uo[O_RHO][c] = new_rho;
        double inv_new_rho = 1. / new_rho;
        for (int q = 1; q < 5; q++) {
          int fq = dU_fields[q];
          Uo[fq][c] = (U0[fq][c] * old_rho + dU[q][c]) * inv_new_rho;
        }
        if (cfg->dual_energy)
          eint_o[c] = (eint0[c] * old_rho + dUe[c]) * inv_new_rho;
      }
  o_floor_energy_and_sync(cfg, Uo, eint_o, mz, my, mx, s + 1);
}

static void update_quantities(const o_config *cfg, double *const *U0,
                              double *const *dU, double **Uo, int mz, int my,
                              int mx, int s) {
  o_config c = *cfg;
  c.dual_energy = 0;
  update_quantities_de(&c, U0, NULL, dU, NULL, Uo, NULL, mz, my, mx, s);
}

/* ------------------------------------------------------------------------ */
/* Constrained transport                                                    */
/* ------------------------------------------------------------------------ */

/* EnzoBfieldMethodCT::correct_reconstructed_bfield
 * (EnzoBfieldMethodCT.cpp:122-166): interior faces of bfieldi[dim] (index
 * f+1) copied into both reconstructed states, region [s, F-s). */
static void correct_reconstructed_bfield(const double *bfi, double *wl,
                                         double *wr, int mz, int my, int mx,
                                         int dim, int s) {
  int fz, fy, fx;
  face_shape(mz, my, mx, dim, &fz, &fy, &fx);
  const int bny = my + (dim == 1), bnx = mx + (dim == 0);
  const int di = (dim == 0), dj = (dim == 1), dk = (dim == 2);
  OMP_FOR
  for (int k = s; k < fz - s; k++)
    for (int j = s; j < fy - s; j++)
      for (int i = s; i < fx - s; i++) {
        double v = bfi[IDX(k + dk, j + dj, i + di, bny, bnx)];
        long f = IDX(k, j, i, fy, fx);
        wl[f] = v;
        wr[f] = v;
      }
}

/* EnzoBfieldMethodCT::identify_upwind (EnzoBfieldMethodCT.cpp:170-216) */
static void identify_upwind(const double *rho_flux, double *w, int mz, int my,
                            int mx, int dim, int s) {
  int fz, fy, fx;
  face_shape(mz, my, mx, dim, &fz, &fy, &fx);
  OMP_FOR
  for (int k = s; k < fz - s; k++)
    for (int j = s; j < fy - s; j++)
      for (int i = s; i < fx - s; i++) {
        long f = IDX(k, j, i, fy, fx);
        double df = rho_flux[f];
        if (df > 0) {
          w[f] = 1.0;
        } else if (df < 0) {
          w[f] = 0.0;
        } else {
          w[f] = 0.5;
        }
      }
}

/* EnzoBfieldMethodCT::compute_center_efield (EnzoBfieldMethodCT.cpp:267-292)
 */
static void center_efield(int dim, double *E, double *const *U, int mz,
                          int my, int mx, int s) {
  int j = (dim + 1) % 3, k = (dim + 2) % 3;
  const double *vj = U[O_VX + j], *vk = U[O_VX + k];
  const double *bj = U[O_BX + j], *bk = U[O_BX + k];
  OMP_FOR
  for (int z = s; z < mz - s; z++)
    for (int y = s; y < my - s; y++)
      for (int x = s; x < mx - s; x++) {
        long c = IDX(z, y, x, my, mx);
        E[c] = (-vj[c] * bk[c] + vk[c] * bj[c]);
      }
}

/* index of position p (x,y,z order) in an array with extents n (x,y,z) */
static inline long at3(const int p[3], const int n[3]) {
  return (((long)p[2]) * n[1] + p[1]) * (long)n[0] + p[0];
}

/* EnzoBfieldMethodCT::compute_edge_efield + compute_edge_<true>
 * (EnzoBfieldMethodCT.cpp:384-557): upwinded edge EMF along axis i at
 * edges (cell a_i along i, face a_j along j, face a_k along k); region
 * a_i in [s+1, m_i-s-1), a_j in [s, m_j-s-1), a_k in [s, m_k-s-1).      */
static void edge_efield(int i, const double *Ec, double *edge,
                        double *const *jflux, double *const *kflux,
                        double *const *weight, int mz, int my, int mx,
                        int s) {
  const int j = (i + 1) % 3, k = (i + 2) % 3;
  const int m[3] = {mx, my, mz};
  int ne[3] = {mx, my, mz};  /* edge array extents */
  ne[j] -= 1; ne[k] -= 1;
  int nj[3] = {mx, my, mz};  /* j-face arrays */
  nj[j] -= 1;
  int nk[3] = {mx, my, mz};  /* k-face arrays */
  nk[k] -= 1;
  const double *Fj = jflux[O_BX + k]; /* jflux(B_k): -E_i on j faces */
  const double *Fk = kflux[O_BX + j]; /* kflux(B_j): +E_i on k faces */
  const double *Wj = weight[j], *Wk = weight[k];

  OMP_FOR
  for (int ak = s; ak < m[k] - s - 1; ak++)
    for (int aj = s; aj < m[j] - s - 1; aj++)
      for (int ai = s + 1; ai < m[i] - s - 1; ai++) {
        int p[3];
        p[i] = ai; p[j] = aj; p[k] = ak;
        int pj1[3] = {p[0], p[1], p[2]}; pj1[j] += 1;
        int pk1[3] = {p[0], p[1], p[2]}; pk1[k] += 1;
        int pjk1[3] = {p[0], p[1], p[2]}; pjk1[j] += 1; pjk1[k] += 1;

        double vEj = Fj[at3(p, nj)];
        double vEj_kp1 = Fj[at3(pk1, nj)];
        double vEk = Fk[at3(p, nk)];
        double vEk_jp1 = Fk[at3(pj1, nk)];
        double vEc = Ec[at3(p, m)];
        double vEc_jp1 = Ec[at3(pj1, m)];
        double vEc_kp1 = Ec[at3(pk1, m)];
        double vEc_jkp1 = Ec[at3(pjk1, m)];
        double vWj = Wj[at3(p, nj)];
        double vWj_kp1 = Wj[at3(pk1, nj)];
        double vWk = Wk[at3(p, nk)];
        double vWk_jp1 = Wk[at3(pj1, nk)];

        double dEdj_r = vWk_jp1 * (vEc_jp1 + vEj) +
                        (1 - vWk_jp1) * (vEc_jkp1 + vEj_kp1);
        double dEdj_l = vWk * (-vEj - vEc) +
                        (1 - vWk) * (-vEj_kp1 - vEc_kp1);
        double dEdk_r = vWj_kp1 * (vEc_kp1 - vEk) +
                        (1 - vWj_kp1) * (vEc_jkp1 - vEk_jp1);
        double dEdk_l = vWj * (vEk - vEc) + (1 - vWj) * (vEk_jp1 - vEc_jp1);

        double Ej_sum = vEj + vEj_kp1;
        Ej_sum *= -1;
        double Ek_sum = vEk + vEk_jp1;
        edge[at3(p, ne)] =
            0.25 * (Ej_sum + Ek_sum + (dEdj_l - dEdj_r) + (dEdk_l - dEdk_r));
      }
}

/* EnzoBfieldMethodCT::update_bfield (EnzoBfieldMethodCT.cpp:617-687): faces
 * X in [s+1, m_i-s) along i, cells [s+1, m-s-1) along j and k:
 * Bout = Bcur - dt/dj (E_k(j+1/2) - E_k(j-1/2)) + dt/dk (E_j(k+1/2) -
 * E_j(k-1/2)), the edges taken at i-position X-1/2.                       */
void o_update_bfield(const double cw[3], int i, double *const *edge,
                     const double *bcur, double *bout, double dt, int mz,
                     int my, int mx, int s) {
  const int j = (i + 1) % 3, k = (i + 2) % 3;
  const int m[3] = {mx, my, mz};
  double dtdj = dt / cw[j];
  double dtdk = dt / cw[k];
  int nb[3] = {mx, my, mz};
  nb[i] += 1;
  int nEj[3] = {mx, my, mz}; /* edge[j]: cell along j, faces along k, i */
  nEj[k] -= 1; nEj[i] -= 1;
  int nEk[3] = {mx, my, mz}; /* edge[k]: cell along k, faces along i, j */
  nEk[i] -= 1; nEk[j] -= 1;
  const double *Ej = edge[j], *Ek = edge[k];
  OMP_FOR
  for (int ak = s + 1; ak < m[k] - s - 1; ak++)
    for (int aj = s + 1; aj < m[j] - s - 1; aj++)
      for (int X = s + 1; X < m[i] - s; X++) {
        int pb[3];
        pb[i] = X; pb[j] = aj; pb[k] = ak;
        int pe[3];
        pe[i] = X - 1; pe[j] = aj; pe[k] = ak;
        int pe_jm[3] = {pe[0], pe[1], pe[2]}; pe_jm[j] -= 1;
        int pe_km[3] = {pe[0], pe[1], pe[2]}; pe_km[k] -= 1;
        double E_k_term = dtdj * (Ek[at3(pe, nEk)] - Ek[at3(pe_jm, nEk)]);
        double E_j_term = dtdk * (Ej[at3(pe, nEj)] - Ej[at3(pe_km, nEj)]);
        long cb = at3(pb, nb);
        bout[cb] = bcur[cb] - E_k_term + E_j_term;
      }
}

/* Exported single-operation entry points (same loops as above), for the
 * unit parity tests of the per-operation device API. */
void o_correct_reconstructed_bfield(const double *bfi, double *wl, double *wr,
                                    int mz, int my, int mx, int dim, int s) {
  correct_reconstructed_bfield(bfi, wl, wr, mz, my, mx, dim, s);
}
void o_identify_upwind(const double *rho_flux, double *w, int mz, int my,
                       int mx, int dim, int s) {
  identify_upwind(rho_flux, w, mz, my, mx, dim, s);
}
void o_compute_center_efield(int dim, double *E, double *const *U, int mz,
                             int my, int mx, int s) {
  center_efield(dim, E, U, mz, my, mx, s);
}
void o_compute_edge_efield(int dim, const double *Ec, double *edge,
                           double *const *jflux, double *const *kflux,
                           double *const *weight, int mz, int my, int mx,
                           int s) {
  edge_efield(dim, Ec, edge, jflux, kflux, weight, mz, my, mx, s);
}
void o_accumulate_flux_component(int dim, double dt, double cw, double **flux,
                                 double **dU, int mz, int my, int mx, int s) {
  accumulate_flux(dim, dt, cw, flux, dU, NULL, NULL, mz, my, mx, s);
}
void o_accumulate_flux_component_de(int dim, double dt, double cw,
                                    double **flux, double **dU,
                                    const double *eflux, double *dUe, int mz,
                                    int my, int mx, int s) {
  accumulate_flux(dim, dt, cw, flux, dU, eflux, dUe, mz, my, mx, s);
}
void o_update_quantities_de(const o_config *cfg, double *const *U0,
                            const double *eint0, double *const *dU,
                            const double *dUe, double **Uo, double *eint_o,
                            int mz, int my, int mx, int s) {
  update_quantities_de(cfg, U0, eint0, dU, dUe, Uo, eint_o, mz, my, mx, s);
}
void o_update_quantities(const o_config *cfg, double *const *U0,
                         double *const *dU, double **Uo, int mz, int my,
                         int mx, int s) {
  update_quantities(cfg, U0, dU, Uo, mz, my, mx, s);
}
void o_primitive_from_integration(const o_config *cfg, double *const *U,
                                  double **W, int mz, int my, int mx, int s) {
  primitive_from_integration(cfg, U, NULL, W, mz, my, mx, s);
}
void o_primitive_from_integration_de(const o_config *cfg, double *const *U,
                                     const double *eint, double **W, int mz,
                                     int my, int mx, int s) {
  primitive_from_integration(cfg, U, eint, W, mz, my, mx, s);
}

/* EnzoBfieldMethodCT::compute_center_bfield (EnzoBfieldMethodCT.cpp:702-728)
 * over cells [s, m-s): Bc = 0.5*(B_{c-1/2} + B_{c+1/2}). */
void o_compute_center_bfield(int dim, double *bc, const double *bi, int mz,
                             int my, int mx, int s) {
  const int bny = my + (dim == 1), bnx = mx + (dim == 0);
  const int di = (dim == 0), dj = (dim == 1), dk = (dim == 2);
  OMP_FOR
  for (int k = s; k < mz - s; k++)
    for (int j = s; j < my - s; j++)
      for (int i = s; i < mx - s; i++) {
        double l = bi[IDX(k, j, i, bny, bnx)];
        double r = bi[IDX(k + dk, j + dj, i + di, bny, bnx)];
        bc[IDX(k, j, i, my, mx)] = 0.5 * (l + r);
      }
}

/* ------------------------------------------------------------------------ */
/* Stage driver                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
  int mx, my, mz, mhd;
  double *temp_cc[O_NCC];
  double *prim[O_NCC];
  double *pl[O_NCC], *pr[O_NCC];
  double *flux[3][O_NCC];
  double *dU[5];
  double *weight[3];
  double *edge[3];
  double *ecenter;
  double *tbfi[3];
  /* dual energy: half-step internal energy, its flux and vi_bar per dim,
     its accumulated change */
  double *temp_eint;
  double *eflux[3], *vbar[3];
  double *dUe;
} o_scratch;

static void scratch_free(o_scratch *sc) {
  if (!sc) return;
  for (int q = 0; q < O_NCC; q++) {
    free(sc->temp_cc[q]); free(sc->prim[q]); free(sc->pl[q]); free(sc->pr[q]);
    for (int d = 0; d < 3; d++) free(sc->flux[d][q]);
  }
  for (int q = 0; q < 5; q++) free(sc->dU[q]);
  for (int d = 0; d < 3; d++) {
    free(sc->weight[d]); free(sc->edge[d]); free(sc->tbfi[d]);
  }
  free(sc->ecenter);
  free(sc->temp_eint);
  for (int d = 0; d < 3; d++) {
    free(sc->eflux[d]); free(sc->vbar[d]);
  }
  free(sc->dUe);
  free(sc);
}

static double *zalloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

static o_scratch *scratch_get(int mz, int my, int mx, int mhd) {
  static __thread o_scratch *cache = NULL;
  if (cache && cache->mx == mx && cache->my == my && cache->mz == mz &&
      cache->mhd == mhd)
    return cache;
  scratch_free(cache);
  o_scratch *sc = (o_scratch *)calloc(1, sizeof(o_scratch));
  sc->mx = mx; sc->my = my; sc->mz = mz; sc->mhd = mhd;
  long ncc = (long)mx * my * mz;
  for (int q = 0; q < O_NCC; q++) {
    sc->temp_cc[q] = zalloc(ncc);
    sc->prim[q] = zalloc(ncc);
    sc->pl[q] = zalloc(ncc);
    sc->pr[q] = zalloc(ncc);
    for (int d = 0; d < 3; d++) sc->flux[d][q] = zalloc(ncc);
  }
  for (int q = 0; q < 5; q++) sc->dU[q] = zalloc(ncc);
  for (int d = 0; d < 3; d++) {
    sc->weight[d] = zalloc(ncc);
    sc->edge[d] = zalloc(ncc);
  }
  sc->ecenter = zalloc(ncc);
  sc->temp_eint = zalloc(ncc);
  for (int d = 0; d < 3; d++) {
    sc->eflux[d] = zalloc(ncc);
    sc->vbar[d] = zalloc(ncc);
  }
  sc->dUe = zalloc(ncc);
  sc->tbfi[0] = zalloc((long)mz * my * (mx + 1));
  sc->tbfi[1] = zalloc((long)mz * (my + 1) * mx);
  sc->tbfi[2] = zalloc((long)(mz + 1) * my * mx);
  cache = sc;
  return sc;
}

/* reconstructor staling rates (EnzoReconstructorNN.hpp:31-35,
 * EnzoReconstructorPLM.hpp:153-157) */
static int total_staling(int kind) { return (kind == O_NN) ? 1 : 2; }
static int immediate_staling(int kind) { return (kind == O_NN) ? 0 : 1; }

/* EnzoMHDIntegratorStageCommands::compute_update_stage
 * (EnzoMHDIntegratorStageCommands.cpp:102-203) with compute_flux_
 * (:207-261) and, for CT, EnzoBfieldMethodCT::update_all_bfield_components
 * (EnzoBfieldMethodCT.cpp:220-263). */
/* EnzoSourceGravity::calculate_source (toolkit/EnzoSourceGravity.cpp:
 * 17-69) with the start-of-step density and velocity, over [s, m-s). */
static void gravity_source(double dt, double *const *U0, double *const *acc,
                           double **dU, int mz, int my, int mx, int s) {
  OMP_FOR
  for (int k = s; k < mz - s; k++)
    for (int j = s; j < my - s; j++)
      for (int i = s; i < mx - s; i++) {
        long c = IDX(k, j, i, my, mx);
        double rho = U0[O_RHO][c];
        double ax = acc[0][c], ay = acc[1][c], az = acc[2][c];
        dU[1][c] += dt * rho * ax;
        dU[2][c] += dt * rho * ay;
        dU[3][c] += dt * rho * az;
        dU[4][c] += dt * rho * ((U0[O_VX][c] * ax) + (U0[O_VY][c] * ay) +
                                (U0[O_VZ][c] * az));
      }
}

static void compute_update_stage(const o_config *cfg, o_block *b,
                                 o_scratch *sc, double **U0, double **Uc,
                                 double **Uo, double *eint0, double *eint_c,
                                 double *eint_o, double **bfi_recon,
                                 double **bfi_out, int stage, double cur_dt,
                                 int stale, double *const *accel) {
  const int mz = b->mz, my = b->my, mx = b->mx;
  const long ncc = (long)mx * my * mz;
  const int kind = cfg->recon[stage];

  /* clear_dUcons_map (EnzoIntegrationQuanUpdate.cpp:83-101) */
  for (int q = 0; q < 5; q++) memset(sc->dU[q], 0, ncc * sizeof(double));
  const int de = cfg->dual_energy;
  if (de) memset(sc->dUe, 0, ncc * sizeof(double));

  primitive_from_integration(cfg, Uc, eint_c, sc->prim, mz, my, mx, stale);

  for (int dim = 0; dim < 3; dim++) {
    /* reconstruct_interface: active keys are the primitive keys; floors on
     * density and pressure (EnzoReconstructorPLM.hpp:234-245) */
    for (int q = 0; q < O_NCC; q++) {
      if (!cfg->mhd && q >= O_BX) continue;
      int use_floor = (q == O_RHO) || (q == O_ETOT);
      double fv = (q == O_RHO) ? cfg->density_floor
                               : ((q == O_ETOT) ? cfg->pressure_floor : 0.);
      o_reconstruct(kind, cfg->theta_limiter, sc->prim[q], sc->pl[q],
                    sc->pr[q], mz, my, mx, dim, stale, use_floor, fv);
    }
    int cur_stale = stale + immediate_staling(kind);
    if (cfg->mhd) {
      correct_reconstructed_bfield(bfi_recon[dim], sc->pl[O_BX + dim],
                                   sc->pr[O_BX + dim], mz, my, mx, dim,
                                   cur_stale);
    }
    riemann_solve_impl(cfg, sc->pl, sc->pr, sc->flux[dim],
                       de ? sc->eflux[dim] : NULL, de ? sc->vbar[dim] : NULL,
                       mz, my, mx, dim, cur_stale);
    accumulate_flux(dim, cur_dt, b->cw[dim], sc->flux[dim], sc->dU,
                    de ? sc->eflux[dim] : NULL, sc->dUe, mz, my, mx,
                    cur_stale);
    /* the internal-energy source of this dim
       (EnzoMHDIntegratorStageCommands.cpp:248-255) */
    if (de)
      o_internal_energy_source(cfg, dim, cur_dt, b->cw[dim], sc->prim[O_ETOT],
                               sc->dUe, sc->vbar[dim], mz, my, mx, cur_stale);
    if (cfg->mhd) {
      identify_upwind(sc->flux[dim][O_RHO], sc->weight[dim], mz, my, mx, dim,
                      cur_stale);
    }
  }
  stale += immediate_staling(kind);

  /* compute_source_terms_ (EnzoMHDIntegratorStageCommands.cpp:265-295):
   * gravity on the stage with index 1 only, from the start-of-step state */
  if (stage == 1 && accel)
    gravity_source(cur_dt, U0, accel, sc->dU, mz, my, mx, stale);

  if (cfg->mhd) {
    /* compute_all_edge_efields (EnzoBfieldMethodCT.cpp:561-590) */
    for (int i = 0; i < 3; i++) {
      center_efield(i, sc->ecenter, Uc, mz, my, mx, stale);
      int jd = (i + 1) % 3, kd = (i + 2) % 3;
      edge_efield(i, sc->ecenter, sc->edge[i], sc->flux[jd], sc->flux[kd],
                  sc->weight, mz, my, mx, stale);
    }
    for (int d = 0; d < 3; d++) {
      o_update_bfield(b->cw, d, sc->edge, b->bfi[d], bfi_out[d], cur_dt, mz,
                      my, mx, stale);
    }
    for (int d = 0; d < 3; d++) {
      o_compute_center_bfield(d, Uo[O_BX + d], bfi_out[d], mz, my, mx, stale);
    }
  }

  update_quantities_de(cfg, U0, eint0, sc->dU, sc->dUe, Uo, eint_o, mz, my,
                       mx, stale);
}

/* EnzoMethodMHDVlct::compute (EnzoMethodMHDVlct.cpp:356-500); with
 * flux_data, store_fluxes_for_corrections_ (:137-141, 360-363, 480-489):
 * the FluxData is allocated zeroed at the start (allocate_FC_flux_buffer_,
 * :333-355) and the final stage's fluxes are saved on the block faces. */
static int vlct_compute_impl(const o_config *cfg, o_block *b, double dt,
                             int ghost, int nf, const int *fields,
                             double *flux_data, double *const *accel);

int o_vlct_compute(const o_config *cfg, o_block *b, double dt) {
  return vlct_compute_impl(cfg, b, dt, 0, 0, NULL, NULL, NULL);
}

int o_vlct_compute_fc(const o_config *cfg, o_block *b, double dt, int ghost,
                      int nf, const int *fields, double *flux_data) {
  /* EnzoMethodMHDVlct.cpp:137-141: hydro only */
  if (cfg->mhd) return -2;
  return vlct_compute_impl(cfg, b, dt, ghost, nf, fields, flux_data, NULL);
}

int o_vlct_compute_accel(const o_config *cfg, o_block *b, double dt,
                         double *const *accel) {
  return vlct_compute_impl(cfg, b, dt, 0, 0, NULL, NULL, accel);
}

static int vlct_compute_impl(const o_config *cfg, o_block *b, double dt,
                             int ghost, int nf, const int *fields,
                             double *flux_data, double *const *accel) {
  if (cfg->mhd && cfg->nstages != 2) return -1; /* EnzoBfieldMethod.cpp:23-27 */
  int req = 0;
  for (int st = 0; st < cfg->nstages; st++) req += total_staling(cfg->recon[st]);
  (void)req;
  o_scratch *sc = scratch_get(b->mz, b->my, b->mx, cfg->mhd);
  int stale = 0;
  for (int stage = 0; stage < cfg->nstages; stage++) {
    const int is_final = (stage + 1) == cfg->nstages;
    const double cur_dt = (!is_final) ? dt / 2. : dt;
    double **Uc = (stage == 0) ? b->cc : sc->temp_cc;
    double **Uo = is_final ? b->cc : sc->temp_cc;
    /* CT: correct_reconstructed_bfield reads bfieldi at stage 0 and the
     * temporary face field at stage 1; updates write the temporary face
     * field at stage 0 and bfieldi at the final stage
     * (EnzoBfieldMethodCT.cpp:128-137, 228-240). */
    double **bfi_recon = (stage == 0) ? b->bfi : sc->tbfi;
    double **bfi_out = is_final ? b->bfi : sc->tbfi;
    if (stage == 0 && flux_data) {
      const int nx = b->mx - 2 * ghost, ny = b->my - 2 * ghost,
                nz = b->mz - 2 * ghost;
      memset(flux_data, 0,
             (size_t)o_flux_data_size(nf, nx, ny, nz) * sizeof(double));
    }
    double *eint_c = (stage == 0) ? b->eint : sc->temp_eint;
    double *eint_o = is_final ? b->eint : sc->temp_eint;
    compute_update_stage(cfg, b, sc, b->cc, Uc, Uo, b->eint, eint_c, eint_o,
                         bfi_recon, bfi_out, stage, cur_dt, stale, accel);
    stale += total_staling(cfg->recon[stage]);
    if (is_final && flux_data) {
      for (int d = 0; d < 3; d++)
        o_save_fluxes_for_corrections(sc->flux[d], d, b->cw[d], cur_dt, b->mx,
                                      b->my, b->mz, ghost, nf, fields,
                                      flux_data);
    }
  }
  return 0;
}
