// vlct_physics.hpp -- per-face / per-cell device functions of the VL+CT step.
//
// Templated on the arithmetic type T (double for the reference-parity path,
// float for the FP32 reconstruction/Riemann stage of the mixed path).  The
// expression trees keep the reference's parenthesisation and left-to-right
// association; the library is built with -ffp-contract=off so the FP64 path
// rounds exactly like the reference's CPU build.
#pragma once

#include <hip/hip_runtime.h>

// The namespace of the device code and of its launch templates: `enzoe` in
// the default build; a unit that compiles the same templates under other
// floating-point flags (the contracted FP64 step) names another one before
// including, so that the two sets of instantiations never share a symbol.
#ifndef ENZOE_KNS
#define ENZOE_KNS enzoe
#endif

namespace ENZOE_KNS {

// field slots used for cell-centred state, primitives and fluxes
enum : int {
  S_RHO = 0, S_VX = 1, S_VY = 2, S_VZ = 3, S_E = 4, S_BX = 5, S_BY = 6,
  S_BZ = 7, S_N = 8
};
// Riemann LUT order, MHD (EnzoRiemannImpl.hpp:40-56)
enum : int { L_D = 0, L_BI, L_BJ, L_BK, L_VI, L_VJ, L_VK, L_E };
// Riemann LUT order, hydro (EnzoRiemannImpl.hpp:21-36)
enum : int { H_D = 0, H_VI, H_VJ, H_VK, H_E };

enum : int { RS_HLLE = 0, RS_HLLC = 1, RS_HLLD = 2 };
enum : int { RC_NN = 0, RC_PLM = 1, RC_PLM_ATHENA = 2 };

// Division and square root: exact IEEE operations in FP64 (reference
// parity); hardware reciprocal / square root (about 1 ulp) in FP32, where
// the mixed-precision path is judged by its L1 error instead.
__device__ __forceinline__ double qdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float qdiv(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
__device__ __forceinline__ double qsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float qsqrt(float a) {
  return __builtin_amdgcn_sqrtf(a);
}

template <typename T> __device__ __forceinline__ T dmin(T a, T b) {
  return (b < a) ? b : a;  // std::min
}
template <typename T> __device__ __forceinline__ T dmax(T a, T b) {
  return (a < b) ? b : a;  // std::max
}
// enzo_utils::min of three values (src/Enzo/utils/utils.hpp:82-89)
template <typename T> __device__ __forceinline__ T dmin3(T a, T b, T c) {
  if (a < b) return (c < a) ? c : a;
  return (c < b) ? c : b;
}
// The same minimum of the magnitudes |a|, |b|, |c| with the hardware
// minimum (IEEE minNum): the two agree on every input, NaN included, except
// where |b| is NaN (the reference's comparisons then return |b|, minNum
// would drop it), which the final select restores.  Used by the PLM limiter,
// whose arguments are differences (never signalling NaNs).
template <typename T> __device__ __forceinline__ T dmin3_abs(T a, T b, T c) {
  const T m = __builtin_fmin(__builtin_fmin(fabs(a), fabs(b)), fabs(c));
  return (b == b) ? m : fabs(b);
}
template <typename T> __device__ __forceinline__ T dsign(T v) {
  return (T)((T(0) < v) - (v < T(0)));
}

// PLM_EnzoRKLimiter (EnzoReconstructorPLM.hpp:287-302)
template <typename T>
__device__ __forceinline__ T limiter_enzo(T vm1, T v, T vp1, T theta) {
  T dv_c = T(0.5) * (vp1 - vm1);
  T dv_l = (v - vm1) * theta;
  T dv_r = (vp1 - v) * theta;
  return (T(0.5) * (dsign(dv_l) + dsign(dv_r))) * dmin3_abs(dv_l, dv_r, dv_c);
}
// The same limiter in FP32 (the mixed path, judged by its L1 distance to
// the FP64 result, not bit for bit), with the sign factor taken from the
// sign bits: equal signs give copysign(m, dv_l), unequal ones 0, and where
// a difference is zero m is 0 as well.  The result equals the form above
// but for the sign of a zero and for how a NaN difference spreads; it
// takes about 12 instructions instead of 20 in the issue-bound FP32 flux
// kernel.
__device__ __forceinline__ float limiter_enzo(float vm1, float v, float vp1, float theta) {
  const float dv_c = 0.5f * (vp1 - vm1);
  const float dv_l = (v - vm1) * theta;
  const float dv_r = (vp1 - v) * theta;
  const float m = __builtin_fminf(__builtin_fminf(fabsf(dv_l), fabsf(dv_r)), fabsf(dv_c));
  const bool same = ((__float_as_uint(dv_l) ^ __float_as_uint(dv_r)) & 0x80000000u) == 0u;
  return same ? copysignf(m, dv_l) : 0.0f;
}

// PLM_AthenaLimiter (EnzoReconstructorPLM.hpp:337-345)
template <typename T>
__device__ __forceinline__ T limiter_athena(T vm1, T v, T vp1) {
  T dv_l = (v - vm1);
  T dv_r = (vp1 - v);
  T temp = dv_l * dv_r;
  if (temp <= T(0)) return T(0);
  return qdiv(T(2) * temp, dv_l + dv_r);
}

// EnzoEOSIdeal::fast_magnetosonic_speed<-1> (EnzoEOSIdeal.hpp:111-140);
// std::pow(x, 2) == x*x exactly.
template <typename T>
__device__ __forceinline__ T fast_ms_speed(T d, T p, T bi, T bj, T bk,
                                           T gamma) {
  const T B2 = ((bi * bi) + ((bj * bj) + (bk * bk)));
  const T cs2 = qdiv(gamma * p, d);
  const T inv_density = qdiv(T(1), d);
  const T va2 = B2 * inv_density;
  const T va2_cos2 = (bi * bi) * inv_density;
  const T sum = cs2 + va2;
  return qsqrt(T(0.5) * (va2 + cs2 + qsqrt(sum * sum - T(4) * cs2 * va2_cos2)));
}

template <typename T> struct Cons1D { T d, mx, my, mz, e, by, bz; };

// enzo_riemann_utils::passive_eint_flux (EnzoRiemannUtils.hpp:224-249),
// dual-energy formalism: the specific internal energy p / ((gamma-1) rho)
// of the side the density flux comes from, selected branch-free as the
// reference writes it, times the density flux.  Split in two: the specific
// energies (EnzoEOSIdeal::specific_eint, EnzoEOSIdeal.hpp:79-82) depend on
// the face states only, so the flux kernel evaluates them before the solve
// and keeps two values live through it instead of four.
template <typename T>
__device__ __forceinline__ T specific_eint(T density, T pressure, T gamma) {
  return qdiv(pressure, (gamma - T(1)) * density);
}
template <typename T>
__device__ __forceinline__ T upwind_eint_flux(T eint_l, T eint_r, T density_flux) {
  const T upwind = T(density_flux > T(0)) * eint_l + T(density_flux <= T(0)) * eint_r;
  return upwind * density_flux;
}

// HLLD (Miyoshi & Kusano 2005) following HLLDKernel::operator()
// (src/Enzo/hydro-mhd/riemann/EnzoRiemannHLLD.hpp:40-404).  In/out in the
// MHD LUT order; flux[L_BI] is 0.  vbar (dual energy, may be null): the
// interface velocity of :418-447 -- the left/right normal velocity outside
// the fan, else the contact speed scaled by the outer state's coefficient.
template <typename T>
__device__ __forceinline__ void hlld(const T wli[8], const T wri[8], T gamma,
                                     T flx[8], T *vbar = nullptr) {
  const T SMALL_NUMBER = T(1.0e-8);
  const T igm1 = qdiv(T(1), gamma - T(1));
  T spd[5];
  Cons1D<T> ul, ur, ulst, uldst, urdst, urst, fl, fr;
  // the outer states' physical fluxes are formed where the fan region is
  // selected: held from the start they would stay live through the whole
  // solve

  const T pressure_l = wli[L_E];
  const T pressure_r = wri[L_E];
  const T bxi = wli[L_BI];

  T bxsq = bxi * bxi;
  T pbl = T(0.5) * (bxsq + (wli[L_BJ] * wli[L_BJ] + wli[L_BK] * wli[L_BK]));
  T pbr = T(0.5) * (bxsq + (wri[L_BJ] * wri[L_BJ] + wri[L_BK] * wri[L_BK]));
  T kel = T(0.5) * wli[L_D] *
          (wli[L_VI] * wli[L_VI] +
           (wli[L_VJ] * wli[L_VJ] + wli[L_VK] * wli[L_VK]));
  T ker = T(0.5) * wri[L_D] *
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

  T cfl = fast_ms_speed(wli[L_D], pressure_l, wli[L_BI], wli[L_BJ], wli[L_BK],
                        gamma);
  T cfr = fast_ms_speed(wri[L_D], pressure_r, wri[L_BI], wri[L_BJ], wri[L_BK],
                        gamma);

  spd[0] = dmin(wli[L_VI] - cfl, wri[L_VI] - cfr);
  spd[4] = dmax(wli[L_VI] + cfl, wri[L_VI] + cfr);

  T ptl = pressure_l + pbl;
  T ptr = pressure_r + pbr;

  T sdl = spd[0] - wli[L_VI];
  T sdr = spd[4] - wri[L_VI];
  spd[2] = qdiv(sdr * ur.mx - sdl * ul.mx + (ptl - ptr),
                sdr * ur.d - sdl * ul.d);

  T sdml = spd[0] - spd[2];
  T sdmr = spd[4] - spd[2];
  if (vbar) {
    // as soon as the outer and contact speeds are known (fewer live
    // registers through the rest of the solve): l_coef = (S_l - vx_L) /
    // (S_l - S_M) = sdl / sdml, r_coef = sdr / sdmr
    const T l_coef = qdiv(sdl, sdml);
    const T r_coef = qdiv(sdr, sdmr);
    if (spd[0] > T(0)) {
      *vbar = wli[L_VI];
    } else if (spd[4] < T(0)) {
      *vbar = wri[L_VI];
    } else if (spd[2] >= T(0)) {
      *vbar = spd[2] * l_coef;
    } else {
      *vbar = spd[2] * r_coef;
    }
  }
  T sdml_inv = qdiv(T(1), sdml);
  T sdmr_inv = qdiv(T(1), sdmr);
  ulst.d = ul.d * sdl * sdml_inv;
  urst.d = ur.d * sdr * sdmr_inv;
  T ulst_d_inv = qdiv(T(1), ulst.d);
  T urst_d_inv = qdiv(T(1), urst.d);
  T sqrtdl = qsqrt(ulst.d);
  T sqrtdr = qsqrt(urst.d);

  spd[1] = spd[2] - qdiv(fabs(bxi), sqrtdl);
  spd[3] = spd[2] + qdiv(fabs(bxi), sqrtdr);

  T ptstl = ptl + ul.d * sdl * (spd[2] - wli[L_VI]);
  T ptstr = ptr + ur.d * sdr * (spd[2] - wri[L_VI]);
  T ptst = T(0.5) * (ptstr + ptstl);

  ulst.mx = ulst.d * spd[2];
  if (fabs(ul.d * sdl * sdml - bxsq) < (SMALL_NUMBER)*ptst) {
    ulst.my = ulst.d * wli[L_VJ];
    ulst.mz = ulst.d * wli[L_VK];
    ulst.by = ul.by;
    ulst.bz = ul.bz;
  } else {
    T tmp = qdiv(bxi * (sdl - sdml), ul.d * sdl * sdml - bxsq);
    ulst.my = ulst.d * (wli[L_VJ] - ul.by * tmp);
    ulst.mz = ulst.d * (wli[L_VK] - ul.bz * tmp);
    tmp = qdiv(ul.d * (sdl * sdl) - bxsq, ul.d * sdl * sdml - bxsq);
    ulst.by = ul.by * tmp;
    ulst.bz = ul.bz * tmp;
  }
  T vbstl =
      (ulst.mx * bxi + (ulst.my * ulst.by + ulst.mz * ulst.bz)) * ulst_d_inv;
  ulst.e = (sdl * ul.e - ptl * wli[L_VI] + ptst * spd[2] +
            bxi * (wli[L_VI] * bxi + (wli[L_VJ] * ul.by + wli[L_VK] * ul.bz) -
                   vbstl)) *
           sdml_inv;

  urst.mx = urst.d * spd[2];
  if (fabs(ur.d * sdr * sdmr - bxsq) < (SMALL_NUMBER)*ptst) {
    urst.my = urst.d * wri[L_VJ];
    urst.mz = urst.d * wri[L_VK];
    urst.by = ur.by;
    urst.bz = ur.bz;
  } else {
    T tmp = qdiv(bxi * (sdr - sdmr), ur.d * sdr * sdmr - bxsq);
    urst.my = urst.d * (wri[L_VJ] - ur.by * tmp);
    urst.mz = urst.d * (wri[L_VK] - ur.bz * tmp);
    tmp = qdiv(ur.d * (sdr * sdr) - bxsq, ur.d * sdr * sdmr - bxsq);
    urst.by = ur.by * tmp;
    urst.bz = ur.bz * tmp;
  }
  T vbstr =
      (urst.mx * bxi + (urst.my * urst.by + urst.mz * urst.bz)) * urst_d_inv;
  urst.e = (sdr * ur.e - ptr * wri[L_VI] + ptst * spd[2] +
            bxi * (wri[L_VI] * bxi + (wri[L_VJ] * ur.by + wri[L_VK] * ur.bz) -
                   vbstr)) *
           sdmr_inv;

  if (T(0.5) * bxsq < (SMALL_NUMBER)*ptst) {
    uldst = ulst;
    urdst = urst;
  } else {
    T invsumd = qdiv(T(1), sqrtdl + sqrtdr);
    T bxsig = (bxi > T(0) ? T(1) : T(-1));
    uldst.d = ulst.d;
    urdst.d = urst.d;
    uldst.mx = ulst.mx;
    urdst.mx = urst.mx;

    T tmp = invsumd * (sqrtdl * (ulst.my * ulst_d_inv) +
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

    tmp = spd[2] * bxi + qdiv(uldst.my * uldst.by + uldst.mz * uldst.bz, uldst.d);
    uldst.e = ulst.e - sqrtdl * bxsig * (vbstl - tmp);
    urdst.e = urst.e + sqrtdr * bxsig * (vbstr - tmp);
  }

  auto flux_l = [&]() {
    fl.d = ul.mx;
    fl.mx = ul.mx * wli[L_VI] + ptl - bxsq;
    fl.my = ul.my * wli[L_VI] - bxi * ul.by;
    fl.mz = ul.mz * wli[L_VI] - bxi * ul.bz;
    fl.e = wli[L_VI] * (ul.e + ptl - bxsq) -
           bxi * (wli[L_VJ] * ul.by + wli[L_VK] * ul.bz);
    fl.by = ul.by * wli[L_VI] - bxi * wli[L_VJ];
    fl.bz = ul.bz * wli[L_VI] - bxi * wli[L_VK];
  };
  auto flux_r = [&]() {
    fr.d = ur.mx;
    fr.mx = ur.mx * wri[L_VI] + ptr - bxsq;
    fr.my = ur.my * wri[L_VI] - bxi * ur.by;
    fr.mz = ur.mz * wri[L_VI] - bxi * ur.bz;
    fr.e = wri[L_VI] * (ur.e + ptr - bxsq) -
           bxi * (wri[L_VJ] * ur.by + wri[L_VK] * ur.bz);
    fr.by = ur.by * wri[L_VI] - bxi * wri[L_VJ];
    fr.bz = ur.bz * wri[L_VI] - bxi * wri[L_VK];
  };
  // Select the region of the fan containing the interface; the state
  // differences of each wave are formed lazily (only the ones used).
  if (spd[0] >= T(0)) {
    flux_l();
    flx[L_D] = fl.d; flx[L_VI] = fl.mx; flx[L_VJ] = fl.my; flx[L_VK] = fl.mz;
    flx[L_E] = fl.e; flx[L_BJ] = fl.by; flx[L_BK] = fl.bz;
  } else if (spd[4] <= T(0)) {
    flux_r();
    flx[L_D] = fr.d; flx[L_VI] = fr.mx; flx[L_VJ] = fr.my; flx[L_VK] = fr.mz;
    flx[L_E] = fr.e; flx[L_BJ] = fr.by; flx[L_BK] = fr.bz;
  } else if (spd[1] >= T(0)) {
    flux_l();
    flx[L_D] = fl.d + spd[0] * (ulst.d - ul.d);
    flx[L_VI] = fl.mx + spd[0] * (ulst.mx - ul.mx);
    flx[L_VJ] = fl.my + spd[0] * (ulst.my - ul.my);
    flx[L_VK] = fl.mz + spd[0] * (ulst.mz - ul.mz);
    flx[L_E] = fl.e + spd[0] * (ulst.e - ul.e);
    flx[L_BJ] = fl.by + spd[0] * (ulst.by - ul.by);
    flx[L_BK] = fl.bz + spd[0] * (ulst.bz - ul.bz);
  } else if (spd[2] >= T(0)) {
    flux_l();
    flx[L_D] = fl.d + spd[0] * (ulst.d - ul.d) + spd[1] * (uldst.d - ulst.d);
    flx[L_VI] = fl.mx + spd[0] * (ulst.mx - ul.mx) + spd[1] * (uldst.mx - ulst.mx);
    flx[L_VJ] = fl.my + spd[0] * (ulst.my - ul.my) + spd[1] * (uldst.my - ulst.my);
    flx[L_VK] = fl.mz + spd[0] * (ulst.mz - ul.mz) + spd[1] * (uldst.mz - ulst.mz);
    flx[L_E] = fl.e + spd[0] * (ulst.e - ul.e) + spd[1] * (uldst.e - ulst.e);
    flx[L_BJ] = fl.by + spd[0] * (ulst.by - ul.by) + spd[1] * (uldst.by - ulst.by);
    flx[L_BK] = fl.bz + spd[0] * (ulst.bz - ul.bz) + spd[1] * (uldst.bz - ulst.bz);
  } else if (spd[3] > T(0)) {
    flux_r();
    flx[L_D] = fr.d + spd[4] * (urst.d - ur.d) + spd[3] * (urdst.d - urst.d);
    flx[L_VI] = fr.mx + spd[4] * (urst.mx - ur.mx) + spd[3] * (urdst.mx - urst.mx);
    flx[L_VJ] = fr.my + spd[4] * (urst.my - ur.my) + spd[3] * (urdst.my - urst.my);
    flx[L_VK] = fr.mz + spd[4] * (urst.mz - ur.mz) + spd[3] * (urdst.mz - urst.mz);
    flx[L_E] = fr.e + spd[4] * (urst.e - ur.e) + spd[3] * (urdst.e - urst.e);
    flx[L_BJ] = fr.by + spd[4] * (urst.by - ur.by) + spd[3] * (urdst.by - urst.by);
    flx[L_BK] = fr.bz + spd[4] * (urst.bz - ur.bz) + spd[3] * (urdst.bz - urst.bz);
  } else {
    flux_r();
    flx[L_D] = fr.d + spd[4] * (urst.d - ur.d);
    flx[L_VI] = fr.mx + spd[4] * (urst.mx - ur.mx);
    flx[L_VJ] = fr.my + spd[4] * (urst.my - ur.my);
    flx[L_VK] = fr.mz + spd[4] * (urst.mz - ur.mz);
    flx[L_E] = fr.e + spd[4] * (urst.e - ur.e);
    flx[L_BJ] = fr.by + spd[4] * (urst.by - ur.by);
    flx[L_BK] = fr.bz + spd[4] * (urst.bz - ur.bz);
  }
  flx[L_BI] = T(0);
}

// conserved state + physical flux, MHD LUT (EnzoRiemannUtils.hpp:48-149)
template <typename T>
__device__ __forceinline__ void mhd_cons_flux(const T w[8], T gamma, T cons[8],
                                              T flux[8]) {
  for (int i = 0; i < 4; i++) cons[i] = w[i];
  for (int i = 4; i < 8; i++) cons[i] = w[i] * w[L_D];
  {
    T density = w[L_D];
    T internal_edens = qdiv(w[L_E], gamma - T(1));
    const T vi = w[L_VI], vj = w[L_VJ], vk = w[L_VK];
    const T kinetic = T(0.5) * density * ((vi * vi) + ((vj * vj) + (vk * vk)));
    T magnetic = T(0.5) * ((w[L_BI] * w[L_BI]) +
                           ((w[L_BJ] * w[L_BJ]) + (w[L_BK] * w[L_BK])));
    cons[L_E] = internal_edens + kinetic + magnetic;
  }
  const T vi = w[L_VI], vj = w[L_VJ], vk = w[L_VK];
  const T Bi = w[L_BI], Bj = w[L_BJ], Bk = w[L_BK];
  const T etot = cons[L_E];
  const T ptot = w[L_E] + T(0.5) * ((Bi * Bi) + ((Bj * Bj) + (Bk * Bk)));
  const T mom_i = cons[L_VI];
  flux[L_D] = mom_i;
  flux[L_VI] = mom_i * vi - Bi * Bi + ptot;
  flux[L_VJ] = mom_i * vj - Bj * Bi;
  flux[L_VK] = mom_i * vk - Bk * Bi;
  flux[L_E] = ((etot + ptot) * vi - (Bi * vi + (Bj * vj + Bk * vk)) * Bi);
  flux[L_BI] = T(0);
  flux[L_BJ] = Bj * vi - Bi * vj;
  flux[L_BK] = Bk * vi - Bi * vk;
}

// HLLE for MHD: HLLKernel<EinfeldtWavespeed<MHDLUT>>
// (EnzoRiemannHLL.hpp:44-172, 238-320); vbar (dual energy, may be null):
// the normal velocity interpolated between the outer speeds (:330-340)
template <typename T>
__device__ __forceinline__ void hlle_mhd(const T wl[8], const T wr[8],
                                         T gamma, T flux[8],
                                         T *vbar = nullptr) {
  T cl[8], cr[8], fl[8], fr[8];
  mhd_cons_flux(wl, gamma, cl, fl);
  mhd_cons_flux(wr, gamma, cr, fr);
  T c_l = fast_ms_speed(wl[L_D], wl[L_E], wl[L_BI], wl[L_BJ], wl[L_BK], gamma);
  T c_r = fast_ms_speed(wr[L_D], wr[L_E], wr[L_BI], wr[L_BJ], wr[L_BK], gamma);
  T left_speed = (wl[L_VI] - c_l);
  T right_speed = (wr[L_VI] + c_r);
  T sqrtrho_l = qsqrt(wl[L_D]);
  T sqrtrho_r = qsqrt(wr[L_D]);
  T inv_sqrtrho_tot = qdiv(T(1), sqrtrho_l + sqrtrho_r);
  T vi_roe = (sqrtrho_l * wl[L_VI] + sqrtrho_r * wr[L_VI]) * inv_sqrtrho_tot;
  T vj_roe = (sqrtrho_l * wl[L_VJ] + sqrtrho_r * wr[L_VJ]) * inv_sqrtrho_tot;
  T vk_roe = (sqrtrho_l * wl[L_VK] + sqrtrho_r * wr[L_VK]) * inv_sqrtrho_tot;
  T v_roe2 = vi_roe * vi_roe + vj_roe * vj_roe + vk_roe * vk_roe;
  T ptot_l = wl[L_E];
  T ptot_r = wr[L_E];
  ptot_l += T(0.5) * ((wl[L_BI] * wl[L_BI]) +
                      ((wl[L_BJ] * wl[L_BJ]) + (wl[L_BK] * wl[L_BK])));
  ptot_r += T(0.5) * ((wr[L_BI] * wr[L_BI]) +
                      ((wr[L_BJ] * wr[L_BJ]) + (wr[L_BK] * wr[L_BK])));
  T h_l = qdiv(cl[L_E] + ptot_l, wl[L_D]);
  T h_r = qdiv(cr[L_E] + ptot_r, wr[L_D]);
  T h_roe = (sqrtrho_l * h_l + sqrtrho_r * h_r) * inv_sqrtrho_tot;
  // roe_cfast_
  T rho_roe = sqrtrho_l * sqrtrho_r;
  T bi_roe = wl[L_BI];
  T bj_roe = (sqrtrho_l * wr[L_BJ] + sqrtrho_r * wl[L_BJ]) * inv_sqrtrho_tot;
  T bk_roe = (sqrtrho_l * wr[L_BK] + sqrtrho_r * wl[L_BK]) * inv_sqrtrho_tot;
  T b_roe2 = bi_roe * bi_roe + bj_roe * bj_roe + bk_roe * bk_roe;
  T gamma_prime = gamma - T(1);
  T dbj = wl[L_BJ] - wr[L_BJ];
  T dbk = wl[L_BK] - wr[L_BK];
  T x_prime =
      ((dbj * dbj + dbk * dbk) * T(0.5) * (gamma_prime - T(1)) * inv_sqrtrho_tot);
  T y_prime = qdiv((gamma_prime - T(1)) * (wl[L_D] + wr[L_D]) * T(0.5), rho_roe);
  T tilde_a2 =
      (gamma_prime * (h_roe - T(0.5) * v_roe2 - qdiv(b_roe2, rho_roe)) - x_prime);
  T tilde_vai2 = qdiv(bi_roe * bi_roe, rho_roe);
  T tilde_va2 = (tilde_vai2 + qdiv((gamma_prime - y_prime) *
                                       (bj_roe * bj_roe + bk_roe * bk_roe),
                                   rho_roe));
  T s = tilde_a2 + tilde_va2;
  T c_roe = qsqrt(T(0.5) * (tilde_a2 + tilde_va2 +
                            qsqrt(s * s - T(4) * tilde_a2 * tilde_vai2)));
  T bp = fmax(vi_roe + c_roe, right_speed);
  T bm = fmin(vi_roe - c_roe, left_speed);
  bp = fmax(bp, T(0));
  bm = fmin(bm, T(0));
  T inv_speed_diff = qdiv(T(1), bp - bm);
#pragma unroll
  for (int f = 0; f < 8; f++) {
    flux[f] = ((bp * fl[f] - bm * fr[f] + (cr[f] - cl[f]) * bp * bm) *
               inv_speed_diff);
  }
  if (vbar) *vbar = (bp * wl[L_VI] - bm * wr[L_VI]) * inv_speed_diff;
}

// HLLC (HLLCKernel::operator(), EnzoRiemannHLLC.hpp:34-172) with
// EinfeldtWavespeed<HydroLUT> (EnzoRiemannHLL.hpp:44-130).  Hydro LUT.
// vbar (dual energy, may be null): the weighted outer velocities of :166-170
template <typename T>
__device__ __forceinline__ void hllc(const T wl[5], const T wr[5], T gamma,
                                     T flux[5], T *vbar = nullptr) {
  const T pressure_l = wl[H_E];
  const T pressure_r = wr[H_E];
  T cons_l[5], cons_r[5];
  {
    const T *w[2] = {wl, wr};
    T *c[2] = {cons_l, cons_r};
#pragma unroll
    for (int s = 0; s < 2; s++) {
      c[s][H_D] = w[s][H_D];
      for (int i = 1; i < 5; i++) c[s][i] = w[s][i] * w[s][H_D];
      T density = w[s][H_D];
      T internal_edens = qdiv(w[s][H_E], gamma - T(1));
      const T vi = w[s][H_VI], vj = w[s][H_VJ], vk = w[s][H_VK];
      const T kinetic =
          T(0.5) * density * ((vi * vi) + ((vj * vj) + (vk * vk)));
      c[s][H_E] = internal_edens + kinetic + T(0);
    }
  }
  T cs_l, cs_r;
  {
    T c_l = qsqrt(qdiv(gamma * pressure_l, wl[H_D]));
    T c_r = qsqrt(qdiv(gamma * pressure_r, wr[H_D]));
    T left_speed = (wl[H_VI] - c_l);
    T right_speed = (wr[H_VI] + c_r);
    T sqrtrho_l = qsqrt(wl[H_D]);
    T sqrtrho_r = qsqrt(wr[H_D]);
    T inv_sqrtrho_tot = qdiv(T(1), sqrtrho_l + sqrtrho_r);
    T vi_roe = (sqrtrho_l * wl[H_VI] + sqrtrho_r * wr[H_VI]) * inv_sqrtrho_tot;
    T vj_roe = (sqrtrho_l * wl[H_VJ] + sqrtrho_r * wr[H_VJ]) * inv_sqrtrho_tot;
    T vk_roe = (sqrtrho_l * wl[H_VK] + sqrtrho_r * wr[H_VK]) * inv_sqrtrho_tot;
    T v_roe2 = vi_roe * vi_roe + vj_roe * vj_roe + vk_roe * vk_roe;
    T h_l = qdiv(cons_l[H_E] + pressure_l, wl[H_D]);
    T h_r = qdiv(cons_r[H_E] + pressure_r, wr[H_D]);
    T h_roe = (sqrtrho_l * h_l + sqrtrho_r * h_r) * inv_sqrtrho_tot;
    T temp = h_roe - T(0.5) * v_roe2;
    T c_roe = qsqrt((gamma - T(1)) * dmax(temp, T(0)));
    cs_r = fmax(vi_roe + c_roe, right_speed);
    cs_l = fmin(vi_roe - c_roe, left_speed);
  }
  T bm = fmin(cs_l, T(0));
  T bp = fmax(cs_r, T(0));
  T tl = (pressure_l - (cs_l - wl[H_VI]) * wl[H_D] * wl[H_VI]);
  T tr = (pressure_r - (cs_r - wr[H_VI]) * wr[H_D] * wr[H_VI]);
  T dl = wl[H_D] * (cs_l - wl[H_VI]);
  T dr = -wr[H_D] * (cs_r - wr[H_VI]);
  T q1 = qdiv(T(1), dl + dr);
  T cw = (tr - tl) * q1;
  T cp = (dl * tr + dr * tl) * q1;
  T sl, sr, sm;
  if (cw >= T(0)) {
    sl = qdiv(cw, cw - bm);
    sr = T(0);
    sm = qdiv(-bm, cw - bm);
  } else {
    sl = T(0);
    sr = qdiv(-cw, bp - cw);
    sm = qdiv(bp, bp - cw);
  }
  if (vbar) *vbar = (sl * (wl[H_VI] - bm) + sr * (wr[H_VI] - bp));
  cp = dmax(cp, T(0));
  T momentumi_l = cons_l[H_VI];
  T momentumi_r = cons_r[H_VI];
  T dfl = momentumi_l - bm * wl[H_D];
  T dfr = momentumi_r - bp * wr[H_D];
  T ufl = momentumi_l * (wl[H_VI] - bm) + pressure_l;
  T ufr = momentumi_r * (wr[H_VI] - bp) + pressure_r;
  T vfl = (wl[H_D] * wl[H_VJ] * (wl[H_VI] - bm));
  T vfr = (wr[H_D] * wr[H_VJ] * (wr[H_VI] - bp));
  T wfl = (wl[H_D] * wl[H_VK] * (wl[H_VI] - bm));
  T wfr = (wr[H_D] * wr[H_VK] * (wr[H_VI] - bp));
  T efl = (cons_l[H_E] * (wl[H_VI] - bm) + pressure_l * wl[H_VI]);
  T efr = (cons_r[H_E] * (wr[H_VI] - bp) + pressure_r * wr[H_VI]);
  flux[H_D] = sl * dfl + sr * dfr;
  flux[H_VI] = sl * ufl + sr * ufr + (sm * cp);
  flux[H_VJ] = sl * vfl + sr * vfr;
  flux[H_VK] = sl * wfl + sr * wfr;
  flux[H_E] = sl * efl + sr * efr + (sm * cp * cw);
}

}  // namespace ENZOE_KNS
