// vlct_kernels.hpp -- gfx950 kernels of one VL+CT stage, the CFL reduction
// and the ghost refresh, launched over a whole list of blocks at once.
//
// A stage (EnzoMHDIntegratorStageCommands::compute_update_stage,
// src/Enzo/hydro-mhd/EnzoMHDIntegratorStageCommands.cpp:102-203) is four
// launches, each covering every block of the batch (grid.y = block):
//   k_flux  : primitives -> reconstruction -> Riemann flux on the three low
//             faces of every cell (one HLLD/HLLE/HLLC solve per face)
//   k_edge  : upwinded edge EMFs (Gardiner & Stone 2009) of constrained
//             transport, the upwind weights taken from the sign of the
//             density flux
//   k_faces : face-centred B update from the edge EMFs
//   k_cells : cell-centred B from faces, flux divergence, conserved update,
//             density and energy floors
// All index regions follow the reference's stale-depth bookkeeping; scratch
// arrays are cell-shaped (a face or edge value is stored at the index of
// its low cell).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "vlct_physics.hpp"

namespace ENZOE_KNS {

// Device descriptor of one block: permanent fields + the scratch the stage
// kernels use for it.  FluxT is double (FP64 path) or float (mixed path).
template <typename FluxT>
struct BlockDesc {
  double *U[S_N];      // density, velocity_x..z, total_energy, bfield_x..z
  double *bfi[3];      // bfieldi_x..z (start of step / final output)
  double *pressure;    // permanent pressure field (may be null)
  double *Ut[S_N];     // half-step state (temp integration map)
  float *Up[S_N];      // mixed path: the half-step state as FP32 primitives
                       // (rho, v, p, B; replaces Ut there)
  double *bft[3];      // half-step face B
  FluxT *F[3][S_N];    // fluxes along x, y, z; slot = field
  double *E[3];        // edge EMFs along x, y, z
  double *Ec[3];       // cell-centred E = -v x B of the stage's state
  double *Ech[3];      // mixed path: centre E of the half-step state (the
                       // first stage's cell update writes it), or null
  double *bf0[3];      // fused CT update: start-of-step face B of the first
                       // stage's update region (the final stage reads its
                       // base values here), or null
  const double *acc[3];  // acceleration_x..z for the gravity source, or null
  // dual-energy formalism (null otherwise): specific internal_energy (start
  // of step / final output), its half-step copy, and per dim the
  // internal-energy flux and the interface velocity vi_bar of every face
  double *ei;
  double *eit;
  double *Fe[3];
  double *Vb[3];
  int mx, my, mz;
  int g;               // ghost depth
  int tile_rows;       // rows of the tiled cell order (0: plain order)
  double cw[3];
};

struct StageParams {
  int stage;           // 0 or 1
  int is_final;        // writes into the permanent fields
  int stale;           // stale depth at the start of the stage
  int cur_stale;       // stale + immediate staling of the reconstructor
  double dt;           // timestep of this stage (dt/2 for the VL half step)
  double gamma;
  double density_floor;
  double pressure_floor;
  double theta;
  const double *dts;   // per-block dt of the step (device) or null: sp.dt
  double dt_factor;    // 0.5 for the partial step, 1 for the final one
  int xcd_remap;       // bit mask over k_flux(1) k_edge(2) k_faces(4)
                       // k_cells(8): XCD-contiguous order, see wg_coords,
                       // with the tiled cell order of cell_coords
  double eta;          // dual-energy eta ("modern" formulation)
  double ggm1;         // gamma (gamma - 1) rounded to float, as the
                       // reference's synchronisation declares it
  unsigned long long *cfl;  // final k_cells: also the CFL minimum of its
                            // cells and their pressure (null: not)
  unsigned long long *dbg;  // debug counters (DBG_*), or null: checks off
};

// Debug checks, off unless a method enables them: counts of the events the
// reference's debug builds abort on -- a non-finite value
// (CHECK_FINITE_ELEMENTS, src/Cello/view_CelloView.hpp:207-217) and a floor
// that has to be applied (RAISE_FLOOR_ERROR in enzo_utils::apply_floor,
// src/Enzo/utils/utils.hpp:102-114; its callers: the PLM reconstruction,
// EnzoReconstructorPLM.hpp:219-220, the density update,
// EnzoIntegrationQuanUpdate.cpp:219, the energy floor,
// EnzoPhysicsFluidProps.cpp:278-285).
enum : int {
  DBG_NONFINITE = 0,       // a non-finite updated cell value
  DBG_DENSITY_FLOOR = 1,   // updated density below the density floor
  DBG_ENERGY_FLOOR = 2,    // energy below its floor after the update
  DBG_RECON_FLOOR = 3,     // a reconstructed density or pressure floored
  DBG_N = 4
};

// Adds the number of lanes of the wave with `cond` to dbg[slot]: one atomic
// per wave, from its lowest lane with the event.
__device__ __forceinline__ void debug_count(unsigned long long *dbg, int slot, bool cond) {
  const unsigned long long mask = __ballot(cond);
  if (mask != 0ull && (int)(threadIdx.x & 63) == __ffsll((long long)mask) - 1)
    atomicAdd(&dbg[slot], (unsigned long long)__popcll(mask));
}

// Energy floor and, with the dual-energy formalism, the synchronisation of
// internal and total energy of one cell
// (EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync,
// EnzoPhysicsFluidProps.cpp:158-290).  rho, v, b2 are the cell's (updated)
// density, velocity and |B|^2 (0 for hydro); etot and eint are updated in
// place.  The internal energy is taken from etot - non-thermal energy where
// the sound speed that implies beats eta times the kinetic and magnetic
// scales (and it exceeds half the tracked value unless eta = 0).
// floored (may be null): set when the floor had to be applied.
template <bool MHD, bool DE>
__device__ __forceinline__ void floor_and_sync(double rho, const double v[3],
                                               double b2, double gamma,
                                               double pfloor, double eta,
                                               double ggm1, double &etot,
                                               double &eint,
                                               bool *floored = nullptr) {
  const double inv_rho = 1. / rho;
  const double eint_floor = pfloor * (1. / (gamma - 1.)) * inv_rho;
  const double v2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  double non_thermal_e = 0.5 * v2;
  if (MHD) non_thermal_e += (0.5 * b2 * inv_rho);
  if (DE) {
    const double half_factor = (eta != 0.) ? 0.5 : 0.;
    const double eint_1 = etot - non_thermal_e;
    double cur_eint = eint;
    const double cs2_1 = fmax(0., ggm1 * eint_1);
    if ((cs2_1 > fmax(eta * v2, eta * (MHD ? b2 : 0.) * inv_rho)) &&
        (eint_1 > half_factor * cur_eint))
      cur_eint = eint_1;
    if (floored) *floored = cur_eint < eint_floor;
    cur_eint = dmax(cur_eint, eint_floor);
    eint = cur_eint;
    etot = cur_eint + non_thermal_e;
  } else {
    const double etot_floor = eint_floor + non_thermal_e;
    if (floored) *floored = etot < etot_floor;
    etot = dmax(etot, etot_floor);
  }
}

__device__ __forceinline__ long idx3(int k, int j, int i, int ny, int nx) {
  return ((long)k * ny + j) * (long)nx + i;
}


// Workgroup coordinates of a (gx, gy) launch.  The hardware hands the
// workgroups of a launch to the 8 XCDs round-robin in linear order, and each
// XCD has its own L2; with xcd_remap the linear range [x T/8, (x+1) T/8)
// goes to XCD x, so the neighbouring rows a stencil re-reads stay in one L2.
// Needs T = gx * gy to be a multiple of 8: the launches pad gx to a multiple
// of 8 and padding workgroups fall outside every box.
__device__ __forceinline__ void wg_coords(int xcd_remap, int &bx, int &by) {
  if (!xcd_remap) {
    bx = blockIdx.x;
    by = blockIdx.y;
    return;
  }
  const unsigned gx = gridDim.x, T = gx * gridDim.y;
  const unsigned L = blockIdx.y * gx + blockIdx.x;
  const unsigned lg = (L & 7u) * (T >> 3) + (L >> 3);
  by = (int)(lg / gx);
  bx = (int)(lg - (unsigned)by * gx);
}

// A zero-length step (a device-driven cycle past its stop time) changes
// nothing, so the stage kernels return at once.
__device__ __forceinline__ bool zero_step(const StageParams &sp, int by) {
  return sp.dts != nullptr && sp.dts[by] == 0.0;
}

// Decodes a linear thread index into (k,j,i) inside the box [lo, lo+n),
// x fastest (32-bit arithmetic when the box allows it).
__device__ __forceinline__ bool box_coords(long t, int lo, int nx, int ny,
                                           int nz, int &k, int &j, int &i) {
  long nxy = (long)nx * ny;
  if (t >= nxy * nz) return false;
  if (nxy * nz < (1L << 31)) {
    const unsigned tt = (unsigned)t, uxy = (unsigned)nxy, ux = (unsigned)nx;
    const unsigned kk = tt / uxy;
    const unsigned r = tt - kk * uxy;
    const unsigned jj = r / ux;
    k = (int)kk + lo;
    j = (int)jj + lo;
    i = (int)(r - jj * ux) + lo;
    return true;
  }
  k = (int)(t / nxy);
  long r = t - (long)k * nxy;
  j = (int)(r / nx);
  i = (int)(r - (long)j * nx);
  k += lo; j += lo; i += lo;
  return true;
}

// The tiled cell order: the box is cut into slabs of `rows` rows (all of
// x), walked slab by slab and, inside a slab, plane by plane, so that
// consecutive workgroups -- one XCD's share under the XCD-contiguous order --
// sweep a narrow slab through z and the planes a z stencil reuses are still
// in that XCD's L2 (in the plain order a plane of a large block is larger
// than the L2, and every z neighbour comes from HBM again).  The last slab
// may be shorter.  32-bit indices: blocks hold fewer than 2^31 cells
// (check_block).
__device__ __forceinline__ bool box_coords_tiled(long t, int lo, int nx, int ny,
                                                 int nz, int rows, int &k,
                                                 int &j, int &i) {
  const unsigned ux = (unsigned)nx;
  const unsigned tt = (unsigned)t;
  if (t >= (long)nx * ny * nz) return false;
  const unsigned full = ux * (unsigned)rows * (unsigned)nz;  // one full slab
  const unsigned slab = tt / full;
  const unsigned y0 = slab * (unsigned)rows;
  const unsigned r = min((unsigned)rows, (unsigned)ny - y0);
  unsigned u = tt - slab * full;
  const unsigned tp = ux * r;  // cells of one plane of this slab
  const unsigned kk = u / tp;
  u -= kk * tp;
  const unsigned jj = u / ux;
  k = (int)kk + lo;
  j = (int)(jj + y0) + lo;
  i = (int)(u - jj * ux) + lo;
  return true;
}

// The cell of linear index t in the box [lo, lo+n): tiled order when
// rows > 0 and the box has more rows than one slab, else x fastest.
__device__ __forceinline__ bool cell_coords(long t, int lo, int nx, int ny, int nz,
                                            int rows, int &k, int &j, int &i) {
  if (rows > 0 && rows < ny) return box_coords_tiled(t, lo, nx, ny, nz, rows, k, j, i);
  return box_coords(t, lo, nx, ny, nz, k, j, i);
}

// Loads through the global address space (the pointers come from a
// descriptor table in memory, so the compiler cannot infer it by itself and
// would otherwise emit flat accesses).
typedef const __attribute__((address_space(1))) double *gdp;
template <typename P>
__device__ __forceinline__ P gld(const P *p, int i) {
  return ((const __attribute__((address_space(1))) P *)p)[i];
}

// Stores through the global address space (see gld).
template <typename P>
__device__ __forceinline__ void gst(P *p, long i, P v) {
  ((__attribute__((address_space(1))) P *)p)[i] = v;
}

// Cell primitives w = (rho, vx, vy, vz, p, bx, by, bz) from the integration
// quantities (EnzoPhysicsFluidProps.cpp:64-138, EnzoComputePressure.cpp:
// 127-189), in FP64; with dual energy the pressure is gm1 rho eint.
// The primitive pressure gm1 (rho (e - ke) - me) of specific total energy
// e (EnzoComputePressure.cpp:177-189).  Never contracted, so that the flux
// kernel's and the mixed path's half-step store round alike.
template <bool MHD>
__device__ __forceinline__ double prim_pressure(double rho, double vx, double vy, double vz,
                                                double e, double bx, double by, double bz,
                                                double gm1) {
#pragma clang fp contract(off)
  const double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
  const double me = MHD ? 0.5 * (bx * bx + by * by + bz * bz) : 0.;
  return gm1 * (rho * (e - ke) - me);
}

template <bool MHD, bool DE = false>
__device__ __forceinline__ void load_prim(double *const U[S_N], int c,
                                          double gm1, double w[S_N],
                                          const double *ei = nullptr) {
  double rho = gld(U[S_RHO], c);
  double vx = gld(U[S_VX], c), vy = gld(U[S_VY], c), vz = gld(U[S_VZ], c);
  double bx = 0., by = 0., bz = 0.;
  w[S_RHO] = rho; w[S_VX] = vx; w[S_VY] = vy; w[S_VZ] = vz;
  if (MHD) {
    bx = gld(U[S_BX], c); by = gld(U[S_BY], c); bz = gld(U[S_BZ], c);
  }
  w[S_BX] = bx; w[S_BY] = by; w[S_BZ] = bz;
  if (DE)
    w[S_E] = gm1 * rho * gld(ei, c);
  else
    w[S_E] = prim_pressure<MHD>(rho, vx, vy, vz, gld(U[S_E], c), bx, by, bz, gm1);
}

// Primitives of one cell in the flux kernel's arithmetic type T: from the
// FP64 integration quantities (pressure computed here), or, with PRIMF,
// loaded from the FP32 primitive half-step state that the previous stage's
// cell update stored (the same FP64 values rounded to FP32 once).
template <typename T, bool MHD, bool DE, bool PRIMF>
struct PrimSource {
  double *const *U;
  const double *ei;
  const float *const *Up;
  double gm1;
  __device__ __forceinline__ void operator()(int c, T w[S_N]) const {
    if constexpr (PRIMF) {
#pragma unroll
      for (int q = 0; q < S_N; q++) w[q] = (MHD || q < S_BX) ? (T)gld(Up[q], c) : T(0);
    } else {
      double d[S_N];
      load_prim<MHD, DE>(U, c, gm1, d, ei);
#pragma unroll
      for (int q = 0; q < S_N; q++) w[q] = (T)d[q];
    }
  }
};

// Centre E_i = -v_j B_k + v_k B_j (EnzoBfieldMethodCT::compute_center_efield,
// EnzoBfieldMethodCT.cpp:267-292); never contracted, so that the flux
// kernel and the mixed path's half-step store give the same bits.
__device__ __forceinline__ double center_efield(double vj, double vk, double bj, double bk) {
#pragma clang fp contract(off)
  return (-vj * bk + vk * bj);
}

// ---------------------------------------------------------------- k_flux

struct FluxConsts {
  double gm1, gamma, theta, dfloor, pfloor;
  unsigned long long *dbg;  // debug counters or null
};

// Flux through the face between cell c and c + e_D (D fixed at compile
// time so the i/j/k permutation costs nothing): reconstruction of both
// sides, the face-centred normal B (EnzoBfieldMethodCT.cpp:122-166) and the
// Riemann solve, written at index c of the flux arrays of dim D.
template <typename T, typename FluxT, int RECON, int RIEMANN, bool MHD, int D,
          bool DE, bool PRIMF, class Src = PrimSource<T, MHD, DE, PRIMF>>
__device__ __forceinline__ void flux_face(const BlockDesc<FluxT> &b, const Src &src,
                                          const double *bfi, int k, int j,
                                          int i, int c, const T wc[S_N],
                                          const FluxConsts &fc) {
  constexpr int DJ = (D + 1) % 3, DK = (D + 2) % 3;
  const int s = (D == 0) ? 1 : ((D == 1) ? b.mx : b.mx * b.my);
  T wl[S_N], wr[S_N];
  if (RECON == RC_NN) {
    T a[S_N];
    src(c + s, a);
#pragma unroll
    for (int q = 0; q < S_N; q++) { wl[q] = wc[q]; wr[q] = a[q]; }
  } else {
    T w0[S_N], w2[S_N], w3[S_N];  // cells c-1, c+1, c+2 along D
    src(c - s, w0);
    src(c + s, w2);
    src(c + 2 * s, w3);
    const T theta = (T)fc.theta;
#pragma unroll
    for (int q = 0; q < S_N; q++) {
      if (!MHD && q >= S_BX) continue;
      if (MHD && q == S_BX + D) continue;  // replaced by the face value
      T a0 = w0[q], a1 = wc[q], a2 = w2[q], a3 = w3[q];
      T dvl = (RECON == RC_PLM) ? limiter_enzo(a0, a1, a2, theta)
                                : limiter_athena(a0, a1, a2);
      T dvr = (RECON == RC_PLM) ? limiter_enzo(a1, a2, a3, theta)
                                : limiter_athena(a1, a2, a3);
      T lv = a1 + dvl * T(0.5);
      T rv = a2 - dvr * T(0.5);
      if (q == S_RHO) {
        if (fc.dbg)
          debug_count(fc.dbg, DBG_RECON_FLOOR, lv < (T)fc.dfloor || rv < (T)fc.dfloor);
        lv = dmax(lv, (T)fc.dfloor);
        rv = dmax(rv, (T)fc.dfloor);
      }
      if (q == S_E) {
        if (fc.dbg)
          debug_count(fc.dbg, DBG_RECON_FLOOR, lv < (T)fc.pfloor || rv < (T)fc.pfloor);
        lv = dmax(lv, (T)fc.pfloor);
        rv = dmax(rv, (T)fc.pfloor);
      }
      wl[q] = lv;
      wr[q] = rv;
    }
  }
  const T gamma = (T)fc.gamma;
  T fo[8];
  T vbar = T(0);
  T *vb = DE ? &vbar : nullptr;
  T eint_l = T(0), eint_r = T(0);
  if (DE) {
    eint_l = specific_eint(wl[S_RHO], wl[S_E], gamma);
    eint_r = specific_eint(wr[S_RHO], wr[S_E], gamma);
  }
  if (MHD) {
    const int bnx = b.mx + (D == 0), bny = b.my + (D == 1);
    const T bface =
        (T)gld(bfi, (k + (D == 2)) * bny * bnx + (j + (D == 1)) * bnx + i + (D == 0));
    T l8[8], r8[8];
    l8[L_D] = wl[S_RHO]; r8[L_D] = wr[S_RHO];
    l8[L_BI] = bface; r8[L_BI] = bface;
    l8[L_BJ] = wl[S_BX + DJ]; r8[L_BJ] = wr[S_BX + DJ];
    l8[L_BK] = wl[S_BX + DK]; r8[L_BK] = wr[S_BX + DK];
    l8[L_VI] = wl[S_VX + D]; r8[L_VI] = wr[S_VX + D];
    l8[L_VJ] = wl[S_VX + DJ]; r8[L_VJ] = wr[S_VX + DJ];
    l8[L_VK] = wl[S_VX + DK]; r8[L_VK] = wr[S_VX + DK];
    l8[L_E] = wl[S_E]; r8[L_E] = wr[S_E];
    if (RIEMANN == RS_HLLD) {
      hlld(l8, r8, gamma, fo, vb);
    } else {
      hlle_mhd(l8, r8, gamma, fo, vb);
    }
    gst(b.F[D][S_RHO], c, (FluxT)fo[L_D]);
    gst(b.F[D][S_VX + D], c, (FluxT)fo[L_VI]);
    gst(b.F[D][S_VX + DJ], c, (FluxT)fo[L_VJ]);
    gst(b.F[D][S_VX + DK], c, (FluxT)fo[L_VK]);
    gst(b.F[D][S_E], c, (FluxT)fo[L_E]);
    gst(b.F[D][S_BX + DJ], c, (FluxT)fo[L_BJ]);
    gst(b.F[D][S_BX + DK], c, (FluxT)fo[L_BK]);
  } else {
    T l5[5], r5[5];
    l5[H_D] = wl[S_RHO]; r5[H_D] = wr[S_RHO];
    l5[H_VI] = wl[S_VX + D]; r5[H_VI] = wr[S_VX + D];
    l5[H_VJ] = wl[S_VX + DJ]; r5[H_VJ] = wr[S_VX + DJ];
    l5[H_VK] = wl[S_VX + DK]; r5[H_VK] = wr[S_VX + DK];
    l5[H_E] = wl[S_E]; r5[H_E] = wr[S_E];
    hllc(l5, r5, gamma, fo, vb);
    gst(b.F[D][S_RHO], c, (FluxT)fo[H_D]);
    gst(b.F[D][S_VX + D], c, (FluxT)fo[H_VI]);
    gst(b.F[D][S_VX + DJ], c, (FluxT)fo[H_VJ]);
    gst(b.F[D][S_VX + DK], c, (FluxT)fo[H_VK]);
    gst(b.F[D][S_E], c, (FluxT)fo[H_E]);
  }
  if (DE) {
    // the dual-energy outputs of the solve (EnzoRiemannHLLD.hpp:404-447,
    // EnzoRiemannHLL.hpp:322-344, EnzoRiemannHLLC.hpp:159-170): the
    // internal-energy flux from the stored density flux (slot 0 of either
    // LUT), and vi_bar
    gst(b.Fe[D], c, (double)upwind_eint_flux(eint_l, eint_r, fo[0]));
    gst(b.Vb[D], c, (double)vbar);
  }
}

// One thread per cell p of the box [cur_stale, m - cur_stale); it solves
// the three faces between p and p + e_d that lie in the flux region of the
// stage (EnzoRiemannImpl.hpp:317-338 with stale_depth = cur_stale: face
// index range [cs, m_d - 1 - cs) along d).
// Minimum waves per SIMD the flux kernel is compiled for (caps its VGPRs);
// the nearest-neighbour variant has its own.
#ifndef ENZOE_FLUX_MIN_WAVES
#define ENZOE_FLUX_MIN_WAVES 3
#endif
#ifndef ENZOE_FLUX_NN_WAVES
#define ENZOE_FLUX_NN_WAVES 3
#endif

// PRIMF: the stage reads the FP32 primitive half-step state of the mixed
// path (b.Up), whose cell update also stored the centre E.
// flux_cells: the work of the flux kernel's workgroup (bx, by); k_flux
// and the persistent stage kernel call it.  dir: -1, the three low faces of
// each cell; 0, 1, 2, only the face along that axis (the face-split launch
// of small blocks: three workgroups per cell range, one face each; the
// first one also writes the centre E).
template <typename T, typename FluxT, int RECON, int RIEMANN, bool MHD, bool DE, bool PRIMF>
__device__ __forceinline__ void flux_cells(const BlockDesc<FluxT> *__restrict__ blocks,
                                           const StageParams &sp, int bx, int by,
                                           int dir = -1) {
  if (zero_step(sp, by)) return;
  const BlockDesc<FluxT> &b = blocks[by];
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int cs = sp.cur_stale;
  int k, j, i;
  const long t = (long)bx * blockDim.x + threadIdx.x;
  if (!cell_coords(t, cs, mx - 2 * cs, my - 2 * cs, mz - 2 * cs,
                   (sp.xcd_remap & 1) ? b.tile_rows : 0, k, j, i))
    return;
  double *const *U = (sp.stage == 0) ? b.U : b.Ut;
  const double *const *bfr = (sp.stage == 0) ? b.bfi : b.bft;
  FluxConsts fc;
  fc.gm1 = sp.gamma - 1.0;
  fc.gamma = sp.gamma;
  fc.theta = sp.theta;
  fc.dfloor = sp.density_floor;
  fc.pfloor = sp.pressure_floor;
  fc.dbg = sp.dbg;
  const int c = (k * my + j) * mx + i;
  const double *ei = DE ? ((sp.stage == 0) ? b.ei : b.eit) : nullptr;
  const PrimSource<T, MHD, DE, PRIMF> src{U, ei, b.Up, fc.gm1};
  T wc[S_N];
  if constexpr (PRIMF) {
    src(c, wc);
  } else {
    double wd[S_N];
    load_prim<MHD, DE>(U, c, fc.gm1, wd, ei);
    if (MHD && dir <= 0) {
      // centre E_i = -v_j B_k + v_k B_j of this cell for the edge kernel
      // (EnzoBfieldMethodCT::compute_center_efield, EnzoBfieldMethodCT.cpp:
      // 267-292); the box of k_flux covers every cell the edges read
      gst(b.Ec[0], c, center_efield(wd[S_VY], wd[S_VZ], wd[S_BY], wd[S_BZ]));
      gst(b.Ec[1], c, center_efield(wd[S_VZ], wd[S_VX], wd[S_BZ], wd[S_BX]));
      gst(b.Ec[2], c, center_efield(wd[S_VX], wd[S_VY], wd[S_BX], wd[S_BY]));
    }
#pragma unroll
    for (int q = 0; q < S_N; q++) wc[q] = (T)wd[q];
  }
  // the z face first: in the tiled order the plane above is the one not yet
  // in L2, so its loads go out with this cell's and the two latencies
  // overlap (each face writes its own arrays; the order changes no value)
  if ((dir < 0 || dir == 2) && k < mz - 1 - cs)
    flux_face<T, FluxT, RECON, RIEMANN, MHD, 2, DE, PRIMF>(b, src, bfr[2], k, j, i, c, wc, fc);
  if ((dir < 0 || dir == 0) && i < mx - 1 - cs)
    flux_face<T, FluxT, RECON, RIEMANN, MHD, 0, DE, PRIMF>(b, src, bfr[0], k, j, i, c, wc, fc);
  if ((dir < 0 || dir == 1) && j < my - 1 - cs)
    flux_face<T, FluxT, RECON, RIEMANN, MHD, 1, DE, PRIMF>(b, src, bfr[1], k, j, i, c, wc, fc);
}

template <typename T, typename FluxT, int RECON, int RIEMANN, bool MHD,
          bool DE = false, bool PRIMF = false>
__global__ void __launch_bounds__(256, RECON == RC_NN ? ENZOE_FLUX_NN_WAVES
                                                      : ENZOE_FLUX_MIN_WAVES)
k_flux(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp) {
  int bx, by;
  wg_coords(sp.xcd_remap & 1, bx, by);
  // a launch with three z slices is the face-split one
  flux_cells<T, FluxT, RECON, RIEMANN, MHD, DE, PRIMF>(blocks, sp, bx, by,
                                                      gridDim.z == 3 ? (int)blockIdx.z : -1);
}

// The primitive state of the neighbour cell a face needs, already at hand
// (the z-pencil kernel's LDS plane or its next-plane load).
template <typename T>
struct GivenPrim {
  const T *w;
  __device__ __forceinline__ void operator()(int, T out[S_N]) const {
#pragma unroll
    for (int q = 0; q < S_N; q++) out[q] = w[q];
  }
};

// Tile of the z-pencil flux kernel: 16 x 16 cells of a plane per
// workgroup (a wave takes 4 rows of 16: four 128-byte lines per FP64 load,
// as a 64-cell row would; square, so that the +x column and +y row past it
// add only 1/8 and partial tiles at the box edges stay small).
constexpr int ZP_TX = 16, ZP_TY = 16;

// The nearest-neighbour flux stage as z pencils: a workgroup takes a 16 x 16
// tile of (x, y) columns of the flux box and marches up through a segment
// of its planes.  Per plane the tile's primitive states (and the +x column
// and +y row past it) sit in LDS, so a cell's x- and y-face neighbours are
// read from there and its z-face neighbour is the thread's own load of the
// next plane, which is its own state at the next step of the march: every
// cell's state is loaded and converted to primitives about 1.1 times
// instead of 4 times (k_flux).  The faces and their arithmetic are
// k_flux's (flux_face, the same neighbour values), so the same bits.
// Launch: grid.x = tiles_x * tiles_y * segments, grid.y = block.
template <typename T, typename FluxT, int RIEMANN, bool MHD, bool DE, bool PRIMF>
__global__ void __launch_bounds__(256, ENZOE_FLUX_NN_WAVES)
k_flux_zp(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp, int zplanes) {
  const int by = blockIdx.y;
  if (zero_step(sp, by)) return;
  const BlockDesc<FluxT> &b = blocks[by];
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int cs = sp.cur_stale;
  const int nx = mx - 2 * cs, ny = my - 2 * cs, nz = mz - 2 * cs;
  const int ntx = (nx + ZP_TX - 1) / ZP_TX, nty = (ny + ZP_TY - 1) / ZP_TY;
  const int tile = (int)blockIdx.x % (ntx * nty), seg = (int)blockIdx.x / (ntx * nty);
  const int tx = (int)threadIdx.x % ZP_TX, ty = (int)threadIdx.x / ZP_TX;
  const int i = cs + (tile % ntx) * ZP_TX + tx, j = cs + (tile / ntx) * ZP_TY + ty;
  const int k0 = cs + seg * zplanes;
  const int k1 = (k0 + zplanes < mz - cs) ? k0 + zplanes : mz - cs;
  const bool mine = i < mx - cs && j < my - cs;  // a cell of the flux box
  double *const *U = (sp.stage == 0) ? b.U : b.Ut;
  const double *const *bfr = (sp.stage == 0) ? b.bfi : b.bft;
  FluxConsts fc;
  fc.gm1 = sp.gamma - 1.0;
  fc.gamma = sp.gamma;
  fc.theta = sp.theta;
  fc.dfloor = sp.density_floor;
  fc.pfloor = sp.pressure_floor;
  fc.dbg = sp.dbg;
  const double *ei = DE ? ((sp.stage == 0) ? b.ei : b.eit) : nullptr;
  const PrimSource<T, MHD, DE, PRIMF> src{U, ei, b.Up, fc.gm1};
  // the halo cell this thread loads per plane: the +y row past the tile
  // (threads 0 .. ZP_TX-1) or the +x column past it (the next ZP_TY)
  const int t = (int)threadIdx.x;
  const int hi_ = (t < ZP_TX) ? cs + (tile % ntx) * ZP_TX + t : cs + (tile % ntx) * ZP_TX + ZP_TX;
  const int hj_ = (t < ZP_TX) ? cs + (tile / ntx) * ZP_TY + ZP_TY
                              : cs + (tile / ntx) * ZP_TY + (t - ZP_TX);
  const bool halo = t < ZP_TX + ZP_TY && hi_ < mx - cs && hj_ < my - cs;
  // a cell's state as primitives, FP64 (kept for the centre E) and as T,
  // exactly as k_flux forms them
  auto load = [&](int cc, double d[S_N], T w[S_N]) {
    if constexpr (PRIMF) {
      src(cc, w);
    } else {
      load_prim<MHD, DE>(U, cc, fc.gm1, d, ei);
#pragma unroll
      for (int q = 0; q < S_N; q++) w[q] = (T)d[q];
    }
  };
  // LDS: the tile's plane with the +x column and +y row
  __shared__ T sw[ZP_TY + 1][ZP_TX + 1][S_N];
  T wc[S_N];
  double wd[S_N];
  if (mine && k0 < k1) load((k0 * my + j) * mx + i, wd, wc);
  for (int k = k0; k < k1; k++) {
    const int c = (k * my + j) * mx + i;
    if (mine) {
#pragma unroll
      for (int q = 0; q < S_N; q++) sw[ty][tx][q] = wc[q];
      if (MHD && !PRIMF) {
        // centre E of this cell for the edge kernel (as k_flux)
        gst(b.Ec[0], c, center_efield(wd[S_VY], wd[S_VZ], wd[S_BY], wd[S_BZ]));
        gst(b.Ec[1], c, center_efield(wd[S_VZ], wd[S_VX], wd[S_BZ], wd[S_BX]));
        gst(b.Ec[2], c, center_efield(wd[S_VX], wd[S_VY], wd[S_BX], wd[S_BY]));
      }
    }
    if (halo) {
      T h[S_N];
      double hd[S_N];
      load((k * my + hj_) * mx + hi_, hd, h);
      const int ly = (t < ZP_TX) ? ZP_TY : t - ZP_TX;
      const int lx = (t < ZP_TX) ? t : ZP_TX;
#pragma unroll
      for (int q = 0; q < S_N; q++) sw[ly][lx][q] = h[q];
    }
    __syncthreads();
    if (mine) {
      if (i < mx - 1 - cs)
        flux_face<T, FluxT, RC_NN, RIEMANN, MHD, 0, DE, PRIMF>(
            b, GivenPrim<T>{sw[ty][tx + 1]}, bfr[0], k, j, i, c, wc, fc);
      if (j < my - 1 - cs)
        flux_face<T, FluxT, RC_NN, RIEMANN, MHD, 1, DE, PRIMF>(
            b, GivenPrim<T>{sw[ty + 1][tx]}, bfr[1], k, j, i, c, wc, fc);
      if (k < mz - 1 - cs) {
        T up[S_N];
        double upd[S_N];
        load(c + mx * my, upd, up);
        flux_face<T, FluxT, RC_NN, RIEMANN, MHD, 2, DE, PRIMF>(
            b, GivenPrim<T>{up}, bfr[2], k, j, i, c, wc, fc);
#pragma unroll
        for (int q = 0; q < S_N; q++) {
          wc[q] = up[q];
          wd[q] = upd[q];
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- k_edge

// identify_upwind weight from the density flux (EnzoBfieldMethodCT.cpp:
// 170-216)
__device__ __forceinline__ double upwind_weight(double f) {
  return (f > 0) ? 1.0 : ((f < 0) ? 0.0 : 0.5);
}

// The centre E of the stage: the mixed path's final stage reads the one its
// first stage's cell update stored with the half-step state.
template <typename FluxT>
__device__ __forceinline__ double *const *center_e(const BlockDesc<FluxT> &b,
                                                   const StageParams &sp) {
  return (sp.stage > 0 && b.Ech[0] != nullptr) ? b.Ech : b.Ec;
}

// The upwinded edge EMF along axis AX at (cell p_AX, faces p_j + 1/2,
// p_k + 1/2), c the cell-shaped index of p, sj / sk the strides of the
// other two axes (EnzoBfieldMethodCT::compute_edge_efield /
// compute_edge_<true>, EnzoBfieldMethodCT.cpp:384-557), from the centre E
// of the four cells around it (compute_center_efield, :267-292) and the
// face fluxes of B and density.  FP64, never contracted: the same bits in
// every kernel and build that evaluates it.
template <int AX, typename FluxT>
__device__ __forceinline__ double edge_emf(const BlockDesc<FluxT> &b, double *const *Ecs,
                                           long c, long sj, long sk) {
#pragma clang fp contract(off)
  constexpr int aj = (AX + 1) % 3, ak = (AX + 2) % 3;
  const double *Eci = Ecs[AX];
  const double Ec = Eci[c], Ec_jp1 = Eci[c + sj], Ec_kp1 = Eci[c + sk],
               Ec_jkp1 = Eci[c + sj + sk];
  // face E: j-faces carry -F_j(B_k), k-faces +F_k(B_j)
  const FluxT *Fj = b.F[aj][S_BX + ak];
  const FluxT *Fk = b.F[ak][S_BX + aj];
  const double Ej = (double)Fj[c], Ej_kp1 = (double)Fj[c + sk];
  const double Ek = (double)Fk[c], Ek_jp1 = (double)Fk[c + sj];
  const double Wj = upwind_weight((double)b.F[aj][S_RHO][c]);
  const double Wj_kp1 = upwind_weight((double)b.F[aj][S_RHO][c + sk]);
  const double Wk = upwind_weight((double)b.F[ak][S_RHO][c]);
  const double Wk_jp1 = upwind_weight((double)b.F[ak][S_RHO][c + sj]);

  double dEdj_r = Wk_jp1 * (Ec_jp1 + Ej) + (1 - Wk_jp1) * (Ec_jkp1 + Ej_kp1);
  double dEdj_l = Wk * (-Ej - Ec) + (1 - Wk) * (-Ej_kp1 - Ec_kp1);
  double dEdk_r = Wj_kp1 * (Ec_kp1 - Ek) + (1 - Wj_kp1) * (Ec_jkp1 - Ek_jp1);
  double dEdk_l = Wj * (Ek - Ec) + (1 - Wj) * (Ek_jp1 - Ec_jp1);
  double Ej_sum = Ej + Ej_kp1;
  Ej_sum *= -1;
  double Ek_sum = Ek + Ek_jp1;
  return 0.25 * (Ej_sum + Ek_sum + (dEdj_l - dEdj_r) + (dEdk_l - dEdk_r));
}

// Whether the edge EMF along axis ax at cell-shaped index p is one the
// stage computes: [cs+1, m_ax-cs-1) along ax, [cs, m-cs-1) across.
__device__ __forceinline__ bool edge_in_region(int ax, const int p[3], const int m[3], int cs) {
  const int aj = (ax + 1) % 3, ak = (ax + 2) % 3;
  return p[ax] >= cs + 1 && p[ax] < m[ax] - cs - 1 && p[aj] >= cs && p[aj] < m[aj] - cs - 1 &&
         p[ak] >= cs && p[ak] < m[ak] - cs - 1;
}

// One thread per position p of the box [cs, m - cs); computes the edge EMF
// along each axis i at (cell p_i, faces p_j + 1/2, p_k + 1/2) when the edge
// lies in [cs+1, m_i-cs-1) x [cs, m_j-cs-1) x [cs, m_k-cs-1)
// (EnzoBfieldMethodCT::compute_edge_efield / compute_edge_<true>,
// EnzoBfieldMethodCT.cpp:384-557); the centre E of the current-stage state
// (compute_center_efield, :267-292) was written by k_flux (or, for the mixed
// path's final stage, by its first stage's cell update).
template <typename FluxT>
__device__ __forceinline__ void edge_cells(const BlockDesc<FluxT> *__restrict__ blocks,
                                           const StageParams &sp, int bx, int by) {
  if (zero_step(sp, by)) return;
  const BlockDesc<FluxT> &b = blocks[by];
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int cs = sp.cur_stale;
  int k, j, i;
  const long t = (long)bx * blockDim.x + threadIdx.x;
  if (!cell_coords(t, cs, mx - 2 * cs, my - 2 * cs, mz - 2 * cs,
                   (sp.xcd_remap & 2) ? b.tile_rows : 0, k, j, i))
    return;
  const int p[3] = {i, j, k};
  const int m[3] = {mx, my, mz};
  const long c = idx3(k, j, i, my, mx);
  const long str[3] = {1, (long)mx, (long)mx * my};
  double *const *Ecs = center_e(b, sp);
  // the three axes unrolled: their loads go out together
  if (edge_in_region(0, p, m, cs)) b.E[0][c] = edge_emf<0>(b, Ecs, c, str[1], str[2]);
  if (edge_in_region(1, p, m, cs)) b.E[1][c] = edge_emf<1>(b, Ecs, c, str[2], str[0]);
  if (edge_in_region(2, p, m, cs)) b.E[2][c] = edge_emf<2>(b, Ecs, c, str[0], str[1]);
}
template <typename FluxT>
__global__ void __launch_bounds__(256)
k_edge(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp) {
  int bx, by;
  wg_coords(sp.xcd_remap & 2, bx, by);
  edge_cells<FluxT>(blocks, sp, bx, by);
}

// --------------------------------------------------------------- k_faces

// The updated face B_i from its start-of-step value and the four edge EMFs
// around it (EnzoBfieldMethodCT::update_bfield, EnzoBfieldMethodCT.cpp:
// 617-687): E_k at j -/+ 1/2 (ek_m, ek_p), E_j at k -/+ 1/2 (ej_m, ej_p).
// Never contracted.
__device__ __forceinline__ double face_b_update(double b0, double dtdj, double dtdk,
                                                double ek_p, double ek_m, double ej_p,
                                                double ej_m) {
#pragma clang fp contract(off)
  double E_k_term = dtdj * (ek_p - ek_m);
  double E_j_term = dtdk * (ej_p - ej_m);
  return b0 - E_k_term + E_j_term;
}

// One thread per position p of the box [cs, m - cs + 1) (faces reach one
// index further); updates face X = p_i of bfieldi_i when
// X in [cs+1, m_i-cs) and p_j, p_k in [cs+1, m-cs-1)
// (EnzoBfieldMethodCT::update_bfield, EnzoBfieldMethodCT.cpp:617-687).  The
// starting value is always the start-of-step face field.
template <typename FluxT>
__device__ __forceinline__ void face_cells(const BlockDesc<FluxT> *__restrict__ blocks,
                                           const StageParams &sp, int bx, int by) {
  if (zero_step(sp, by)) return;
  const BlockDesc<FluxT> &b = blocks[by];
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int cs = sp.cur_stale;
  int k, j, i;
  const long t = (long)bx * blockDim.x + threadIdx.x;
  const int lo = cs + 1;
  if (!cell_coords(t, lo, mx - cs - lo, my - cs - lo, mz - cs - lo,
                   (sp.xcd_remap & 4) ? b.tile_rows : 0, k, j, i))
    return;
  const double dt = sp.dts ? sp.dts[by] * sp.dt_factor : sp.dt;
  const int p[3] = {i, j, k};
  const int m[3] = {mx, my, mz};
  const long str[3] = {1, (long)mx, (long)mx * my};
// the three axes unrolled: their loads go out together
#pragma unroll
  for (int ax = 0; ax < 3; ax++) {
    const int aj = (ax + 1) % 3, ak = (ax + 2) % 3;
    if (p[aj] >= m[aj] - cs - 1 || p[ak] >= m[ak] - cs - 1) continue;
    // p[ax] in [cs+1, m_ax - cs) by construction of the box
    const double dtdj = dt / b.cw[aj];
    const double dtdk = dt / b.cw[ak];
    // edges at i-position X - 1/2: cell-shaped index of p - e_i
    const long e = idx3(k, j, i, my, mx) - str[ax];
    const double *Ej = b.E[aj], *Ek = b.E[ak];
    const int bnx = mx + (ax == 0), bny = my + (ax == 1);
    const long cb = idx3(k, j, i, bny, bnx);
    double *out = sp.is_final ? b.bfi[ax] : b.bft[ax];
    out[cb] = face_b_update(b.bfi[ax][cb], dtdj, dtdk, Ek[e], Ek[e - str[aj]], Ej[e],
                            Ej[e - str[ak]]);
  }
}
template <typename FluxT>
__global__ void __launch_bounds__(256)
k_faces(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp) {
  int bx, by;
  wg_coords(sp.xcd_remap & 4, bx, by);
  face_cells<FluxT>(blocks, sp, bx, by);
}

// ------------------------------------------------------- CFL of one cell

// EnzoMHDIntegratorStageCommands::timestep (EnzoMHDIntegratorStageCommands.cpp:
// 299-366) for one cell without the dual-energy formalism: its pressure
// (EnzoComputePressure.cpp:177-189) and min_i dx_i / (|v_i| + c_f) with the
// fast magnetosonic speed (EnzoEOSIdeal.hpp:111-140) or the sound speed.
// Shared by k_dt and the final k_cells so both give the same bits; never
// contracted into FMAs, also where the including unit allows it (the mixed
// build), so that both round alike.
template <bool MHD>
__device__ __forceinline__ double cell_pressure(double d, double vx, double vy, double vz,
                                                double e, double bx, double by, double bz,
                                                double gamma) {
#pragma clang fp contract(off)
  const double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
  const double me = MHD ? 0.5 * (bx * bx + by * by + bz * bz) : 0.;
  return (gamma - 1.0) * (d * (e - ke) - me);
}
template <bool MHD>
__device__ __forceinline__ double cfl_speed(double d, double vx, double vy, double vz,
                                            double bx, double by, double bz, double p,
                                            double gamma, const double cw[3]) {
#pragma clang fp contract(off)
  double sig;
  if (MHD) {
    double B2 = ((bx * bx) + ((by * by) + (bz * bz)));
    double cs2 = gamma * p / d;
    double va2 = B2 / d;
    sig = sqrt(va2 + cs2);
  } else {
    sig = sqrt(gamma * p / d);
  }
  return dmin3(cw[0] / (fabs(vx) + sig), cw[1] / (fabs(vy) + sig),
               cw[2] / (fabs(vz) + sig));
}
template <bool MHD>
__device__ __forceinline__ double cfl_cell(double d, double vx, double vy, double vz,
                                           double e, double bx, double by, double bz,
                                           double gamma, const double cw[3], double &p) {
  p = cell_pressure<MHD>(d, vx, vy, vz, e, bx, by, bz, gamma);
  return cfl_speed<MHD>(d, vx, vy, vz, bx, by, bz, p, gamma, cw);
}

// Min of `local` over the workgroup into *result, as the bit pattern of a
// non-negative double (an integer atomicMin is then an exact double min; a
// NaN never wins, as std::min in the reference ignores it): wave64
// butterfly, LDS across the waves, one atomic per workgroup.  Every thread
// of the workgroup must call it.
__device__ __forceinline__ void block_min_to(unsigned long long *result, double local) {
  unsigned long long bits = __double_as_longlong(local);
  if (!(local >= 0.0)) bits = 0x7FF8000000000000ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(bits, off, 64);
    bits = (o < bits) ? o : bits;
  }
  __shared__ unsigned long long wmin[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wmin[wave] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmin[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = (wmin[w] < m) ? wmin[w] : m;
    // one address for every workgroup of a launch (hundreds of thousands
    // in a batched cell update): most find a smaller minimum there already
    // and skip the atomic.  The value only decreases, so a stale read is
    // larger and never skips a needed update.
    if (m < __hip_atomic_load(result, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(result, m);
  }
}

// --------------------------------------------------------------- k_cells

// One thread per cell of [lo, m - lo) with lo = cur_stale for the final
// stage (cell-centred B is rebuilt there, EnzoBfieldMethodCT.cpp:702-728)
// and cur_stale + 1 otherwise.  Cells in [cs+1, m-cs-1) also get the flux
// divergence (EnzoIntegrationQuanUpdate::accumulate_flux_component,
// EnzoIntegrationQuanUpdate.cpp:105-147), the conserved update with the
// density floor (update_quantities, :183-238) and the energy floor
// (EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync, :162-290).
// With the dual-energy formalism the internal energy is updated as a
// specific quantity: per dim its flux divergence, then the source
// -dt/dx p (vi_bar_r - vi_bar_l) with the stage's cell pressure
// (EnzoSourceInternalEnergy::calculate_source, EnzoSourceInternalEnergy.cpp:
// 16-92, called after each dim's divergence,
// EnzoMHDIntegratorStageCommands.cpp:248-255), and the energies are
// synchronised instead of floored.
// The per-cell work of k_cells for the cell (k, j, i) of the block `b`;
// in_launch / in_box: the cell lies in the launch box / in [lo, m - lo).
// face_b(k, j, i, f[6]) gives the cell's six face values of the stage's
// face-centred B (low and high face along x, y, z).  Returns the cell's CFL
// term (DBL_MAX when it takes none).
template <typename FluxT, bool MHD, bool DE, typename FaceB>
__device__ __forceinline__ double cell_stage(const BlockDesc<FluxT> &b, const StageParams &sp,
                                             int by, int k, int j, int i, bool in_launch,
                                             bool in_box, const FaceB &face_b) {
  // the FP64 update is never contracted into FMAs, also where the including
  // unit allows it: every kernel that runs it (k_cells, k_ctu, k_stage)
  // rounds alike
#pragma clang fp contract(off)
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int cs = sp.cur_stale;
  const bool cfl = sp.cfl != nullptr;
  double local = 1.7976931348623157e308;
  // the mixed path keeps the half-step state as FP32 primitives (b.Up)
  // plus the FP64 centre E of the next stage's edge EMFs
  constexpr bool HALF_F32 = std::is_same<FluxT, float>::value && !DE;
  const bool to_f32 = HALF_F32 && !sp.is_final;
  auto body = [&]() {
#pragma clang fp contract(off)
    const long c = idx3(k, j, i, my, mx);
    double *const *Uo = sp.is_final ? b.U : b.Ut;
    double bc[3] = {0., 0., 0.};
    if (MHD) {
      double f[6];
      face_b(k, j, i, f);
      bc[0] = 0.5 * (f[0] + f[1]);
      bc[1] = 0.5 * (f[2] + f[3]);
      bc[2] = 0.5 * (f[4] + f[5]);
      if (!to_f32) {
        Uo[S_BX][c] = bc[0];
        Uo[S_BY][c] = bc[1];
        Uo[S_BZ][c] = bc[2];
      }
    }
    const int ul = cs + 1;
    if (k < ul || j < ul || i < ul || k >= mz - ul || j >= my - ul ||
        i >= mx - ul) {
      if (cfl) {
        // outside the update, inside the cell-centred B rebuild: the
        // start-of-step state with the new centre B
        const double d = Uo[S_RHO][c];
        const double v[3] = {Uo[S_VX][c], Uo[S_VY][c], Uo[S_VZ][c]};
        double p;
        if (DE) {
          // the timestep's in-place synchronisation first
          // (EnzoMethodMHDVlct.cpp:563-568), as k_dt does it
          const double b2 = MHD ? (bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]) : 0.;
          double etot = Uo[S_E][c], eint = b.ei[c];
          floor_and_sync<MHD, true>(d, v, b2, sp.gamma, sp.pressure_floor, sp.eta, sp.ggm1,
                                    etot, eint);
          Uo[S_E][c] = etot;
          b.ei[c] = eint;
          p = (sp.gamma - 1.0) * d * eint;
          local = cfl_speed<MHD>(d, v[0], v[1], v[2], bc[0], bc[1], bc[2], p, sp.gamma, b.cw);
        } else {
          local = cfl_cell<MHD>(d, v[0], v[1], v[2], Uo[S_E][c], bc[0], bc[1], bc[2],
                                sp.gamma, b.cw, p);
        }
        if (b.pressure) b.pressure[c] = p;
      }
      return;
    }
    const double dt = sp.dts ? sp.dts[by] * sp.dt_factor : sp.dt;
    const long str[3] = {1, (long)mx, (long)mx * my};
    const int fq[5] = {S_RHO, S_VX, S_VY, S_VZ, S_E};
    double dU[5];
  #pragma unroll
    for (int q = 0; q < 5; q++) dU[q] = 0.;
    double dUe = 0.;
    double pc = 0.;
    if (DE) {
      // the stage's primitive pressure, floored (EnzoSourceInternalEnergy.cpp:
      // 75-88)
      double *const *Us = (sp.stage == 0) ? b.U : b.Ut;
      const double *es = (sp.stage == 0) ? b.ei : b.eit;
      pc = dmax((sp.gamma - 1.0) * Us[S_RHO][c] * es[c], sp.pressure_floor);
    }
  #pragma unroll
    for (int d = 0; d < 3; d++) {
      const double dtdx = dt / b.cw[d];
  #pragma unroll
      for (int q = 0; q < 5; q++) {
        const FluxT *F = b.F[d][fq[q]];
        double fr = (double)F[c], fl = (double)F[c - str[d]];
        dU[q] -= dtdx * (fr - fl);
      }
      if (DE) {
        dUe -= dtdx * (b.Fe[d][c] - b.Fe[d][c - str[d]]);
        dUe -= dtdx * pc * (b.Vb[d][c] - b.Vb[d][c - str[d]]);
      }
    }
    // gravity (EnzoSourceGravity::calculate_source, EnzoSourceGravity.cpp:
    // 17-69) on the stage with index 1, from the start-of-step state, after
    // the flux divergence (EnzoMHDIntegratorStageCommands.cpp:180-182,
    // 265-295)
    if (sp.stage == 1 && b.acc[0] != nullptr) {
      const double rho = b.U[S_RHO][c];
      const double ax = b.acc[0][c], ay = b.acc[1][c], az = b.acc[2][c];
      dU[1] += dt * rho * ax;
      dU[2] += dt * rho * ay;
      dU[3] += dt * rho * az;
      dU[4] += dt * rho * ((b.U[S_VX][c] * ax) + (b.U[S_VY][c] * ay) +
                           (b.U[S_VZ][c] * az));
    }
    // the update base is always the start-of-step state
    double old_rho = b.U[S_RHO][c];
    double new_rho = b.U[S_RHO][c] + dU[0];
    if (sp.dbg) debug_count(sp.dbg, DBG_DENSITY_FLOOR, new_rho < sp.density_floor);
    new_rho = dmax(new_rho, sp.density_floor);
    double inv_new_rho = 1. / new_rho;
    double v[3];
  #pragma unroll
    for (int q = 1; q < 4; q++)
      v[q - 1] = (b.U[fq[q]][c] * old_rho + dU[q]) * inv_new_rho;
    double etot = (b.U[S_E][c] * old_rho + dU[4]) * inv_new_rho;
    double eint = 0.;
    if (DE) eint = (b.ei[c] * old_rho + dUe) * inv_new_rho;
    // energy floor (and synchronisation)
    const double b2 = MHD ? (bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]) : 0.;
    bool floored = false;
    floor_and_sync<MHD, DE>(new_rho, v, b2, sp.gamma, sp.pressure_floor, sp.eta, sp.ggm1,
                            etot, eint, sp.dbg ? &floored : nullptr);
    if (sp.dbg) {
      debug_count(sp.dbg, DBG_ENERGY_FLOOR, floored);
      bool finite = isfinite(new_rho) && isfinite(v[0]) && isfinite(v[1]) &&
                    isfinite(v[2]) && isfinite(etot) && isfinite(b2);
      if (DE) finite = finite && isfinite(eint);
      debug_count(sp.dbg, DBG_NONFINITE, !finite);
    }
    if (to_f32) {
      float *const *Up = b.Up;
      gst(Up[S_RHO], c, (float)new_rho);
      gst(Up[S_VX], c, (float)v[0]);
      gst(Up[S_VY], c, (float)v[1]);
      gst(Up[S_VZ], c, (float)v[2]);
      gst(Up[S_E], c, (float)prim_pressure<MHD>(new_rho, v[0], v[1], v[2], etot, bc[0], bc[1],
                                                bc[2], sp.gamma - 1.0));
      if (MHD) {
        gst(Up[S_BX], c, (float)bc[0]);
        gst(Up[S_BY], c, (float)bc[1]);
        gst(Up[S_BZ], c, (float)bc[2]);
        gst(b.Ech[0], c, center_efield(v[1], v[2], bc[1], bc[2]));
        gst(b.Ech[1], c, center_efield(v[2], v[0], bc[2], bc[0]));
        gst(b.Ech[2], c, center_efield(v[0], v[1], bc[0], bc[1]));
      }
      return;
    }
    Uo[S_RHO][c] = new_rho;
    Uo[S_VX][c] = v[0];
    Uo[S_VY][c] = v[1];
    Uo[S_VZ][c] = v[2];
    if (cfl) {
      double p;
      if (DE) {
        // the timestep synchronises the energies of the cell once more, in
        // place, before its pressure (EnzoMethodMHDVlct.cpp:563-568)
        floor_and_sync<MHD, DE>(new_rho, v, b2, sp.gamma, sp.pressure_floor, sp.eta,
                                sp.ggm1, etot, eint);
        p = (sp.gamma - 1.0) * new_rho * eint;
        local = cfl_speed<MHD>(new_rho, v[0], v[1], v[2], bc[0], bc[1], bc[2], p, sp.gamma,
                               b.cw);
      } else {
        local = cfl_cell<MHD>(new_rho, v[0], v[1], v[2], etot, bc[0], bc[1], bc[2],
                              sp.gamma, b.cw, p);
      }
      if (b.pressure) b.pressure[c] = p;
    }
    Uo[S_E][c] = etot;
    if (DE) (sp.is_final ? b.ei : b.eit)[c] = eint;
  };
  if (in_box) {
    body();
  } else if (in_launch && cfl) {
    // the outer ghost layers: the timestep's work on the state as it is
    // (k_dt's per-cell steps)
    const long c = idx3(k, j, i, my, mx);
    const double d = b.U[S_RHO][c];
    const double v[3] = {b.U[S_VX][c], b.U[S_VY][c], b.U[S_VZ][c]};
    double bxc = 0., byc = 0., bzc = 0.;
    if (MHD) {
      bxc = b.U[S_BX][c]; byc = b.U[S_BY][c]; bzc = b.U[S_BZ][c];
    }
    double p;
    if (DE) {
      const double b2 = MHD ? (bxc * bxc + byc * byc + bzc * bzc) : 0.;
      double etot = b.U[S_E][c], eint = b.ei[c];
      floor_and_sync<MHD, true>(d, v, b2, sp.gamma, sp.pressure_floor, sp.eta, sp.ggm1,
                                etot, eint);
      b.U[S_E][c] = etot;
      b.ei[c] = eint;
      p = (sp.gamma - 1.0) * d * eint;
      local = cfl_speed<MHD>(d, v[0], v[1], v[2], bxc, byc, bzc, p, sp.gamma, b.cw);
    } else {
      local = cfl_cell<MHD>(d, v[0], v[1], v[2], b.U[S_E][c], bxc, byc, bzc, sp.gamma,
                            b.cw, p);
    }
    if (b.pressure) b.pressure[c] = p;
  }
  return local;
}

// The centre-B faces of a cell from the stage's face field in memory.
template <typename FluxT>
struct FaceBGlobal {
  const BlockDesc<FluxT> &b;
  double *const *bf;
  __device__ __forceinline__ void operator()(int k, int j, int i, double f[6]) const {
    const int mx = b.mx, my = b.my;
    f[0] = bf[0][idx3(k, j, i, my, mx + 1)];
    f[1] = bf[0][idx3(k, j, i + 1, my, mx + 1)];
    f[2] = bf[1][idx3(k, j, i, my + 1, mx)];
    f[3] = bf[1][idx3(k, j + 1, i, my + 1, mx)];
    f[4] = bf[2][idx3(k, j, i, my, mx)];
    f[5] = bf[2][idx3(k + 1, j, i, my, mx)];
  }
};

// update_cells: the work of the cell-update kernel's workgroup (bx, by);
// with the CFL minimum every thread of the workgroup reaches its barrier.
template <typename FluxT, bool MHD, bool DE>
__device__ __forceinline__ void update_cells(const BlockDesc<FluxT> *__restrict__ blocks,
                                             const StageParams &sp, int bx, int by) {
  if (zero_step(sp, by)) return;
  const BlockDesc<FluxT> &b = blocks[by];
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int cs = sp.cur_stale;
  const int lo = sp.is_final ? cs : cs + 1;
  int k, j, i;
  const long t = (long)bx * blockDim.x + threadIdx.x;
  // with the CFL minimum the launch covers the whole block (the cells
  // outside [lo, m - lo) only take their CFL term) and every thread takes
  // part in the reduction
  const bool cfl = sp.cfl != nullptr;
  const int blo = cfl ? 0 : lo;
  const bool in_launch = cell_coords(t, blo, mx - 2 * blo, my - 2 * blo, mz - 2 * blo,
                                     (sp.xcd_remap & 8) ? b.tile_rows : 0, k, j, i);
  if (!in_launch && !cfl) return;
  const bool in_box = in_launch && k >= lo && j >= lo && i >= lo && k < mz - lo &&
                      j < my - lo && i < mx - lo;
  const FaceBGlobal<FluxT> fb{b, sp.is_final ? b.bfi : b.bft};
  const double local = cell_stage<FluxT, MHD, DE>(b, sp, by, k, j, i, in_launch, in_box, fb);
  if (cfl) block_min_to(sp.cfl, local);
}
template <typename FluxT, bool MHD, bool DE = false>
__global__ void __launch_bounds__(256)
k_cells(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp) {
  int bx, by;
  wg_coords(sp.xcd_remap & 8, bx, by);
  update_cells<FluxT, MHD, DE>(blocks, sp, bx, by);
}

// ------------------------------------------------ persistent stage kernel

// A barrier across every workgroup of a launch whose workgroups are all
// resident together (the launch holds no more workgroups than the device
// runs at once).  count: arrivals at the current barrier; gen: barriers
// passed so far; abort: set once a wait exceeds its poll limit, after which
// every barrier returns at once, so that the launch ends (the host reads
// the flag and reports the failure, enzoe_vlct_check_persistent).  The
// agent-scope fences around the arrival make every store of the launch
// before the barrier visible to every load after it, on every XCD.
struct GridBarrier {
  unsigned count;
  unsigned gen;
  unsigned abort;
  unsigned pad;
};

// thread 0 of the workgroup keeps `gen`, the barriers it has passed
__device__ __forceinline__ void grid_barrier(GridBarrier *gb, unsigned nwg, unsigned &gen,
                                             unsigned limit) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (__hip_atomic_load(&gb->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
      if (atomicAdd(&gb->count, 1u) == nwg - 1u) {
        // the last arrival: the count starts again from 0 before any
        // workgroup can see the barrier passed
        __hip_atomic_store(&gb->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&gb->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        unsigned polls = 0;
        while (__hip_atomic_load(&gb->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
          if (++polls > limit) {
            atomicExch(&gb->abort, 1u);
            break;
          }
          if (__hip_atomic_load(&gb->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
            break;
          __builtin_amdgcn_s_sleep(2);
        }
      }
    }
    gen++;
    __threadfence();
  }
  __syncthreads();
}

// The phases of one persistent stage launch: per block, the workgroups of
// the flux, edge, face and cell kernels (each phase's box as its own
// launch would cover it).
struct PersistParams {
  int nb;         // blocks of the table
  int nwg[4];     // workgroups per block: flux, edge EMF, face B, cells
  GridBarrier *bar;
  unsigned limit; // polls a barrier wait may take before it aborts
};

// One stage of the VL+CT step (k_flux, then for MHD k_edge and k_faces,
// then k_cells) as one launch: the launch's workgroups stay resident and
// take each phase's workgroups in turn, with a grid barrier between the
// phases where the separate launches had a kernel boundary.  The same
// per-workgroup work in the same cell order as the separate kernels (their
// plain order), so the same bits.  Workgroup w takes, in every phase, the
// virtual workgroups of one contiguous eighth of the phase (w % 8: the XCD
// the hardware deals it to), strided by the launch's workgroups per XCD.
template <typename T, typename FluxT, int RECON, int RIEMANN, bool MHD, bool PRIMF>
__global__ void __launch_bounds__(256, RECON == RC_NN ? ENZOE_FLUX_NN_WAVES
                                                      : ENZOE_FLUX_MIN_WAVES)
k_stage(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp, PersistParams pp) {
  unsigned gen = 0;
  if (threadIdx.x == 0)
    gen = __hip_atomic_load(&pp.bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned G = gridDim.x;  // a multiple of 8
  const unsigned x = blockIdx.x & 7u, r = blockIdx.x >> 3, G8 = G >> 3;
  auto phase = [&](int nwg, auto &&body) {
    const unsigned nv = (unsigned)nwg * (unsigned)pp.nb;
    const unsigned per = (nv + 7u) >> 3;
    const unsigned v0 = x * per;
    const unsigned v1 = (v0 + per < nv) ? v0 + per : nv;
    for (unsigned v = v0 + r; v < v1; v += G8) body((int)(v % (unsigned)nwg), (int)(v / (unsigned)nwg));
  };
  phase(pp.nwg[0], [&](int bx, int by) {
    flux_cells<T, FluxT, RECON, RIEMANN, MHD, false, PRIMF>(blocks, sp, bx, by);
  });
  grid_barrier(pp.bar, G, gen, pp.limit);
  if (MHD) {
    phase(pp.nwg[1], [&](int bx, int by) { edge_cells<FluxT>(blocks, sp, bx, by); });
    grid_barrier(pp.bar, G, gen, pp.limit);
    phase(pp.nwg[2], [&](int bx, int by) { face_cells<FluxT>(blocks, sp, bx, by); });
    grid_barrier(pp.bar, G, gen, pp.limit);
  }
  const bool cfl = sp.cfl != nullptr;
  phase(pp.nwg[3], [&](int bx, int by) {
    update_cells<FluxT, MHD, false>(blocks, sp, bx, by);
    // the CFL reduction's LDS words are read by thread 0 before the next
    // virtual workgroup writes them
    if (cfl) __syncthreads();
  });
}

// ----------------------------------------------------------------- k_ctu

// Constrained transport and the cell update of one stage in one launch:
// k_edge, k_faces and k_cells fused over tiles of TX x TY x TZ cells of
// the k_cells launch box, one cell per thread, with the same
// expressions in the same order (the same bits).  A workgroup
//  1. computes into LDS the edge EMFs its faces need: E_a over the tile
//     along a, and over the tile plus one edge below across a;
//  2. computes into LDS the face-centred B of every face of its cells (low
//     and high faces), from the edges in LDS; it writes the updated faces
//     it owns (the low faces of its cells and, in the last tile along an
//     axis, the face past the last cell); a face outside the update region
//     keeps its start-of-step value;
//  3. updates its cells as k_cells does, the centre B from the faces in LDS.
// Neighbouring tiles read the start-of-step value of a face that its owner
// may already have overwritten in place in the final stage: the first
// stage stores a copy (bf0) of the faces it updates, a region that holds
// the final stage's, and the final stage reads its base values there.
// The tile: TX x TY x TZ cells, one thread each.
template <int TX, int TY, int TZ>
struct CtuTile {
  static constexpr int NT = TX * TY * TZ;
  __host__ __device__ static constexpr int t(int d) { return d == 0 ? TX : (d == 1 ? TY : TZ); }
  // extent of the LDS edge array along axis a, direction d
  __host__ __device__ static constexpr int eext(int a, int d) { return t(d) + (d != a); }
  // extent of the LDS face array of B_a, direction d
  __host__ __device__ static constexpr int fext(int a, int d) { return t(d) + (d == a); }
  __host__ __device__ static constexpr int ecount(int a) {
    return eext(a, 0) * eext(a, 1) * eext(a, 2);
  }
  __host__ __device__ static constexpr int fcount(int a) {
    return fext(a, 0) * fext(a, 1) * fext(a, 2);
  }
  __host__ __device__ static constexpr int ebase(int a) {
    return a == 0 ? 0 : (a == 1 ? ecount(0) : ecount(0) + ecount(1));
  }
  __host__ __device__ static constexpr int fbase(int a) {
    return a == 0 ? 0 : (a == 1 ? fcount(0) : fcount(0) + fcount(1));
  }
  static constexpr int NE = ecount(0) + ecount(1) + ecount(2);
  static constexpr int NF = fcount(0) + fcount(1) + fcount(2);
  // LDS index of the edge along a at cell-shaped index p (tile origin o)
  __device__ static int eidx(int a, const int p[3], const int o[3]) {
    const int l0 = p[0] - o[0] + (a != 0), l1 = p[1] - o[1] + (a != 1),
              l2 = p[2] - o[2] + (a != 2);
    return ebase(a) + (l2 * eext(a, 1) + l1) * eext(a, 0) + l0;
  }
  // LDS index of face B_a at cell-shaped index (px, py, pz)
  __device__ static int fidx(int a, int px, int py, int pz, const int o[3]) {
    return fbase(a) + ((pz - o[2]) * fext(a, 1) + (py - o[1])) * fext(a, 0) + (px - o[0]);
  }
};

// The tile shapes of k_ctu the launch chooses from (enzoe_vlct_set_fused_ct
// / the autotuner): 1 = 16 x 4 x 4, 2 = 32 x 4 x 2, 3 = 8 x 8 x 4 (edges in
// LDS); 4 = 16 x 4 x 4 after k_edge (faces and cells fused).
typedef CtuTile<16, 4, 4> CtuTile1;
typedef CtuTile<32, 4, 2> CtuTile2;
typedef CtuTile<8, 8, 4> CtuTile3;

// Tiles of the launch box [blo, m - blo) along each axis.
template <class Tile>
__host__ __device__ inline void ctu_tiles(int mx, int my, int mz, int blo, int nt[3]) {
  const int n[3] = {mx - 2 * blo, my - 2 * blo, mz - 2 * blo};
  for (int d = 0; d < 3; d++) nt[d] = n[d] > 0 ? (n[d] + Tile::t(d) - 1) / Tile::t(d) : 0;
}

template <class Tile>
struct FaceBLds {
  const double *sF;
  const int *o;
  __device__ __forceinline__ void operator()(int k, int j, int i, double f[6]) const {
    f[0] = sF[Tile::fidx(0, i, j, k, o)];
    f[1] = sF[Tile::fidx(0, i + 1, j, k, o)];
    f[2] = sF[Tile::fidx(1, i, j, k, o)];
    f[3] = sF[Tile::fidx(1, i, j + 1, k, o)];
    f[4] = sF[Tile::fidx(2, i, j, k, o)];
    f[5] = sF[Tile::fidx(2, i, j, k + 1, o)];
  }
};

// The edges along AX of the tile into LDS, each thread a fixed set of them
// (the iterations unrolled, so their loads go out together).
template <int AX, class Tile, typename FluxT>
__device__ __forceinline__ void ctu_edges(const BlockDesc<FluxT> &b, double *const *Ecs,
                                          const int o[3], const int m[3], int cs,
                                          const long str[3], double *sE) {
  constexpr int n = Tile::ecount(AX);
  constexpr int ex = Tile::eext(AX, 0), ey = Tile::eext(AX, 1);
  constexpr int aj = (AX + 1) % 3, ak = (AX + 2) % 3;
#pragma unroll
  for (int it = 0; it < (n + Tile::NT - 1) / Tile::NT; it++) {
    const int r = threadIdx.x + it * Tile::NT;
    if (r < n) {
      const int p[3] = {o[0] - (AX != 0) + r % ex, o[1] - (AX != 1) + (r / ex) % ey,
                        o[2] - (AX != 2) + r / (ex * ey)};
      double val = 0.;
      if (edge_in_region(AX, p, m, cs))
        val = edge_emf<AX>(b, Ecs, idx3(p[2], p[1], p[0], m[1], m[0]), str[aj], str[ak]);
      sE[Tile::ebase(AX) + r] = val;
    }
  }
}

// The faces B_AX of the tile into LDS (and the owned updated ones to
// memory), the edge EMFs from LDS (EDGES_LDS) or from the edge arrays k_edge
// wrote.
template <int AX, class Tile, bool EDGES_LDS, typename FluxT>
__device__ __forceinline__ void ctu_faces(const BlockDesc<FluxT> &b, const StageParams &sp,
                                          double dt, const int o[3], const int m[3],
                                          const int hi[3], int cs, const double *sE,
                                          double *sF) {
  constexpr int n = Tile::fcount(AX);
  constexpr int fx = Tile::fext(AX, 0), fy = Tile::fext(AX, 1);
  constexpr int aj = (AX + 1) % 3, ak = (AX + 2) % 3;
  const int bnx = m[0] + (AX == 0), bny = m[1] + (AX == 1);
  const double dtdj = dt / b.cw[aj], dtdk = dt / b.cw[ak];
#pragma unroll
  for (int it = 0; it < (n + Tile::NT - 1) / Tile::NT; it++) {
    const int r = threadIdx.x + it * Tile::NT;
    if (r < n) {
      const int p[3] = {o[0] + r % fx, o[1] + (r / fx) % fy, o[2] + r / (fx * fy)};
      const bool upd = p[AX] >= cs + 1 && p[AX] < m[AX] - cs && p[aj] >= cs + 1 &&
                       p[aj] < m[aj] - cs - 1 && p[ak] >= cs + 1 && p[ak] < m[ak] - cs - 1;
      double val = 0.;
      if (upd) {
        const long cb = idx3(p[2], p[1], p[0], bny, bnx);
        const double b0 = sp.is_final ? b.bf0[AX][cb] : b.bfi[AX][cb];
        // edges at AX-position p_AX - 1/2 (cell-shaped index p - e_AX)
        int q[3] = {p[0], p[1], p[2]};
        q[AX] -= 1;
        int qj[3] = {q[0], q[1], q[2]}, qk[3] = {q[0], q[1], q[2]};
        qj[aj] -= 1;
        qk[ak] -= 1;
        if constexpr (EDGES_LDS) {
          val = face_b_update(b0, dtdj, dtdk, sE[Tile::eidx(ak, q, o)],
                              sE[Tile::eidx(ak, qj, o)], sE[Tile::eidx(aj, q, o)],
                              sE[Tile::eidx(aj, qk, o)]);
        } else {
          const long e = idx3(q[2], q[1], q[0], m[1], m[0]);
          const long str_j = aj == 0 ? 1 : (aj == 1 ? (long)m[0] : (long)m[0] * m[1]);
          const long str_k = ak == 0 ? 1 : (ak == 1 ? (long)m[0] : (long)m[0] * m[1]);
          const double *Ej = b.E[aj], *Ek = b.E[ak];
          val = face_b_update(b0, dtdj, dtdk, Ek[e], Ek[e - str_j], Ej[e], Ej[e - str_k]);
        }
        if (p[AX] < o[AX] + Tile::t(AX) || p[AX] == hi[AX]) {
          (sp.is_final ? b.bfi : b.bft)[AX][cb] = val;
          if (!sp.is_final) b.bf0[AX][cb] = b0;
        }
      } else if (sp.is_final && p[AX] <= m[AX] && p[aj] < m[aj] && p[ak] < m[ak]) {
        // outside the update region: the start-of-step value, never written
        // in this stage
        val = b.bfi[AX][idx3(p[2], p[1], p[0], bny, bnx)];
      }
      sF[Tile::fbase(AX) + r] = val;
    }
  }
}

// EDGES_LDS: the edge EMFs are computed here (k_edge + k_faces + k_cells
// fused); otherwise k_edge has written them and this is k_faces + k_cells.
template <typename FluxT, bool DE, class Tile, bool EDGES_LDS = true>
__global__ void __launch_bounds__(Tile::NT)
k_ctu(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp) {
  int bx, by;
  wg_coords(sp.xcd_remap & 8, bx, by);
  if (zero_step(sp, by)) return;
  const BlockDesc<FluxT> &b = blocks[by];
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int m[3] = {mx, my, mz};
  const int cs = sp.cur_stale;
  const int lo = sp.is_final ? cs : cs + 1;
  const bool cfl = sp.cfl != nullptr;
  const int blo = cfl ? 0 : lo;
  int nt[3];
  ctu_tiles<Tile>(mx, my, mz, blo, nt);
  // padding workgroups (launch rounded up to the XCDs) take part in nothing
  if (bx >= nt[0] * nt[1] * nt[2]) return;
  const int tix = bx % nt[0], tiy = (bx / nt[0]) % nt[1], tiz = bx / (nt[0] * nt[1]);
  const int o[3] = {blo + tix * Tile::t(0), blo + tiy * Tile::t(1), blo + tiz * Tile::t(2)};
  const int hi[3] = {mx - blo, my - blo, mz - blo};  // end of the launch box
  const long str[3] = {1, (long)mx, (long)mx * my};
  const double dt = sp.dts ? sp.dts[by] * sp.dt_factor : sp.dt;

  __shared__ double sE[EDGES_LDS ? Tile::NE : 1];
  __shared__ double sF[Tile::NF];

  // 1. edges
  if constexpr (EDGES_LDS) {
    double *const *Ecs = center_e(b, sp);
    ctu_edges<0, Tile>(b, Ecs, o, m, cs, str, sE);
    ctu_edges<1, Tile>(b, Ecs, o, m, cs, str, sE);
    ctu_edges<2, Tile>(b, Ecs, o, m, cs, str, sE);
    __syncthreads();
  }
  // 2. faces
  ctu_faces<0, Tile, EDGES_LDS>(b, sp, dt, o, m, hi, cs, sE, sF);
  ctu_faces<1, Tile, EDGES_LDS>(b, sp, dt, o, m, hi, cs, sE, sF);
  ctu_faces<2, Tile, EDGES_LDS>(b, sp, dt, o, m, hi, cs, sE, sF);
  __syncthreads();
  // 3. cells
  const int tid = threadIdx.x;
  const int i = o[0] + tid % Tile::t(0), j = o[1] + (tid / Tile::t(0)) % Tile::t(1),
            k = o[2] + tid / (Tile::t(0) * Tile::t(1));
  const bool in_launch = i < hi[0] && j < hi[1] && k < hi[2];
  const bool in_box = in_launch && k >= lo && j >= lo && i >= lo && k < mz - lo &&
                      j < my - lo && i < mx - lo;
  const FaceBLds<Tile> fb{sF, o};
  const double local = cell_stage<FluxT, true, DE>(b, sp, by, k, j, i, in_launch, in_box, fb);
  if (cfl) block_min_to(sp.cfl, local);
}

// ----------------------------------------------------------------- k_dt

// CFL timestep of every cell (ghosts included): writes the pressure field
// and min-reduces dx_i/(|v_i|+c_f) into *result, stored as the bit pattern
// of a non-negative double so that an integer atomicMin is an exact double
// min (a NaN local value never wins, as std::min in the reference ignores
// it).  EnzoMHDIntegratorStageCommands::timestep
// (EnzoMHDIntegratorStageCommands.cpp:299-366).  Grid-stride loop; wave64
// butterfly, then LDS across the workgroup, then one atomic per workgroup.
// Dual-energy parameters of the CFL kernel's synchronisation.  time (device,
// may be null) and t_stop: a device-driven cycle that starts at or past the
// stop time is zero-length and changes nothing: the kernel leaves the
// energies and the pressure field alone and only takes the minimum.
struct DualParams {
  double pressure_floor, eta, ggm1;
  const double *time;
  double t_stop;
};

template <typename FluxT, bool MHD, bool DE = false>
__global__ void __launch_bounds__(256)
k_dt(const BlockDesc<FluxT> *__restrict__ blocks, double gamma,
     unsigned long long *result, DualParams dp) {
  const BlockDesc<FluxT> &b = blocks[blockIdx.y];
  const long n = (long)b.mx * b.my * b.mz;
  double local = 1.7976931348623157e308;
  const bool live = dp.time == nullptr || *dp.time < dp.t_stop;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long)gridDim.x * blockDim.x) {
    double d = b.U[S_RHO][c];
    double vx = b.U[S_VX][c], vy = b.U[S_VY][c], vz = b.U[S_VZ][c];
    double bx = 0., by = 0., bz = 0.;
    if (MHD) {
      bx = b.U[S_BX][c]; by = b.U[S_BY][c]; bz = b.U[S_BZ][c];
    }
    double p;
    if (DE) {
      // the block's energies are synchronised first, in place
      // (EnzoMethodMHDVlct.cpp:563-568), then p = gm1 rho eint
      // (EnzoComputePressure.cpp:127-134)
      const double v[3] = {vx, vy, vz};
      const double b2 = MHD ? (bx * bx + by * by + bz * bz) : 0.;
      double etot = b.U[S_E][c], eint = b.ei[c];
      if (live) {
        floor_and_sync<MHD, true>(d, v, b2, gamma, dp.pressure_floor, dp.eta, dp.ggm1,
                                  etot, eint);
        b.U[S_E][c] = etot;
        b.ei[c] = eint;
      }
      p = (gamma - 1.0) * d * eint;
    } else {
      p = cell_pressure<MHD>(d, vx, vy, vz, b.U[S_E][c], bx, by, bz, gamma);
    }
    if (b.pressure && live) b.pressure[c] = p;
    local = dmin(local, cfl_speed<MHD>(d, vx, vy, vz, bx, by, bz, p, gamma, b.cw));
  }
  block_min_to(result, local);
}

static __global__ void k_dt_init(unsigned long long *result) {
  *result = 0x7FEFFFFFFFFFFFFFull;  // numeric_limits<double>::max()
}

// dt of the cycle on the device: courant * min, clipped so the cycle ends
// at t_stop (control_stopping.cpp:83-86, as the host driver clips), copied
// into dts[0..n) for the compute launches; the time advances by dt.  A
// cycle that starts at or past t_stop is zero-length: the clipped cycle's
// t + (t_stop - t) can round above t_stop, and the difference of the next
// one would then be a (tiny) negative step.
// reset: the minimum word starts again from the largest double once read
// (the cycle's compute then takes the next cycle's minimum into it).
static __global__ void k_dt_finalize(unsigned long long *bits, double courant, double *time,
                                     double t_stop, double *dts, int n, int reset) {
  const double t = *time;
  double dt = __longlong_as_double((long long)*bits) * courant;
  if (t < t_stop && t_stop < t + dt) dt = t_stop - t;
  const double rem = t_stop - t;
  dt = (rem > 0.0) ? (dt < rem ? dt : rem) : 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) dts[i] = dt;
  __syncthreads();
  if (threadIdx.x == 0) {
    *time = t + dt;
    if (reset) *bits = 0x7FEFFFFFFFFFFFFFull;
  }
}

// -------------------------------------------------------------- refresh

// Ghost-zone copy for a uniform grid of equally sized blocks.  One thread
// per element of one array (cell- or face-centred along `fc`); interior
// elements return at once.
struct RefreshGrid {
  int nb[3];     // blocks per axis
  int n[3];      // active cells per block
  int g;         // ghost depth
  int bc[3];     // 0 periodic, 1 outflow
  int face0;     // table index of bfieldi_x (then _y, _z); every other
                 // field is cell-centred
  const double *dt;  // device dt of the cycle, or null: a cycle whose dt is
                     // 0 (past its stop time) refreshes nothing
};

__device__ __forceinline__ bool refresh_axis(const RefreshGrid &G, int a,
                                             int blk, int idx, int centred,
                                             int &sblk, int &sidx) {
  const int n = G.n[a], nb = G.nb[a], g = G.g;
  const int N = n * nb;
  if (!centred) {
    if (idx >= g && idx < g + n) { sblk = blk; sidx = idx; return false; }
    int gi = blk * n + idx - g;
    if (G.bc[a] == 0) {
      gi = ((gi % N) + N) % N;
    } else {
      gi = (gi < 0) ? 0 : ((gi >= N) ? N - 1 : gi);
    }
    sblk = gi / n;
    sidx = gi % n + g;
    return true;
  }
  if (idx >= g && idx <= g + n) { sblk = blk; sidx = idx; return false; }
  int F = blk * n + idx - g;
  if (G.bc[a] == 0) {
    F = ((F % N) + N) % N;
    sblk = F / n;
    sidx = F % n + g;
  } else if (F < 0) {
    sblk = 0; sidx = g;
  } else {
    sblk = nb - 1; sidx = g + n;
  }
  return true;
}

// Decodes the t-th ghost element of an (az, ay, ax) array whose interior is
// [lo, hi) along each axis: z slabs first, then y slabs of the interior z
// range, then x slabs of the interior (z, y) range.  Returns false past the
// last ghost element.
__device__ __forceinline__ bool ghost_coords(long t, const int ext[3],
                                             const int lo[3], const int hi[3],
                                             int &k, int &j, int &i) {
  const int ax = ext[0], ay = ext[1], az = ext[2];
  const int nxi = hi[0] - lo[0], nyi = hi[1] - lo[1], nzi = hi[2] - lo[2];
  const long plane = (long)ay * ax;
  const long cz = (long)(az - nzi) * plane;
  if (t < cz) {
    int pz = (int)(t / plane);
    long r = t - (long)pz * plane;
    k = pz < lo[2] ? pz : hi[2] + (pz - lo[2]);
    j = (int)(r / ax);
    i = (int)(r - (long)j * ax);
    return true;
  }
  t -= cz;
  const long row = (long)(ay - nyi) * ax;
  const long cy = (long)nzi * row;
  if (t < cy) {
    int kk = (int)(t / row);
    long r = t - (long)kk * row;
    int py = (int)(r / ax);
    k = lo[2] + kk;
    j = py < lo[1] ? py : hi[1] + (py - lo[1]);
    i = (int)(r - (long)py * ax);
    return true;
  }
  t -= cy;
  const int gx = ax - nxi;
  const long cx = (long)nzi * nyi * gx;
  if (t >= cx) return false;
  long q = t / gx;
  int px = (int)(t - q * gx);
  k = lo[2] + (int)(q / nyi);
  j = lo[1] + (int)(q % nyi);
  i = px < lo[0] ? px : hi[0] + (px - lo[0]);
  return true;
}

// table[f * nb + b] = device pointer of field f of block b; fields
// face0..face0+2 are the face-centred bfieldi_x..z, the others cell-centred.
// grid: x = ghost elements, y = block, z = field.
static __global__ void __launch_bounds__(256)
k_refresh(double *const *__restrict__ table, RefreshGrid G) {
  if (G.dt != nullptr && *G.dt == 0.0) return;
  const int nb = G.nb[0] * G.nb[1] * G.nb[2];
  const int bidx = blockIdx.y;
  const int f = blockIdx.z;
  const int fc = (f >= G.face0 && f < G.face0 + 3) ? f - G.face0 : -1;
  const int bx = bidx % G.nb[0];
  const int by = (bidx / G.nb[0]) % G.nb[1];
  const int bz = bidx / (G.nb[0] * G.nb[1]);
  int ext[3], lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    ext[a] = G.n[a] + 2 * G.g + (fc == a);
    lo[a] = G.g;
    hi[a] = G.g + G.n[a] + (fc == a);
  }
  int k, j, i;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!ghost_coords(t, ext, lo, hi, k, j, i)) return;
  int sb[3], si[3];
  refresh_axis(G, 0, bx, i, fc == 0, sb[0], si[0]);
  refresh_axis(G, 1, by, j, fc == 1, sb[1], si[1]);
  refresh_axis(G, 2, bz, k, fc == 2, sb[2], si[2]);
  const int sbid = (sb[2] * G.nb[1] + sb[1]) * G.nb[0] + sb[0];
  double *const *tab = table + (size_t)f * nb;
  tab[bidx][idx3(k, j, i, ext[1], ext[0])] =
      tab[sbid][idx3(si[2], si[1], si[0], ext[1], ext[0])];
}

// ------------------------------------------------------- flux corrections

// One block's FluxData as a single array (src/Cello/data_FluxData.cpp:
// 57-87): for each field i_f, axis, face the block's face fluxes, then the
// neighbour's, m[axis] values each.  The face array of axis 0 is (nz, ny)
// with y fastest, of axis 1 (nz, nx), of axis 2 (ny, nx)
// (FaceFluxes::flux_array, src/Cello/data_FaceFluxes.cpp:114-126).
struct FluxLayout {
  int n[3];        // active cells along x, y, z
  long m[3];       // face sizes
  long per_field;  // values per field (4 faces' worth per axis)

  __host__ __device__ void init(int nx, int ny, int nz) {
    n[0] = nx; n[1] = ny; n[2] = nz;
    m[0] = (long)ny * nz;
    m[1] = (long)nx * nz;
    m[2] = (long)nx * ny;
    per_field = 4 * (m[0] + m[1] + m[2]);
  }
  __host__ __device__ long offset(int i_f, int axis, int face, int nbr) const {
    long o = (long)i_f * per_field;
    for (int a = 0; a < axis; a++) o += 4 * m[a];
    return o + (long)face * 2 * m[axis] + (nbr ? m[axis] : 0);
  }
  // entry t -> (field, axis, face, neighbour part, index in the face array)
  __host__ __device__ void decode(long t, int &i_f, int &axis, int &face,
                                  int &nbr, long &idx) const {
    i_f = (int)(t / per_field);
    long r = t - (long)i_f * per_field;
    axis = 0;
    while (axis < 2 && r >= 4 * m[axis]) {
      r -= 4 * m[axis];
      axis++;
    }
    face = (int)(r / (2 * m[axis]));
    r -= (long)face * 2 * m[axis];
    nbr = r >= m[axis];
    idx = r - (nbr ? m[axis] : 0);
  }
  // active-region cell (x, y, z), from 0, of entry idx of a face array of
  // axis: its two in-face coordinates; the normal one is left at 0
  __host__ __device__ void face_cell(int axis, long idx, int p[3]) const {
    const int a0 = (axis == 0) ? 1 : 0;          // fastest in-face axis
    const int a1 = (axis == 2) ? 1 : 2;          // slowest in-face axis
    p[axis] = 0;
    p[a0] = (int)(idx % n[a0]);
    p[a1] = (int)(idx / n[a0]);
  }
};

// Fields saved in the FluxData, in the order of the "conserved" group.
struct StoreSlots {
  int nf;
  int slot[8];
};

// After the final stage: dt/dx * F of that stage's fluxes through the six
// faces of the active region (EnzoMethodMHDVlct::save_fluxes_for_
// corrections_, EnzoMethodMHDVlct.cpp:250-330: the faces between cells g-1
// and g, and m-g-1 and m-g) into the block part of the FluxData, and zeros
// into the neighbour part (allocate_FC_flux_buffer_, :333-355, allocates the
// FluxData zeroed every cycle).  One thread per FluxData entry.
template <typename FluxT>
__global__ void __launch_bounds__(256)
k_store_fluxes(const BlockDesc<FluxT> *__restrict__ blocks,
               double *const *__restrict__ fd, StoreSlots sl, double dt,
               const double *__restrict__ dts) {
  const int by = blockIdx.y;
  const BlockDesc<FluxT> &b = blocks[by];
  const int g = b.g;
  FluxLayout L;
  L.init(b.mx - 2 * g, b.my - 2 * g, b.mz - 2 * g);
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.per_field * sl.nf) return;
  int i_f, axis, face, nbr;
  long idx;
  L.decode(t, i_f, axis, face, nbr, idx);
  double v = 0.0;
  if (!nbr) {
    int p[3];
    L.face_cell(axis, idx, p);
    const int m[3] = {b.mx, b.my, b.mz};
    for (int a = 0; a < 3; a++) p[a] += g;
    p[axis] = face ? m[axis] - g - 1 : g - 1;
    const int c = (p[2] * b.my + p[1]) * b.mx + p[0];
    // dt_dxi = dt/cell_width (EnzoMethodMHDVlct.cpp:265)
    const double dt_dxi = (dts ? dts[by] : dt) / b.cw[axis];
    v = dt_dxi * (double)gld(b.F[axis][sl.slot[i_f]], c);
  }
  gst(fd[by], t, v);
}

}  // namespace ENZOE_KNS
