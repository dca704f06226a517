// vlct_ct.hpp -- constrained transport and the conserved update of one VL+CT
// stage as two z-marching kernels:
//
//   k_ctface  : edge EMFs of one (x, y) tile per plane, exchanged through
//               LDS, and the face-B update of the tile's faces (edges and
//               face update of EnzoBfieldMethodCT::update_all_bfield_
//               components, EnzoBfieldMethodCT.cpp:220-263, 384-687);
//   k_cells_z : cell-centred B from the faces, flux divergence, conserved
//               update and floors (as k_cells) down a column of cells.
//
// A workgroup walks its tile or columns through a chunk of consecutive z
// planes, so the z-neighbours a stencil touches were loaded by the same CU
// one plane earlier (L1/L2 hits instead of re-fetches from beyond L2), and
// the edge EMFs never go to HBM.  Every value is computed with the same
// expression as in k_edge / k_faces / k_cells, so results are identical.
#pragma once

#include "vlct_kernels.hpp"

namespace enzoe {

// Edge EMF along `ax` at (cell p_ax, faces p_j + 1/2, p_k + 1/2) with the
// cell-shaped index c of p (EnzoBfieldMethodCT.cpp:384-557; centre E from
// the stage's state, :267-292).
template <typename FluxT>
__device__ __forceinline__ double edge_emf(const BlockDesc<FluxT> &b,
                                           double *const *U, int ax, long c,
                                           const long str[3]) {
  const int aj = (ax + 1) % 3, ak = (ax + 2) % 3;
  const long sj = str[aj], sk = str[ak];
  const double *vj = U[S_VX + aj], *vk = U[S_VX + ak];
  const double *bj = U[S_BX + aj], *bk = U[S_BX + ak];
  auto ecen = [&](long cc) { return (-vj[cc] * bk[cc] + vk[cc] * bj[cc]); };
  const double Ec = ecen(c), Ec_jp1 = ecen(c + sj), Ec_kp1 = ecen(c + sk),
               Ec_jkp1 = ecen(c + sj + sk);
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

// Launch shape of k_ctface: tiles of tx * ty threads (tx * ty <= 256) over
// the face box [lo, lo + n) in x and y, z chunks of zc planes over
// [z0, z0 + nz).
struct CtTile {
  int tx, ty, ntx, nty;
  int zc, nzc;
  int x0, y0, z0;
  int nx, ny, nz;
};

constexpr int kCtLds = 360;  // >= (tx + 1) * (ty + 1) for tx * ty <= 256

// Faces of one stage (box of the reference's update_bfield):
//   B_x at X in [cs+1, mx-cs), y, z in [cs+1, m-cs-1);  likewise B_y, B_z.
// Plane z of the march updates B_x(z), B_y(z) (edges of planes z and z-1)
// and B_z(z+1) (edges of plane z); the tile's edges plus the halo at
// x0 - 1 and y0 - 1 are computed per plane into LDS.
template <typename FluxT>
__global__ void __launch_bounds__(256)
k_ctface(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp, CtTile T) {
  __shared__ double sEx[3][kCtLds];
  __shared__ double sEy[3][kCtLds];
  __shared__ double sEz[2][kCtLds];
  const BlockDesc<FluxT> &b = blocks[blockIdx.y];
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int cs = sp.cur_stale;
  int w = blockIdx.x;
  const int per_chunk = T.ntx * T.nty;
  const int zci = w / per_chunk;
  w -= zci * per_chunk;
  const int tyi = w / T.ntx, txi = w - tyi * T.ntx;
  const int lxo = threadIdx.x % T.tx, lyo = threadIdx.x / T.tx;
  const int xb = T.x0 + txi * T.tx, yb = T.y0 + tyi * T.ty;
  const int x = xb + lxo, y = yb + lyo;
  const bool active = lyo < T.ty && x < T.x0 + T.nx && y < T.y0 + T.ny;
  const int W = T.tx + 1;
  const int zs = T.z0 + zci * T.zc;
  const int ze = min(zs + T.zc, T.z0 + T.nz);
  double *const *U = (sp.stage == 0) ? b.U : b.Ut;
  double *const *out = sp.is_final ? b.bfi : b.bft;
  const long str[3] = {1, (long)mx, (long)mx * my};
  const int m[3] = {mx, my, mz};
  const double dt = sp.dts ? sp.dts[blockIdx.y] * sp.dt_factor : sp.dt;
  const double dtdx = dt / b.cw[0], dtdy = dt / b.cw[1], dtdz = dt / b.cw[2];

  auto valid_edge = [&](int ax, const int p[3]) {
    const int aj = (ax + 1) % 3, ak = (ax + 2) % 3;
    return p[ax] >= cs + 1 && p[ax] < m[ax] - cs - 1 && p[aj] >= cs &&
           p[aj] < m[aj] - cs - 1 && p[ak] >= cs && p[ak] < m[ak] - cs - 1;
  };
  auto put = [&](int px, int py, int pz, int l, int b3, int b2, bool ez) {
    const int p[3] = {px, py, pz};
    const long c = idx3(pz, py, px, my, mx);
    sEx[b3][l] = valid_edge(0, p) ? edge_emf(b, U, 0, c, str) : 0.;
    sEy[b3][l] = valid_edge(1, p) ? edge_emf(b, U, 1, c, str) : 0.;
    if (ez) sEz[b2][l] = valid_edge(2, p) ? edge_emf(b, U, 2, c, str) : 0.;
  };

  const int l = (lyo + 1) * W + (lxo + 1);
  for (int z = zs - 1; z < ze; z++) {
    const int rel = z - (zs - 1);
    const int b3 = rel % 3, b2 = rel & 1;
    const bool ez = z >= zs;  // plane zs - 1 only feeds B_x, B_y of plane zs
    if (active && z >= cs) {
      put(x, y, z, l, b3, b2, ez);
      if (lxo == 0) put(x - 1, y, z, l - 1, b3, b2, ez);
      if (lyo == 0) put(x, y - 1, z, l - W, b3, b2, ez);
      if (lxo == 0 && lyo == 0) put(x - 1, y - 1, z, l - W - 1, b3, b2, ez);
    }
    __syncthreads();
    if (!active || z < zs) continue;
    const int p3 = (b3 + 2) % 3;
    const int lxm = l - 1, lym = l - W, lxym = l - W - 1;
    const bool zin = z >= cs + 1 && z < mz - cs - 1;
    if (zin && x >= cs + 1 && x < mx - cs && y >= cs + 1 && y < my - cs - 1) {
      const double E_k_term = dtdy * (sEz[b2][lxm] - sEz[b2][lxym]);
      const double E_j_term = dtdz * (sEy[b3][lxm] - sEy[p3][lxm]);
      const long cb = idx3(z, y, x, my, mx + 1);
      out[0][cb] = b.bfi[0][cb] - E_k_term + E_j_term;
    }
    if (zin && y >= cs + 1 && y < my - cs && x >= cs + 1 && x < mx - cs - 1) {
      const double E_k_term = dtdz * (sEx[b3][lym] - sEx[p3][lym]);
      const double E_j_term = dtdx * (sEz[b2][lym] - sEz[b2][lxym]);
      const long cb = idx3(z, y, x, my + 1, mx);
      out[1][cb] = b.bfi[1][cb] - E_k_term + E_j_term;
    }
    if (z >= cs && z < mz - cs - 1 && x >= cs + 1 && x < mx - cs - 1 &&
        y >= cs + 1 && y < my - cs - 1) {
      const double E_k_term = dtdx * (sEy[b3][l] - sEy[b3][lxm]);
      const double E_j_term = dtdy * (sEx[b3][l] - sEx[b3][lym]);
      const long cb = idx3(z + 1, y, x, my, mx);
      out[2][cb] = b.bfi[2][cb] - E_k_term + E_j_term;
    }
  }
}

// Launch shape of k_cells_z: one thread per column (x, y) of [lo, m - lo)
// (linear, x fastest), chunks of zc planes of [lo, mz - lo).
struct ColShape {
  int lo, ncol_x, ncol_y, ncol;
  int zc, nzc;
};

// k_cells on a column (same regions and expressions): the z-neighbours it
// reads -- the z flux below the cell and the B_z face above it -- are the
// values it loaded for the previous plane.
template <typename FluxT, bool MHD>
__global__ void __launch_bounds__(256)
k_cells_z(const BlockDesc<FluxT> *__restrict__ blocks, StageParams sp, ColShape S) {
  const BlockDesc<FluxT> &b = blocks[blockIdx.y];
  const int mx = b.mx, my = b.my, mz = b.mz;
  const int cs = sp.cur_stale;
  const int lo = S.lo;
  const int per_chunk = (S.ncol + 255) / 256;
  const int zci = blockIdx.x / per_chunk;
  const int t = (blockIdx.x - zci * per_chunk) * 256 + threadIdx.x;
  if (t >= S.ncol) return;
  const int yy = t / S.ncol_x;
  const int x = lo + (t - yy * S.ncol_x), y = lo + yy;
  const int zs = lo + zci * S.zc;
  const int ze = min(zs + S.zc, mz - lo);
  double *const *Uo = sp.is_final ? b.U : b.Ut;
  double *const *bf = sp.is_final ? b.bfi : b.bft;
  const double dt = sp.dts ? sp.dts[blockIdx.y] * sp.dt_factor : sp.dt;
  const long str[3] = {1, (long)mx, (long)mx * my};
  const int fq[5] = {S_RHO, S_VX, S_VY, S_VZ, S_E};
  const int ul = cs + 1;
  const bool col_in = x >= ul && y >= ul && x < mx - ul && y < my - ul;
  const double dtdx[3] = {dt / b.cw[0], dt / b.cw[1], dt / b.cw[2]};
  double fzl[5];  // z flux of the face below the cell
  double bz_lo = 0.;
  {
    const long c0 = idx3(zs, y, x, my, mx);
    // the face below plane zs (plane 0 has none and is never updated)
    const long cm = (zs >= 1) ? c0 - str[2] : c0;
#pragma unroll
    for (int q = 0; q < 5; q++) fzl[q] = (double)b.F[2][fq[q]][cm];
    if (MHD) bz_lo = bf[2][c0];
  }
  for (int z = zs; z < ze; z++) {
    const long c = idx3(z, y, x, my, mx);
    double bc[3] = {0., 0., 0.};
    if (MHD) {
      const double bz_hi = bf[2][c + str[2]];
      bc[0] = 0.5 * (bf[0][idx3(z, y, x, my, mx + 1)] +
                     bf[0][idx3(z, y, x + 1, my, mx + 1)]);
      bc[1] = 0.5 * (bf[1][idx3(z, y, x, my + 1, mx)] +
                     bf[1][idx3(z, y + 1, x, my + 1, mx)]);
      bc[2] = 0.5 * (bz_lo + bz_hi);
      bz_lo = bz_hi;
      Uo[S_BX][c] = bc[0];
      Uo[S_BY][c] = bc[1];
      Uo[S_BZ][c] = bc[2];
    }
    double fzr[5];
#pragma unroll
    for (int q = 0; q < 5; q++) fzr[q] = (double)b.F[2][fq[q]][c];
    if (col_in && z >= ul && z < mz - ul) {
      double dU[5];
#pragma unroll
      for (int q = 0; q < 5; q++) dU[q] = 0.;
#pragma unroll
      for (int d = 0; d < 2; d++) {
#pragma unroll
        for (int q = 0; q < 5; q++) {
          const FluxT *F = b.F[d][fq[q]];
          double fr = (double)F[c], fl = (double)F[c - str[d]];
          dU[q] -= dtdx[d] * (fr - fl);
        }
      }
#pragma unroll
      for (int q = 0; q < 5; q++) dU[q] -= dtdx[2] * (fzr[q] - fzl[q]);
      double old_rho = b.U[S_RHO][c];
      double new_rho = b.U[S_RHO][c] + dU[0];
      new_rho = dmax(new_rho, sp.density_floor);
      double inv_new_rho = 1. / new_rho;
      double v[3];
#pragma unroll
      for (int q = 1; q < 4; q++)
        v[q - 1] = (b.U[fq[q]][c] * old_rho + dU[q]) * inv_new_rho;
      double etot = (b.U[S_E][c] * old_rho + dU[4]) * inv_new_rho;
      double inv_rho = 1. / new_rho;
      double eint_floor = sp.pressure_floor * (1. / (sp.gamma - 1.)) * inv_rho;
      double v2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      double non_thermal_e = 0.5 * v2;
      if (MHD) {
        double b2 = (bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
        non_thermal_e += (0.5 * b2 * inv_rho);
      }
      double etot_floor = eint_floor + non_thermal_e;
      etot = dmax(etot, etot_floor);
      Uo[S_RHO][c] = new_rho;
      Uo[S_VX][c] = v[0];
      Uo[S_VY][c] = v[1];
      Uo[S_VZ][c] = v[2];
      Uo[S_E][c] = etot;
    }
#pragma unroll
    for (int q = 0; q < 5; q++) fzl[q] = fzr[q];
  }
}

}  // namespace enzoe
