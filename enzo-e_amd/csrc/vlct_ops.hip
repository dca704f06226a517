// vlct_ops.hip -- single-operation entry points of the C ABI.  Each runs one
// reference plugin call (a reconstructor, a Riemann solver, one CT step) on
// device arrays in the reference's own face-trimmed layout, so the C++
// drop-in classes can be used one call at a time exactly like the CPU ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <string>

#include "../../include/enzoe_vlct.h"
#include "vlct_kernels.hpp"

using namespace enzoe;

namespace {

std::string lower(const char *s) {
  std::string r = s ? s : "";
  std::transform(r.begin(), r.end(), r.begin(),
                 [](unsigned char c) { return (char)std::tolower(c); });
  return r;
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

struct Ptr8 { double *p[8]; };

// Reconstruction on face-trimmed arrays (EnzoReconstructorNN.cpp:14-49,
// EnzoReconstructorPLM.hpp:166-248): one thread per cell of the
// stale-trimmed box.
template <int RECON>
__global__ void k_op_reconstruct(Ptr8 w, Ptr8 wl, Ptr8 wr, int nkeys, int mz,
                                 int my, int mx, int dim, int s, double theta,
                                 double dfloor, double pfloor) {
  const int di = dim == 0, dj = dim == 1, dk = dim == 2;
  const int fy = my - dj, fx = mx - di;
  int k, j, i;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!box_coords(t, s, mx - 2 * s, my - 2 * s, mz - 2 * s, k, j, i)) return;
  const long cs = (dim == 0) ? 1 : (dim == 1 ? (long)mx : (long)mx * my);
  const long c = idx3(k, j, i, my, mx);
  const int p[3] = {i, j, k};
  const int m[3] = {mx, my, mz};
  for (int q = 0; q < nkeys; q++) {
    const double *W = w.p[q];
    if (RECON == RC_NN) {
      if (p[dim] >= m[dim] - s - 1) continue;
      long f = idx3(k, j, i, fy, fx);
      wl.p[q][f] = W[c];
      wr.p[q][f] = W[c + cs];
    } else {
      if (p[dim] < s + 1 || p[dim] >= m[dim] - s - 1) continue;
      double val = W[c];
      double dv = (RECON == RC_PLM) ? limiter_enzo(W[c - cs], val, W[c + cs], theta)
                                    : limiter_athena(W[c - cs], val, W[c + cs]);
      double half_dv = dv * 0.5;
      double rv = val - half_dv, lv = val + half_dv;
      if (q == S_RHO) { rv = dmax(rv, dfloor); lv = dmax(lv, dfloor); }
      if (q == S_E) { rv = dmax(rv, pfloor); lv = dmax(lv, pfloor); }
      wr.p[q][idx3(k - dk, j - dj, i - di, fy, fx)] = rv;
      wl.p[q][idx3(k, j, i, fy, fx)] = lv;
    }
  }
}

// Riemann solve on face-trimmed arrays (EnzoRiemannImpl.hpp:317-338).
template <int RIEMANN, bool MHD>
__global__ void k_op_riemann(Ptr8 pl, Ptr8 pr, Ptr8 fx8, int mz, int my,
                             int mx, int dim, int s, double gamma) {
  const int fz = mz - (dim == 2), fy = my - (dim == 1), fx = mx - (dim == 0);
  int k, j, i;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!box_coords(t, 0, fx, fy, fz, k, j, i)) return;
  if (k < s || j < s || i < s || k >= fz - s || j >= fy - s || i >= fx - s)
    return;
  const long f = idx3(k, j, i, fy, fx);
  const int dj = (dim + 1) % 3, dk = (dim + 2) % 3;
  if (MHD) {
    double l[8], r[8], o[8];
    l[L_D] = pl.p[S_RHO][f]; r[L_D] = pr.p[S_RHO][f];
    l[L_BI] = pl.p[S_BX + dim][f]; r[L_BI] = pr.p[S_BX + dim][f];
    l[L_BJ] = pl.p[S_BX + dj][f]; r[L_BJ] = pr.p[S_BX + dj][f];
    l[L_BK] = pl.p[S_BX + dk][f]; r[L_BK] = pr.p[S_BX + dk][f];
    l[L_VI] = pl.p[S_VX + dim][f]; r[L_VI] = pr.p[S_VX + dim][f];
    l[L_VJ] = pl.p[S_VX + dj][f]; r[L_VJ] = pr.p[S_VX + dj][f];
    l[L_VK] = pl.p[S_VX + dk][f]; r[L_VK] = pr.p[S_VX + dk][f];
    l[L_E] = pl.p[S_E][f]; r[L_E] = pr.p[S_E][f];
    if (RIEMANN == RS_HLLD) hlld(l, r, gamma, o);
    else hlle_mhd(l, r, gamma, o);
    fx8.p[S_RHO][f] = o[L_D];
    fx8.p[S_BX + dim][f] = o[L_BI];
    fx8.p[S_BX + dj][f] = o[L_BJ];
    fx8.p[S_BX + dk][f] = o[L_BK];
    fx8.p[S_VX + dim][f] = o[L_VI];
    fx8.p[S_VX + dj][f] = o[L_VJ];
    fx8.p[S_VX + dk][f] = o[L_VK];
    fx8.p[S_E][f] = o[L_E];
  } else {
    double l[5], r[5], o[5];
    l[H_D] = pl.p[S_RHO][f]; r[H_D] = pr.p[S_RHO][f];
    l[H_VI] = pl.p[S_VX + dim][f]; r[H_VI] = pr.p[S_VX + dim][f];
    l[H_VJ] = pl.p[S_VX + dj][f]; r[H_VJ] = pr.p[S_VX + dj][f];
    l[H_VK] = pl.p[S_VX + dk][f]; r[H_VK] = pr.p[S_VX + dk][f];
    l[H_E] = pl.p[S_E][f]; r[H_E] = pr.p[S_E][f];
    hllc(l, r, gamma, o);
    fx8.p[S_RHO][f] = o[H_D];
    fx8.p[S_VX + dim][f] = o[H_VI];
    fx8.p[S_VX + dj][f] = o[H_VJ];
    fx8.p[S_VX + dk][f] = o[H_VK];
    fx8.p[S_E][f] = o[H_E];
  }
}

struct Ptr3 { double *p[3]; };

// EnzoBfieldMethodCT::update_bfield on the reference's edge-array layout.
__global__ void k_op_update_bfield(Ptr3 edge, const double *bcur, double *bout,
                                   double dtdj, double dtdk, int mz, int my,
                                   int mx, int ax, int s) {
  const int aj = (ax + 1) % 3, ak = (ax + 2) % 3;
  const int m[3] = {mx, my, mz};
  int nb[3] = {mx, my, mz};
  nb[ax] += 1;
  int k, j, i;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!box_coords(t, 0, nb[0], nb[1], nb[2], k, j, i)) return;
  int p[3] = {i, j, k};
  if (p[ax] < s + 1 || p[ax] >= m[ax] - s) return;
  if (p[aj] < s + 1 || p[aj] >= m[aj] - s - 1) return;
  if (p[ak] < s + 1 || p[ak] >= m[ak] - s - 1) return;
  int nEj[3] = {mx, my, mz}; nEj[ak] -= 1; nEj[ax] -= 1;
  int nEk[3] = {mx, my, mz}; nEk[ax] -= 1; nEk[aj] -= 1;
  int pe[3] = {p[0], p[1], p[2]}; pe[ax] -= 1;
  int pjm[3] = {pe[0], pe[1], pe[2]}; pjm[aj] -= 1;
  int pkm[3] = {pe[0], pe[1], pe[2]}; pkm[ak] -= 1;
  auto at = [](const int q[3], const int n[3]) {
    return ((long)q[2] * n[1] + q[1]) * (long)n[0] + q[0];
  };
  double E_k_term = dtdj * (edge.p[ak][at(pe, nEk)] - edge.p[ak][at(pjm, nEk)]);
  double E_j_term = dtdk * (edge.p[aj][at(pe, nEj)] - edge.p[aj][at(pkm, nEj)]);
  long cb = at(p, nb);
  bout[cb] = bcur[cb] - E_k_term + E_j_term;
}

__global__ void k_op_center_bfield(double *bc, const double *bi, int mz,
                                   int my, int mx, int dim, int s) {
  int k, j, i;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!box_coords(t, s, mx - 2 * s, my - 2 * s, mz - 2 * s, k, j, i)) return;
  const int bny = my + (dim == 1), bnx = mx + (dim == 0);
  double l = bi[idx3(k, j, i, bny, bnx)];
  double r = bi[idx3(k + (dim == 2), j + (dim == 1), i + (dim == 0), bny, bnx)];
  bc[idx3(k, j, i, my, mx)] = 0.5 * (l + r);
}

}  // namespace

extern "C" {

int enzoe_reconstructor_staling(const char *name, int *total_rate,
                                int *immediate_rate) {
  std::string n = lower(name);
  if (!total_rate || !immediate_rate) return ENZOE_ERR_ARG;
  if (n == "nn") { *total_rate = 1; *immediate_rate = 0; return ENZOE_OK; }
  if (n == "plm" || n == "plm_enzo" || n == "plm_athena") {
    *total_rate = 2; *immediate_rate = 1; return ENZOE_OK;
  }
  return ENZOE_ERR_CONFIG;
}

int enzoe_reconstruct_interface(const char *name, double theta_limiter,
                                double density_floor, double pressure_floor,
                                double *const prim[8], double *priml[8],
                                double *primr[8], int nkeys, int mz, int my,
                                int mx, int dim, int stale_depth,
                                enzoe_stream_t stream) {
  std::string n = lower(name);
  if (!(theta_limiter >= 1.0 && theta_limiter <= 2.0)) return ENZOE_ERR_CONFIG;
  if (!prim || !priml || !primr || (nkeys != 5 && nkeys != 8) || dim < 0 ||
      dim > 2 || stale_depth < 0 || 2 * stale_depth + 2 >= std::min(mx, std::min(my, mz)))
    return ENZOE_ERR_ARG;
  Ptr8 w{}, l{}, r{};
  for (int q = 0; q < nkeys; q++) {
    if (!prim[q] || !priml[q] || !primr[q]) return ENZOE_ERR_ARG;
    w.p[q] = prim[q]; l.p[q] = priml[q]; r.p[q] = primr[q];
  }
  hipStream_t s = (hipStream_t)stream;
  long nt = (long)(mx - 2 * stale_depth) * (my - 2 * stale_depth) * (mz - 2 * stale_depth);
  if (n == "nn") {
    hipLaunchKernelGGL(k_op_reconstruct<RC_NN>, dim3(nblk(nt)), dim3(256), 0, s, w, l, r,
                       nkeys, mz, my, mx, dim, stale_depth, theta_limiter,
                       density_floor, pressure_floor);
  } else if (n == "plm" || n == "plm_enzo") {
    hipLaunchKernelGGL(k_op_reconstruct<RC_PLM>, dim3(nblk(nt)), dim3(256), 0, s, w, l, r,
                       nkeys, mz, my, mx, dim, stale_depth, theta_limiter,
                       density_floor, pressure_floor);
  } else if (n == "plm_athena") {
    hipLaunchKernelGGL(k_op_reconstruct<RC_PLM_ATHENA>, dim3(nblk(nt)), dim3(256), 0, s, w,
                       l, r, nkeys, mz, my, mx, dim, stale_depth, theta_limiter,
                       density_floor, pressure_floor);
  } else {
    return ENZOE_ERR_CONFIG;
  }
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int enzoe_riemann_solve(const char *solver, int mhd, double gamma,
                        double *const priml[8], double *const primr[8],
                        double *flux[8], int mz, int my, int mx, int dim,
                        int stale_depth, enzoe_stream_t stream) {
  std::string n = lower(solver);
  if (!priml || !primr || !flux || dim < 0 || dim > 2 || stale_depth < 0)
    return ENZOE_ERR_ARG;
  const int nk = mhd ? 8 : 5;
  Ptr8 l{}, r{}, f{};
  for (int q = 0; q < nk; q++) {
    if (!priml[q] || !primr[q] || !flux[q]) return ENZOE_ERR_ARG;
    l.p[q] = priml[q]; r.p[q] = primr[q]; f.p[q] = flux[q];
  }
  hipStream_t s = (hipStream_t)stream;
  long nt = (long)(mx - (dim == 0)) * (my - (dim == 1)) * (mz - (dim == 2));
  if (n == "hlld" && mhd) {
    hipLaunchKernelGGL((k_op_riemann<RS_HLLD, true>), dim3(nblk(nt)), dim3(256), 0, s, l,
                       r, f, mz, my, mx, dim, stale_depth, gamma);
  } else if (n == "hlle" && mhd) {
    hipLaunchKernelGGL((k_op_riemann<RS_HLLE, true>), dim3(nblk(nt)), dim3(256), 0, s, l,
                       r, f, mz, my, mx, dim, stale_depth, gamma);
  } else if (n == "hllc" && !mhd) {
    hipLaunchKernelGGL((k_op_riemann<RS_HLLC, false>), dim3(nblk(nt)), dim3(256), 0, s, l,
                       r, f, mz, my, mx, dim, stale_depth, gamma);
  } else {
    // EnzoRiemann::construct_riemann: HLLD needs B, HLLC cannot take B,
    // hydro HLLE and Davis "hll" are not available
    return ENZOE_ERR_CONFIG;
  }
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int enzoe_ct_update_bfield(const double cell_width[3], int dim,
                           double *const edge[3], const double *bfieldi_cur,
                           double *bfieldi_out, double dt, int mz, int my,
                           int mx, int stale_depth, enzoe_stream_t stream) {
  if (!cell_width || !edge || !bfieldi_cur || !bfieldi_out || dim < 0 || dim > 2)
    return ENZOE_ERR_ARG;
  Ptr3 e{};
  for (int a = 0; a < 3; a++) {
    if (!edge[a]) return ENZOE_ERR_ARG;
    e.p[a] = edge[a];
  }
  const int aj = (dim + 1) % 3, ak = (dim + 2) % 3;
  double dtdj = dt / cell_width[aj];
  double dtdk = dt / cell_width[ak];
  long nt = (long)(mx + (dim == 0)) * (my + (dim == 1)) * (mz + (dim == 2));
  hipLaunchKernelGGL(k_op_update_bfield, dim3(nblk(nt)), dim3(256), 0,
                     (hipStream_t)stream, e, bfieldi_cur, bfieldi_out, dtdj,
                     dtdk, mz, my, mx, dim, stale_depth);
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int enzoe_ct_compute_center_bfield(int dim, double *bfieldc,
                                   const double *bfieldi, int mz, int my,
                                   int mx, int stale_depth,
                                   enzoe_stream_t stream) {
  if (!bfieldc || !bfieldi || dim < 0 || dim > 2 || stale_depth < 0)
    return ENZOE_ERR_ARG;
  long nt = (long)(mx - 2 * stale_depth) * (my - 2 * stale_depth) * (mz - 2 * stale_depth);
  hipLaunchKernelGGL(k_op_center_bfield, dim3(nblk(nt)), dim3(256), 0,
                     (hipStream_t)stream, bfieldc, bfieldi, mz, my, mx, dim,
                     stale_depth);
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

}  // extern "C"
