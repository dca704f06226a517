// vlct_ops.hip -- single-operation entry points of the C ABI.  Each runs one
// reference plugin call (a reconstructor, a Riemann solver, one CT step, one
// quantity-update step) on device views with arbitrary strides, so the C++
// drop-in classes can pass the reference's own views (stale-trimmed,
// face-trimmed with cell-centred strides, 4-D backing arrays) unchanged.
//
// Index convention of every kernel: a position p = (x, y, z) is taken in the
// coordinates of each view (element (0,0,0) of the view is its corner); the
// stale-depth regions are those of the reference loops, which all address
// their views through `get(key, stale_depth)` / `subarray(stale_slc, ...)`
// (src/Cello/view_ViewMap.hpp:310-327) -- an offset common to every view.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <cctype>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/enzoe_vlct.h"
#include "vlct_kernels.hpp"
#include "vlct_launch.hpp"

using namespace enzoe;

namespace {

std::string lower(const char *s) {
  std::string r = s ? s : "";
  std::transform(r.begin(), r.end(), r.begin(),
                 [](unsigned char c) { return (char)std::tolower(c); });
  return r;
}

// View of one array, axes in (x, y, z) order.
struct KArr {
  double *p;
  long s[3];
  int n[3];
};

// Same-shape views of a map, indexed by slot (ENZOE_SLOT_*).
struct KMap {
  double *p[ENZOE_MAP_SLOTS];
  long s[3];
  int n[3];
};

// Iteration box [lo, lo + n) in (x, y, z).
struct Box {
  int lo[3];
  int n[3];
};

__device__ __forceinline__ long off(const long s[3], const int p[3]) {
  return p[2] * s[2] + p[1] * s[1] + p[0] * s[0];
}

__device__ __forceinline__ bool box_pos(const Box &b, int p[3]) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nxy = (long)b.n[0] * b.n[1];
  if (t >= nxy * b.n[2]) return false;
  int z = (int)(t / nxy);
  long r = t - (long)z * nxy;
  int y = (int)(r / b.n[0]);
  int x = (int)(r - (long)y * b.n[0]);
  p[0] = x + b.lo[0];
  p[1] = y + b.lo[1];
  p[2] = z + b.lo[2];
  return true;
}

// ------------------------------------------------------------- host helpers

bool to_karr(const enzoe_farray *a, KArr &k) {
  if (!a || !a->data) return false;
  k.p = a->data;
  for (int d = 0; d < 3; d++) {
    k.n[d] = a->shape[2 - d];
    k.s[d] = (long)a->stride[2 - d];
    if (k.n[d] < 0) return false;
  }
  return true;
}

// required: bitmask of slots that must be present
bool to_kmap(const enzoe_fmap *m, unsigned required, KMap &k) {
  if (!m) return false;
  for (int q = 0; q < ENZOE_MAP_SLOTS; q++) {
    k.p[q] = m->data[q];
    if ((required >> q & 1u) && !k.p[q]) return false;
  }
  for (int d = 0; d < 3; d++) {
    k.n[d] = m->shape[2 - d];
    k.s[d] = (long)m->stride[2 - d];
    if (k.n[d] < 0) return false;
  }
  return true;
}

unsigned present(const KMap &k) {
  unsigned r = 0;
  for (int q = 0; q < ENZOE_MAP_SLOTS; q++)
    if (k.p[q]) r |= 1u << q;
  return r;
}

bool same_shape(const int a[3], const int b[3]) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// true if `face` is `cell` with the extent along dim changed by `delta`
bool shifted(const int face[3], const int cell[3], int dim, int delta) {
  for (int d = 0; d < 3; d++)
    if (face[d] != cell[d] + (d == dim ? delta : 0)) return false;
  return true;
}

Box make_box(const int lo[3], const int hi[3], long &count) {
  Box b;
  count = 1;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = lo[d];
    b.n[d] = std::max(0, hi[d] - lo[d]);
    count *= b.n[d];
  }
  return b;
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

int launched() {
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

constexpr unsigned kMhdSlots = 0xFFu;
constexpr unsigned kHydroSlots = 0x1Fu;
constexpr unsigned kUpdateSlots = 0x1Fu;  // density, velocity, total_energy
constexpr unsigned kPassiveSlots = 0xFF00u;
constexpr unsigned kEintSlot = 1u << ENZOE_SLOT_INTERNAL_ENERGY;

// ------------------------------------------------------------------ kernels

// EnzoReconstructorNN::reconstruct_interface (EnzoReconstructorNN.cpp:14-49):
// over cells p of [s, n-s) with p_dim < n_dim - s - 1:
//   wl(p) = w(p), wr(p) = w(p + e_dim).
// EnzoReconstructorPLM<L>::reconstruct_interface
// (EnzoReconstructorPLM.hpp:166-248): over cells c with c_dim in
// [s+1, n_dim-s-1): dv = L(w(c-1), w(c), w(c+1)); wr(c - e_dim) =
// floor(w - dv/2), wl(c) = floor(w + dv/2), floors on density and pressure.
template <int RECON>
__global__ void k_op_reconstruct(KMap w, KMap wl, KMap wr, unsigned slots,
                                 Box box, int dim, int s, double theta,
                                 double dfloor, double pfloor) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long c = off(w.s, p);
  const long sd = w.s[dim];
  if (RECON == RC_NN) {
    if (p[dim] >= w.n[dim] - s - 1) return;
    const long fl = off(wl.s, p), fr = off(wr.s, p);
    for (int q = 0; q < ENZOE_MAP_SLOTS; q++) {
      if (!(slots >> q & 1u)) continue;
      wl.p[q][fl] = w.p[q][c];
      wr.p[q][fr] = w.p[q][c + sd];
    }
  } else {
    if (p[dim] < s + 1 || p[dim] >= w.n[dim] - s - 1) return;
    const long fl = off(wl.s, p), fr = off(wr.s, p) - wr.s[dim];
    for (int q = 0; q < ENZOE_MAP_SLOTS; q++) {
      if (!(slots >> q & 1u)) continue;
      const double *W = w.p[q];
      double val = W[c];
      double dv = (RECON == RC_PLM) ? limiter_enzo(W[c - sd], val, W[c + sd], theta)
                                    : limiter_athena(W[c - sd], val, W[c + sd]);
      double half_dv = dv * 0.5;
      double rv = val - half_dv, lv = val + half_dv;
      if (q == ENZOE_SLOT_DENSITY) { rv = dmax(rv, dfloor); lv = dmax(lv, dfloor); }
      if (q == ENZOE_SLOT_ENERGY) { rv = dmax(rv, pfloor); lv = dmax(lv, pfloor); }
      wr.p[q][fr] = rv;
      wl.p[q][fl] = lv;
    }
  }
}

// EnzoRiemannImpl<K>::solve_ (EnzoRiemannImpl.hpp:317-338): one face per
// thread over [s, n-s) of the flux view; states permuted into the i/j/k
// order of the kernels (EnzoRiemannLUT, KernelConfig).
// DE: the dual-energy outputs -- the internal-energy flux into the flux
// map's internal_energy slot and vi_bar into `vel` (EnzoRiemannImpl.hpp:
// 270-301 with calculate_internal_energy_flux_).
template <int RIEMANN, bool MHD, bool DE = false>
__global__ void k_op_riemann(KMap pl, KMap pr, KMap fx8, Box box, int dim,
                             double gamma, unsigned pmask, KArr vel = KArr()) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long a = off(pl.s, p), b = off(pr.s, p), f = off(fx8.s, p);
  const int dj = (dim + 1) % 3, dk = (dim + 2) % 3;
  double vbar = 0.;
  double *vb = DE ? &vbar : nullptr;
  double eint_l = 0., eint_r = 0.;
  if (MHD) {
    const int src[8] = {S_RHO, S_BX + dim, S_BX + dj, S_BX + dk,
                        S_VX + dim, S_VX + dj, S_VX + dk, S_E};
    double l[8], r[8], o[8];
    for (int q = 0; q < 8; q++) { l[q] = pl.p[src[q]][a]; r[q] = pr.p[src[q]][b]; }
    if (DE) {
      eint_l = specific_eint(l[L_D], l[L_E], gamma);
      eint_r = specific_eint(r[L_D], r[L_E], gamma);
    }
    if (RIEMANN == RS_HLLD) hlld(l, r, gamma, o, vb);
    else hlle_mhd(l, r, gamma, o, vb);
    for (int q = 0; q < 8; q++) fx8.p[src[q]][f] = o[q];
  } else {
    const int src[5] = {S_RHO, S_VX + dim, S_VX + dj, S_VX + dk, S_E};
    double l[5], r[5], o[5];
    for (int q = 0; q < 5; q++) { l[q] = pl.p[src[q]][a]; r[q] = pr.p[src[q]][b]; }
    if (DE) {
      eint_l = specific_eint(l[H_D], l[H_E], gamma);
      eint_r = specific_eint(r[H_D], r[H_E], gamma);
    }
    hllc(l, r, gamma, o, vb);
    for (int q = 0; q < 5; q++) fx8.p[src[q]][f] = o[q];
  }
  if (DE) {
    fx8.p[ENZOE_SLOT_INTERNAL_ENERGY][f] = upwind_eint_flux(eint_l, eint_r, fx8.p[S_RHO][f]);
    vel.p[off(vel.s, p)] = vbar;
  }
  // passive scalars: upwind mass fraction times the density flux
  // (calc_passive_scalar_flux_, EnzoRiemannUtils.hpp:224-232, branch-free
  // form kept for identical rounding)
  const double dens_flux = fx8.p[S_RHO][f];
  for (int q = ENZOE_SLOT_PASSIVE0; q < ENZOE_MAP_SLOTS; q++) {
    if (!(pmask >> q & 1u)) continue;
    const double wl = pl.p[q][a], wr = pr.p[q][b];
    const double upwind = (double)(dens_flux > 0) * wl + (double)(dens_flux <= 0) * wr;
    fx8.p[q][f] = upwind * dens_flux;
  }
}

// EnzoBfieldMethodCT::correct_reconstructed_bfield
// (EnzoBfieldMethodCT.cpp:122-166): face f of the trimmed states takes
// bfieldi(f + e_dim), over [s, n-s) of the interior-face view.
__global__ void k_op_correct_bfield(KArr bfi, KMap wl, KMap wr, Box box,
                                    int dim) {
  int p[3];
  if (!box_pos(box, p)) return;
  double v = bfi.p[off(bfi.s, p) + bfi.s[dim]];
  wl.p[S_BX + dim][off(wl.s, p)] = v;
  wr.p[S_BX + dim][off(wr.s, p)] = v;
}

// EnzoBfieldMethodCT::identify_upwind (EnzoBfieldMethodCT.cpp:170-216)
__global__ void k_op_upwind(KMap flux, KArr w, Box box) {
  int p[3];
  if (!box_pos(box, p)) return;
  w.p[off(w.s, p)] = upwind_weight(flux.p[S_RHO][off(flux.s, p)]);
}

// EnzoBfieldMethodCT::compute_center_efield (EnzoBfieldMethodCT.cpp:267-292)
__global__ void k_op_center_efield(KArr E, KMap U, Box box, int dim) {
  int p[3];
  if (!box_pos(box, p)) return;
  const int aj = (dim + 1) % 3, ak = (dim + 2) % 3;
  const long c = off(U.s, p);
  E.p[off(E.s, p)] = (-U.p[S_VX + aj][c] * U.p[S_BX + ak][c] +
                      U.p[S_VX + ak][c] * U.p[S_BX + aj][c]);
}

// EnzoBfieldMethodCT::compute_edge_efield + compute_edge_<true>
// (EnzoBfieldMethodCT.cpp:384-557): edge EMF along ax at (cell p_ax, faces
// p_j + 1/2, p_k + 1/2) over p_ax in [s+1, n_ax-s-1), p_j, p_k in
// [s, n-s-1) of the centre-E view.
struct W3 {
  KArr w[3];
};

__global__ void k_op_edge_efield(KArr Ec, KArr edge, KMap jflux, KMap kflux,
                                 W3 W, Box box, int ax) {
  int p[3];
  if (!box_pos(box, p)) return;
  const int aj = (ax + 1) % 3, ak = (ax + 2) % 3;
  const double *Fj = jflux.p[S_BX + ak];  // -E_ax on j faces
  const double *Fk = kflux.p[S_BX + aj];  // +E_ax on k faces
  const long ej = off(jflux.s, p), ek = off(kflux.s, p), ec = off(Ec.s, p);
  const long wj = off(W.w[aj].s, p), wk = off(W.w[ak].s, p);
  const double *Wj = W.w[aj].p, *Wk = W.w[ak].p;
  const double vEj = Fj[ej], vEj_kp1 = Fj[ej + jflux.s[ak]];
  const double vEk = Fk[ek], vEk_jp1 = Fk[ek + kflux.s[aj]];
  const double vEc = Ec.p[ec], vEc_jp1 = Ec.p[ec + Ec.s[aj]],
               vEc_kp1 = Ec.p[ec + Ec.s[ak]],
               vEc_jkp1 = Ec.p[ec + Ec.s[aj] + Ec.s[ak]];
  const double vWj = Wj[wj], vWj_kp1 = Wj[wj + W.w[aj].s[ak]];
  const double vWk = Wk[wk], vWk_jp1 = Wk[wk + W.w[ak].s[aj]];
  double dEdj_r = vWk_jp1 * (vEc_jp1 + vEj) + (1 - vWk_jp1) * (vEc_jkp1 + vEj_kp1);
  double dEdj_l = vWk * (-vEj - vEc) + (1 - vWk) * (-vEj_kp1 - vEc_kp1);
  double dEdk_r = vWj_kp1 * (vEc_kp1 - vEk) + (1 - vWj_kp1) * (vEc_jkp1 - vEk_jp1);
  double dEdk_l = vWj * (vEk - vEc) + (1 - vWj) * (vEk_jp1 - vEc_jp1);
  double Ej_sum = vEj + vEj_kp1;
  Ej_sum *= -1;
  double Ek_sum = vEk + vEk_jp1;
  edge.p[off(edge.s, p)] =
      0.25 * (Ej_sum + Ek_sum + (dEdj_l - dEdj_r) + (dEdk_l - dEdk_r));
}

// EnzoBfieldMethodCT::update_bfield (EnzoBfieldMethodCT.cpp:617-687): face
// X = p_ax in [s+1, nb_ax-s-1), p_j, p_k in [s+1, nb-s-1) of the face view;
// the edges are taken at ax-position X-1.
__global__ void k_op_update_bfield(KArr Ej, KArr Ek, KArr bcur, KArr bout,
                                   Box box, int ax, double dtdj, double dtdk) {
  int p[3];
  if (!box_pos(box, p)) return;
  const int aj = (ax + 1) % 3, ak = (ax + 2) % 3;
  int pe[3] = {p[0], p[1], p[2]};
  pe[ax] -= 1;
  const long k0 = off(Ek.s, pe), j0 = off(Ej.s, pe);
  double E_k_term = dtdj * (Ek.p[k0] - Ek.p[k0 - Ek.s[aj]]);
  double E_j_term = dtdk * (Ej.p[j0] - Ej.p[j0 - Ej.s[ak]]);
  bout.p[off(bout.s, p)] = bcur.p[off(bcur.s, p)] - E_k_term + E_j_term;
}

// EnzoBfieldMethodCT::compute_center_bfield (EnzoBfieldMethodCT.cpp:702-728)
__global__ void k_op_center_bfield(KArr bc, KArr bi, Box box, int dim) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long f = off(bi.s, p);
  bc.p[off(bc.s, p)] = 0.5 * (bi.p[f] + bi.p[f + bi.s[dim]]);
}

// EnzoIntegrationQuanUpdate::clear_dUcons_map (EnzoIntegrationQuanUpdate.cpp:
// 83-101), a kernel rather than a memset so it may sit in a captured graph.
__global__ void k_op_fill(KMap m, unsigned slots, Box box, double v) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long c = off(m.s, p);
  for (int q = 0; q < ENZOE_MAP_SLOTS; q++)
    if (slots >> q & 1u) m.p[q][c] = v;
}

// EnzoIntegrationQuanUpdate::accumulate_flux_component
// (EnzoIntegrationQuanUpdate.cpp:105-147): cells with p_dim in
// [s+1, n_dim-s-1), [s, n-s) transverse: dU(c) -= dtdx (F(c) - F(c - e_dim)).
__global__ void k_op_accumulate(KMap F, KMap dU, unsigned slots, Box box,
                                int dim, double dtdx) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long c = off(dU.s, p), fr = off(F.s, p), fl = fr - F.s[dim];
  for (int q = 0; q < ENZOE_MAP_SLOTS; q++)
    if (slots >> q & 1u) dU.p[q][c] -= dtdx * (F.p[q][fr] - F.p[q][fl]);
}

// EnzoIntegrationQuanUpdate::update_quantities (:183-238) over [s+1, n-s-1)
// and EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync
// (EnzoPhysicsFluidProps.cpp:162-290, dual energy off) at stale s+1, which
// uses the cell-centred B already written into the output map.
__global__ void k_op_update(KMap U0, KMap dU, KMap Uo, Box box, int mhd,
                            double gamma, double dfloor, double pfloor,
                            unsigned pmask) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long a = off(U0.s, p), d = off(dU.s, p), o = off(Uo.s, p);
  // passive scalars in conserved form (update_passive_scalars_,
  // EnzoIntegrationQuanUpdate.cpp:242-268)
  for (int q = ENZOE_SLOT_PASSIVE0; q < ENZOE_MAP_SLOTS; q++)
    if (pmask >> q & 1u) Uo.p[q][o] = U0.p[q][a] + dU.p[q][d];
  double old_rho = U0.p[S_RHO][a];
  double new_rho = dmax(U0.p[S_RHO][a] + dU.p[S_RHO][d], dfloor);
  double inv_new_rho = 1. / new_rho;
  double v[4];
  for (int q = 1; q < 5; q++)
    v[q - 1] = (U0.p[q][a] * old_rho + dU.p[q][d]) * inv_new_rho;
  Uo.p[S_RHO][o] = new_rho;
  for (int q = 1; q < 5; q++) Uo.p[q][o] = v[q - 1];
  double inv_rho = 1. / new_rho;
  double eint_floor = pfloor * (1. / (gamma - 1.)) * inv_rho;
  double non_thermal_e = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (mhd) {
    double bx = Uo.p[S_BX][o], by = Uo.p[S_BY][o], bz = Uo.p[S_BZ][o];
    non_thermal_e += (0.5 * (bx * bx + by * by + bz * bz) * inv_rho);
  }
  double etot_floor = eint_floor + non_thermal_e;
  Uo.p[S_E][o] = dmax(v[3], etot_floor);
}

// EnzoSourceInternalEnergy::calculate_source
// (EnzoSourceInternalEnergy.cpp:16-92): cells with p_dim in [s+1, n-s-1),
// [s, n-s) across: deint(c) -= dtdx max(p(c), p_floor) (vbar(c+1/2) -
// vbar(c-1/2)); vbar is face-trimmed along dim (entry p is face p+1/2).
__global__ void k_op_source_eint(KMap W, KMap dU, KArr V, Box box, int dim,
                                 double dtdx, double pfloor) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long vr = off(V.s, p), vl = vr - V.s[dim];
  const double pc = dmax(W.p[S_E][off(W.s, p)], pfloor);
  dU.p[ENZOE_SLOT_INTERNAL_ENERGY][off(dU.s, p)] -= dtdx * pc * (V.p[vr] - V.p[vl]);
}

// update_quantities with the internal energy as a specific quantity and the
// dual-energy synchronisation (EnzoIntegrationQuanUpdate.cpp:183-238,
// EnzoPhysicsFluidProps.cpp:158-290 at stale s+1).
template <bool MHD>
__global__ void k_op_update_de(KMap U0, KMap dU, KMap Uo, Box box, double gamma,
                               double dfloor, double pfloor, double eta, double ggm1,
                               unsigned pmask) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long a = off(U0.s, p), d = off(dU.s, p), o = off(Uo.s, p);
  for (int q = ENZOE_SLOT_PASSIVE0; q < ENZOE_MAP_SLOTS; q++)
    if (pmask >> q & 1u) Uo.p[q][o] = U0.p[q][a] + dU.p[q][d];
  const int E = ENZOE_SLOT_INTERNAL_ENERGY;
  double old_rho = U0.p[S_RHO][a];
  double new_rho = dmax(U0.p[S_RHO][a] + dU.p[S_RHO][d], dfloor);
  double inv_new_rho = 1. / new_rho;
  double v[3];
  for (int q = 1; q < 4; q++)
    v[q - 1] = (U0.p[q][a] * old_rho + dU.p[q][d]) * inv_new_rho;
  double etot = (U0.p[S_E][a] * old_rho + dU.p[S_E][d]) * inv_new_rho;
  double eint = (U0.p[E][a] * old_rho + dU.p[E][d]) * inv_new_rho;
  double b2 = 0.;
  if (MHD) {
    double bx = Uo.p[S_BX][o], by = Uo.p[S_BY][o], bz = Uo.p[S_BZ][o];
    b2 = (bx * bx + by * by + bz * bz);
  }
  floor_and_sync<MHD, true>(new_rho, v, b2, gamma, pfloor, eta, ggm1, etot, eint);
  Uo.p[S_RHO][o] = new_rho;
  for (int q = 1; q < 4; q++) Uo.p[q][o] = v[q - 1];
  Uo.p[S_E][o] = etot;
  Uo.p[E][o] = eint;
}

// apply_floor_to_energy_and_sync in place over a box of an integration map
template <bool MHD, bool DE>
__global__ void k_op_sync(KMap U, Box box, double gamma, double pfloor, double eta,
                          double ggm1) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long c = off(U.s, p);
  const double v[3] = {U.p[S_VX][c], U.p[S_VY][c], U.p[S_VZ][c]};
  double b2 = 0.;
  if (MHD) {
    double bx = U.p[S_BX][c], by = U.p[S_BY][c], bz = U.p[S_BZ][c];
    b2 = (bx * bx + by * by + bz * bz);
  }
  double etot = U.p[S_E][c];
  double eint = DE ? U.p[ENZOE_SLOT_INTERNAL_ENERGY][c] : 0.;
  floor_and_sync<MHD, DE>(U.p[S_RHO][c], v, b2, gamma, pfloor, eta, ggm1, etot, eint);
  U.p[S_E][c] = etot;
  if (DE) U.p[ENZOE_SLOT_INTERNAL_ENERGY][c] = eint;
}

// p = (gamma-1) rho eint (EnzoComputePressure.cpp:127-134), dual energy
__global__ void k_op_primitive_de(KMap U, KMap W, Box box, int mhd, double gm1,
                                  unsigned pmask) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long c = off(U.s, p), w = off(W.s, p);
  const int nq = mhd ? 8 : 5;
  for (int q = 0; q < nq; q++)
    if (q != S_E) W.p[q][w] = U.p[q][c];
  for (int q = ENZOE_SLOT_PASSIVE0; q < ENZOE_MAP_SLOTS; q++)
    if (pmask >> q & 1u) W.p[q][w] = U.p[q][c] / U.p[S_RHO][c];
  W.p[S_E][w] = gm1 * U.p[S_RHO][c] * U.p[ENZOE_SLOT_INTERNAL_ENERGY][c];
}

__global__ void k_op_pressure_de(KMap U, KArr P, Box box, double gm1) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long c = off(U.s, p);
  P.p[off(P.s, p)] = gm1 * U.p[S_RHO][c] * U.p[ENZOE_SLOT_INTERNAL_ENERGY][c];
}

// p = (gamma-1) (rho (e_tot - v^2/2) - B^2/2)
// (EnzoComputePressure::compute_pressure, EnzoComputePressure.cpp:177-189)
__device__ __forceinline__ double pressure_at(const KMap &U, long c, int mhd,
                                              double gm1) {
  double vx = U.p[S_VX][c], vy = U.p[S_VY][c], vz = U.p[S_VZ][c];
  double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
  double me_den = 0.;
  if (mhd) {
    double bx = U.p[S_BX][c], by = U.p[S_BY][c], bz = U.p[S_BZ][c];
    me_den = 0.5 * (bx * bx + by * by + bz * bz);
  }
  return gm1 * (U.p[S_RHO][c] * (U.p[S_E][c] - ke) - me_den);
}

// EnzoPhysicsFluidProps::primitive_from_integration
// (EnzoPhysicsFluidProps.cpp:64-138) over [s, n-s): copies density, velocity
// and B, pressure into the pressure slot.
__global__ void k_op_primitive(KMap U, KMap W, Box box, int mhd, double gm1,
                               unsigned pmask) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long c = off(U.s, p), w = off(W.s, p);
  const int nq = mhd ? 8 : 5;
  for (int q = 0; q < nq; q++)
    if (q != S_E) W.p[q][w] = U.p[q][c];
  // passive densities -> mass fractions (EnzoPhysicsFluidProps.cpp:121-134)
  for (int q = ENZOE_SLOT_PASSIVE0; q < ENZOE_MAP_SLOTS; q++)
    if (pmask >> q & 1u) W.p[q][w] = U.p[q][c] / U.p[S_RHO][c];
  W.p[S_E][w] = pressure_at(U, c, mhd, gm1);
}

// EnzoPhysicsFluidProps::pressure_from_integration (:142-156)
__global__ void k_op_pressure(KMap U, KArr P, Box box, int mhd, double gm1) {
  int p[3];
  if (!box_pos(box, p)) return;
  P.p[off(P.s, p)] = pressure_at(U, off(U.s, p), mhd, gm1);
}


// EnzoMHDIntegratorStageCommands::timestep
// (EnzoMHDIntegratorStageCommands.cpp:299-366): min over every cell of the
// views of dx_i/(|v_i| + c), c = fast magnetosonic (MHD) or sound speed from
// the given pressure; reduced as the bit pattern of a non-negative double
// (integer atomicMin == exact double min; a NaN never wins, as std::min).
template <bool MHD>
__global__ void __launch_bounds__(256)
k_op_dt(KMap U, KArr P, double gamma, double dx, double dy, double dz,
        unsigned long long *result) {
  const long nxy = (long)U.n[0] * U.n[1];
  const long n = nxy * U.n[2];
  double local = 1.7976931348623157e308;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    int p[3];
    p[2] = (int)(t / nxy);
    long r = t - (long)p[2] * nxy;
    p[1] = (int)(r / U.n[0]);
    p[0] = (int)(r - (long)p[1] * U.n[0]);
    const long c = off(U.s, p);
    double d = U.p[S_RHO][c], pr = P.p[off(P.s, p)];
    double sig;
    if (MHD) {
      double bx = U.p[S_BX][c], by = U.p[S_BY][c], bz = U.p[S_BZ][c];
      double B2 = ((bx * bx) + ((by * by) + (bz * bz)));
      double cs2 = gamma * pr / d;
      double va2 = B2 / d;
      sig = sqrt(va2 + cs2);
    } else {
      sig = sqrt(gamma * pr / d);
    }
    double l = dmin3(dx / (fabs(U.p[S_VX][c]) + sig),
                     dy / (fabs(U.p[S_VY][c]) + sig),
                     dz / (fabs(U.p[S_VZ][c]) + sig));
    local = dmin(local, l);
  }
  unsigned long long bits = __double_as_longlong(local);
  if (!(local >= 0.0)) bits = 0x7FF8000000000000ull;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long v = __shfl_xor(bits, o, 64);
    bits = (v < bits) ? v : bits;
  }
  __shared__ unsigned long long wmin[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wmin[wave] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wmin[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = (wmin[w] < m) ? wmin[w] : m;
    atomicMin(result, m);
  }
}

// The 8-byte reduction target of enzoe_integrator_timestep: one per host
// thread and device, so that callers on different threads (each with its
// own method and stream, or all on the null stream) never initialise, reduce
// into or read back each other's minimum.  A thread's targets go back to a
// pool when the thread ends and are reused by later threads (device memory
// is never freed: a thread may end after the runtime has shut down).
struct DtTargetPool {
  std::mutex mtx;
  std::map<int, std::vector<unsigned long long *>> free_list;
};
DtTargetPool &dt_pool() {
  static DtTargetPool *pool = new DtTargetPool;  // outlives every thread
  return *pool;
}
struct ThreadDtTargets {
  std::map<int, unsigned long long *> own;
  ~ThreadDtTargets() {
    DtTargetPool &p = dt_pool();
    std::lock_guard<std::mutex> lock(p.mtx);
    for (auto &kv : own) p.free_list[kv.first].push_back(kv.second);
  }
};
unsigned long long *dt_target() {
  thread_local ThreadDtTargets mine;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  auto it = mine.own.find(dev);
  if (it != mine.own.end()) return it->second;
  unsigned long long *p = nullptr;
  {
    DtTargetPool &pool = dt_pool();
    std::lock_guard<std::mutex> lock(pool.mtx);
    std::vector<unsigned long long *> &fl = pool.free_list[dev];
    if (!fl.empty()) {
      p = fl.back();
      fl.pop_back();
    }
  }
  if (!p && hipMalloc((void **)&p, sizeof(unsigned long long)) != hipSuccess)
    return nullptr;
  mine.own[dev] = p;
  return p;
}

int recon_kind(const std::string &n) {
  if (n == "nn") return RC_NN;
  if (n == "plm" || n == "plm_enzo") return RC_PLM;
  if (n == "plm_athena") return RC_PLM_ATHENA;
  return -1;
}

bool valid_dim(int dim) { return dim >= 0 && dim <= 2; }

void stale_box(const int n[3], int s, int lo[3], int hi[3]) {
  for (int d = 0; d < 3; d++) {
    lo[d] = s;
    hi[d] = n[d] - s;
  }
}

}  // namespace

extern "C" {

int enzoe_reconstructor_staling(const char *name, int *total_rate,
                                int *immediate_rate) {
  int kind = recon_kind(lower(name));
  if (kind < 0) return ENZOE_ERR_CONFIG;
  // EnzoReconstructorNN.hpp:31-35, EnzoReconstructorPLM.hpp:153-157
  if (total_rate) *total_rate = (kind == RC_NN) ? 1 : 2;
  if (immediate_rate) *immediate_rate = (kind == RC_NN) ? 0 : 1;
  return ENZOE_OK;
}

int enzoe_reconstruct_interface(const char *name, double theta_limiter,
                                double density_floor, double pressure_floor,
                                const enzoe_fmap *prim,
                                const enzoe_fmap *priml,
                                const enzoe_fmap *primr, int dim,
                                int stale_depth, enzoe_stream_t stream) {
  int kind = recon_kind(lower(name));
  if (kind < 0) return ENZOE_ERR_CONFIG;
  KMap w, wl, wr;
  if (!valid_dim(dim) || stale_depth < 0 || !to_kmap(prim, 0, w))
    return ENZOE_ERR_ARG;
  const unsigned slots = present(w);
  if (!slots || !to_kmap(priml, slots, wl) || !to_kmap(primr, slots, wr))
    return ENZOE_ERR_ARG;
  if (!shifted(wl.n, w.n, dim, -1) || !same_shape(wl.n, wr.n))
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(w.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipStream_t st = (hipStream_t)stream;
  if (kind == RC_NN)
    hipLaunchKernelGGL(k_op_reconstruct<RC_NN>, dim3(nblk(count)), dim3(256), 0,
                       st, w, wl, wr, slots, box, dim, stale_depth,
                       theta_limiter, density_floor, pressure_floor);
  else if (kind == RC_PLM)
    hipLaunchKernelGGL(k_op_reconstruct<RC_PLM>, dim3(nblk(count)), dim3(256),
                       0, st, w, wl, wr, slots, box, dim, stale_depth,
                       theta_limiter, density_floor, pressure_floor);
  else
    hipLaunchKernelGGL(k_op_reconstruct<RC_PLM_ATHENA>, dim3(nblk(count)),
                       dim3(256), 0, st, w, wl, wr, slots, box, dim,
                       stale_depth, theta_limiter, density_floor,
                       pressure_floor);
  return launched();
}

int enzoe_riemann_solve(const char *solver, int mhd, double gamma,
                        const enzoe_fmap *priml, const enzoe_fmap *primr,
                        const enzoe_fmap *flux, int dim, int stale_depth,
                        enzoe_stream_t stream) {
  std::string s = lower(solver);
  int kind;
  // EnzoRiemann::construct_riemann (riemann/EnzoRiemann.cpp:26-68)
  if (s == "hlld" && mhd) kind = RS_HLLD;
  else if (s == "hlle" && mhd) kind = RS_HLLE;
  else if (s == "hllc" && !mhd) kind = RS_HLLC;
  else return ENZOE_ERR_CONFIG;
  const unsigned need = mhd ? kMhdSlots : kHydroSlots;
  KMap l, r, f;
  if (!valid_dim(dim) || stale_depth < 0 || !to_kmap(priml, need, l) ||
      !to_kmap(primr, need, r) || !to_kmap(flux, need, f))
    return ENZOE_ERR_ARG;
  if (!same_shape(l.n, r.n) || !same_shape(l.n, f.n)) return ENZOE_ERR_ARG;
  const unsigned pmask = present(l) & kPassiveSlots;
  if ((present(r) & kPassiveSlots) != pmask || (present(f) & pmask) != pmask)
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(f.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipStream_t st = (hipStream_t)stream;
  if (kind == RS_HLLD)
    hipLaunchKernelGGL((k_op_riemann<RS_HLLD, true>), dim3(nblk(count)),
                       dim3(256), 0, st, l, r, f, box, dim, gamma, pmask);
  else if (kind == RS_HLLE)
    hipLaunchKernelGGL((k_op_riemann<RS_HLLE, true>), dim3(nblk(count)),
                       dim3(256), 0, st, l, r, f, box, dim, gamma, pmask);
  else
    hipLaunchKernelGGL((k_op_riemann<RS_HLLC, false>), dim3(nblk(count)),
                       dim3(256), 0, st, l, r, f, box, dim, gamma, pmask);
  return launched();
}

int enzoe_ct_correct_reconstructed_bfield(int dim, const enzoe_farray *bfieldi,
                                          const enzoe_fmap *priml,
                                          const enzoe_fmap *primr,
                                          int stale_depth,
                                          enzoe_stream_t stream) {
  if (!valid_dim(dim) || stale_depth < 0) return ENZOE_ERR_ARG;
  KArr b;
  KMap l, r;
  const unsigned need = 1u << (ENZOE_SLOT_BFIELD_X + dim);
  if (!to_karr(bfieldi, b) || !to_kmap(priml, need, l) ||
      !to_kmap(primr, need, r))
    return ENZOE_ERR_ARG;
  // interior faces: CSlice(1,-1) of bfieldi along dim
  if (!shifted(b.n, l.n, dim, 2) || !same_shape(l.n, r.n)) return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(l.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_correct_bfield, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, b, l, r, box, dim);
  return launched();
}

int enzoe_ct_identify_upwind(int dim, const enzoe_fmap *flux,
                             const enzoe_farray *weight, int stale_depth,
                             enzoe_stream_t stream) {
  KMap f;
  KArr w;
  if (!valid_dim(dim) || stale_depth < 0 ||
      !to_kmap(flux, 1u << ENZOE_SLOT_DENSITY, f) || !to_karr(weight, w) ||
      !same_shape(f.n, w.n))
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(f.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_upwind, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, f, w, box);
  return launched();
}

int enzoe_ct_compute_center_efield(int dim, const enzoe_farray *efield,
                                   const enzoe_fmap *integration,
                                   int stale_depth, enzoe_stream_t stream) {
  KArr e;
  KMap u;
  if (!valid_dim(dim) || stale_depth < 0 || !to_karr(efield, e) ||
      !to_kmap(integration, kMhdSlots, u) || !same_shape(e.n, u.n))
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(e.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_center_efield, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, e, u, box, dim);
  return launched();
}

int enzoe_ct_compute_edge_efield(int dim, const enzoe_farray *center_efield,
                                 const enzoe_farray *edge_efield,
                                 const enzoe_fmap *jflux,
                                 const enzoe_fmap *kflux,
                                 const enzoe_farray weight[3],
                                 int stale_depth, enzoe_stream_t stream) {
  if (!valid_dim(dim) || stale_depth < 0 || !weight) return ENZOE_ERR_ARG;
  const int aj = (dim + 1) % 3, ak = (dim + 2) % 3;
  KArr ec, ed;
  KMap fj, fk;
  W3 w;
  if (!to_karr(center_efield, ec) || !to_karr(edge_efield, ed) ||
      !to_kmap(jflux, 1u << (ENZOE_SLOT_BFIELD_X + ak), fj) ||
      !to_kmap(kflux, 1u << (ENZOE_SLOT_BFIELD_X + aj), fk))
    return ENZOE_ERR_ARG;
  for (int d = 0; d < 3; d++)
    if (!to_karr(&weight[d], w.w[d]) || !shifted(w.w[d].n, ec.n, d, -1))
      return ENZOE_ERR_ARG;
  // edge array: extents minus one along j and k (EnzoBfieldMethodCT.cpp:70-72)
  for (int d = 0; d < 3; d++)
    if (ed.n[d] != ec.n[d] - ((d == aj || d == ak) ? 1 : 0)) return ENZOE_ERR_ARG;
  if (!shifted(fj.n, ec.n, aj, -1) || !shifted(fk.n, ec.n, ak, -1))
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = stale_depth + (d == dim ? 1 : 0);
    hi[d] = ec.n[d] - stale_depth - 1;
  }
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_edge_efield, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, ec, ed, fj, fk, w, box, dim);
  return launched();
}

int enzoe_ct_update_bfield(const double cell_width[3], int dim,
                           const enzoe_farray edge_efield[3],
                           const enzoe_farray *cur_bfieldi,
                           const enzoe_farray *out_bfieldi, double dt,
                           int stale_depth, enzoe_stream_t stream) {
  if (!cell_width || !edge_efield || !valid_dim(dim) || stale_depth < 0)
    return ENZOE_ERR_ARG;
  const int aj = (dim + 1) % 3, ak = (dim + 2) % 3;
  KArr ej, ek, bc, bo;
  if (!to_karr(&edge_efield[aj], ej) || !to_karr(&edge_efield[ak], ek) ||
      !to_karr(cur_bfieldi, bc) || !to_karr(out_bfieldi, bo) ||
      !same_shape(bc.n, bo.n))
    return ENZOE_ERR_ARG;
  // cell extents from the face array; edge[j] lacks one along k and dim,
  // edge[k] lacks one along dim and j
  int m[3] = {bc.n[0], bc.n[1], bc.n[2]};
  m[dim] -= 1;
  for (int d = 0; d < 3; d++) {
    if (ej.n[d] != m[d] - ((d == ak || d == dim) ? 1 : 0)) return ENZOE_ERR_ARG;
    if (ek.n[d] != m[d] - ((d == dim || d == aj) ? 1 : 0)) return ENZOE_ERR_ARG;
  }
  // dt/dj, dt/dk (EnzoBfieldMethodCT.cpp:627-628)
  double dtdj = dt / cell_width[aj];
  double dtdk = dt / cell_width[ak];
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = stale_depth + 1;
    hi[d] = bc.n[d] - stale_depth - 1;
  }
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_update_bfield, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, ej, ek, bc, bo, box, dim, dtdj, dtdk);
  return launched();
}

int enzoe_ct_compute_center_bfield(int dim, const enzoe_farray *bfieldc,
                                   const enzoe_farray *bfieldi,
                                   int stale_depth, enzoe_stream_t stream) {
  KArr c, f;
  if (!valid_dim(dim) || stale_depth < 0 || !to_karr(bfieldc, c) ||
      !to_karr(bfieldi, f) || !shifted(f.n, c.n, dim, 1))
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(c.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_center_bfield, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, c, f, box, dim);
  return launched();
}

int enzoe_fill_map(const enzoe_fmap *map, double value,
                   enzoe_stream_t stream) {
  KMap m;
  if (!to_kmap(map, 0, m)) return ENZOE_ERR_ARG;
  const unsigned slots = present(m);
  const int lo[3] = {0, 0, 0};
  long count;
  Box box = make_box(lo, m.n, count);
  if (count == 0 || !slots) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_fill, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, m, slots, box, value);
  return launched();
}

int enzoe_accumulate_flux_component(int dim, double dt, double cell_width,
                                    const enzoe_fmap *flux,
                                    const enzoe_fmap *dUcons, int stale_depth,
                                    enzoe_stream_t stream) {
  KMap f, d;
  if (!valid_dim(dim) || stale_depth < 0 || !to_kmap(dUcons, 0, d))
    return ENZOE_ERR_ARG;
  const unsigned slots = present(d);
  if (!to_kmap(flux, slots, f) || !shifted(f.n, d.n, dim, -1))
    return ENZOE_ERR_ARG;
  // dtdx_i = dt/cell_width (EnzoIntegrationQuanUpdate.cpp:111)
  double dtdx = dt / cell_width;
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = stale_depth + (a == dim ? 1 : 0);
    hi[a] = d.n[a] - stale_depth - (a == dim ? 1 : 0);
  }
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0 || !slots) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_accumulate, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, f, d, slots, box, dim, dtdx);
  return launched();
}

int enzoe_update_quantities(const enzoe_fmap *initial,
                            const enzoe_fmap *dUcons, const enzoe_fmap *out,
                            int mhd, double gamma, double density_floor,
                            double pressure_floor, int stale_depth,
                            enzoe_stream_t stream) {
  KMap a, d, o;
  const unsigned need_out = mhd ? kMhdSlots : kUpdateSlots;
  if (stale_depth < 0 || !to_kmap(initial, kUpdateSlots, a) ||
      !to_kmap(dUcons, kUpdateSlots, d) || !to_kmap(out, need_out, o) ||
      !same_shape(a.n, d.n) || !same_shape(a.n, o.n))
    return ENZOE_ERR_ARG;
  const unsigned pmask = present(d) & kPassiveSlots;
  if ((present(a) & pmask) != pmask || (present(o) & pmask) != pmask)
    return ENZOE_ERR_ARG;
  // an internal-energy change belongs to the dual-energy update
  if (present(d) & kEintSlot) return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(a.n, stale_depth + 1, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_update, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, a, d, o, box, mhd, gamma,
                     density_floor, pressure_floor, pmask);
  return launched();
}

int enzoe_primitive_from_integration(const enzoe_fmap *integration,
                                     const enzoe_fmap *primitive, int mhd,
                                     double gamma, int stale_depth,
                                     enzoe_stream_t stream) {
  KMap u, w;
  const unsigned need = mhd ? kMhdSlots : kHydroSlots;
  if (stale_depth < 0 || !to_kmap(integration, need, u) ||
      !to_kmap(primitive, need, w) || !same_shape(u.n, w.n))
    return ENZOE_ERR_ARG;
  const unsigned pmask = present(u) & present(w) & kPassiveSlots;
  int lo[3], hi[3];
  stale_box(u.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_primitive, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, u, w, box, mhd, gamma - 1.0, pmask);
  return launched();
}

int enzoe_pressure_from_integration(const enzoe_fmap *integration,
                                    const enzoe_farray *pressure, int mhd,
                                    double gamma, int stale_depth,
                                    enzoe_stream_t stream) {
  KMap u;
  KArr p;
  const unsigned need = mhd ? kMhdSlots : kHydroSlots;
  if (stale_depth < 0 || !to_kmap(integration, need, u) ||
      !to_karr(pressure, p) || !same_shape(u.n, p.n))
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(u.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_pressure, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, u, p, box, mhd, gamma - 1.0);
  return launched();
}

int enzoe_riemann_solve_dual_energy(const char *solver, int mhd, double gamma,
                                    const enzoe_fmap *priml,
                                    const enzoe_fmap *primr,
                                    const enzoe_fmap *flux,
                                    const enzoe_farray *interface_velocity,
                                    int dim, int stale_depth,
                                    enzoe_stream_t stream) {
  std::string s = lower(solver);
  int kind;
  if (s == "hlld" && mhd) kind = RS_HLLD;
  else if (s == "hlle" && mhd) kind = RS_HLLE;
  else if (s == "hllc" && !mhd) kind = RS_HLLC;
  else return ENZOE_ERR_CONFIG;
  const unsigned need = mhd ? kMhdSlots : kHydroSlots;
  KMap l, r, f;
  KArr v;
  // EnzoRiemannUtils.hpp:381-384: the interface velocity must be given
  if (!valid_dim(dim) || stale_depth < 0 || !to_kmap(priml, need, l) ||
      !to_kmap(primr, need, r) || !to_kmap(flux, need | kEintSlot, f) ||
      !to_karr(interface_velocity, v))
    return ENZOE_ERR_ARG;
  if (!same_shape(l.n, r.n) || !same_shape(l.n, f.n) || !same_shape(l.n, v.n))
    return ENZOE_ERR_ARG;
  const unsigned pmask = present(l) & kPassiveSlots;
  if ((present(r) & kPassiveSlots) != pmask || (present(f) & pmask) != pmask)
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(f.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipStream_t st = (hipStream_t)stream;
  if (kind == RS_HLLD)
    hipLaunchKernelGGL((k_op_riemann<RS_HLLD, true, true>), dim3(nblk(count)),
                       dim3(256), 0, st, l, r, f, box, dim, gamma, pmask, v);
  else if (kind == RS_HLLE)
    hipLaunchKernelGGL((k_op_riemann<RS_HLLE, true, true>), dim3(nblk(count)),
                       dim3(256), 0, st, l, r, f, box, dim, gamma, pmask, v);
  else
    hipLaunchKernelGGL((k_op_riemann<RS_HLLC, false, true>), dim3(nblk(count)),
                       dim3(256), 0, st, l, r, f, box, dim, gamma, pmask, v);
  return launched();
}

int enzoe_source_internal_energy(int dim, double dt, double cell_width,
                                 const enzoe_fmap *primitive,
                                 const enzoe_fmap *dUcons,
                                 const enzoe_farray *interface_velocity,
                                 double pressure_floor, int stale_depth,
                                 enzoe_stream_t stream) {
  KMap w, d;
  KArr v;
  if (!valid_dim(dim) || stale_depth < 0 ||
      !to_kmap(primitive, 1u << ENZOE_SLOT_ENERGY, w) || !to_kmap(dUcons, kEintSlot, d) ||
      !to_karr(interface_velocity, v) || !same_shape(w.n, d.n) ||
      !shifted(v.n, d.n, dim, -1))
    return ENZOE_ERR_ARG;
  // dtdx = dt/cell_width (EnzoSourceInternalEnergy.cpp:34)
  const double dtdx = dt / cell_width;
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = stale_depth + (a == dim ? 1 : 0);
    hi[a] = d.n[a] - stale_depth - (a == dim ? 1 : 0);
  }
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_source_eint, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, w, d, v, box, dim, dtdx, pressure_floor);
  return launched();
}

int enzoe_update_quantities_dual_energy(const enzoe_fmap *initial,
                                        const enzoe_fmap *dUcons,
                                        const enzoe_fmap *out, int mhd,
                                        double gamma, double density_floor,
                                        double pressure_floor, double eta,
                                        int stale_depth, enzoe_stream_t stream) {
  KMap a, d, o;
  const unsigned need = kUpdateSlots | kEintSlot;
  const unsigned need_out = (mhd ? kMhdSlots : kUpdateSlots) | kEintSlot;
  if (stale_depth < 0 || !(eta >= 0) || !to_kmap(initial, need, a) ||
      !to_kmap(dUcons, need, d) || !to_kmap(out, need_out, o) ||
      !same_shape(a.n, d.n) || !same_shape(a.n, o.n))
    return ENZOE_ERR_ARG;
  const unsigned pmask = present(d) & kPassiveSlots;
  if ((present(a) & pmask) != pmask || (present(o) & pmask) != pmask)
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(a.n, stale_depth + 1, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  const double ggm1 = sync_ggm1(gamma);
  if (mhd)
    hipLaunchKernelGGL(k_op_update_de<true>, dim3(nblk(count)), dim3(256), 0,
                       (hipStream_t)stream, a, d, o, box, gamma, density_floor,
                       pressure_floor, eta, ggm1, pmask);
  else
    hipLaunchKernelGGL(k_op_update_de<false>, dim3(nblk(count)), dim3(256), 0,
                       (hipStream_t)stream, a, d, o, box, gamma, density_floor,
                       pressure_floor, eta, ggm1, pmask);
  return launched();
}

int enzoe_primitive_from_integration_dual_energy(const enzoe_fmap *integration,
                                                 const enzoe_fmap *primitive,
                                                 int mhd, double gamma,
                                                 int stale_depth,
                                                 enzoe_stream_t stream) {
  KMap u, w;
  const unsigned need = mhd ? kMhdSlots : kHydroSlots;
  if (stale_depth < 0 || !to_kmap(integration, need | kEintSlot, u) ||
      !to_kmap(primitive, need, w) || !same_shape(u.n, w.n))
    return ENZOE_ERR_ARG;
  const unsigned pmask = present(u) & present(w) & kPassiveSlots;
  int lo[3], hi[3];
  stale_box(u.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_primitive_de, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, u, w, box, mhd, gamma - 1.0, pmask);
  return launched();
}

int enzoe_pressure_from_integration_dual_energy(const enzoe_fmap *integration,
                                                const enzoe_farray *pressure,
                                                double gamma, int stale_depth,
                                                enzoe_stream_t stream) {
  KMap u;
  KArr p;
  if (stale_depth < 0 || !to_kmap(integration, 1u | kEintSlot, u) ||
      !to_karr(pressure, p) || !same_shape(u.n, p.n))
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(u.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_pressure_de, dim3(nblk(count)), dim3(256), 0,
                     (hipStream_t)stream, u, p, box, gamma - 1.0);
  return launched();
}

int enzoe_apply_floor_to_energy_and_sync(const enzoe_fmap *integration,
                                         int mhd, double gamma,
                                         double pressure_floor, int dual_energy,
                                         double eta, int stale_depth,
                                         enzoe_stream_t stream) {
  KMap u;
  const unsigned need = (mhd ? kMhdSlots : kUpdateSlots) | (dual_energy ? kEintSlot : 0u);
  if (stale_depth < 0 || (dual_energy && !(eta >= 0)) || !to_kmap(integration, need, u))
    return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  stale_box(u.n, stale_depth, lo, hi);
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  const double ggm1 = sync_ggm1(gamma);
  hipStream_t st = (hipStream_t)stream;
  if (mhd && dual_energy)
    hipLaunchKernelGGL((k_op_sync<true, true>), dim3(nblk(count)), dim3(256), 0, st, u,
                       box, gamma, pressure_floor, eta, ggm1);
  else if (mhd)
    hipLaunchKernelGGL((k_op_sync<true, false>), dim3(nblk(count)), dim3(256), 0, st, u,
                       box, gamma, pressure_floor, eta, ggm1);
  else if (dual_energy)
    hipLaunchKernelGGL((k_op_sync<false, true>), dim3(nblk(count)), dim3(256), 0, st, u,
                       box, gamma, pressure_floor, eta, ggm1);
  else
    hipLaunchKernelGGL((k_op_sync<false, false>), dim3(nblk(count)), dim3(256), 0, st, u,
                       box, gamma, pressure_floor, eta, ggm1);
  return launched();
}

int enzoe_integrator_timestep(const enzoe_fmap *integration,
                              const enzoe_farray *pressure, int mhd,
                              double gamma, double dx, double dy, double dz,
                              enzoe_stream_t stream, double *dt_out) {
  KMap u;
  KArr p;
  const unsigned need = mhd ? kMhdSlots : kHydroSlots;
  if (!dt_out || !to_kmap(integration, need, u) || !to_karr(pressure, p) ||
      !same_shape(u.n, p.n))
    return ENZOE_ERR_ARG;
  const long n = (long)u.n[0] * u.n[1] * u.n[2];
  if (n == 0) {
    *dt_out = 1.7976931348623157e308;
    return ENZOE_OK;
  }
  unsigned long long *target = dt_target();
  if (!target) return ENZOE_ERR_ALLOC;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, st, target);
  const unsigned wg = (unsigned)std::min<long>(nblk(n), 1024);
  if (mhd)
    hipLaunchKernelGGL(k_op_dt<true>, dim3(wg), dim3(256), 0, st, u, p, gamma,
                       dx, dy, dz, target);
  else
    hipLaunchKernelGGL(k_op_dt<false>, dim3(wg), dim3(256), 0, st, u, p, gamma,
                       dx, dy, dz, target);
  if (hipGetLastError() != hipSuccess) return ENZOE_ERR_HIP;
  unsigned long long bits = 0;
  if (hipMemcpyAsync(&bits, target, sizeof(bits), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return ENZOE_ERR_HIP;
  double dt;
  std::memcpy(&dt, &bits, sizeof(dt));
  *dt_out = dt;
  return ENZOE_OK;
}

}  // extern "C"

// ------------------------------------------------------- flux corrections

namespace {

// save_fluxes_for_corrections_ for one dim (EnzoMethodMHDVlct.cpp:250-330):
// one thread per (field, face, face-array entry) of the block part.
__global__ void k_op_save_fluxes(KMap F, StoreSlots sl, int dim, double dt_dxi,
                                 int g, FluxLayout L, double *fd) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long m = L.m[dim];
  if (t >= 2 * m * sl.nf) return;
  const int i_f = (int)(t / (2 * m));
  const long r = t - (long)i_f * 2 * m;
  const int face = (int)(r / m);
  const long idx = r - face * m;
  int p[3];
  L.face_cell(dim, idx, p);
  for (int a = 0; a < 3; a++) p[a] += g;
  // the flux view is the cell shape minus one along dim
  p[dim] = face ? F.n[dim] - g : g - 1;
  const double v = dt_dxi * F.p[sl.slot[i_f]][off(F.s, p)];
  gst(fd, L.offset(i_f, dim, face, 0) + idx, v);
}

// Block::refresh_load_flux_face_ (control_refresh.cpp:1458-1500): the fine
// block's fluxes on (axis, face) coarsened as FaceFluxes::coarsen does for a
// rank-3 cell-centred face (data_FaceFluxes.cpp:128-250: every fine value
// added four times with weight 0.25 * 0.125, fine faces in loop order) into
// quadrant `child` of the coarse face, then summed into the coarse block's
// neighbour fluxes on (axis, 1 - face) (FluxData::sum_neighbor_fluxes,
// data_FluxData.hpp:136-145).  One thread per (field, coarse entry).
struct Child {
  int c[3];
};

// quadrant_only: leave the entries outside the child's quadrant untouched
// instead of adding the coarsened array's zeros there (the batched launch,
// where the four siblings of one coarse face run concurrently; the result
// differs from the sequential sum only in the sign of a zero).
__device__ __forceinline__ void send_to_coarse(const double *__restrict__ fine,
                                               double *__restrict__ coarse,
                                               const FluxLayout &L, int axis,
                                               int face, const int child[3],
                                               int i_f, long idx,
                                               bool quadrant_only) {
  const int a0 = (axis == 0) ? 1 : 0, a1 = (axis == 2) ? 1 : 2;
  const int nu = L.n[a0], nv = L.n[a1];
  const int u = (int)(idx % nu), v = (int)(idx / nu);
  const int hu = nu / 2, hv = nv / 2;
  const int cu = child[a0], cv = child[a1];
  double acc = 0.0;
  const bool inside =
      u >= cu * hu && u < cu * hu + hu && v >= cv * hv && v < cv * hv + hv;
  if (inside) {
    const double *src = fine + L.offset(i_f, axis, face, 0);
    const int u2 = 2 * (u - cu * hu), v2 = 2 * (v - cv * hv);
    const double w = 0.25 * 0.125;
    for (int dv = 0; dv < 2; dv++)
      for (int du = 0; du < 2; du++) {
        const double f = src[(long)(v2 + dv) * nu + (u2 + du)];
        acc += w * f;
        acc += w * f;
        acc += w * f;
        acc += w * f;
      }
  } else if (quadrant_only) {
    return;
  }
  double *dst = coarse + L.offset(i_f, axis, 1 - face, 1);
  dst[idx] = dst[idx] + acc;
}

__global__ void k_op_send_to_coarse(const double *__restrict__ fine,
                                    double *__restrict__ coarse, int nf,
                                    FluxLayout L, int axis, int face,
                                    Child ch) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long m = L.m[axis];
  if (t >= m * nf) return;
  const int i_f = (int)(t / m);
  send_to_coarse(fine, coarse, L, axis, face, ch.c, i_f, t - (long)i_f * m,
                 false);
}

// one link per blockIdx.y
__global__ void k_op_send_to_coarse_batch(const enzoe_flux_link *__restrict__ links,
                                          int nf, FluxLayout L) {
  const enzoe_flux_link &k = links[blockIdx.y];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long m = L.m[k.axis];
  if (t >= m * nf) return;
  const int i_f = (int)(t / m);
  send_to_coarse(k.fine_flux_data, k.coarse_flux_data, L, k.axis, k.face,
                 k.child, i_f, t - (long)i_f * m, true);
}

// MethodFluxCorrect::flux_correct_ (problem_MethodFluxCorrect.cpp:425-560)
// for one leaf block, one thread per cell of the shell of the active
// region: density first (its start value kept), then every other field,
// through rho * q for fields that must be made conservative; a cell on
// several corrected faces takes them in axis, then face, order
// (flux_correct_helper_, :275-370), and is divided by the corrected density
// once (divide_flux_corrected_cells_by_density_, :372-423).
struct CorrectArgs {
  double *field[16];
  int conservative[16];
  int nf, i_density;
  int correct[6];
};

// the fields' shared parameters of a batched correction
struct CorrectShared {
  int conservative[16];
  int nf, i_density;
};

__device__ __forceinline__ void flux_correct_cell(
    double *const *field, const int *conservative, int nf, int i_density,
    const int *correct, const double *__restrict__ fd, const FluxLayout &L,
    int g, int mx, int my, long t) {
  const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
  // shell: the two x faces, then the y faces without the x boundary, then
  // the z faces without the x and y boundaries
  const long nA = 2L * ny * nz, nB = 2L * (nx - 2) * nz,
             nC = 2L * (nx - 2) * (ny - 2);
  int p[3];
  if (t < nA) {
    const long r = t % ((long)ny * nz);
    p[0] = (t < (long)ny * nz) ? 0 : nx - 1;
    p[1] = (int)(r % ny);
    p[2] = (int)(r / ny);
  } else if (t < nA + nB) {
    const long u = t - nA, half = (long)(nx - 2) * nz;
    const long r = u % half;
    p[1] = (u < half) ? 0 : ny - 1;
    p[0] = 1 + (int)(r % (nx - 2));
    p[2] = (int)(r / (nx - 2));
  } else if (t < nA + nB + nC) {
    const long u = t - nA - nB, half = (long)(nx - 2) * (ny - 2);
    const long r = u % half;
    p[2] = (u < half) ? 0 : nz - 1;
    p[0] = 1 + (int)(r % (nx - 2));
    p[1] = 1 + (int)(r / (nx - 2));
  } else {
    return;
  }
  bool on[6];
  long ib[3];
  bool any = false;
  for (int ax = 0; ax < 3; ax++) {
    const int a0 = (ax == 0) ? 1 : 0, a1 = (ax == 2) ? 1 : 2;
    ib[ax] = p[a0] + (long)L.n[a0] * p[a1];
    for (int f = 0; f < 2; f++) {
      on[2 * ax + f] = correct[2 * ax + f] && p[ax] == (f ? L.n[ax] - 1 : 0);
      any = any || on[2 * ax + f];
    }
  }
  if (!any) return;
  const long c = (p[0] + g) + (long)mx * ((p[1] + g) + (long)my * (p[2] + g));
  double rho_old = 0.0, rho_new = 0.0;
  if (i_density >= 0) {
    rho_old = field[i_density][c];
    double r = rho_old;
    for (int k = 0; k < 6; k++) {
      if (!on[k]) continue;
      const int ax = k >> 1, f = k & 1;
      const double *bf = fd + L.offset(i_density, ax, f, 0);
      const double *nb = fd + L.offset(i_density, ax, f, 1);
      r += (double)(2 * f - 1) * (bf[ib[ax]] - nb[ib[ax]]);
    }
    rho_new = r;
    field[i_density][c] = r;
  }
  for (int i_f = 0; i_f < nf; i_f++) {
    if (i_f == i_density) continue;
    double q = conservative[i_f] ? rho_old * field[i_f][c] : field[i_f][c];
    for (int k = 0; k < 6; k++) {
      if (!on[k]) continue;
      const int ax = k >> 1, f = k & 1;
      const double *bf = fd + L.offset(i_f, ax, f, 0);
      const double *nb = fd + L.offset(i_f, ax, f, 1);
      q += (double)(2 * f - 1) * (bf[ib[ax]] - nb[ib[ax]]);
    }
    field[i_f][c] = conservative[i_f] ? q / rho_new : q;
  }
}

__global__ void k_op_flux_correct(CorrectArgs a, const double *__restrict__ fd,
                                  FluxLayout L, int g, int mx, int my) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  flux_correct_cell(a.field, a.conservative, a.nf, a.i_density, a.correct, fd,
                    L, g, mx, my, t);
}

// one block per blockIdx.y
__global__ void k_op_flux_correct_batch(
    const enzoe_flux_correct_block *__restrict__ blocks, CorrectShared sh,
    FluxLayout L, int g, int mx, int my) {
  const enzoe_flux_correct_block &b = blocks[blockIdx.y];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  flux_correct_cell(b.fields, sh.conservative, sh.nf, sh.i_density, b.correct,
                    b.flux_data, L, g, mx, my, t);
}

// ------------------------------------------------------- gravity source

// EnzoSourceGravity::calculate_source (toolkit/EnzoSourceGravity.cpp:17-69)
// over the box: momentum += dt rho a, energy density += dt rho (v . a)
__global__ void k_op_gravity(KMap U, KMap dU, KMap A, Box box, double dt) {
  int p[3];
  if (!box_pos(box, p)) return;
  const long cu = off(U.s, p), cd = off(dU.s, p), ca = off(A.s, p);
  const double rho = U.p[ENZOE_SLOT_DENSITY][cu];
  const double ax = A.p[1][ca], ay = A.p[2][ca], az = A.p[3][ca];
  dU.p[1][cd] += dt * rho * ax;
  dU.p[2][cd] += dt * rho * ay;
  dU.p[3][cd] += dt * rho * az;
  dU.p[ENZOE_SLOT_ENERGY][cd] +=
      dt * rho * ((U.p[1][cu] * ax) + (U.p[2][cu] * ay) + (U.p[3][cu] * az));
}

// ------------------------------------------------ coarse/fine transfers

struct Int3 {
  int v[3];
};

// ProlongLinear::apply_ (problem_ProlongLinear.cpp:85-330): source index and
// weights of fine cell i_f of n_f along one axis; gc = 1 when the coarse
// region has no ghost cells around the fine region (one-sided weights at
// its two ends)
__device__ __forceinline__ void prolong_axis(int i_f, int n_f, int gc, int &ic,
                                             double &w0, double &w1) {
  int c = ((i_f + 1) >> 1) - gc;
  int wa = 1, wb = 3;
  if (i_f == 0) c += gc;
  if (i_f == n_f - 1) c -= gc;
  if (i_f == 0 || i_f == n_f - 1) {
    wa += 4 * gc;
    wb -= 4 * gc;
  }
  ic = c;
  w0 = 0.25 * ((i_f & 1) ? wb : wa);
  w1 = 0.25 * ((i_f & 1) ? wa : wb);
}

// one fine cell t of the region; the sums are formed in the reference's
// order (8, 4 or 2 terms left to right)
__device__ __forceinline__ void prolong_cell(double *__restrict__ vf,
                                             const int *mf, const int *of,
                                             const int *nf,
                                             const double *__restrict__ vc,
                                             const int *mc, const int *oc,
                                             const int *gc, int rank,
                                             int accumulate, long t) {
  const long nxy = (long)nf[0] * nf[1];
  if (t >= nxy * nf[2]) return;
  const int ifz = (int)(t / nxy);
  const long r = t - (long)ifz * nxy;
  const int ify = (int)(r / nf[0]);
  const int ifx = (int)(r - (long)ify * nf[0]);
  const long dcx = 1, dcy = mc[0], dcz = (long)mc[0] * mc[1];
  int icx, icy = 0, icz = 0;
  double wx0, wx1, wy0 = 0., wy1 = 0., wz0 = 0., wz1 = 0.;
  prolong_axis(ifx, nf[0], gc[0], icx, wx0, wx1);
  if (rank >= 2) prolong_axis(ify, nf[1], gc[1], icy, wy0, wy1);
  if (rank >= 3) prolong_axis(ifz, nf[2], gc[2], icz, wz0, wz1);
  const long i_c = (oc[0] + icx) +
                   (long)mc[0] * ((oc[1] + icy) + (long)mc[1] * (oc[2] + icz));
  const long i_f = (of[0] + ifx) +
                   (long)mf[0] * ((of[1] + ify) + (long)mf[1] * (of[2] + ifz));
  double v;
  if (rank == 1) {
    v = wx0 * vc[i_c] + wx1 * vc[i_c + dcx];
  } else if (rank == 2) {
    v = wx0 * wy0 * vc[i_c] + wx1 * wy0 * vc[i_c + dcx] +
        wx0 * wy1 * vc[i_c + dcy] + wx1 * wy1 * vc[i_c + dcx + dcy];
  } else {
    v = wx0 * wy0 * wz0 * vc[i_c] + wx1 * wy0 * wz0 * vc[i_c + dcx] +
        wx0 * wy1 * wz0 * vc[i_c + dcy] + wx1 * wy1 * wz0 * vc[i_c + dcx + dcy] +
        wx0 * wy0 * wz1 * vc[i_c + dcz] + wx1 * wy0 * wz1 * vc[i_c + dcx + dcz] +
        wx0 * wy1 * wz1 * vc[i_c + dcy + dcz] +
        wx1 * wy1 * wz1 * vc[i_c + dcx + dcy + dcz];
  }
  vf[i_f] = accumulate ? vf[i_f] + v : v;
}

// RestrictLinear::apply_ (problem_RestrictLinear.cpp:60-198): one coarse
// cell t, the mean of its 2^rank children summed in the reference's order
__device__ __forceinline__ void restrict_cell(double *__restrict__ vc,
                                              const int *ndc, const int *imc,
                                              const int *nc,
                                              const double *__restrict__ vf,
                                              const int *ndf, const int *imf,
                                              int rank, int accumulate, long t) {
  const long nxy = (long)nc[0] * nc[1];
  if (t >= nxy * nc[2]) return;
  const int izc = (int)(t / nxy);
  const long r = t - (long)izc * nxy;
  const int iyc = (int)(r / nc[0]);
  const int ixc = (int)(r - (long)iyc * nc[0]);
  const long dx = 1, dy = ndf[0], dz = (long)ndf[0] * ndf[1];
  const long i_c = (imc[0] + ixc) +
                   (long)ndc[0] * ((imc[1] + iyc) + (long)ndc[1] * (imc[2] + izc));
  const long i_f = (imf[0] + 2 * ixc) +
                   (long)ndf[0] * ((imf[1] + 2 * iyc) + (long)ndf[1] * (imf[2] + 2 * izc));
  double v;
  if (rank == 1) {
    v = 0.5 * (vf[i_f] + vf[i_f + dx]);
  } else if (rank == 2) {
    v = 0.25 * (vf[i_f] + vf[i_f + dx] + vf[i_f + dy] + vf[i_f + dx + dy]);
  } else {
    v = 0.125 * (vf[i_f] + vf[i_f + dx] + vf[i_f + dy] + vf[i_f + dx + dy] +
                 vf[i_f + dz] + vf[i_f + dx + dz] + vf[i_f + dy + dz] +
                 vf[i_f + dx + dy + dz]);
  }
  vc[i_c] = accumulate ? vc[i_c] + v : v;
}

// --------------------------------------------- EnzoProlong (3A, 2A, 2B)
// EnzoProlong::apply_ (src/Enzo/mesh/EnzoProlong.cpp:112-190) calls
// interpolate (src/Enzo/mesh/interpolate.F) with refinement 2, pstart 2 and
// pend = nf + 1: the coarse region holds the parents of the fine region and
// one cell of stencil on each side.  One thread per parent: it writes its 8
// children, with the Fortran routines' operation order (tscint3d, interp3d,
// int_lin3d in src/Enzo/hydro-mhd/ppm_fortran/).  pc points at the coarse
// region's first cell, gf at the fine region's; Fortran indices are kept
// 1-based (PRO_P, PRO_G).

#define PRO_P(ip, jp, kp) pc[((ip) - 1) + (long)mc.v[0] * (((jp) - 1) + (long)mc.v[1] * ((kp) - 1))]
#define PRO_G(ig, jg, kg) gf[((ig) - 1) + (long)mf.v[0] * (((jg) - 1) + (long)mf.v[1] * ((kg) - 1))]

__device__ __forceinline__ double pro_min(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double pro_max(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double pro_sign(double a, double b) { return copysign(fabs(a), b); }

// the corner average work(i, j, k) of interp3d / int_lin3d (pstart = 1)
__device__ __forceinline__ double pro_corner(const double *__restrict__ pc, Int3 mc, int i, int j,
                                             int k) {
  return 0.125 * (PRO_P(i, j, k) + PRO_P(i, j, k + 1) + PRO_P(i, j + 1, k) +
                  PRO_P(i, j + 1, k + 1) + PRO_P(i + 1, j, k) + PRO_P(i + 1, j, k + 1) +
                  PRO_P(i + 1, j + 1, k) + PRO_P(i + 1, j + 1, k + 1));
}

// the children fbar + (i1 - 1/2)/2 fx + (j1 - 1/2)/2 fy + (k1 - 1/2)/2 fz of
// parent (i, j, k)
__device__ __forceinline__ void pro_children(double *__restrict__ gf, Int3 mf, int i, int j,
                                             int k, double fbar, double fx, double fy,
                                             double fz) {
  for (int k1 = 0; k1 < 2; k1++)
    for (int j1 = 0; j1 < 2; j1++)
      for (int i1 = 0; i1 < 2; i1++)
        PRO_G(1 + (i - 1) * 2 + i1, 1 + (j - 1) * 2 + j1, 1 + (k - 1) * 2 + k1) =
            (fbar + ((double)i1 + 0.5 - 0.5 * 2.0) / 2.0 * fx +
             ((double)j1 + 0.5 - 0.5 * 2.0) / 2.0 * fy +
             ((double)k1 + 0.5 - 0.5 * 2.0) / 2.0 * fz);
}

// tscint3d ("3A"): triangular-shaped-cloud weights, the conservation
// correction, then two passes of the routine's monotonicity step along x, y
// and z; its second bound, max(max(min(d1right/d2right, 0), 1), fraction),
// is at least 1 and the first leaves fraction <= 1, so every fraction is 1
// (finite data) and each pass sets a parent's children to their mean along
// the axis, all within one parent.
__global__ void k_op_prolong_3a(double *__restrict__ gf, Int3 mf, const double *__restrict__ pc,
                                Int3 mc, Int3 pd) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)pd.v[0] * pd.v[1] * pd.v[2]) return;
  const int i = (int)(t % pd.v[0]) + 1, j = (int)((t / pd.v[0]) % pd.v[1]) + 1,
            k = (int)(t / ((long)pd.v[0] * pd.v[1])) + 1;
  double cm[3], cp[3], c0[3];
  for (int q = 0; q < 2; q++) {
    cm[q + 1] = (double)((2 - q) * (2 - q) * (2 - q) - (1 - q) * (1 - q) * (1 - q)) / 48.0;
    cp[q + 1] = (double)(3 * q * q + 3 * q + 1) / 48.0;
    c0[q + 1] = 1.0 / 2 - cm[q + 1] - cp[q + 1];
  }
  const int ip = i + 1, jp = j + 1, kp = k + 1;  // parent of the children
  double g[2][2][2];  // [k1][j1][i1]
  for (int k1 = 0; k1 < 2; k1++) {
    const double wm3 = cm[k1 + 1], w03 = c0[k1 + 1], wp3 = cp[k1 + 1];
    for (int j1 = 0; j1 < 2; j1++) {
      const double wm2 = cm[j1 + 1], w02 = c0[j1 + 1], wp2 = cp[j1 + 1];
      for (int i1 = 0; i1 < 2; i1++) {
        const double wm1 = cm[i1 + 1], w01 = c0[i1 + 1], wp1 = cp[i1 + 1];
        g[k1][j1][i1] =
            wm1 * wm2 * wm3 * PRO_P(ip - 1, jp - 1, kp - 1) +
            w01 * wm2 * wm3 * PRO_P(ip, jp - 1, kp - 1) +
            wp1 * wm2 * wm3 * PRO_P(ip + 1, jp - 1, kp - 1) +
            wm1 * w02 * wm3 * PRO_P(ip - 1, jp, kp - 1) +
            w01 * w02 * wm3 * PRO_P(ip, jp, kp - 1) +
            wp1 * w02 * wm3 * PRO_P(ip + 1, jp, kp - 1) +
            wm1 * wp2 * wm3 * PRO_P(ip - 1, jp + 1, kp - 1) +
            w01 * wp2 * wm3 * PRO_P(ip, jp + 1, kp - 1) +
            wp1 * wp2 * wm3 * PRO_P(ip + 1, jp + 1, kp - 1) +
            wm1 * wm2 * w03 * PRO_P(ip - 1, jp - 1, kp) +
            w01 * wm2 * w03 * PRO_P(ip, jp - 1, kp) +
            wp1 * wm2 * w03 * PRO_P(ip + 1, jp - 1, kp) +
            wm1 * w02 * w03 * PRO_P(ip - 1, jp, kp) +
            w01 * w02 * w03 * PRO_P(ip, jp, kp) +
            wp1 * w02 * w03 * PRO_P(ip + 1, jp, kp) +
            wm1 * wp2 * w03 * PRO_P(ip - 1, jp + 1, kp) +
            w01 * wp2 * w03 * PRO_P(ip, jp + 1, kp) +
            wp1 * wp2 * w03 * PRO_P(ip + 1, jp + 1, kp) +
            wm1 * wm2 * wp3 * PRO_P(ip - 1, jp - 1, kp + 1) +
            w01 * wm2 * wp3 * PRO_P(ip, jp - 1, kp + 1) +
            wp1 * wm2 * wp3 * PRO_P(ip + 1, jp - 1, kp + 1) +
            wm1 * w02 * wp3 * PRO_P(ip - 1, jp, kp + 1) +
            w01 * w02 * wp3 * PRO_P(ip, jp, kp + 1) +
            wp1 * w02 * wp3 * PRO_P(ip + 1, jp, kp + 1) +
            wm1 * wp2 * wp3 * PRO_P(ip - 1, jp + 1, kp + 1) +
            w01 * wp2 * wp3 * PRO_P(ip, jp + 1, kp + 1) +
            wp1 * wp2 * wp3 * PRO_P(ip + 1, jp + 1, kp + 1);
      }
    }
  }
  double temp = 0.0;
  for (int k1 = 0; k1 < 2; k1++)
    for (int j1 = 0; j1 < 2; j1++)
      for (int i1 = 0; i1 < 2; i1++) temp = temp + g[k1][j1][i1];
  temp = (PRO_P(1 + i, 1 + j, 1 + k) - temp) / (double)(2 * 2 * 2);
  for (int k1 = 0; k1 < 2; k1++)
    for (int j1 = 0; j1 < 2; j1++)
      for (int i1 = 0; i1 < 2; i1++) g[k1][j1][i1] = g[k1][j1][i1] + temp;
  const double frac = 1.0;
  for (int iter = 0; iter < 2; iter++) {
    for (int k1 = 0; k1 < 2; k1++)
      for (int j1 = 0; j1 < 2; j1++) {
        double tm = 0.0;
        for (int i1 = 0; i1 < 2; i1++) tm = tm + g[k1][j1][i1] / 2.0;
        for (int i1 = 0; i1 < 2; i1++) g[k1][j1][i1] = frac * tm + (1.0 - frac) * g[k1][j1][i1];
      }
    for (int k1 = 0; k1 < 2; k1++)
      for (int i1 = 0; i1 < 2; i1++) {
        double tm = 0.0;
        for (int j1 = 0; j1 < 2; j1++) tm = tm + g[k1][j1][i1] / 2.0;
        for (int j1 = 0; j1 < 2; j1++) g[k1][j1][i1] = frac * tm + (1.0 - frac) * g[k1][j1][i1];
      }
    for (int j1 = 0; j1 < 2; j1++)
      for (int i1 = 0; i1 < 2; i1++) {
        double tm = 0.0;
        for (int k1 = 0; k1 < 2; k1++) tm = tm + g[k1][j1][i1] / 2.0;
        for (int k1 = 0; k1 < 2; k1++) g[k1][j1][i1] = frac * tm + (1.0 - frac) * g[k1][j1][i1];
      }
  }
  for (int k1 = 0; k1 < 2; k1++)
    for (int j1 = 0; j1 < 2; j1++)
      for (int i1 = 0; i1 < 2; i1++)
        PRO_G(1 + (i - 1) * 2 + i1, 1 + (j - 1) * 2 + j1, 1 + (k - 1) * 2 + k1) =
            g[k1][j1][i1] * 2.0 * 2.0 * 2.0;
}

// interp3d's first loop over the (pd+1)^3 corners: flags[0] counts corners
// or cells that are not positive (nneg), flags[1] a NaN coarse value
__global__ void k_op_prolong_2a_scan(const double *__restrict__ pc, Int3 mc, Int3 pd,
                                     int *__restrict__ flags) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int n0 = pd.v[0] + 1, n1 = pd.v[1] + 1, n2 = pd.v[2] + 1;
  if (t >= (long)n0 * n1 * n2) return;
  const int i = (int)(t % n0) + 1, j = (int)((t / n0) % n1) + 1, k = (int)(t / ((long)n0 * n1)) + 1;
  const double w = pro_corner(pc, mc, i, j, k);
  const double pv = PRO_P(1 + i, 1 + j, 1 + k);
  if (pv != pv) atomicOr(&flags[1], 1);
  if (w <= 0.0 || pv <= 0.0) atomicAdd(&flags[0], 1);
}

// interp3d ("2A"): min-limited corner slopes along the four diagonals; with
// flags from the scan (nothing is written after a NaN coarse value, as the
// routine returns there); flags[2] is set where the routine would return
// with its error (a non-positive child of a positive field, a NaN slope).
__global__ void k_op_prolong_2a(double *__restrict__ gf, Int3 mf, const double *__restrict__ pc,
                                Int3 mc, Int3 pd, int *__restrict__ flags) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)pd.v[0] * pd.v[1] * pd.v[2]) return;
  if (flags[1]) return;
  const int nneg = flags[0];
  const int i = (int)(t % pd.v[0]) + 1, j = (int)((t / pd.v[0]) % pd.v[1]) + 1,
            k = (int)(t / ((long)pd.v[0] * pd.v[1])) + 1;
  const double w000 = pro_corner(pc, mc, i, j, k), w111 = pro_corner(pc, mc, i + 1, j + 1, k + 1),
               w100 = pro_corner(pc, mc, i + 1, j, k), w011 = pro_corner(pc, mc, i, j + 1, k + 1),
               w010 = pro_corner(pc, mc, i, j + 1, k), w101 = pro_corner(pc, mc, i + 1, j, k + 1),
               w001 = pro_corner(pc, mc, i, j, k + 1), w110 = pro_corner(pc, mc, i + 1, j + 1, k);
  const double fbar = PRO_P(1 + i, 1 + j, 1 + k);
  double d0 = pro_min(fabs(fbar - w000), fabs(fbar - w111)) * pro_sign(1.0, fbar - w000);
  if ((w111 - fbar) * (fbar - w000) <= 0.0) d0 = 0.0;
  double d1 = pro_min(fabs(fbar - w100), fabs(fbar - w011)) * pro_sign(1.0, fbar - w100);
  if ((w011 - fbar) * (fbar - w100) <= 0.0) d1 = 0.0;
  double d2 = pro_min(fabs(fbar - w010), fabs(fbar - w101)) * pro_sign(1.0, fbar - w010);
  if ((w101 - fbar) * (fbar - w010) <= 0.0) d2 = 0.0;
  double d3 = pro_min(fabs(fbar - w001), fabs(fbar - w110)) * pro_sign(1.0, fbar - w001);
  if ((w110 - fbar) * (fbar - w001) <= 0.0) d3 = 0.0;
  double s, sp;
  if (d0 == 0.0) {
    d0 = 1.0e-5 * pro_sign(1.0e-20, (d1 + d2 + d3));
    sp = 1.0;
  } else {
    s = (d1 + d2 + d3) / d0;
    sp = pro_min(pro_max(1.0 * s, 0.0), 1.0);
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
  frac = pro_min(frac, 1.0);
  if (c1 + c2 + c3 == 0) frac = 0.0;
  frac = pro_max(frac, 0.0);
  const double d1n = frac * c1 * d1 + (1.0 - c1) * d1;
  const double d2n = frac * c2 * d2 + (1.0 - c2) * d2;
  const double d3n = frac * c3 * d3 + (1.0 - c3) * d3;
  s = d1n + d2n + d3n;
  if ((nneg == 0 && (fbar + s <= 0.0 || fbar - s <= 0.0)) || s != s) atomicOr(&flags[2], 1);
  pro_children(gf, mf, i, j, k, fbar, d2n + d3n, d1n + d3n, d1n + d2n);
}

// int_lin3d ("2B"): corner-difference slopes, positivity (iposflag = 1)
// limiting them to 0.8 of the parent value
__global__ void k_op_prolong_2b(double *__restrict__ gf, Int3 mf, const double *__restrict__ pc,
                                Int3 mc, Int3 pd, int positive) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)pd.v[0] * pd.v[1] * pd.v[2]) return;
  const int i = (int)(t % pd.v[0]) + 1, j = (int)((t / pd.v[0]) % pd.v[1]) + 1,
            k = (int)(t / ((long)pd.v[0] * pd.v[1])) + 1;
  const double fbar = PRO_P(1 + i, 1 + j, 1 + k);
  double d0 = 0.5 * (pro_corner(pc, mc, i + 1, j + 1, k + 1) - pro_corner(pc, mc, i, j, k));
  double d1 = 0.5 * (pro_corner(pc, mc, i, j + 1, k + 1) - pro_corner(pc, mc, i + 1, j, k));
  double d2 = 0.5 * (pro_corner(pc, mc, i + 1, j, k + 1) - pro_corner(pc, mc, i, j + 1, k));
  double d3 = 0.5 * (pro_corner(pc, mc, i + 1, j + 1, k) - pro_corner(pc, mc, i, j, k + 1));
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
  if (d0 == 0.0) d0 = pro_sign(1.0e-20, d0 + d1 + d2);
  const double s = (d1 + d2 + d3) / d0;
  const double sp = pro_min(pro_max(s, 0.0), 1.0);
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
                (c1 * d1 + c2 * d2 + c3 * d3 + 1.0e-20);
  frac = pro_min(frac, 1.0);
  const double d1n = frac * c1 * d1 + (1.0 - c1) * d1;
  const double d2n = frac * c2 * d2 + (1.0 - c2) * d2;
  const double d3n = frac * c3 * d3 + (1.0 - c3) * d3;
  pro_children(gf, mf, i, j, k, fbar, d2n + d3n, d1n + d3n, d1n + d2n);
}

#undef PRO_P
#undef PRO_G

__global__ void k_op_prolong_linear(double *__restrict__ vf, Int3 mf, Int3 of,
                                    Int3 nf, const double *__restrict__ vc,
                                    Int3 mc, Int3 oc, Int3 gc, int rank,
                                    int accumulate) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  prolong_cell(vf, mf.v, of.v, nf.v, vc, mc.v, oc.v, gc.v, rank, accumulate, t);
}

__global__ void k_op_restrict_linear(double *__restrict__ vc, Int3 ndc, Int3 imc,
                                     Int3 nc, const double *__restrict__ vf,
                                     Int3 ndf, Int3 imf, int rank,
                                     int accumulate) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  restrict_cell(vc, ndc.v, imc.v, nc.v, vf, ndf.v, imf.v, rank, accumulate, t);
}

// one transfer per grid row
__global__ void k_op_transfer_batch(const enzoe_transfer *__restrict__ items) {
  const enzoe_transfer &it = items[blockIdx.y];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (it.kind == ENZOE_TRANSFER_PROLONG) {
    const int rank = (it.m_dst[1] == 1) ? 1 : ((it.m_dst[2] == 1) ? 2 : 3);
    int gc[3];
    for (int a = 0; a < 3; a++) gc[a] = (it.n_dst[a] == 2 * it.n_src[a]) ? 1 : 0;
    prolong_cell(it.dst, it.m_dst, it.o_dst, it.n_dst, it.src, it.m_src, it.o_src, gc,
                 rank, it.accumulate, t);
  } else {
    const int rank = (it.m_src[1] == 1) ? 1 : ((it.m_src[2] == 1) ? 2 : 3);
    restrict_cell(it.dst, it.m_dst, it.o_dst, it.n_dst, it.src, it.m_src, it.o_src, rank,
                  it.accumulate, t);
  }
}

bool region_in(const int m[3], const int o[3], const int n[3]) {
  for (int a = 0; a < 3; a++)
    if (m[a] < 1 || o[a] < 0 || n[a] < 1 || o[a] + n[a] > m[a]) return false;
  return true;
}

// region bounds, the size relation the reference asserts
// (problem_ProlongLinear.cpp:97-103) and a two-cell stencil inside the
// coarse region
bool prolong_ok(const int mf3[3], const int of3[3], const int nf3[3],
                const int mc3[3], const int oc3[3], const int nc3[3]) {
  if (!region_in(mf3, of3, nf3) || !region_in(mc3, oc3, nc3)) return false;
  const int rank = (mf3[1] == 1) ? 1 : ((mf3[2] == 1) ? 2 : 3);
  for (int a = 0; a < rank; a++) {
    if (!(nf3[a] == 2 * nc3[a] || nf3[a] == 2 * (nc3[a] - 2))) return false;
    if (nc3[a] < 2) return false;
  }
  return true;
}

bool restrict_ok(const int nd3_c[3], const int im3_c[3], const int n3_c[3],
                 const int nd3_f[3], const int im3_f[3]) {
  const int rank = (nd3_f[1] == 1) ? 1 : ((nd3_f[2] == 1) ? 2 : 3);
  int fine_n[3];
  for (int a = 0; a < 3; a++) fine_n[a] = a < rank ? 2 * n3_c[a] : 1;
  return region_in(nd3_c, im3_c, n3_c) && region_in(nd3_f, im3_f, fine_n);
}

bool even_extents(int nx, int ny, int nz) {
  return nx >= 2 && ny >= 2 && nz >= 2 && !(nx & 1) && !(ny & 1) && !(nz & 1);
}

}  // namespace

extern "C" {

long long enzoe_flux_data_size(int nf, int nx, int ny, int nz) {
  if (nf < 0 || nx < 1 || ny < 1 || nz < 1) return -1;
  FluxLayout L;
  L.init(nx, ny, nz);
  return (long long)(L.per_field * nf);
}

int enzoe_save_fluxes_for_corrections(const enzoe_fmap *flux, int dim,
                                      double cell_width, double dt, int ghost,
                                      int nf, const int *fields,
                                      double *flux_data,
                                      enzoe_stream_t stream) {
  KMap f;
  if (!valid_dim(dim) || ghost < 1 || nf < 1 || nf > 8 || !fields || !flux_data)
    return ENZOE_ERR_ARG;
  StoreSlots sl;
  sl.nf = nf;
  unsigned need = 0;
  for (int i = 0; i < 8; i++) sl.slot[i] = 0;
  for (int i = 0; i < nf; i++) {
    if (fields[i] < 0 || fields[i] >= ENZOE_MAP_SLOTS) return ENZOE_ERR_ARG;
    sl.slot[i] = fields[i];
    need |= 1u << fields[i];
  }
  if (!to_kmap(flux, need, f)) return ENZOE_ERR_ARG;
  int cell[3];
  for (int a = 0; a < 3; a++) cell[a] = f.n[a] + (a == dim ? 1 : 0);
  FluxLayout L;
  L.init(cell[0] - 2 * ghost, cell[1] - 2 * ghost, cell[2] - 2 * ghost);
  if (L.n[0] < 1 || L.n[1] < 1 || L.n[2] < 1) return ENZOE_ERR_ARG;
  // dt_dxi = dt/cell_width (EnzoMethodMHDVlct.cpp:265)
  const double dt_dxi = dt / cell_width;
  const long work = 2 * L.m[dim] * nf;
  hipLaunchKernelGGL(k_op_save_fluxes, dim3(nblk(work)), dim3(256), 0,
                     (hipStream_t)stream, f, sl, dim, dt_dxi, ghost, L,
                     flux_data);
  return launched();
}

int enzoe_flux_data_send_to_coarse(const double *fine_flux_data,
                                   double *coarse_flux_data, int nf, int nx,
                                   int ny, int nz, int axis, int face,
                                   const int child[3], enzoe_stream_t stream) {
  if (!fine_flux_data || !coarse_flux_data || nf < 1 || !valid_dim(axis) ||
      (face != 0 && face != 1) || !child || !even_extents(nx, ny, nz))
    return ENZOE_ERR_ARG;
  Child ch;
  for (int a = 0; a < 3; a++) {
    if (child[a] != 0 && child[a] != 1) return ENZOE_ERR_ARG;
    ch.c[a] = child[a];
  }
  FluxLayout L;
  L.init(nx, ny, nz);
  const long work = L.m[axis] * nf;
  hipLaunchKernelGGL(k_op_send_to_coarse, dim3(nblk(work)), dim3(256), 0,
                     (hipStream_t)stream, fine_flux_data, coarse_flux_data, nf,
                     L, axis, face, ch);
  return launched();
}

int enzoe_vlct_flux_send_to_coarse_batch(enzoe_vlct_t m,
                                         const enzoe_flux_link *links,
                                         int nlinks, int nf, int nx, int ny,
                                         int nz, enzoe_stream_t stream) {
  // one link per grid row (at most 65535)
  if (!m || !links || nlinks < 0 || nlinks > 65535 || nf < 1 ||
      !even_extents(nx, ny, nz))
    return ENZOE_ERR_ARG;
  if (nlinks == 0) return ENZOE_OK;
  for (int i = 0; i < nlinks; i++) {
    const enzoe_flux_link &k = links[i];
    if (!k.fine_flux_data || !k.coarse_flux_data || !valid_dim(k.axis) ||
        (k.face != 0 && k.face != 1))
      return ENZOE_ERR_ARG;
    for (int a = 0; a < 3; a++)
      if (k.child[a] != 0 && k.child[a] != 1) return ENZOE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (int rc = upload_table(&m->d_fltab, &m->fltab_bytes, &m->last_fltab, links,
                            (size_t)nlinks * sizeof(enzoe_flux_link), s))
    return rc;
  FluxLayout L;
  L.init(nx, ny, nz);
  const long work = std::max(L.m[0], std::max(L.m[1], L.m[2])) * nf;
  {
    KernelTimer kt(s, "k_send_to_coarse");
    hipLaunchKernelGGL(k_op_send_to_coarse_batch, dim3(nblk(work), nlinks), dim3(256), 0,
                       s, (const enzoe_flux_link *)m->d_fltab, nf, L);
  }
  return launched();
}

int enzoe_vlct_flux_correct_blocks(enzoe_vlct_t m,
                                   const enzoe_flux_correct_block *blocks,
                                   int n, int nf, int i_density,
                                   const int *conservative, int nx, int ny,
                                   int nz, int ghost, enzoe_stream_t stream) {
  // one block per grid row (at most 65535)
  if (!m || !blocks || n < 0 || n > 65535 || nf < 0 || nf > 16 || !conservative ||
      ghost < 0 || i_density >= nf || !even_extents(nx, ny, nz))
    return ENZOE_ERR_ARG;
  if (n == 0 || nf == 0) return ENZOE_OK;
  CorrectShared sh;
  std::memset(&sh, 0, sizeof(sh));
  sh.nf = nf;
  sh.i_density = i_density < 0 ? -1 : i_density;
  for (int i = 0; i < nf; i++) {
    sh.conservative[i] = (conservative[i] && i != sh.i_density) ? 1 : 0;
    // problem_MethodFluxCorrect.cpp:520-524: needs the density field
    if (sh.conservative[i] && sh.i_density < 0) return ENZOE_ERR_ARG;
  }
  for (int b = 0; b < n; b++) {
    if (!blocks[b].flux_data) return ENZOE_ERR_ARG;
    for (int i = 0; i < nf; i++)
      if (!blocks[b].fields[i]) return ENZOE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (int rc = upload_table(&m->d_fbtab, &m->fbtab_bytes, &m->last_fbtab, blocks,
                            (size_t)n * sizeof(enzoe_flux_correct_block), s))
    return rc;
  FluxLayout L;
  L.init(nx, ny, nz);
  const long shell = 2L * ny * nz + 2L * (nx - 2) * nz + 2L * (nx - 2) * (ny - 2);
  {
    KernelTimer kt(s, "k_flux_correct");
    hipLaunchKernelGGL(k_op_flux_correct_batch, dim3(nblk(shell), n), dim3(256), 0, s,
                       (const enzoe_flux_correct_block *)m->d_fbtab, sh, L, ghost,
                       nx + 2 * ghost, ny + 2 * ghost);
  }
  return launched();
}

int enzoe_source_gravity(double dt, const enzoe_fmap *integration,
                         const enzoe_fmap *dUcons, const enzoe_fmap *accel,
                         int stale_depth, enzoe_stream_t stream) {
  KMap u, d, a;
  if (stale_depth < 0 || !to_kmap(integration, 0xFu, u) || !to_kmap(dUcons, 0x1Eu, d) ||
      !to_kmap(accel, 0xEu, a))
    return ENZOE_ERR_ARG;
  if (!same_shape(u.n, d.n) || !same_shape(u.n, a.n)) return ENZOE_ERR_ARG;
  int lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = stale_depth;
    hi[k] = u.n[k] - stale_depth;
  }
  long count;
  Box box = make_box(lo, hi, count);
  if (count == 0) return ENZOE_OK;
  hipLaunchKernelGGL(k_op_gravity, dim3(nblk(count)), dim3(256), 0, (hipStream_t)stream,
                     u, d, a, box, dt);
  return launched();
}

int enzoe_prolong_linear(double *values_f, const int mf3[3], const int of3[3],
                         const int nf3[3], const double *values_c,
                         const int mc3[3], const int oc3[3], const int nc3[3],
                         int accumulate, enzoe_stream_t stream) {
  if (!values_f || !values_c || !mf3 || !of3 || !nf3 || !mc3 || !oc3 || !nc3)
    return ENZOE_ERR_ARG;
  if (!prolong_ok(mf3, of3, nf3, mc3, oc3, nc3)) return ENZOE_ERR_ARG;
  const int rank = (mf3[1] == 1) ? 1 : ((mf3[2] == 1) ? 2 : 3);
  Int3 mf, of, nf, mc, oc, gc;
  for (int a = 0; a < 3; a++) {
    mf.v[a] = mf3[a]; of.v[a] = of3[a]; nf.v[a] = nf3[a];
    mc.v[a] = mc3[a]; oc.v[a] = oc3[a];
    gc.v[a] = (nf3[a] == 2 * nc3[a]) ? 1 : 0;
  }
  const long work = (long)nf3[0] * nf3[1] * nf3[2];
  hipLaunchKernelGGL(k_op_prolong_linear, dim3(nblk(work)), dim3(256), 0,
                     (hipStream_t)stream, values_f, mf, of, nf, values_c, mc, oc, gc,
                     rank, accumulate ? 1 : 0);
  return launched();
}

int enzoe_enzo_prolong(int method, int positive, double *values_f, const int mf3[3],
                       const int of3[3], const int nf3[3], const double *values_c,
                       const int mc3[3], const int oc3[3], const int nc3[3], int *ierror,
                       enzoe_stream_t stream) {
  if (!values_f || !values_c || !mf3 || !of3 || !nf3 || !mc3 || !oc3 || !nc3)
    return ENZOE_ERR_ARG;
  if (method < 0 || method > 2) return ENZOE_ERR_CONFIG;
  for (int a = 0; a < 3; a++) {
    if (nf3[a] < 2 || nf3[a] % 2 != 0 || nc3[a] != nf3[a] / 2 + 2) return ENZOE_ERR_ARG;
    if (oc3[a] < 0 || oc3[a] + nc3[a] > mc3[a]) return ENZOE_ERR_ARG;
    if (of3[a] < 0 || of3[a] + nf3[a] > mf3[a]) return ENZOE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  Int3 mf, mc, pd;
  for (int a = 0; a < 3; a++) {
    mf.v[a] = mf3[a];
    mc.v[a] = mc3[a];
    pd.v[a] = nf3[a] / 2;
  }
  const double *pc = values_c + oc3[0] + (long)mc3[0] * (oc3[1] + (long)mc3[1] * oc3[2]);
  double *gf = values_f + of3[0] + (long)mf3[0] * (of3[1] + (long)mf3[1] * of3[2]);
  const long parents = (long)pd.v[0] * pd.v[1] * pd.v[2];
  if (ierror) *ierror = 0;
  if (method == 0) {
    hipLaunchKernelGGL(k_op_prolong_3a, dim3(nblk(parents)), dim3(256), 0, s, gf, mf, pc, mc, pd);
    return launched();
  }
  if (method == 2) {
    hipLaunchKernelGGL(k_op_prolong_2b, dim3(nblk(parents)), dim3(256), 0, s, gf, mf, pc, mc, pd,
                       positive == 1 ? 1 : 0);
    return launched();
  }
  // 2A: a scan of the corners first (the routine's count of non-positive
  // values decides its error test; a NaN coarse value stops it)
  int *flags = nullptr;
  if (hipMalloc((void **)&flags, 3 * sizeof(int)) != hipSuccess) return ENZOE_ERR_ALLOC;
  int rc = ENZOE_OK;
  if (hipMemsetAsync(flags, 0, 3 * sizeof(int), s) != hipSuccess) rc = ENZOE_ERR_HIP;
  if (!rc) {
    const long corners = (long)(pd.v[0] + 1) * (pd.v[1] + 1) * (pd.v[2] + 1);
    hipLaunchKernelGGL(k_op_prolong_2a_scan, dim3(nblk(corners)), dim3(256), 0, s, pc, mc, pd,
                       flags);
    hipLaunchKernelGGL(k_op_prolong_2a, dim3(nblk(parents)), dim3(256), 0, s, gf, mf, pc, mc, pd,
                       flags);
    rc = launched();
  }
  int host_flags[3] = {0, 0, 0};
  if (!rc && (hipMemcpyAsync(host_flags, flags, sizeof(host_flags), hipMemcpyDeviceToHost, s) !=
                  hipSuccess ||
              hipStreamSynchronize(s) != hipSuccess))
    rc = ENZOE_ERR_HIP;
  (void)hipStreamSynchronize(s);
  (void)hipFree(flags);
  if (!rc && ierror) *ierror = (host_flags[1] || host_flags[2]) ? 1 : 0;
  return rc;
}

int enzoe_vlct_transfer_batch(enzoe_vlct_t m, const enzoe_transfer *items,
                              int n, enzoe_stream_t stream) {
  // one transfer per grid row (at most 65535)
  if (!m || !items || n < 0 || n > 65535) return ENZOE_ERR_ARG;
  if (n == 0) return ENZOE_OK;
  long work = 0;
  int kinds = 0;
  for (int i = 0; i < n; i++) {
    const enzoe_transfer &it = items[i];
    if (!it.dst || !it.src) return ENZOE_ERR_ARG;
    if (it.kind == ENZOE_TRANSFER_PROLONG) {
      if (!prolong_ok(it.m_dst, it.o_dst, it.n_dst, it.m_src, it.o_src, it.n_src))
        return ENZOE_ERR_ARG;
    } else if (it.kind == ENZOE_TRANSFER_RESTRICT) {
      if (!restrict_ok(it.m_dst, it.o_dst, it.n_dst, it.m_src, it.o_src))
        return ENZOE_ERR_ARG;
    } else {
      return ENZOE_ERR_ARG;
    }
    kinds |= 1 << it.kind;
    work = std::max(work, (long)it.n_dst[0] * it.n_dst[1] * it.n_dst[2]);
  }
  hipStream_t s = (hipStream_t)stream;
  // restriction-only and other batches keep separate cached tables
  const int slot = (kinds == (1 << ENZOE_TRANSFER_RESTRICT)) ? 1 : 0;
  if (int rc = upload_table(&m->d_trtab[slot], &m->trtab_bytes[slot], &m->last_trtab[slot],
                            items, (size_t)n * sizeof(enzoe_transfer), s))
    return rc;
  {
    KernelTimer kt(s, slot ? "k_restrict_batch" : "k_prolong_batch");
    hipLaunchKernelGGL(k_op_transfer_batch, dim3(nblk(work), n), dim3(256), 0, s,
                       (const enzoe_transfer *)m->d_trtab[slot]);
  }
  return launched();
}

int enzoe_restrict_linear(double *values_c, const int nd3_c[3],
                          const int im3_c[3], const int n3_c[3],
                          const double *values_f, const int nd3_f[3],
                          const int im3_f[3], const int n3_f[3],
                          int accumulate, enzoe_stream_t stream) {
  if (!values_c || !values_f || !nd3_c || !im3_c || !n3_c || !nd3_f || !im3_f || !n3_f)
    return ENZOE_ERR_ARG;
  const int rank = (nd3_f[1] == 1) ? 1 : ((nd3_f[2] == 1) ? 2 : 3);
  if (!restrict_ok(nd3_c, im3_c, n3_c, nd3_f, im3_f)) return ENZOE_ERR_ARG;
  Int3 ndc, imc, nc, ndf, imf;
  for (int a = 0; a < 3; a++) {
    ndc.v[a] = nd3_c[a]; imc.v[a] = im3_c[a]; nc.v[a] = n3_c[a];
    ndf.v[a] = nd3_f[a]; imf.v[a] = im3_f[a];
  }
  const long work = (long)n3_c[0] * n3_c[1] * n3_c[2];
  hipLaunchKernelGGL(k_op_restrict_linear, dim3(nblk(work)), dim3(256), 0,
                     (hipStream_t)stream, values_c, ndc, imc, nc, values_f, ndf, imf, rank,
                     accumulate ? 1 : 0);
  return launched();
}

int enzoe_flux_correct(double *const *fields, int nf, int i_density,
                       const int *conservative, const double *flux_data,
                       int nx, int ny, int nz, int ghost, const int correct[6],
                       enzoe_stream_t stream) {
  if (!fields || nf < 0 || nf > 16 || !conservative || !flux_data || !correct ||
      ghost < 0 || i_density >= nf || !even_extents(nx, ny, nz))
    return ENZOE_ERR_ARG;
  if (nf == 0) return ENZOE_OK;
  CorrectArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nf = nf;
  a.i_density = i_density < 0 ? -1 : i_density;
  bool any = false;
  for (int k = 0; k < 6; k++) {
    a.correct[k] = correct[k] ? 1 : 0;
    any = any || a.correct[k];
  }
  for (int i = 0; i < nf; i++) {
    if (!fields[i]) return ENZOE_ERR_ARG;
    a.field[i] = fields[i];
    a.conservative[i] = (conservative[i] && i != a.i_density) ? 1 : 0;
    // problem_MethodFluxCorrect.cpp:520-524: needs the density field
    if (a.conservative[i] && a.i_density < 0) return ENZOE_ERR_ARG;
  }
  if (!any) return ENZOE_OK;
  FluxLayout L;
  L.init(nx, ny, nz);
  const long shell = 2L * ny * nz + 2L * (nx - 2) * nz + 2L * (nx - 2) * (ny - 2);
  hipLaunchKernelGGL(k_op_flux_correct, dim3(nblk(shell)), dim3(256), 0,
                     (hipStream_t)stream, a, flux_data, L, ghost,
                     nx + 2 * ghost, ny + 2 * ghost);
  return launched();
}

}  // extern "C"
