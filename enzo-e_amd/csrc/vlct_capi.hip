// vlct_capi.hip -- C ABI of the gfx950 VL+CT library (declared in
// include/enzoe_vlct.h): method configuration, timestep, compute, refresh,
// profiling.  Built with -ffp-contract=off: the FP64 path rounds exactly
// like the reference's CPU build.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "vlct_launch.hpp"

using namespace enzoe;

namespace enzoe {

bool g_profiling = false;
std::map<std::string, ProfEntry> g_prof;
std::vector<PendingTiming> g_pending;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pool_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void resolve_pending() {
  for (PendingTiming &p : g_pending) {
    (void)hipEventSynchronize(p.b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    ProfEntry &e = g_prof[p.name];
    e.total_ms += ms;
    e.count += 1;
    g_event_pool.push_back(p.a);
    g_event_pool.push_back(p.b);
  }
  g_pending.clear();
}

}  // namespace enzoe

namespace {

const char *kVersion = "enzoe-vlct 0.2 (gfx950)";

std::string lower(const char *s, const char *dflt) {
  std::string r = s ? s : dflt;
  std::transform(r.begin(), r.end(), r.begin(),
                 [](unsigned char c) { return (char)std::tolower(c); });
  return r;
}

// ------------------------------------------------- allocation-free compute
// The workspace of enzoe_vlct_compute_noalloc: one fixed array per device
// (each GPU the library runs on holds it until the process ends).  Host
// code never uses the array's name as an address: noalloc_workspace takes
// its device address once per device, with that device current.
constexpr size_t NOALLOC_BYTES = (size_t)64 << 20;
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

std::mutex g_noalloc_mtx;
std::map<int, char *> g_noalloc_ptr;

// The current device's address of the workspace.
int noalloc_workspace(char **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ENZOE_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_noalloc_mtx);
  auto it = g_noalloc_ptr.find(dev);
  if (it == g_noalloc_ptr.end()) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return ENZOE_ERR_HIP;
    it = g_noalloc_ptr.emplace(dev, (char *)p).first;
  }
  *out = it->second;
  return ENZOE_OK;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// ------------------------------------------------------------ diagnostics
// Accumulators of enzoe_vlct_diagnostics.  Minima are kept as order-
// preserving integer images of doubles (so an integer atomicMin is a double
// min for either sign); maxima of non-negative values as their bit patterns.
// The conserved totals are not summed with atomics (their order, and so
// their rounding, would change from run to run, by far more than the
// scheme's own conservation error over a large grid): every workgroup
// writes compensated partial sums, which the host adds in a fixed order.
struct DiagAcc {
  unsigned long long min_rho, min_p, max_divb, max_bface, nonfinite;
};
constexpr int DIAG_NSUM = 5;   // rho, rho v_x, rho v_y, rho v_z, rho e_tot
constexpr int DIAG_WG = 64;    // workgroups per block (grid-stride over cells)

__host__ __device__ inline unsigned long long double_to_ordered(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double ordered_to_double(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  return __builtin_bit_cast(double, b);
}

// s + c += x, Neumaier's compensated summation
__host__ __device__ inline void comp_add(double &s, double &c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}
// (s, c) += (s2, c2): Knuth's two-sum of the leading parts, the
// compensations added to the error term
__host__ __device__ inline void comp_merge(double &s, double &c, double s2, double c2) {
  const double t = s + s2;
  const double bp = t - s;
  const double err = (s - (t - bp)) + (s2 - bp);
  c = (c + c2) + err;
  s = t;
}

__global__ void k_diag_init(DiagAcc *a) {
  a->min_rho = a->min_p = double_to_ordered(1.7976931348623157e308);
  a->max_divb = a->max_bface = 0ull;
  a->nonfinite = 0ull;
}

__device__ inline unsigned long long wave_umin(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
__device__ inline unsigned long long wave_umax(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// DIAG_WG workgroups of 256 threads per block (grid.y = block), grid-stride
// over the block's cells, ghosts included: counts non-finite values of
// every field at the cell (and of the face fields' last layer at the
// block's high ends); over the active cells the totals of the conserved
// quantities (compensated, one partial pair per workgroup and quantity into
// part[((block * DIAG_WG + wg) * DIAG_NSUM + q) * 2 + {0, 1}]), the minima
// of density and pressure, the largest |div B| (times the smallest cell
// width) and the largest |face B|.  Every thread reaches the reductions.
template <bool MHD>
__global__ void __launch_bounds__(256)
k_diagnostics(const BlockDesc<double> *__restrict__ blocks, double gamma, DiagAcc *acc,
              double *__restrict__ part) {
  const BlockDesc<double> &b = blocks[blockIdx.y];
  const long ncell = (long)b.mx * b.my * b.mz;
  double s[DIAG_NSUM], cs[DIAG_NSUM];
#pragma unroll
  for (int q = 0; q < DIAG_NSUM; q++) s[q] = cs[q] = 0.;
  double rmin = 1.7976931348623157e308, pmin = rmin, divb = 0., bmax = 0.;
  unsigned long long bad = 0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += (long)gridDim.x * blockDim.x) {
    const int i = (int)(c % b.mx), j = (int)((c / b.mx) % b.my), k = (int)(c / ((long)b.mx * b.my));
    const int nf = MHD ? S_N : 5;
    for (int q = 0; q < nf; q++) bad += !isfinite(b.U[q][c]);
    if (MHD) {
      const long fx = idx3(k, j, i, b.my, b.mx + 1), fy = idx3(k, j, i, b.my + 1, b.mx),
                 fz = idx3(k, j, i, b.my, b.mx);
      bad += !isfinite(b.bfi[0][fx]) + !isfinite(b.bfi[1][fy]) + !isfinite(b.bfi[2][fz]);
      if (i == b.mx - 1) bad += !isfinite(b.bfi[0][fx + 1]);
      if (j == b.my - 1) bad += !isfinite(b.bfi[1][fy + b.mx]);
      if (k == b.mz - 1) bad += !isfinite(b.bfi[2][fz + (long)b.mx * b.my]);
    }
    const int g = b.g;
    if (i >= g && j >= g && k >= g && i < b.mx - g && j < b.my - g && k < b.mz - g) {
      const double d = b.U[S_RHO][c];
      const double vx = b.U[S_VX][c], vy = b.U[S_VY][c], vz = b.U[S_VZ][c];
      const double e = b.U[S_E][c];
      comp_add(s[0], cs[0], d);
      comp_add(s[1], cs[1], d * vx);
      comp_add(s[2], cs[2], d * vy);
      comp_add(s[3], cs[3], d * vz);
      comp_add(s[4], cs[4], d * e);
      rmin = fmin(rmin, d);
      double bx = 0., by = 0., bz = 0.;
      if (MHD) { bx = b.U[S_BX][c]; by = b.U[S_BY][c]; bz = b.U[S_BZ][c]; }
      pmin = fmin(pmin, cell_pressure<MHD>(d, vx, vy, vz, e, bx, by, bz, gamma));
      if (MHD) {
        const long fx = idx3(k, j, i, b.my, b.mx + 1), fy = idx3(k, j, i, b.my + 1, b.mx),
                   fz = idx3(k, j, i, b.my, b.mx);
        const double dv = (b.bfi[0][fx + 1] - b.bfi[0][fx]) / b.cw[0] +
                          (b.bfi[1][fy + b.mx] - b.bfi[1][fy]) / b.cw[1] +
                          (b.bfi[2][fz + (long)b.mx * b.my] - b.bfi[2][fz]) / b.cw[2];
        const double dvb = fabs(dv) * fmin(b.cw[0], fmin(b.cw[1], b.cw[2]));
        // NaN never wins a max: a non-finite cell is counted instead
        if (isfinite(dvb)) divb = fmax(divb, dvb);
        const double bm =
            fmax(fabs(b.bfi[0][fx]), fmax(fabs(b.bfi[1][fy]), fabs(b.bfi[2][fz])));
        if (isfinite(bm)) bmax = fmax(bmax, bm);
      }
    }
  }
  // compensated sums: wave butterfly, then the waves in order through LDS
#pragma unroll
  for (int q = 0; q < DIAG_NSUM; q++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double os = __shfl_xor(s[q], off, 64), oc = __shfl_xor(cs[q], off, 64);
      // both lanes of a pair merge in the same order (lower lane first), so
      // every lane ends with the same bits
      if ((threadIdx.x & off) == 0)
        comp_merge(s[q], cs[q], os, oc);
      else {
        double ls = os, lc = oc;
        comp_merge(ls, lc, s[q], cs[q]);
        s[q] = ls;
        cs[q] = lc;
      }
    }
  }
  __shared__ double wsum[4][DIAG_NSUM][2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < DIAG_NSUM; q++) {
      wsum[wave][q][0] = s[q];
      wsum[wave][q][1] = cs[q];
    }
  const unsigned long long r = wave_umin(double_to_ordered(rmin));
  const unsigned long long p = wave_umin(double_to_ordered(pmin));
  const unsigned long long dvb = wave_umax(__double_as_longlong(divb));
  const unsigned long long bm = wave_umax(__double_as_longlong(bmax));
  unsigned long long nb = bad;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) nb += __shfl_xor(nb, off, 64);
  if (lane == 0) {
    atomicMin(&acc->min_rho, r);
    atomicMin(&acc->min_p, p);
    atomicMax(&acc->max_divb, dvb);
    atomicMax(&acc->max_bface, bm);
    if (nb) atomicAdd(&acc->nonfinite, nb);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *out = part + ((long)blockIdx.y * gridDim.x + blockIdx.x) * DIAG_NSUM * 2;
    for (int q = 0; q < DIAG_NSUM; q++) {
      double ts = wsum[0][q][0], tc = wsum[0][q][1];
      for (int w = 1; w < (int)(blockDim.x >> 6); w++)
        comp_merge(ts, tc, wsum[w][q][0], wsum[w][q][1]);
      out[2 * q] = ts;
      out[2 * q + 1] = tc;
    }
  }
}

// Per-field L1 sums of a state against another (enzoe_vlct_state_l1):
// DIAG_WG workgroups per block pair (grid.y), grid-stride over the active
// cells, compensated per thread, then per wave and workgroup as the
// diagnostics sums; part[(pair * DIAG_WG + wg) * L1_NPART + q * 2 + {0,1}]
// with q = diff 0..7, ref 8..15, |v| 16, |B| 17; then at
// L1_NSUM * 2 + q the largest per-cell relative difference of field q
// (|a - b| over |b|, for a velocity or B component over the cell's |b_vec|;
// a NaN difference is kept).
constexpr int L1_NSUM = 18;
constexpr int L1_NPART = L1_NSUM * 2 + 8;
__device__ __forceinline__ double nan_max(double a, double b) {
  return (b > a || b != b) ? b : a;
}
template <bool MHD>
__global__ void __launch_bounds__(256)
k_state_l1(const BlockDesc<double> *__restrict__ da, const BlockDesc<double> *__restrict__ db,
           double *__restrict__ part) {
  const BlockDesc<double> &a = da[blockIdx.y];
  const BlockDesc<double> &b = db[blockIdx.y];
  const int g = b.g;
  const int nx = b.mx - 2 * g, ny = b.my - 2 * g, nz = b.mz - 2 * g;
  const long nact = (long)nx * ny * nz;
  constexpr int NQ = MHD ? S_N : 5;
  double s[L1_NSUM], cs[L1_NSUM], mr[8];
#pragma unroll
  for (int q = 0; q < L1_NSUM; q++) s[q] = cs[q] = 0.;
#pragma unroll
  for (int q = 0; q < 8; q++) mr[q] = 0.;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nact;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % nx) + g, j = (int)((t / nx) % ny) + g,
              k = (int)(t / ((long)nx * ny)) + g;
    const long c = idx3(k, j, i, b.my, b.mx);
    double vb[S_N], df[S_N];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      vb[q] = b.U[q][c];
      df[q] = fabs(a.U[q][c] - vb[q]);
      comp_add(s[q], cs[q], df[q]);
      comp_add(s[8 + q], cs[8 + q], fabs(vb[q]));
    }
    const double vmag = sqrt(vb[S_VX] * vb[S_VX] + vb[S_VY] * vb[S_VY] + vb[S_VZ] * vb[S_VZ]);
    comp_add(s[16], cs[16], vmag);
    double bmag = 0.;
    if (MHD) {
      bmag = sqrt(vb[S_BX] * vb[S_BX] + vb[S_BY] * vb[S_BY] + vb[S_BZ] * vb[S_BZ]);
      comp_add(s[17], cs[17], bmag);
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const double den = (q >= S_VX && q <= S_VZ) ? vmag
                         : (q >= S_BX)            ? bmag
                                                  : fabs(vb[q]);
      mr[q] = nan_max(mr[q], df[q] == 0. ? 0. : df[q] / den);
    }
  }
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mr[q] = nan_max(mr[q], __shfl_xor(mr[q], off, 64));
#pragma unroll
  for (int q = 0; q < L1_NSUM; q++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double os = __shfl_xor(s[q], off, 64), oc = __shfl_xor(cs[q], off, 64);
      if ((threadIdx.x & off) == 0)
        comp_merge(s[q], cs[q], os, oc);
      else {
        double ls = os, lc = oc;
        comp_merge(ls, lc, s[q], cs[q]);
        s[q] = ls;
        cs[q] = lc;
      }
    }
  }
  __shared__ double wsum[4][L1_NSUM][2];
  __shared__ double wmax[4][8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    for (int q = 0; q < L1_NSUM; q++) {
      wsum[wave][q][0] = s[q];
      wsum[wave][q][1] = cs[q];
    }
    for (int q = 0; q < 8; q++) wmax[wave][q] = mr[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double *out = part + ((long)blockIdx.y * gridDim.x + blockIdx.x) * L1_NPART;
    for (int q = 0; q < L1_NSUM; q++) {
      double ts = wsum[0][q][0], tc = wsum[0][q][1];
      for (int w = 1; w < (int)(blockDim.x >> 6); w++)
        comp_merge(ts, tc, wsum[w][q][0], wsum[w][q][1]);
      out[2 * q] = ts;
      out[2 * q + 1] = tc;
    }
    for (int q = 0; q < 8; q++) {
      double mx = wmax[0][q];
      for (int w = 1; w < (int)(blockDim.x >> 6); w++) mx = nan_max(mx, wmax[w][q]);
      out[L1_NSUM * 2 + q] = mx;
    }
  }
}

__global__ void k_calibration_copy(double *__restrict__ dst,
                                   const double *__restrict__ src, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

extern "C" {

const char *enzoe_version(void) { return kVersion; }

int enzoe_vlct_create(const enzoe_vlct_params *p, enzoe_vlct_t *out) {
  if (!p || !out) return ENZOE_ERR_ARG;
  *out = nullptr;
  enzoe_vlct_method *m = new enzoe_vlct_method();
  // parse_mhdchoice_pack_pair_ (EnzoMethodMHDVlct.cpp:38-86)
  const std::string ts = lower(p->time_scheme, "vl");
  const std::string rn = lower(p->reconstruct_method, "plm");
  int r1;
  if (parse_recon(rn, &r1)) { delete m; return ENZOE_ERR_CONFIG; }
  if (ts == "vl") {
    m->nstages = 2;
    m->recon[0] = RC_NN;
    m->recon[1] = r1;
  } else if (ts == "euler") {
    m->nstages = 1;
    m->recon[0] = r1;
  } else {
    delete m;
    return ENZOE_ERR_CONFIG;
  }
  // EnzoMHDIntegratorStageCommands::parse_bfield_choice_ (:76-98)
  if (!p->mhd_choice) { delete m; return ENZOE_ERR_CONFIG; }
  const std::string mc = lower(p->mhd_choice, "");
  if (mc == "constrained_transport") m->mhd = 1;
  else if (mc == "no_bfield") m->mhd = 0;
  else { delete m; return ENZOE_ERR_CONFIG; }
  // EnzoRiemann::construct_riemann (EnzoRiemann.cpp:26-68)
  const std::string rs = lower(p->riemann_solver, "hlld");
  if (rs == "hlld") {
    if (!m->mhd) { delete m; return ENZOE_ERR_CONFIG; }
    m->riemann = RS_HLLD;
  } else if (rs == "hlle") {
    if (!m->mhd) { delete m; return ENZOE_ERR_CONFIG; }  // hydro HLLE errors
    m->riemann = RS_HLLE;
  } else if (rs == "hllc") {
    if (m->mhd) { delete m; return ENZOE_ERR_CONFIG; }
    m->riemann = RS_HLLC;
  } else {
    delete m;  // "hll" (Davis) aborts in the reference
    return ENZOE_ERR_CONFIG;
  }
  // EnzoReconstructor::construct_reconstructor: 1 <= theta <= 2
  m->theta = (p->theta_limiter < 0) ? 1.5 : p->theta_limiter;
  if (!(m->theta >= 1.0 && m->theta <= 2.0)) { delete m; return ENZOE_ERR_CONFIG; }
  m->courant = (p->courant < 0) ? (m->nstages == 2 ? 0.3 : 1.0) : p->courant;
  m->gamma = p->gamma;
  if (!(m->gamma > 1.0)) { delete m; return ENZOE_ERR_CONFIG; }
  // floors must exist (EnzoMHDIntegratorStageCommands.cpp:35-40)
  if (!(p->density_floor > 0) || !(p->pressure_floor > 0)) {
    delete m;
    return ENZOE_ERR_CONFIG;
  }
  m->dfloor = p->density_floor;
  m->pfloor = p->pressure_floor;
  if (p->precision != ENZOE_PREC_DOUBLE && p->precision != ENZOE_PREC_MIXED &&
      p->precision != ENZOE_PREC_DOUBLE_FMA) {
    delete m;
    return ENZOE_ERR_CONFIG;
  }
  m->precision = p->precision;
  // the dual-energy formalism: "modern" or disabled
  // (EnzoMHDIntegratorStageCommands.cpp:31-34, EnzoConfig.cpp:980-1008)
  {
    const std::string de = lower(p->dual_energy, "disabled");
    if (de == "modern") {
      m->dual = 1;
      m->eta = (p->dual_energy_eta < 0) ? 0.001 : p->dual_energy_eta;
      // the FP32 flux path carries no internal-energy terms
      if (m->precision == ENZOE_PREC_MIXED) { delete m; return ENZOE_ERR_CONFIG; }
    } else if (de != "disabled") {
      delete m;
      return ENZOE_ERR_CONFIG;
    }
  }
  // CT requires exactly 2 partial timesteps (EnzoBfieldMethod.cpp:23-27)
  if (m->mhd && m->nstages != 2) { delete m; return ENZOE_ERR_CONFIG; }
  {
    const char *e = std::getenv("ENZOE_XCD_REMAP");
    if (e && *e) m->xcd_remap = std::atoi(e) & 15;
  }
  // no device call here: device memory is taken on first use, so a
  // configuration can be validated on a host without a GPU
  *out = m;
  return ENZOE_OK;
}

int enzoe_vlct_destroy(enzoe_vlct_t m) {
  if (!m) return ENZOE_ERR_ARG;
  if (m->scratch) (void)hipFree(m->scratch);
  if (m->d_desc) (void)hipFree(m->d_desc);
  if (m->d_dt) (void)hipFree(m->d_dt);
  if (m->d_ptrtab) (void)hipFree(m->d_ptrtab);
  if (m->d_dts) (void)hipFree(m->d_dts);
  if (m->d_rtab) (void)hipFree(m->d_rtab);
  for (void *p : {m->d_fdtab, m->d_fltab, m->d_fbtab, m->d_trtab[0], m->d_trtab[1],
                  (void *)m->d_dbg, m->d_bar})
    if (p) (void)hipFree(p);
  delete m;
  return ENZOE_OK;
}

double enzoe_vlct_courant(enzoe_vlct_t m) { return m ? m->courant : -1.0; }

namespace {

// blocks grouped into runs of equal shape, each run one compute_group
int compute_runs(enzoe_vlct_method *m, const enzoe_block *blocks, int n,
                 double dt, const double *d_dts, hipStream_t s,
                 const StoreArgs *st = nullptr,
                 const double *const *accel = nullptr,
                 unsigned long long *cfl = nullptr) {
  for (int i = 0; i < n; i++)
    if (int rc = check_block(m, blocks[i])) return rc;
  int b0 = 0;
  while (b0 < n) {
    int b1 = b0 + 1;
    while (b1 < n && blocks[b1].mx == blocks[b0].mx &&
           blocks[b1].my == blocks[b0].my && blocks[b1].mz == blocks[b0].mz)
      b1++;
    const double *dts = d_dts ? d_dts + b0 : nullptr;
    StoreArgs run_st;
    if (st) {
      run_st = *st;
      run_st.d_fdtab = st->d_fdtab + b0;
    }
    const StoreArgs *rs = st ? &run_st : nullptr;
    const double *const *ra = accel ? accel + 3 * b0 : nullptr;
    int rc;
    if (m->precision == ENZOE_PREC_DOUBLE && m->dual) {
      rc = m->mhd ? compute_group<double, true, true>(m, blocks + b0, b1 - b0, dt, dts, s,
                                                      rs, ra, cfl)
                  : compute_group<double, false, true>(m, blocks + b0, b1 - b0, dt, dts, s,
                                                       rs, ra, cfl);
    } else if (m->precision == ENZOE_PREC_DOUBLE) {
      rc = m->mhd ? compute_group<double, true>(m, blocks + b0, b1 - b0, dt, dts, s, rs, ra,
                                                cfl)
                  : compute_group<double, false>(m, blocks + b0, b1 - b0, dt, dts, s, rs,
                                                 ra, cfl);
    } else if (m->precision == ENZOE_PREC_DOUBLE_FMA) {
      rc = compute_contracted(m, blocks + b0, b1 - b0, dt, dts, s, rs, ra, cfl);
    } else {
      rc = compute_mixed(m, blocks + b0, b1 - b0, dt, dts, s, rs, ra, cfl);
    }
    if (rc) return rc;
    b0 = b1;
  }
  return ENZOE_OK;
}

}  // namespace

int enzoe_vlct_compute(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                       double dt, enzoe_stream_t stream) {
  if (!m || !blocks || n <= 0) return ENZOE_ERR_ARG;
  return compute_runs(m, blocks, n, dt, nullptr, (hipStream_t)stream);
}

int enzoe_vlct_compute_noalloc(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                               double dt, enzoe_stream_t stream) {
  if (!m || !blocks || n <= 0) return ENZOE_ERR_ARG;
  char *ws = nullptr;
  if (int rc = noalloc_workspace(&ws)) return rc;
  m->arena = ws;
  m->arena_bytes = NOALLOC_BYTES;
  const int rc = compute_runs(m, blocks, n, dt, nullptr, (hipStream_t)stream);
  m->arena = nullptr;
  m->arena_bytes = 0;
  return rc;
}

size_t enzoe_vlct_noalloc_bytes(void) { return NOALLOC_BYTES; }

int enzoe_vlct_compute_dt(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                          const double *dt, int dt_on_device,
                          enzoe_stream_t stream) {
  if (!m || !blocks || n <= 0 || !dt) return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const double *d_dts = dt;
  if (!dt_on_device) {
    for (int i = 0; i < n; i++)
      if (!(dt[i] >= 0.0)) return ENZOE_ERR_ARG;
    if (int rc = ensure_bytes(&m->d_dts, &m->dts_bytes, n * sizeof(double))) return rc;
    // the copy completes before the host array may change
    if (hipMemcpyAsync(m->d_dts, dt, n * sizeof(double), hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return ENZOE_ERR_HIP;
    d_dts = (const double *)m->d_dts;
  }
  return compute_runs(m, blocks, n, 0.0, d_dts, s);
}

int enzoe_vlct_compute_gravity(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                               double dt, double *const *accel,
                               enzoe_stream_t stream) {
  if (!m || !blocks || n <= 0 || !accel) return ENZOE_ERR_ARG;
  for (int i = 0; i < 3 * n; i++)
    if (!accel[i]) return ENZOE_ERR_ARG;
  return compute_runs(m, blocks, n, dt, nullptr, (hipStream_t)stream, nullptr,
                      (const double *const *)accel);
}

int enzoe_vlct_compute_store_fluxes(enzoe_vlct_t m, const enzoe_block *blocks,
                                    int n, double dt, int nf,
                                    const int *fields,
                                    double *const *flux_data,
                                    enzoe_stream_t stream) {
  if (!m || !blocks || n <= 0 || !fields || !flux_data) return ENZOE_ERR_ARG;
  // EnzoMethodMHDVlct.cpp:137-141: flux corrections are hydro-only
  if (m->mhd) return ENZOE_ERR_CONFIG;
  if (nf < 1 || nf > 8) return ENZOE_ERR_ARG;
  StoreArgs st;
  st.sl.nf = nf;
  for (int i = 0; i < 8; i++) st.sl.slot[i] = 0;
  for (int i = 0; i < nf; i++) {
    if (fields[i] < ENZOE_SLOT_DENSITY || fields[i] > ENZOE_SLOT_ENERGY)
      return ENZOE_ERR_ARG;
    st.sl.slot[i] = fields[i];
  }
  for (int i = 0; i < n; i++) {
    const enzoe_block &b = blocks[i];
    const int ext[3] = {b.mx - 2 * b.ghost, b.my - 2 * b.ghost, b.mz - 2 * b.ghost};
    // FaceFluxes needs even block dimensions (data_FaceFluxes.cpp:38-45)
    for (int a = 0; a < 3; a++)
      if (ext[a] < 2 || (ext[a] & 1)) return ENZOE_ERR_ARG;
    if (!flux_data[i]) return ENZOE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  if (int rc = upload_table(&m->d_fdtab, &m->fdtab_bytes, &m->last_fdtab, flux_data,
                            n * sizeof(double *), s))
    return rc;
  st.d_fdtab = (double *const *)m->d_fdtab;
  return compute_runs(m, blocks, n, dt, nullptr, s, &st);
}

}  // extern "C"

namespace {

// Launches the CFL reduction of the blocks into m->d_dt (bit pattern of
// the unscaled min); the descriptor table is uploaded only when the block
// list changed, so the call makes no host round trip otherwise.
// d_time / t_stop: the device time of a device-driven cycle and its stop
// time (a cycle past it leaves the dual-energy fields alone), or null.
// Descriptors of the CFL kernel for blocks [0, n): only U, pressure, the
// internal energy, shape and widths are used.
void fill_dt_desc(const enzoe_vlct_method *m, const enzoe_block *blocks, int n,
                  BlockDesc<double> *out) {
  for (int i = 0; i < n; i++) {
    BlockDesc<double> &d = out[i];
    std::memset(&d, 0, sizeof(d));
    const enzoe_block &b = blocks[i];
    d.U[S_RHO] = b.density;
    for (int a = 0; a < 3; a++) d.U[S_VX + a] = b.velocity[a];
    d.U[S_E] = b.total_energy;
    if (m->mhd)
      for (int a = 0; a < 3; a++) d.U[S_BX + a] = b.bfield[a];
    d.pressure = b.pressure;
    if (m->dual) d.ei = b.internal_energy;
    d.mx = b.mx; d.my = b.my; d.mz = b.mz;
    for (int a = 0; a < 3; a++) d.cw[a] = b.cell_width[a];
  }
}

// The CFL reduction of n blocks whose descriptors are table[0 .. n) (device)
// into *target: its initialisation, then the kernel.
int launch_dt_on(const enzoe_vlct_method *m, const BlockDesc<double> *table, int n,
                 long maxcells, unsigned long long *target, hipStream_t s,
                 const double *d_time, double t_stop) {
  hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, s, target);
  KernelTimer kt(s, "k_dt");
  // enough workgroups to fill the device; each thread strides over cells
  const unsigned wg = (unsigned)std::max<long>(1, std::min<long>(nblk(maxcells, 256),
                                                                 std::max<long>(1, 2048 / n)));
  dim3 grid(wg, n);
  const DualParams dp = {m->pfloor, m->eta, sync_ggm1(m->gamma), d_time, t_stop};
  if (m->mhd && m->dual)
    hipLaunchKernelGGL((k_dt<double, true, true>), grid, dim3(256), 0, s, table, m->gamma,
                       target, dp);
  else if (m->mhd)
    hipLaunchKernelGGL((k_dt<double, true>), grid, dim3(256), 0, s, table, m->gamma, target,
                       dp);
  else if (m->dual)
    hipLaunchKernelGGL((k_dt<double, false, true>), grid, dim3(256), 0, s, table, m->gamma,
                       target, dp);
  else
    hipLaunchKernelGGL((k_dt<double, false>), grid, dim3(256), 0, s, table, m->gamma, target,
                       dp);
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int launch_dt(enzoe_vlct_method *m, const enzoe_block *blocks, int n, hipStream_t s,
              const double *d_time = nullptr, double t_stop = 0.0) {
  for (int i = 0; i < n; i++)
    if (int rc = check_block(m, blocks[i])) return rc;
  if (!m->d_dt &&
      hipMalloc((void **)&m->d_dt, sizeof(unsigned long long)) != hipSuccess) {
    m->d_dt = nullptr;
    return ENZOE_ERR_ALLOC;
  }
  std::vector<BlockDesc<double>> desc(n);
  fill_dt_desc(m, blocks, n, desc.data());
  const size_t dbytes = desc.size() * sizeof(BlockDesc<double>);
  const char *raw = (const char *)desc.data();
  if (m->last_dt_desc.size() != dbytes ||
      std::memcmp(m->last_dt_desc.data(), raw, dbytes) != 0) {
    if (int rc = ensure_bytes(&m->d_ptrtab, &m->ptrtab_bytes, dbytes)) return rc;
    if (hipMemcpyAsync(m->d_ptrtab, raw, dbytes, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return ENZOE_ERR_HIP;
    m->last_dt_desc.assign(raw, raw + dbytes);
  }
  // with dual energy the reduction also writes the energies: a dry call
  // (graph set-up) uploads the table only
  if (m->dry) return ENZOE_OK;
  long maxcells = 0;
  for (int i = 0; i < n; i++)
    maxcells = std::max<long>(maxcells, (long)blocks[i].mx * blocks[i].my * blocks[i].mz);
  return launch_dt_on(m, (const BlockDesc<double> *)m->d_ptrtab, n, maxcells, m->d_dt, s,
                      d_time, t_stop);
}

}  // namespace

extern "C" {

int enzoe_vlct_timestep(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                        enzoe_stream_t stream, double *dt_out) {
  if (!m || !blocks || n <= 0 || !dt_out) return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (int rc = launch_dt(m, blocks, n, s)) return rc;
  unsigned long long bits = 0;
  if (hipMemcpyAsync(&bits, m->d_dt, sizeof(bits), hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return ENZOE_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return ENZOE_ERR_HIP;
  double dtb;
  std::memcpy(&dtb, &bits, sizeof(dtb));
  *dt_out = dtb * m->courant;
  return ENZOE_OK;
}

int enzoe_vlct_compute_timestep(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                                double dt, enzoe_stream_t stream, double *dt_next) {
  if (!m || !blocks || n <= 0 || !dt_next) return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n; i++)
    if (int rc = check_block(m, blocks[i])) return rc;
  if (!m->d_dt &&
      hipMalloc((void **)&m->d_dt, sizeof(unsigned long long)) != hipSuccess) {
    m->d_dt = nullptr;
    return ENZOE_ERR_ALLOC;
  }
  // the final k_cells covers every cell: the update, and the CFL term of
  // each cell of the block into m->d_dt
  hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, s, m->d_dt);
  if (int rc = compute_runs(m, blocks, n, dt, nullptr, s, nullptr, nullptr,
                            (unsigned long long *)m->d_dt))
    return rc;
  unsigned long long bits = 0;
  if (hipMemcpyAsync(&bits, m->d_dt, sizeof(bits), hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return ENZOE_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return ENZOE_ERR_HIP;
  double dtb;
  std::memcpy(&dtb, &bits, sizeof(dtb));
  *dt_next = dtb * m->courant;
  return ENZOE_OK;
}

int enzoe_vlct_timestep_device(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                               double *d_time, double t_stop, double *d_dts,
                               int ndts, enzoe_stream_t stream) {
  if (!m || !blocks || n <= 0 || !d_time || !d_dts || ndts <= 0)
    return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (int rc = launch_dt(m, blocks, n, s, d_time, t_stop)) return rc;
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, s, m->d_dt, m->courant, d_time,
                     t_stop, d_dts, ndts, 0);
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int enzoe_vlct_autotune(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                        double dt, enzoe_stream_t stream, int *mask_out) {
  if (!m || !blocks || n <= 0 || !(dt > 0.0)) return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n; i++)
    if (int rc = check_block(m, blocks[i])) return rc;
  const int mx = blocks[0].mx, my = blocks[0].my, mz = blocks[0].mz;
  // sample: the leading blocks of the first shape, at most ~2 GiB of fields
  const size_t cells = (size_t)mx * my * mz;
  const size_t faces = (size_t)(mx + 1) * my * mz + (size_t)mx * (my + 1) * mz +
                       (size_t)mx * my * (mz + 1);
  const size_t per =
      ((6 + (m->dual ? 1 : 0)) * cells + (m->mhd ? 3 * cells + faces : 0)) * sizeof(double);
  int ns = 0;
  while (ns < n && blocks[ns].mx == mx && blocks[ns].my == my && blocks[ns].mz == mz &&
         (ns == 0 || (size_t)(ns + 1) * per <= ((size_t)2 << 30)))
    ns++;
  char *pool = nullptr;
  if (hipMalloc((void **)&pool, per * ns) != hipSuccess) return ENZOE_ERR_ALLOC;
  std::vector<enzoe_block> copy(blocks, blocks + ns);
  int rc = ENZOE_OK;
  {
    char *p = pool;
    auto take = [&](double *&dst, const double *src, size_t count) {
      double *d = (double *)p;
      p += count * sizeof(double);
      if (src && rc == ENZOE_OK &&
          hipMemcpyAsync(d, src, count * sizeof(double), hipMemcpyDeviceToDevice, s) !=
              hipSuccess)
        rc = ENZOE_ERR_HIP;
      dst = d;
    };
    for (int i = 0; i < ns; i++) {
      enzoe_block &b = copy[i];
      take(b.density, blocks[i].density, cells);
      for (int a = 0; a < 3; a++) take(b.velocity[a], blocks[i].velocity[a], cells);
      take(b.total_energy, blocks[i].total_energy, cells);
      take(b.pressure, blocks[i].pressure, cells);
      if (m->dual) take(b.internal_energy, blocks[i].internal_energy, cells);
      if (m->mhd) {
        for (int a = 0; a < 3; a++) take(b.bfield[a], blocks[i].bfield[a], cells);
        for (int a = 0; a < 3; a++)
          take(b.bfieldi[a], blocks[i].bfieldi[a],
               (size_t)(mx + (a == 0)) * (my + (a == 1)) * (mz + (a == 2)));
      }
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == ENZOE_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = ENZOE_ERR_HIP;
  const bool prof = g_profiling;
  g_profiling = false;
  int best = m->xcd_remap, best_ctu = 0;
  float best_ms = 1e30f;
  // the launch structures: the three CT/update launches with every
  // combination of XCD-contiguous order over the four stage kernels, and
  // (MHD, setting "auto") the fused launch with each tile shape and the
  // orders of its two kernels (k_flux bit 0, k_ctu bit 3)
  std::vector<std::pair<int, int>> cand;  // (launch structure, mask)
  const int fixed = m->ctu;
  const bool ct_fused = m->mhd && m->nstages == 2;
  for (int v = 0; v <= 5; v++) {
    if (fixed >= 0 && v != fixed) continue;
    if (v >= 1 && v <= 4 && !ct_fused) continue;
    if (v == 5) {
      // the persistent stage launch walks its own order
      if (persistent_ok(m)) cand.push_back({5, 0});
      continue;
    }
    for (int mask = 0; mask < 16; mask++)
      if (v == 0 || (mask & (v == 4 ? 4 : 6)) == 0) cand.push_back({v, mask});
  }
  if (persistent_ok(m) && (fixed == -1 || fixed == 5))
    if (int rc2 = ensure_barrier(m)) rc = rc2;
  // the time of one step of the sample under the forced settings: a timed
  // warm-up run sets the repetitions (about 2 ms of work, 3 to 40 runs, so
  // that small blocks are not timed on launch noise), then their average
  auto time_runs = [&](float *ms_per_run) -> int {
    (void)hipEventRecord(e0, s);
    if (int r = compute_runs(m, copy.data(), ns, dt, nullptr, s)) return r;
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) return ENZOE_ERR_HIP;
    float warm = 0.f;
    (void)hipEventElapsedTime(&warm, e0, e1);
    const int reps = (int)std::min(40.f, std::max(3.f, std::ceil(2.f / std::max(warm, 1e-3f))));
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; r++)
      if (int rr = compute_runs(m, copy.data(), ns, dt, nullptr, s)) return rr;
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) return ENZOE_ERR_HIP;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_run = ms / reps;
    return ENZOE_OK;
  };
  for (size_t ci = 0; ci < cand.size() && rc == ENZOE_OK; ci++) {
    const int mask = cand[ci].second;
    m->force_mask = mask;
    m->force_ctu = cand[ci].first;
    float ms = 0.f;
    rc = time_runs(&ms);
    if (rc) break;
    if (cand[ci].first == 5) {
      // a persistent launch whose barrier wait ran out is no candidate
      GridBarrier gb;
      if (hipMemcpy(&gb, m->d_bar, sizeof(gb), hipMemcpyDeviceToHost) != hipSuccess) {
        rc = ENZOE_ERR_HIP;
        break;
      }
      if (gb.abort != 0u || gb.count != 0u) {
        (void)hipMemset(m->d_bar, 0, sizeof(GridBarrier));
        continue;
      }
    }
    if (ms < best_ms) {
      best_ms = ms;
      best = mask;
      best_ctu = cand[ci].first;
    }
  }
  m->force_mask = -1;
  m->force_ctu = -1;
  // then the rows of the tiled cell order, for the chosen structure and
  // order (when that order tiles at all)
  const int rows0 = tile_rows_of(m, mx, my, mz);
  int best_rows = rows0;
  if (best != 0 && rc == ENZOE_OK) {
    m->force_mask = best;
    m->force_ctu = best_ctu;
    for (int rows : {0, 8, 16, 32, 64}) {
      if (rows == rows0 || rows >= my) continue;
      m->force_rows = rows;
      float ms = 0.f;
      rc = time_runs(&ms);
      if (rc) break;
      if (ms < best_ms) {
        best_ms = ms;
        best_rows = rows;
      }
    }
    m->force_mask = -1;
    m->force_ctu = -1;
    m->force_rows = -1;
  }
  // then the other shapes of the flux launch (one face per thread: more
  // workgroups for small blocks; z pencils for the nearest-neighbour
  // stage), for the chosen structure, order and rows
  int best_shape = 0;
  if (best_ctu != 5 && m->flux_shape < 0 && rc == ENZOE_OK) {
    m->force_mask = best;
    m->force_ctu = best_ctu;
    m->force_rows = best_rows;
    for (int shape = 1; shape <= 2 && rc == ENZOE_OK; shape++) {
      m->force_shape = shape;
      float ms = 0.f;
      rc = time_runs(&ms);
      if (rc == ENZOE_OK && ms < best_ms) {
        best_ms = ms;
        best_shape = shape;
      }
    }
    m->force_mask = -1;
    m->force_ctu = -1;
    m->force_rows = -1;
    m->force_shape = -1;
  }
  g_profiling = prof;
  (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(pool);
  if (rc) return rc;
  m->tuned_mask[shape_key(mx, my, mz)] = best;
  m->tuned_ctu[shape_key(mx, my, mz)] = best_ctu;
  m->tuned_rows[shape_key(mx, my, mz)] = best_rows;
  m->tuned_shape[shape_key(mx, my, mz)] = best_shape;
  if (mask_out) *mask_out = best;
  return ENZOE_OK;
}

int enzoe_vlct_launch_choice(enzoe_vlct_t m, int mx, int my, int mz, int *mask, int *fused_ct,
                             int *tile_rows, int *flux_shape) {
  if (!m) return ENZOE_ERR_ARG;
  if (flux_shape) *flux_shape = flux_shape_of(m, mx, my, mz);
  if (mask) *mask = launch_mask(m, mx, my, mz);
  if (tile_rows) *tile_rows = tile_rows_of(m, mx, my, mz);
  if (fused_ct) {
    const int v = ctu_variant(m, mx, my, mz);
    *fused_ct = (v == 5) ? (persistent_ok(m) ? 5 : 0)
                         : ((m->mhd && ctu_possible(m)) ? v : 0);
  }
  return ENZOE_OK;
}

// The refresh of a uniform grid; cycle_dt (device, may be null): the dt of
// a device-driven cycle, the refresh does nothing when it is 0.
static int grid_refresh(enzoe_vlct_t m, const enzoe_block *blocks, int nbx, int nby,
                        int nbz, const int bc[3], hipStream_t s,
                        const double *cycle_dt) {
  if (!m || !blocks || nbx <= 0 || nby <= 0 || nbz <= 0 || !bc)
    return ENZOE_ERR_ARG;
  const int nb = nbx * nby * nbz;
  RefreshGrid G;
  G.dt = cycle_dt;
  G.nb[0] = nbx; G.nb[1] = nby; G.nb[2] = nbz;
  G.g = blocks[0].ghost;
  G.n[0] = blocks[0].mx - 2 * G.g;
  G.n[1] = blocks[0].my - 2 * G.g;
  G.n[2] = blocks[0].mz - 2 * G.g;
  for (int a = 0; a < 3; a++) G.bc[a] = bc[a];
  for (int i = 0; i < nb; i++) {
    if (blocks[i].mx != blocks[0].mx || blocks[i].my != blocks[0].my ||
        blocks[i].mz != blocks[0].mz || blocks[i].ghost != G.g)
      return ENZOE_ERR_ARG;
    if (m->dual && !blocks[i].internal_energy) return ENZOE_ERR_ARG;
  }
  // one pointer table per field: the cell-centred fields, then for MHD the
  // 3 face fields, then internal_energy with dual energy
  const int ncc = m->mhd ? 8 : 5;
  G.face0 = m->mhd ? ncc : (1 << 20);
  const int nf = ncc + (m->mhd ? 3 : 0) + (m->dual ? 1 : 0);
  std::vector<double *> tab((size_t)nf * nb);
  for (int i = 0; i < nb; i++) {
    const enzoe_block &b = blocks[i];
    double *f[12] = {b.density, b.velocity[0], b.velocity[1], b.velocity[2],
                     b.total_energy, b.bfield[0], b.bfield[1], b.bfield[2],
                     b.bfieldi[0], b.bfieldi[1], b.bfieldi[2], b.internal_energy};
    int q = 0;
    for (int slot = 0; slot < 12; slot++) {
      if (!m->mhd && slot >= 5 && slot < 11) continue;
      if (!m->dual && slot == 11) continue;
      tab[(size_t)q * nb + i] = f[slot];
      q++;
    }
  }
  const size_t bytes = tab.size() * sizeof(double *);
  // the pointer table is uploaded only when the block list changes
  if (m->last_rtab != tab) {
    if (int rc = ensure_bytes(&m->d_rtab, &m->rtab_bytes, bytes)) return rc;
    if (hipMemcpyAsync(m->d_rtab, tab.data(), bytes, hipMemcpyHostToDevice, s) !=
        hipSuccess)
      return ENZOE_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return ENZOE_ERR_HIP;
    m->last_rtab = tab;
  }
  if (m->dry) return ENZOE_OK;  // graph set-up: the table only
  KernelTimer kt(s, "k_refresh");
  // largest ghost-element count over the field kinds (a face field has one
  // more interior layer and one more layer of extent along its axis)
  long nghost = 0;
  for (int fc = -1; fc < 3; fc++) {
    long tot = 1, in = 1;
    for (int a = 0; a < 3; a++) {
      tot *= (G.n[a] + 2 * G.g + (fc == a));
      in *= (G.n[a] + (fc == a));
    }
    nghost = std::max(nghost, tot - in);
  }
  hipLaunchKernelGGL(k_refresh, dim3(nblk(nghost, 256), nb, nf), dim3(256), 0, s,
                     (double *const *)m->d_rtab, G);
  if (hipGetLastError() != hipSuccess) return ENZOE_ERR_HIP;
  return ENZOE_OK;
}

int enzoe_grid_refresh(enzoe_vlct_t m, const enzoe_block *blocks, int nbx,
                       int nby, int nbz, const int bc[3],
                       enzoe_stream_t stream) {
  return grid_refresh(m, blocks, nbx, nby, nbz, bc, (hipStream_t)stream, nullptr);
}

int enzoe_vlct_set_fused_ct(enzoe_vlct_t m, int variant) {
  if (!m || variant < -1 || variant > 5) return ENZOE_ERR_ARG;
  m->ctu = variant;
  return ENZOE_OK;
}

int enzoe_vlct_set_flux_shape(enzoe_vlct_t m, int shape) {
  if (!m || shape < -1 || shape > 2) return ENZOE_ERR_ARG;
  m->flux_shape = shape;
  return ENZOE_OK;
}

int enzoe_vlct_check_persistent(enzoe_vlct_t m) {
  if (!m) return ENZOE_ERR_ARG;
  if (!m->d_bar) return ENZOE_OK;
  GridBarrier gb;
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(&gb, m->d_bar, sizeof(gb), hipMemcpyDeviceToHost) != hipSuccess)
    return ENZOE_ERR_HIP;
  if (gb.abort == 0u && gb.count == 0u) return ENZOE_OK;
  // a barrier wait ran out: the launch's results are not valid; the barrier
  // starts afresh for the next launch
  (void)hipMemset(m->d_bar, 0, sizeof(GridBarrier));
  (void)hipDeviceSynchronize();
  return ENZOE_ERR_HIP;
}

int enzoe_vlct_set_debug_checks(enzoe_vlct_t m, int on) {
  if (!m) return ENZOE_ERR_ARG;
  if (!on) {
    // work already queued may still count: the counters live until destroy
    if (m->d_dbg) {
      (void)hipDeviceSynchronize();
      (void)hipFree(m->d_dbg);
    }
    m->d_dbg = nullptr;
    return ENZOE_OK;
  }
  if (m->d_dbg) return ENZOE_OK;
  unsigned long long *p = nullptr;
  if (hipMalloc((void **)&p, DBG_N * sizeof(unsigned long long)) != hipSuccess)
    return ENZOE_ERR_ALLOC;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, nullptr, p, DBG_N);
  if (hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(p);
    return ENZOE_ERR_HIP;
  }
  m->d_dbg = p;
  return ENZOE_OK;
}

int enzoe_vlct_debug_counts(enzoe_vlct_t m, enzoe_stream_t stream,
                            unsigned long long counts[ENZOE_DEBUG_COUNTS]) {
  if (!m || !counts) return ENZOE_ERR_ARG;
  if (!m->d_dbg) return ENZOE_ERR_STATE;
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(counts, m->d_dbg, DBG_N * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ENZOE_ERR_HIP;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, m->d_dbg, DBG_N);
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int enzoe_vlct_diagnostics(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                           enzoe_stream_t stream, enzoe_diagnostics *out) {
  if (!m || !blocks || n <= 0 || !out) return ENZOE_ERR_ARG;
  for (int i = 0; i < n; i++) {
    const enzoe_block &b = blocks[i];
    if (b.mx < 1 || b.my < 1 || b.mz < 1 || b.ghost < 0 || 2 * b.ghost >= b.mx ||
        2 * b.ghost >= b.my || 2 * b.ghost >= b.mz || !b.density || !b.velocity[0] ||
        !b.velocity[1] || !b.velocity[2] || !b.total_energy)
      return ENZOE_ERR_ARG;
    if (m->mhd)
      for (int d = 0; d < 3; d++)
        if (!b.bfield[d] || !b.bfieldi[d]) return ENZOE_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<BlockDesc<double>> desc(n);
  for (int i = 0; i < n; i++) {
    BlockDesc<double> &d = desc[i];
    std::memset(&d, 0, sizeof(d));
    const enzoe_block &b = blocks[i];
    d.U[S_RHO] = b.density;
    for (int a = 0; a < 3; a++) d.U[S_VX + a] = b.velocity[a];
    d.U[S_E] = b.total_energy;
    if (m->mhd)
      for (int a = 0; a < 3; a++) {
        d.U[S_BX + a] = b.bfield[a];
        d.bfi[a] = b.bfieldi[a];
      }
    d.mx = b.mx; d.my = b.my; d.mz = b.mz;
    d.g = b.ghost;
    for (int a = 0; a < 3; a++) d.cw[a] = b.cell_width[a];
  }
  // a diagnostic: its own buffers, taken and returned on every call
  void *d_desc = nullptr, *d_acc = nullptr, *d_part = nullptr;
  const size_t dbytes = desc.size() * sizeof(BlockDesc<double>);
  const size_t nparts = (size_t)n * DIAG_WG * DIAG_NSUM * 2;
  if (hipMalloc(&d_desc, dbytes) != hipSuccess) return ENZOE_ERR_ALLOC;
  if (hipMalloc(&d_acc, sizeof(DiagAcc)) != hipSuccess) {
    (void)hipFree(d_desc);
    return ENZOE_ERR_ALLOC;
  }
  if (hipMalloc(&d_part, nparts * sizeof(double)) != hipSuccess) {
    (void)hipFree(d_desc);
    (void)hipFree(d_acc);
    return ENZOE_ERR_ALLOC;
  }
  int rc = ENZOE_OK;
  DiagAcc acc;
  std::vector<double> parts(nparts);
  if (hipMemcpyAsync(d_desc, desc.data(), dbytes, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = ENZOE_ERR_HIP;
  if (rc == ENZOE_OK) {
    hipLaunchKernelGGL(k_diag_init, dim3(1), dim3(1), 0, s, (DiagAcc *)d_acc);
    dim3 grid(DIAG_WG, n);
    if (m->mhd)
      hipLaunchKernelGGL((k_diagnostics<true>), grid, dim3(256), 0, s,
                         (const BlockDesc<double> *)d_desc, m->gamma, (DiagAcc *)d_acc,
                         (double *)d_part);
    else
      hipLaunchKernelGGL((k_diagnostics<false>), grid, dim3(256), 0, s,
                         (const BlockDesc<double> *)d_desc, m->gamma, (DiagAcc *)d_acc,
                         (double *)d_part);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&acc, d_acc, sizeof(acc), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(parts.data(), d_part, nparts * sizeof(double), hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = ENZOE_ERR_HIP;
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_desc);
  (void)hipFree(d_acc);
  (void)hipFree(d_part);
  if (rc) return rc;
  // the partial pairs in a fixed order (block, workgroup)
  double tot[DIAG_NSUM], comp[DIAG_NSUM];
  for (int q = 0; q < DIAG_NSUM; q++) tot[q] = comp[q] = 0.;
  for (size_t w = 0; w < (size_t)n * DIAG_WG; w++)
    for (int q = 0; q < DIAG_NSUM; q++)
      comp_merge(tot[q], comp[q], parts[(w * DIAG_NSUM + q) * 2],
                 parts[(w * DIAG_NSUM + q) * 2 + 1]);
  out->mass = tot[0] + comp[0];
  for (int a = 0; a < 3; a++) out->momentum[a] = tot[1 + a] + comp[1 + a];
  out->energy = tot[4] + comp[4];
  out->min_density = ordered_to_double(acc.min_rho);
  out->min_pressure = ordered_to_double(acc.min_p);
  out->max_abs_divb = __builtin_bit_cast(double, acc.max_divb);
  out->max_abs_bface = __builtin_bit_cast(double, acc.max_bface);
  out->nonfinite = (long long)acc.nonfinite;
  return ENZOE_OK;
}

int enzoe_vlct_state_l1(enzoe_vlct_t m, const enzoe_block *a, const enzoe_block *b,
                        int n, enzoe_stream_t stream, enzoe_state_l1 *out) {
  if (!m || !a || !b || n <= 0 || !out) return ENZOE_ERR_ARG;
  std::vector<BlockDesc<double>> desc(2 * (size_t)n);
  for (int side = 0; side < 2; side++)
    for (int i = 0; i < n; i++) {
      const enzoe_block &x = side ? b[i] : a[i];
      const enzoe_block &y = b[i];
      if (x.mx != y.mx || x.my != y.my || x.mz != y.mz || x.ghost != y.ghost ||
          x.mx < 1 || x.my < 1 || x.mz < 1 || x.ghost < 0 || 2 * x.ghost >= x.mx ||
          2 * x.ghost >= x.my || 2 * x.ghost >= x.mz || !x.density || !x.velocity[0] ||
          !x.velocity[1] || !x.velocity[2] || !x.total_energy)
        return ENZOE_ERR_ARG;
      if (m->mhd && (!x.bfield[0] || !x.bfield[1] || !x.bfield[2])) return ENZOE_ERR_ARG;
      BlockDesc<double> &d = desc[(size_t)side * n + i];
      std::memset(&d, 0, sizeof(d));
      d.U[S_RHO] = x.density;
      for (int c = 0; c < 3; c++) d.U[S_VX + c] = x.velocity[c];
      d.U[S_E] = x.total_energy;
      if (m->mhd)
        for (int c = 0; c < 3; c++) d.U[S_BX + c] = x.bfield[c];
      d.mx = x.mx; d.my = x.my; d.mz = x.mz;
      d.g = x.ghost;
    }
  hipStream_t s = (hipStream_t)stream;
  void *d_desc = nullptr, *d_part = nullptr;
  const size_t dbytes = desc.size() * sizeof(BlockDesc<double>);
  const size_t nparts = (size_t)n * DIAG_WG * L1_NPART;
  if (hipMalloc(&d_desc, dbytes) != hipSuccess) return ENZOE_ERR_ALLOC;
  if (hipMalloc(&d_part, nparts * sizeof(double)) != hipSuccess) {
    (void)hipFree(d_desc);
    return ENZOE_ERR_ALLOC;
  }
  std::vector<double> parts(nparts);
  int rc = ENZOE_OK;
  if (hipMemcpyAsync(d_desc, desc.data(), dbytes, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = ENZOE_ERR_HIP;
  if (rc == ENZOE_OK) {
    const BlockDesc<double> *da = (const BlockDesc<double> *)d_desc;
    dim3 grid(DIAG_WG, n);
    if (m->mhd)
      hipLaunchKernelGGL((k_state_l1<true>), grid, dim3(256), 0, s, da, da + n,
                         (double *)d_part);
    else
      hipLaunchKernelGGL((k_state_l1<false>), grid, dim3(256), 0, s, da, da + n,
                         (double *)d_part);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(parts.data(), d_part, nparts * sizeof(double), hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = ENZOE_ERR_HIP;
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_desc);
  (void)hipFree(d_part);
  if (rc) return rc;
  double tot[L1_NSUM], comp[L1_NSUM], mx[8];
  for (int q = 0; q < L1_NSUM; q++) tot[q] = comp[q] = 0.;
  for (int q = 0; q < 8; q++) mx[q] = 0.;
  for (size_t w = 0; w < (size_t)n * DIAG_WG; w++) {
    const double *pw = parts.data() + w * L1_NPART;
    for (int q = 0; q < L1_NSUM; q++) comp_merge(tot[q], comp[q], pw[q * 2], pw[q * 2 + 1]);
    for (int q = 0; q < 8; q++) {
      const double v = pw[L1_NSUM * 2 + q];
      if (v > mx[q] || v != v) mx[q] = v;
    }
  }
  for (int q = 0; q < 8; q++) {
    out->diff[q] = tot[q] + comp[q];
    out->ref[q] = tot[8 + q] + comp[8 + q];
    out->max_rel[q] = mx[q];
  }
  out->velocity_magnitude = tot[16] + comp[16];
  out->bfield_magnitude = tot[17] + comp[17];
  return ENZOE_OK;
}

int enzoe_set_profiling(int on) {
  resolve_pending();
  g_profiling = on != 0;
  if (on) g_prof.clear();
  return ENZOE_OK;
}

int enzoe_calibration_copy(double *dst, const double *src, size_t n,
                           enzoe_stream_t stream) {
  if (!dst || !src) return ENZOE_ERR_ARG;
  KernelTimer kt((hipStream_t)stream, "k_calibration_copy");
  hipLaunchKernelGGL(k_calibration_copy, dim3(nblk((long)n, 256)), dim3(256), 0,
                     (hipStream_t)stream, dst, src, (long)n);
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int enzoe_profile_report(char *names, int name_len, double *ms, int *count,
                         int cap) {
  resolve_pending();
  int i = 0;
  for (auto &kv : g_prof) {
    if (i >= cap) break;
    std::snprintf(names + (size_t)i * name_len, name_len, "%s", kv.first.c_str());
    ms[i] = kv.second.total_ms / std::max<long>(1, kv.second.count);
    count[i] = (int)kv.second.count;
    i++;
  }
  return i;
}

}  // extern "C"

// The device buffers a method's launches read, moved in and out of the
// method so that a graph owns a private set (a later eager call may
// re-upload or regrow the method's own ones while the graph is alive).
// ------------------------------------------------------ sub-cycled levels

// dt of a coarse step of sub-cycled levels on the device: dt_l = courant *
// CFL minimum of level l (as enzoe_vlct_timestep takes it), dt_0 = min_l
// 2^l dt_l, and every entry of every substep's block list gets dt_0 /
// 2^(its level); the time state[0] advances by dt_0, state[1] holds it.
__global__ void k_levels_dt(const unsigned long long *__restrict__ cfl, int nlev,
                            double courant, const int *__restrict__ lev,
                            double *__restrict__ dts, long ndts, double *state) {
  double dt0 = 0.;
  for (int l = 0; l < nlev; l++) {
    const double d =
        (double)(1LL << l) * (__longlong_as_double((long long)cfl[l]) * courant);
    dt0 = (l == 0 || d < dt0) ? d : dt0;
  }
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ndts) dts[t] = dt0 / (double)(1LL << lev[t]);
  if (t == 0) {
    state[1] = dt0;
    state[0] = state[0] + dt0;
  }
}

// One launch run of a substep: consecutive blocks of one shape.
struct LevelRun {
  int first = 0, n = 0, chunk = 1;
  size_t desc_off = 0;  // byte offset of its descriptor table
};
// One substep: the blocks of the levels that step at it, coarsest first.
struct LevelSub {
  std::vector<enzoe_block> blocks;
  std::vector<int> level;
  long dts_off = 0;  // first entry of its per-block dt
  std::vector<LevelRun> runs;
};

struct MethodBuffers {
  void *scratch = nullptr, *d_desc = nullptr, *d_dt = nullptr,
       *d_ptrtab = nullptr, *d_rtab = nullptr;
  size_t scratch_bytes = 0, desc_bytes = 0, ptrtab_bytes = 0, rtab_bytes = 0;
  std::vector<char> last_desc, last_dt_desc;
  std::vector<double *> last_rtab;

  void swap_with(enzoe_vlct_method *m) {
    std::swap(scratch, m->scratch);
    std::swap(scratch_bytes, m->scratch_bytes);
    std::swap(d_desc, m->d_desc);
    std::swap(desc_bytes, m->desc_bytes);
    std::swap(last_desc, m->last_desc);
    void *dt = (void *)m->d_dt;
    std::swap(d_dt, dt);
    m->d_dt = (unsigned long long *)dt;
    std::swap(d_ptrtab, m->d_ptrtab);
    std::swap(ptrtab_bytes, m->ptrtab_bytes);
    std::swap(last_dt_desc, m->last_dt_desc);
    std::swap(d_rtab, m->d_rtab);
    std::swap(rtab_bytes, m->rtab_bytes);
    std::swap(last_rtab, m->last_rtab);
  }
  void release() {
    for (void *p : {scratch, d_desc, d_dt, d_ptrtab, d_rtab})
      if (p) (void)hipFree(p);
    scratch = d_desc = d_dt = d_ptrtab = d_rtab = nullptr;
  }
};

struct enzoe_vlct_levels_s {
  enzoe_vlct_t m = nullptr;
  int nlev = 0;
  std::vector<std::vector<enzoe_block>> lev;  // blocks of each level
  std::vector<LevelSub> sub;                   // 2^(nlev-1) substeps
  void *scratch = nullptr;                     // stage scratch (shared by the runs)
  void *d_desc = nullptr;                      // descriptor tables of the runs
  void *d_dtdesc = nullptr;                    // CFL-kernel tables of the levels
  std::vector<long> dtdesc_first;              // first descriptor of each level
  std::vector<long> level_maxcells;
  double *d_dts = nullptr;                     // per-block dt of every substep
  int *d_lev = nullptr;                        // the level of each of those entries
  long ndts = 0;
  unsigned long long *d_cfl = nullptr;         // CFL minimum of each level
  double *d_state = nullptr;                   // [0] time, [1] dt_0 of the last step
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void *p : {scratch, d_desc, d_dtdesc, (void *)d_dts, (void *)d_lev, (void *)d_cfl,
                    (void *)d_state})
      if (p) (void)hipFree(p);
    exec = nullptr;
    graph = nullptr;
    scratch = d_desc = d_dtdesc = nullptr;
    d_dts = nullptr;
    d_lev = nullptr;
    d_cfl = nullptr;
    d_state = nullptr;
  }
};

namespace {

// One coarse step of the levels: the CFL reduction of every level, the dt
// kernel, then every substep's runs; kernels only (capturable).
int levels_launch(enzoe_vlct_levels_s *L, hipStream_t s) {
  enzoe_vlct_method *m = L->m;
  for (int l = 0; l < L->nlev; l++)
    if (int rc = launch_dt_on(m, (const BlockDesc<double> *)L->d_dtdesc + L->dtdesc_first[l],
                              (int)L->lev[l].size(), L->level_maxcells[l], L->d_cfl + l, s,
                              nullptr, 0.0))
      return rc;
  hipLaunchKernelGGL(k_levels_dt, dim3(nblk(std::max<long>(1, L->ndts), 256)), dim3(256), 0, s,
                     L->d_cfl, L->nlev, m->courant, L->d_lev, L->d_dts, L->ndts, L->d_state);
  for (const LevelSub &sb : L->sub)
    for (const LevelRun &r : sb.runs) {
      const enzoe_block *bl = sb.blocks.data() + r.first;
      const void *table = (const char *)L->d_desc + r.desc_off;
      const double *dts = L->d_dts + sb.dts_off + r.first;
      int rc;
      if (m->precision == ENZOE_PREC_DOUBLE) {
        if (m->mhd)
          rc = m->dual ? launch_chunks<double, true, true>(m, bl, r.n, r.chunk,
                                                           (const BlockDesc<double> *)table, 0.0,
                                                           dts, s, nullptr, nullptr, true)
                       : launch_chunks<double, true, false>(m, bl, r.n, r.chunk,
                                                            (const BlockDesc<double> *)table,
                                                            0.0, dts, s, nullptr, nullptr, true);
        else
          rc = m->dual ? launch_chunks<double, false, true>(m, bl, r.n, r.chunk,
                                                            (const BlockDesc<double> *)table,
                                                            0.0, dts, s, nullptr, nullptr, true)
                       : launch_chunks<double, false, false>(m, bl, r.n, r.chunk,
                                                             (const BlockDesc<double> *)table,
                                                             0.0, dts, s, nullptr, nullptr, true);
      } else if (m->precision == ENZOE_PREC_DOUBLE_FMA) {
        rc = launch_prepared_contracted(m, bl, r.n, r.chunk, table, dts, s);
      } else {
        rc = launch_prepared_mixed(m, bl, r.n, r.chunk, table, dts, s);
      }
      if (rc) return rc;
    }
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

}  // namespace

struct enzoe_vlct_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double *d_state = nullptr;  // [0] time, [1 .. n] dt of every block
  int n = 0;
  MethodBuffers own;
};

extern "C" {

int enzoe_vlct_graph_create(enzoe_vlct_t m, const enzoe_block *blocks, int nbx,
                            int nby, int nbz, const int bc[3], double t0,
                            double t_stop, int ncycles,
                            enzoe_vlct_graph_t *out) {
  if (!m || !blocks || !bc || !out || nbx <= 0 || nby <= 0 || nbz <= 0 ||
      ncycles <= 0)
    return ENZOE_ERR_ARG;
  *out = nullptr;
  const int n = nbx * nby * nbz;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
    return ENZOE_ERR_HIP;
  enzoe_vlct_graph_s *g = new enzoe_vlct_graph_s;
  g->n = n;
  int rc = ENZOE_OK;
  bool swapped = false;
  auto fail = [&](int code) {
    (void)hipStreamSynchronize(s);
    if (swapped) g->own.swap_with(m);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_state) (void)hipFree(g->d_state);
    g->own.release();
    delete g;
    (void)hipStreamDestroy(s);
    return code;
  };
  if (hipMalloc((void **)&g->d_state, (n + 1) * sizeof(double)) != hipSuccess) {
    g->d_state = nullptr;
    return fail(ENZOE_ERR_ALLOC);
  }
  if (hipMemcpy(g->d_state, &t0, sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(ENZOE_ERR_HIP);
  // the graph's private buffers: take every buffer and upload every table
  // now, so the captured work allocates and uploads nothing
  if (hipDeviceSynchronize() != hipSuccess) return fail(ENZOE_ERR_HIP);
  g->own.swap_with(m);  // the method now holds empty buffers
  swapped = true;
  // (dry: nothing is launched, so the first replayed cycle sees the fields
  // exactly as an eager cycle would, stale ghost zones included)
  m->dry = true;
  rc = grid_refresh(m, blocks, nbx, nby, nbz, bc, s, nullptr);
  if (!rc) rc = launch_dt(m, blocks, n, s);
  if (!rc) rc = compute_runs(m, blocks, n, 0.0, g->d_state + 1, s);
  m->dry = false;
  if (rc) return fail(rc);
  if (hipStreamSynchronize(s) != hipSuccess) return fail(ENZOE_ERR_HIP);
  const bool prof = g_profiling;
  g_profiling = false;  // no timing events inside the graph
  if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
    g_profiling = prof;
    return fail(ENZOE_ERR_HIP);
  }
  for (int c = 0; c < ncycles && rc == ENZOE_OK; c++) {
    rc = enzoe_vlct_timestep_device(m, blocks, n, g->d_state, t_stop, g->d_state + 1,
                                    n, s);
    // a cycle past t_stop (dt 0) leaves the ghost zones alone too
    if (!rc) rc = grid_refresh(m, blocks, nbx, nby, nbz, bc, s, g->d_state + 1);
    if (!rc) rc = compute_runs(m, blocks, n, 0.0, g->d_state + 1, s);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &graph);
  g_profiling = prof;
  g->graph = graph;
  if (rc) return fail(rc);
  if (ec != hipSuccess || !graph) return fail(ENZOE_ERR_HIP);
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
    g->exec = nullptr;
    return fail(ENZOE_ERR_HIP);
  }
  g->own.swap_with(m);  // the graph keeps its buffers, the method its own
  swapped = false;
  (void)hipStreamDestroy(s);
  *out = g;
  return ENZOE_OK;
}

int enzoe_vlct_run_cycles(enzoe_vlct_t m, const enzoe_block *blocks, int nbx, int nby,
                          int nbz, const int bc[3], double *d_state, double t_stop,
                          int ncycles, enzoe_stream_t stream) {
  if (!m || !blocks || !bc || !d_state || nbx <= 0 || nby <= 0 || nbz <= 0 || ncycles <= 0)
    return ENZOE_ERR_ARG;
  const int n = nbx * nby * nbz;
  hipStream_t s = (hipStream_t)stream;
  // the first cycle's CFL minimum from the state as it is; then per cycle
  // its dt (clipped at t_stop, the time advanced, the minimum word reset),
  // the refresh, and the compute whose final cell update takes the next
  // cycle's minimum (the dt a CFL pass over its result would give)
  if (int rc = launch_dt(m, blocks, n, s, d_state, t_stop)) return rc;
  for (int c = 0; c < ncycles; c++) {
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, s, m->d_dt, m->courant, d_state,
                       t_stop, d_state + 1, n, 1);
    if (int rc = grid_refresh(m, blocks, nbx, nby, nbz, bc, s, d_state + 1)) return rc;
    if (int rc = compute_runs(m, blocks, n, 0.0, d_state + 1, s, nullptr, nullptr, m->d_dt))
      return rc;
  }
  return hipGetLastError() == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int enzoe_vlct_graph_launch(enzoe_vlct_graph_t g, enzoe_stream_t stream) {
  if (!g || !g->exec) return ENZOE_ERR_ARG;
  return hipGraphLaunch(g->exec, (hipStream_t)stream) == hipSuccess ? ENZOE_OK
                                                                     : ENZOE_ERR_HIP;
}

int enzoe_vlct_graph_time(enzoe_vlct_graph_t g, enzoe_stream_t stream, double *t) {
  if (!g || !t) return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(t, g->d_state, sizeof(double), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ENZOE_ERR_HIP;
  return ENZOE_OK;
}

int enzoe_vlct_levels_create(enzoe_vlct_t m, const enzoe_block *blocks, const int *level_of,
                             int n, enzoe_vlct_levels_t *out) {
  if (!m || !blocks || !level_of || n <= 0 || !out) return ENZOE_ERR_ARG;
  *out = nullptr;
  int nlev = 0;
  for (int i = 0; i < n; i++) {
    if (level_of[i] < 0 || level_of[i] > 20) return ENZOE_ERR_ARG;
    if (int rc = check_block(m, blocks[i])) return rc;
    nlev = std::max(nlev, level_of[i] + 1);
  }
  // the persistent stage launch's barrier (the autotuner may choose it)
  if (persistent_ok(m))
    if (int rc = ensure_barrier(m)) return rc;
  std::unique_ptr<enzoe_vlct_levels_s> L(new enzoe_vlct_levels_s);
  L->m = m;
  L->nlev = nlev;
  L->lev.resize(nlev);
  for (int i = 0; i < n; i++) L->lev[level_of[i]].push_back(blocks[i]);
  for (int l = 0; l < nlev; l++)
    if (L->lev[l].empty()) return ENZOE_ERR_ARG;  // every level holds blocks
  // substep s advances the levels l with s % 2^(L-1-l) == 0
  const int nsub = 1 << (nlev - 1);
  L->sub.resize(nsub);
  std::vector<int> lev_of_entry;
  for (int st = 0; st < nsub; st++) {
    LevelSub &sb = L->sub[st];
    for (int l = 0; l < nlev; l++)
      if (st % (1 << (nlev - 1 - l)) == 0)
        for (const enzoe_block &b : L->lev[l]) {
          sb.blocks.push_back(b);
          sb.level.push_back(l);
        }
    sb.dts_off = L->ndts;
    L->ndts += (long)sb.blocks.size();
    lev_of_entry.insert(lev_of_entry.end(), sb.level.begin(), sb.level.end());
    for (int b0 = 0; b0 < (int)sb.blocks.size();) {
      int b1 = b0 + 1;
      while (b1 < (int)sb.blocks.size() && sb.blocks[b1].mx == sb.blocks[b0].mx &&
             sb.blocks[b1].my == sb.blocks[b0].my && sb.blocks[b1].mz == sb.blocks[b0].mz)
        b1++;
      LevelRun r;
      r.first = b0;
      r.n = b1 - b0;
      sb.runs.push_back(r);
      b0 = b1;
    }
  }
  // one scratch pool for every run (they run one after the other), each
  // run in chunks that fit the method's budget
  const bool mixed = m->precision == ENZOE_PREC_MIXED;
  const size_t budget = scratch_budget(m);
  size_t pool = 0;
  for (LevelSub &sb : L->sub)
    for (LevelRun &r : sb.runs) {
      const enzoe_block &b = sb.blocks[r.first];
      const size_t per = mixed ? scratch_per_block<float>(b.mx, b.my, b.mz, m->mhd, m->dual,
                                                          ctu_possible(m))
                               : scratch_per_block<double>(b.mx, b.my, b.mz, m->mhd, m->dual,
                                                           ctu_possible(m));
      r.chunk = (int)std::max<size_t>(1, std::min<size_t>(r.n, budget / per));
      pool = std::max(pool, per * r.chunk);
    }
  auto fail = [&](int code) {
    L->release();
    return code;
  };
  if (hipMalloc(&L->scratch, pool) != hipSuccess) {
    L->scratch = nullptr;
    return fail(ENZOE_ERR_ALLOC);
  }
  std::vector<char> tables;
  for (LevelSub &sb : L->sub)
    for (LevelRun &r : sb.runs) {
      r.desc_off = tables.size();
      const enzoe_block *bl = sb.blocks.data() + r.first;
      if (mixed) {
        std::vector<BlockDesc<float>> d(r.n);
        for (int c0 = 0; c0 < r.n; c0 += r.chunk)
          fill_desc_at<float>(m, (char *)L->scratch, bl, c0, std::min(r.n, c0 + r.chunk),
                              d.data() + c0);
        tables.insert(tables.end(), (const char *)d.data(),
                      (const char *)(d.data() + r.n));
      } else {
        std::vector<BlockDesc<double>> d(r.n);
        for (int c0 = 0; c0 < r.n; c0 += r.chunk)
          fill_desc_at<double>(m, (char *)L->scratch, bl, c0, std::min(r.n, c0 + r.chunk),
                               d.data() + c0);
        tables.insert(tables.end(), (const char *)d.data(),
                      (const char *)(d.data() + r.n));
      }
      // 256-byte aligned tables
      tables.resize((tables.size() + 255) & ~(size_t)255);
    }
  std::vector<BlockDesc<double>> dtdesc;
  for (int l = 0; l < nlev; l++) {
    L->dtdesc_first.push_back((long)dtdesc.size());
    const size_t at = dtdesc.size();
    dtdesc.resize(at + L->lev[l].size());
    fill_dt_desc(m, L->lev[l].data(), (int)L->lev[l].size(), dtdesc.data() + at);
    long mc = 0;
    for (const enzoe_block &b : L->lev[l]) mc = std::max<long>(mc, (long)b.mx * b.my * b.mz);
    L->level_maxcells.push_back(mc);
  }
  const double state0[2] = {0.0, 0.0};
  if (hipMalloc(&L->d_desc, tables.size()) != hipSuccess ||
      hipMalloc(&L->d_dtdesc, dtdesc.size() * sizeof(BlockDesc<double>)) != hipSuccess ||
      hipMalloc((void **)&L->d_dts, L->ndts * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&L->d_lev, L->ndts * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&L->d_cfl, nlev * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc((void **)&L->d_state, 2 * sizeof(double)) != hipSuccess)
    return fail(ENZOE_ERR_ALLOC);
  if (hipMemcpy(L->d_desc, tables.data(), tables.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(L->d_dtdesc, dtdesc.data(), dtdesc.size() * sizeof(BlockDesc<double>),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(L->d_lev, lev_of_entry.data(), L->ndts * sizeof(int),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(L->d_state, state0, sizeof(state0), hipMemcpyHostToDevice) != hipSuccess)
    return fail(ENZOE_ERR_HIP);
  *out = L.release();
  return ENZOE_OK;
}

int enzoe_vlct_levels_step(enzoe_vlct_levels_t L, enzoe_stream_t stream) {
  if (!L) return ENZOE_ERR_ARG;
  return levels_launch(L, (hipStream_t)stream);
}

int enzoe_vlct_levels_capture(enzoe_vlct_levels_t L, int nsteps) {
  if (!L || nsteps <= 0) return ENZOE_ERR_ARG;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return ENZOE_ERR_HIP;
  if (L->exec) (void)hipGraphExecDestroy(L->exec);
  if (L->graph) (void)hipGraphDestroy(L->graph);
  L->exec = nullptr;
  L->graph = nullptr;
  const bool prof = g_profiling;
  g_profiling = false;  // no timing events inside the graph
  int rc = ENZOE_OK;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
    rc = ENZOE_ERR_HIP;
  } else {
    for (int i = 0; i < nsteps && rc == ENZOE_OK; i++) rc = levels_launch(L, s);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    L->graph = graph;
    if (!rc && (ec != hipSuccess || !graph)) rc = ENZOE_ERR_HIP;
    if (!rc && hipGraphInstantiate(&L->exec, L->graph, nullptr, nullptr, 0) != hipSuccess) {
      L->exec = nullptr;
      rc = ENZOE_ERR_HIP;
    }
  }
  g_profiling = prof;
  (void)hipStreamDestroy(s);
  return rc;
}

int enzoe_vlct_levels_replay(enzoe_vlct_levels_t L, enzoe_stream_t stream) {
  if (!L || !L->exec) return ENZOE_ERR_STATE;
  return hipGraphLaunch(L->exec, (hipStream_t)stream) == hipSuccess ? ENZOE_OK : ENZOE_ERR_HIP;
}

int enzoe_vlct_levels_state(enzoe_vlct_levels_t L, enzoe_stream_t stream, double *time,
                            double *dt0) {
  if (!L) return ENZOE_ERR_ARG;
  double st[2];
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(st, L->d_state, sizeof(st), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ENZOE_ERR_HIP;
  if (time) *time = st[0];
  if (dt0) *dt0 = st[1];
  return ENZOE_OK;
}

int enzoe_vlct_levels_destroy(enzoe_vlct_levels_t L) {
  if (!L) return ENZOE_ERR_ARG;
  (void)hipDeviceSynchronize();
  L->release();
  delete L;
  return ENZOE_OK;
}

int enzoe_vlct_graph_destroy(enzoe_vlct_graph_t g) {
  if (!g) return ENZOE_ERR_ARG;
  (void)hipDeviceSynchronize();
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->d_state) (void)hipFree(g->d_state);
  g->own.release();
  delete g;
  return ENZOE_OK;
}

}  // extern "C"
