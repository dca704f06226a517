// vlct_capi.hip -- C ABI of the gfx950 VL+CT library (declared in
// include/enzoe_vlct.h).  Owns the method configuration, the device scratch
// pool and the launch sequence of one VL+CT step over a list of blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/enzoe_vlct.h"
#include "vlct_kernels.hpp"

using namespace enzoe;

namespace {

const char *kVersion = "enzoe-vlct 0.1 (gfx950)";

std::string lower(const char *s, const char *dflt) {
  std::string r = s ? s : dflt;
  std::transform(r.begin(), r.end(), r.begin(),
                 [](unsigned char c) { return (char)std::tolower(c); });
  return r;
}

// ---------------------------------------------------------------- profiling

struct ProfEntry {
  double total_ms = 0;
  long count = 0;
};
bool g_profiling = false;
std::map<std::string, ProfEntry> g_prof;
// launches recorded but not yet resolved: events are read only when a
// report is asked for, so profiling adds no host synchronisation
struct PendingTiming {
  const char *name;
  hipEvent_t a, b;
};
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

struct KernelTimer {
  hipStream_t s;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  KernelTimer(hipStream_t s_, const char *n) : s(s_), name(n) {
    if (g_profiling) {
      a = pool_event();
      b = pool_event();
      (void)hipEventRecord(a, s);
    }
  }
  ~KernelTimer() {
    if (g_profiling) {
      (void)hipEventRecord(b, s);
      g_pending.push_back({name, a, b});
      if (g_pending.size() > 4096) resolve_pending();
    }
  }
};

__global__ void k_calibration_copy(double *__restrict__ dst,
                                   const double *__restrict__ src, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ------------------------------------------------------------------- method

struct enzoe_vlct_method {
  int mhd = 1;
  int riemann = RS_HLLD;
  int recon[2] = {RC_NN, RC_PLM};
  int nstages = 2;
  int precision = ENZOE_PREC_DOUBLE;
  double theta = 1.5, courant = 0.3, gamma = 5.0 / 3.0;
  double dfloor = 1e-200, pfloor = 1e-200;

  // scratch pool (grown on demand)
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // descriptor table
  void *d_desc = nullptr;
  size_t desc_bytes = 0;
  std::vector<char> last_desc;
  unsigned long long *d_dt = nullptr;
  void *d_ptrtab = nullptr;
  size_t ptrtab_bytes = 0;
};

namespace {

int total_staling(int r) { return r == RC_NN ? 1 : 2; }
int immediate_staling(int r) { return r == RC_NN ? 0 : 1; }

int parse_recon(const std::string &n, int *out) {
  if (n == "nn") { *out = RC_NN; return 0; }
  if (n == "plm" || n == "plm_enzo") { *out = RC_PLM; return 0; }
  if (n == "plm_athena") { *out = RC_PLM_ATHENA; return 0; }
  return ENZOE_ERR_CONFIG;
}

// bytes of scratch per block
template <typename FluxT>
size_t scratch_per_block(int mx, int my, int mz, int mhd) {
  size_t ncc = (size_t)mx * my * mz;
  size_t nq = mhd ? 8 : 5;
  size_t bytes = 0;
  bytes += nq * ncc * sizeof(double);                  // Ut
  if (mhd) {
    bytes += ((size_t)mz * my * (mx + 1) + (size_t)mz * (my + 1) * mx +
              (size_t)(mz + 1) * my * mx) * sizeof(double);  // bft
    bytes += 3 * ncc * sizeof(double);                 // E
  }
  bytes += 3 * (mhd ? 7 : 5) * ncc * sizeof(FluxT);    // F
  return bytes;
}

int ensure_bytes(void **ptr, size_t *have, size_t need) {
  if (*have >= need) return 0;
  // work queued earlier may still read the old buffer
  if (*ptr) {
    (void)hipDeviceSynchronize();
    (void)hipFree(*ptr);
  }
  *ptr = nullptr;
  *have = 0;
  if (hipMalloc(ptr, need) != hipSuccess) return ENZOE_ERR_ALLOC;
  *have = need;
  return 0;
}

int check_block(const enzoe_vlct_method *m, const enzoe_block &b) {
  if (b.mx < 7 || b.my < 7 || b.mz < 7) return ENZOE_ERR_ARG;
  // EnzoMethodMHDVlct.cpp:124-133: ghost depth >= sum of staling rates
  int req = 0;
  for (int s = 0; s < m->nstages; s++) req += total_staling(m->recon[s]);
  if (b.ghost < req) return ENZOE_ERR_ARG;
  if (!b.density || !b.velocity[0] || !b.velocity[1] || !b.velocity[2] ||
      !b.total_energy)
    return ENZOE_ERR_ARG;
  if (m->mhd) {
    for (int d = 0; d < 3; d++)
      if (!b.bfield[d] || !b.bfieldi[d]) return ENZOE_ERR_ARG;
  }
  return 0;
}

// Fills descriptors for blocks [b0, b1) using the scratch pool.
template <typename FluxT>
void fill_desc(const enzoe_vlct_method *m, const enzoe_block *blocks, int b0,
               int b1, std::vector<BlockDesc<FluxT>> &out) {
  out.clear();
  char *base = (char *)m->scratch;
  for (int bi = b0; bi < b1; bi++) {
    const enzoe_block &b = blocks[bi];
    BlockDesc<FluxT> d;
    std::memset(&d, 0, sizeof(d));
    d.U[S_RHO] = b.density;
    for (int a = 0; a < 3; a++) d.U[S_VX + a] = b.velocity[a];
    d.U[S_E] = b.total_energy;
    if (m->mhd)
      for (int a = 0; a < 3; a++) {
        d.U[S_BX + a] = b.bfield[a];
        d.bfi[a] = b.bfieldi[a];
      }
    d.pressure = b.pressure;
    d.mx = b.mx; d.my = b.my; d.mz = b.mz;
    for (int a = 0; a < 3; a++) d.cw[a] = b.cell_width[a];
    const size_t ncc = (size_t)b.mx * b.my * b.mz;
    const int nq = m->mhd ? 8 : 5;
    for (int q = 0; q < S_N; q++) {
      if (!m->mhd && q >= S_BX) continue;
      d.Ut[q] = (double *)base;
      base += ncc * sizeof(double);
    }
    (void)nq;
    if (m->mhd) {
      d.bft[0] = (double *)base; base += (size_t)b.mz * b.my * (b.mx + 1) * sizeof(double);
      d.bft[1] = (double *)base; base += (size_t)b.mz * (b.my + 1) * b.mx * sizeof(double);
      d.bft[2] = (double *)base; base += (size_t)(b.mz + 1) * b.my * b.mx * sizeof(double);
      for (int a = 0; a < 3; a++) {
        d.E[a] = (double *)base;
        base += ncc * sizeof(double);
      }
    }
    for (int dd = 0; dd < 3; dd++) {
      for (int q = 0; q < S_N; q++) {
        if (!m->mhd && q >= S_BX) continue;
        if (m->mhd && q == S_BX + dd) continue;  // normal-B flux is 0
        d.F[dd][q] = (FluxT *)base;
        base += ncc * sizeof(FluxT);
      }
    }
    out.push_back(d);
  }
}

template <typename FluxT, int RECON, int RIEMANN, bool MHD>
void launch_flux(const BlockDesc<FluxT> *d, int nb, long nthreads,
                 StageParams sp, hipStream_t s) {
  dim3 grid(nblk(nthreads, 256), nb);
  hipLaunchKernelGGL((k_flux<FluxT, FluxT, RECON, RIEMANN, MHD>), grid,
                     dim3(256), 0, s, d, sp);
}

template <typename FluxT, bool MHD>
void dispatch_flux(const enzoe_vlct_method *m, int recon,
                   const BlockDesc<FluxT> *d, int nb, long nthreads,
                   StageParams sp, hipStream_t s) {
  if (MHD) {
    if (m->riemann == RS_HLLD) {
      if (recon == RC_NN) launch_flux<FluxT, RC_NN, RS_HLLD, MHD>(d, nb, nthreads, sp, s);
      else if (recon == RC_PLM) launch_flux<FluxT, RC_PLM, RS_HLLD, MHD>(d, nb, nthreads, sp, s);
      else launch_flux<FluxT, RC_PLM_ATHENA, RS_HLLD, MHD>(d, nb, nthreads, sp, s);
    } else {
      if (recon == RC_NN) launch_flux<FluxT, RC_NN, RS_HLLE, MHD>(d, nb, nthreads, sp, s);
      else if (recon == RC_PLM) launch_flux<FluxT, RC_PLM, RS_HLLE, MHD>(d, nb, nthreads, sp, s);
      else launch_flux<FluxT, RC_PLM_ATHENA, RS_HLLE, MHD>(d, nb, nthreads, sp, s);
    }
  } else {
    if (recon == RC_NN) launch_flux<FluxT, RC_NN, RS_HLLC, MHD>(d, nb, nthreads, sp, s);
    else if (recon == RC_PLM) launch_flux<FluxT, RC_PLM, RS_HLLC, MHD>(d, nb, nthreads, sp, s);
    else launch_flux<FluxT, RC_PLM_ATHENA, RS_HLLC, MHD>(d, nb, nthreads, sp, s);
  }
}

long cube(int mx, int my, int mz, int lo, int hi_trim) {
  long a = mx - lo - hi_trim, b = my - lo - hi_trim, c = mz - lo - hi_trim;
  if (a <= 0 || b <= 0 || c <= 0) return 0;
  return a * b * c;
}

// One VL+CT step (EnzoMethodMHDVlct::compute) on blocks of equal shape.
template <typename FluxT, bool MHD>
int compute_group(enzoe_vlct_method *m, const enzoe_block *blocks, int n,
                  double dt, hipStream_t s) {
  const int mx = blocks[0].mx, my = blocks[0].my, mz = blocks[0].mz;
  const size_t per = scratch_per_block<FluxT>(mx, my, mz, m->mhd);
  // chunk size: as many blocks as fit a scratch budget of ~4 GiB (at least 1)
  const size_t budget = (size_t)4 << 30;
  int chunk = (int)std::max<size_t>(1, std::min<size_t>(n, budget / per));
  if (int rc = ensure_bytes(&m->scratch, &m->scratch_bytes, per * chunk)) return rc;

  std::vector<BlockDesc<FluxT>> desc;
  for (int c0 = 0; c0 < n; c0 += chunk) {
    const int c1 = std::min(n, c0 + chunk);
    const int nb = c1 - c0;
    fill_desc<FluxT>(m, blocks, c0, c1, desc);
    const size_t dbytes = desc.size() * sizeof(BlockDesc<FluxT>);
    // upload the descriptor table only when it changed
    const char *raw = (const char *)desc.data();
    bool same = (m->last_desc.size() == dbytes) &&
                std::memcmp(m->last_desc.data(), raw, dbytes) == 0;
    if (!same) {
      if (int rc = ensure_bytes(&m->d_desc, &m->desc_bytes, dbytes)) return rc;
      if (hipMemcpyAsync(m->d_desc, raw, dbytes, hipMemcpyHostToDevice, s) !=
          hipSuccess)
        return ENZOE_ERR_HIP;
      // the host copy must stay valid until the copy has been consumed
      if (hipStreamSynchronize(s) != hipSuccess) return ENZOE_ERR_HIP;
      m->last_desc.assign(raw, raw + dbytes);
    }
    const BlockDesc<FluxT> *d = (const BlockDesc<FluxT> *)m->d_desc;

    int stale = 0;
    for (int stage = 0; stage < m->nstages; stage++) {
      const bool is_final = (stage + 1) == m->nstages;
      const int recon = m->recon[stage];
      StageParams sp;
      sp.stage = stage;
      sp.is_final = is_final ? 1 : 0;
      sp.stale = stale;
      sp.cur_stale = stale + immediate_staling(recon);
      // EnzoMethodMHDVlct.cpp:462: dt/2 for the partial step
      sp.dt = (!is_final) ? dt / 2. : dt;
      sp.gamma = m->gamma;
      sp.density_floor = m->dfloor;
      sp.pressure_floor = m->pfloor;
      sp.theta = m->theta;
      const int cs = sp.cur_stale;
      {
        KernelTimer kt(s, recon == RC_NN ? "k_flux_nn" : "k_flux_plm");
        dispatch_flux<FluxT, MHD>(m, recon, d, nb, cube(mx, my, mz, cs, cs), sp, s);
      }
      if (MHD) {
        {
          KernelTimer kt(s, "k_edge");
          hipLaunchKernelGGL((k_edge<FluxT>), dim3(nblk(cube(mx, my, mz, cs, cs), 256), nb),
                             dim3(256), 0, s, d, sp);
        }
        {
          KernelTimer kt(s, "k_faces");
          hipLaunchKernelGGL((k_faces<FluxT>),
                             dim3(nblk(cube(mx, my, mz, cs + 1, cs), 256), nb), dim3(256),
                             0, s, d, sp);
        }
      }
      {
        KernelTimer kt(s, "k_cells");
        const int lo = is_final ? cs : cs + 1;
        hipLaunchKernelGGL((k_cells<FluxT, MHD>),
                           dim3(nblk(cube(mx, my, mz, lo, lo), 256), nb), dim3(256), 0,
                           s, d, sp);
      }
      stale += total_staling(recon);
    }
  }
  if (hipGetLastError() != hipSuccess) return ENZOE_ERR_HIP;
  return 0;
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
  if (p->precision != ENZOE_PREC_DOUBLE && p->precision != ENZOE_PREC_MIXED) {
    delete m;
    return ENZOE_ERR_CONFIG;
  }
  m->precision = p->precision;
  // CT requires exactly 2 partial timesteps (EnzoBfieldMethod.cpp:23-27)
  if (m->mhd && m->nstages != 2) { delete m; return ENZOE_ERR_CONFIG; }
  if (hipMalloc((void **)&m->d_dt, sizeof(unsigned long long)) != hipSuccess) {
    delete m;
    return ENZOE_ERR_ALLOC;
  }
  *out = m;
  return ENZOE_OK;
}

int enzoe_vlct_destroy(enzoe_vlct_t m) {
  if (!m) return ENZOE_ERR_ARG;
  if (m->scratch) (void)hipFree(m->scratch);
  if (m->d_desc) (void)hipFree(m->d_desc);
  if (m->d_dt) (void)hipFree(m->d_dt);
  if (m->d_ptrtab) (void)hipFree(m->d_ptrtab);
  delete m;
  return ENZOE_OK;
}

double enzoe_vlct_courant(enzoe_vlct_t m) { return m ? m->courant : -1.0; }

int enzoe_vlct_compute(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                       double dt, enzoe_stream_t stream) {
  if (!m || !blocks || n <= 0) return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n; i++)
    if (int rc = check_block(m, blocks[i])) return rc;
  // group consecutive blocks of equal shape
  int b0 = 0;
  while (b0 < n) {
    int b1 = b0 + 1;
    while (b1 < n && blocks[b1].mx == blocks[b0].mx &&
           blocks[b1].my == blocks[b0].my && blocks[b1].mz == blocks[b0].mz)
      b1++;
    int rc;
    if (m->precision == ENZOE_PREC_DOUBLE) {
      rc = m->mhd ? compute_group<double, true>(m, blocks + b0, b1 - b0, dt, s)
                  : compute_group<double, false>(m, blocks + b0, b1 - b0, dt, s);
    } else {
      rc = m->mhd ? compute_group<float, true>(m, blocks + b0, b1 - b0, dt, s)
                  : compute_group<float, false>(m, blocks + b0, b1 - b0, dt, s);
    }
    if (rc) return rc;
    b0 = b1;
  }
  return ENZOE_OK;
}

int enzoe_vlct_timestep(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                        enzoe_stream_t stream, double *dt_out) {
  if (!m || !blocks || n <= 0 || !dt_out) return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n; i++)
    if (int rc = check_block(m, blocks[i])) return rc;
  hipLaunchKernelGGL(k_dt_init, dim3(1), dim3(1), 0, s, m->d_dt);
  // descriptors for the dt kernel: only U, pressure, shape and widths used
  std::vector<BlockDesc<double>> desc(n);
  for (int i = 0; i < n; i++) {
    BlockDesc<double> &d = desc[i];
    std::memset(&d, 0, sizeof(d));
    const enzoe_block &b = blocks[i];
    d.U[S_RHO] = b.density;
    for (int a = 0; a < 3; a++) d.U[S_VX + a] = b.velocity[a];
    d.U[S_E] = b.total_energy;
    if (m->mhd)
      for (int a = 0; a < 3; a++) d.U[S_BX + a] = b.bfield[a];
    d.pressure = b.pressure;
    d.mx = b.mx; d.my = b.my; d.mz = b.mz;
    for (int a = 0; a < 3; a++) d.cw[a] = b.cell_width[a];
  }
  const size_t dbytes = desc.size() * sizeof(BlockDesc<double>);
  if (int rc = ensure_bytes(&m->d_ptrtab, &m->ptrtab_bytes, dbytes)) return rc;
  if (hipMemcpyAsync(m->d_ptrtab, desc.data(), dbytes, hipMemcpyHostToDevice,
                     s) != hipSuccess)
    return ENZOE_ERR_HIP;
  long maxcells = 0;
  for (int i = 0; i < n; i++)
    maxcells = std::max<long>(maxcells, (long)blocks[i].mx * blocks[i].my * blocks[i].mz);
  {
    KernelTimer kt(s, "k_dt");
    // enough workgroups to fill the device; each thread strides over cells
    const unsigned wg = (unsigned)std::max<long>(1, std::min<long>(nblk(maxcells, 256),
                                                                   std::max<long>(1, 2048 / n)));
    dim3 grid(wg, n);
    if (m->mhd)
      hipLaunchKernelGGL((k_dt<double, true>), grid, dim3(256), 0, s,
                         (const BlockDesc<double> *)m->d_ptrtab, m->gamma, m->d_dt);
    else
      hipLaunchKernelGGL((k_dt<double, false>), grid, dim3(256), 0, s,
                         (const BlockDesc<double> *)m->d_ptrtab, m->gamma, m->d_dt);
  }
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

int enzoe_grid_refresh(enzoe_vlct_t m, const enzoe_block *blocks, int nbx,
                       int nby, int nbz, const int bc[3],
                       enzoe_stream_t stream) {
  if (!m || !blocks || nbx <= 0 || nby <= 0 || nbz <= 0 || !bc)
    return ENZOE_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int nb = nbx * nby * nbz;
  RefreshGrid G;
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
  }
  // one pointer table per field: 8 cc + 3 face
  const int nf = m->mhd ? 11 : 5;
  std::vector<double *> tab((size_t)nf * nb);
  for (int i = 0; i < nb; i++) {
    const enzoe_block &b = blocks[i];
    double *f[11] = {b.density, b.velocity[0], b.velocity[1], b.velocity[2],
                     b.total_energy, b.bfield[0], b.bfield[1], b.bfield[2],
                     b.bfieldi[0], b.bfieldi[1], b.bfieldi[2]};
    for (int q = 0; q < nf; q++) {
      int slot = m->mhd ? q : q;  // hydro: first 5
      tab[(size_t)q * nb + i] = f[slot];
    }
  }
  const size_t bytes = tab.size() * sizeof(double *);
  // the pointer table is uploaded only when the block list changes
  static thread_local void *d_tab = nullptr;
  static thread_local size_t d_tab_bytes = 0;
  static thread_local std::vector<double *> last_tab;
  if (last_tab != tab) {
    if (int rc = ensure_bytes(&d_tab, &d_tab_bytes, bytes)) return rc;
    if (hipMemcpyAsync(d_tab, tab.data(), bytes, hipMemcpyHostToDevice, s) !=
        hipSuccess)
      return ENZOE_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return ENZOE_ERR_HIP;
    last_tab = tab;
  }
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
                     (double *const *)d_tab, G);
  if (hipGetLastError() != hipSuccess) return ENZOE_ERR_HIP;
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
