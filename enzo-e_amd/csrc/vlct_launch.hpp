// vlct_launch.hpp -- host-side launch sequence of the VL+CT step, shared by
// the FP64 translation unit (vlct_capi.hip, built without FMA contraction
// for bitwise reference parity) and the mixed-precision one
// (vlct_mixed.hip, FP32 reconstruction and Riemann fluxes).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <string>
#include <cmath>
#include <type_traits>
#include <vector>

#include "../../include/enzoe_vlct.h"
#include "vlct_kernels.hpp"

namespace enzoe {

// ----------------------------------------------------------- profiling
// Launches are bracketed by HIP events on their stream while profiling is
// on; the events are resolved only when a report is asked for.
struct ProfEntry {
  double total_ms = 0;
  long count = 0;
};
struct PendingTiming {
  const char *name;
  hipEvent_t a, b;
};
extern bool g_profiling;
extern std::map<std::string, ProfEntry> g_prof;
extern std::vector<PendingTiming> g_pending;
extern std::vector<hipEvent_t> g_event_pool;

hipEvent_t pool_event();
void resolve_pending();

struct KernelTimer {
  hipStream_t s;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  bool on;
  KernelTimer(hipStream_t s_, const char *n, bool allowed = true)
      : s(s_), name(n), on(allowed && g_profiling) {
    if (on) {
      a = pool_event();
      b = pool_event();
      (void)hipEventRecord(a, s);
    }
  }
  ~KernelTimer() {
    if (on) {
      (void)hipEventRecord(b, s);
      g_pending.push_back({name, a, b});
      if (g_pending.size() > 4096) resolve_pending();
    }
  }
};

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
// workgroup count padded to a multiple of the 8 XCDs (see wg_coords)
inline unsigned nblk8(long n, int bs) { return (nblk(n, bs) + 7u) & ~7u; }

}  // namespace enzoe

// --------------------------------------------------------------- method

struct enzoe_vlct_method {
  int mhd = 1;
  int riemann = ENZOE_KNS::RS_HLLD;
  int recon[2] = {ENZOE_KNS::RC_NN, ENZOE_KNS::RC_PLM};
  int nstages = 2;
  int precision = ENZOE_PREC_DOUBLE;
  // kernels launched in XCD-contiguous workgroup order (StageParams mask);
  // measured on the 256^3 Orszag-Tang block it pays for k_edge only
  int xcd_remap = 2;
  double theta = 1.5, courant = 0.3, gamma = 5.0 / 3.0;
  double dfloor = 1e-200, pfloor = 1e-200;
  // launch structure of the stages: for MHD constrained transport and the
  // cell update as three launches (0: k_edge, k_faces, k_cells) or fused
  // launches over tiles (1..4, ctu_variant); 5: each stage as one
  // persistent launch (k_stage, MHD or hydro); same bits every way; -1
  // (the default): the autotuner's choice per block shape, three launches
  // until it has run
  int ctu = -1;
  int force_ctu = -1;                    // set while the autotuner times one
  std::map<long long, int> tuned_ctu;    // the autotuner's choice per shape
  // rows of the tiled cell order per block shape (the autotuner's choice;
  // tile_rows_for otherwise) and the value forced while it times one
  std::map<long long, int> tuned_rows;
  int force_rows = -1;
  // the shape of the flux launch for blocks of this shape (flux_shape_of:
  // 0 three faces per thread, 1 one face per thread, 2 z pencils for the
  // nearest-neighbour stage): the autotuner's choice, the value forced
  // while it times one, and the method's setting (-1: the autotuner's)
  std::map<long long, int> tuned_shape;
  int force_shape = -1;
  int flux_shape = -1;
  // the persistent stage kernel's grid barrier (a GridBarrier in device
  // memory, zeroed once), or null until first needed
  void *d_bar = nullptr;
  // dual-energy formalism ("modern" formulation) and its eta
  int dual = 0;
  double eta = 0.001;

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
  std::vector<char> last_dt_desc;  // host copy of the dt descriptor table
  void *d_rtab = nullptr;          // refresh pointer table
  size_t rtab_bytes = 0;
  std::vector<double *> last_rtab;
  size_t scratch_budget = 0;       // bytes of the chunk scratch pool
  bool dry = false;                // prepare buffers and tables, launch nothing
  // fixed workspace of the allocation-free entry point while it runs (the
  // descriptor table, then the stage scratch), or null
  char *arena = nullptr;
  size_t arena_bytes = 0;
  // launch-order mask chosen by enzoe_vlct_autotune per block shape, and a
  // mask forced while the autotuner times a candidate (-1: none)
  std::map<long long, int> tuned_mask;
  int force_mask = -1;
  void *d_dts = nullptr;  // per-block dt uploaded by enzoe_vlct_compute_dt
  size_t dts_bytes = 0;
  // device tables of the flux-correction calls (FluxData pointers of
  // enzoe_vlct_compute_store_fluxes, links of the batched send, blocks of
  // the batched correction), re-uploaded only when they change
  void *d_fdtab = nullptr, *d_fltab = nullptr, *d_fbtab = nullptr;
  size_t fdtab_bytes = 0, fltab_bytes = 0, fbtab_bytes = 0;
  std::vector<char> last_fdtab, last_fltab, last_fbtab;
  // debug counters (DBG_* in vlct_kernels.hpp): device memory while the
  // checks are on (enzoe_vlct_set_debug_checks), null otherwise
  unsigned long long *d_dbg = nullptr;
  // transfer tables of enzoe_vlct_transfer_batch: two, so that the
  // restriction and prolongation batches of a cycle both stay cached
  void *d_trtab[2] = {nullptr, nullptr};
  size_t trtab_bytes[2] = {0, 0};
  std::vector<char> last_trtab[2];
};

namespace ENZOE_KNS {
// the profiling helpers above stay in `enzoe` (one set for the library)
using namespace ::enzoe;


inline long long shape_key(int mx, int my, int mz) {
  return ((long long)mx << 40) | ((long long)my << 20) | (long long)mz;
}

// The launch structure of the stages for blocks of this shape: 0 =
// k_flux, then k_edge + k_faces + k_cells; for MHD 1..3 = k_ctu with tile
// CtuTile1..3, 4 = k_edge + k_ctu without its edge phase (faces and cells
// fused); 5 = the persistent k_stage (persistent_ok); the method's
// setting, or the autotuner's choice for the shape when the setting is
// "auto" (-1).
inline int ctu_variant(const enzoe_vlct_method *m, int mx, int my, int mz) {
  if (m->force_ctu >= 0) return m->force_ctu;
  if (m->ctu >= 0) return m->ctu;
  auto it = m->tuned_ctu.find(shape_key(mx, my, mz));
  return it != m->tuned_ctu.end() ? it->second : 0;
}

// The fused CT update needs the two stages of the VL scheme (its final
// stage reads the first stage's copy of the start-of-step faces); the
// scratch holds that copy whenever the fused launch may be chosen.
inline bool ctu_possible(const enzoe_vlct_method *m) {
  return m->mhd && m->nstages == 2 && (m->ctu != 0 || m->force_ctu > 0);
}

// The persistent stage launch (structure 5) exists for HLLD MHD and for
// hydro without the dual-energy formalism, and not on the allocation-free
// path (its barrier is a buffer of the method).
inline bool persistent_ok(const enzoe_vlct_method *m) {
  return !m->dual && (!m->mhd || m->riemann == RS_HLLD) && m->arena == nullptr;
}

// The method's grid barrier, allocated and zeroed on first use (not while
// a graph is being captured: compute_group takes it before any launch,
// also in the dry run that precedes a capture).
inline int ensure_barrier(enzoe_vlct_method *m) {
  if (m->d_bar) return 0;
  void *p = nullptr;
  if (hipMalloc(&p, sizeof(GridBarrier)) != hipSuccess) return ENZOE_ERR_ALLOC;
  if (hipMemset(p, 0, sizeof(GridBarrier)) != hipSuccess) {
    (void)hipFree(p);
    return ENZOE_ERR_HIP;
  }
  m->d_bar = p;
  return 0;
}

// Workgroups of a persistent stage kernel the current device holds at once
// (occupancy x CUs), cached per device; 0 on failure.
template <typename FluxT, int RECON, int RIEMANN, bool MHD, bool PRIMF>
int stage_capacity() {
  static std::atomic<int> cap[16];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 0;
  int c = cap[dev].load(std::memory_order_relaxed);
  if (c == 0) {
    int per = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, k_stage<FluxT, FluxT, RECON, RIEMANN, MHD, PRIMF>, 256, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    c = per * cus;
    cap[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

// One stage as one persistent launch: as many workgroups as the largest
// phase needs, at most what the device holds at once (a multiple of 8).
template <typename FluxT, int RECON, int RIEMANN, bool MHD, bool PRIMF>
int launch_stage(const enzoe_vlct_method *m, const BlockDesc<FluxT> *d, int nb,
                 const int nwg[4], StageParams sp, hipStream_t s) {
  const int cap = stage_capacity<FluxT, RECON, RIEMANN, MHD, PRIMF>() & ~7;
  if (cap < 8 || !m->d_bar) return ENZOE_ERR_HIP;
  long need = 0;
  for (int q = 0; q < 4; q++) need = std::max(need, (long)nwg[q] * nb);
  const int G = (int)std::min<long>(cap, std::max<long>(8, (need + 7) & ~7L));
  PersistParams pp;
  pp.nb = nb;
  for (int q = 0; q < 4; q++) pp.nwg[q] = nwg[q];
  pp.bar = (GridBarrier *)m->d_bar;
  pp.limit = 1u << 22;
  sp.xcd_remap = 0;  // the phases walk their boxes in the plain order
  hipLaunchKernelGGL((k_stage<FluxT, FluxT, RECON, RIEMANN, MHD, PRIMF>), dim3(G), dim3(256),
                     0, s, d, sp, pp);
  return 0;
}

template <typename FluxT, bool MHD>
int dispatch_stage(const enzoe_vlct_method *m, int recon, bool primf,
                   const BlockDesc<FluxT> *d, int nb, const int nwg[4],
                   const StageParams &sp, hipStream_t s) {
  constexpr int RS = MHD ? RS_HLLD : RS_HLLC;
  if constexpr (std::is_same<FluxT, float>::value) {
    if (primf) {
      if (recon == RC_NN) return launch_stage<FluxT, RC_NN, RS, MHD, true>(m, d, nb, nwg, sp, s);
      if (recon == RC_PLM)
        return launch_stage<FluxT, RC_PLM, RS, MHD, true>(m, d, nb, nwg, sp, s);
      return launch_stage<FluxT, RC_PLM_ATHENA, RS, MHD, true>(m, d, nb, nwg, sp, s);
    }
  }
  if (recon == RC_NN) return launch_stage<FluxT, RC_NN, RS, MHD, false>(m, d, nb, nwg, sp, s);
  if (recon == RC_PLM) return launch_stage<FluxT, RC_PLM, RS, MHD, false>(m, d, nb, nwg, sp, s);
  return launch_stage<FluxT, RC_PLM_ATHENA, RS, MHD, false>(m, d, nb, nwg, sp, s);
}
template <typename FluxT, bool DE, class Tile, bool EDGES_LDS = true>
void launch_ctu(const BlockDesc<FluxT> *d, int nb, int mx, int my, int mz, int lo,
                StageParams sp, hipStream_t s) {
  int nt[3];
  ctu_tiles<Tile>(mx, my, mz, lo, nt);
  const long tiles = (long)nt[0] * nt[1] * nt[2];
  hipLaunchKernelGGL((k_ctu<FluxT, DE, Tile, EDGES_LDS>),
                     dim3(nblk8(tiles * Tile::NT, Tile::NT), nb), dim3(Tile::NT), 0, s, d, sp);
}

inline int total_staling(int r) { return r == RC_NN ? 1 : 2; }
inline int immediate_staling(int r) { return r == RC_NN ? 0 : 1; }

inline int parse_recon(const std::string &n, int *out) {
  if (n == "nn") { *out = RC_NN; return 0; }
  if (n == "plm" || n == "plm_enzo") { *out = RC_PLM; return 0; }
  if (n == "plm_athena") { *out = RC_PLM_ATHENA; return 0; }
  return ENZOE_ERR_CONFIG;
}

// The mixed path (FP32 fluxes, no dual energy) keeps its half-step state as
// FP32 primitives.
template <typename FluxT>
inline constexpr bool half_f32(int dual) {
  return std::is_same<FluxT, float>::value && !dual;
}

// bytes of scratch per block
template <typename FluxT>
inline size_t scratch_per_block(int mx, int my, int mz, int mhd, int dual = 0, int ctu = 0) {
  size_t ncc = (size_t)mx * my * mz;
  size_t nq = mhd ? 8 : 5;
  size_t bytes = 0;
  // half-step state: FP64 integration quantities (Ut), or in the mixed path
  // FP32 primitives (Up)
  bytes += nq * ncc * (half_f32<FluxT>(dual) ? sizeof(float) : sizeof(double));
  if (mhd) {
    const size_t faces = ((size_t)mz * my * (mx + 1) + (size_t)mz * (my + 1) * mx +
                          (size_t)(mz + 1) * my * mx) * sizeof(double);
    bytes += faces;                                    // bft
    if (ctu) bytes += faces;                           // bf0
    bytes += 6 * ncc * sizeof(double);                 // E, Ec
    if (half_f32<FluxT>(dual)) bytes += 3 * ncc * sizeof(double);  // Ech
  }
  bytes += 3 * (mhd ? 7 : 5) * ncc * sizeof(FluxT);    // F
  if (dual) bytes += 7 * ncc * sizeof(double);         // eit, Fe, Vb
  return bytes;
}

inline int ensure_bytes(void **ptr, size_t *have, size_t need) {
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

// Copies a host table to a device buffer unless the buffer already holds
// exactly these bytes; the stream is synchronised after a copy, so the
// host table may change as soon as the call returns.
inline int upload_table(void **dptr, size_t *have, std::vector<char> *last,
                        const void *host, size_t bytes, hipStream_t s) {
  const char *raw = (const char *)host;
  if (*dptr && last->size() == bytes && std::memcmp(last->data(), raw, bytes) == 0)
    return 0;
  if (int rc = ensure_bytes(dptr, have, bytes)) return rc;
  if (hipMemcpyAsync(*dptr, raw, bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ENZOE_ERR_HIP;
  last->assign(raw, raw + bytes);
  return 0;
}

inline int check_block(const enzoe_vlct_method *m, const enzoe_block &b) {
  if (b.mx < 7 || b.my < 7 || b.mz < 7) return ENZOE_ERR_ARG;
  // cell indices of the tiled order are 32-bit
  if ((long long)(b.mx + 1) * (b.my + 1) * (b.mz + 1) >= (1LL << 31)) return ENZOE_ERR_ARG;
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
  // the dual-energy formalism integrates the internal_energy field
  if (m->dual && !b.internal_energy) return ENZOE_ERR_ARG;
  return 0;
}

// Rows of the tiled cell order for a block with planes of mx x my cells:
// the largest of 64, 32, 16, 8 rows whose slab planes hold the flux
// kernel's z stencil (about 5 planes of 11 words per cell) in 2 MB, half an
// XCD's L2; 0 (plain order) when a whole plane of that already fits in 4 MB.
inline int tile_rows_for(int mx, int my) {
  const double row = (double)mx * 11 * 8 * 5;
  if (row * my <= double(4 << 20)) return 0;
  int rows = 64;
  while (rows > 8 && row * rows > double(2 << 20)) rows /= 2;
  return rows;
}


// The flux launch for blocks of this shape: 0, three faces per thread
// (k_flux); 1, one face per thread (k_flux over three z slices, for blocks
// too small to fill the device); 2, the nearest-neighbour stage as z
// pencils with LDS-staged planes (k_flux_zp; the other stages as 0).  The
// setting, else the autotuner's choice, else 0.
inline int flux_shape_of(const enzoe_vlct_method *m, int mx, int my, int mz) {
  if (m->force_shape >= 0) return m->force_shape;
  if (m->flux_shape >= 0) return m->flux_shape;
  auto it = m->tuned_shape.find(shape_key(mx, my, mz));
  return it != m->tuned_shape.end() ? it->second : 0;
}

// Rows of the tiled cell order for blocks of this shape.
inline int tile_rows_of(const enzoe_vlct_method *m, int mx, int my, int mz) {
  if (m->force_rows >= 0) return m->force_rows;
  auto it = m->tuned_rows.find(shape_key(mx, my, mz));
  return it != m->tuned_rows.end() ? it->second : tile_rows_for(mx, my);
}

// Fills descriptors for blocks [b0, b1) into out[0 .. b1-b0), their scratch
// laid out from `base`.
// accel: 3 acceleration pointers per block of the call (x, y, z), or null
template <typename FluxT>
void fill_desc_at(const enzoe_vlct_method *m, char *base, const enzoe_block *blocks,
                  int b0, int b1, BlockDesc<FluxT> *out,
                  const double *const *accel = nullptr) {
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
    d.g = b.ghost;
    d.tile_rows = tile_rows_of(m, b.mx, b.my, b.mz);
    if (accel)
      for (int a = 0; a < 3; a++) d.acc[a] = accel[3 * bi + a];
    for (int a = 0; a < 3; a++) d.cw[a] = b.cell_width[a];
    const size_t ncc = (size_t)b.mx * b.my * b.mz;
    const int nq = m->mhd ? 8 : 5;
    for (int q = 0; q < S_N; q++) {
      if (!m->mhd && q >= S_BX) continue;
      if (half_f32<FluxT>(m->dual)) {
        d.Up[q] = (float *)base;
        base += ncc * sizeof(float);
      } else {
        d.Ut[q] = (double *)base;
        base += ncc * sizeof(double);
      }
    }
    (void)nq;
    if (m->mhd) {
      d.bft[0] = (double *)base; base += (size_t)b.mz * b.my * (b.mx + 1) * sizeof(double);
      d.bft[1] = (double *)base; base += (size_t)b.mz * (b.my + 1) * b.mx * sizeof(double);
      d.bft[2] = (double *)base; base += (size_t)(b.mz + 1) * b.my * b.mx * sizeof(double);
      if (ctu_possible(m)) {
        d.bf0[0] = (double *)base; base += (size_t)b.mz * b.my * (b.mx + 1) * sizeof(double);
        d.bf0[1] = (double *)base; base += (size_t)b.mz * (b.my + 1) * b.mx * sizeof(double);
        d.bf0[2] = (double *)base; base += (size_t)(b.mz + 1) * b.my * b.mx * sizeof(double);
      }
      for (int a = 0; a < 3; a++) {
        d.E[a] = (double *)base;
        base += ncc * sizeof(double);
      }
      for (int a = 0; a < 3; a++) {
        d.Ec[a] = (double *)base;
        base += ncc * sizeof(double);
      }
      if (half_f32<FluxT>(m->dual))
        for (int a = 0; a < 3; a++) {
          d.Ech[a] = (double *)base;
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
    if (m->dual) {
      d.ei = b.internal_energy;
      d.eit = (double *)base;
      base += ncc * sizeof(double);
      for (int a = 0; a < 3; a++) {
        d.Fe[a] = (double *)base;
        base += ncc * sizeof(double);
        d.Vb[a] = (double *)base;
        base += ncc * sizeof(double);
      }
    }
    out[bi - b0] = d;
  }
}

// fill_desc_at over the method's scratch pool, into a vector
template <typename FluxT>
void fill_desc(const enzoe_vlct_method *m, const enzoe_block *blocks, int b0,
               int b1, std::vector<BlockDesc<FluxT>> &out,
               const double *const *accel = nullptr) {
  out.resize(b1 - b0);
  fill_desc_at<FluxT>(m, (char *)m->scratch, blocks, b0, b1, out.data(), accel);
}

// cells of the box [lo, m - hi_trim) of a block
inline long cube(int mx, int my, int mz, int lo, int hi_trim) {
  long a = mx - lo - hi_trim, b = my - lo - hi_trim, c = mz - lo - hi_trim;
  if (a <= 0 || b <= 0 || c <= 0) return 0;
  return a * b * c;
}

// flux launch of one stage; shape as flux_shape_of (2 applies to the
// nearest-neighbour reconstruction only)
template <typename FluxT, int RECON, int RIEMANN, bool MHD, bool DE, bool PRIMF>
void launch_flux(const BlockDesc<FluxT> *d, int nb, int mx, int my, int mz,
                 StageParams sp, hipStream_t s, int shape) {
  const int cs = sp.cur_stale;
  if (RECON == RC_NN && shape == 2) {
    // z pencils: ZP_TX x ZP_TY tiles, each marching through a segment of
    // planes; about 2048 workgroups per block
    const int nx = mx - 2 * cs, ny = my - 2 * cs, nz = mz - 2 * cs;
    const int tiles = ((nx + ZP_TX - 1) / ZP_TX) * ((ny + ZP_TY - 1) / ZP_TY);
    const int want = std::max(1, std::min(nz, (2048 + tiles - 1) / tiles));
    const int zplanes = (nz + want - 1) / want;
    const int nseg = (nz + zplanes - 1) / zplanes;
    hipLaunchKernelGGL((k_flux_zp<FluxT, FluxT, RIEMANN, MHD, DE, PRIMF>),
                       dim3(tiles * nseg, nb), dim3(256), 0, s, d, sp, zplanes);
    return;
  }
  const long nthreads = cube(mx, my, mz, cs, cs);
  dim3 grid(nblk8(nthreads, 256), nb, shape == 1 ? 3 : 1);
  hipLaunchKernelGGL((k_flux<FluxT, FluxT, RECON, RIEMANN, MHD, DE, PRIMF>), grid,
                     dim3(256), 0, s, d, sp);
}

template <typename FluxT, bool MHD, bool DE, bool PRIMF>
void dispatch_flux_p(const enzoe_vlct_method *m, int recon,
                     const BlockDesc<FluxT> *d, int nb, int mx, int my, int mz,
                     StageParams sp, hipStream_t s, int shape) {
  constexpr int RS = MHD ? RS_HLLD : RS_HLLC;
  if (MHD && m->riemann != RS_HLLD) {
    if (recon == RC_NN)
      launch_flux<FluxT, RC_NN, RS_HLLE, MHD, DE, PRIMF>(d, nb, mx, my, mz, sp, s, shape);
    else if (recon == RC_PLM)
      launch_flux<FluxT, RC_PLM, RS_HLLE, MHD, DE, PRIMF>(d, nb, mx, my, mz, sp, s, shape);
    else
      launch_flux<FluxT, RC_PLM_ATHENA, RS_HLLE, MHD, DE, PRIMF>(d, nb, mx, my, mz, sp, s, shape);
    return;
  }
  if (recon == RC_NN)
    launch_flux<FluxT, RC_NN, RS, MHD, DE, PRIMF>(d, nb, mx, my, mz, sp, s, shape);
  else if (recon == RC_PLM)
    launch_flux<FluxT, RC_PLM, RS, MHD, DE, PRIMF>(d, nb, mx, my, mz, sp, s, shape);
  else
    launch_flux<FluxT, RC_PLM_ATHENA, RS, MHD, DE, PRIMF>(d, nb, mx, my, mz, sp, s, shape);
}

// primf: the stage reads the mixed path's FP32 primitive half-step state;
// shape: flux_shape_of
template <typename FluxT, bool MHD, bool DE>
void dispatch_flux(const enzoe_vlct_method *m, int recon, bool primf,
                   const BlockDesc<FluxT> *d, int nb, int mx, int my, int mz,
                   StageParams sp, hipStream_t s, int shape) {
  if constexpr (half_f32<FluxT>(DE ? 1 : 0)) {
    if (primf) {
      dispatch_flux_p<FluxT, MHD, DE, true>(m, recon, d, nb, mx, my, mz, sp, s, shape);
      return;
    }
  }
  dispatch_flux_p<FluxT, MHD, DE, false>(m, recon, d, nb, mx, my, mz, sp, s, shape);
}

// gamma (gamma - 1) as the reference's synchronisation holds it: a float
// (EnzoPhysicsFluidProps.cpp:236)
inline double sync_ggm1(double gamma) {
  return (double)(float)(gamma * (gamma - 1.));
}


// XCD-order mask of the stage kernels for blocks of this shape
inline int launch_mask(const enzoe_vlct_method *m, int mx, int my, int mz) {
  if (m->force_mask >= 0) return m->force_mask;
  auto it = m->tuned_mask.find(shape_key(mx, my, mz));
  return it != m->tuned_mask.end() ? it->second : m->xcd_remap;
}

// Saving the final stage's face fluxes for flux corrections: the fields and
// the device table of one FluxData pointer per block of the call.
struct StoreArgs {
  StoreSlots sl;
  double *const *d_fdtab;
};

// Bytes of the chunk scratch pool: the larger of 4 GiB and a third of the
// free device memory, at most 16 GiB (taken once per method).
inline size_t scratch_budget(enzoe_vlct_method *m) {
  if (m->scratch_budget == 0) {
    size_t budget = (size_t)4 << 30;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      budget = std::max(budget, std::min((size_t)16 << 30, free_b / 3));
    m->scratch_budget = budget;
  }
  return m->scratch_budget;
}

// The launch sequence of compute_group over chunks of `chunk` blocks whose
// descriptors are table[0 .. n) in device memory.  timers: bracket the
// kernels with profiling events (when profiling is on).
template <typename FluxT, bool MHD, bool DE>
int launch_chunks(enzoe_vlct_method *m, const enzoe_block *blocks, int n, int chunk,
                  const BlockDesc<FluxT> *table, double dt, const double *dts,
                  hipStream_t s, const StoreArgs *st, unsigned long long *cfl,
                  bool timers) {
  const int mx = blocks[0].mx, my = blocks[0].my, mz = blocks[0].mz;
  for (int c0 = 0; c0 < n; c0 += chunk) {
    const int c1 = std::min(n, c0 + chunk);
    const int nb = c1 - c0;
    const BlockDesc<FluxT> *d = table + c0;

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
      sp.dts = dts ? dts + c0 : nullptr;
      sp.dt_factor = (!is_final) ? 0.5 : 1.0;
      sp.gamma = m->gamma;
      sp.density_floor = m->dfloor;
      sp.pressure_floor = m->pfloor;
      sp.theta = m->theta;
      sp.xcd_remap = launch_mask(m, mx, my, mz);
      sp.eta = m->eta;
      sp.ggm1 = sync_ggm1(m->gamma);
      sp.cfl = is_final ? cfl : nullptr;
      sp.dbg = m->d_dbg;
      const int cs = sp.cur_stale;
      const bool primf = stage > 0 && half_f32<FluxT>(DE ? 1 : 0);
      // with the fused CFL the final update's launch covers the block
      const int lo = sp.cfl ? 0 : (is_final ? cs : cs + 1);
      const int structure = ctu_variant(m, mx, my, mz);
      if constexpr (!DE) {
        if (structure == 5 && persistent_ok(m) && m->d_bar) {
          // the stage's four kernels as one persistent launch
          const int nwg[4] = {(int)nblk(cube(mx, my, mz, cs, cs), 256),
                              (int)nblk(cube(mx, my, mz, cs, cs), 256),
                              (int)nblk(cube(mx, my, mz, cs + 1, cs), 256),
                              (int)nblk(cube(mx, my, mz, lo, lo), 256)};
          KernelTimer kt(s, recon == RC_NN ? "k_stage_nn" : "k_stage_plm", timers);
          if (int rc = dispatch_stage<FluxT, MHD>(m, recon, primf, d, nb, nwg, sp, s)) return rc;
          stale += total_staling(recon);
          continue;
        }
      }
      {
        KernelTimer kt(s, recon == RC_NN ? "k_flux_nn" : "k_flux_plm", timers);
        dispatch_flux<FluxT, MHD, DE>(m, recon, primf, d, nb, mx, my, mz, sp, s,
                                      flux_shape_of(m, mx, my, mz));
      }
      const int variant =
          (MHD && ctu_possible(m) && structure >= 1 && structure <= 4) ? structure : 0;
      if (variant == 4) {
        {
          KernelTimer kt(s, "k_edge", timers);
          hipLaunchKernelGGL((k_edge<FluxT>), dim3(nblk8(cube(mx, my, mz, cs, cs), 256), nb),
                             dim3(256), 0, s, d, sp);
        }
        KernelTimer kt(s, "k_fc", timers);
        launch_ctu<FluxT, DE, CtuTile1, false>(d, nb, mx, my, mz, lo, sp, s);
      } else if (variant > 0) {
        KernelTimer kt(s, "k_ctu", timers);
        if (variant == 2)
          launch_ctu<FluxT, DE, CtuTile2>(d, nb, mx, my, mz, lo, sp, s);
        else if (variant == 3)
          launch_ctu<FluxT, DE, CtuTile3>(d, nb, mx, my, mz, lo, sp, s);
        else
          launch_ctu<FluxT, DE, CtuTile1>(d, nb, mx, my, mz, lo, sp, s);
      } else {
        if (MHD) {
          {
            KernelTimer kt(s, "k_edge", timers);
            hipLaunchKernelGGL((k_edge<FluxT>), dim3(nblk8(cube(mx, my, mz, cs, cs), 256), nb),
                               dim3(256), 0, s, d, sp);
          }
          {
            KernelTimer kt(s, "k_faces", timers);
            hipLaunchKernelGGL((k_faces<FluxT>),
                               dim3(nblk8(cube(mx, my, mz, cs + 1, cs), 256), nb), dim3(256),
                               0, s, d, sp);
          }
        }
        KernelTimer kt(s, "k_cells", timers);
        hipLaunchKernelGGL((k_cells<FluxT, MHD, DE>),
                           dim3(nblk8(cube(mx, my, mz, lo, lo), 256), nb), dim3(256), 0,
                           s, d, sp);
      }
      stale += total_staling(recon);
    }
    if (st) {
      // the chunk's flux scratch still holds the final stage's fluxes
      long work = 0;
      for (int bi = c0; bi < c1; bi++) {
        const int g = blocks[bi].ghost;
        FluxLayout L;
        L.init(mx - 2 * g, my - 2 * g, mz - 2 * g);
        work = std::max(work, L.per_field * st->sl.nf);
      }
      KernelTimer kt(s, "k_store_fluxes", timers);
      hipLaunchKernelGGL((k_store_fluxes<FluxT>), dim3(nblk8(work, 256), nb),
                         dim3(256), 0, s, d, st->d_fdtab + c0, st->sl, dt,
                         dts ? dts + c0 : nullptr);
    }
  }
  if (hipGetLastError() != hipSuccess) return ENZOE_ERR_HIP;
  return 0;
}

// One VL+CT step (EnzoMethodMHDVlct::compute) on blocks of equal shape.
// dts: per-block dt in device memory (n values) or null (all blocks use dt).
// st: store the final stage's face fluxes (hydro only) or null.
// DE: the dual-energy formalism (FP64 path only).
// cfl (non-null): the final stage's k_cells also takes the CFL minimum of
// its cells and writes their pressure.
template <typename FluxT, bool MHD, bool DE = false>
int compute_group(enzoe_vlct_method *m, const enzoe_block *blocks, int n,
                  double dt, const double *dts, hipStream_t s,
                  const StoreArgs *st = nullptr,
                  const double *const *accel = nullptr,
                  unsigned long long *cfl = nullptr) {
  const int mx = blocks[0].mx, my = blocks[0].my, mz = blocks[0].mz;
  if (DE != (m->dual != 0)) return ENZOE_ERR_CONFIG;
  const size_t per = scratch_per_block<FluxT>(mx, my, mz, m->mhd, m->dual, ctu_possible(m));
  if (m->arena) {
    // the allocation-free path: descriptor table and scratch in the fixed
    // workspace; nothing is allocated, nothing cached across calls
    const size_t dbytes = ((size_t)n * sizeof(BlockDesc<FluxT>) + 255) & ~(size_t)255;
    if (dbytes >= m->arena_bytes || m->arena_bytes - dbytes < per) return ENZOE_ERR_SIZE;
    const int chunk =
        (int)std::max<size_t>(1, std::min<size_t>(n, (m->arena_bytes - dbytes) / per));
    std::vector<BlockDesc<FluxT>> desc(n);
    for (int c0 = 0; c0 < n; c0 += chunk)
      fill_desc_at<FluxT>(m, m->arena + dbytes, blocks, c0, std::min(n, c0 + chunk),
                          desc.data() + c0, accel);
    if (hipMemcpyAsync(m->arena, desc.data(), n * sizeof(BlockDesc<FluxT>),
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return ENZOE_ERR_HIP;
    return launch_chunks<FluxT, MHD, DE>(m, blocks, n, chunk, (const BlockDesc<FluxT> *)m->arena,
                                         dt, dts, s, st, cfl, false);
  }
  // chunk size: as many blocks as fit the scratch budget (at least 1); the
  // chunks run one after the other on the stream and share the pool
  const size_t budget = scratch_budget(m);
  int chunk = (int)std::max<size_t>(1, std::min<size_t>(n, budget / per));
  if (int rc = ensure_bytes(&m->scratch, &m->scratch_bytes, per * chunk)) return rc;

  // descriptor table of every chunk, uploaded once (and only when changed)
  std::vector<BlockDesc<FluxT>> desc, part;
  for (int c0 = 0; c0 < n; c0 += chunk) {
    fill_desc<FluxT>(m, blocks, c0, std::min(n, c0 + chunk), part, accel);
    desc.insert(desc.end(), part.begin(), part.end());
  }
  {
    const size_t dbytes = desc.size() * sizeof(BlockDesc<FluxT>);
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
  }
  if (!DE && persistent_ok(m) && (m->ctu == -1 || m->ctu == 5 || m->force_ctu == 5))
    if (int rc = ensure_barrier(m)) return rc;
  if (m->dry) return 0;
  return launch_chunks<FluxT, MHD, DE>(m, blocks, n, chunk,
                                       (const BlockDesc<FluxT> *)m->d_desc, dt, dts, s, st,
                                       cfl, true);
}

// The launch sequence of n blocks (chunks of `chunk`) of the FP32 flux path
// over a descriptor table prepared in device memory, per-block dt in dts
// (device); compiled in vlct_mixed.hip.
int launch_prepared_mixed(enzoe_vlct_method *m, const enzoe_block *blocks, int n, int chunk,
                          const void *table, const double *dts, hipStream_t s);

// FP32 flux path, compiled in vlct_mixed.hip
int compute_mixed(enzoe_vlct_method *m, const enzoe_block *blocks, int n,
                  double dt, const double *dts, hipStream_t s,
                  const StoreArgs *st = nullptr,
                  const double *const *accel = nullptr,
                  unsigned long long *cfl = nullptr);

}  // namespace ENZOE_KNS

namespace enzoe {

// FP64 step with multiply-adds contracted into FMAs, compiled in
// vlct_fp64c.hip: compute_runs' and the prepared launch sequence's
// counterparts (st: the StoreArgs of the flux store, or null)
int compute_contracted(enzoe_vlct_method *m, const enzoe_block *blocks, int n, double dt,
                       const double *dts, hipStream_t s, const void *st,
                       const double *const *accel, unsigned long long *cfl);
int launch_prepared_contracted(enzoe_vlct_method *m, const enzoe_block *blocks, int n,
                               int chunk, const void *table, const double *dts,
                               hipStream_t s);

}  // namespace enzoe
