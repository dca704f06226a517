/*
 * vlct_harness.c -- CPU test harness around the oracle (test infrastructure).
 *
 * Provides what the reference's regression decks need around the VL+CT
 * method: a uniform grid of blocks, ghost refresh (periodic / outflow), the
 * Cello cycle (dt reduction then compute), and the initial conditions of the
 * reference's linear-wave and shock-tube problems.  The control flow follows
 * src/Cello/control_*.cpp; the initial conditions follow
 * src/Enzo/initial/EnzoInitialInclinedWave.cpp and EnzoInitialShockTube.cpp
 * (recipes only, written here from scratch).
 */
#include "vlct_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(k, j, i, ny, nx) \
  ((((long)(k)) * (long)(ny) + (long)(j)) * (long)(nx) + (long)(i))

static const double CELLO_PI = 3.14159265358979324; /* cello.hpp:623 */

/* ------------------------------------------------------------------------ */
/* Grid                                                                     */
/* ------------------------------------------------------------------------ */

/* Block::lower / Block::upper (src/Cello/mesh_Block.cpp:832-880) and
 * FieldData::cell_width (data_FieldData.cpp:254-264). */
static void block_extent(const o_grid *gr, const int ib[3], double lo[3],
                         double hi[3]) {
  const int nb[3] = {gr->nbx, gr->nby, gr->nbz};
  for (int a = 0; a < 3; a++) {
    double al = 1.0 * ib[a] / nb[a];
    double ah = 1.0 * (ib[a] + 1) / nb[a];
    lo[a] = (1.0 - al) * gr->lower[a] + al * gr->upper[a];
    hi[a] = (1.0 - ah) * gr->lower[a] + ah * gr->upper[a];
  }
}

int o_grid_alloc(o_grid *gr) {
  const int nb = gr->nbx * gr->nby * gr->nbz;
  gr->blocks = (o_block *)calloc((size_t)nb, sizeof(o_block));
  if (!gr->blocks) return -1;
  const int n[3] = {gr->nx, gr->ny, gr->nz};
  for (int bz = 0; bz < gr->nbz; bz++)
    for (int by = 0; by < gr->nby; by++)
      for (int bx = 0; bx < gr->nbx; bx++) {
        o_block *b = &gr->blocks[(bz * gr->nby + by) * gr->nbx + bx];
        b->mx = gr->nx + 2 * gr->g;
        b->my = gr->ny + 2 * gr->g;
        b->mz = gr->nz + 2 * gr->g;
        int ib[3] = {bx, by, bz};
        double lo[3], hi[3];
        block_extent(gr, ib, lo, hi);
        for (int a = 0; a < 3; a++) b->cw[a] = (hi[a] - lo[a]) / n[a];
        long ncc = (long)b->mx * b->my * b->mz;
        for (int q = 0; q < O_NCC; q++)
          b->cc[q] = (double *)calloc((size_t)ncc, sizeof(double));
        b->pressure = (double *)calloc((size_t)ncc, sizeof(double));
        b->eint = (double *)calloc((size_t)ncc, sizeof(double));
        b->bfi[0] = (double *)calloc((size_t)b->mz * b->my * (b->mx + 1), sizeof(double));
        b->bfi[1] = (double *)calloc((size_t)b->mz * (b->my + 1) * b->mx, sizeof(double));
        b->bfi[2] = (double *)calloc((size_t)(b->mz + 1) * b->my * b->mx, sizeof(double));
      }
  return 0;
}

void o_grid_free(o_grid *gr) {
  if (!gr->blocks) return;
  const int nb = gr->nbx * gr->nby * gr->nbz;
  for (int i = 0; i < nb; i++) {
    o_block *b = &gr->blocks[i];
    for (int q = 0; q < O_NCC; q++) free(b->cc[q]);
    free(b->pressure);
    free(b->eint);
    for (int d = 0; d < 3; d++) free(b->bfi[d]);
  }
  free(gr->blocks);
  gr->blocks = NULL;
}

/* Map a ghost index along one axis to (source block, source local index).
 * centred: 1 if the array is face-centred along this axis.  Returns 0 if the
 * index is interior (no copy along this axis needed).                      */
static int map_axis(const o_grid *gr, int axis, int blk, int idx, int centred,
                    int *sblk, int *sidx) {
  const int n = (axis == 0) ? gr->nx : ((axis == 1) ? gr->ny : gr->nz);
  const int nb = (axis == 0) ? gr->nbx : ((axis == 1) ? gr->nby : gr->nbz);
  const int g = gr->g;
  const int N = n * nb;
  if (!centred) {
    if (idx >= g && idx < g + n) { *sblk = blk; *sidx = idx; return 0; }
    int G = blk * n + idx - g;
    if (gr->bc[axis] == O_BC_PERIODIC) {
      G = ((G % N) + N) % N;
    } else { /* outflow: copy the nearest active cell
                (src/Enzo/enzo-core/EnzoBoundary.cpp:356-470) */
      G = (G < 0) ? 0 : ((G >= N) ? N - 1 : G);
    }
    *sblk = G / n;
    *sidx = G % n + g;
    return 1;
  }
  /* face-centred axis: faces [g, g+n] belong to this block */
  if (idx >= g && idx <= g + n) { *sblk = blk; *sidx = idx; return 0; }
  int F = blk * n + idx - g;
  if (gr->bc[axis] == O_BC_PERIODIC) {
    F = ((F % N) + N) % N;
    *sblk = F / n;
    *sidx = F % n + g;
  } else {
    if (F < 0) {
      *sblk = 0; *sidx = g;
    } else { /* F > N: upper boundary face of the last block */
      *sblk = nb - 1; *sidx = g + n;
    }
  }
  return 1;
}

static void refresh_array(const o_grid *gr, int which /* -1 cc, 0..2 face */,
                          int q /* cc slot, O_NCC = internal_energy */) {
  const int nbx = gr->nbx, nby = gr->nby, nbz = gr->nbz;
  const int ce[3] = {which == 0, which == 1, which == 2};
  for (int bz = 0; bz < nbz; bz++)
    for (int by = 0; by < nby; by++)
      for (int bx = 0; bx < nbx; bx++) {
        o_block *b = &gr->blocks[(bz * nby + by) * nbx + bx];
        const int ax = b->mx + ce[0], ay = b->my + ce[1], az = b->mz + ce[2];
        double *dst = (which < 0) ? (q == O_NCC ? b->eint : b->cc[q])
                                  : b->bfi[which];
        for (int k = 0; k < az; k++)
          for (int j = 0; j < ay; j++)
            for (int i = 0; i < ax; i++) {
              int sb[3], si[3], ghost = 0;
              ghost |= map_axis(gr, 0, bx, i, ce[0], &sb[0], &si[0]);
              ghost |= map_axis(gr, 1, by, j, ce[1], &sb[1], &si[1]);
              ghost |= map_axis(gr, 2, bz, k, ce[2], &sb[2], &si[2]);
              if (!ghost) continue;
              const o_block *s =
                  &gr->blocks[(sb[2] * nby + sb[1]) * nbx + sb[0]];
              const double *src = (which < 0)
                                       ? (q == O_NCC ? s->eint : s->cc[q])
                                       : s->bfi[which];
              dst[IDX(k, j, i, ay, ax)] = src[IDX(si[2], si[1], si[0], ay, ax)];
            }
      }
}

/* Refresh of every field (EnzoMethodMHDVlct.cpp:146-151 adds all fields;
 * internal_energy is one of them with dual energy). */
void o_grid_refresh(const o_config *cfg, o_grid *gr) {
  for (int q = 0; q < O_NCC; q++) {
    if (!cfg->mhd && q >= O_BX) continue;
    refresh_array(gr, -1, q);
  }
  if (cfg->dual_energy) refresh_array(gr, -1, O_NCC);
  if (cfg->mhd)
    for (int d = 0; d < 3; d++) refresh_array(gr, d, 0);
}

/* One cycle as in Cello: stopping phase (Block::stopping_begin_,
 * control_stopping.cpp:44-141: min of Method::timestep over blocks, clipped
 * to the stop time) on the current data, then the compute phase (refresh,
 * control_compute.cpp:42-68, followed by Method::compute on every block). */
double o_grid_cycle(const o_config *cfg, o_grid *gr, double t, double t_stop,
                    int nthreads) {
  const int nb = gr->nbx * gr->nby * gr->nbz;
  double dt = 1.7976931348623157e308;
  for (int i = 0; i < nb; i++) {
    double dtb = o_vlct_timestep(cfg, &gr->blocks[i]);
    /* ScheduleList::update_timestep for an output at t_stop
       (io_ScheduleList.cpp:133-159) */
    if (t < t_stop && t_stop < t + dtb) dtb = t_stop - t;
    double rem = t_stop - t;
    dtb = (dtb < rem) ? dtb : rem; /* MIN(dt_block, time_stop - time_curr) */
    dt = (dtb < dt) ? dtb : dt;    /* CkReduction::min_double */
  }
  dt *= 1.0; /* Method::courant_global */
  o_grid_refresh(cfg, gr);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) if (nb > 1)
#endif
  for (int i = 0; i < nb; i++) o_vlct_compute(cfg, &gr->blocks[i], dt);
  (void)nthreads;
  return dt;
}

int o_grid_evolve(const o_config *cfg, o_grid *gr, double t0, double t_stop,
                  int nthreads, double *dt_log, int dt_log_cap) {
  double t = t0;
  int cycles = 0;
  /* Stopping::complete (problem_Stopping.cpp:12-27): stop when !(t < t_stop) */
  while (t < t_stop) {
    double dt = o_grid_cycle(cfg, gr, t, t_stop, nthreads);
    if (dt_log && cycles < dt_log_cap) dt_log[cycles] = dt;
    t = t + dt; /* Block::compute_end_ (control_compute.cpp:146-148) */
    cycles++;
  }
  return cycles;
}

/* ------------------------------------------------------------------------ */
/* Inclined linear waves (EnzoInitialInclinedWave.cpp recipe)               */
/* ------------------------------------------------------------------------ */

typedef struct { double m[3][3]; } rotation;

static rotation make_rotation(double a, double b) {
  rotation r;
  r.m[0][0] = cos(a) * cos(b);
  r.m[0][1] = cos(a) * sin(b);
  r.m[0][2] = sin(a);
  r.m[1][0] = -1. * sin(b);
  r.m[1][1] = cos(b);
  r.m[1][2] = 0.;
  r.m[2][0] = -1. * sin(a) * cos(b);
  r.m[2][1] = -1. * sin(a) * sin(b);
  r.m[2][2] = cos(a);
  return r;
}
static void rot_apply(const rotation *r, double v0, double v1, double v2,
                      double *o0, double *o1, double *o2) {
  *o0 = r->m[0][0] * v0 + r->m[0][1] * v1 + r->m[0][2] * v2;
  *o1 = r->m[1][0] * v0 + r->m[1][1] * v1 + r->m[1][2] * v2;
  *o2 = r->m[2][0] * v0 + r->m[2][1] * v1 + r->m[2][2] * v2;
}
static void rot_inv(const rotation *r, double v0, double v1, double v2,
                    double *o0, double *o1, double *o2) {
  *o0 = r->m[0][0] * v0 + r->m[1][0] * v1 + r->m[2][0] * v2;
  *o1 = r->m[0][1] * v0 + r->m[1][1] * v1 + r->m[2][1] * v2;
  *o2 = r->m[0][2] * v0 + r->m[1][2] * v1 + r->m[2][2] * v2;
}

typedef struct {
  rotation rot;
  double amplitude, lambda;
  /* scalar/vector linear initialisers */
  double d_back, d_ev, e_back, e_ev;
  double m_back[3], m_ev[3];
  /* vector potential */
  int has_a;
  double b_back[3], b1_ev, b2_ev;
} wave_setup;

/* LinearScalarInit::operator() with use_cosine = true */
static double lin_scalar(const wave_setup *w, double back, double ev,
                         double x, double y, double z) {
  double x0, x1, x2;
  rot_apply(&w->rot, x, y, z, &x0, &x1, &x2);
  double tmp = x0 * 2. * CELLO_PI / w->lambda;
  return (back + w->amplitude * ev * (1 * cos(tmp) + (1 - 1) * sin(tmp)));
}

/* RotatedVectorInit(LinearVectorInit) */
static void lin_vector(const wave_setup *w, double x, double y, double z,
                       double *v0, double *v1, double *v2) {
  double x0, x1, x2, r0, r1, r2;
  rot_apply(&w->rot, x, y, z, &x0, &x1, &x2);
  double tmp = x0 * 2. * CELLO_PI / w->lambda;
  double trig = (1 * cos(tmp) + (1 - 1) * sin(tmp));
  r0 = w->m_back[0] + w->amplitude * w->m_ev[0] * trig;
  r1 = w->m_back[1] + w->amplitude * w->m_ev[1] * trig;
  r2 = w->m_back[2] + w->amplitude * w->m_ev[2] * trig;
  rot_inv(&w->rot, r0, r1, r2, v0, v1, v2);
}

/* RotatedVectorInit(LinearVectorPotentialInit) */
static void vec_potential(const wave_setup *w, double x, double y, double z,
                          double *a0, double *a1, double *a2) {
  double x0, x1, x2, r0, r1, r2;
  rot_apply(&w->rot, x, y, z, &x0, &x1, &x2);
  r0 = (x2 * w->amplitude * w->b1_ev * cos(2. * CELLO_PI * x0 / w->lambda) -
        x1 * w->amplitude * w->b2_ev * cos(2. * CELLO_PI * x0 / w->lambda));
  r1 = w->b_back[2] * x0;
  r2 = w->b_back[0] * x1 - w->b_back[1] * x0;
  rot_inv(&w->rot, r0, r1, r2, a0, a1, a2);
}

/* EnzoBfieldMethodCT::compute_center_bfield over the whole block */
static void center_bfield_all(o_block *b) {
  for (int d = 0; d < 3; d++)
    o_compute_center_bfield(d, b->cc[O_BX + d], b->bfi[d], b->mz, b->my,
                            b->mx, 0);
}

int o_init_inclined_wave(const o_config *cfg, o_grid *gr, const char *wave,
                         double alpha, double beta, double amplitude,
                         double lambda, int positive_vel) {
  wave_setup w;
  memset(&w, 0, sizeof(w));
  w.rot = make_rotation(alpha, beta);
  w.amplitude = amplitude;
  w.lambda = lambda;
  const double gamma = cfg->gamma;
  double wsign = positive_vel ? 1. : -1;

  int mhd_wave = (!strcmp(wave, "fast") || !strcmp(wave, "alfven") ||
                  !strcmp(wave, "slow") || !strcmp(wave, "mhd_entropy"));
  if (mhd_wave) {
    /* prepare_MHD_initializers_ (EnzoInitialInclinedWave.cpp:820-915) */
    if (!cfg->mhd) return -1;
    double density_back = 1, mom0_back = 0, mom1_back = 0, mom2_back = 0;
    double b0_back = 1., b1_back = 1.5, b2_back = 0.0;
    double etot_back = (1. / gamma) / (gamma - 1.) + 1.625;
    if (!strcmp(wave, "mhd_entropy")) {
      mom0_back = wsign;
      etot_back += 0.5;
    }
    double density_ev, mom0_ev, mom1_ev, mom2_ev, etot_ev, b1_ev, b2_ev;
    if (!strcmp(wave, "fast")) {
      double coef = 0.5 / sqrt(5.);
      density_ev = 2. * coef;
      mom0_ev = wsign * 4. * coef;
      mom1_ev = -1. * wsign * 2. * coef;
      mom2_ev = 0;
      etot_ev = 9. * coef;
      b1_ev = 4. * coef;
      b2_ev = 0;
    } else if (!strcmp(wave, "alfven")) {
      density_ev = 0; mom0_ev = 0; mom1_ev = 0;
      mom2_ev = -1. * wsign * 1;
      etot_ev = 0; b1_ev = 0; b2_ev = 1.;
    } else if (!strcmp(wave, "slow")) {
      double coef = 0.5 / sqrt(5.);
      density_ev = 4. * coef;
      mom0_ev = wsign * 2. * coef;
      mom1_ev = wsign * 4. * coef;
      mom2_ev = 0;
      etot_ev = 3. * coef;
      b1_ev = -2. * coef;
      b2_ev = 0;
    } else {
      double coef = 0.5;
      density_ev = 2. * coef;
      mom0_ev = 2. * coef * wsign;
      mom1_ev = 0; mom2_ev = 0;
      etot_ev = 1. * coef;
      b1_ev = 0; b2_ev = 0;
    }
    w.has_a = 1;
    w.b_back[0] = b0_back; w.b_back[1] = b1_back; w.b_back[2] = b2_back;
    w.b1_ev = b1_ev; w.b2_ev = b2_ev;
    w.d_back = density_back; w.d_ev = density_ev;
    w.e_back = etot_back; w.e_ev = etot_ev;
    w.m_back[0] = mom0_back; w.m_back[1] = mom1_back; w.m_back[2] = mom2_back;
    w.m_ev[0] = mom0_ev; w.m_ev[1] = mom1_ev; w.m_ev[2] = mom2_ev;
  } else {
    /* prepare_HD_initializers_ (EnzoInitialInclinedWave.cpp:919-1040) */
    double v0_back = 0, v1_back = 0, v2_back = 0;
    if (strcmp(wave, "sound") != 0) v0_back = wsign;
    double squared_v_back = v0_back * v0_back + v1_back * v1_back + v2_back * v2_back;
    double density_ev, mom0_ev, mom1_ev, mom2_ev, etot_ev;
    if (!strcmp(wave, "sound")) {
      double h_back = 1 / (gamma - 1.) + 0.5 * squared_v_back;
      double signed_cs = wsign * 1;
      density_ev = 1;
      mom0_ev = v0_back + signed_cs;
      mom1_ev = v1_back;
      mom2_ev = v2_back;
      etot_ev = h_back + v0_back * signed_cs;
    } else if (!strcmp(wave, "hd_entropy")) {
      density_ev = 1; mom0_ev = v0_back; mom1_ev = v1_back; mom2_ev = v2_back;
      etot_ev = 0.5 * squared_v_back;
    } else if (!strcmp(wave, "hd_transv_entropy_v1")) {
      density_ev = 0; mom0_ev = 0; mom1_ev = 1; mom2_ev = 0;
      etot_ev = v1_back;
    } else if (!strcmp(wave, "hd_transv_entropy_v2")) {
      density_ev = 0; mom0_ev = 0; mom1_ev = 0; mom2_ev = 1;
      etot_ev = v2_back;
    } else {
      return -2;
    }
    w.d_back = 1; w.d_ev = density_ev;
    w.e_back = ((1. / gamma) / (gamma - 1.) + 0.5 * squared_v_back);
    w.e_ev = etot_ev;
    w.m_back[0] = v0_back; w.m_back[1] = v1_back; w.m_back[2] = v2_back;
    w.m_ev[0] = mom0_ev; w.m_ev[1] = mom1_ev; w.m_ev[2] = mom2_ev;
  }

  const int g = gr->g;
  for (int bz = 0; bz < gr->nbz; bz++)
    for (int by = 0; by < gr->nby; by++)
      for (int bx = 0; bx < gr->nbx; bx++) {
        o_block *b = &gr->blocks[(bz * gr->nby + by) * gr->nbx + bx];
        int ib[3] = {bx, by, bz};
        double lo[3], hi[3];
        block_extent(gr, ib, lo, hi);
        const double dx = b->cw[0], dy = b->cw[1], dz = b->cw[2];
        const int mx = b->mx, my = b->my, mz = b->mz;
/* MeshPos (EnzoInitialInclinedWave.cpp:455-500) */
#define PX(i) (lo[0] + dx * (0.5 + (double)((i) - g)))
#define PY(j) (lo[1] + dy * (0.5 + (double)((j) - g)))
#define PZ(k) (lo[2] + dz * (0.5 + (double)((k) - g)))
#define FX(i) (lo[0] + dx * (double)((i) - g))
#define FY(j) (lo[1] + dy * (double)((j) - g))
#define FZ(k) (lo[2] + dz * (double)((k) - g))
        if (cfg->mhd) {
          if (w.has_a) {
            /* setup_bfield (EnzoInitialInclinedWave.cpp:506-565) */
            double *Ax = (double *)calloc((size_t)(mz + 1) * (my + 1) * mx, sizeof(double));
            double *Ay = (double *)calloc((size_t)(mz + 1) * my * (mx + 1), sizeof(double));
            double *Az = (double *)calloc((size_t)mz * (my + 1) * (mx + 1), sizeof(double));
            for (int k = 0; k < mz + 1; k++)
              for (int j = 0; j < my + 1; j++)
                for (int i = 0; i < mx + 1; i++) {
                  double a0, a1, a2;
                  if (i != mx) {
                    vec_potential(&w, PX(i), FY(j), FZ(k), &a0, &a1, &a2);
                    Ax[IDX(k, j, i, my + 1, mx)] = a0;
                  }
                  if (j != my) {
                    vec_potential(&w, FX(i), PY(j), FZ(k), &a0, &a1, &a2);
                    Ay[IDX(k, j, i, my, mx + 1)] = a1;
                  }
                  if (k != mz) {
                    vec_potential(&w, FX(i), FY(j), PZ(k), &a0, &a1, &a2);
                    Az[IDX(k, j, i, my + 1, mx + 1)] = a2;
                  }
                }
            /* EnzoInitialBCenter::initialize_bfield_interface /
               bfieldi_helper_ (EnzoInitialBCenter.cpp:43-127) */
            for (int k = 0; k < mz; k++)
              for (int j = 0; j < my; j++)
                for (int i = 0; i < mx + 1; i++) {
                  double akr = Az[IDX(k, j + 1, i, my + 1, mx + 1)];
                  double akl = Az[IDX(k, j, i, my + 1, mx + 1)];
                  double ajr = Ay[IDX(k + 1, j, i, my, mx + 1)];
                  double ajl = Ay[IDX(k, j, i, my, mx + 1)];
                  b->bfi[0][IDX(k, j, i, my, mx + 1)] =
                      ((akr - akl) / dy - (ajr - ajl) / dz);
                }
            for (int k = 0; k < mz; k++)
              for (int j = 0; j < my + 1; j++)
                for (int i = 0; i < mx; i++) {
                  double akr = Ax[IDX(k + 1, j, i, my + 1, mx)];
                  double akl = Ax[IDX(k, j, i, my + 1, mx)];
                  double ajr = Az[IDX(k, j, i + 1, my + 1, mx + 1)];
                  double ajl = Az[IDX(k, j, i, my + 1, mx + 1)];
                  b->bfi[1][IDX(k, j, i, my + 1, mx)] =
                      ((akr - akl) / dz - (ajr - ajl) / dx);
                }
            for (int k = 0; k < mz + 1; k++)
              for (int j = 0; j < my; j++)
                for (int i = 0; i < mx; i++) {
                  double akr = Ay[IDX(k, j, i + 1, my, mx + 1)];
                  double akl = Ay[IDX(k, j, i, my, mx + 1)];
                  double ajr = Ax[IDX(k, j + 1, i, my + 1, mx)];
                  double ajl = Ax[IDX(k, j, i, my + 1, mx)];
                  b->bfi[2][IDX(k, j, i, my, mx)] =
                      ((akr - akl) / dx - (ajr - ajl) / dy);
                }
            free(Ax); free(Ay); free(Az);
          }
          center_bfield_all(b);
        }
        /* setup_fluid_, conserved form (EnzoInitialInclinedWave.cpp:627-660) */
        for (int k = 0; k < mz; k++)
          for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
              double x = PX(i), y = PY(j), z = PZ(k);
              long c = IDX(k, j, i, my, mx);
              double rho = lin_scalar(&w, w.d_back, w.d_ev, x, y, z);
              b->cc[O_RHO][c] = rho;
              double px, py, pz;
              lin_vector(&w, x, y, z, &px, &py, &pz);
              b->cc[O_VX][c] = px / rho;
              b->cc[O_VY][c] = py / rho;
              b->cc[O_VZ][c] = pz / rho;
              double etot_dens = lin_scalar(&w, w.e_back, w.e_ev, x, y, z);
              b->cc[O_ETOT][c] = etot_dens / rho;
            }
        /* setup_eint_ (:495-538): with dual energy the blocks carry
           eint = etot - kinetic - magnetic / density */
        if (cfg->dual_energy) {
          long ncc = (long)mx * my * mz;
          for (long c = 0; c < ncc; c++) {
            double vx = b->cc[O_VX][c], vy = b->cc[O_VY][c], vz = b->cc[O_VZ][c];
            double kinetic = 0.5 * (vx * vx + vy * vy + vz * vz);
            double magnetic = 0.;
            if (cfg->mhd) {
              double bx = b->cc[O_BX][c], by = b->cc[O_BY][c], bz = b->cc[O_BZ][c];
              magnetic = 0.5 * (bx * bx + by * by + bz * bz) / b->cc[O_RHO][c];
            }
            b->eint[c] = b->cc[O_ETOT][c] - kinetic - magnetic;
          }
        }
#undef PX
#undef PY
#undef PZ
#undef FX
#undef FY
#undef FZ
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* RJ2a shock tube (EnzoInitialShockTube.cpp recipe)                        */
/* ------------------------------------------------------------------------ */

/* left / right states (EnzoInitialShockTube.cpp:36-86): density, pressure,
 * velocity_0..2 (0 along the tube), bfield_1, bfield_2; and bfield_0 */
static const double TUBE_L[2][7] = {
    {1.08, 0.95, 1.2, 0.01, 0.5, 1.0155412503859613, 0.5641895835477563},
    {1.0, 1.0, 0., 0., 0., 0., 0.}};
static const double TUBE_R[2][7] = {
    {1., 1.0, 0.0, 0.0, 0.0, 1.1283791670955126, 0.5641895835477563},
    {0.125, 0.1, 0., 0., 0., 0., 0.}};
static const double TUBE_B0[2] = {0.5641895835477563, 0.0};

int o_init_shock_tube(const o_config *cfg, o_grid *gr, int setup, int ax,
                      double axis_velocity_p, double trans_velocity_p,
                      int flipped) {
  if (setup < 0 || setup > 1) return -1;
  if (setup == 0 && !cfg->mhd) return -1;
  /* setup_maps_ (EnzoInitialShockTube.cpp:161-176): the flipped tube swaps
     the states and negates every velocity and bfield entry */
  double L[7], R[7];
  for (int q = 0; q < 7; q++) {
    const double f = (flipped && q >= 2) ? -1. : 1.;
    L[q] = f * (flipped ? TUBE_R[setup][q] : TUBE_L[setup][q]);
    R[q] = f * (flipped ? TUBE_L[setup][q] : TUBE_R[setup][q]);
  }
  const double flip = flipped ? -1. : 1.;
  const double b0 = flip * TUBE_B0[setup];
  const double axis_velocity = flip * axis_velocity_p;
  const double trans_velocity = flip * trans_velocity_p;
  const double gamma = cfg->gamma;
  const int ai = ax, aj = (ax + 1) % 3, ak = (ax + 2) % 3;
  const int g = gr->g;
  for (int bz = 0; bz < gr->nbz; bz++)
    for (int by = 0; by < gr->nby; by++)
      for (int bx = 0; bx < gr->nbx; bx++) {
        o_block *b = &gr->blocks[(bz * gr->nby + by) * gr->nbx + bx];
        int ib[3] = {bx, by, bz};
        double lo[3], hi[3];
        block_extent(gr, ib, lo, hi);
        const int m[3] = {b->mx, b->my, b->mz};
        /* prep_aligned_slices_ (EnzoInitialShockTube.cpp:318-362) */
        double di = b->cw[ai];
        int shock_ind = (int)ceil((0.5 - lo[ai]) / di - 0.5 + (double)g);
        if (shock_ind < 0) shock_ind = 0;
        if (shock_ind > m[ai]) shock_ind = m[ai];
        for (int side = 0; side < 2; side++) {
          const double *S = side ? R : L;
          int s0 = side ? shock_ind : 0, s1 = side ? m[ai] : shock_ind;
          if (s0 >= s1) continue;
          double velocity_0 = S[2] + axis_velocity,
                 velocity_1 = S[3] + trans_velocity, velocity_2 = S[4];
          double eint = (S[1] / ((gamma - 1.) * S[0]));
          double v2 = (velocity_0 * velocity_0 + velocity_1 * velocity_1 +
                       velocity_2 * velocity_2);
          double b2 = (b0 * b0 + S[5] * S[5] + S[6] * S[6]);
          double etot = (eint + 0.5 * (v2 + b2 / S[0]));
          /* cell-centred fields and transverse face fields */
          for (int f = 0; f < 3; f++) { /* f: 0 = cc, 1 = bfieldi_j, 2 = bfieldi_k */
            if (f > 0 && !cfg->mhd) continue;
            int ext[3] = {b->mx, b->my, b->mz};
            if (f == 1) ext[aj] += 1;
            if (f == 2) ext[ak] += 1;
            for (int k = 0; k < ext[2]; k++)
              for (int j = 0; j < ext[1]; j++)
                for (int i = 0; i < ext[0]; i++) {
                  int p[3] = {i, j, k};
                  if (p[ai] < s0 || p[ai] >= s1) continue;
                  long c = IDX(k, j, i, ext[1], ext[0]);
                  if (f == 0) {
                    b->cc[O_RHO][c] = S[0];
                    b->cc[O_VX + ai][c] = velocity_0;
                    b->cc[O_VX + aj][c] = velocity_1;
                    b->cc[O_VX + ak][c] = velocity_2;
                    /* (optionally) the specific internal energy (:255-261) */
                    if (cfg->dual_energy) b->eint[c] = eint;
                    b->cc[O_ETOT][c] = etot;
                  } else if (f == 1) {
                    b->bfi[aj][c] = S[5];
                  } else {
                    b->bfi[ak][c] = S[6];
                  }
                }
          }
        }
        if (cfg->mhd) {
          int ext[3] = {b->mx, b->my, b->mz};
          ext[ai] += 1;
          long n = (long)ext[0] * ext[1] * ext[2];
          for (long c = 0; c < n; c++) b->bfi[ai][c] = b0;
          center_bfield_all(b);
        }
      }
  return 0;
}

int o_init_shock_tube_rj2a(const o_config *cfg, o_grid *gr, int ax) {
  return o_init_shock_tube(cfg, gr, 0, ax, 0.0, 0.0, 0);
}

/* ------------------------------------------------------------------------ */
/* Gather active cells (what the reference's HDF5 output would hold)        */
/* ------------------------------------------------------------------------ */

void o_grid_gather(const o_config *cfg, const o_grid *gr, int field,
                   double *out) {
  const int g = gr->g;
  const int NX = gr->nx * gr->nbx, NY = gr->ny * gr->nby;
  for (int bz = 0; bz < gr->nbz; bz++)
    for (int by = 0; by < gr->nby; by++)
      for (int bx = 0; bx < gr->nbx; bx++) {
        const o_block *b = &gr->blocks[(bz * gr->nby + by) * gr->nbx + bx];
        double *p = NULL;
        if (field == 8) {
          /* derived "pressure" computed for output
             (EnzoComputePressure::compute, EnzoComputePressure.cpp:40-78) */
          p = (double *)malloc(sizeof(double) * (size_t)b->mx * b->my * b->mz);
          o_compute_pressure(cfg, b, p, 0);
        }
        const double *src =
            (field == 8) ? p : ((field == 9) ? b->eint : b->cc[field]);
        for (int k = 0; k < gr->nz; k++)
          for (int j = 0; j < gr->ny; j++)
            for (int i = 0; i < gr->nx; i++) {
              long G = IDX(bz * gr->nz + k, by * gr->ny + j, bx * gr->nx + i,
                           NY, NX);
              out[G] = src[IDX(k + g, j + g, i + g, b->my, b->mx)];
            }
        free(p);
      }
}
