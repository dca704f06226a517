/*
 * sanitize_main.c -- drives the oracle (test infrastructure) through every
 * entry point on small grids, for a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (Makefile.sanitize).  The run is clean when the
 * sanitizers report nothing and every result is finite; it exits non-zero
 * otherwise.  Not part of the product.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vlct_oracle.h"

static int failures = 0;

static void expect(int ok, const char *what) {
  if (!ok) {
    fprintf(stderr, "FAILED: %s\n", what);
    failures++;
  }
}

static int all_finite(const o_grid *gr, int mhd) {
  const int nb = gr->nbx * gr->nby * gr->nbz;
  for (int b = 0; b < nb; b++) {
    const o_block *k = &gr->blocks[b];
    const long n = (long)k->mx * k->my * k->mz;
    for (int q = 0; q < (mhd ? O_NCC : 5); q++)
      for (long i = 0; i < n; i++)
        if (!isfinite(k->cc[q][i])) return 0;
  }
  return 1;
}

static o_config config(int riemann, int recon1, int mhd, int dual) {
  o_config c;
  memset(&c, 0, sizeof(c));
  c.gamma = 5.0 / 3.0;
  c.density_floor = 1e-200;
  c.pressure_floor = 1e-200;
  c.theta_limiter = 2.0;
  c.riemann = riemann;
  c.recon[0] = O_NN;
  c.recon[1] = recon1;
  c.nstages = 2;
  c.mhd = mhd;
  c.courant = 0.4;
  c.dual_energy = dual;
  c.eta = dual ? 0.001 : 0.0;
  return c;
}

static void run_grid(const char *name, o_config c, const char *wave, int setup) {
  o_grid gr;
  memset(&gr, 0, sizeof(gr));
  gr.nbx = 2; gr.nby = 1; gr.nbz = 1;
  gr.nx = 8; gr.ny = 10; gr.nz = 6;
  gr.g = 3;
  gr.upper[0] = 3.0; gr.upper[1] = 1.5; gr.upper[2] = 1.5;
  if (o_grid_alloc(&gr) != 0) {
    expect(0, "o_grid_alloc");
    return;
  }
  int rc = wave ? o_init_inclined_wave(&c, &gr, wave, 0.7, 1.1, 1e-3, 1.0, 1)
                : o_init_shock_tube(&c, &gr, setup, 0, 0.0, 0.0, 0);
  expect(rc == 0, name);
  o_grid_refresh(&c, &gr);
  double t = 0.0;
  for (int i = 0; i < 3; i++) t += o_grid_cycle(&c, &gr, t, 10.0, 2);
  expect(t > 0.0 && all_finite(&gr, c.mhd), name);
  double out[8 * 10 * 6 * 2];
  o_grid_gather(&c, &gr, 8, out);
  expect(isfinite(out[0]), name);
  o_grid_free(&gr);
}

static void faces(void) {
  double wl[8] = {1.0, 0.7, 0.3, -0.2, 0.4, 0.1, 0.0, 0.9};
  double wr[8] = {0.5, 0.7, -0.4, 0.1, -0.3, 0.0, 0.2, 0.4};
  double f[8], vb[2];
  o_hlld_face(wl, wr, 5.0 / 3.0, f);
  o_hlle_mhd_face(wl, wr, 5.0 / 3.0, f);
  o_hlld_face_de(wl, wr, 5.0 / 3.0, f, vb);
  o_hlle_mhd_face_de(wl, wr, 5.0 / 3.0, f, vb);
  double hl[5] = {1.0, 0.4, 0.1, 0.0, 0.9}, hr[5] = {0.5, -0.3, 0.0, 0.2, 0.4};
  o_hllc_face(hl, hr, 1.4, f);
  o_hllc_face_de(hl, hr, 1.4, f, vb);
  expect(isfinite(f[0]) && isfinite(vb[0]), "faces");
  expect(isfinite(o_limiter_enzo(0.1, 0.3, 0.2, 1.5)) &&
             isfinite(o_limiter_athena(0.1, 0.3, 0.2)),
         "limiters");
}

static void flux_corrections(void) {
  const int n = 4, g = 3, m = n + 2 * g, nf = 5;
  const int fields[5] = {O_RHO, O_VX, O_VY, O_VZ, O_ETOT};
  o_config c = config(O_HLLC, O_PLM, 0, 0);
  o_block b;
  memset(&b, 0, sizeof(b));
  b.mx = b.my = b.mz = m;
  b.cw[0] = b.cw[1] = b.cw[2] = 0.25;
  const long cells = (long)m * m * m;
  double *mem = calloc((size_t)cells * 7, sizeof(double));
  for (int q = 0; q < 5; q++) b.cc[q] = mem + q * cells;
  b.pressure = mem + 5 * cells;
  b.eint = mem + 6 * cells;
  for (long i = 0; i < cells; i++) {
    b.cc[O_RHO][i] = 1.0 + 0.1 * sin(0.3 * i);
    b.cc[O_VX][i] = 0.2 * cos(0.2 * i);
    b.cc[O_ETOT][i] = 2.0;
  }
  const long fsz = o_flux_data_size(nf, n, n, n);
  expect(fsz > 0 && o_flux_data_offset(nf - 1, 2, 1, 1, n, n, n) < fsz, "flux data layout");
  double *fd = calloc((size_t)fsz, sizeof(double));
  double *coarse = calloc((size_t)fsz, sizeof(double));
  o_vlct_timestep(&c, &b);
  expect(o_vlct_compute_fc(&c, &b, 0.01, g, nf, fields, fd) == 0, "compute_fc");
  const int child[3] = {1, 0, 1};
  o_flux_send_to_coarse(fd, coarse, nf, n, n, n, 0, 1, child);
  double *flds[5];
  for (int q = 0; q < 5; q++) flds[q] = b.cc[q];
  const int cons[5] = {0, 1, 1, 1, 1}, corr[6] = {1, 0, 0, 1, 1, 0};
  expect(o_flux_correct(flds, nf, 0, cons, coarse, n, n, n, g, corr) == 0, "flux_correct");
  free(fd);
  free(coarse);
  free(mem);
}

static void transfers(void) {
  const int mc[3] = {8, 8, 8}, mf[3] = {12, 12, 12};
  double *vc = malloc(sizeof(double) * 8 * 8 * 8), *vf = calloc(12 * 12 * 12, sizeof(double));
  for (int i = 0; i < 512; i++) vc[i] = 0.01 * i;
  const int oc[3] = {1, 1, 1}, nc[3] = {6, 6, 6}, of[3] = {0, 0, 0}, nf3[3] = {8, 8, 8};
  expect(o_prolong_linear(vf, mf, of, nf3, vc, mc, oc, nc, 0) == 0, "prolong");
  const int ic[3] = {0, 0, 0}, n3c[3] = {4, 4, 4}, iff[3] = {0, 0, 0}, n3f[3] = {8, 8, 8};
  expect(o_restrict_linear(vc, mc, ic, n3c, vf, mf, iff, n3f, 0) == 0, "restrict");
  free(vc);
  free(vf);
}

int main(void) {
  faces();
  run_grid("mhd hlld plm", config(O_HLLD, O_PLM, 1, 0), "fast", 0);
  run_grid("mhd hlle plm_athena", config(O_HLLE, O_PLM_ATHENA, 1, 0), "alfven", 0);
  run_grid("mhd hlld nn", config(O_HLLD, O_NN, 1, 0), "slow", 0);
  run_grid("hd hllc plm", config(O_HLLC, O_PLM, 0, 0), "sound", 0);
  run_grid("mhd hlld dual energy rj2a", config(O_HLLD, O_PLM, 1, 1), NULL, 0);
  run_grid("hd hllc dual energy sod", config(O_HLLC, O_PLM, 0, 1), NULL, 1);
  flux_corrections();
  transfers();
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("oracle sanitizer run clean\n");
  return 0;
}
