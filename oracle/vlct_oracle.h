/*
 * vlct_oracle.h -- CPU oracle for the Enzo-E VL+CT MHD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, double-precision restatement
 * of the reference algorithm (src/Enzo/hydro-mhd/EnzoMethodMHDVlct.cpp and the
 * toolkit/riemann files it drives), written operation by operation so that the
 * floating-point evaluation order follows the reference expressions.  It is
 * the checker for the HIP library, never the product: only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - end to end: the reference's golden L1 errors for the inclined MHD and HD
 *     linear waves (input/vlct/run_MHD_linear_wave_test.py:178-188,
 *     input/vlct/run_HD_linear_wave_test.py:70-80) and the RJ2a shock tube
 *     (input/vlct/run_MHD_shock_tube_test.py:82-87) at rtol 1e-13 /
 *     atol 7e-14 (input/vlct/testing_utils.py:275-292); the dual-energy
 *     Sod shock tube (input/vlct/run_dual_energy_shock_tube_test.py:52-90);
 *   - per function: closed forms, invariants and brute-force cases.
 *
 * Layout: every array is dense, row-major (z, y, x), x fastest, exactly as a
 * Cello field with ghost zones (src/Cello/view_CelloView.hpp:507-512).
 */
#ifndef VLCT_ORACLE_H
#define VLCT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* cell-centred integration quantities (Cello field names) */
enum {
  O_RHO = 0, O_VX, O_VY, O_VZ, O_ETOT, O_BX, O_BY, O_BZ, O_NCC
};

enum { O_HLLE = 0, O_HLLC = 1, O_HLLD = 2 };
enum { O_NN = 0, O_PLM = 1, O_PLM_ATHENA = 2 };

typedef struct {
  double gamma;
  double density_floor;
  double pressure_floor;
  double theta_limiter;
  int riemann;      /* O_HLLE / O_HLLC / O_HLLD                        */
  int recon[2];     /* reconstructor per stage                         */
  int nstages;      /* 2 = "vl", 1 = "euler"                            */
  int mhd;          /* 1 = constrained_transport, 0 = no_bfield         */
  double courant;
  int dual_energy;  /* 1 = "modern" dual-energy formalism
                       (EnzoDualEnergyConfig.hpp:40-47), 0 = disabled     */
  double eta;       /* its eta (Physics:fluid_props:dual_energy:eta)      */
} o_config;

/* One Cello block: ghost-extended fields owned by the caller. */
typedef struct {
  int mx, my, mz;          /* cell-centred extents including ghosts   */
  double cw[3];            /* cell widths along x, y, z               */
  double *cc[O_NCC];       /* (mz, my, mx) each                       */
  double *bfi[3];          /* bfieldi_x (mz,my,mx+1), _y (mz,my+1,mx),
                              _z (mz+1,my,mx)                          */
  double *pressure;        /* permanent "pressure" field (mz,my,mx)   */
  double *eint;            /* specific "internal_energy" (mz,my,mx);
                              read and written with dual energy only     */
} o_block;

/* ---- whole-method entry points ---------------------------------------- */

/* EnzoMethodMHDVlct::timestep (EnzoMethodMHDVlct.cpp:551-588): writes the
 * pressure field over the whole block and returns courant * dt_block.      */
double o_vlct_timestep(const o_config *cfg, o_block *b);

/* EnzoMethodMHDVlct::compute (EnzoMethodMHDVlct.cpp:356-500) for one block
 * and one timestep dt.  Returns 0 on success.                              */
int o_vlct_compute(const o_config *cfg, o_block *b, double dt);

/* Flux corrections at refinement boundaries (fluxcorr_oracle.c).
 *
 * One block's FluxData as a single array (src/Cello/data_FluxData.cpp:
 * 14-110): for each listed field i_f, axis 0..2, face 0..1, the block's face
 * fluxes then the neighbour's, each n_j*n_k values; the face array of axis 0
 * is (nz, ny) with y fastest, axis 1 (nz, nx), axis 2 (ny, nx)
 * (FaceFluxes::flux_array, data_FaceFluxes.cpp:114-126).  fields[] are cc
 * slots O_RHO..O_ETOT; nx, ny, nz are active cells.                        */
long o_flux_data_size(int nf, int nx, int ny, int nz);
long o_flux_data_offset(int i_f, int axis, int face, int neighbor, int nx,
                        int ny, int nz);
/* EnzoMethodMHDVlct::compute with store_fluxes_for_corrections (hydro only;
 * returns -2 for MHD as EnzoMethodMHDVlct.cpp:137-141 asserts).           */
int o_vlct_compute_fc(const o_config *cfg, o_block *b, double dt, int ghost,
                      int nf, const int *fields, double *flux_data);
/* EnzoMethodMHDVlct::save_fluxes_for_corrections_ (EnzoMethodMHDVlct.cpp:
 * 250-330) for one dim; flux[q] are dense face-trimmed flux arrays of dim
 * (shape (mz,my,mx) minus one along dim).                                  */
void o_save_fluxes_for_corrections(double *const *flux, int dim,
                                   double cell_width, double dt, int mx,
                                   int my, int mz, int ghost, int nf,
                                   const int *fields, double *flux_data);
/* FaceFluxes::coarsen (data_FaceFluxes.cpp:128-250), rank 3, cell-centred:
 * the face array `fine` of a block (n x n x n active, normal `axis`) into
 * `out` (same size), the quadrant of child (cx, cy, cz) filled.           */
void o_face_fluxes_coarsen(const double *fine, double *out, int axis, int nx,
                           int ny, int nz, int cx, int cy, int cz);
/* Block::refresh_load_flux_face_ (control_refresh.cpp:1458-1500) with
 * FluxData::sum_neighbor_fluxes (data_FluxData.hpp:136-145): the fine
 * block's fluxes on (axis, face) coarsened into the coarse block's
 * neighbour fluxes on (axis, 1 - face), for every field.                   */
void o_flux_send_to_coarse(const double *fine_fd, double *coarse_fd, int nf,
                           int nx, int ny, int nz, int axis, int face,
                           const int child[3]);
/* MethodFluxCorrect::flux_correct_ (problem_MethodFluxCorrect.cpp:425-560)
 * on one leaf block, rank 3: fields[i_f] are (mz,my,mx) arrays with ghost
 * depth g; correct[2*axis+face] marks faces whose neighbour is finer;
 * conservative[i_f] marks fields of "make_field_conservative".  Returns -1
 * when a conservative field is listed without density.                     */
int o_flux_correct(double *const *fields, int nf, int i_density,
                   const int *conservative, const double *flux_data, int nx,
                   int ny, int nz, int g, const int correct[6]);

/* Coarse/fine transfers (mesh_oracle.c).  ProlongLinear::apply_
 * (src/Cello/problem_ProlongLinear.cpp:85-370): region (of3, nf3) of the
 * fine array (dims mf3, x fastest) from region (oc3, nc3) of the coarse
 * array; nf3 = 2*nc3 (coarse ghosts available) or 2*(nc3-2).  Returns -1
 * for other sizes (the reference asserts).  RestrictLinear::apply_
 * (problem_RestrictLinear.cpp:60-198): the mean of the 2^rank children.    */
int o_prolong_linear(double *values_f, const int mf3[3], const int of3[3],
                     const int nf3[3], const double *values_c,
                     const int mc3[3], const int oc3[3], const int nc3[3],
                     int accumulate);
int o_restrict_linear(double *values_c, const int nd3_c[3], const int im3_c[3],
                      const int n3_c[3], const double *values_f,
                      const int nd3_f[3], const int im3_f[3],
                      const int n3_f[3], int accumulate);

/* EnzoProlong (src/Enzo/mesh/EnzoProlong.cpp:112-190 -> interpolate.F),
 * rank 3, refinement 2, accumulate = false, use_linear = false (the other
 * cases call ProlongLinear): method 0 "3A" (tscint3d), 1 "2A" (interp3d),
 * 2 "2B" (int_lin3d with iposflag = positive).  Region (of3, nf3) of the
 * fine array from the coarse region (oc3, nc3), nc3 = nf3/2 + 2 (one coarse
 * cell of stencil on each side).  Returns 0; -1 for sizes outside that
 * contract; -2 where interp3d returns with its error flag
 * (prolong_oracle.c).                                                       */
int o_enzo_prolong(int method, int positive, double *values_f, const int mf3[3],
                   const int of3[3], const int nf3[3], const double *values_c,
                   const int mc3[3], const int oc3[3], const int nc3[3]);

/* EnzoMethodMHDVlct::compute with acceleration fields (accel[0..2] =
 * acceleration_x..z, (mz,my,mx)): the gravity source of
 * EnzoSourceGravity::calculate_source (toolkit/EnzoSourceGravity.cpp:17-69),
 * added on the stage with index 1 (EnzoMHDIntegratorStageCommands.cpp:
 * 265-295).                                                                 */
int o_vlct_compute_accel(const o_config *cfg, o_block *b, double dt,
                         double *const *accel);

/* EnzoComputePressure::compute_pressure (EnzoComputePressure.cpp:82-198)
 * over [stale, m-stale) into p.                                            */
void o_compute_pressure(const o_config *cfg, const o_block *b, double *p,
                        int stale_depth);

/* ---- single-face Riemann solvers (LUT order in/out) -------------------- */
/* MHD LUT: density, bfield_i, bfield_j, bfield_k, velocity_i, velocity_j,
 * velocity_k, pressure|total_energy (EnzoRiemannImpl.hpp:40-56).
 * Hydro LUT: density, velocity_i..k, pressure|total_energy (:21-36).      */
void o_hlld_face(const double wl[8], const double wr[8], double gamma,
                 double flux[8]);
void o_hlle_mhd_face(const double wl[8], const double wr[8], double gamma,
                     double flux[8]);
void o_hllc_face(const double wl[5], const double wr[5], double gamma,
                 double flux[5]);

/* Dual-energy outputs of the single-face solvers: the internal-energy flux
 * of the passively advected specific internal energy
 * (enzo_riemann_utils::passive_eint_flux, EnzoRiemannUtils.hpp:224-249)
 * and the interface velocity vi_bar (EnzoRiemannHLLD.hpp:404-450,
 * EnzoRiemannHLL.hpp:322-344, EnzoRiemannHLLC.hpp:159-170):
 * de[0] = internal-energy flux, de[1] = vi_bar.                            */
void o_hlld_face_de(const double wl[8], const double wr[8], double gamma,
                    double flux[8], double de[2]);
void o_hlle_mhd_face_de(const double wl[8], const double wr[8], double gamma,
                        double flux[8], double de[2]);
void o_hllc_face_de(const double wl[5], const double wr[5], double gamma,
                    double flux[5], double de[2]);
/* EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync
 * (EnzoPhysicsFluidProps.cpp:158-290) over [s, m-s) of one block's
 * integration quantities (U: O_* slots, eint used with dual energy).       */
void o_floor_energy_and_sync(const o_config *cfg, double *const *U,
                             double *eint, int mz, int my, int mx, int s);
/* EnzoSourceInternalEnergy::calculate_source
 * (toolkit/EnzoSourceInternalEnergy.cpp:16-92): cells [s+1, m_d-s-1) along
 * dim, [s, m-s) across: deint(c) -= dt/dx * max(p(c), p_floor) *
 * (vbar(c) - vbar(c - e_dim)); vbar is face-trimmed along dim.             */
void o_internal_energy_source(const o_config *cfg, int dim, double dt,
                              double cw, const double *p, double *deint,
                              const double *vbar, int mz, int my, int mx,
                              int s);

/* slope limiters (EnzoReconstructorPLM.hpp:287-346) */
double o_limiter_enzo(double vm1, double v, double vp1, double theta);
double o_limiter_athena(double vm1, double v, double vp1);

/* ---- per-operation kernels (exposed for unit tests) -------------------- */
/* Each takes dense arrays; see vlct_oracle.c for the region semantics.     */
void o_reconstruct(int kind, double theta, const double *w, double *wl,
                   double *wr, int mz, int my, int mx, int dim,
                   int stale_depth, int use_floor, double floor_val);
void o_riemann_solve(const o_config *cfg, double *const *pl,
                     double *const *pr, double **flux, int mz, int my,
                     int mx, int dim, int stale_depth);
void o_update_bfield(const double cw[3], int dim, double *const *edge,
                     const double *bcur, double *bout, double dt, int mz,
                     int my, int mx, int stale_depth);
void o_compute_center_bfield(int dim, double *bc, const double *bi, int mz,
                             int my, int mx, int stale_depth);

/* single-operation restatements, flux/face arrays face-trimmed along dim;
 * flux and dU maps indexed by O_* slots (dU: density, vx, vy, vz, etot) */
void o_correct_reconstructed_bfield(const double *bfi, double *wl, double *wr,
                                    int mz, int my, int mx, int dim, int s);
void o_identify_upwind(const double *rho_flux, double *w, int mz, int my,
                       int mx, int dim, int s);
void o_compute_center_efield(int dim, double *E, double *const *U, int mz,
                             int my, int mx, int s);
void o_compute_edge_efield(int dim, const double *Ec, double *edge,
                           double *const *jflux, double *const *kflux,
                           double *const *weight, int mz, int my, int mx,
                           int s);
void o_accumulate_flux_component(int dim, double dt, double cw, double **flux,
                                 double **dU, int mz, int my, int mx, int s);
void o_update_quantities(const o_config *cfg, double *const *U0,
                         double *const *dU, double **Uo, int mz, int my,
                         int mx, int s);
void o_primitive_from_integration(const o_config *cfg, double *const *U,
                                  double **W, int mz, int my, int mx, int s);
/* the same with the dual-energy formalism: the internal-energy flux and
 * vi_bar of every face, the "internal_energy" key of the flux divergence
 * (eflux into dUe), its update as a specific quantity, pressure from eint */
void o_riemann_solve_de(const o_config *cfg, double *const *pl,
                        double *const *pr, double **flux, double *eflux,
                        double *vbar, int mz, int my, int mx, int dim, int s);
void o_accumulate_flux_component_de(int dim, double dt, double cw,
                                    double **flux, double **dU,
                                    const double *eflux, double *dUe, int mz,
                                    int my, int mx, int s);
void o_update_quantities_de(const o_config *cfg, double *const *U0,
                            const double *eint0, double *const *dU,
                            const double *dUe, double **Uo, double *eint_o,
                            int mz, int my, int mx, int s);
void o_primitive_from_integration_de(const o_config *cfg, double *const *U,
                                     const double *eint, double **W, int mz,
                                     int my, int mx, int s);


/* ---- multi-block harness ------------------------------------------------ */

enum { O_BC_PERIODIC = 0, O_BC_OUTFLOW = 1 };

/* A uniform grid of equally sized blocks covering [lower, upper].          */
typedef struct {
  int nbx, nby, nbz;       /* blocks per axis                           */
  int nx, ny, nz;          /* active cells per block                    */
  int g;                   /* ghost depth                               */
  double lower[3], upper[3];
  int bc[3];               /* boundary type per axis                    */
  o_block *blocks;         /* nbx*nby*nbz, x fastest                    */
} o_grid;

int  o_grid_alloc(o_grid *gr);
void o_grid_free(o_grid *gr);
/* Fill ghost zones of every block from its neighbours / boundaries.        */
void o_grid_refresh(const o_config *cfg, o_grid *gr);
/* One Cello cycle: refresh, dt (min over blocks, clipped to t_stop), compute
 * on every block.  Returns the dt taken.                                   */
double o_grid_cycle(const o_config *cfg, o_grid *gr, double t, double t_stop,
                    int nthreads);
/* Run from t0 to t_stop; returns number of cycles.                         */
int o_grid_evolve(const o_config *cfg, o_grid *gr, double t0, double t_stop,
                  int nthreads, double *dt_log, int dt_log_cap);

/* Initial conditions written by the harness (recipes of the reference's
 * initializers): see vlct_harness.c.                                       */
int o_init_inclined_wave(const o_config *cfg, o_grid *gr, const char *wave,
                         double alpha, double beta, double amplitude,
                         double lambda, int positive_vel);
int o_init_shock_tube_rj2a(const o_config *cfg, o_grid *gr, int aligned_ax);
/* setup 0 = "rj2a", 1 = "sod" with axis_velocity, transverse_velocity and
 * flip_initialize (EnzoInitialShockTube.cpp:92-292); the specific internal
 * energy is written when the blocks carry it (cfg->dual_energy).          */
int o_init_shock_tube(const o_config *cfg, o_grid *gr, int setup,
                      int aligned_ax, double axis_velocity,
                      double transverse_velocity, int flip);

/* Copy the active region of one field of the grid into a dense global array
 * (nz_tot, ny_tot, nx_tot).  field: 0..7 cc fields, 8 = derived pressure,
 * 9 = internal_energy.                                                     */
void o_grid_gather(const o_config *cfg, const o_grid *gr, int field,
                   double *out);

#ifdef __cplusplus
}
#endif
#endif
