/*
 * enzoe_vlct.h -- C ABI of the MI355X (gfx950) VL+CT MHD library.
 *
 * Drop-in boundary for Enzo-E's VL+CT method (EnzoMethodMHDVlct) and the
 * plugin objects it is built from.  Plain C types only: device pointers,
 * sizes and integer status codes; no exceptions cross this boundary.
 *
 * Array layout (identical to Cello fields, src/Cello/view_CelloView.hpp:
 * 507-512): every field is a dense row-major (z, y, x) array with x fastest,
 * ghost zones included.  Cell-centred fields have extents (mz, my, mx);
 * bfieldi_x / _y / _z are face-centred and hold one more entry along their
 * own axis (input/vlct/vlct.incl:14-31).  All pointers in an enzoe_block are
 * device pointers (hipMalloc'd or resident in HBM).
 *
 * Status codes: 0 on success, otherwise one of ENZOE_ERR_* (the C++ host
 * shim turns a nonzero status into the reference's ERROR behaviour,
 * src/Cello/error_Error.cpp:86).
 */
#ifndef ENZOE_VLCT_H
#define ENZOE_VLCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ENZOE_OK 0
#define ENZOE_ERR_CONFIG 1  /* invalid parameter / unknown solver name     */
#define ENZOE_ERR_ARG 2     /* null pointer, bad shape, ghost depth < 3    */
#define ENZOE_ERR_HIP 3     /* a HIP runtime call failed                   */
#define ENZOE_ERR_ALLOC 4   /* device allocation failed                    */
#define ENZOE_ERR_STATE 5   /* call out of order (CT state machine)        */
#define ENZOE_ERR_SIZE 6    /* input larger than the fixed workspace of
                               enzoe_vlct_compute_noalloc                  */

/* stream handle: a hipStream_t passed as an opaque pointer (NULL = default) */
typedef void *enzoe_stream_t;

/* Arithmetic used by a method instance. */
#define ENZOE_PREC_DOUBLE 0 /* everything in FP64 (reference parity)       */
#define ENZOE_PREC_MIXED 1  /* FP32 reconstruction + Riemann fluxes; FP64
                               flux divergence, CT EMF/face-B update and dt */
#define ENZOE_PREC_DOUBLE_FMA 2 /* FP64 with multiply-adds contracted into
                               fused multiply-adds in the stage kernels: not
                               bit-identical to the reference, fewer FP64
                               instructions; the CT terms and the CFL
                               expressions round as in ENZOE_PREC_DOUBLE  */

/* Parameters of the "Method:mhd_vlct" group (EnzoMethodMHDVlct.cpp:38-101)
 * plus the fluid properties the method reads (gamma, floors:
 * src/Enzo/fluid-props/EnzoPhysicsFluidProps.cpp).  A NULL string or a
 * negative number selects the reference default.                          */
typedef struct {
  const char *mhd_choice;         /* "constrained_transport" | "no_bfield" */
  const char *riemann_solver;     /* "hlld" (default) | "hlle" | "hllc"     */
  const char *reconstruct_method; /* "plm" (default) | "plm_enzo" |
                                     "plm_athena" | "nn"                    */
  const char *time_scheme;        /* "vl" (default) | "euler"               */
  double theta_limiter;           /* default 1.5, must be in [1, 2]         */
  double courant;                 /* default 0.3 ("vl") / 1.0 ("euler")     */
  double gamma;                   /* adiabatic index, > 1                   */
  double density_floor;           /* > 0                                    */
  double pressure_floor;          /* > 0                                    */
  int precision;                  /* ENZOE_PREC_DOUBLE | ENZOE_PREC_MIXED |
                                     ENZOE_PREC_DOUBLE_FMA                  */
  /* Physics:fluid_props:dual_energy (EnzoConfig.cpp:952-1040,
   * EnzoDualEnergyConfig.hpp:17-80): "disabled" (NULL) or "modern"; the
   * VL+CT method rejects "bryan95" (EnzoMHDIntegratorStageCommands.cpp:
   * 31-34).  With "modern" every block must carry internal_energy, and the
   * precision must be ENZOE_PREC_DOUBLE or ENZOE_PREC_DOUBLE_FMA.          */
  const char *dual_energy;
  double dual_energy_eta;         /* "modern" eta; negative = 0.001         */
} enzoe_vlct_params;

/* Device view of one block's permanent fields (a Cello Block's FieldData,
 * src/Cello/data_FieldData.cpp:319-395).  Hydro mode leaves the bfield
 * pointers NULL.                                                            */
typedef struct {
  int mx, my, mz;          /* cell-centred extents including ghost zones    */
  int ghost;               /* ghost depth on every side (>= 3)              */
  double cell_width[3];    /* EnzoBlock::CellWidth (EnzoBlock.cpp:238-245)  */
  double *density;
  double *velocity[3];
  double *total_energy;    /* specific total energy                         */
  double *bfield[3];       /* cell-centred B                                */
  double *bfieldi[3];      /* face-centred B: (mz,my,mx+1), (mz,my+1,mx),
                              (mz+1,my,mx)                                   */
  double *pressure;        /* permanent "pressure" field (may be NULL)      */
  double *internal_energy; /* specific internal energy: required with the
                              dual-energy formalism, ignored otherwise     */
} enzoe_block;

typedef struct enzoe_vlct_method *enzoe_vlct_t;

/* ---- EnzoMethodMHDVlct ------------------------------------------------- */

/* EnzoMethodMHDVlct::EnzoMethodMHDVlct(ParameterGroup, bool)
 * (src/Enzo/hydro-mhd/EnzoMethodMHDVlct.cpp:90-152).  Validates the names
 * exactly as parse_mhdchoice_pack_pair_ (:38-86),
 * EnzoRiemann::construct_riemann (riemann/EnzoRiemann.cpp:26-68) and
 * EnzoReconstructor::construct_reconstructor
 * (toolkit/EnzoReconstructor.cpp:14-44) do.                                 */
int enzoe_vlct_create(const enzoe_vlct_params *params, enzoe_vlct_t *out);
int enzoe_vlct_destroy(enzoe_vlct_t m);

/* Courant factor in use (Method::courant(), problem_Method.hpp:90-94).      */
double enzoe_vlct_courant(enzoe_vlct_t m);

/* EnzoMethodMHDVlct::timestep(Block*) (EnzoMethodMHDVlct.cpp:551-588):
 * writes the pressure field over the whole block (if non-NULL) and returns
 * courant * min over every cell (ghosts included) of dx_i/(|v_i| + c_f).
 * Blocks: the Cello min-reduction over blocks (control_stopping.cpp:94-106)
 * is applied across the n blocks.  Synchronises the stream.                 */
int enzoe_vlct_timestep(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                        enzoe_stream_t stream, double *dt_out);

/* EnzoMethodMHDVlct::compute (EnzoMethodMHDVlct.cpp:356-500) followed by
 * the next cycle's EnzoMethodMHDVlct::timestep (:551-588) on the blocks it
 * leaves -- Cello's compute, then the CFL reduction of the next cycle
 * (control_compute.cpp:102, control_stopping.cpp:67-69) with nothing in
 * between -- in fewer passes over memory: the final stage's cell update,
 * launched over the whole block, also writes the pressure field and takes
 * the CFL minimum of every cell (no separate CFL kernel runs).
 * *dt_next and every field equal what enzoe_vlct_compute and then
 * enzoe_vlct_timestep give, bit for bit; with the dual-energy formalism the
 * updated cells' energies are synchronised a second time there, as the
 * timestep does it in place.  Synchronises the stream.                      */
int enzoe_vlct_compute_timestep(enzoe_vlct_t m, const enzoe_block *blocks,
                                int n, double dt, enzoe_stream_t stream,
                                double *dt_next);

/* EnzoMethodMHDVlct::compute(Block*) (EnzoMethodMHDVlct.cpp:356-500) on n
 * blocks with one timestep dt: both VL stages (dt/2 with nearest-neighbour
 * reconstruction, then dt with the configured reconstructor), constrained
 * transport and the conserved update.  Asynchronous on `stream`.           */
int enzoe_vlct_compute(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                       double dt, enzoe_stream_t stream);

/* enzoe_vlct_compute that allocates nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it).  No counterpart in the reference.  The block descriptor
 * table and the stage scratch (half-step state, fluxes, EMFs) live in one
 * fixed device array of the library, enzoe_vlct_noalloc_bytes() bytes
 * (64 MiB) on every GPU it runs on, held until the process ends; blocks
 * are taken in chunks that fit it.  ENZOE_ERR_SIZE, before anything is
 * launched, when one block's scratch plus the descriptor table does not
 * fit (a 32^3 FP64 MHD block needs about 17 MB, a 64^3 one about 110 MB).
 * Every call on a GPU shares that GPU's array: calls that overlap on one
 * GPU (two streams, two threads) race.  Results equal enzoe_vlct_compute's
 * bit for bit.  Synchronises the stream (the descriptor upload).           */
int enzoe_vlct_compute_noalloc(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                               double dt, enzoe_stream_t stream);
size_t enzoe_vlct_noalloc_bytes(void);

/* As enzoe_vlct_compute with one dt per block (blocks of different
 * refinement levels advanced together under sub-cycling, each by its own
 * level's dt; the reference advances every block by the global dt,
 * control_stopping.cpp:94-141).  dt[n] is host memory, or device memory when
 * dt_on_device is nonzero (e.g. written by a previous kernel, so a whole
 * step can run without a host round trip).                                  */
int enzoe_vlct_compute_dt(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                          const double *dt, int dt_on_device,
                          enzoe_stream_t stream);

/* Launch-shape autotuner (no counterpart in the reference): times the
 * launch structures on a scratch copy of the leading blocks (at most ~2 GiB;
 * the caller's fields are not touched), advancing them by dt, and keeps the
 * fastest for this block shape; later compute calls on blocks of that shape
 * use it.  The candidates: every combination of XCD-contiguous workgroup
 * order over the four stage kernels (k_flux, k_edge, k_faces, k_cells)
 * and, with the launch-structure setting "auto" (enzoe_vlct_set_fused_ct),
 * for MHD the fused CT + update launch with each of its tile shapes and the
 * orders of its two kernels, and (HLLD MHD or hydro, no dual energy) the
 * persistent launch of a whole stage; then, for the chosen one, the rows of
 * the tiled cell order (slabs of 8, 16, 32 or 64 rows, or the plain
 * order), and the other shapes of the flux launch (enzoe_vlct_set_flux_shape).
 * *mask_out (optional) gets the chosen mask: bit 0 k_flux, 1 k_edge,
 * 2 k_faces, 3 k_cells (or the fused launch).  enzoe_vlct_launch_choice:
 * the mask, the launch structure (0 = four launches per stage, 1..5 as
 * enzoe_vlct_set_fused_ct), the tile rows and the flux-launch shape (0..2)
 * calls on blocks of this shape use.                                       */
int enzoe_vlct_autotune(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                        double dt, enzoe_stream_t stream, int *mask_out);
int enzoe_vlct_launch_choice(enzoe_vlct_t m, int mx, int my, int mz, int *mask,
                             int *fused_ct, int *tile_rows, int *flux_shape);

/* EnzoMethodMHDVlct::timestep + the clip of control_stopping.cpp:83-86
 * without a host round trip: reads the time from *d_time (device), writes
 * dt = min(courant * min_cfl, t_stop - t) into d_dts[0..ndts) (device) and
 * advances *d_time by dt.  For device-driven cycles (graph capture); a
 * cycle that starts at t_stop is zero-length and changes nothing: the
 * pressure field is not written and, with dual energy, the energies are
 * left unsynchronised.                                                      */
int enzoe_vlct_timestep_device(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                               double *d_time, double t_stop, double *d_dts,
                               int ndts, enzoe_stream_t stream);

/* Sub-cycled refinement levels advanced together on the device (no
 * counterpart in the reference, which takes one global dt): blocks[i] lies
 * on level level_of[i] (0 = coarsest; every level below the finest must
 * hold blocks).  One coarse step: the CFL minimum of every level (the
 * pressure field written, as enzoe_vlct_timestep), dt_0 = min_l 2^l
 * courant * min_l, then 2^(L-1) substeps; substep s advances the levels l
 * with s % 2^(L-1-l) == 0, each block with dt_0 / 2^l, as one batched
 * launch sequence per run of equal block shapes (each block equal bit for
 * bit to enzoe_vlct_compute_dt of the same list with those dt).  Every
 * table and buffer is prepared at create, so a step launches kernels only:
 * enzoe_vlct_levels_capture records `nsteps` steps into a HIP graph that
 * enzoe_vlct_levels_replay launches with no host work.  The blocks must
 * stay valid and unmoved; the ghost zones are not refreshed (each block is
 * advanced as a self-contained Method input).  levels_state: the device
 * time (the sum of the steps' dt_0) and the last dt_0; synchronises.      */
typedef struct enzoe_vlct_levels_s *enzoe_vlct_levels_t;
int enzoe_vlct_levels_create(enzoe_vlct_t m, const enzoe_block *blocks, const int *level_of,
                             int n, enzoe_vlct_levels_t *out);
int enzoe_vlct_levels_step(enzoe_vlct_levels_t levels, enzoe_stream_t stream);
int enzoe_vlct_levels_capture(enzoe_vlct_levels_t levels, int nsteps);
int enzoe_vlct_levels_replay(enzoe_vlct_levels_t levels, enzoe_stream_t stream);
int enzoe_vlct_levels_state(enzoe_vlct_levels_t levels, enzoe_stream_t stream, double *time,
                            double *dt0);
int enzoe_vlct_levels_destroy(enzoe_vlct_levels_t levels);

/* A HIP graph of `ncycles` Cello cycles of a uniform periodic/outflow grid
 * of blocks (device dt + refresh + compute), captured once and replayed
 * with no host work per cycle.  The time lives on the device, starts at t0
 * and stops at t_stop: cycles past t_stop are zero-length and change
 * nothing.  The block list and its buffers must stay valid; the method's
 * other calls must not run concurrently with a replay.                      */
typedef struct enzoe_vlct_graph_s *enzoe_vlct_graph_t;
int enzoe_vlct_graph_create(enzoe_vlct_t m, const enzoe_block *blocks, int nbx,
                            int nby, int nbz, const int bc[3], double t0,
                            double t_stop, int ncycles,
                            enzoe_vlct_graph_t *out);
int enzoe_vlct_graph_launch(enzoe_vlct_graph_t g, enzoe_stream_t stream);
/* Synchronises `stream` and reads the device time.                          */
int enzoe_vlct_graph_time(enzoe_vlct_graph_t g, enzoe_stream_t stream,
                          double *t);
int enzoe_vlct_graph_destroy(enzoe_vlct_graph_t g);

/* The same `ncycles` cycles enqueued on `stream` without a graph (no host
 * synchronisation either), with the CFL minimum of each next cycle taken
 * by the final cell update (as enzoe_vlct_compute_timestep) instead of a
 * separate pass: the same fields and dt as the graph's cycles, except
 * that the pressure field ends as the pressure of the final state (a CFL
 * pass over it would write the same).  d_state is n + 1 doubles of device
 * memory, [0] the time (read and advanced by each cycle), [1 .. n] the
 * per-block dt the cycles write.  On this platform a captured graph of
 * these short kernels replays more slowly than launches enqueued directly. */
int enzoe_vlct_run_cycles(enzoe_vlct_t m, const enzoe_block *blocks, int nbx, int nby,
                          int nbz, const int bc[3], double *d_state, double t_stop,
                          int ncycles, enzoe_stream_t stream);

/* EnzoMethodMHDVlct::compute on blocks with acceleration fields
 * (EnzoMethodMHDVlct.cpp:219-231, 410-414): the step of enzoe_vlct_compute
 * plus the gravity source of EnzoSourceGravity::calculate_source
 * (toolkit/EnzoSourceGravity.cpp:17-69) on the stage with index 1, from the
 * start-of-step density and velocity (EnzoMHDIntegratorStageCommands.cpp:
 * 265-295).  accel[3*i + a]: device acceleration_x/y/z of block i, cell
 * shaped with ghosts.                                                       */
int enzoe_vlct_compute_gravity(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                               double dt, double *const *accel,
                               enzoe_stream_t stream);

/* ---- ghost refresh for a uniform grid of blocks ------------------------- */

#define ENZOE_BC_PERIODIC 0
#define ENZOE_BC_OUTFLOW 1

/* Fills the ghost zones of every field of a uniform nbx x nby x nbz grid of
 * equally sized blocks (x fastest in `blocks`) from neighbouring blocks or
 * the domain boundary: Cello's refresh (src/Cello/control_refresh.cpp:
 * 43-357, data_FieldFace.cpp:214-440) with periodic boundaries or Enzo's
 * outflow boundary (src/Enzo/enzo-core/EnzoBoundary.cpp:287-470).         */
int enzoe_grid_refresh(enzoe_vlct_t m, const enzoe_block *blocks, int nbx,
                       int nby, int nbz, const int bc[3],
                       enzoe_stream_t stream);

/* ---- plugin objects used one call at a time -----------------------------
 * These mirror the reference's plugin methods; each call runs on device
 * views and follows the reference's stale-depth regions exactly.  A view is
 * the device analogue of CelloView<double,3> (src/Cello/view_CelloView.hpp:
 * 507-512): a pointer to element (0,0,0), a (z,y,x) shape and (z,y,x)
 * element strides, so trimmed views and slices of 4-D backing arrays are
 * passed without copies.  A map (EnzoEFltArrayMap = ViewMap<double>,
 * enzo_typedefs.hpp:36; view_ViewMap.hpp:43-395) is a set of same-shape,
 * same-stride views addressed by slot instead of by key:                    */
#define ENZOE_MAP_SLOTS 17
#define ENZOE_SLOT_DENSITY 0     /* "density"                               */
#define ENZOE_SLOT_VELOCITY_X 1  /* "velocity_x"                            */
#define ENZOE_SLOT_VELOCITY_Y 2  /* "velocity_y"                            */
#define ENZOE_SLOT_VELOCITY_Z 3  /* "velocity_z"                            */
#define ENZOE_SLOT_ENERGY 4      /* "total_energy" (integration, dU, flux)
                                    or "pressure" (primitive, priml/primr)   */
#define ENZOE_SLOT_BFIELD_X 5    /* "bfield_x"                              */
#define ENZOE_SLOT_BFIELD_Y 6    /* "bfield_y"                              */
#define ENZOE_SLOT_BFIELD_Z 7    /* "bfield_z"                              */
#define ENZOE_SLOT_PASSIVE0 8    /* passively advected scalars, slots 8..15:
                                    conserved form (densities) in integration,
                                    dU and flux maps, specific form (mass
                                    fractions) in primitive maps            */
#define ENZOE_PASSIVE_SLOTS 8    /* how many passive slots there are        */
#define ENZOE_SLOT_INTERNAL_ENERGY 16 /* "internal_energy" (dual energy):
                                    specific in integration maps, conserved
                                    change in dUcons, flux in flux maps     */

typedef struct enzoe_farray {
  double *data;           /* element (0,0,0) of the view                      */
  int shape[3];           /* (mz, my, mx)                                     */
  long long stride[3];    /* (z, y, x) strides in elements                    */
} enzoe_farray;

typedef struct enzoe_fmap {
  double *data[ENZOE_MAP_SLOTS]; /* NULL for a key the map does not hold     */
  int shape[3];                  /* shared (mz, my, mx)                      */
  long long stride[3];           /* shared (z, y, x) strides in elements     */
} enzoe_fmap;

/* Face-trimmed views along dim (reconstructed states, fluxes, CT weights)
 * have the cell-centred shape minus one along dim; entry f is the face
 * between cells f and f+1.  All functions return ENZOE_ERR_ARG when a view
 * has the wrong shape or a required slot is NULL, and launch nothing.       */

/* EnzoReconstructor::reconstruct_interface for "nn", "plm"/"plm_enzo",
 * "plm_athena" (toolkit/EnzoReconstructorNN.cpp:14-49,
 * toolkit/EnzoReconstructorPLM.hpp:166-248; names as
 * EnzoReconstructor::construct_reconstructor, EnzoReconstructor.cpp:14-44).
 * Reconstructs every slot present in prim; PLM floors density and pressure.*/
int enzoe_reconstruct_interface(const char *name, double theta_limiter,
                                double density_floor, double pressure_floor,
                                const enzoe_fmap *prim,
                                const enzoe_fmap *priml,
                                const enzoe_fmap *primr, int dim,
                                int stale_depth, enzoe_stream_t stream);

/* EnzoReconstructor::total_staling_rate / immediate_staling_rate
 * (EnzoReconstructor.hpp:86-94; NN 1/0, PLM 2/1).                            */
int enzoe_reconstructor_staling(const char *name, int *total_rate,
                                int *immediate_rate);

/* EnzoRiemann::construct_riemann + EnzoRiemannImpl::solve
 * (riemann/EnzoRiemann.cpp:26-68, riemann/EnzoRiemannImpl.hpp:265-338):
 * "hlld", "hlle" (MHD) or "hllc" (hydro) over [stale, n - stale) of the
 * flux view, then the upwinded flux of every passive slot present in both
 * states (enzo_riemann_utils::solve_passive_advection,
 * riemann/EnzoRiemannUtils.hpp:224-314).  Other names or combinations:
 * ENZOE_ERR_CONFIG.                                                         */
int enzoe_riemann_solve(const char *solver, int mhd, double gamma,
                        const enzoe_fmap *priml, const enzoe_fmap *primr,
                        const enzoe_fmap *flux, int dim, int stale_depth,
                        enzoe_stream_t stream);

/* EnzoBfieldMethodCT::correct_reconstructed_bfield
 * (toolkit/EnzoBfieldMethodCT.cpp:122-166): the interior faces of bfieldi
 * (full face-centred view, +1 along dim) into the bfield_dim slot of both
 * reconstructed states.                                                     */
int enzoe_ct_correct_reconstructed_bfield(int dim, const enzoe_farray *bfieldi,
                                          const enzoe_fmap *priml,
                                          const enzoe_fmap *primr,
                                          int stale_depth,
                                          enzoe_stream_t stream);

/* EnzoBfieldMethodCT::identify_upwind (:170-216): weight 1 / 0 / 0.5 from
 * the sign of the density flux; weight has the flux view's shape.           */
int enzoe_ct_identify_upwind(int dim, const enzoe_fmap *flux,
                             const enzoe_farray *weight, int stale_depth,
                             enzoe_stream_t stream);

/* EnzoBfieldMethodCT::compute_center_efield (:267-292):
 * E_dim = -v_j B_k + v_k B_j of the integration map.                        */
int enzoe_ct_compute_center_efield(int dim, const enzoe_farray *efield,
                                   const enzoe_fmap *integration,
                                   int stale_depth, enzoe_stream_t stream);

/* EnzoBfieldMethodCT::compute_edge_efield (:480-557): upwinded edge EMF
 * along dim from the centre E, the j-flux map (bfield_k slot) and the
 * k-flux map (bfield_j slot) and the weights of the three dims.  The edge
 * view has the cell shape minus one along j and k (:70-72).                 */
int enzoe_ct_compute_edge_efield(int dim, const enzoe_farray *center_efield,
                                 const enzoe_farray *edge_efield,
                                 const enzoe_fmap *jflux,
                                 const enzoe_fmap *kflux,
                                 const enzoe_farray weight[3],
                                 int stale_depth, enzoe_stream_t stream);

/* EnzoBfieldMethodCT::update_bfield (:617-687): out = cur - dt/dj dE_k +
 * dt/dk dE_j on the interior faces; bfieldi views are +1 along dim.         */
int enzoe_ct_update_bfield(const double cell_width[3], int dim,
                           const enzoe_farray edge_efield[3],
                           const enzoe_farray *cur_bfieldi,
                           const enzoe_farray *out_bfieldi, double dt,
                           int stale_depth, enzoe_stream_t stream);

/* EnzoBfieldMethodCT::compute_center_bfield (:702-728).                      */
int enzoe_ct_compute_center_bfield(int dim, const enzoe_farray *bfieldc,
                                   const enzoe_farray *bfieldi,
                                   int stale_depth, enzoe_stream_t stream);

/* EnzoIntegrationQuanUpdate::clear_dUcons_map
 * (toolkit/EnzoIntegrationQuanUpdate.cpp:83-101): every present slot over
 * the whole view set to value (a kernel, no memset).                        */
int enzoe_fill_map(const enzoe_fmap *map, double value,
                   enzoe_stream_t stream);

/* EnzoIntegrationQuanUpdate::accumulate_flux_component (:105-147):
 * dU(c) -= dt/dx (F(c+1/2) - F(c-1/2)) for every slot present in dUcons.   */
int enzoe_accumulate_flux_component(int dim, double dt, double cell_width,
                                    const enzoe_fmap *flux,
                                    const enzoe_fmap *dUcons, int stale_depth,
                                    enzoe_stream_t stream);

/* EnzoIntegrationQuanUpdate::update_quantities (:183-238) followed by
 * EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync
 * (EnzoPhysicsFluidProps.cpp:162-290) at stale_depth + 1.  initial/dUcons
 * need slots 0-4; out's bfield slots (MHD) must already hold the updated
 * cell-centred B (the reference runs CT first for this reason).  Passive
 * slots of dUcons (update_passive_scalars_, :242-268) need the same slots in
 * initial and out.                                                          */
int enzoe_update_quantities(const enzoe_fmap *initial,
                            const enzoe_fmap *dUcons, const enzoe_fmap *out,
                            int mhd, double gamma, double density_floor,
                            double pressure_floor, int stale_depth,
                            enzoe_stream_t stream);

/* EnzoPhysicsFluidProps::primitive_from_integration
 * (EnzoPhysicsFluidProps.cpp:64-138) with the pressure of
 * EnzoComputePressure::compute_pressure (EnzoComputePressure.cpp:177-189);
 * passive slots present in both maps become mass fractions (density / rho).*/
int enzoe_primitive_from_integration(const enzoe_fmap *integration,
                                     const enzoe_fmap *primitive, int mhd,
                                     double gamma, int stale_depth,
                                     enzoe_stream_t stream);

/* ---- the dual-energy formalism ("modern" formulation) on views ----------
 * EnzoRiemannImpl::solve with calculate_internal_energy_flux_
 * (riemann/EnzoRiemannImpl.hpp:240-301): enzoe_riemann_solve plus the
 * internal-energy flux into the flux map's internal_energy slot
 * (passive_eint_flux, EnzoRiemannUtils.hpp:224-249) and vi_bar into
 * interface_velocity (the flux view's shape), as EnzoRiemannHLLD.hpp:404-447,
 * EnzoRiemannHLL.hpp:322-344, EnzoRiemannHLLC.hpp:159-170 compute them.    */
int enzoe_riemann_solve_dual_energy(const char *solver, int mhd, double gamma,
                                    const enzoe_fmap *priml,
                                    const enzoe_fmap *primr,
                                    const enzoe_fmap *flux,
                                    const enzoe_farray *interface_velocity,
                                    int dim, int stale_depth,
                                    enzoe_stream_t stream);

/* EnzoSourceInternalEnergy::calculate_source
 * (toolkit/EnzoSourceInternalEnergy.cpp:16-92): cells with index along dim
 * in [stale+1, n-stale-1) and [stale, n-stale) across:
 * dUcons[internal_energy](c) -= dt/dx * max(p(c), pressure_floor) *
 * (vi_bar(c+1/2) - vi_bar(c-1/2)); p is the primitive map's pressure slot,
 * interface_velocity the face-trimmed view along dim.                       */
int enzoe_source_internal_energy(int dim, double dt, double cell_width,
                                 const enzoe_fmap *primitive,
                                 const enzoe_fmap *dUcons,
                                 const enzoe_farray *interface_velocity,
                                 double pressure_floor, int stale_depth,
                                 enzoe_stream_t stream);

/* enzoe_update_quantities with the internal energy as one more specific
 * quantity (EnzoIntegrationQuanUpdate.cpp:183-238; slot internal_energy of
 * initial, dUcons and out required) and the synchronisation of
 * apply_floor_to_energy_and_sync (EnzoPhysicsFluidProps.cpp:158-290) with
 * the given eta in place of the energy floor.                               */
int enzoe_update_quantities_dual_energy(const enzoe_fmap *initial,
                                        const enzoe_fmap *dUcons,
                                        const enzoe_fmap *out, int mhd,
                                        double gamma, double density_floor,
                                        double pressure_floor, double eta,
                                        int stale_depth, enzoe_stream_t stream);

/* primitive_from_integration / pressure_from_integration with dual energy:
 * p = (gamma-1) rho eint (EnzoComputePressure.cpp:127-134).                 */
int enzoe_primitive_from_integration_dual_energy(const enzoe_fmap *integration,
                                                 const enzoe_fmap *primitive,
                                                 int mhd, double gamma,
                                                 int stale_depth,
                                                 enzoe_stream_t stream);
int enzoe_pressure_from_integration_dual_energy(const enzoe_fmap *integration,
                                                const enzoe_farray *pressure,
                                                double gamma, int stale_depth,
                                                enzoe_stream_t stream);

/* EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync
 * (EnzoPhysicsFluidProps.cpp:158-290) over [stale, n-stale) of an
 * integration map, in place: the total-energy floor (dual_energy = 0) or the
 * synchronisation of internal and total energy with eta (dual_energy = 1,
 * internal_energy slot required).                                           */
int enzoe_apply_floor_to_energy_and_sync(const enzoe_fmap *integration,
                                         int mhd, double gamma,
                                         double pressure_floor, int dual_energy,
                                         double eta, int stale_depth,
                                         enzoe_stream_t stream);

/* EnzoMHDIntegratorStageCommands::timestep
 * (EnzoMHDIntegratorStageCommands.cpp:299-366): min over every cell of the
 * views of dx_i/(|v_i| + c_fast) (c_s for hydro) from the given pressure,
 * without the courant factor.  Synchronises `stream`.                       */
int enzoe_integrator_timestep(const enzoe_fmap *integration,
                              const enzoe_farray *pressure, int mhd,
                              double gamma, double dx, double dy, double dz,
                              enzoe_stream_t stream, double *dt_out);

/* EnzoSourceGravity::calculate_source (toolkit/EnzoSourceGravity.cpp:17-69)
 * over [stale, n - stale): dUcons momentum (velocity slots) += dt rho a,
 * dUcons energy += dt rho (v . a), with rho and v from `integration` and
 * acceleration_x..z in slots 1..3 of `accel` (the velocity slots).        */
int enzoe_source_gravity(double dt, const enzoe_fmap *integration,
                         const enzoe_fmap *dUcons, const enzoe_fmap *accel,
                         int stale_depth, enzoe_stream_t stream);

/* EnzoPhysicsFluidProps::pressure_from_integration (:142-156).               */
int enzoe_pressure_from_integration(const enzoe_fmap *integration,
                                    const enzoe_farray *pressure, int mhd,
                                    double gamma, int stale_depth,
                                    enzoe_stream_t stream);

/* ---- diagnostics ---------------------------------------------------------- */


/* ---- flux corrections at refinement boundaries (hydro) ------------------
 *
 * One block's FluxData is a single device array of doubles
 * (src/Cello/data_FluxData.cpp:14-110, allocate with single_array): for each
 * field i_f of the list, axis 0..2, face 0..1 (low, high), the block's face
 * fluxes followed by the neighbour's, each n_j*n_k values; the face array of
 * axis 0 is (nz, ny) with y fastest, of axis 1 (nz, nx), of axis 2 (ny, nx)
 * (FaceFluxes::flux_array, data_FaceFluxes.cpp:114-126).  nx, ny, nz are
 * the active cells of the block (even, as FaceFluxes requires,
 * data_FaceFluxes.cpp:38-45).  Field lists hold ENZOE_SLOT_* numbers in the
 * order of the "conserved" group, e.g. density, total_energy, velocity_x,
 * velocity_y, velocity_z (input/FluxCorrect/inclined_contact.incl).         */

/* Number of doubles of one block's FluxData (-1 for invalid sizes).         */
long long enzoe_flux_data_size(int nf, int nx, int ny, int nz);

/* EnzoMethodMHDVlct::compute with store_fluxes_for_corrections
 * (EnzoMethodMHDVlct.cpp:137-141, 333-363, 480-489): the step of
 * enzoe_vlct_compute, then, from the final stage's fluxes, dt/dx * F through
 * the six faces of each block's active region into the block parts of
 * flux_data[b] (a device pointer per block), with the neighbour parts set
 * to zero (the reference allocates the FluxData zeroed every cycle).
 * Hydro only: ENZOE_ERR_CONFIG for a constrained_transport method, as the
 * reference asserts.  fields: nf slots among density, velocity, energy.     */
int enzoe_vlct_compute_store_fluxes(enzoe_vlct_t m, const enzoe_block *blocks,
                                    int n, double dt, int nf,
                                    const int *fields,
                                    double *const *flux_data,
                                    enzoe_stream_t stream);

/* EnzoMethodMHDVlct::save_fluxes_for_corrections_ (EnzoMethodMHDVlct.cpp:
 * 250-330) for one dim, from a flux map of that dim (the cell shape minus
 * one along dim; entry f is the face between cells f and f+1): writes
 * dt/cell_width * F at the faces between cells ghost-1 and ghost, and
 * m-ghost-1 and m-ghost, into the block parts of flux_data.                 */
int enzoe_save_fluxes_for_corrections(const enzoe_fmap *flux, int dim,
                                      double cell_width, double dt, int ghost,
                                      int nf, const int *fields,
                                      double *flux_data,
                                      enzoe_stream_t stream);

/* Block::refresh_load_flux_face_ (src/Cello/control_refresh.cpp:1458-1500)
 * toward a coarser neighbour, with FaceFluxes::coarsen (rank 3, cell-
 * centred; data_FaceFluxes.cpp:128-250) and FluxData::sum_neighbor_fluxes
 * (data_FluxData.hpp:136-145): for each of the nf fields, the fine block's
 * fluxes on (axis, face) are coarsened into quadrant `child` (the fine
 * block's child index in its parent) and added to the coarse block's
 * neighbour fluxes on (axis, 1 - face).  Both blocks have nx*ny*nz active
 * cells.                                                                    */
int enzoe_flux_data_send_to_coarse(const double *fine_flux_data,
                                   double *coarse_flux_data, int nf, int nx,
                                   int ny, int nz, int axis, int face,
                                   const int child[3], enzoe_stream_t stream);

/* A fine block's face toward a coarser neighbour.                          */
typedef struct {
  const double *fine_flux_data;  /* the fine block's FluxData              */
  double *coarse_flux_data;      /* the coarse neighbour's FluxData        */
  int axis, face;                /* the fine block's face                  */
  int child[3];                  /* its child index in its parent          */
} enzoe_flux_link;

/* enzoe_flux_data_send_to_coarse for nlinks links in one launch, blocks of
 * nx*ny*nz active cells.  The links into one coarse face must be distinct
 * children; each writes only its own quadrant, concurrently (the one-link
 * form also adds the zeros of the other quadrants, as the reference does:
 * the results differ at most in the sign of a zero).  The link table is
 * copied through the method's device buffer: calls on one handle are
 * ordered by their stream.                                                 */
int enzoe_vlct_flux_send_to_coarse_batch(enzoe_vlct_t m,
                                         const enzoe_flux_link *links,
                                         int nlinks, int nf, int nx, int ny,
                                         int nz, enzoe_stream_t stream);

/* One leaf block of a batched flux correction.                              */
typedef struct {
  double *fields[16];            /* the nf fields (device, ghosts included) */
  const double *flux_data;       /* its FluxData                           */
  int correct[6];                /* [2*axis + face]: the neighbour is finer */
} enzoe_flux_correct_block;

/* MethodFluxCorrect::flux_correct_ on n blocks of one size in one launch
 * (arguments as enzoe_flux_correct).                                        */
int enzoe_vlct_flux_correct_blocks(enzoe_vlct_t m,
                                   const enzoe_flux_correct_block *blocks,
                                   int n, int nf, int i_density,
                                   const int *conservative, int nx, int ny,
                                   int nz, int ghost, enzoe_stream_t stream);

/* MethodFluxCorrect::flux_correct_ (src/Cello/problem_MethodFluxCorrect.cpp:
 * 425-560) on one leaf block: fields[i] is the device (mz, my, mx) array,
 * ghosts included, of field i of the FluxData list; i_density its density
 * entry (-1 if absent); conservative[i] marks the "make_field_conservative"
 * fields, corrected as rho * q and divided by the corrected density;
 * correct[2*axis + face] marks the faces whose neighbour is finer
 * (face_level > level).  ENZOE_ERR_ARG for a conservative field without
 * density (the reference asserts).                                          */
int enzoe_flux_correct(double *const *fields, int nf, int i_density,
                       const int *conservative, const double *flux_data,
                       int nx, int ny, int nz, int ghost, const int correct[6],
                       enzoe_stream_t stream);

/* ---- coarse/fine transfers (ghost zones across refinement levels) -------
 * Arrays are dense, x fastest, dims (x, y, z); rank 1 if dims[1] == 1,
 * 2 if dims[2] == 1, else 3.  FP64 (enzo_float).                           */

/* ProlongLinear::apply (src/Cello/problem_ProlongLinear.cpp:25-370): region
 * (of3, nf3) of the fine array values_f (dims mf3) from region (oc3, nc3)
 * of the coarse array values_c (dims mc3) by linear interpolation (weights
 * 1/4, 3/4 per axis), with nf3 = 2*(nc3-2) when the coarse region carries a
 * layer of ghost cells around the fine one, or nf3 = 2*nc3 (one-sided
 * weights 5/4, -1/4 at the ends); other sizes: ENZOE_ERR_ARG (the reference
 * asserts).  accumulate adds to values_f.  EnzoProlong with use_linear
 * (src/Enzo/mesh/EnzoProlong.cpp:49-97) is this call with the coarse region
 * shrunk by one cell on every side.                                         */
int enzoe_prolong_linear(double *values_f, const int mf3[3], const int of3[3],
                         const int nf3[3], const double *values_c,
                         const int mc3[3], const int oc3[3], const int nc3[3],
                         int accumulate, enzoe_stream_t stream);

/* EnzoProlong::apply with use_linear = false and accumulate = false
 * (src/Enzo/mesh/EnzoProlong.cpp:64-74, apply_ :112-190): Enzo's
 * interpolate() (src/Enzo/mesh/interpolate.F) with refinement 2, rank 3.
 * method 0 = "3A" (tscint3d: triangular-shaped-cloud weights, a
 * conservation correction, then the routine's monotonicity passes, which
 * for finite data leave every child at its parent's value), 1 = "2A"
 * (interp3d: min-limited slopes along the four cell diagonals from the
 * corner averages), 2 = "2B" (int_lin3d: corner-difference slopes;
 * positive = 1 limits them to 0.8 of the parent value).  Region (of3, nf3)
 * of the fine array from the coarse region (oc3, nc3) that holds the
 * parents and one cell of stencil on each side: nf3 even, nc3 = nf3/2 + 2
 * (ENZOE_ERR_ARG otherwise; the reference checks less and reads past the
 * region).  ierror (may be NULL; 2A only, which then synchronises the
 * stream): 1 where interp3d sets its error flag -- a NaN coarse value (no
 * child written, as the routine returns first) or a child of an
 * all-positive field that would not be positive (every child is written
 * here, where the routine stops at that row); EnzoProlong ignores the flag.
 * The other cases of EnzoProlong::apply are enzoe_prolong_linear.           */
int enzoe_enzo_prolong(int method, int positive, double *values_f, const int mf3[3],
                       const int of3[3], const int nf3[3], const double *values_c,
                       const int mc3[3], const int oc3[3], const int nc3[3], int *ierror,
                       enzoe_stream_t stream);

/* RestrictLinear::apply (src/Cello/problem_RestrictLinear.cpp:20-198):
 * region (im3_c, n3_c) of values_c (dims nd3_c) = mean of the 2^rank fine
 * children starting at im3_f of values_f (dims nd3_f); accumulate adds.    */
int enzoe_restrict_linear(double *values_c, const int nd3_c[3],
                          const int im3_c[3], const int n3_c[3],
                          const double *values_f, const int nd3_f[3],
                          const int im3_f[3], const int n3_f[3],
                          int accumulate, enzoe_stream_t stream);

/* One transfer of a batch: kind ENZOE_TRANSFER_PROLONG (arguments of
 * enzoe_prolong_linear: dst = values_f with m_dst/o_dst/n_dst = mf3/of3/nf3,
 * src = values_c with m_src/o_src/n_src = mc3/oc3/nc3) or
 * ENZOE_TRANSFER_RESTRICT (enzoe_restrict_linear: dst = values_c with
 * nd3_c/im3_c/n3_c, src = values_f with nd3_f/im3_f/n3_f).                 */
#define ENZOE_TRANSFER_PROLONG 0
#define ENZOE_TRANSFER_RESTRICT 1
typedef struct {
  int kind, accumulate;
  double *dst;
  const double *src;
  int m_dst[3], o_dst[3], n_dst[3];
  int m_src[3], o_src[3], n_src[3];
} enzoe_transfer;

/* n transfers in one launch (grid row = transfer), each validated as the
 * one-transfer entry points validate theirs.  The transfers of one call run
 * concurrently: their destination regions must not overlap each other, and
 * no transfer may read what another writes (order restrictions before the
 * prolongations that read them by making two calls).  The table is copied
 * through the method's device buffer (calls on one handle are ordered by
 * their stream).                                                            */
int enzoe_vlct_transfer_batch(enzoe_vlct_t m, const enzoe_transfer *items,
                              int n, enzoe_stream_t stream);

/* ---- debug checks and diagnostics ----------------------------------------
 * The reference's debug builds abort where a field value is not finite
 * (CHECK_FINITE_ELEMENTS, src/Cello/view_CelloView.hpp:207-217) or where a
 * floor has to be applied (RAISE_FLOOR_ERROR in enzo_utils::apply_floor,
 * src/Enzo/utils/utils.hpp:102-114).  With debug checks on, the method's
 * kernels count those events instead (one device atomic per wave and event;
 * off, they cost one uniform branch): non-finite updated cell values, the
 * density floor of the update (EnzoIntegrationQuanUpdate.cpp:219), the
 * energy floor (EnzoPhysicsFluidProps.cpp:278-285) and the reconstruction's
 * density/pressure floors (EnzoReconstructorPLM.hpp:219-220).  The C++ layer
 * turns a nonzero count into the reference's ERROR.                         */
/* Launch structure of the stages (no reference counterpart: the reference
 * runs every operation as its own loop).  variant 0: the flux launch, then
 * (MHD) three launches (edge EMFs, face update, cell update) through
 * memory; for MHD, 1, 2, 3: the constrained transport and the cell update
 * of a stage as one fused launch over tiles of 16x4x4, 32x4x2 or 8x8x4
 * cells, the edge EMFs and face-centred B of a tile held in LDS; 4: the
 * edge EMFs through memory, then the face update and the cell update as
 * one launch over 16x4x4 tiles (face B in LDS); 5: each stage as one
 * persistent launch whose resident workgroups run the flux, edge, face and
 * cell phases in turn with a device-wide barrier between them (HLLD MHD or
 * hydro without the dual-energy formalism; elsewhere, and on the
 * allocation-free entry point, the four launches); -1 (the default): the
 * autotuner's choice per block shape (the four launches until it has run).
 * Every variant gives the same bits in FP64 (the mixed and contracted
 * builds: up to the rounding of their contracted flux arithmetic, which the
 * compiler may fuse differently in different kernels).  Not to be changed
 * while work of the method is queued.
 * enzoe_vlct_check_persistent: synchronises and reports ENZOE_ERR_HIP if a
 * barrier wait of a persistent launch ran out (the workgroups of a launch
 * could not all be resident: results since the last check are not valid;
 * the barrier is reset), ENZOE_OK otherwise.                               */
int enzoe_vlct_set_fused_ct(enzoe_vlct_t m, int variant);
int enzoe_vlct_check_persistent(enzoe_vlct_t m);

/* The shape of the flux launch (no reference counterpart): 0, one thread
 * per cell solving its three faces; 1, one face per thread (three times
 * the workgroups, for blocks too small to fill the device); 2, the
 * nearest-neighbour stage as z pencils: 16 x 16 tiles of columns march
 * through the planes with each plane's primitive states staged in LDS, so
 * a cell's state is loaded and converted about 1.1 times instead of 4
 * (the reconstructing stage as 0); -1 (the default), the autotuner's
 * choice per block shape (0 until it has run).  Every shape gives the same
 * bits in FP64; the mixed and contracted builds fuse their flux
 * multiply-adds as the compiler finds them in each kernel, so there the
 * shapes agree to the rounding of that arithmetic.                         */
int enzoe_vlct_set_flux_shape(enzoe_vlct_t m, int shape);

#define ENZOE_DEBUG_NONFINITE 0
#define ENZOE_DEBUG_DENSITY_FLOOR 1
#define ENZOE_DEBUG_ENERGY_FLOOR 2
#define ENZOE_DEBUG_RECONSTRUCTION_FLOOR 3
#define ENZOE_DEBUG_COUNTS 4
int enzoe_vlct_set_debug_checks(enzoe_vlct_t m, int on);
/* Copies the counters accumulated since the last call (or since the checks
 * were switched on) into counts and clears them; synchronises the stream.
 * ENZOE_ERR_STATE while the checks are off.                                 */
int enzoe_vlct_debug_counts(enzoe_vlct_t m, enzoe_stream_t stream,
                            unsigned long long counts[ENZOE_DEBUG_COUNTS]);

/* Device-side summary of n blocks (no counterpart in the reference, whose
 * diagnostics read fields on the host): over the active cells the sums of
 * density, momentum and total energy density (rho e_tot), the smallest
 * density and pressure, the largest |div B| of the face field times the
 * smallest cell width and the largest |face B|; over every cell, ghosts
 * included, the number of non-finite field values.  Sums are compensated
 * per workgroup and added in a fixed order on the host (the same bits on
 * every run, within a few units in the last place of the exact sum).
 * Synchronises.                                                             */
typedef struct {
  double mass, momentum[3], energy;
  double min_density, min_pressure;
  double max_abs_divb, max_abs_bface;
  long long nonfinite;
} enzoe_diagnostics;
int enzoe_vlct_diagnostics(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                           enzoe_stream_t stream, enzoe_diagnostics *out);

/* Difference of two states of the same block list (a[i] against b[i], equal
 * shapes): over the active cells, per cell-centred field (density,
 * velocity_x..z, total_energy, bfield_x..z; the B slots are 0 for hydro)
 * the sums of |a - b| and of |b|, and the sums of the magnitudes |v_b| and
 * |B_b| of b's velocity and cell-centred B.  The L1 error of a field
 * against b is diff/ref, or, for a vector component, diff over the sum of
 * its vector's magnitude (a component that is zero in b has no scale of its
 * own).  Compensated per workgroup, added in a fixed order on the host.
 * max_rel: per field the largest per-cell relative difference |a - b| / |b|,
 * for a vector component over the cell's |b_vec| (0 where a equals b, NaN if
 * any difference is NaN).  Does not correspond to a reference function (the reference computes L1
 * norms in Python, tools/l1_error_norm.py:195-234).  Synchronises.          */
typedef struct {
  double diff[8], ref[8];
  double velocity_magnitude, bfield_magnitude;
  double max_rel[8];
} enzoe_state_l1;
int enzoe_vlct_state_l1(enzoe_vlct_t m, const enzoe_block *a, const enzoe_block *b,
                        int n, enzoe_stream_t stream, enzoe_state_l1 *out);

/* Per-kernel average duration (ms) of the launches made while profiling was
 * on (enzoe_set_profiling(1)), measured with HIP events recorded on the
 * launch stream around each kernel; the events are read when the report is
 * asked for, so profiling adds no host synchronisation.  names holds cap
 * strings of name_len bytes.  Returns the number of kernels reported.       */
int enzoe_set_profiling(int on);
int enzoe_profile_report(char *names, int name_len, double *ms, int *count,
                         int cap);

/* Plain device copy of n doubles (a kernel of known traffic, n*8 bytes read
 * and written), used to calibrate hardware byte counters.                   */
int enzoe_calibration_copy(double *dst, const double *src, size_t n,
                           enzoe_stream_t stream);

/* Library version string. */
const char *enzoe_version(void);

#ifdef __cplusplus
}
#endif
#endif
