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

/* stream handle: a hipStream_t passed as an opaque pointer (NULL = default) */
typedef void *enzoe_stream_t;

/* Arithmetic used by a method instance. */
#define ENZOE_PREC_DOUBLE 0 /* everything in FP64 (reference parity)       */
#define ENZOE_PREC_MIXED 1  /* FP32 reconstruction + Riemann fluxes; FP64
                               flux divergence, CT EMF/face-B update and dt */

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
  int precision;                  /* ENZOE_PREC_DOUBLE | ENZOE_PREC_MIXED   */
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

/* EnzoMethodMHDVlct::compute(Block*) (EnzoMethodMHDVlct.cpp:356-500) on n
 * blocks with one timestep dt: both VL stages (dt/2 with nearest-neighbour
 * reconstruction, then dt with the configured reconstructor), constrained
 * transport and the conserved update.  Asynchronous on `stream`.           */
int enzoe_vlct_compute(enzoe_vlct_t m, const enzoe_block *blocks, int n,
                       double dt, enzoe_stream_t stream);

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
 * These mirror the reference's factories; each call runs on the device
 * arrays it is given and follows the reference's stale-depth regions.
 * Arrays are dense (z,y,x) of the cell-centred extents (mz,my,mx) unless
 * stated; a face array along `dim` has the cell-centred extents with the
 * `dim` extent reduced by one (index f = face between cells f and f+1).
 * Primitive / flux arrays are passed as 8 pointers in the order
 * density, velocity_x, velocity_y, velocity_z, pressure|total_energy,
 * bfield_x, bfield_y, bfield_z (hydro uses the first 5).                    */

/* EnzoReconstructor::reconstruct_interface for "nn", "plm"/"plm_enzo",
 * "plm_athena" (toolkit/EnzoReconstructorNN.cpp:14-49,
 * EnzoReconstructorPLM.hpp:166-248). nkeys = 5 (hydro) or 8 (MHD).          */
int enzoe_reconstruct_interface(const char *name, double theta_limiter,
                                double density_floor, double pressure_floor,
                                double *const prim[8], double *priml[8],
                                double *primr[8], int nkeys, int mz, int my,
                                int mx, int dim, int stale_depth,
                                enzoe_stream_t stream);

/* Staling rates of a reconstructor (EnzoReconstructor.hpp:86-94).           */
int enzoe_reconstructor_staling(const char *name, int *total_rate,
                                int *immediate_rate);

/* EnzoRiemann::construct_riemann + solve (riemann/EnzoRiemann.cpp:26-68,
 * EnzoRiemannImpl.hpp:265-338) for "hlld", "hlle" (MHD) or "hllc" (hydro). */
int enzoe_riemann_solve(const char *solver, int mhd, double gamma,
                        double *const priml[8], double *const primr[8],
                        double *flux[8], int mz, int my, int mx, int dim,
                        int stale_depth, enzoe_stream_t stream);

/* EnzoBfieldMethodCT::update_bfield (toolkit/EnzoBfieldMethodCT.cpp:
 * 617-687).  edge[i] has the cell-centred extents minus one along the two
 * axes other than i.                                                        */
int enzoe_ct_update_bfield(const double cell_width[3], int dim,
                           double *const edge[3], const double *bfieldi_cur,
                           double *bfieldi_out, double dt, int mz, int my,
                           int mx, int stale_depth, enzoe_stream_t stream);

/* EnzoBfieldMethodCT::compute_center_bfield (:702-728).                      */
int enzoe_ct_compute_center_bfield(int dim, double *bfieldc,
                                   const double *bfieldi, int mz, int my,
                                   int mx, int stale_depth,
                                   enzoe_stream_t stream);

/* ---- diagnostics ---------------------------------------------------------- */

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
