/* enzoe_vlct_host.h -- C entry points of the C++ drop-in layer
 * (enzo-e_amd/host/enzoe_host.hpp), for bindings that cannot use C++:
 * they construct EnzoMethodMHDVlct from a parameter group and call its
 * Method interface on blocks whose fields are device memory.
 *
 * Construction errors behave as the reference's ERROR/ASSERT (message on
 * stderr, exit(1)): EnzoMethodMHDVlct::EnzoMethodMHDVlct,
 * src/Enzo/hydro-mhd/EnzoMethodMHDVlct.cpp:89-150.                          */
#ifndef ENZOE_VLCT_HOST_H
#define ENZOE_VLCT_HOST_H

#include "enzoe_vlct.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct enzoe_host_method_s *enzoe_host_method_t;

/* Sets the fluid properties every method reads (EnzoPhysicsFluidProps:
 * ideal-gas gamma and the density / pressure floors).                       */
void enzoe_host_set_fluid_props(double gamma, double density_floor,
                                double pressure_floor);

/* new EnzoMethodMHDVlct(ParameterGroup(parameters), false); `parameters`
 * holds "name = value" items separated by ';' or newlines with the names of
 * the reference's "Method:mhd_vlct" group (mhd_choice, riemann_solver,
 * reconstruct_method, theta_limiter, time_scheme, courant) plus
 * device_fused (true|false) and precision (double|mixed).                   */
enzoe_host_method_t enzoe_host_method_create(const char *parameters);
void enzoe_host_method_destroy(enzoe_host_method_t m);

/* Method::name(), Method::courant()                                         */
const char *enzoe_host_method_name(enzoe_host_method_t m);
double enzoe_host_method_courant(enzoe_host_method_t m);

/* EnzoMethodMHDVlct::compute(Block*) on each of n blocks with the given dt
 * (control_compute.cpp:102); the fused path advances them in one batched
 * launch sequence.  Returns the number of compute_done() calls made.        */
int enzoe_host_method_compute(enzoe_host_method_t m, const enzoe_block *blocks,
                              int n, double dt, enzoe_stream_t stream);

/* EnzoMethodMHDVlct::compute(Block*) on one block carrying `npassive`
 * passively advected scalars (densities, cell-centred, the block's shape;
 * the reference's "color" field group): names[i] and device arrays
 * passive[i].  Needs device_fused = false.  Returns compute_done() calls.  */
int enzoe_host_method_compute_passive(enzoe_host_method_t m,
                                      const enzoe_block *block, int npassive,
                                      const char *const *names,
                                      double *const *passive, double dt,
                                      enzoe_stream_t stream);

/* EnzoMethodMHDVlct::timestep(Block*) (control_stopping.cpp:67-69); writes
 * the block's pressure field.                                               */
double enzoe_host_method_timestep(enzoe_host_method_t m,
                                  const enzoe_block *block,
                                  enzoe_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ENZOE_VLCT_HOST_H */
