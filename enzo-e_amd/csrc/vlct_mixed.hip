// vlct_mixed.hip -- the mixed-precision VL+CT step: FP32 reconstruction and
// Riemann fluxes (fast FP32 reciprocal and square root), FP64 primitives,
// flux divergence, CT EMF / face-B update and dt.  Built with
// -ffp-contract=fast: this path is checked against the FP64 result by its
// L1 error, not bit by bit.
#include <hip/hip_runtime.h>

#include "vlct_launch.hpp"

namespace enzoe {

int compute_mixed(enzoe_vlct_method *m, const enzoe_block *blocks, int n,
                  double dt, const double *dts, hipStream_t s,
                  const StoreArgs *st, const double *const *accel,
                  unsigned long long *cfl) {
  return m->mhd ? compute_group<float, true>(m, blocks, n, dt, dts, s, st, accel, cfl)
                : compute_group<float, false>(m, blocks, n, dt, dts, s, st, accel, cfl);
}

int launch_prepared_mixed(enzoe_vlct_method *m, const enzoe_block *blocks, int n, int chunk,
                          const void *table, const double *dts, hipStream_t s) {
  return m->mhd ? launch_chunks<float, true, false>(m, blocks, n, chunk,
                                                    (const BlockDesc<float> *)table, 0.0, dts,
                                                    s, nullptr, nullptr, true)
                : launch_chunks<float, false, false>(m, blocks, n, chunk,
                                                     (const BlockDesc<float> *)table, 0.0, dts,
                                                     s, nullptr, nullptr, true);
}

}  // namespace enzoe
