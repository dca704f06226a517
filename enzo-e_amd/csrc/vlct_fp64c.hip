// vlct_fp64c.hip -- the FP64 VL+CT step with multiply-adds contracted into
// fused multiply-adds (ENZOE_PREC_DOUBLE_FMA).  The same kernel templates as
// the reference-parity build, instantiated in their own namespace and built
// with -ffp-contract=fast-honor-pragmas: every product feeding a sum rounds
// once instead of twice, so the flux kernels issue fewer FP64 instructions.
// The result is no longer bit-identical to the reference's CPU build; it is
// checked against it per cell (relative difference), not bit by bit.  The
// CT terms and the CFL expressions keep their uncontracted form (their
// pragmas), so div(B) and the dt reduction round as in the FP64 build.
#include <hip/hip_runtime.h>

#define ENZOE_KNS enzoe_fma
#include "vlct_launch.hpp"

namespace enzoe {

int compute_contracted(enzoe_vlct_method *m, const enzoe_block *blocks, int n, double dt,
                       const double *dts, hipStream_t s, const void *st,
                       const double *const *accel, unsigned long long *cfl) {
  using namespace enzoe_fma;
  const StoreArgs *rs = (const StoreArgs *)st;
  if (m->mhd)
    return m->dual ? compute_group<double, true, true>(m, blocks, n, dt, dts, s, rs, accel, cfl)
                   : compute_group<double, true>(m, blocks, n, dt, dts, s, rs, accel, cfl);
  return m->dual ? compute_group<double, false, true>(m, blocks, n, dt, dts, s, rs, accel, cfl)
                 : compute_group<double, false>(m, blocks, n, dt, dts, s, rs, accel, cfl);
}

int launch_prepared_contracted(enzoe_vlct_method *m, const enzoe_block *blocks, int n,
                               int chunk, const void *table, const double *dts,
                               hipStream_t s) {
  using namespace enzoe_fma;
  const BlockDesc<double> *t = (const BlockDesc<double> *)table;
  if (m->mhd)
    return m->dual ? launch_chunks<double, true, true>(m, blocks, n, chunk, t, 0.0, dts, s,
                                                       nullptr, nullptr, true)
                   : launch_chunks<double, true, false>(m, blocks, n, chunk, t, 0.0, dts, s,
                                                        nullptr, nullptr, true);
  return m->dual ? launch_chunks<double, false, true>(m, blocks, n, chunk, t, 0.0, dts, s,
                                                      nullptr, nullptr, true)
                 : launch_chunks<double, false, false>(m, blocks, n, chunk, t, 0.0, dts, s,
                                                       nullptr, nullptr, true);
}

}  // namespace enzoe
