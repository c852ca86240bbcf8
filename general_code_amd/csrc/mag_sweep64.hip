// mag_sweep64.hip -- the FP64 instantiations of the link-form half-sweep
// kernels (mag_sweep_link.h).  Compiled with -ffp-contract=off: no
// multiply-add is fused, so each site update is the sequence of IEEE
// operations written in the source -- the sequence the oracle's C
// restatement (also built without contraction) performs -- and the FP64
// link sweeps agree with it bit for bit.
#include "mag_sweep_link.h"

namespace mag {

template void launch_su2_half<double>(double*, double*, const Geom&, int, double, double*,
                                      const JobCtl*, Shape, hipStream_t,
                                     const HalfOpt&);
template void launch_su3_half<double>(double*, double*, const Geom&, int, double, double*,
                                      const JobCtl*, Shape, hipStream_t,
                                     const HalfOpt&);

void launch_half_batch_f64(int group, Shape sh, const BatchJob* jobs, int njobs,
                           int nblocks_total, int parity, hipStream_t st) {
  launch_half_batch_t<double>(group, sh, jobs, njobs, nblocks_total, parity, st, 0, HALF_PLAIN);
}

}  // namespace mag
