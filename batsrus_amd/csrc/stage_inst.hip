// stage_inst.hip -- instantiates the fused stage kernel for one block size
// and one face flux.  Built once per (BATS_NB, BATS_FLUX) pair, e.g.
//   hipcc -DBATS_NB=16 -DBATS_FLUX=hlld -c stage_inst.hip
// so that the heavy instantiations compile in parallel.
#include "stage_kernel.hpp"

#if !defined(BATS_NB) || !defined(BATS_FLUX)
#error "define BATS_NB (8 or 16) and BATS_FLUX (rusanov, linde or hlld)"
#endif

#define BATS_CAT_(a, b, c) a##b##_##c
#define BATS_CAT(a, b, c) BATS_CAT_(a, b, c)
#define BATS_FLUX_rusanov FLUX_RUSANOV
#define BATS_FLUX_linde FLUX_LINDE
#define BATS_FLUX_hlld FLUX_HLLD
#define BATS_FLUX_CODE_(f) BATS_FLUX_##f
#define BATS_FLUX_CODE(f) BATS_FLUX_CODE_(f)

namespace bats {
hipError_t BATS_CAT(launch_stage_, BATS_NB, BATS_FLUX)(const StageArgs &a,
                                                       bool powell,
                                                       hipStream_t stream) {
  return launch_stage_kernel<BATS_NB, BATS_FLUX_CODE(BATS_FLUX)>(a, powell,
                                                                 stream);
}
} // namespace bats
