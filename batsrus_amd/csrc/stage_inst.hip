// stage_inst.hip -- instantiates the fused stage kernel for one block size,
// one face flux and one storage type.  Built once per combination, e.g.
//   hipcc -DBATS_NB=16 -DBATS_FLUX=hlld -DBATS_REAL=f64 -c stage_inst.hip
// so that the heavy instantiations compile in parallel.
#include "stage_kernel.hpp"

#if !defined(BATS_NB) || !defined(BATS_FLUX) || !defined(BATS_REAL)
#error "define BATS_NB (8, 16), BATS_FLUX (rusanov, linde, hlld), BATS_REAL (f64, f32)"
#endif

#define BATS_FLUX_rusanov FLUX_RUSANOV
#define BATS_FLUX_linde FLUX_LINDE
#define BATS_FLUX_hlld FLUX_HLLD
#define BATS_REAL_f64 double
#define BATS_REAL_f32 float
#define BATS_PASTE_(a, b) a##b
#define BATS_PASTE(a, b) BATS_PASTE_(a, b)
#define BATS_NAME_(nb, fl, ty) launch_stage_##nb##_##fl##_##ty
#define BATS_NAME(nb, fl, ty) BATS_NAME_(nb, fl, ty)

namespace bats {
hipError_t BATS_NAME(BATS_NB, BATS_FLUX, BATS_REAL)(const StageArgs &a,
                                                    int opt,
                                                    hipStream_t stream) {
  return launch_stage_kernel<BATS_NB, BATS_PASTE(BATS_FLUX_, BATS_FLUX),
                             BATS_PASTE(BATS_REAL_, BATS_REAL)>(a, opt,
                                                                stream);
}
#define BATS_FNAME_(nb, fl, ty) launch_fineflux_##nb##_##fl##_##ty
#define BATS_FNAME(nb, fl, ty) BATS_FNAME_(nb, fl, ty)
hipError_t BATS_FNAME(BATS_NB, BATS_FLUX, BATS_REAL)(const AmrArgs &a,
                                                     hipStream_t stream) {
  return launch_amr_fineflux<BATS_NB, BATS_PASTE(BATS_FLUX_, BATS_FLUX),
                             BATS_PASTE(BATS_REAL_, BATS_REAL)>(a, stream);
}
} // namespace bats
