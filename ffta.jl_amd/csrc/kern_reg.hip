// kern_reg.hip -- the register line-pass kernel of one precision (built with
// -DKRT=float|double -DKTAG=f32|f64); exports reg_fn_<KTAG>(mode).
#include "device.h"
#include "kernels.h"

#if !defined(KRT) || !defined(KTAG)
#error "build with -DKRT=<float|double> -DKTAG=<f32|f64>"
#endif

#define FFTAHIP_CAT2_(a, b) a##b
#define FFTAHIP_CAT2(a, b) FFTAHIP_CAT2_(a, b)

namespace fftahip {

const void* FFTAHIP_CAT2(reg_fn_, KTAG)(int mode) {
  if (mode == MODE_PLAIN) return reinterpret_cast<const void*>(&reg_pass_group_kernel<KRT, MODE_PLAIN>);
  if (mode == MODE_PRETW) return reinterpret_cast<const void*>(&reg_pass_group_kernel<KRT, MODE_PRETW>);
  return reinterpret_cast<const void*>(&reg_pass_group_kernel<KRT, MODE_FULL>);
}

}  // namespace fftahip
