// kern_lds.hip -- one instantiation of the in-LDS line-pass kernel per
// (precision, workgroup size), so that the variants compile in parallel.
// Built once per variant with -DKRT=float|double -DKTAG=f32|f64
// -DKBLOCK=256|512|1024; exports lds_fn_<KTAG>_<KBLOCK>(mode), the kernel
// entry of each pass mode (plain, pre-twiddle, full).
#include "device.h"
#include "kernels.h"

#if !defined(KRT) || !defined(KTAG) || !defined(KBLOCK)
#error "build with -DKRT=<float|double> -DKTAG=<f32|f64> -DKBLOCK=<256|512|1024>"
#endif

#define FFTAHIP_CAT4_(a, b, c, d) a##b##c##d
#define FFTAHIP_CAT4(a, b, c, d) FFTAHIP_CAT4_(a, b, c, d)

namespace fftahip {

const void* FFTAHIP_CAT4(lds_fn_, KTAG, _, KBLOCK)(int mode) {
  if (mode == MODE_PLAIN) return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_PLAIN>);
  if (mode == MODE_PRETW) return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_PRETW>);
  return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_FULL>);
}

}  // namespace fftahip
