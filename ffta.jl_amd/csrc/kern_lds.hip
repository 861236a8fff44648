// kern_lds.hip -- one instantiation of the in-LDS line-pass kernel per
// (precision, workgroup size), so that the variants compile in parallel.
// Built once per variant with -DKRT=float|double -DKTAG=f32|f64
// -DKBLOCK=256|512|1024; exports lds_fn_<KTAG>_<KBLOCK>(variant), the kernel
// entry of each pass variant (plain, pre-twiddle, full; each with and
// without direct-DFT stages of prime radices >= 17).
#include "device.h"
#include "kernels.h"

#if !defined(KRT) || !defined(KTAG) || !defined(KBLOCK)
#error "build with -DKRT=<float|double> -DKTAG=<f32|f64> -DKBLOCK=<256|512|1024>"
#endif

#define FFTAHIP_CAT4_(a, b, c, d) a##b##c##d
#define FFTAHIP_CAT4(a, b, c, d) FFTAHIP_CAT4_(a, b, c, d)

namespace fftahip {

const void* FFTAHIP_CAT4(lds_fn_, KTAG, _, KBLOCK)(int variant) {
  switch (variant) {
    case MODE_PLAIN: return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_PLAIN, false>);
    case MODE_PRETW: return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_PRETW, false>);
    case MODE_FULL: return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_FULL, false>);
    case kBigVariant + MODE_PLAIN:
      return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_PLAIN, true>);
    case kBigVariant + MODE_PRETW:
      return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_PRETW, true>);
    case kBigVariant + MODE_FULL:
      return reinterpret_cast<const void*>(&line_pass_group_kernel<KRT, 16, KBLOCK, MODE_FULL, true>);
    default: return nullptr;
  }
}

}  // namespace fftahip
