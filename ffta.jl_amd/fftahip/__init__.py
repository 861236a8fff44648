"""fftahip -- FFTA.jl's complex fft/bfft plan interface on MI355X (gfx950).

Host-side mirror of FFTA.jl's AbstractFFTs plan boundary (src/plan.jl),
calling the C ABI of libfftahip.so through ctypes.  Arrays follow Julia's
conventions: shapes and regions are Julia's (1-based dims, column-major
memory).  numpy arrays are taken as column-major (Fortran order) data;
device arrays are passed as (pointer, dims).

Reference correspondence:
  plan_fft / plan_bfft      src/plan.jl:71-100
  Plan.size(i)              src/plan.jl:36-47
  mul_(y, p, x)             LinearAlgebra.mul!, src/plan.jl:157-285
  p * x                     Base.:*, src/plan.jl:374-384
  plan_ifft / ifft          AbstractFFTs ScaledPlan: bfft with 1/prod(size(x, region))
                            applied by the pass that writes the result
  fft / bfft                AbstractFFTs convenience front-ends (plan, then apply)
  plan_rfft / plan_brfft    real transforms, src/plan.jl:102-150, 386-583
  rfft / brfft / irfft
Errors map to DimensionMismatch, ArgumentError and DomainError with the
reference's messages.
"""
from .server import Server  # noqa: F401
from .plan import (  # noqa: F401
    FFT_FORWARD,
    FFT_BACKWARD,
    DEFAULT_BLUESTEIN_CUTOFF,
    ArgumentError,
    DimensionMismatch,
    DomainError,
    FFTAHipError,
    Plan,
    Batch,
    plan_fft,
    plan_bfft,
    plan_ifft,
    RPlan,
    plan_rfft,
    plan_brfft,
    rfft,
    brfft,
    irfft,
    fft,
    bfft,
    ifft,
    mul_,
    device_info,
    version,
)
