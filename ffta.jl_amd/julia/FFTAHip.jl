"""
    FFTAHip

FFTA.jl's complex `fft`/`bfft` plan interface running on an MI355X through the
C ABI of `libfftahip.so` (include/fftahip.h).  Mirrors FFTA's
`src/plan.jl`: `plan_fft`/`plan_bfft` (`:71-100`), `size(p[, i])` (`:36-47`),
`mul!(y, p, x)` (`:157-285`) and `p * x` (`:374-384`), with the same error
types and messages (`DimensionMismatch`, `DomainError`).  Arrays are Julia
`Array{ComplexF32/ComplexF64}` (column-major, interleaved complex), which is
exactly the library's layout; `mul!` stages them through device memory.

Set `ENV["FFTAHIP_LIB"]` to the path of `libfftahip.so` before loading.
"""
module FFTAHip

using AbstractFFTs: AbstractFFTs
using LinearAlgebra: LinearAlgebra

const libfftahip = get(ENV, "FFTAHIP_LIB", joinpath(@__DIR__, "..", "lib", "libfftahip.so"))

const FFT_FORWARD = Cint(-1)
const FFT_BACKWARD = Cint(1)
const DEFAULT_BLUESTEIN_CUTOFF = 73

# status codes of fftahip_status
const SUCCESS = Cint(0)
const ERR_DIMENSION_MISMATCH = Cint(1)
const ERR_ARGUMENT = Cint(2)
const ERR_DOMAIN = Cint(3)

last_error() = unsafe_string(ccall((:fftahip_last_error, libfftahip), Cstring, ()))

function check(st::Cint, arg=nothing)
    st == SUCCESS && return nothing
    msg = last_error()
    st == ERR_DIMENSION_MISMATCH && throw(DimensionMismatch(msg))
    st == ERR_DOMAIN && throw(DomainError(arg, msg))
    st == ERR_ARGUMENT && throw(ArgumentError(msg))
    error("fftahip: " * msg)
end

precision_of(::Type{ComplexF32}) = Cint(32)
precision_of(::Type{ComplexF64}) = Cint(64)

mutable struct FFTAHipPlan{T<:Union{ComplexF32,ComplexF64},N} <: AbstractFFTs.Plan{T}
    handle::Ptr{Cvoid}
    region::Union{Int,Vector{Int}}
    dir::Cint
end

# A region as FFTA accepts it (an Int, a tuple, a vector or a range), sorted
# as _sort does (src/plan.jl:51-69).  collect(2) is a 0-dimensional array,
# so an Int is wrapped rather than collected.
region_vector(region::Integer) = Int64[region]
region_vector(region) = sort!(Int64[r for r in region])

function FFTAHipPlan{T,N}(x::AbstractArray{T}, region, dir::Cint, cutoff::Int) where {T,N}
    dims = Int64[size(x)...]
    reg = region_vector(region)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fftahip_plan_create, libfftahip), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Cint, Cint, Ptr{Int64}, Cint, Ptr{Int64}, Int64),
                h, precision_of(T), dir, length(dims), dims, length(reg), reg, cutoff))
    r = N == 1 ? Int(reg[1]) : Vector{Int}(reg)
    p = FFTAHipPlan{T,N}(h[], r, dir)
    finalizer(p) do q
        q.handle == C_NULL || ccall((:fftahip_plan_destroy, libfftahip), Cint, (Ptr{Cvoid},), q.handle)
        q.handle = C_NULL
    end
    return p
end

# src/plan.jl:71-100
_plan(x::AbstractArray{T}, region, dir; BLUESTEIN_CUTOFF=DEFAULT_BLUESTEIN_CUTOFF, _kwargs...) where {T} =
    FFTAHipPlan{T,length(region_vector(region))}(x, region, dir, BLUESTEIN_CUTOFF)
AbstractFFTs.plan_fft(x::AbstractArray{T}, region; kwargs...) where {T<:Union{ComplexF32,ComplexF64}} =
    _plan(x, region, FFT_FORWARD; kwargs...)
AbstractFFTs.plan_bfft(x::AbstractArray{T}, region; kwargs...) where {T<:Union{ComplexF32,ComplexF64}} =
    _plan(x, region, FFT_BACKWARD; kwargs...)

# AbstractFFTs.plan_ifft = plan_bfft scaled by 1/prod(size(x, region)); the
# factor is applied by the pass that writes the result (one sweep less than
# scaling afterwards)
function AbstractFFTs.plan_ifft(x::AbstractArray{T}, region; kwargs...) where {T<:Union{ComplexF32,ComplexF64}}
    p = _plan(x, region, FFT_BACKWARD; kwargs...)
    check(ccall((:fftahip_plan_set_scale, libfftahip), Cint, (Ptr{Cvoid}, Cdouble), p.handle,
                inv(prod(size(x, d) for d in region))))
    return p
end

# src/plan.jl:36-47
function Base.size(p::FFTAHipPlan{T,N}, i::Int) where {T,N}
    out = Ref{Int64}(0)
    check(ccall((:fftahip_plan_size, libfftahip), Cint, (Ptr{Cvoid}, Cint, Ptr{Int64}), p.handle, i, out), i)
    return Int(out[])
end
Base.size(p::FFTAHipPlan{T,N}) where {T,N} = ntuple(i -> size(p, i), Val(N))

"""
    device_pointer(x) -> Union{Ptr{Cvoid},Nothing}

Device address of the data of `x` when `x` is a dense array in GPU memory,
else `nothing` (the default).  A GPU array package adds a method for its
array type, e.g. for AMDGPU.jl

    FFTAHip.device_pointer(x::ROCArray) = Ptr{Cvoid}(UInt(pointer(x)))
    FFTAHip.stream_handle(x::ROCArray) = AMDGPU.stream().stream

and `mul!`/`*` then run `fftahip_execute` on the device arrays directly
(asynchronous on that stream, no host staging).
"""
device_pointer(x) = nothing
"""hipStream_t (as `Ptr{Cvoid}`) the transform of device array `x` is issued on; C_NULL = default stream."""
stream_handle(x) = C_NULL

# Dense column-major host data (what the library reads through a plain pointer).
is_dense(x::DenseArray) = true
is_dense(x::StridedArray) = strides(x) == Base.size_to_strides(1, size(x)...)
is_dense(x) = false

# src/plan.jl:157-285: argument checks in the library, then the transform.
# Any AbstractArray is accepted, as FFTA's mul! methods do: device arrays
# run in place on the device, dense host arrays are staged through device
# memory by fftahip_execute_host, other host arrays (non-unit-stride views,
# lazy arrays) through a dense copy.
function LinearAlgebra.mul!(y::AbstractArray{T}, p::FFTAHipPlan{T}, x::AbstractArray{T}) where {T}
    Base.require_one_based_indexing(x, y)
    reg = region_vector(p.region)
    dx = Int64[size(x)...]
    dy = Int64[size(y)...]
    check(ccall((:fftahip_plan_check, libfftahip), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Int64}, Cint, Ptr{Int64}, Cint, Ptr{Int64}),
                p.handle, length(reg), reg, length(dx), dx, length(dy), dy))
    px, py = device_pointer(x), device_pointer(y)
    if px !== nothing && py !== nothing
        GC.@preserve x y begin
            check(ccall((:fftahip_execute, libfftahip), Cint,
                        (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                        p.handle, length(dx), dx, px, py, stream_handle(x)))
        end
        return y
    end
    (px === nothing && py === nothing) ||
        throw(ArgumentError("x and y must both be device arrays or both host arrays"))
    xs = is_dense(x) ? x : Array(x)
    ys = is_dense(y) ? y : Array{T}(undef, size(y))
    GC.@preserve xs ys begin
        check(ccall((:fftahip_execute_host, libfftahip), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}),
                    p.handle, length(dx), dx, pointer(xs), pointer(ys)))
    end
    ys === y || copyto!(y, ys)
    return y
end

# src/plan.jl:374-384
function Base.:*(p::FFTAHipPlan{T}, x::AbstractArray{T}) where {T}
    y = similar(x)
    LinearAlgebra.mul!(y, p, x)
    return y
end

# ---------------------------------------------------------------- real transforms
# plan_rfft / plan_brfft (src/plan.jl:102-150) and p * x (src/plan.jl:386-583):
# one library plan per input shape (FFTA's real plans are applied to the
# array they were made from, slice by slice).

mutable struct FFTAHipRPlan{T,N} <: AbstractFFTs.Plan{T}
    handle::Ptr{Cvoid}
    region::Any
    dir::Cint
    flen::Int
    in_size::NTuple{N,Int}
    out_size::NTuple{N,Int}
end

function _rplan(::Type{T}, sz::NTuple{N,Int}, region, dir::Cint, len::Int, cutoff::Int) where {T,N}
    reg = Int64[region...]
    dims = Int64[sz...]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    prec = real(T) == Float32 ? Cint(32) : Cint(64)
    check(ccall((:fftahip_rplan_create, libfftahip), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Cint, Cint, Ptr{Int64}, Int64, Cint, Ptr{Int64}, Int64),
                h, prec, dir, N, dims, len, length(reg), reg, cutoff))
    nd = Ref{Cint}(0)
    out = zeros(Int64, max(N, maximum(reg)))
    check(ccall((:fftahip_rplan_output_dims, libfftahip), Cint, (Ptr{Cvoid}, Ptr{Cint}, Ptr{Int64}), h[], nd, out))
    p = FFTAHipRPlan{T,N}(h[], region, dir, dir == FFT_FORWARD ? sz[first(region)] : len, sz,
                          ntuple(i -> Int(out[i]), Val(N)))
    finalizer(p) do q
        q.handle == C_NULL || ccall((:fftahip_rplan_destroy, libfftahip), Cint, (Ptr{Cvoid},), q.handle)
        q.handle = C_NULL
    end
    return p
end

AbstractFFTs.plan_rfft(x::AbstractArray{T,N}, region; BLUESTEIN_CUTOFF=DEFAULT_BLUESTEIN_CUTOFF,
                       _kwargs...) where {T<:Union{Float32,Float64},N} =
    _rplan(T, size(x), region_vector(region), FFT_FORWARD, 0, BLUESTEIN_CUTOFF)

AbstractFFTs.plan_brfft(x::AbstractArray{T,N}, len::Int, region; BLUESTEIN_CUTOFF=DEFAULT_BLUESTEIN_CUTOFF,
                        _kwargs...) where {T<:Union{ComplexF32,ComplexF64},N} =
    _rplan(T, size(x), region_vector(region), FFT_BACKWARD, len, BLUESTEIN_CUTOFF)

function Base.:*(p::FFTAHipRPlan{T,N}, x::AbstractArray{T,N}) where {T,N}
    Base.require_one_based_indexing(x)
    size(x) == p.in_size || throw(DimensionMismatch("plan and input array axes do not match"))
    y = p.dir == FFT_FORWARD ? Array{complex(T)}(undef, p.out_size) : Array{real(T)}(undef, p.out_size)
    xs = is_dense(x) ? x : Array(x)
    GC.@preserve xs y begin
        check(ccall((:fftahip_rplan_execute_host, libfftahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    p.handle, pointer(xs), pointer(y)))
    end
    return y
end

"""
    fft_batch!(y, x, lengths; directions) -> y

Many independent 1-D transforms of uneven lengths packed back to back in `x`
(one work-list launch per pass on the device).
"""
function fft_batch!(y::Vector{T}, x::Vector{T}, lengths::Vector{Int}, directions::Vector{Cint};
                    BLUESTEIN_CUTOFF=DEFAULT_BLUESTEIN_CUTOFF) where {T<:Union{ComplexF32,ComplexF64}}
    b = Ref{Ptr{Cvoid}}(C_NULL)
    L = Int64.(lengths)
    check(ccall((:fftahip_batch_create, libfftahip), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cint}, Int64),
                b, precision_of(T), length(L), L, C_NULL, directions, BLUESTEIN_CUTOFF))
    try
        nbytes = sizeof(x)
        din = Ref{Ptr{Cvoid}}(C_NULL); dout = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:fftahip_malloc, libfftahip), Cint, (Ptr{Ptr{Cvoid}}, Csize_t), din, nbytes))
        check(ccall((:fftahip_malloc, libfftahip), Cint, (Ptr{Ptr{Cvoid}}, Csize_t), dout, nbytes))
        GC.@preserve x y begin
            check(ccall((:fftahip_memcpy_htod, libfftahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                        din[], pointer(x), nbytes))
            check(ccall((:fftahip_batch_execute, libfftahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                        b[], din[], dout[], C_NULL))
            check(ccall((:fftahip_memcpy_dtoh, libfftahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                        pointer(y), dout[], nbytes))
        end
        ccall((:fftahip_free, libfftahip), Cint, (Ptr{Cvoid},), din[])
        ccall((:fftahip_free, libfftahip), Cint, (Ptr{Cvoid},), dout[])
    finally
        ccall((:fftahip_batch_destroy, libfftahip), Cint, (Ptr{Cvoid},), b[])
    end
    return y
end

end # module
