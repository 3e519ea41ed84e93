"""
    HighOrderMRIHIP

Julia binding of libhoe (include/hoe.h): an MI355X (gfx950) implementation of
HighOrderMRI.jl's Explicit encoding operator behind the same operator API.

`HIPExplicitOp <: HighOrderExplicitOp` carries the LinearOperators fields and
`prod!`/`ctprod!` closures of the reference's resident operators
(src/EncodingOperator/Explicit/explicit_gpu_operator.jl:89-104), so
`op * x`, `adjoint(op) * y`, `normalOperator(op, D)`, `normalOperator(W ∘ op)`,
`recon_HOOp` and RegularizedLeastSquares solvers work unchanged. Vectors may be
host `Array`s (copied by the library) or device arrays whose `pointer` is a
HIP device address on the operator's GPU (e.g. AMDGPU.jl `ROCArray`).
"""
module HighOrderMRIHIP

using LinearAlgebra
using LinearOperators
using LinearOperatorCollection
using ChainRulesCore
using AMDGPU
import HighOrderMRI
using HighOrderMRI: Grid, HighOrderExplicitOp

export HIPExplicitOp, hip_explicit_loss, hip_explicit_grad, hip_FindDelay, hip_apply_girf, set_kspha!,
       set_kspha_dt!, explicit_gpu_normal, route!

const libhoe = get(ENV, "HOE_LIBRARY", joinpath(@__DIR__, "..", "..", "..", "libhoe.so"))

const HOE_DEVICE_PTRS = Cint(1)
const HOE_CURRENT_WEIGHTS = Int64(-1)
const BACKENDS = Dict(:direct => Int32(0), :mixed => Int32(1), :mixed_fp16 => Int32(1), :mixed_bf16 => Int32(2),
                      :fp64 => Int32(3))

# Field order and types of hoe_model_desc.
struct ModelDesc
    n_x::Int32; n_y::Int32; n_z::Int32
    x::Ptr{Float32}; y::Ptr{Float32}; z::Ptr{Float32}
    mask::Ptr{UInt8}
    n_term::Int32; n_sam::Int32
    kspha::Ptr{Float32}; times::Ptr{Float32}; fieldmap::Ptr{Float32}
    n_cha::Int32
    csm::Ptr{ComplexF32}; concomitant::Ptr{Float32}
    recon_terms::Ptr{UInt8}; k_nominal::Ptr{Float32}
    dwell::Float64
    backend::Int32; device::Int32
end

# Field order and types of hoe_model_desc_f64 (the FP64 backend).
struct ModelDescF64
    n_x::Int32; n_y::Int32; n_z::Int32
    x::Ptr{Float64}; y::Ptr{Float64}; z::Ptr{Float64}
    mask::Ptr{UInt8}
    n_term::Int32; n_sam::Int32
    kspha::Ptr{Float64}; times::Ptr{Float64}; fieldmap::Ptr{Float64}
    n_cha::Int32
    csm::Ptr{ComplexF64}; concomitant::Ptr{Float64}
    recon_terms::Ptr{UInt8}; k_nominal::Ptr{Float64}
    device::Int32
end

function check(rc::Integer, h::Ptr{Cvoid}=C_NULL)
    rc == 0 && return nothing
    throw_status(rc, unsafe_string(ccall((:hoe_last_error, libhoe), Cstring, (Ptr{Cvoid},), h)))
end

function throw_status(rc::Integer, msg::AbstractString)
    rc == 1 && throw(ArgumentError(msg))
    rc == 2 && throw(DimensionMismatch(msg))
    rc == 3 && throw(AssertionError(msg))
    rc == 4 && throw(OutOfMemoryError())
    rc == 6 && throw(ArgumentError(msg))
    error("libhoe: ", msg)
end

ondevice(x) = !(x isa Array)
flags(x) = ondevice(x) ? HOE_DEVICE_PTRS : Cint(0)

# libhoe reads device vectors on its own stream: work queued on the caller's AMDGPU
# stream that writes them must be finished first.
function sync_for(xs...)
    any(ondevice, xs) && AMDGPU.synchronize()
    return nothing
end

# Identity of a weights vector handed to libhoe (objectid can be reused by a new array
# at the address of a collected one, a counter cannot).
const WEIGHTS_GEN = Threads.Atomic{UInt}(0)
next_weights_id() = Threads.atomic_add!(WEIGHTS_GEN, UInt(1)) + UInt(1)

# T = ComplexF32 (backends :direct, :mixed, :mixed_bf16) or ComplexF64 (:fp64, built from
# Grid{Float64} inputs).
mutable struct HIPExplicitOp{T,F,G} <: HighOrderExplicitOp{T}
    nrow::Int
    ncol::Int
    symmetric::Bool
    hermitian::Bool
    prod!::F
    tprod!::Nothing
    ctprod!::G
    nprod::Int
    ntprod::Int
    nctprod::Int
    Mv::Vector{T}
    Mtu::Vector{T}
    handle::Ptr{Cvoid}
    nSam::Int
    nCha::Int
    nTerm::Int
    weights_id::UInt
end

LinearOperators.storage_type(op::HIPExplicitOp) = typeof(op.Mv)

handle(op::HIPExplicitOp) = op.handle == C_NULL ? throw(ArgumentError("operator is closed")) : op.handle

"""
    HIPExplicitOp(grid, kspha, times; fieldmap, csm, mask, concomitant, kspha_dt,
                  recon_terms, k_nominal, gpus=[0], backend=:mixed, dwell=0)

Same inputs and conventions as `HighOrderExplicitOp(...; implementation=:kernel)`
(kernel_operator.jl:199-221): `kspha` is `(9|16, nSam)` in cycles/m^p, `times` in
seconds, `fieldmap` in Hz, `csm` `(nX,nY,nZ,nCha)`, `mask` Bool, `concomitant`
`(4, nSam)`. `backend` is `:direct` (FP32 VALU), `:mixed`/`:mixed_fp16` or
`:mixed_bf16` (MFMA). `dwell` is the coefficient sampling interval used by delays.

With `kspha_dt` (`size(kspha)`, dk/dt per term) the operator is the derivative
operator of `explicit_array_derivative` (array_operator.jl:187-235, as built by
FindDelay.jl:184-187): `op * x` is `B x`, `op' * y` stays `Eᴴ y`, concomitant
terms are dropped (array_operator.jl:139) and `normalOperator`/`hip_explicit_grad`
raise `ArgumentError`.
"""
function HIPExplicitOp(grid::Grid{Float32}, kspha::AbstractMatrix{Float32}, times::AbstractVector{Float32};
                       fieldmap=zeros(Float32, grid.matrixSize...), csm=ones(ComplexF32, grid.matrixSize..., 1),
                       mask=trues(grid.matrixSize...), concomitant=nothing, kspha_dt=nothing,
                       recon_terms=nothing, k_nominal=nothing, gpus=[0], backend::Symbol=:mixed, dwell::Real=0)
    length(gpus) == 1 || throw(ArgumentError("HIPExplicitOp runs on one GPU"))
    if kspha_dt !== nothing
        @assert size(kspha_dt) == size(kspha) "kspha_dt must have same size as kspha"
        concomitant = nothing
    end
    haskey(BACKENDS, backend) || throw(ArgumentError("unknown explicit backend $backend"))
    nTerm, nSam = size(kspha)
    nCha = size(csm)[end]
    x, y, z = Vector{Float32}(grid.x), Vector{Float32}(grid.y), Vector{Float32}(grid.z)
    m = Vector{UInt8}(vec(mask))
    k = Matrix{Float32}(kspha)
    t = Vector{Float32}(times)
    f = Vector{Float32}(vec(fieldmap))
    c = Vector{ComplexF32}(vec(csm))                       # grid fastest, then coils
    conc = concomitant === nothing ? nothing : Matrix{Float32}(concomitant)
    kn = k_nominal === nothing ? nothing : Matrix{Float32}(k_nominal)
    rt = recon_terms === nothing ? nothing : Vector{UInt8}(codeunits(recon_terms * "\0"))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x y z m k t f c conc kn rt begin
        d = ModelDesc(grid.nX, grid.nY, grid.nZ, pointer(x), pointer(y), pointer(z), pointer(m),
                      nTerm, nSam, pointer(k), pointer(t), pointer(f), nCha, pointer(c),
                      conc === nothing ? C_NULL : pointer(conc), rt === nothing ? C_NULL : pointer(rt),
                      kn === nothing ? C_NULL : pointer(kn), Float64(dwell), BACKENDS[backend], Int32(first(gpus)))
        check(ccall((:hoe_create, libhoe), Cint, (Ref{ModelDesc}, Ref{Ptr{Cvoid}}), d, h))
    end
    if kspha_dt !== nothing
        kdt = Matrix{Float32}(kspha_dt)
        rc = GC.@preserve kdt ccall((:hoe_set_kspha_dt, libhoe), Cint, (Ptr{Cvoid}, Ptr{Float32}),
                                    h[], pointer(kdt))
        if rc != 0
            msg = unsafe_string(ccall((:hoe_last_error, libhoe), Cstring, (Ptr{Cvoid},), h[]))
            ccall((:hoe_destroy, libhoe), Cvoid, (Ptr{Cvoid},), h[])
            throw_status(rc, msg)
        end
    end
    nTermEff = nTerm + (concomitant === nothing ? 0 : 4)
    return wrap_handle(ComplexF32, h[], nSam, nCha, prod(grid.matrixSize), nTermEff)
end

"""
    HIPExplicitOp(grid::Grid{Float64}, kspha::AbstractMatrix{Float64}, times::AbstractVector{Float64};
                  fieldmap, csm, mask, concomitant, recon_terms, k_nominal, gpus=[0])

The Float64 operator (backend `:fp64`, `hoe_create_f64`): FP64 model, phase, sincos and
contraction, `ComplexF64` vectors; the reference's operators with `Grid{Float64}`
(test_explicit_gpu.jl:51). The derivative product and delays are not provided in FP64.
"""
function HIPExplicitOp(grid::Grid{Float64}, kspha::AbstractMatrix{Float64}, times::AbstractVector{Float64};
                       fieldmap=zeros(Float64, grid.matrixSize...), csm=ones(ComplexF64, grid.matrixSize..., 1),
                       mask=trues(grid.matrixSize...), concomitant=nothing, kspha_dt=nothing,
                       recon_terms=nothing, k_nominal=nothing, gpus=[0], backend::Symbol=:fp64, dwell::Real=0)
    length(gpus) == 1 || throw(ArgumentError("HIPExplicitOp runs on one GPU"))
    kspha_dt === nothing || throw(ArgumentError("kspha_dt (the derivative product) is not provided in Float64"))
    backend in (:fp64, :mixed, :direct, :mixed_fp16, :mixed_bf16) ||
        throw(ArgumentError("unknown explicit backend $backend"))
    nTerm, nSam = size(kspha)
    nCha = size(csm)[end]
    x, y, z = Vector{Float64}(grid.x), Vector{Float64}(grid.y), Vector{Float64}(grid.z)
    m = Vector{UInt8}(vec(mask))
    k = Matrix{Float64}(kspha)
    t = Vector{Float64}(times)
    f = Vector{Float64}(vec(fieldmap))
    c = Vector{ComplexF64}(vec(csm))
    conc = concomitant === nothing ? nothing : Matrix{Float64}(concomitant)
    kn = k_nominal === nothing ? nothing : Matrix{Float64}(k_nominal)
    rt = recon_terms === nothing ? nothing : Vector{UInt8}(codeunits(recon_terms * "\0"))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve x y z m k t f c conc kn rt begin
        d = ModelDescF64(grid.nX, grid.nY, grid.nZ, pointer(x), pointer(y), pointer(z), pointer(m),
                         nTerm, nSam, pointer(k), pointer(t), pointer(f), nCha, pointer(c),
                         conc === nothing ? C_NULL : pointer(conc), rt === nothing ? C_NULL : pointer(rt),
                         kn === nothing ? C_NULL : pointer(kn), Int32(first(gpus)))
        check(ccall((:hoe_create_f64, libhoe), Cint, (Ref{ModelDescF64}, Ref{Ptr{Cvoid}}), d, h))
    end
    nTermEff = nTerm + (concomitant === nothing ? 0 : 4)
    return wrap_handle(ComplexF64, h[], nSam, nCha, prod(grid.matrixSize), nTermEff)
end

# The LinearOperators fields and prod!/ctprod! closures around a libhoe handle.
function wrap_handle(::Type{T}, h::Ptr{Cvoid}, nSam, nCha, nGrid, nTermEff) where {T<:Union{ComplexF32,ComplexF64}}
    op = HIPExplicitOp{T,Any,Any}(nSam * nCha, nGrid, false, false, nothing, nothing, nothing,
                                  0, 0, 0, T[], T[], h, nSam, nCha, nTermEff, UInt(0))
    op.prod! = (res, xv) -> begin
        sync_for(xv, res)
        check(ccall((:hoe_forward, libhoe), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Cint),
                    handle(op), pointer(xv), pointer(res), flags(xv)), op.handle)
        res
    end
    op.ctprod! = (res, yv) -> begin
        sync_for(yv, res)
        check(ccall((:hoe_adjoint, libhoe), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Cint),
                    handle(op), pointer(yv), pointer(res), flags(yv)), op.handle)
        res
    end
    finalizer(close, op)
    return op
end

# ---- the reference's family constructor (HighOrderExplicitOp.jl:44-51) ----

# Implementations of HighOrderExplicitOp(...; implementation) that build a HIPExplicitOp
# besides :hip (route!).
const ROUTED = Dict{Symbol,Bool}(:kernel => false, :array => false)
const ROUTE_BACKEND = Ref(:mixed)

"""
    route!(; kernel=true, array=false, backend=:mixed)

Make the reference's `HighOrderExplicitOp(grid, kspha, times; implementation, ...)` build a
`HIPExplicitOp` (with `backend`) for `implementation=:kernel` and/or `:array`, so callers
that construct operators themselves (`FindDelay` with its default `:array`,
demo/Recon_Benchmark.jl and demo/CompareMeasuredNominal.jl with `:kernel`) run on the GPU
unchanged. `implementation=:hip` always does.
"""
function route!(; kernel::Bool=true, array::Bool=false, backend::Symbol=:mixed)
    haskey(BACKENDS, backend) || throw(ArgumentError("unknown explicit backend $backend"))
    ROUTED[:kernel] = kernel
    ROUTED[:array] = array
    ROUTE_BACKEND[] = backend
    return nothing
end

# Keywords of HIPExplicitOp; the reference's execution keywords (nBlock, arrayType,
# verbose, block sizes, phase/GEMM options) have no meaning here and are dropped.
const HIP_KEYWORDS = (:fieldmap, :csm, :mask, :concomitant, :kspha_dt, :recon_terms, :k_nominal, :gpus, :dwell)

function HighOrderMRI.HighOrderExplicitOp(grid::Grid{T}, kspha::Matrix{T}, times::Vector{T};
                                          implementation::Symbol=:array, kwargs...) where {T<:Union{Float32,Float64}}
    if implementation === :hip || get(ROUTED, implementation, false)
        kw = Dict{Symbol,Any}(k => v for (k, v) in kwargs if k in HIP_KEYWORDS)
        b = get(kwargs, :backend, ROUTE_BACKEND[])
        kw[:backend] = haskey(BACKENDS, b) ? b : ROUTE_BACKEND[]
        return HIPExplicitOp(grid, kspha, times; kw...)
    end
    implementation === :array && return HighOrderMRI.ArrayExplicitOp(grid, kspha, times; kwargs...)
    implementation === :kernel && return HighOrderMRI.CUDAExplicitOp(grid, kspha, times; kwargs...)
    throw(ArgumentError("implementation must be :array, :kernel or :hip"))
end

function Base.adjoint(op::HIPExplicitOp{T}) where {T}
    return LinearOperator{T}(op.ncol, op.nrow, false, false, op.ctprod!, nothing, op.prod!)
end

function Base.close(op::HIPExplicitOp)
    op.handle == C_NULL && return nothing
    ccall((:hoe_destroy, libhoe), Cvoid, (Ptr{Cvoid},), op.handle)
    op.handle = C_NULL
    return nothing
end

# normalOperator(W ∘ op) supplies WᴴW; normalOperator(op, D) takes the squared diagonal,
# real within 10 eps, nonnegative and finite (explicit_gpu_operator.jl:399-416).
function normal_weights(op::HIPExplicitOp, raw)
    raw === nothing && return nothing
    A = Array(raw)
    isempty(A) && throw(DimensionMismatch("normal weights"))
    all(isfinite, A) || throw(ArgumentError("normal weights must be finite"))
    if eltype(A) <: Complex
        scale = max(1, maximum(abs, A))
        maximum(abs ∘ imag, A) <= 10 * eps(Float32) * scale ||
            throw(ArgumentError("normal weights must already be real nonnegative WᴴW"))
    end
    D = Vector{real(eltype(op.Mv))}(real.(vec(A)))
    minimum(D) >= 0 || throw(ArgumentError("normal weights must be nonnegative"))
    length(D) in (op.nSam, op.nrow) || throw(DimensionMismatch("normal weights"))
    return D
end

set_weights_ccall(h, D::Union{Nothing,Vector{Float32}}) =
    GC.@preserve D ccall((:hoe_set_weights, libhoe), Cint, (Ptr{Cvoid}, Ptr{Float32}, Int64), h,
                         D === nothing ? C_NULL : pointer(D), D === nothing ? 0 : length(D))
set_weights_ccall(h, D::Vector{Float64}) =
    GC.@preserve D ccall((:hoe_set_weights_f64, libhoe), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), h,
                         pointer(D), length(D))

function LinearOperatorCollection.normalOperator(op::HIPExplicitOp{T}, weights=nothing; kwargs...) where {T}
    raw = weights isa LinearOperatorCollection.WeightingOp ? weights.weights : weights
    D = normal_weights(op, raw)
    id = next_weights_id()
    normal = (res, xv) -> begin
        h = handle(op)
        if op.weights_id != id
            check(set_weights_ccall(h, D), h)
            op.weights_id = id
        end
        sync_for(xv, res)
        check(ccall((:hoe_normal, libhoe), Cint, (Ptr{Cvoid}, Ptr{Float32}, Int64, Ptr{T}, Ptr{T}, Cint),
                    h, C_NULL, HOE_CURRENT_WEIGHTS, pointer(xv), pointer(res), flags(xv)), h)
        res
    end
    return LinearOperator{T}(op.ncol, op.ncol, false, true, normal, nothing, normal)
end

"""
    explicit_gpu_normal(op::HIPExplicitOp; weights=nothing)

Normal operator from raw weights `W`: `abs2.(weights)` is squared once here
(explicit_gpu_operator.jl:418-433).
"""
explicit_gpu_normal(op::HIPExplicitOp; weights=nothing) =
    LinearOperatorCollection.normalOperator(op, weights === nothing ? nothing : abs2.(Array(weights)))

"""
    hip_explicit_loss(op, h, τ, x, y)

`‖E(h, τ) x − y‖²` with coefficient traces `h` `(nTerm_eff, nSam)` and delay `τ`
(seconds); differentiable through `ChainRulesCore.rrule` with respect to `h` and `τ`.
"""
function hip_explicit_loss(op::HIPExplicitOp{ComplexF32}, h::AbstractMatrix{Float32}, τ::Real, x, y)
    loss, _, _ = hip_explicit_grad(op, h, τ, x, y)
    return loss
end

"""
    hip_explicit_grad(op, h, τ, x, y) -> (loss, g_h, g_τ)

Sets the traces `h` `(nTerm_eff, nSam)` and the delay `τ`, and returns `‖E(h, τ) x − y‖²`,
its gradient with respect to `h` and its derivative with respect to `τ`. `x` and `y` are
both host `Array`s or both device arrays; with device arrays `h` may be a device array
too and `g_h` is one (no host round trip).
"""
function hip_explicit_grad(op::HIPExplicitOp{ComplexF32}, h::AbstractMatrix{Float32}, τ::Real, x, y)
    ondevice(x) == ondevice(y) || throw(ArgumentError("x and y must both be host or both be device arrays"))
    size(h) == (op.nTerm, op.nSam) || throw(DimensionMismatch("coefficients"))
    hd = handle(op)
    H = ondevice(h) ? h : Matrix{Float32}(h)
    sync_for(H, x, y)
    GC.@preserve H check(ccall((:hoe_set_coefficients, libhoe), Cint, (Ptr{Cvoid}, Ptr{Float32}, Float64, Cint),
                               hd, pointer(H), Float64(τ), flags(H)), hd)
    g = ondevice(x) ? ROCArray{Float32}(undef, op.nTerm, op.nSam) : Matrix{Float32}(undef, op.nTerm, op.nSam)
    gτ, loss = Ref{Float64}(0), Ref{Float64}(0)
    check(ccall((:hoe_grad, libhoe), Cint,
                (Ptr{Cvoid}, Ptr{ComplexF32}, Ptr{ComplexF32}, Ptr{Float32}, Ref{Float64}, Ref{Float64}, Cint),
                hd, pointer(x), pointer(y), pointer(g), gτ, loss, flags(x)), hd)
    return loss[], g, gτ[]
end

function ChainRulesCore.rrule(::typeof(hip_explicit_loss), op::HIPExplicitOp, h, τ, x, y)
    loss, g, gτ = hip_explicit_grad(op, h, τ, x, y)
    pullback(Δ) = (NoTangent(), NoTangent(), Δ .* g, Δ * gτ, NoTangent(), NoTangent())
    return loss, pullback
end

"""
    set_kspha!(op, kspha; k_nominal=nothing)

Replace the operator's coefficients, prepared as at construction (`recon_terms`,
`k_nominal`; prep_kspha.jl:1-56): the operator then equals one built anew with `kspha`.
"""
function set_kspha!(op::HIPExplicitOp, kspha::AbstractMatrix; k_nominal=nothing)
    k = Matrix{Float32}(kspha)
    kn = k_nominal === nothing ? nothing : Matrix{Float32}(k_nominal)
    hd = handle(op)
    GC.@preserve k kn check(ccall((:hoe_set_kspha, libhoe), Cint, (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}),
                                  hd, pointer(k), kn === nothing ? C_NULL : pointer(kn)), hd)
    return op
end

"""
    set_kspha_dt!(op, kspha_dt)

Switch `op * x` to the derivative product `B x` of array_operator.jl:187-235
(`kspha_dt` of `size(kspha)`), or back to `E x` with `nothing`.
"""
function set_kspha_dt!(op::HIPExplicitOp, kspha_dt)
    hd = handle(op)
    if kspha_dt === nothing
        check(ccall((:hoe_set_kspha_dt, libhoe), Cint, (Ptr{Cvoid}, Ptr{Float32}), hd, C_NULL), hd)
    else
        kdt = Matrix{Float32}(kspha_dt)
        GC.@preserve kdt check(ccall((:hoe_set_kspha_dt, libhoe), Cint, (Ptr{Cvoid}, Ptr{Float32}),
                                     hd, pointer(kdt)), hd)
    end
    return op
end

"""
    hip_FindDelay(gridding, data, kspha, datatime, StartTime, dt; JumpFact=3, Δτ_min,
                  intermode=AkimaMonotonicInterpolation(), fieldmap, csm, recon_terms="111",
                  solver=CGNR, reg=L2Regularization(zero(T)), iter_max=10, backend=:mixed)

`FindDelay` (FindDelay.jl:139-214) with every operator product and the inner CGNR's
products on the GPU through one `HIPExplicitOp`: each trial delay replaces the
coefficients with `set_kspha!` instead of building two operators, and `B x` is taken
with `set_kspha_dt!`. Resampling and density weights are the reference's own
`InterpTrajTime` and `samplingDensity`.
"""
function hip_FindDelay(gridding::Grid{T}, data::AbstractArray{Complex{T},2}, kspha::AbstractArray{T,2},
                       datatime::AbstractVector{T}, StartTime::T, dt::T; JumpFact=3, Δτ_min::T=T(0.005),
                       intermode=HighOrderMRI.AkimaMonotonicInterpolation(),
                       fieldmap=zeros(T, gridding.nX, gridding.nY),
                       csm=ones(Complex{T}, gridding.nX, gridding.nY, 1), recon_terms::String="111",
                       solver=HighOrderMRI.CGNR, reg=HighOrderMRI.L2Regularization(zero(T)),
                       iter_max::Int=10, backend::Symbol=:mixed) where {T<:AbstractFloat}
    @assert size(data, 2) == size(csm, 3) "data and csm must have the same number of coil channels"
    @assert size(data, 1) == size(datatime, 1) "data and datatime must have the same number of spatial points"
    nSample = size(data, 1)
    # 5-tap derivative with zeros beyond both ends (conv(kspha, kernel)[3:end-2, :] / dt)
    c = T.((1/8, 1/4, 0, -1/4, -1/8))
    n = size(kspha, 1)
    kz(i, j) = 1 <= i <= n ? kspha[i, j] : zero(T)
    kspha_dt = [sum(c[m] * kz(i + 3 - m, j) for m in 1:5) / dt for i in 1:n, j in 1:size(kspha, 2)]
    recParams = Dict{Symbol,Any}(:reconSize => (gridding.nX, gridding.nY), :reg => reg,
                                 :iterations => iter_max, :solver => solver)
    τ = 0.0; nIter = 1; Δτ_prev = Δτ = Inf
    op = nothing
    try
        while abs(Δτ) > Δτ_min
            kspha_τ = T.(HighOrderMRI.InterpTrajTime(kspha, dt, τ + StartTime, datatime;
                                                     intermode=intermode)[1:nSample, :]')
            kspha_dt_τ = T.(HighOrderMRI.InterpTrajTime(kspha_dt, dt, τ + StartTime, datatime;
                                                        intermode=intermode)[1:nSample, :]')
            weight = HighOrderMRI.samplingDensity(kspha_τ[2:3, :], (gridding.nX, gridding.nY))
            if op === nothing
                op = HIPExplicitOp(gridding, kspha_τ, datatime; recon_terms=recon_terms, csm=csm,
                                   fieldmap=fieldmap, backend=backend)
            else
                set_kspha!(op, kspha_τ)
            end
            x = HighOrderMRI.recon_HOOp(op, data, weight, recParams)
            y1 = vec(data) - op * vec(x)          # Y - Aₚxₚ
            set_kspha_dt!(op, kspha_dt_τ)
            y2 = op * vec(x)                      # Bₚxₚ
            set_kspha_dt!(op, nothing)
            Δτ = JumpFact * real(y2 \ y1)
            τ += Δτ
            @info "Iteration $nIter: Δτ = $(round(Δτ/dt, digits=5)) [dt] | τ = $(round(τ/dt, digits=5)) [dt] | JumpFact = $JumpFact"
            if nIter > 1 && sign(Δτ_prev) != sign(Δτ)
                JumpFact = max(1, JumpFact / 2)
            end
            Δτ_prev = Δτ
            nIter += 1
        end
    finally
        op === nothing || close(op)
    end
    return T(τ - Δτ / JumpFact * (JumpFact - 1))
end

"""
    hip_apply_girf(G_DCS_nom, Hw; dim_spatial=1, dim_time=2, rbw=1.0, device=0)

`apply_girf` (src/GIRF/apply_girf.jl:39-109) on the GPU: hipFFT transforms and HIP
kernels for the transfer product, batched over every waveform. Same arguments, shapes,
result and FFT-length / response-expansion / rbw conventions; `Float32` gradients are
computed in FP32, `Float64` in FP64. Also accepts a `GIRFModel`.
"""
function hip_apply_girf(G::AbstractArray{T,N}, Hw::AbstractArray; dim_spatial::Int=1, dim_time::Int=2,
                        rbw::Real=1.0, device::Integer=0) where {T<:Union{Float32,Float64},N}
    size(G, dim_spatial) == 3 || throw(ArgumentError("Input gradient must have 3 channels along dim_spatial."))
    nFreq, nIn, nOut = size(Hw)
    @assert nIn == 3 "GIRF model must have 3 input channels (X, Y, Z)"
    nPoint = size(G, dim_time)
    other = Tuple(setdiff(1:N, (dim_spatial, dim_time)))
    perm = (dim_spatial, dim_time, other...)
    g = Array{T}(permutedims(G, perm))                 # (3, nPoint, others...), channel fastest
    nOther = length(g) ÷ (3 * nPoint)
    h = Array{Complex{T}}(Hw)
    out = Array{T}(undef, nOut, nPoint, (size(G, d) for d in other)...)
    check(girf_ccall(g, nPoint, nOther, h, nFreq, nOut, Float64(rbw), Int32(device), out))
    return permutedims(out, invperm(perm))
end

girf_ccall(g::Array{Float32}, nPoint, nOther, h::Array{ComplexF32}, nFreq, nOut, rbw, device, out::Array{Float32}) =
    ccall((:hoe_apply_girf, libhoe), Cint,
          (Ptr{Float32}, Int64, Int64, Ptr{ComplexF32}, Int32, Int32, Float64, Int32, Ptr{Float32}, Cint),
          g, nPoint, nOther, h, nFreq, nOut, rbw, device, out, Cint(0))
girf_ccall(g::Array{Float64}, nPoint, nOther, h::Array{ComplexF64}, nFreq, nOut, rbw, device, out::Array{Float64}) =
    ccall((:hoe_apply_girf_f64, libhoe), Cint,
          (Ptr{Float64}, Int64, Int64, Ptr{ComplexF64}, Int32, Int32, Float64, Int32, Ptr{Float64}, Cint),
          g, nPoint, nOther, h, nFreq, nOut, rbw, device, out, Cint(0))
hip_apply_girf(G::AbstractArray, model::HighOrderMRI.GIRFModel; kwargs...) = hip_apply_girf(G, model.Hw; kwargs...)

end # module
