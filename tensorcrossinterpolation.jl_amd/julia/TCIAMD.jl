# TCIAMD.jl — Julia binding of libtci_hip.so for TensorCrossInterpolation.jl.
#
# Adds methods (no new names a user has to learn) so that the TCI2 hot path
# runs on an MI355X:
#
#   TCI.rrlu(A::Matrix{Float64}; maxrank, reltol, abstol, leftorthogonal)
#       → tci_rrlu_f64                                 (matrixlu.jl:217-225)
#   DeviceFunctor{Float64} <: TCI.BatchEvaluator{Float64}
#       (f)(Iset, Jset, Val(M)) → tci_pi_eval_f64      (batcheval.jl:67-83)
#   TCI.crossinterpolate2(Float64, f::DeviceFunctor, localdims, initialpivots; kw...)
#       → tci_solve_submit / tci_batch_wait            (tensorci2.jl:943-953)
#   TCI.optimize!(tci::TensorCI2{Float64}, f::DeviceFunctor; kw...)
#       → tci_solve_submit with a tci_state            (tensorci2.jl:700-850)
#   TCI.TensorCI2{Float64}(f::DeviceFunctor, localdims, Iset, Jset)
#       → a zero-iteration solve from the sets         (tensorci2.jl:58-72)
#   leftright(A; kw...) → left/right of MatrixLUCI     (matrixluci.jl:40-84)
#   TCI.compress!(tt::TensorTrain{Float64,3}, :LU | :CI; kw...)
#       → tci_tt_compress_f64                          (tensortrain.jl:147-183)
#   TCI.TensorCI2{Float64}(tt::TensorTrain{Float64,3}; kw...)
#       → tci_tt_sweep_f64 per sweep                   (conversion.jl:73-176)
#   TCI.estimatetrueerror(tt::TensorTrain{Float64,3}, f::DeviceFunctor; kw...)
#       → tci_tt_floatingzone_f64                      (globalsearch.jl:18-95)
#
# Every entry point is a plain ccall on the C ABI of include/tci.h; arrays
# are passed as pointers under GC.@preserve.  Nonzero status codes become
# the reference's exceptions (ErrorException with the reference's message,
# ArgumentError, DimensionMismatch, SingularException).
#
# This file is not exercised in this repository's test suite (the build
# image has no Julia); the Python package mirrors the same calls and is
# what the tests run.
module TCIAMD

import TensorCrossInterpolation as TCI
using LinearAlgebra: SingularException

const libtci = get(ENV, "TCI_HIP_LIB", joinpath(@__DIR__, "..", "libtci_hip.so"))
const UNLIMITED = typemax(Int64)

# ---------------------------------------------------------------- errors
function throwstatus(st::Integer, what::AbstractString="")
    st == 0 && return nothing
    msg = unsafe_string(ccall((:tci_last_error, libtci), Cstring, ()))
    st == 1 && error("lu.L contains NaNs")
    st == 2 && error("lu.U contains NaNs")
    st == 3 && throw(ArgumentError(msg))
    st == 4 && error("maxsamplevalue is zero!")
    st == 5 && error(msg)                      # "Error: NaN in tensor T[b]"
    st == 6 && error(msg)                      # "Pivot matrix at bond b is not square!"
    st == 7 && throw(DimensionMismatch(msg))
    st == 9 && throw(SingularException(0))
    error("libtci_hip ($what): status $st: $msg")
end

# ---------------------------------------------------------------- context
mutable struct Context
    handle::Ptr{Cvoid}
    function Context(device::Integer=0; hbm_budget::Integer=0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        throwstatus(ccall((:tci_ctx_create, libtci), Cint, (Ref{Ptr{Cvoid}}, Int32, UInt64),
                          h, device, hbm_budget), "tci_ctx_create")
        c = new(h[])
        finalizer(c) do c
            ccall((:tci_ctx_destroy, libtci), Cint, (Ptr{Cvoid},), c.handle)
        end
        return c
    end
end

const _ctx = Ref{Union{Nothing,Context}}(nothing)
defaultcontext() = something(_ctx[], (_ctx[] = Context(0)))

# ---------------------------------------------------------------- rrLU
"""
    TCI.rrlu(A::Matrix{Float64}; maxrank, reltol, abstol, leftorthogonal)

Full-pivot rank-revealing LU on the GPU; the pivots, L, U and error are
bit-identical to the reference `rrlu` (matrixlu.jl:217-225).
"""
function TCI.rrlu(A::Matrix{Float64}; maxrank::Int=UNLIMITED, reltol::Number=1e-14,
                  abstol::Number=0.0, leftorthogonal::Bool=true)
    m, n = size(A)
    cap = max(min(m, n, maxrank), 1)
    rp = Vector{Int64}(undef, m)
    cp = Vector{Int64}(undef, n)
    L = zeros(Float64, m, cap)
    U = zeros(Float64, cap, n)
    p = Ref{Int64}(0)
    err = Ref{Float64}(0.0)
    GC.@preserve A rp cp L U begin
        st = ccall((:tci_rrlu_f64, libtci), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64, Int64, Int64, Float64, Float64, Int32,
                    Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
                    Ref{Int64}, Ref{Float64}),
                   defaultcontext().handle, A, m, m, n, maxrank, reltol, abstol, leftorthogonal,
                   rp, cp, L, m, U, cap, p, err)
    end
    throwstatus(st, "tci_rrlu_f64")
    np = Int(p[])
    return TCI.rrLU(Vector{Int}(rp), Vector{Int}(cp), L[:, 1:np], U[1:np, :],
                    leftorthogonal, np, err[])
end

# ---------------------------------------------------------------- functors
# Mirrors tci_functor of include/tci.h; the kinds and parameter layouts are
# those of include/tci_functors.h.
struct CFunctor
    kind::Int32
    nparams::Int32
    params::Ptr{Float64}
    niparams::Int32
    iparams::Ptr{Int32}
end

"""
    DeviceFunctor{Float64}(kind, params, iparams=Int32[])

A function given by a device functor kind and its parameters.  Scalar calls
and batch calls both evaluate on the GPU.
"""
struct DeviceFunctor{T} <: TCI.BatchEvaluator{T}
    kind::Int32
    params::Vector{Float64}
    iparams::Vector{Int32}
    localdims::Vector{Int}
end

DeviceFunctor{Float64}(kind::Integer, params, localdims; iparams=Int32[]) =
    DeviceFunctor{Float64}(Int32(kind), Vector{Float64}(params), Vector{Int32}(iparams),
                           collect(localdims))

lorentzian(localdims; a=1.0) = DeviceFunctor{Float64}(1, [a], localdims)

_cf(f::DeviceFunctor) = CFunctor(f.kind, length(f.params), pointer(f.params),
                                 length(f.iparams), pointer(f.iparams))

function _packindices(set::Vector{Vector{Int}}, len::Int)
    out = Vector{UInt8}(undef, length(set) * len)
    for (i, x) in enumerate(set), t in 1:len
        out[(i-1)*len+t] = UInt8(x[t])
    end
    return out
end

# (f)(Iset, Jset, Val(M)) of the BatchEvaluator contract (batcheval.jl:67-83)
function (f::DeviceFunctor{Float64})(leftindexset::Vector{Vector{Int}},
                                     rightindexset::Vector{Vector{Int}}, ::Val{M}) where {M}
    nI, nJ = length(leftindexset), length(rightindexset)
    L = length(f.localdims)
    nl = nI > 0 ? length(first(leftindexset)) : 0
    nr = nJ > 0 ? length(first(rightindexset)) : L - nl - M
    dims = (nI, f.localdims[nl+1:nl+M]..., nJ)
    out = zeros(Float64, dims...)
    (nI == 0 || nJ == 0) && return out
    left = _packindices(leftindexset, nl)
    right = _packindices(rightindexset, nr)
    ld = Int32.(f.localdims)
    mx = Ref{Float64}(0.0)
    GC.@preserve f left right ld out begin
        cf = Ref(_cf(f))
        st = ccall((:tci_pi_eval_f64, libtci), Cint,
                   (Ptr{Cvoid}, Ref{CFunctor}, Int32, Ptr{Int32}, Ptr{UInt8}, Int64, Int32,
                    Ptr{UInt8}, Int64, Int32, Ptr{Float64}, Ref{Float64}),
                   defaultcontext().handle, cf, L, ld, left, nI, nl, right, nJ, nr, out, mx)
    end
    throwstatus(st, "tci_pi_eval_f64")
    return out
end

(f::DeviceFunctor{Float64})(x::Vector{Int}) =
    f([x[1:1]], [x[2:end]], Val(0))[1, 1]

# ---------------------------------------------------------------- LUCI
"""
    leftright(A::Matrix{Float64}; maxrank, reltol, abstol, leftorthogonal)

`(left(luci), right(luci), luci_rowindices, luci_colindices)` of
`MatrixLUCI(A; kwargs...)` (matrixluci.jl:5-7, 40-84) with the factorization
and both factors computed on the device (tci_luci_f64).
"""
function leftright(A::Matrix{Float64}; maxrank::Int=UNLIMITED, reltol::Number=1e-14,
                   abstol::Number=0.0, leftorthogonal::Bool=true)
    m, n = size(A)
    cap = max(min(m, n, maxrank), 0)
    rp = zeros(Int64, max(m, 1)); cp = zeros(Int64, max(n, 1))
    lb = zeros(Float64, max(m, 1), max(cap, 1)); rb = zeros(Float64, max(cap, 1), max(n, 1))
    p = Ref{Int64}(0); err = Ref{Float64}(0.0)
    GC.@preserve A rp cp lb rb begin
        throwstatus(ccall((:tci_luci_f64, libtci), Cint,
                          (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64, Int64, Int64, Float64, Float64, Int32,
                           Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ref{Int64},
                           Ref{Float64}),
                          defaultcontext().handle, A, max(m, 1), m, n, maxrank, reltol, abstol, leftorthogonal,
                          rp, cp, lb, max(m, 1), rb, max(cap, 1), p, err), "tci_luci_f64")
    end
    k = p[]
    return lb[1:m, 1:k], rb[1:k, 1:n], rp[1:k], cp[1:k]
end

# ---------------------------------------------------------------- solves
struct COptions
    tolerance::Float64
    maxbonddim::Int64
    maxiter::Int32
    sweepstrategy::Int32
    normalizeerror::Int32
    ncheckhistory::Int32
    maxnglobalpivot::Int32
    nsearchglobalpivot::Int32
    tolmarginglobalsearch::Float64
    strictlynested::Int32
    checkconvglobalpivot::Int32
    seed::UInt64
    chi_cap::Int32
    flags::Int32
    cache_log2::Int32
    pivotsearch::Int32
end

const FLAG_NO_FINAL_SWEEP = Int32(1)

# tci_state: a solve's starting pivot sets (and the last history entry)
struct CState
    nI::Ptr{Int32}
    Iset::Ptr{UInt8}
    nJ::Ptr{Int32}
    Jset::Ptr{UInt8}
    nI_hist::Ptr{Int32}
    Iset_hist::Ptr{UInt8}
    nJ_hist::Ptr{Int32}
    Jset_hist::Ptr{UInt8}
    maxsamplevalue::Float64
end

struct CProblem
    L::Int32
    localdims::Ptr{Int32}
    f::CFunctor
    npivots::Int32
    initialpivots::Ptr{Int32}
    opts::COptions
    state::Ptr{CState}
end

struct CResult
    status::Int32
    niter::Int32
    L::Int32
    npivoterrors::Int32
    maxsamplevalue::Float64
    workstats::NTuple{6,Float64}
end

const _sweep = Dict(:backandforth => 0, :forward => 1, :backward => 2)

# counts and bytes of one family of sets (site b: entries of length len(b))
function _packsets(sets::Vector{Vector{Vector{Int}}}, len)
    cnt = Int32[length(s) for s in sets]
    bytes = UInt8[]
    for (b, s) in enumerate(sets), x in s
        length(x) == len(b) || throw(ArgumentError("index set $b has an entry of length $(length(x))"))
        append!(bytes, UInt8.(x))
    end
    return cnt, isempty(bytes) ? UInt8[0] : bytes
end

# optimize! keywords (tensorci2.jl:700-720) as the device options
function _options(; tolerance::Union{Float64,Nothing}=nothing, pivottolerance::Union{Float64,Nothing}=nothing,
                  maxbonddim::Int=typemax(Int), maxiter::Int=20, sweepstrategy::Symbol=:backandforth,
                  pivotsearch::Symbol=:full, verbosity::Int=0, loginterval::Int=10,
                  normalizeerror::Bool=true, ncheckhistory::Int=3, maxnglobalpivot::Int=5,
                  nsearchglobalpivot::Int=5, tolmarginglobalsearch::Float64=10.0,
                  strictlynested::Bool=false, checkbatchevaluatable::Bool=false,
                  checkconvglobalpivot::Bool=true, seed::Integer=0, chi_cap::Integer=0, flags::Integer=0,
                  cache_log2::Integer=0)
    # tolerance / pivottolerance resolution (tensorci2.jl:734-746)
    if pivottolerance !== nothing
        if tolerance !== nothing && tolerance != pivottolerance
            throw(ArgumentError("Got different values for pivottolerance and tolerance in optimize!(TCI2). " *
                                "For TCI2, both of these options have the same meaning. Please assign only `tolerance`."))
        end
        @warn "The option `pivottolerance` of `optimize!(tci::TensorCI2, f)` is deprecated. Please update your " *
              "code to use `tolerance`, as `pivottolerance` will be removed in the future."
        tolerance = pivottolerance
    end
    tol = something(tolerance, 1e-8)
    pivotsearch in (:full, :rook) ||
        throw(ArgumentError("Unknown pivot search strategy $pivotsearch. Choose from :rook, :full."))
    # verbosity / loginterval only print in the reference; checkbatchevaluatable
    # holds trivially for a DeviceFunctor (a BatchEvaluator)
    return COptions(tol, maxbonddim, maxiter, _sweep[sweepstrategy], normalizeerror, ncheckhistory,
                    maxnglobalpivot, nsearchglobalpivot, tolmarginglobalsearch, strictlynested,
                    checkconvglobalpivot, UInt64(seed), chi_cap, flags, cache_log2, pivotsearch === :rook ? 1 : 0)
end

# one device solve; a solve that outgrows the capacity chosen at submit is
# resubmitted at twice the capacity (tci_capacity_plan) unless chi_cap was set
function _solve(f::DeviceFunctor{Float64}, localdims::Vector{Int}, piv::Vector{Vector{Int}},
                opts::COptions, state::Union{Nothing,NTuple{9,Any}})
    L = length(localdims)
    ld = Int32.(localdims)
    pv = Int32.(reduce(vcat, piv))
    while true
        batch = Ref{Ptr{Cvoid}}(C_NULL)
        cst = state === nothing ? nothing :
              Ref(CState(pointer(state[1]), pointer(state[2]), pointer(state[3]), pointer(state[4]),
                         state[5] === nothing ? C_NULL : pointer(state[5]),
                         state[6] === nothing ? C_NULL : pointer(state[6]),
                         state[7] === nothing ? C_NULL : pointer(state[7]),
                         state[8] === nothing ? C_NULL : pointer(state[8]), state[9]))
        GC.@preserve f ld pv state cst begin
            sp = cst === nothing ? Ptr{CState}(C_NULL) : Base.unsafe_convert(Ptr{CState}, cst)
            prob = Ref(CProblem(L, pointer(ld), _cf(f), length(piv), pointer(pv), opts, sp))
            throwstatus(ccall((:tci_solve_submit, libtci), Cint,
                              (Ptr{Cvoid}, Int32, Ref{CProblem}, Ref{Ptr{Cvoid}}),
                              defaultcontext().handle, 1, prob, batch), "tci_solve_submit")
            b = batch[]
            try
                throwstatus(ccall((:tci_batch_wait, libtci), Cint, (Ptr{Cvoid},), b), "tci_batch_wait")
                r = Ref{CResult}()
                ccall((:tci_batch_result, libtci), Cint, (Ptr{Cvoid}, Int32, Ref{CResult}), b, 0, r)
                if r[].status == 10 && opts.chi_cap == 0
                    used = Ref{Int64}(0); limit = Ref{Int64}(0)
                    throwstatus(ccall((:tci_capacity_plan, libtci), Cint, (Ref{CProblem}, Ref{Int64}, Ref{Int64}),
                                      prob, used, limit), "tci_capacity_plan")
                    if used[] < limit[]
                        o = opts
                        opts = COptions(o.tolerance, o.maxbonddim, o.maxiter, o.sweepstrategy, o.normalizeerror,
                                        o.ncheckhistory, o.maxnglobalpivot, o.nsearchglobalpivot,
                                        o.tolmarginglobalsearch, o.strictlynested, o.checkconvglobalpivot, o.seed,
                                        Int32(min(2 * used[], limit[], 8192)), o.flags, o.cache_log2, 0)
                        continue
                    end
                end
                throwstatus(r[].status, "solve")
                return _totci(b, r[], localdims)
            finally
                ccall((:tci_batch_free, libtci), Cint, (Ptr{Cvoid},), b)
            end
        end
    end
end

"""
    TCI.crossinterpolate2(Float64, f::DeviceFunctor, localdims, initialpivots; kwargs...)

The whole TCI2 solve (optimize! and the final sweep1site!) on the GPU, one
workgroup per solve.  Returns `(tci::TensorCI2{Float64}, ranks, errors)` like
the reference (tensorci2.jl:943-953) and takes its keyword arguments.  The
global pivot search draws from the library's PRNG (seed keyword) instead of
Julia's default RNG.
"""
function TCI.crossinterpolate2(::Type{Float64}, f::DeviceFunctor{Float64},
                               localdims::Union{Vector{Int},NTuple{N,Int}},
                               initialpivots::Vector{Vector{Int}}=[ones(Int, length(localdims))];
                               kwargs...) where {N}
    return _solve(f, collect(localdims), initialpivots, _options(; kwargs...), nothing)
end

_statebuffers(tci, L, maxsample) = begin
    nI, I = _packsets(tci.Iset, b -> b - 1)
    nJ, J = _packsets(tci.Jset, b -> L - b)
    if isempty(tci.Iset_history)
        (nI, I, nJ, J, nothing, nothing, nothing, nothing, Float64(maxsample))
    else
        nIh, Ih = _packsets(tci.Iset_history[end], b -> b - 1)
        nJh, Jh = _packsets(tci.Jset_history[end], b -> L - b)
        (nI, I, nJ, J, nIh, Ih, nJh, Jh, Float64(maxsample))
    end
end

"""
    TCI.optimize!(tci::TensorCI2{Float64}, f::DeviceFunctor; kwargs...)

optimize! (tensorci2.jl:700-850) continuing from tci's pivot sets, last
history entry and maxsamplevalue, on the device.  Updates tci in place and
returns `(ranks, errors)`.
"""
function TCI.optimize!(tci::TCI.TensorCI2{Float64}, f::DeviceFunctor{Float64}; kwargs...)
    L = length(tci.localdims)
    st = _statebuffers(tci, L, tci.maxsamplevalue)
    t, ranks, errors = _solve(f, collect(tci.localdims), [ones(Int, L)], _options(; kwargs...), st)
    for fld in (:Iset, :Jset, :sitetensors, :pivoterrors, :bonderrors, :maxsamplevalue)
        setfield!(tci, fld, getfield(t, fld))
    end
    append!(tci.Iset_history, t.Iset_history)
    append!(tci.Jset_history, t.Jset_history)
    return ranks, errors
end

"""
    TCI.TensorCI2{Float64}(f::DeviceFunctor, localdims, Iset, Jset)

The constructor from local pivot lists (tensorci2.jl:58-72): maxsamplevalue
is the maximum of |f| over every Iset[b] ⊕ σ ⊕ Jset[b], evaluated on the device.
"""
function TCI.TensorCI2{Float64}(f::DeviceFunctor{Float64}, localdims::Union{Vector{Int},NTuple{N,Int}},
                                Iset::Vector{Vector{Vector{Int}}}, Jset::Vector{Vector{Vector{Int}}}) where {N}
    L = length(localdims)
    nI, I = _packsets(Iset, b -> b - 1)
    nJ, J = _packsets(Jset, b -> L - b)
    st = (nI, I, nJ, J, nothing, nothing, nothing, nothing, 0.0)
    t, _, _ = _solve(f, collect(localdims), [ones(Int, L)],
                     _options(; maxiter=0, flags=FLAG_NO_FINAL_SWEEP), st)
    empty!(t.Iset_history); empty!(t.Jset_history)
    return t
end

# ---------------------------------------------------------------- tensor trains
const _factor = Dict(:LU => Int32(0), :CI => Int32(1), :SVD => Int32(2))

function _packtt(tt::TCI.TensorTrain{Float64,3})
    T = tt.sitetensors
    links = Int64[size(T[1], 1); [size(t, 3) for t in T]]
    ld = Int32[size(t, 2) for t in T]
    return links, ld, reduce(vcat, vec.(T))
end

function _unpacktt!(tt::TCI.TensorTrain{Float64,3}, links, ld, flat)
    o = 0
    for b in eachindex(ld)
        sz = links[b] * ld[b] * links[b+1]
        tt.sitetensors[b] = reshape(flat[o+1:o+sz], links[b], ld[b], links[b+1])
        o += sz
    end
    return tt
end

"""
    TCI.compress!(tt::TensorTrain{Float64,3}, method=:LU; tolerance, maxbonddim, normalizeerror)

compress! (tensortrain.jl:147-183) on the device (tci_tt_compress_f64): the
train is uploaded once, swept left to right without truncation and right to
left with the truncation, and written back.  :LU and :CI; :SVD throws.
"""
function TCI.compress!(tt::TCI.TensorTrain{Float64,3}, method::Symbol=:LU; tolerance::Float64=1e-12,
                       maxbonddim::Int=typemax(Int), normalizeerror::Bool=true)
    haskey(_factor, method) || error("Not implemented yet.")
    links, ld, flat = _packtt(tt)
    GC.@preserve links ld flat begin
        throwstatus(ccall((:tci_tt_compress_f64, libtci), Cint,
                          (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}, Int32, Float64, Int64, Int32),
                          defaultcontext().handle, length(ld), links, ld, flat, _factor[method], tolerance,
                          maxbonddim, normalizeerror), "tci_tt_compress_f64")
    end
    _unpacktt!(tt, links, ld, flat)
    nothing
end

# sweep1sitegetindices! (conversion.jl:73-156): the MatrixLUCI sweep on the
# device (tci_tt_sweep_f64), the index bookkeeping here
function _sweep1sitegetindices!(tt::TCI.TensorTrain{Float64,3}, forwardsweep::Bool,
                                spectatorindices::Vector{Vector{Vector{Int}}}=Vector{Vector{Int}}[];
                                maxbonddim=typemax(Int), tolerance=0.0)
    L = length(tt)
    links, ld, flat = _packtt(tt)
    piv = zeros(Int32, max(2 * sum(links[2:L]), 1))
    pe = zeros(Float64, maximum(links[2:L]) + 1)
    GC.@preserve links ld flat piv pe begin
        throwstatus(ccall((:tci_tt_sweep_f64, libtci), Cint,
                          (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}, Int32, Int32, Int64, Float64,
                           Float64, Ptr{Int32}, Ptr{Float64}),
                          defaultcontext().handle, L, links, ld, flat, Int32(1), forwardsweep, maxbonddim, 1e-14,
                          tolerance, piv, pe), "tci_tt_sweep_f64")
    end
    _unpacktt!(tt, links, ld, flat)
    rows = Vector{Vector{Int}}(undef, L - 1); cols = Vector{Vector{Int}}(undef, L - 1)
    q = 0
    for b in 1:L-1
        p = links[b+1]
        rows[b] = Int.(piv[q+1:q+p]); cols[b] = Int.(piv[q+p+1:q+2p])
        q += 2p
    end
    indexset = Vector{Vector{Int}}[Vector{Int}[Int[]]]
    for i in 1:L-1
        if forwardsweep
            push!(indexset, TCI.kronecker(last(indexset), Int(ld[i]))[rows[i]])
            isempty(spectatorindices) || (spectatorindices[i] = spectatorindices[i][cols[i]])
        else
            b = L - i                       # bond b, site b+1
            push!(indexset, TCI.kronecker(Int(ld[b+1]), last(indexset))[cols[b]])
            isempty(spectatorindices) || (spectatorindices[b+1] = spectatorindices[b+1][rows[b]])
        end
    end
    return (forwardsweep ? indexset : reverse(indexset)), pe
end

"""
    TCI.TensorCI2{Float64}(tt::TensorTrain{Float64,3}; tolerance, maxbonddim, maxiter)

The TT → TensorCI2 conversion (conversion.jl:158-176) with every sweep on
the device; tt is swept in place as in the reference.
"""
function TCI.TensorCI2{Float64}(tt::TCI.TensorTrain{Float64,3}; tolerance=1e-12, maxbonddim=typemax(Int),
                                maxiter=3)
    Iset, = _sweep1sitegetindices!(tt, true; maxbonddim, tolerance)
    Jset, pivoterrors = _sweep1sitegetindices!(tt, false; maxbonddim, tolerance)
    for iter in 3:maxiter
        if isodd(iter)
            Isetnew, pivoterrors = _sweep1sitegetindices!(tt, true, Jset)
            Isetnew == Iset && break
        else
            Jsetnew, pivoterrors = _sweep1sitegetindices!(tt, false, Iset)
            Jsetnew == Jset && break
        end
    end
    tci2 = TCI.TensorCI2{Float64}([size(t, 2) for t in tt.sitetensors])
    tci2.Iset = Iset
    tci2.Jset = Jset
    tci2.sitetensors = tt.sitetensors
    tci2.pivoterrors = pivoterrors
    tci2.maxsamplevalue = maximum(maximum.(abs, tci2.sitetensors))
    return tci2
end

"""
    TCI.estimatetrueerror(tt::TensorTrain{Float64,3}, f::DeviceFunctor; nsearch, initialpoints)

estimatetrueerror (globalsearch.jl:18-41) with every _floatingzone search
on the device, one workgroup per start point (tci_tt_floatingzone_f64).
"""
function TCI.estimatetrueerror(tt::TCI.TensorTrain{Float64,3}, f::DeviceFunctor{Float64}; nsearch::Int=100,
                               initialpoints::Union{Nothing,AbstractVector{Vector{Int}}}=nothing)
    if nsearch <= 0 && initialpoints === nothing
        error("No search is performed")
    end
    nsearch >= 0 || error("nsearch must be non-negative")
    localdims = [size(t, 2) for t in tt.sitetensors]
    if initialpoints === nothing
        initialpoints = [[rand(1:d) for d in localdims] for _ in 1:nsearch]
    end
    L = length(tt)
    n = length(initialpoints)
    pts = Vector{UInt8}(undef, n * L)
    for (q, p) in enumerate(initialpoints), t in 1:L
        pts[(q-1)*L+t] = UInt8(p[t])
    end
    errs = zeros(Float64, n)
    links, ld, flat = _packtt(tt)
    GC.@preserve f links ld flat pts errs begin
        cf = Ref(_cf(f))
        throwstatus(ccall((:tci_tt_floatingzone_f64, libtci), Cint,
                          (Ptr{Cvoid}, Ref{CFunctor}, Int32, Ptr{Int64}, Ptr{Int32}, Ptr{Float64}, Int64, Ptr{UInt8},
                           Ptr{Float64}, Int64, Float64),
                          defaultcontext().handle, cf, L, links, ld, flat, n, pts, errs, typemax(Int64), Inf),
                    "tci_tt_floatingzone_f64")
    end
    pivoterror = [(Int.(pts[(q-1)*L+1:q*L]), errs[q]) for q in 1:n]
    p = sortperm([e for (_, e) in pivoterror], rev=true)
    return unique(pivoterror[p])
end

# one set family of a finished solve (which: 0 Iset, 1 Jset, 2/3 the last
# history entry of each)
function _indexsets(b::Ptr{Cvoid}, L::Int, which::Integer)
    sets = Vector{Vector{Vector{Int}}}(undef, L)
    for site in 1:L
        len = iseven(which) ? site - 1 : L - site
        cnt = Ref{Int32}(0)
        # the first call reports the count (and a too-small-buffer status)
        ccall((:tci_batch_indexset, libtci), Cint,
              (Ptr{Cvoid}, Int32, Int32, Int32, Ref{Int32}, Ptr{UInt8}, Int64),
              b, 0, which, site, cnt, C_NULL, 0)
        buf = Vector{UInt8}(undef, max(Int(cnt[]) * len, 1))
        throwstatus(ccall((:tci_batch_indexset, libtci), Cint,
                          (Ptr{Cvoid}, Int32, Int32, Int32, Ref{Int32}, Ptr{UInt8}, Int64),
                          b, 0, which, site, cnt, buf, length(buf)), "tci_batch_indexset")
        sets[site] = [Int.(buf[(i-1)*len+1:i*len]) for i in 1:cnt[]]
    end
    return sets
end

function _totci(b::Ptr{Cvoid}, r::CResult, localdims::Vector{Int})
    L = length(localdims)
    tci = TCI.TensorCI2{Float64}(localdims)
    tci.Iset = _indexsets(b, L, 0)
    tci.Jset = _indexsets(b, L, 1)
    # Iset_history / Jset_history: the device keeps the last entry, the only
    # one sweep2site! reads (tensorci2.jl:874-877)
    hI, hJ = _indexsets(b, L, 2), _indexsets(b, L, 3)
    if any(!isempty, hI)
        push!(tci.Iset_history, hI)
        push!(tci.Jset_history, hJ)
    end
    for site in 1:L
        d3 = zeros(Int64, 3)
        ccall((:tci_batch_sitetensor, libtci), Cint,
              (Ptr{Cvoid}, Int32, Int32, Ptr{Int64}, Ptr{Float64}, Int64), b, 0, site, d3, C_NULL, 0)
        T = zeros(Float64, d3...)
        ccall((:tci_batch_sitetensor, libtci), Cint,
              (Ptr{Cvoid}, Int32, Int32, Ptr{Int64}, Ptr{Float64}, Int64), b, 0, site, d3, T, length(T))
        tci.sitetensors[site] = T
    end
    pe = zeros(Float64, r.npivoterrors)
    ccall((:tci_batch_pivoterrors, libtci), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64),
          b, 0, pe, length(pe))
    tci.pivoterrors = pe
    be = zeros(Float64, L - 1)
    ccall((:tci_batch_bonderrors, libtci), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), b, 0, be)
    tci.bonderrors = be
    tci.maxsamplevalue = r.maxsamplevalue
    ranks = zeros(Int64, r.niter)
    errors = zeros(Float64, r.niter)
    ccall((:tci_batch_history, libtci), Cint,
          (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Float64}, Ptr{Int64}), b, 0, ranks, errors, C_NULL)
    return tci, Vector{Int}(ranks), errors
end

end # module
