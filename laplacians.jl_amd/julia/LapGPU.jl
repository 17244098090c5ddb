# LapGPU.jl — Julia host side of the MI355X solve phase of approxchol_lap.
#
# The randomized elimination that builds LDLinv stays in Laplacians.jl's own
# code (approxChol(LLmatp(a)), src/approxChol.jl:864-977, or the builder its
# ApproxCholParams names) with the caller's seed; only the solve f(b) moves to the GPU, through ccall into liblapgpu.so
# (include/lapgpu.h).  The returned closure has the signature and keyword
# semantics of the reference's (src/approxChol.jl:1624), and the component
# handling is the reference's own lapWrapComponents
# (src/solverInterface.jl:212-275).
#
# Usage:
#     using Laplacians; include("LapGPU.jl"); using .LapGPU
#     f = approxchol_lap_gpu(a; tol=1e-8)
#     x = f(b; pcgIts=its)
module LapGPU

using Laplacians
using SparseArrays
using Random

export approxchol_lap_gpu, approxchol_lap2_gpu, approxchol_sddm_gpu, harmonic_interp_gpu

const libpath = joinpath(@__DIR__, "..", "lapgpu", "liblapgpu.so")

mutable struct Ctx
    h::Ptr{Cvoid}
end

function lasterr(h::Ptr{Cvoid})
    unsafe_string(ccall((:lapgpu_last_error, libpath), Cstring, (Ptr{Cvoid},), h))
end

check(ctx::Ctx, rc::Cint, what) = rc == 0 || error("$what: ", lasterr(ctx.h))

function create(device::Integer = 0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:lapgpu_create, libpath), Cint, (Ref{Ptr{Cvoid}}, Cint), r, device)
    rc == 0 || error("lapgpu_create: ", lasterr(C_NULL))
    ctx = Ctx(r[])
    finalizer(c -> ccall((:lapgpu_destroy, libpath), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    return ctx
end

# la = lap(a), 1-based Int64 CSC exactly as Julia stores it
function upload_lap!(ctx::Ctx, la::SparseMatrixCSC{Float64,Int64})
    check(ctx, ccall((:lapgpu_upload_csc, libpath), Cint,
                     (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Cint),
                     ctx.h, la.n, la.colptr, la.rowval, la.nzval, 1), "lapgpu_upload_csc")
end

# LDLinv{Int64,Float64} (src/approxCholTypes.jl:73-79), 1-based
function upload_ldli!(ctx::Ctx, ldli)
    n = length(ldli.d)
    check(ctx, ccall((:lapgpu_upload_ldli, libpath), Cint,
                     (Ptr{Cvoid}, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                      Ptr{Float64}, Ptr{Float64}, Cint),
                     ctx.h, n, length(ldli.fval), ldli.col, ldli.colptr, ldli.rowval,
                     ldli.fval, ldli.d, 1), "lapgpu_upload_ldli")
end

function gpu_solve(ctx::Ctx, b::Vector{Float64}, tol, maxits, maxtime, verbose,
                   pcgIts, stag_test)
    x = zeros(length(b))
    its = Ref{Int64}(0)
    nr = Ref{Float64}(0.0)
    stop = Ref{Cint}(0)
    mi = isinf(maxits) ? Int64(-1) : Int64(maxits)
    check(ctx, ccall((:lapgpu_solve, libpath), Cint,
                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Int64,
                      Float64, Cint, Cint, Ref{Int64}, Ref{Float64}, Ref{Cint}),
                     ctx.h, b, x, length(b), tol, mi, Float64(maxtime), stag_test,
                     verbose ? 1 : 0, its, nr, stop), "lapgpu_solve")
    if length(pcgIts) > 0
        pcgIts[1] = its[]
    end
    return x
end

# The reference's own factor builders for each ApproxCholParams
# (approxchol_lapGreedy / Wdeg / Random / Given, src/approxChol.jl:1581-1790),
# fixed-order factors relabelled to the graph's vertices as the reference's
# ldlip is
function build_ldli(a::SparseMatrixCSC, params)
    if params.order == :deg
        if params.split >= 1 && params.merge < 1
            return Laplacians.approxChol(Laplacians.LLmatp(a, params.split), params.split)
        elseif params.split >= 1 && params.merge >= 1
            return Laplacians.approxChol(Laplacians.LLmatp(a, params.split), params.split,
                                         params.merge)
        end
        return Laplacians.approxChol(Laplacians.LLmatp(a))
    elseif params.order == :wdeg || params.order == :random
        n = size(a, 1)
        p = params.order == :wdeg ? sortperm(vec(sum(a, dims = 1)) .* (1 .+ rand(n))) :
            randperm(n)
        ldli = Laplacians.approxChol(Laplacians.LLMatOrd(a, p))
        return Laplacians.LDLinv(p[ldli.col], ldli.colptr, p[ldli.rowval], ldli.fval,
                                 ldli.d[invperm(p)])
    end
    return Laplacians.approxChol(Laplacians.LLMatOrd(a))
end

# approxchol_lap1 for a connected graph: the reference's factor on the host,
# the solve on the GPU (src/approxChol.jl:1791-1851)
function approxchol_lap1_gpu(a::SparseMatrixCSC; tol::Real = 1e-6, maxits = 1000,
                             maxtime = Inf, verbose = false, pcgIts = Int[],
                             params = ApproxCholParams(), seed::Integer = 1,
                             device::Integer = 0)
    la = lap(a)
    Random.seed!(seed)
    ldli = build_ldli(a, params)
    ctx = create(device)
    upload_lap!(ctx, la)
    upload_ldli!(ctx, ldli)
    tol_, maxits_, maxtime_, verbose_, pcgIts_ = tol, maxits, maxtime, verbose, pcgIts
    f(b; tol = tol_, maxits = maxits_, maxtime = maxtime_, verbose = verbose_,
      pcgIts = pcgIts_) = gpu_solve(ctx, b, tol, maxits, maxtime, verbose, pcgIts,
                                    params.stag_test)
    return f
end

"""
    solver = approxchol_lap_gpu(a; tol=1e-6, maxits=1000, maxtime=Inf, verbose=false, pcgIts=Int[], params=ApproxCholParams())

Drop-in for `approxchol_lap` (src/approxChol.jl:1522-1545) whose solves run on an MI355X.
"""
function approxchol_lap_gpu(a::SparseMatrixCSC{Tv,Ti}; tol::Real = 1e-6, maxits = 1000,
                            maxtime = Inf, verbose = false, pcgIts = Int[],
                            params = ApproxCholParams(), kwargs...) where {Tv,Ti}
    if minimum(a.nzval) < 0
        error("Adjacency matrix can not have negative edge weights")
    end
    return Laplacians.lapWrapComponents((a; kw...) -> approxchol_lap1_gpu(a; kw..., kwargs...), a,
                                        verbose = verbose, tol = tol, maxits = maxits,
                                        maxtime = maxtime, pcgIts = pcgIts, params = params)
end

# approxchol_lap2 (src/approxChol.jl:1559-1579), with the same parameters
function approxchol_lap2_gpu(a::SparseMatrixCSC{Tv,Ti}; tol::Real = 1e-6, maxits = 1000,
                             maxtime = Inf, verbose = false, pcgIts = Int[],
                             params = ApproxCholParams(), kwargs...) where {Tv,Ti}
    if minimum(a.nzval) < 0
        error("Adjacency matrix can not have negative edge weights")
    end
    return Laplacians.lapWrapComponents((a; kw...) -> approxchol_lap1_gpu(a; kw..., kwargs...), a,
                                        verbose = verbose, tol = tol, maxits = maxits,
                                        maxtime = maxtime, pcgIts = pcgIts,
                                        params = ApproxCholParams(:deg, params.stag_test, 2))
end

approxchol_sddm_gpu = Laplacians.sddmWrapLap(approxchol_lap_gpu)

# harmonic_interp (src/harmonic.jl:16-33) with the SDDM solve on the GPU
function harmonic_interp_gpu(adj_mat, S::Vector, vals::Vector; tol = 1e-6)
    n = size(adj_mat, 1)
    b = zeros(n)
    b[S] = vals
    inds = ones(Bool, n)
    inds[S] .= false
    la = lap(adj_mat)
    f = approxchol_sddm_gpu(la[inds, inds]; tol = tol)
    x = copy(b)
    x[inds] = f((-la * b)[inds])
    return x
end

# ---- entry points beyond the reference (include/lapgpu.h) ----------------

# k right-hand sides (the columns of B) in every launch, each with the
# closure's stopping rules; returns X and per-column iteration counts
function solve_multi(ctx::Ctx, B::Matrix{Float64}; tol = 1e-6, maxits = 1000,
                     maxtime = Inf, stag_test = 5, flags = 0)
    n, k = size(B)
    X = zeros(n, k)
    its = zeros(Int64, k)
    nr = zeros(Float64, k)
    stop = zeros(Cint, k)
    mi = isinf(maxits) ? Int64(-1) : Int64(maxits)
    check(ctx, ccall((:lapgpu_solve_multi, libpath), Cint,
                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Int64, Float64, Int64,
                      Float64, Cint, Cint, Ptr{Int64}, Ptr{Float64}, Ptr{Cint}),
                     ctx.h, B, X, n, k, tol, mi, Float64(maxtime), stag_test, flags, its, nr,
                     stop), "lapgpu_solve_multi")
    return X, its
end

# FP32 SpMV and LDLinv apply with FP64 dots, alpha/beta and a residual
# replacement every `replace_every` iterations (mode 1), or FP64 (mode 0)
set_precision!(ctx::Ctx, mode::Integer, replace_every::Integer = 20) =
    check(ctx, ccall((:lapgpu_set_precision, libpath), Cint, (Ptr{Cvoid}, Cint, Cint),
                     ctx.h, mode, replace_every), "lapgpu_set_precision")

# per-graph schedule autotuning (k > 1: also the k-column multi-RHS kernels;
# groups: also the level-group budget); returns the apply times before and
# after (ms)
function autotune!(ctx::Ctx; k::Integer = 0, groups::Bool = false)
    t = zeros(Float64, 4)
    check(ctx, ccall((:lapgpu_autotune, libpath), Cint, (Ptr{Cvoid}, Int64, Cint, Ptr{Float64}),
                     ctx.h, k, groups ? 1 : 0, t), "lapgpu_autotune")
    return t
end

# level-group budget of the schedule (regroups an uploaded factor)
set_group_budget!(ctx::Ctx, depth::Integer, growth::Real, extra::Integer) =
    check(ctx, ccall((:lapgpu_set_group_budget, libpath), Cint,
                     (Ptr{Cvoid}, Cint, Float64, Int64), ctx.h, depth, growth, extra),
          "lapgpu_set_group_budget")

# the solving thread sleeps between event queries (blocking = true)
set_wait_mode!(ctx::Ctx, blocking::Bool) =
    check(ctx, ccall((:lapgpu_set_wait_mode, libpath), Cint, (Ptr{Cvoid}, Cint), ctx.h,
                     blocking ? 1 : 0), "lapgpu_set_wait_mode")

end # module
