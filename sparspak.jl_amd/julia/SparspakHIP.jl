# SparspakHIP — the reference-side binding a Sparspak.jl maintainer would add
# to route the numeric phase of SparseSolver / SparseSpdSolver to
# libsparspak_hip.so (MI355X). Julia is not available in the build image, so
# this file is not exercised by the tests here; its ccall signatures follow
# include/sparspak_hip.h one to one.
module SparspakHIP

using Sparspak.SpkProblem: Problem

const libspk = get(ENV, "SPARSPAK_HIP_LIB",
                   joinpath(@__DIR__, "..", "sparspak_hip", "libsparspak_hip.so"))

const SPKHIP_KIND_LU = 0
const SPKHIP_KIND_SPD = 1
const SPKHIP_FP64 = 0
const SPKHIP_FP32 = 1
const SPKHIP_FP32_BF16 = 2
const SPKHIP_FP32_FP16 = 3
const SPKHIP_FP32_BF16X3 = 4

check(rc, what) = rc == 0 ? nothing :
    rc == -1 ? error("$what: zero pivot in a diagonal block") :
    rc == -3 ? error("$what: sequence error") :
    rc <= -1000 ? error("$what: HIP error $(-1000 - rc)") : error("$what: error $rc")

# ---- stateless drop-ins: same arguments as the reference's internal calls ----

# replaces _lufactor!(n, nsuper, xsuper, snode, xlindx, lindx, xlnz, lnz,
#                     xunz, unz, ipvt)   (SpkLUFactor.jl:60)
function lufactor!(n, nsuper, xsuper::Vector{Int64}, snode::Vector{Int64},
                   xlindx::Vector{Int64}, lindx::Vector{Int64}, xlnz::Vector{Int64},
                   lnz::Vector{Float64}, xunz::Vector{Int64}, unz::Vector{Float64},
                   ipvt::Vector{Int64})
    GC.@preserve xsuper snode xlindx lindx xlnz lnz xunz unz ipvt begin
        ccall((:spkhip_lufactor_f64, libspk), Int64,
              (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
               Ptr{Float64}, Ptr{Int64}, Ptr{Float64}, Ptr{Int64}),
              n, nsuper, xsuper, snode, xlindx, lindx, xlnz, lnz, xunz, unz, ipvt)
    end
end

# replaces _lulsolve!(nsuper, xsuper, xlindx, lindx, xlnz, lnz, ipiv, rhs)
function lulsolve!(nsuper, xsuper, xlindx, lindx, xlnz, lnz::Vector{Float64}, ipiv,
                   rhs::Vector{Float64})
    GC.@preserve xsuper xlindx lindx xlnz lnz ipiv rhs begin
        ccall((:spkhip_lulsolve_f64, libspk), Int64,
              (Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64},
               Ptr{Int64}, Ptr{Float64}),
              nsuper, xsuper, xlindx, lindx, xlnz, lnz, ipiv, rhs)
    end
end

# replaces _luusolve!(n, nsuper, xsuper, xlindx, lindx, xlnz, lnz, xunz, unz, rhs)
function luusolve!(n, nsuper, xsuper, xlindx, lindx, xlnz, lnz::Vector{Float64}, xunz,
                   unz::Vector{Float64}, rhs::Vector{Float64})
    GC.@preserve xsuper xlindx lindx xlnz lnz xunz unz rhs begin
        ccall((:spkhip_luusolve_f64, libspk), Int64,
              (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
               Ptr{Float64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}),
              n, nsuper, xsuper, xlindx, lindx, xlnz, lnz, xunz, unz, rhs)
    end
end

# replaces _ldltfactor! / _ldltsolve! (SpkLDLtFactor.jl:58, 266); correct
# LDL^T numerics in the reference's layout
function ldltfactor!(n, nsuper, xsuper, snode, xlindx, lindx, xlnz, lnz::Vector{Float64})
    GC.@preserve xsuper snode xlindx lindx xlnz lnz begin
        ccall((:spkhip_ldltfactor_f64, libspk), Int64,
              (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
               Ptr{Float64}), n, nsuper, xsuper, snode, xlindx, lindx, xlnz, lnz)
    end
end

function ldltsolve!(nsuper, xsuper, xlindx, lindx, xlnz, lnz::Vector{Float64},
                    rhs::Vector{Float64})
    GC.@preserve xsuper xlindx lindx xlnz lnz rhs begin
        ccall((:spkhip_ldltsolve_f64, libspk), Int64,
              (Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64},
               Ptr{Float64}), nsuper, xsuper, xlindx, lindx, xlnz, lnz, rhs)
    end
end

# ---- stateful handle: resident symbolic factor, values and factor ----

mutable struct Handle
    ptr::Ptr{Cvoid}
end

function Handle(device::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:spkhip_create, libspk), Int64, (Ptr{Ptr{Cvoid}}, Int64), r, device),
          "spkhip_create")
    h = Handle(r[])
    finalizer(h) do x
        x.ptr == C_NULL || ccall((:spkhip_destroy, libspk), Int64, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    return h
end

# after _symbolicfactor!(s.slvr) of a SparseSolver built on a SparseMatrixCSC m
function upload!(h::Handle, s, m; kind=SPKHIP_KIND_LU)
    b = s.slvr
    xunz = kind == SPKHIP_KIND_LU ? b.xunz : zeros(Int64, b.n + 1)
    GC.@preserve b xunz m begin
        rc = ccall((:spkhip_symbolic_upload, libspk), Int64,
                   (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                    Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64},
                    Ptr{Int64}),
                   h.ptr, kind, b.n, b.nsuper, b.xsuper, b.snode, b.xlindx, b.lindx, b.xlnz,
                   xunz, b.order.rperm, b.order.rinvp, m.colptr, m.rowval)
    end
    rc == -1 && error("No space for matrix element.")
    check(rc, "spkhip_symbolic_upload")
end

inmatrix!(h::Handle, m) = GC.@preserve m check(
    ccall((:spkhip_inmatrix, libspk), Int64, (Ptr{Cvoid}, Ptr{Float64}), h.ptr, m.nzval),
    "spkhip_inmatrix")

# A reference Problem (SpkProblem.jl:105-139) as the CSC arrays the upload
# takes: each column's linked list from head[c] through link, rows already
# ascending (inaij! keeps them sorted), values as stored.
function problem_csc(p)
    n = p.ncols
    colptr = ones(Int64, n + 1)
    rowval = Int64[]
    nzval = Float64[]
    sizehint!(rowval, p.nedges)
    sizehint!(nzval, p.nedges)
    for c in 1:n
        e = p.head[c]
        while e > 0
            push!(rowval, p.rowsubs[e])
            push!(nzval, p.values[e])
            e = p.link[e]
        end
        colptr[c + 1] = length(rowval) + 1
    end
    return colptr, rowval, nzval
end

# SparseSolver(p::Problem) / SparseSpdSolver(p::Problem) after
# symbolicfactor!: the symbolic factor and the Problem's pattern to the device
function upload!(h::Handle, s, p::Problem; kind=SPKHIP_KIND_LU)
    colptr, rowval, _ = problem_csc(p)
    upload!(h, s, (colptr=colptr, rowval=rowval); kind=kind)
end

# the Problem's values (the lists walked in the same order as the upload)
function inmatrix!(h::Handle, p::Problem)
    _, _, nzval = problem_csc(p)
    inmatrix!(h, (nzval=nzval,))
end

# sparsespdcondest (SpkSparseSpdSolver.jl:433-459, commented in the
# reference): (estimate of cond_1(A), estimate of ||A^-1||_1) of an SPD factor
function condest(h::Handle)
    c = Ref(0.0)
    a = Ref(0.0)
    check(ccall((:spkhip_condest, libspk), Int64, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}),
                h.ptr, c, a), "spkhip_condest")
    return c[], a[]
end

# Value sets: nvs matrices of the uploaded pattern (same ordering and
# symbolic factor, other values) factored and solved in the same launches,
# the batched form of sparspaklu!'s refactorisation (SparseCSCInterface.jl:
# 247-268). After it, inmatrix! takes the sets' matrices (their nzval
# concatenated) and solve! the sets' right-hand sides stacked (n * nvs rows).
setvaluesets!(h::Handle, nvs::Integer) = check(
    ccall((:spkhip_set_value_sets, libspk), Int64, (Ptr{Cvoid}, Int64), h.ptr, nvs),
    "spkhip_set_value_sets")

function inmatrix!(h::Handle, ms::AbstractVector)
    nz = reduce(vcat, [m.nzval for m in ms])
    GC.@preserve nz check(
        ccall((:spkhip_inmatrix, libspk), Int64, (Ptr{Cvoid}, Ptr{Float64}), h.ptr, nz),
        "spkhip_inmatrix")
end

factor!(h::Handle; precision=SPKHIP_FP64) = check(
    ccall((:spkhip_factor, libspk), Int64, (Ptr{Cvoid}, Int64), h.ptr, precision),
    "An empty problem. No matrix.")   # the message _factor! raises (SpkSparseBase.jl:387)

# spkhip_stats_t (field order of include/sparspak_hip.h)
mutable struct Stats
    n::Int64; nsuper::Int64; nfront::Int64; nlevel::Int64
    nnz_lnz::Int64; nnz_unz::Int64
    first_fail_supernode::Int64; refine_iters::Int64
    factor_flops::Float64; solve_flops::Float64; front_bytes::Float64
    t_upload_ms::Float64; t_inmatrix_ms::Float64; t_factor_ms::Float64
    t_solve_ms::Float64; t_refine_ms::Float64
    backward_error::Float64
end
Stats() = Stats(zeros(Int64, 8)..., zeros(Float64, 9)...)

function stats(h::Handle)
    st = Ref(Stats())
    check(ccall((:spkhip_stats, libspk), Int64, (Ptr{Cvoid}, Ref{Stats}), h.ptr, st),
          "spkhip_stats")
    return st[]
end

# plan knobs and the autotuner (spkhip_tuning_t; no counterpart in the reference)
mutable struct Tuning
    small_lmax::Int64
    block_cols::Int64
    upd_group::Int64
    small_wave_min::Int64
    factor_ms::Float64
end
Tuning() = Tuning(0, 0, 0, 0, 0.0)

function tuning(h::Handle)
    t = Ref(Tuning())
    check(ccall((:spkhip_get_tuning, libspk), Int64, (Ptr{Cvoid}, Ref{Tuning}), h.ptr, t),
          "spkhip_get_tuning")
    return t[]
end

settuning!(h::Handle, t::Tuning) = check(
    ccall((:spkhip_set_tuning, libspk), Int64, (Ptr{Cvoid}, Ref{Tuning}), h.ptr, Ref(t)),
    "spkhip_set_tuning")

function autotune!(h::Handle; precision=SPKHIP_FP32, reps=2)
    t = Ref(Tuning())
    check(ccall((:spkhip_autotune, libspk), Int64, (Ptr{Cvoid}, Int64, Int64, Ref{Tuning}),
                h.ptr, precision, reps, t), "spkhip_autotune")
    return t[]
end

function solve!(h::Handle, B::StridedVecOrMat{Float64}; refine=false, tol=1e-14, maxiter=10)
    nrhs = size(B, 2)
    ldb = stride(B, 2)
    GC.@preserve B begin
        rc = refine ?
            ccall((:spkhip_solve_refine, libspk), Int64,
                  (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64, Float64, Int64),
                  h.ptr, nrhs, B, ldb, tol, maxiter) :
            ccall((:spkhip_solve, libspk), Int64, (Ptr{Cvoid}, Int64, Ptr{Float64}, Int64),
                  h.ptr, nrhs, B, ldb)
    end
    if rc == -5
        st = stats(h)
        @warn "iterative refinement stopped at backward error $(st.backward_error), above tol=$tol"
    elseif rc == -6
        throw(DomainError(rc, "Triangular Solve: the solution is not finite (NaN or Inf)"))
    else
        check(rc, "Triangular Solve.")
    end
    return B
end

end # module
