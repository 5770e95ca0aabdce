# VQCHip.jl -- Julia binding of libvqc_hip.so (include/vqc_hip.h) for VQC.jl.
#
# Drops in under VQC's StateVector / DensityMatrix / QCircuit / expectation API:
# states live in HBM behind a library handle, every kernel call goes through
# ccall, and Zygote reaches the hand-written adjoint kernel through
# ChainRulesCore rrules. Written against VQC.jl and QuantumCircuits.jl as used by
# the reference; there is no Julia toolchain in the build image, so this file is
# not exercised by the tests (the Python mirror in ../vqchip exercises the same
# C ABI).
module VQCHip

using LinearAlgebra
using ChainRulesCore
using QuantumCircuits, QuantumCircuits.Gates
import VQC

const libvqc = joinpath(@__DIR__, "..", "libvqc_hip.so")

struct VQCError <: Exception
    code::Cint
    msg::String
end

function check(rc::Cint)
    rc == 0 && return nothing
    throw(VQCError(rc, unsafe_string(ccall((:vqc_last_error, libvqc), Cstring, ()))))
end

# ---- C structs (layouts of vqc_op_t, vqc_term_t) ----------------------------
struct COp
    kind::Int32
    nq::Int32
    pos::NTuple{5,Int32}
    nparams::Int32
    slot::NTuple{3,Int32}
    reserved::Int32
    mat::Int64
    aux::Int64
    gen::NTuple{3,Int64}
end

struct CTerm
    nq::Int32
    pos::NTuple{5,Int32}
    mat::Int64
end

pad5(v) = ntuple(i -> i <= length(v) ? Int32(v[i]) : Int32(0), 5)

# ---- device states ----------------------------------------------------------
mutable struct GPUState
    h::Ptr{Cvoid}
    nqubits::Int
    isdm::Bool
    function GPUState(n::Int, isdm::Bool)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((isdm ? :vqc_dm_create : :vqc_sv_create, libvqc), Cint, (Cint, Ptr{Ptr{Cvoid}}), n, r))
        s = new(r[], n, isdm)
        finalizer(x -> ccall((:vqc_state_destroy, libvqc), Cint, (Ptr{Cvoid},), x.h), s)
        return s
    end
end

nbits(s::GPUState) = s.isdm ? 2 * s.nqubits : s.nqubits

"Upload a VQC.StateVector / DensityMatrix (statevector.jl:1-10, densitymatrix.jl:1-10)."
function GPUState(x::VQC.StateVector)
    s = GPUState(VQC.nqubits(x), false)
    d = Vector{ComplexF64}(VQC.storage(x))
    check(ccall((:vqc_state_upload, libvqc), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, UInt64), s.h, d, length(d)))
    return s
end
function GPUState(x::VQC.DensityMatrix)
    s = GPUState(VQC.nqubits(x), true)
    d = Vector{ComplexF64}(x.data)
    check(ccall((:vqc_state_upload, libvqc), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, UInt64), s.h, d, length(d)))
    return s
end

function download(s::GPUState)
    d = Vector{ComplexF64}(undef, 1 << nbits(s))
    check(ccall((:vqc_state_download, libvqc), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, UInt64), s.h, d, length(d)))
    return s.isdm ? VQC.DensityMatrix(d, s.nqubits) : VQC.StateVector(d, s.nqubits)
end

Base.copy(s::GPUState) = (t = GPUState(s.nqubits, s.isdm);
                          check(ccall((:vqc_state_copy, libvqc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), t.h, s.h)); t)

# ---- apply! (generic.jl:6-31; apply_qmaps.jl:8-16) ---------------------------
function apply!(g::Gate, s::GPUState)
    pos = Int32.(collect(ordered_positions(g)) .- 1)
    m = Matrix{ComplexF64}(ordered_mat(g))
    if s.isdm
        check(ccall((:vqc_dm_apply_gate, libvqc), Cint, (Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{ComplexF64}), s.h, length(pos), pos, m))
    else
        check(ccall((:vqc_sv_apply_dense, libvqc), Cint, (Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{ComplexF64}), s.h, length(pos), pos, m))
    end
    return s
end

# named kernels (src/applygates/specific/*.jl): permutations and phases, no matrix
const NAMED = Dict{DataType,Tuple{Cint,Bool}}(  # kind, has angle
    XGate => (1, false), PHASEGate => (2, true), CZGate => (3, false), CNOTGate => (4, false),
    CPHASEGate => (5, true), SWAPGate => (6, false), iSWAPGate => (7, false), FREDKINGate => (9, false),
    CCPHASEGate => (10, true))

function apply!(g::Union{XGate,PHASEGate,CZGate,CNOTGate,CPHASEGate,SWAPGate,iSWAPGate,FREDKINGate,CCPHASEGate},
                s::GPUState)
    s.isdm && return invoke(apply!, Tuple{Gate,GPUState}, g, s)
    kind, has_angle = NAMED[typeof(g)]
    pos = Int32.(collect(positions(g)) .- 1)   # the gate's own order: control first (czgate.jl:30-40)
    phi = has_angle ? Float64(parameters(g)[1]) : 0.0
    check(ccall((:vqc_sv_apply_named, libvqc), Cint, (Ptr{Cvoid}, Cint, Ptr{Int32}, Cdouble), s.h, kind, pos, phi))
    return s
end

function apply!(x::QuantumMap, s::GPUState)
    s.isdm || throw(ArgumentError("channels act on density matrices"))
    pos = Int32.(collect(ordered_positions(x)) .- 1)
    S = Matrix{ComplexF64}(ordered_supermat(x))
    check(ccall((:vqc_dm_apply_superop, libvqc), Cint, (Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{ComplexF64}, Cint),
                s.h, length(pos), pos, S, 0))
    return s
end

# circuit -> (ops, pool, nparams); generators G = U' dU/dθ = item' for item in
# differentiate(m) (circuitdiff.jl:29 evaluates real(<y|item|Δ>) = real(<Δ|item'|y>))
function pack(circuit::QCircuit, isdm::Bool)
    ops = COp[]
    pool = ComplexF64[]
    slot = 0
    addm(m) = (off = length(pool); append!(pool, vec(Matrix{ComplexF64}(m))); off)
    for g in circuit
        pos = collect(ordered_positions(g)) .- 1
        if g isa QuantumMap
            S = Matrix{ComplexF64}(ordered_supermat(g))
            push!(ops, COp(1, length(pos), pad5(pos), 0, (-1, -1, -1), 0, addm(S), addm(inv(S)), (-1, -1, -1)))
        else
            items = nparameters(g) > 0 ? differentiate(g) : []
            gens = Int64[addm(adjoint(Matrix{ComplexF64}(it))) for it in items]
            slots = Int32[]
            for _ in gens
                push!(slots, slot)
                slot += 1
            end
            push!(ops, COp(0, length(pos), pad5(pos), length(gens), ntuple(i -> i <= length(slots) ? slots[i] : Int32(-1), 3),
                           0, addm(ordered_mat(g)), -1, ntuple(i -> i <= length(gens) ? gens[i] : Int64(-1), 3)))
        end
    end
    return ops, pool, slot
end

# QubitsOperator -> merged keys of at most 5 sites (hamiltonian/util.jl:6-15), and
# the longer terms as product terms (vqc_qterm_t: coefficient, sites, 2x2 factors)
struct CQTerm
    nsites::Int32
    reserved::Int32
    pos::Int64
    mat::Int64
    re::Float64
    im::Float64
end

function pack(h::QubitsOperator)
    terms = CTerm[]
    pool = ComplexF64[]
    for (k, v) in h.data
        length(k) <= 5 || continue
        m = VQC._get_mat(length(k), v)
        push!(terms, CTerm(length(k), pad5(collect(k) .- 1), length(pool)))
        append!(pool, vec(Matrix{ComplexF64}(m)))
    end
    return terms, pool
end

# the terms of more than 5 sites, one by one (the per-term route of
# _apply_qterm_util!, apply_qterms.jl:20-25,47-59)
function pack_long(h::QubitsOperator)
    qt = CQTerm[]
    pos = Int32[]
    pool = ComplexF64[]
    for (k, v) in h.data
        length(k) > 5 || continue
        for item in v  # (oplist, coeff), as apply_qterms.jl:22 reads them
            c = ComplexF64(item[2])
            push!(qt, CQTerm(length(k), 0, length(pos), length(pool), real(c), imag(c)))
            append!(pos, Int32.(collect(k) .- 1))
            for m in item[1]
                append!(pool, vec(Matrix{ComplexF64}(m)))
            end
        end
    end
    return qt, pos, pool
end

function apply!(circuit::QCircuit, s::GPUState)
    ops, pool, _ = pack(circuit, s.isdm)
    isempty(ops) && return s
    check(ccall((:vqc_apply_circuit, libvqc), Cint, (Ptr{Cvoid}, Cint, Ptr{COp}, Ptr{ComplexF64}), s.h, length(ops), ops, pool))
    return s
end

Base.:*(circuit::QCircuit, s::GPUState) = apply!(circuit, copy(s))

# ---- expectation (expecs.jl:12-42, expecs_dm.jl:7-27) -------------------------
# Keys of at most 5 sites take one dense sweep each; longer terms the per-term
# route (dot(psi, m * psi), expecs.jl:40; tr(m * rho), expecs_dm.jl:25) in the
# product-term kernels, which take Pauli and sigma+- strings in one pass each.
function VQC.expectation(h::QubitsOperator, s::GPUState)
    r = zero(ComplexF64)
    out = zeros(Float64, 2)
    for (k, v) in h.data
        length(k) <= 5 || continue
        pos = Int32.(collect(k) .- 1)
        m = Matrix{ComplexF64}(VQC._get_mat(length(k), v))
        if s.isdm
            check(ccall((:vqc_dm_expec_dense, libvqc), Cint, (Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{ComplexF64}, Ptr{Float64}),
                        s.h, length(pos), pos, m, out))
        else
            check(ccall((:vqc_sv_expec_dense, libvqc), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Int32}, Ptr{ComplexF64}, Ptr{Float64}),
                        s.h, C_NULL, length(pos), pos, m, out))
        end
        r += complex(out[1], out[2])
    end
    qt, qpos, qpool = pack_long(h)
    if !isempty(qt)
        if s.isdm
            check(ccall((:vqc_dm_expec_qterms, libvqc), Cint,
                        (Ptr{Cvoid}, Cint, Ptr{CQTerm}, Ptr{Int32}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{Float64}),
                        s.h, length(qt), qt, qpos, qpool, C_NULL, out))
        else
            check(ccall((:vqc_sv_expec_qterms, libvqc), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{CQTerm}, Ptr{Int32}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{Float64}),
                        s.h, C_NULL, length(qt), qt, qpos, qpool, C_NULL, out))
        end
        r += complex(out[1], out[2])
    end
    return r
end
VQC.expectation(h::QubitsTerm, s::GPUState) = VQC.expectation(QubitsOperator(h), s)

# H * psi ((m::QubitsOperator)(v), apply_qterms.jl:13-28; on a density matrix the
# keys act on the row bits, apply_qterms_dm.jl:11-27); long terms accumulate on
# top of the dense keys
function Base.:*(h::QubitsOperator, s::GPUState)
    terms, pool = pack(h)
    out = GPUState(s.nqubits, s.isdm)
    check(ccall((:vqc_sv_apply_terms, libvqc), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{CTerm}, Ptr{ComplexF64}, Ptr{Float64}),
                s.h, out.h, length(terms), terms, pool, C_NULL))
    qt, qpos, qpool = pack_long(h)
    if !isempty(qt)
        check(ccall((:vqc_sv_apply_qterms, libvqc), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{CQTerm}, Ptr{Int32}, Ptr{ComplexF64}, Cint),
                    s.h, out.h, length(qt), qt, qpos, qpool, 1))
    end
    return out
end
Base.:*(h::QubitsTerm, s::GPUState) = QubitsOperator(h) * s

# y <- a x + y on the device (vqc_state_axpy), and the sum Zygote takes of two
# tangents of one state
axpy!(a::Number, x::GPUState, y::GPUState) = (c = ComplexF64(a);
    check(ccall((:vqc_state_axpy, libvqc), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Ptr{Cvoid}), y.h, real(c), imag(c), x.h)); y)
scale!(s::GPUState, a::Number) = (c = ComplexF64(a);
    check(ccall((:vqc_state_scale, libvqc), Cint, (Ptr{Cvoid}, Cdouble, Cdouble), s.h, real(c), imag(c))); s)
Base.:+(a::GPUState, b::GPUState) = axpy!(1.0, b, copy(a))

# The expectation pullback (@adjoint expectation, expecdiff.jl:4-68): for a state
# vector (conj(z) H + z H') psi (expecdiff.jl:37-56 for long terms: the same
# per-term sweeps); for a density matrix z H' applied to the identity
# (expecdiff.jl:24-25,64-65). With the rrule of * below, Zygote differentiates
# the reference's own loss c -> real(expectation(H, c * psi0)) unchanged.
function expectation_pullback(h::QubitsOperator, s::GPUState, z::Number)
    z = ComplexF64(z)
    if s.isdm
        ident = GPUState(s.nqubits, true)
        check(ccall((:vqc_dm_identity, libvqc), Cint, (Ptr{Cvoid},), ident.h))
        return scale!(h' * ident, z)
    end
    out = scale!(h * s, conj(z))
    return axpy!(z, h' * s, out)
end

function ChainRulesCore.rrule(::typeof(VQC.expectation), h::QubitsOperator, s::GPUState)
    pullback(z) = (NoTangent(), NoTangent(), expectation_pullback(h, s, unthunk(z)))
    return VQC.expectation(h, s), pullback
end
function ChainRulesCore.rrule(::typeof(VQC.expectation), h::QubitsTerm, s::GPUState)
    pullback(z) = (NoTangent(), NoTangent(), expectation_pullback(QubitsOperator(h), s, unthunk(z)))
    return VQC.expectation(h, s), pullback
end

# Pauli strings beyond 5 sites (the dot(psi, H*psi) fallback of expecs.jl:40)
struct CPauli
    x::UInt64
    z::UInt64
    re::Float64
    im::Float64
end

function pauli_expectation(strings::Vector{Tuple{Dict{Int,Char},ComplexF64}}, s::GPUState)
    ps = CPauli[]
    for (ops, c) in strings
        x = UInt64(0); z = UInt64(0)
        for (q, p) in ops
            b = UInt64(1) << (q - 1)
            p in ('X', 'Y') && (x |= b)
            p in ('Z', 'Y') && (z |= b)
        end
        push!(ps, CPauli(x, z, real(c), imag(c)))
    end
    per = zeros(Float64, 2 * length(ps))
    tot = zeros(Float64, 2)
    check(ccall((:vqc_sv_expec_pauli, libvqc), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{CPauli}, Ptr{Float64}, Ptr{Float64}),
                s.h, C_NULL, length(ps), ps, per, tot))
    return complex(tot[1], tot[2])
end

# ---- energy and adjoint gradient -----------------------------------------------
"E = expectation(h, circuit * s0) and dRe(E)/dθ for the active parameters, in circuit order."
function energy_and_gradient(h::QubitsOperator, circuit::QCircuit, s0::GPUState)
    if !isempty(pack_long(h)[1])
        # terms of more than 5 sites: the reference's chain, y = circuit * s0, the
        # expectation pullback with z = 1 and the circuit pullback of it
        y = circuit * s0
        g, _ = circuit_vjp(circuit, s0, expectation_pullback(h, y, 1.0))
        return VQC.expectation(h, y), g
    end
    ops, pool, np = pack(circuit, s0.isdm)
    terms, hpool = pack(h)
    e = zeros(Float64, 2)
    g = zeros(Float64, max(np, 1))
    f = s0.isdm ? :vqc_dm_energy_grad : :vqc_sv_energy_grad
    check(ccall((f, libvqc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{COp}, Ptr{ComplexF64}, Cint, Ptr{CTerm}, Ptr{ComplexF64}, Cint, Ptr{Float64},
                 Ptr{Float64}, Ptr{Cvoid}, Cint),
                s0.h, length(ops), ops, pool, length(terms), terms, hpool, np, e, g, C_NULL, 0))
    return complex(e[1], e[2]), g[1:np]
end

# Zygote sees the loss c -> real(expectation(h, c * s0)) as one primitive whose
# pullback is the adjoint kernel (replaces @adjoint *(circuit, x) at
# circuitdiff.jl:2-8 and @adjoint expectation at expecdiff.jl:28,68 for this loss).
energy(h::QubitsOperator, circuit::QCircuit, s0::GPUState) = real(energy_and_gradient(h, circuit, s0)[1])

function ChainRulesCore.rrule(::typeof(energy), h::QubitsOperator, circuit::QCircuit, s0::GPUState)
    E, g = energy_and_gradient(h, circuit, s0)
    pullback(Δ) = (NoTangent(), NoTangent(), Δ .* g, NoTangent())
    return real(E), pullback
end

# The pullback of y = circuit * x for any loss: replaces @adjoint *(circuit, x)
# with back_propagate (circuitdiff.jl:2-8, 23-78). Returns the circuit's flat
# gradient (active parameters in circuit order, as the reference's pullback does)
# and the cotangent on x, circuit' * Δy, as a GPUState.
function circuit_vjp(circuit::QCircuit, x::GPUState, Δy::GPUState)
    (Δy.nqubits == x.nqubits && Δy.isdm == x.isdm) || throw(DimensionMismatch("Δy must have the shape of x"))
    ops, pool, np = pack(circuit, x.isdm)
    g = zeros(Float64, max(np, 1))
    Δx = GPUState(x.nqubits, x.isdm)
    check(ccall((:vqc_circuit_vjp, libvqc), Cint,
                (Ptr{Cvoid}, Cint, Ptr{COp}, Ptr{ComplexF64}, Cint, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                x.h, length(ops), ops, pool, np, Δy.h, g, Δx.h, C_NULL, 0))
    return g[1:np], Δx
end

function ChainRulesCore.rrule(::typeof(*), circuit::QCircuit, x::GPUState)
    y = circuit * x
    pullback(Δy::GPUState) = begin
        g, Δx = circuit_vjp(circuit, x, Δy)
        (NoTangent(), g, Δx)
    end
    return y, pullback
end

# ---- batches and serving (no counterpart in the reference) ---------------------
struct CInstance
    nqubits::Int32
    nops::Int32
    op_begin::Int64
    nterms::Int32
    nparams::Int32
    term_begin::Int64
    grad_begin::Int64
end

"Energies and gradients of many (circuit, H, n) instances in one persistent-kernel launch."
function batch_energy_and_gradient(instances::Vector{<:Tuple{QCircuit,QubitsOperator,Int}})
    ops = COp[]; terms = CTerm[]; pool = ComplexF64[]; hpool = ComplexF64[]; inst = CInstance[]
    g = 0
    for (c, h, n) in instances
        o, p, np = pack(c, false)
        t, hp = pack(h)
        shift(x) = x < 0 ? x : x + length(pool)
        append!(ops, [COp(a.kind, a.nq, a.pos, a.nparams, a.slot, a.reserved, shift(a.mat), shift(a.aux), map(shift, a.gen)) for a in o])
        append!(terms, [CTerm(a.nq, a.pos, a.mat + length(hpool)) for a in t])
        push!(inst, CInstance(n, length(o), length(ops) - length(o), length(t), np, length(terms) - length(t), g))
        append!(pool, p); append!(hpool, hp)
        g += np
    end
    b = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:vqc_batch_create, libvqc), Cint,
                (Cint, Ptr{CInstance}, Ptr{COp}, Int64, Ptr{CTerm}, Int64, Int64, Int64, Cint, Ptr{Ptr{Cvoid}}),
                length(inst), inst, ops, length(ops), terms, length(terms), length(pool), length(hpool), 0, b))
    e = zeros(Float64, 2 * length(inst)); gr = zeros(Float64, max(g, 1))
    try
        check(ccall((:vqc_batch_run, libvqc), Cint, (Ptr{Cvoid}, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{Float64}, Ptr{Float64}),
                    b[], pool, hpool, e, gr))
    finally
        ccall((:vqc_batch_destroy, libvqc), Cint, (Ptr{Cvoid},), b[])
    end
    return [complex(e[2i-1], e[2i]) for i in 1:length(inst)], gr[1:g]
end

end # module
