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

# QubitsOperator -> merged keys (hamiltonian/util.jl:6-15)
function pack(h::QubitsOperator)
    terms = CTerm[]
    pool = ComplexF64[]
    for (k, v) in h.data
        length(k) <= 5 || throw(ArgumentError("terms longer than 5 sites are not supported on this path"))
        m = VQC._get_mat(length(k), v)
        push!(terms, CTerm(length(k), pad5(collect(k) .- 1), length(pool)))
        append!(pool, vec(Matrix{ComplexF64}(m)))
    end
    return terms, pool
end

function apply!(circuit::QCircuit, s::GPUState)
    ops, pool, _ = pack(circuit, s.isdm)
    isempty(ops) && return s
    check(ccall((:vqc_apply_circuit, libvqc), Cint, (Ptr{Cvoid}, Cint, Ptr{COp}, Ptr{ComplexF64}), s.h, length(ops), ops, pool))
    return s
end

Base.:*(circuit::QCircuit, s::GPUState) = apply!(circuit, copy(s))

# ---- expectation (expecs.jl:12-42, expecs_dm.jl:7-27) -------------------------
function VQC.expectation(h::QubitsOperator, s::GPUState)
    r = zero(ComplexF64)
    out = zeros(Float64, 2)
    for (k, v) in h.data
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
    return r
end

# ---- energy and adjoint gradient -----------------------------------------------
"E = expectation(h, circuit * s0) and dRe(E)/dθ for the active parameters, in circuit order."
function energy_and_gradient(h::QubitsOperator, circuit::QCircuit, s0::GPUState)
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

end # module
