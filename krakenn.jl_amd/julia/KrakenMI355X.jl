# KrakenMI355X.jl — Julia side of the MI355X backend for Kraken.jl's stream–collide path.
#
# A thin `ccall` layer over libkraken_mi355x.so (include/kraken_mi355x.h).  It defines the
# backend value `MI355XBackend()` that Kraken's drivers receive as `backend=`, and the
# methods of the hot-path drivers for it.  No KernelAbstractions, no AMDGPU.jl: every
# time step runs in the library's HIP kernels.  Julia is not available in this build's
# container, so this file is exercised only through the identical C ABI from Python.
module KrakenMI355X

using Kraken
import Kraken: run_cavity_2d, LBMConfig, D2Q9, D3Q19

const libkmi = get(ENV, "KRAKEN_MI355X_LIB",
                   joinpath(@__DIR__, "..", "kraken_mi355x", "libkraken_mi355x.so"))

# ---- C structs (layout of kmi_face / kmi_desc) --------------------------------------
struct KmiFace
    kind::Int32
    thermal::Int32
    u::NTuple{3,Float64}
    rho::Float64
    T::Float64
end

struct KmiDesc
    stencil::Int32; storage::Int32; collision::Int32; thermal::Int32
    nx::Int64; ny::Int64; nz::Int64
    omega::Float64; s_plus::Float64; s_minus::Float64
    s_e::Float64; s_eps::Float64; s_q::Float64; s_nu::Float64
    omega_T::Float64; beta_g::Float64; T_ref::Float64
    coupling::Int32; n_bc_order::Int32; bc_order::NTuple{6,Int32}
    faces::NTuple{6,KmiFace}
    forcing::Int32; force::NTuple{3,Float64}
    solid_mode::Int32
end

const FACE_PERIODIC, FACE_WALL, FACE_MOVING_WALL, FACE_ZOUHE_VELOCITY, FACE_ZOUHE_PRESSURE = 0, 1, 2, 3, 4
const FACE_ZOUHE_VELOCITY_PROFILE, FACE_ZOUHE_PRESSURE_PROFILE = 5, 6
const STORAGE = Dict(Float64 => Int32(0), Float32 => Int32(1))
const FORCE_NONE, FORCE_GUO, FORCE_GUO_PLAIN_MOMENTS, FORCE_GUO_FIELD = 0, 1, 2, 3
const HOST_F64, HOST_F32, HOST_STORAGE = 0, 1, 2

lasterr(ctx) = unsafe_string(ccall((:kmi_last_error, libkmi), Cstring, (Ptr{Cvoid},), ctx))

function check(rc, ctx=C_NULL)
    rc == 0 && return nothing
    msg = lasterr(ctx)
    # kmi_status 1 (bad argument) and 5 (unsupported) are the caller's fault
    rc in (1, 5) ? throw(ArgumentError(msg)) : error(msg)
end

# ---- the backend value ----------------------------------------------------------------
"""
    MI355XBackend(device=0)

Backend value for `run_simulation(...; backend=MI355XBackend())` and the drivers.  Raises
`ArgumentError` when no MI355X is visible — never a silent CPU run
(src/simulation_runner.jl:123-149).
"""
struct MI355XBackend
    ctx::Ptr{Cvoid}
end

function MI355XBackend(device::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:kmi_ctx_create, libkmi), Cint, (Cint, Ref{Ptr{Cvoid}}), device, r)
    rc == 0 || throw(ArgumentError("Backend mi355x is unavailable: " * lasterr(C_NULL)))
    return MI355XBackend(r[])
end

mutable struct Lattice
    h::Ptr{Cvoid}
    backend::MI355XBackend
    dims::NTuple{3,Int}
    Q::Int
end

function Lattice(b::MI355XBackend, d::KmiDesc)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:kmi_lattice_create, libkmi), Cint, (Ptr{Cvoid}, Ref{KmiDesc}, Ref{Ptr{Cvoid}}),
                b.ctx, d, r), b.ctx)
    lat = Lattice(r[], b, (Int(d.nx), Int(d.ny), Int(d.nz)), d.stencil == 0 ? 9 : 19)
    finalizer(l -> ccall((:kmi_lattice_destroy, libkmi), Cint, (Ptr{Cvoid},), l.h), lat)
    return lat
end

step!(l::Lattice, n::Integer) =
    check(ccall((:kmi_step, libkmi), Cint, (Ptr{Cvoid}, Int64), l.h, n), l.backend.ctx)

upload!(l::Lattice, f::Array{Float64}) =
    check(ccall((:kmi_upload, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32),
                l.h, 0, f, 0), l.backend.ctx)

function download(l::Lattice)
    f = Array{Float64}(undef, l.dims[1], l.dims[2], (l.Q == 19 ? (l.dims[3],) : ())..., l.Q)
    check(ccall((:kmi_download, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int32),
                l.h, 0, f, 0), l.backend.ctx)
    return f
end

function moments(l::Lattice)
    n = l.Q == 9 ? (l.dims[1], l.dims[2]) : l.dims
    ρ, ux, uy = zeros(n), zeros(n), zeros(n)
    uz = l.Q == 19 ? zeros(n) : nothing
    check(ccall((:kmi_moments, libkmi), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                l.h, ρ, ux, uy, uz === nothing ? C_NULL : uz, C_NULL), l.backend.ctx)
    return (ρ=ρ, ux=ux, uy=uy, uz=uz)
end

# Per-node force of a FORCE_GUO_FIELD lattice: the Fx_field, Fy_field(, Fz_field) arrays of
# collide_guo_field_2d!/_3d! (collide_guo_2d.jl:93-160, collide_guo_3d.jl:199-332).
function set_force_field!(l::Lattice, Fx::Array{Float64}, Fy::Array{Float64},
                          Fz::Union{Nothing,Array{Float64}}=nothing)
    check(ccall((:kmi_set_force_field, libkmi), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                l.h, Fx, Fy, Fz === nothing ? C_NULL : Fz), l.backend.ctx)
end

# The stored populations verbatim (F64 storage: Float64, F32: Float32), for checkpoints of a
# lattice (src/platform/state.jl) that restart bit for bit; set_steps! restores the cycle.
function download_raw(l::Lattice, ::Type{E}) where {E<:Union{Float64,Float32}}
    f = Array{E}(undef, l.dims[1], l.dims[2], (l.Q == 19 ? (l.dims[3],) : ())..., l.Q)
    check(ccall((:kmi_download, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{E}, Int32),
                l.h, 0, f, HOST_STORAGE), l.backend.ctx)
    return f
end
set_steps!(l::Lattice, n::Integer) =
    check(ccall((:kmi_lattice_set_steps, libkmi), Cint, (Ptr{Cvoid}, Int64), l.h, n), l.backend.ctx)

# Launch choice for the lattice (results do not depend on it): the autotuner times the
# single-step shapes and, on fully periodic D3Q19 BGK/TRT lattices, the two- and three-step
# kernels; set_temporal! fixes a multi-step tile by hand (0: single steps).
function autotune!(l::Lattice, steps_per_trial::Integer=64)
    ms = Ref{Float64}(0.0)
    check(ccall((:kmi_lattice_autotune, libkmi), Cint, (Ptr{Cvoid}, Int64, Ref{Float64}),
                l.h, steps_per_trial, ms), l.backend.ctx)
    return ms[]
end
set_temporal!(l::Lattice, tile::Integer, zchunks::Integer=1) =
    check(ccall((:kmi_lattice_set_temporal, libkmi), Cint, (Ptr{Cvoid}, Int32, Int32),
                l.h, tile, zchunks), l.backend.ctx)

# ---- batched cavities (kmi_batch_*) -------------------------------------------------------
# Hundreds of square D2Q9 lid-driven cavities of uneven side and viscosity, advanced by one
# launch per step (or per two steps) and each run to its own residual: the ensemble form of
# run_cavity_2d / the MRT loop of test/test_mrt.jl:57-63.
struct KmiBatchDesc
    storage::Int32; collision::Int32; lid::Int32; pad::Int32
    u_lid::Float64; s_e::Float64; s_eps::Float64; s_q::Float64; Lambda::Float64
end

const CAV_RUNNING, CAV_CONVERGED, CAV_DIVERGED, CAV_MAX_STEPS = 0, 1, 2, 3

mutable struct Batch
    h::Ptr{Cvoid}
    backend::MI355XBackend
    n::Vector{Int32}
end

function Batch(b::MI355XBackend, n::AbstractVector{<:Integer}, ν::AbstractVector{<:Real};
               storage::Integer=2, collision::Integer=2, u_lid::Real=0.1, zou_he_lid::Bool=true)
    d = KmiBatchDesc(storage, collision, zou_he_lid ? FACE_ZOUHE_VELOCITY : FACE_MOVING_WALL, 0,
                     u_lid, 1.4, 1.4, 1.2, 3 / 16)
    nn, νν = Int32.(collect(n)), Float64.(collect(ν))
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:kmi_batch_create, libkmi), Cint,
                (Ptr{Cvoid}, Ref{KmiBatchDesc}, Int32, Ptr{Int32}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
                b.ctx, d, length(nn), nn, νν, r), b.ctx)
    bt = Batch(r[], b, nn)
    finalizer(x -> ccall((:kmi_batch_destroy, libkmi), Cint, (Ptr{Cvoid},), x.h), bt)
    return bt
end

function autotune!(bt::Batch, steps_per_trial::Integer=8)
    ms = Ref{Float64}(0.0)
    check(ccall((:kmi_batch_autotune, libkmi), Cint, (Ptr{Cvoid}, Int64, Ref{Float64}),
                bt.h, steps_per_trial, ms), bt.backend.ctx)
    return ms[]
end

set_temporal!(bt::Batch, steps_per_launch::Integer) =
    check(ccall((:kmi_batch_set_temporal, libkmi), Cint, (Ptr{Cvoid}, Int32), bt.h, steps_per_launch),
          bt.backend.ctx)

# Warm start of cavity c (1-based): populations set to feq of the given fields (Nc × Nc).
init_equilibrium!(bt::Batch, c::Integer, ρ::Matrix{Float64}, ux::Matrix{Float64}, uy::Matrix{Float64}) =
    check(ccall((:kmi_batch_init_equilibrium, libkmi), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                bt.h, c - 1, ρ, ux, uy), bt.backend.ctx)

# Step every running cavity to r < tol (checked every `check_every` steps, r over the last
# `window` steps), divergence or max_steps; returns the device milliseconds.
function run!(bt::Batch, tol::Real; check_every::Integer=100, window::Integer=1,
              max_steps::Integer=10_000_000)
    ms = Ref{Float64}(0.0)
    check(ccall((:kmi_batch_run, libkmi), Cint, (Ptr{Cvoid}, Float64, Int64, Int64, Int64, Ref{Float64}),
                bt.h, tol, check_every, window, max_steps, ms), bt.backend.ctx)
    return ms[]
end

function status(bt::Batch)
    k = length(bt.n)
    st, steps, r = zeros(Int32, k), zeros(Int64, k), zeros(Float64, k)
    check(ccall((:kmi_batch_status, libkmi), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int64}, Ptr{Float64}),
                bt.h, st, steps, r), bt.backend.ctx)
    return (status=st, steps=steps, residual=r)
end

function moments(bt::Batch, c::Integer)
    m = Int(bt.n[c])
    ρ, ux, uy = zeros(m, m), zeros(m, m), zeros(m, m)
    check(ccall((:kmi_batch_moments, libkmi), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                bt.h, c - 1, ρ, ux, uy), bt.backend.ctx)
    return (ρ=ρ, ux=ux, uy=uy)
end

function desc_2d(Nx, Ny, ν, T; faces)
    w = KmiFace(FACE_WALL, 0, (0.0, 0.0, 0.0), 1.0, 0.0)
    fs = [w, w, w, w, w, w]
    for (k, v) in faces
        fs[k + 1] = v
    end
    sp = 1 / (3ν + 0.5); sm = 1 / ((3 / 16) / (3ν) + 0.5)
    KmiDesc(0, STORAGE[T], 0, 0, Nx, Ny, 1, 1 / (3ν + 0.5), sp, sm, 1.4, 1.4, 1.2,
            1 / (3ν + 0.5), 1 / (3ν + 0.5), 0.0, 0.0, 0, 0, (0, 0, 0, 0, 0, 0), Tuple(fs),
            0, (0.0, 0.0, 0.0), 0)
end

# ---- driver methods for the backend (src/drivers/basic.jl:112-144) ------------------------
"""
    run_cavity_2d(config::LBMConfig{D2Q9}; backend::MI355XBackend, T=Float64)

Same result as the reference driver: Zou–He lid, halfway bounce-back walls, BGK, rest start.
"""
function run_cavity_2d(config::LBMConfig{D2Q9}; backend::MI355XBackend, T=Float64)
    lid = KmiFace(FACE_ZOUHE_VELOCITY, 0, (config.u_lid, 0.0, 0.0), 1.0, 0.0)
    d = desc_2d(config.Nx, config.Ny, config.ν, T; faces=Dict(3 => lid))
    lat = Lattice(backend, d)
    step!(lat, config.max_steps)
    m = moments(lat)
    return (ρ=m.ρ, ux=m.ux, uy=m.uy, config=config)
end

end # module
