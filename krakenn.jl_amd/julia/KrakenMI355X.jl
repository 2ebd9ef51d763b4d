# KrakenMI355X.jl — Julia side of the MI355X backend for Kraken.jl's stream–collide path.
#
# A thin `ccall` layer over libkraken_mi355x.so (include/kraken_mi355x.h): no
# KernelAbstractions, no AMDGPU.jl — every time step runs in the library's HIP kernels.
#
# Dispatch.  Kraken's drivers take the backend as a keyword (`run_cavity_2d(config;
# backend=...)`, src/drivers/basic.jl:112-113), and Julia does not dispatch on keywords.  So
# this module never redefines a Kraken method: it adds methods that take the backend as their
# FIRST POSITIONAL argument,
#
#     run_cavity_2d(b::MI355XBackend, config::LBMConfig{D2Q9}; T=Float64)
#     run_cavity_3d(b::MI355XBackend, config::LBMConfig{D3Q19}; T=Float64)
#     run_rayleigh_benard_2d(b::MI355XBackend; <the reference's keywords>)
#     run_natural_convection_3d(b::MI355XBackend; <the reference's keywords>)
#     run_simulation(b::MI355XBackend, setup::SimulationSetup; T, callback, callback_every)
#
# and Kraken forwards to them with one line at the top of each driver,
#
#     backend isa KernelAbstractions.Backend || return run_cavity_2d(backend, config; T=T)
#
# (in run_simulation right after `_resolve_run_backend`, src/simulation_runner.jl:245-246).
# The `.krk` directive `Backend mi355x` resolves through `krk_backend()` below.
#
# Julia is not installed where this package is built and tested; the identical C ABI is
# exercised from Python (kraken_mi355x/), and a test checks every `ccall` of this file
# (symbol, argument count and types) and the struct layouts against the C header.
module KrakenMI355X

using Kraken
using Kraken: LBMConfig, D2Q9, D3Q19, SimulationSetup, evaluate, is_spatial, is_time_dependent

export MI355XBackend, Lattice, Batch, WarmStartNet

const libkmi = get(ENV, "KRAKEN_MI355X_LIB",
                   joinpath(@__DIR__, "..", "kraken_mi355x", "libkraken_mi355x.so"))

# ---- C structs (layouts of kmi_face, kmi_desc, kmi_batch_desc, kmi_nn_desc) ------------------
struct KmiFace
    kind::Int32
    thermal::Int32
    u::NTuple{3,Float64}
    rho::Float64
    T::Float64
end

struct KmiDesc
    stencil::Int32
    storage::Int32
    collision::Int32
    thermal::Int32
    nx::Int64
    ny::Int64
    nz::Int64
    omega::Float64
    s_plus::Float64
    s_minus::Float64
    s_e::Float64
    s_eps::Float64
    s_q::Float64
    s_nu::Float64
    omega_T::Float64
    beta_g::Float64
    T_ref::Float64
    coupling::Int32
    n_bc_order::Int32
    bc_order::NTuple{6,Int32}
    faces::NTuple{6,KmiFace}
    forcing::Int32
    force::NTuple{3,Float64}
    solid_mode::Int32
end

struct KmiBatchDesc
    storage::Int32
    collision::Int32
    lid::Int32
    pad::Int32
    u_lid::Float64
    s_e::Float64
    s_eps::Float64
    s_q::Float64
    Lambda::Float64
end

struct KmiNNDesc
    in_dim::Int32
    hidden::Int32
    latent_hw::Int32
    latent_c::Int32
    c1::Int32
    c2::Int32
    out_c::Int32
    max_batch::Int32
    seed::UInt64
end

# ---- enums of the header ------------------------------------------------------------------
const KMI_D2Q9, KMI_D3Q19 = Int32(0), Int32(1)
const STORAGE = Dict(:f64 => Int32(0), :f32 => Int32(1), :f32s => Int32(2), :f16s => Int32(3))
const COLLISION = Dict(:bgk => Int32(0), :trt => Int32(1), :mrt => Int32(2))
const FACE_KIND = Dict(:periodic => Int32(0), :wall => Int32(1), :moving_wall => Int32(2),
                       :zouhe_velocity => Int32(3), :zouhe_pressure => Int32(4),
                       :zouhe_velocity_profile => Int32(5), :zouhe_pressure_profile => Int32(6),
                       :zouhe_velocity_rebuild => Int32(7), :zouhe_pressure_rebuild => Int32(8))
const FACE_ID = Dict(:west => Int32(0), :east => Int32(1), :south => Int32(2), :north => Int32(3),
                     :bottom => Int32(4), :top => Int32(5))
const THERMAL_FACE = Dict(:adiabatic => Int32(0), :dirichlet => Int32(1))
const COUPLING = Dict(:split => Int32(0), :fused => Int32(1))
const FORCING = Dict(:none => Int32(0), :guo => Int32(1), :guo_plain_moments => Int32(2),
                     :guo_field => Int32(3))
const REDUCTION = Dict(:mass => Int32(0), :momentum => Int32(1), :kinetic_energy => Int32(2),
                       :thermal => Int32(3), :nusselt => Int32(4), :max_speed => Int32(5),
                       :drag_mea => Int32(6), :drag_halfway => Int32(7))
const FIELD_F, FIELD_G = Int32(0), Int32(1)
const HOST_F64, HOST_F32, HOST_STORAGE = Int32(0), Int32(1), Int32(2)
const CAV_RUNNING, CAV_CONVERGED, CAV_DIVERGED, CAV_MAX_STEPS = 0, 1, 2, 3

storage_of(::Type{Float64}) = :f64
storage_of(::Type{Float32}) = :f32
storage_of(T) = throw(ArgumentError("MI355X backend: element type $T is not supported (Float64, Float32)"))
host_dtype(::Type{Float64}) = HOST_F64
host_dtype(::Type{Float32}) = HOST_F32

# ---- errors ---------------------------------------------------------------------------------
lasterr(ctx::Ptr{Cvoid}) = unsafe_string(ccall((:kmi_last_error, libkmi), Cstring, (Ptr{Cvoid},), ctx))

# kmi_status 1 (bad argument) and 5 (unsupported) are the caller's: ArgumentError, as the
# reference raises for invalid setups; the rest (device, state, memory) are ErrorExceptions.
function check(rc::Integer, ctx::Ptr{Cvoid}=C_NULL)
    rc == 0 && return nothing
    msg = lasterr(ctx)
    rc in (1, 5) ? throw(ArgumentError(msg)) : error(msg)
end

# ---- the backend value ----------------------------------------------------------------------
"""
    MI355XBackend(device=0)

Backend value for Kraken's drivers and `run_simulation` on one MI355X (one `kmi_ctx`: a
device and a HIP stream).  Raises `ArgumentError` naming the backend when no device is
visible — never a silent CPU run (src/simulation_runner.jl:123-149).

The context is reference-counted: the backend holds one reference and every lattice, batch
or network built on it holds one more, so the context is destroyed by whichever finalizer
runs last.
"""
mutable struct MI355XBackend
    ctx::Ptr{Cvoid}
    device::Int
    refs::Threads.Atomic{Int}
end

function MI355XBackend(device::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:kmi_ctx_create, libkmi), Cint, (Cint, Ref{Ptr{Cvoid}}), device, r)
    rc == 0 || throw(ArgumentError("Backend 'mi355x' is unavailable: " * lasterr(C_NULL)))
    b = MI355XBackend(r[], Int(device), Threads.Atomic{Int}(1))
    finalizer(_release_ctx!, b)
    return b
end

_retain_ctx!(b::MI355XBackend) = (Threads.atomic_add!(b.refs, 1); b)

function _release_ctx!(b::MI355XBackend)
    if Threads.atomic_sub!(b.refs, 1) == 1 && b.ctx != C_NULL
        ccall((:kmi_ctx_destroy, libkmi), Cint, (Ptr{Cvoid},), b.ctx)
        b.ctx = C_NULL
    end
    return nothing
end

synchronize(b::MI355XBackend) =
    check(ccall((:kmi_ctx_synchronize, libkmi), Cint, (Ptr{Cvoid},), b.ctx), b.ctx)

"""
    krk_backend() -> MI355XBackend

Resolver of the `.krk` directive `Backend mi355x` (the `:mi355x` branch of
`_backend_from_krk_symbol`, src/simulation_runner.jl:110-121): the device's backend, or an
`ArgumentError` naming the backend when the library or the device is unavailable.
"""
function krk_backend()
    isfile(libkmi) || throw(ArgumentError(
        "Backend 'mi355x' requested via .krk 'Backend' directive, but $libkmi is missing. " *
        "Use `Backend cpu` instead."))
    n = Ref{Cint}(0)
    ccall((:kmi_device_count, libkmi), Cint, (Ref{Cint},), n)
    n[] > 0 || throw(ArgumentError(
        "Backend 'mi355x' requested via .krk 'Backend' directive, but no HIP device was " *
        "detected on this machine. Use `Backend cpu` instead."))
    return MI355XBackend(0)
end

# ---- descriptors ----------------------------------------------------------------------------
face(kind::Symbol; thermal::Symbol=:adiabatic, u=(0.0, 0.0, 0.0), rho::Real=1.0, T::Real=0.0) =
    KmiFace(FACE_KIND[kind], THERMAL_FACE[thermal], Float64.(Tuple(u)), Float64(rho), Float64(T))

# A copy of `d` with the named fields replaced (KmiDesc is immutable, as the C struct is plain).
function with_fields(d::KmiDesc; kw...)
    names = fieldnames(KmiDesc)
    vals = Any[getfield(d, f) for f in names]
    for (k, v) in kw
        i = findfirst(==(k), names)
        i === nothing && throw(ArgumentError("kmi_desc has no field $k"))
        vals[i] = convert(fieldtype(KmiDesc, i), v)
    end
    return KmiDesc(vals...)
end

function trt_rates(ν::Real, Λ::Real=3 / 16)
    sp, sm = Ref{Float64}(0.0), Ref{Float64}(0.0)
    ccall((:kmi_trt_rates, libkmi), Cvoid, (Float64, Float64, Ref{Float64}, Ref{Float64}), ν, Λ, sp, sm)
    return sp[], sm[]
end

"""
    make_desc(stencil, nx, ny, nz, ν; storage, collision, faces, ...) -> KmiDesc

Descriptor of one lattice: `stencil` is `:d2q9` or `:d3q19`; `faces` maps a face symbol
(`:west` … `:top`) to a `face(...)`; unnamed faces are halfway bounce-back walls.  Rates
follow the reference: ω = 1/(3ν + ½) (basic.jl:30), TRT s± from `trt_rates` (Λ), MRT
s_e = s_ε = 1.4, s_q = 1.2, s_ν = ω (collide_mrt_2d.jl:118).
"""
function make_desc(stencil::Symbol, nx::Integer, ny::Integer, nz::Integer, ν::Real;
                   storage::Symbol=:f64, collision::Symbol=:bgk, faces=Dict{Symbol,KmiFace}(),
                   Λ::Real=3 / 16, thermal::Bool=false, omega_T::Real=0.0, beta_g::Real=0.0,
                   T_ref::Real=0.0, coupling::Symbol=:split, forcing::Symbol=:none,
                   force=(0.0, 0.0, 0.0), bc_order::Vector{Symbol}=Symbol[])
    st = stencil === :d2q9 ? KMI_D2Q9 : stencil === :d3q19 ? KMI_D3Q19 :
         throw(ArgumentError("unknown stencil $stencil"))
    r = Ref{KmiDesc}()
    ccall((:kmi_desc_init, libkmi), Cvoid, (Ref{KmiDesc}, Int32, Int64, Int64, Int64, Float64),
          r, st, nx, ny, nz, ν)
    fs = collect(r[].faces)
    for (k, v) in faces
        fs[FACE_ID[k] + 1] = v
    end
    sp, sm = trt_rates(ν, Λ)
    order = zeros(Int32, 6)
    for (n, f) in enumerate(bc_order)
        order[n] = FACE_ID[f]
    end
    d = with_fields(r[]; storage=STORAGE[storage], collision=COLLISION[collision],
                    s_plus=sp, s_minus=sm, thermal=Int32(thermal), omega_T=omega_T,
                    beta_g=beta_g, T_ref=T_ref, coupling=COUPLING[coupling],
                    n_bc_order=Int32(length(bc_order)), bc_order=Tuple(order), faces=Tuple(fs),
                    forcing=FORCING[forcing], force=Float64.(Tuple(force)))
    check(ccall((:kmi_desc_validate, libkmi), Cint, (Ref{KmiDesc},), d))
    return d
end

# ---- one lattice ----------------------------------------------------------------------------
"""
    Lattice(b::MI355XBackend, d::KmiDesc)

Populations of one lattice resident in HBM, at the rest equilibrium (initialize_2d/3d,
basic.jl:51-101).  Host arrays use Kraken's own layout: `f[i, j, (k,) q]`, column-major.
"""
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
    _retain_ctx!(b)
    lat = Lattice(r[], b, (Int(d.nx), Int(d.ny), Int(d.nz)), d.stencil == KMI_D2Q9 ? 9 : 19)
    finalizer(_destroy!, lat)
    return lat
end

function _destroy!(l::Lattice)
    if l.h != C_NULL
        ccall((:kmi_lattice_destroy, libkmi), Cint, (Ptr{Cvoid},), l.h)
        l.h = C_NULL
        _release_ctx!(l.backend)
    end
    return nothing
end

cells(l::Lattice) = l.Q == 9 ? (l.dims[1], l.dims[2]) : l.dims
_ctx(l::Lattice) = l.backend.ctx
_f64(a) = a === nothing ? C_NULL : convert(Array{Float64}, a)

step!(l::Lattice, n::Integer) =
    check(ccall((:kmi_step, libkmi), Cint, (Ptr{Cvoid}, Int64), l.h, n), _ctx(l))

"""Advance `n` steps; returns r_u = ‖u^k − u^(k−1)‖₂ / ‖u^k‖₂ of the last one
(gnf_periodic_steady.jl:186-196)."""
function step_residual!(l::Lattice, n::Integer)
    r = Ref{Float64}(0.0)
    check(ccall((:kmi_step_residual, libkmi), Cint, (Ptr{Cvoid}, Int64, Ref{Float64}), l.h, n, r), _ctx(l))
    return r[]
end

function upload!(l::Lattice, f::Array{E}; field::Int32=FIELD_F) where {E<:Union{Float64,Float32}}
    length(f) == prod(cells(l)) * l.Q || throw(DimensionMismatch("population array of $(size(f))"))
    check(ccall((:kmi_upload, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int32),
                l.h, field, f, host_dtype(E)), _ctx(l))
end

function populations(l::Lattice, ::Type{E}=Float64; field::Int32=FIELD_F) where {E<:Union{Float64,Float32}}
    f = Array{E}(undef, cells(l)..., l.Q)
    check(ccall((:kmi_download, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int32),
                l.h, field, f, host_dtype(E)), _ctx(l))
    return f
end

"""f_q = feq_q(ρ, u) per cell; `nothing` stands for ρ = 1 or u = 0."""
init_equilibrium!(l::Lattice, ρ, ux, uy, uz=nothing) =
    check(ccall((:kmi_init_equilibrium, libkmi), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                l.h, _f64(ρ), _f64(ux), _f64(uy), _f64(uz)), _ctx(l))

init_temperature!(l::Lattice, T::AbstractArray) =
    check(ccall((:kmi_init_temperature, libkmi), Cint, (Ptr{Cvoid}, Ptr{Float64}), l.h, _f64(T)), _ctx(l))

set_mask!(l::Lattice, is_solid::AbstractArray{Bool}) =
    check(ccall((:kmi_set_mask, libkmi), Cint, (Ptr{Cvoid}, Ptr{UInt8}), l.h, convert(Array{UInt8}, is_solid)),
          _ctx(l))

set_face!(l::Lattice, f::Symbol, v::KmiFace) =
    check(ccall((:kmi_set_face, libkmi), Cint, (Ptr{Cvoid}, Int32, Ref{KmiFace}), l.h, FACE_ID[f], v), _ctx(l))

set_face_profile!(l::Lattice, f::Symbol, a::AbstractVector, b=nothing) =
    check(ccall((:kmi_set_face_profile, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}),
                l.h, FACE_ID[f], _f64(a), _f64(b)), _ctx(l))

# Per-node force of a :guo_field lattice: the Fx_field, Fy_field(, Fz_field) arrays of
# collide_guo_field_2d!/_3d! (collide_guo_2d.jl:93-160, collide_guo_3d.jl:199-332).
set_force_field!(l::Lattice, Fx, Fy, Fz=nothing) =
    check(ccall((:kmi_set_force_field, libkmi), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                l.h, _f64(Fx), _f64(Fy), _f64(Fz)), _ctx(l))

"""
    moments(l; temperature=false) -> (ρ, ux, uy[, uz][, Temp])

Moments of the stored (post-collision) populations, as compute_macroscopic_2d/3d!
(macroscopic.jl:5-80).  Kraken's fused kernels (fused_bgk_step!, fused_trt_step!,
fused_bgk_2d.jl:26-56) instead write the pre-collision moments and ρ = 1, u = 0 on solid
cells: collisions conserve ρ and ρu, so fluid cells agree up to the rounding of the
collision, and solid cells differ.
"""
function moments(l::Lattice; temperature::Bool=false)
    n = cells(l)
    ρ, ux, uy = zeros(n), zeros(n), zeros(n)
    uz = l.Q == 19 ? zeros(n) : nothing
    Tm = temperature ? zeros(n) : nothing
    check(ccall((:kmi_moments, libkmi), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                l.h, ρ, ux, uy, _f64(uz), _f64(Tm)), _ctx(l))
    base = l.Q == 19 ? (ρ=ρ, ux=ux, uy=uy, uz=uz) : (ρ=ρ, ux=ux, uy=uy)
    return temperature ? merge(base, (Temp=Tm,)) : base
end

"""FP64 reductions: :mass, :momentum (3 values), :kinetic_energy, :thermal, :max_speed,
:drag_mea / :drag_halfway (3 values), :nusselt (hot wall `face`)."""
function reduction(l::Lattice, kind::Symbol; face::Symbol=:south)
    out = zeros(3)
    check(ccall((:kmi_reduce, libkmi), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}),
                l.h, REDUCTION[kind], FACE_ID[face], out), _ctx(l))
    return kind in (:momentum, :drag_mea, :drag_halfway) ? out : out[1]
end

# The stored populations verbatim (F64 storage: Float64, F32: Float32), for checkpoints of a
# lattice (src/platform/state.jl) that restart bit for bit; set_steps! restores the cycle.
function download_raw(l::Lattice, ::Type{E}) where {E<:Union{Float64,Float32}}
    f = Array{E}(undef, cells(l)..., l.Q)
    check(ccall((:kmi_download, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Int32),
                l.h, FIELD_F, f, HOST_STORAGE), _ctx(l))
    return f
end
set_steps!(l::Lattice, n::Integer) =
    check(ccall((:kmi_lattice_set_steps, libkmi), Cint, (Ptr{Cvoid}, Int64), l.h, n), _ctx(l))

# Launch choice (results do not depend on it): the autotuner times the single-step shapes and
# the multi-step kernels where they apply; set_temporal! fixes a multi-step tile (0: off).
function autotune!(l::Lattice, steps_per_trial::Integer=64)
    ms = Ref{Float64}(0.0)
    check(ccall((:kmi_lattice_autotune, libkmi), Cint, (Ptr{Cvoid}, Int64, Ref{Float64}),
                l.h, steps_per_trial, ms), _ctx(l))
    return ms[]
end
set_temporal!(l::Lattice, tile::Integer, zchunks::Integer=1) =
    check(ccall((:kmi_lattice_set_temporal, libkmi), Cint, (Ptr{Cvoid}, Int32, Int32),
                l.h, tile, zchunks), _ctx(l))

# Small 2D lattices are bound by launch latency: over a long run the tuner's trials (a few
# thousand steps, state restored) pay for themselves.
_tune!(l::Lattice, steps::Integer) = (l.Q == 9 && steps >= 2000) ? autotune!(l, 64) : nothing

# ---- drivers (backend as first positional argument) -----------------------------------------
cavity_desc(Nx, Ny, ν, u_lid, T) =
    make_desc(:d2q9, Nx, Ny, 1, ν; storage=storage_of(T),
              faces=Dict(:north => face(:zouhe_velocity; u=(u_lid, 0.0, 0.0))))

"""
    run_cavity_2d(b::MI355XBackend, config::LBMConfig{D2Q9}; T=Float64)

run_cavity_2d (basic.jl:112-144) on the device: rest start, halfway bounce-back walls,
Zou–He lid, BGK, `max_steps` single-pass steps; returns `(ρ, ux, uy, config)` on the host.
"""
function Kraken.run_cavity_2d(b::MI355XBackend, config::LBMConfig{D2Q9}; T=Float64)
    lat = Lattice(b, cavity_desc(config.Nx, config.Ny, config.ν, config.u_lid, T))
    _tune!(lat, config.max_steps)
    step!(lat, config.max_steps)
    m = moments(lat)
    _destroy!(lat)
    return (ρ=m.ρ, ux=m.ux, uy=m.uy, config=config)
end

"""
    run_cavity_3d(b::MI355XBackend, config::LBMConfig{D3Q19}; T=Float64)

run_cavity_3d (basic.jl:153-176): halfway bounce-back on five faces, Zou–He lid (u_lid, 0, 0)
on the top face, BGK; returns `(ρ, ux, uy, uz, config)`.
"""
function Kraken.run_cavity_3d(b::MI355XBackend, config::LBMConfig{D3Q19}; T=Float64)
    d = make_desc(:d3q19, config.Nx, config.Ny, config.Nz, config.ν; storage=storage_of(T),
                  faces=Dict(:top => face(:zouhe_velocity; u=(config.u_lid, 0.0, 0.0))))
    lat = Lattice(b, d)
    step!(lat, config.max_steps)
    m = moments(lat)
    _destroy!(lat)
    return (ρ=m.ρ, ux=m.ux, uy=m.uy, uz=m.uz, config=config)
end

"""
    run_rayleigh_benard_2d(b::MI355XBackend; Nx=128, Ny=32, Ra=2000.0, Pr=1.0, T_hot=1.0,
                           T_cold=0.0, max_steps=20000, nu=nothing, alpha=nothing,
                           orientation=:vertical, FT=Float64)

run_rayleigh_benard_2d (thermal.jl:31-141) on the device, one launch per step: the unfused
chain's coupling (thermal collision with the post-stream velocity, Boussinesq
F_y = β_g (T − T_ref)).  :vertical — hot south / cold north anti-bounce-back, periodic x,
linear T plus the reference's sine seed; :horizontal — hot west / cold east, walls on all
faces, linear T without a seed.
"""
function Kraken.run_rayleigh_benard_2d(b::MI355XBackend; Nx=128, Ny=32, Ra=2000.0, Pr=1.0,
                                       T_hot=1.0, T_cold=0.0, max_steps=20000, nu=nothing,
                                       alpha=nothing, orientation::Symbol=:vertical, FT=Float64)
    orientation in (:vertical, :horizontal) || throw(ArgumentError(
        "run_rayleigh_benard_2d: orientation ':$(orientation)' is not supported. " *
        "Use :vertical (hot south / cold north, periodic x) or :horizontal (hot west / cold east, solid walls)."))
    ΔT = T_hot - T_cold
    iszero(ΔT) && throw(ArgumentError(
        "run_rayleigh_benard_2d: T_hot and T_cold are both $(T_hot). " *
        "A zero temperature difference leaves the buoyancy scaling undefined."))
    H = orientation === :vertical ? Ny : Nx
    ν = nu === nothing ? 0.05 : Float64(nu)
    α = alpha === nothing ? ν / Pr : Float64(alpha)
    ν > 0 || throw(ArgumentError("run_rayleigh_benard_2d: nu must be positive, got $ν."))
    α > 0 || throw(ArgumentError("run_rayleigh_benard_2d: alpha must be positive, got $α."))
    Pr = alpha === nothing ? Pr : ν / α
    β_g = Ra * ν * α / (ΔT * H^3)
    hot, cold = face(:wall; thermal=:dirichlet, T=T_hot), face(:wall; thermal=:dirichlet, T=T_cold)
    faces = orientation === :vertical ?
            Dict(:west => face(:periodic), :east => face(:periodic), :south => hot, :north => cold) :
            Dict(:west => hot, :east => cold)
    d = make_desc(:d2q9, Nx, Ny, 1, ν; storage=storage_of(FT), thermal=true,
                  omega_T=1 / (3α + 0.5), beta_g=β_g, T_ref=(T_hot + T_cold) / 2,
                  coupling=:split, faces=faces)
    Tinit = zeros(Nx, Ny)
    for j in 1:Ny, i in 1:Nx
        # the reference's profile, rounded to FT as its g_cpu is (thermal.jl:74-91)
        t = orientation === :vertical ?
            FT(T_hot - ΔT * (j - 0.5) / Ny) + FT(0.01 * ΔT) * sin(FT(2π * i / Nx)) * sin(FT(π * j / Ny)) :
            FT(T_hot - ΔT * (i - 0.5) / Nx)
        Tinit[i, j] = Float64(t)
    end
    lat = Lattice(b, d)
    init_temperature!(lat, Tinit)
    _tune!(lat, max_steps)
    step!(lat, max_steps)
    m = moments(lat; temperature=true)
    _destroy!(lat)
    config = LBMConfig(D2Q9(); Nx=Nx, Ny=Ny, ν=ν, u_lid=0.0, max_steps=max_steps)
    return (ρ=m.ρ, ux=m.ux, uy=m.uy, Temp=m.Temp, config=config, Ra=Ra, Pr=Pr, ν=ν, α=α,
            T_hot=T_hot, T_cold=T_cold, orientation=orientation)
end

"""
    run_natural_convection_3d(b::MI355XBackend; N=32, Ra=1e3, Pr=0.71, T_hot=1.0,
                              T_cold=0.0, max_steps=10000, FT=Float64)

run_natural_convection_3d (thermal.jl:455-538): hot west / cold east anti-bounce-back with
the edge rule, adiabatic elsewhere, Boussinesq in y; Nu from the hot-wall estimator averaged
without the edge rows (a device reduction).
"""
function Kraken.run_natural_convection_3d(b::MI355XBackend; N=32, Ra=1e3, Pr=0.71, T_hot=1.0,
                                          T_cold=0.0, max_steps=10000, FT=Float64)
    ΔT = T_hot - T_cold
    ν = 0.05
    α = ν / Pr
    β_g = Ra * ν * α / (ΔT * Float64(N)^3)
    d = make_desc(:d3q19, N, N, N, ν; storage=storage_of(FT), thermal=true, omega_T=1 / (3α + 0.5),
                  beta_g=β_g, T_ref=(T_hot + T_cold) / 2, coupling=:split,
                  faces=Dict(:west => face(:wall; thermal=:dirichlet, T=T_hot),
                             :east => face(:wall; thermal=:dirichlet, T=T_cold)))
    Tinit = [Float64(FT(T_hot - ΔT * (i - 0.5) / N)) for i in 1:N, j in 1:N, k in 1:N]
    lat = Lattice(b, d)
    init_temperature!(lat, Tinit)
    step!(lat, max_steps)
    m = moments(lat; temperature=true)
    Nu = reduction(lat, :nusselt; face=:west)
    _destroy!(lat)
    config = LBMConfig(D3Q19(); Nx=N, Ny=N, Nz=N, ν=ν, u_lid=0.0, max_steps=max_steps)
    return (ρ=m.ρ, ux=m.ux, uy=m.uy, uz=m.uz, Temp=m.Temp, Nu=Nu, config=config, Ra=Ra, Pr=Pr,
            ν=ν, α=α)
end

# ---- run_simulation: the default single-phase D2Q9 loop (simulation_runner.jl:279-443) -------
"""
    run_simulation(b::MI355XBackend, setup::SimulationSetup; T=Float64, callback=nothing,
                   callback_every=100)

The reference's default 2D path on the device.  The host-side setup is Kraken's own
(`sanity_check`, `_apply_geometry!`, `_apply_initial_conditions!`, `_fill_bc_arrays!`); the
per-step chain — stream (periodic or halfway bounce-back, `_select_streaming_kernel`), the
boundary handlers in setup order, BGK / Guo BGK / fused TRT, moments — is one launch.
Time-dependent faces are refreshed on the device before every step with the reference's
t = step.  Paths other than the default loop raise ArgumentError naming the backend.
"""
function Kraken.run_simulation(b::MI355XBackend, setup::SimulationSetup; T=Float64,
                               callback::Union{Nothing,Function}=nothing, callback_every::Int=100)
    Kraken.sanity_check(setup)
    other = setup.sensitivity !== nothing || setup.mesh !== nothing || !isempty(setup.modules) ||
            !isempty(setup.refinements) || !isempty(setup.rheology) || setup.lattice !== :D2Q9
    other && throw(ArgumentError(
        "Backend 'mi355x' runs the default single-phase D2Q9 path of run_simulation; this setup " *
        "(modules $(setup.modules), lattice $(setup.lattice)) needs another runner. Use `Backend cpu`."))
    # wall=libb changes the loop only where an STL obstacle carries it (simulation_runner.jl:299-300)
    wall_bc = Kraken._normalize_wall_bc_selector(setup.wall_bc)
    wall_bc === :libb && Kraken._has_stl_libb_obstacle(Kraken._setup_with_global_stl_libb(setup)) &&
        throw(ArgumentError("Backend 'mi355x': STL obstacles with wall=libb in the generic .krk " *
                            "runner are not supported. Use `Backend cpu`."))
    collision = Kraken._normalize_collision_selector(setup.collision)
    dom = setup.domain
    Nx, Ny = dom.Nx, dom.Ny
    dx, dy = dom.Lx / Nx, dom.Ly / Ny
    ν = Float64(setup.physics.params[:nu])

    # streaming selection (_select_streaming_kernel): periodic x with a periodic west or east
    # face; periodic y with a periodic south or north face and periodic x; walls otherwise
    types = Dict(bc.face => bc.type for bc in setup.boundaries)
    per_x = get(types, :west, :wall) === :periodic || get(types, :east, :wall) === :periodic
    per_y = per_x && (get(types, :south, :wall) === :periodic || get(types, :north, :wall) === :periodic)
    faces = Dict{Symbol,KmiFace}()
    per_x && (faces[:west] = faces[:east] = face(:periodic))
    per_y && (faces[:south] = faces[:north] = face(:periodic))

    has_force = !isempty(setup.physics.body_force)
    collision === :trt && Kraken._ensure_generic_trt_supported!(setup, has_force, false)

    # boundary handlers in setup order (handlers.jl:26-256); walls are the pull's bounce-back
    handlers = Tuple{Kraken.BoundarySetup,Symbol,Bool}[]
    order = Symbol[]
    for bc in setup.boundaries
        (bc.type === :periodic || bc.type === :wall) && continue
        get(faces, bc.face, face(:wall)).kind == FACE_KIND[:periodic] && continue
        td = any(is_time_dependent, values(bc.values))
        if bc.type === :velocity
            bc.face === :east && throw(ArgumentError(
                "Backend 'mi355x': a velocity condition on the east face is not applied by the " *
                "generic 2D route (handlers.jl:228-247); it is rejected rather than dropped."))
            # spatial or time-dependent velocities take the per-node route (handlers.jl:40-47)
            sp = any(k -> haskey(bc.values, k) && (is_spatial(bc.values[k]) || is_time_dependent(bc.values[k])),
                     (:ux, :uy))
            if sp
                faces[bc.face] = face(:zouhe_velocity_profile)
            else
                ux = haskey(bc.values, :ux) ? Float64(evaluate(bc.values[:ux]; t=0.0)) : 0.0
                faces[bc.face] = face(:zouhe_velocity; u=(ux, 0.0, 0.0))
            end
            push!(handlers, (bc, sp ? :profile : :scalar, td))
        elseif bc.type === :pressure
            bc.face in (:west, :east) || throw(ArgumentError(
                "Pressure boundary condition on face ':$(bc.face)' is not supported by the generic " *
                "2D runner. Supported pressure faces: :west, :east."))
            ρw = haskey(bc.values, :rho) ? Float64(evaluate(bc.values[:rho]; t=0.0)) : 1.0
            faces[bc.face] = face(:zouhe_pressure; rho=ρw)
            push!(handlers, (bc, :pressure, td))
        else
            throw(ArgumentError("unknown boundary type $(bc.type)"))
        end
        push!(order, bc.face)
    end

    force = (0.0, 0.0, 0.0)
    if has_force
        Fx = haskey(setup.physics.body_force, :Fx) ? Float64(T(evaluate(setup.physics.body_force[:Fx]))) : 0.0
        Fy = haskey(setup.physics.body_force, :Fy) ? Float64(T(evaluate(setup.physics.body_force[:Fy]))) : 0.0
        force = (Fx, Fy, 0.0)
    end
    d = make_desc(:d2q9, Nx, Ny, 1, ν; storage=storage_of(T), collision=collision, faces=faces,
                  bc_order=order, forcing=has_force ? :guo : :none, force=force)

    lat = Lattice(b, d)
    if !isempty(setup.regions)
        is_solid = zeros(Bool, Nx, Ny)
        Kraken._apply_geometry!(is_solid, setup, dx, dy)
        set_mask!(lat, is_solid)
    end
    if setup.initial !== nothing
        f0, f1 = zeros(T, Nx, Ny, 9), zeros(T, Nx, Ny, 9)
        Kraken._apply_initial_conditions!(f0, f1, setup, dx, dy, T)
        upload!(lat, f0)
    end
    # face values at t = 0 (and, for time-dependent faces, before each step at t = step)
    function refresh!(t::Float64, only_td::Bool)
        for (bc, kind, td) in handlers
            (only_td && !td) && continue
            if kind === :profile
                n = bc.face in (:north, :south) ? Nx : Ny
                ua, va = zeros(T, n), zeros(T, n)
                ux_fn = haskey(bc.values, :ux) ? bc.values[:ux].func : nothing
                uy_fn = haskey(bc.values, :uy) ? bc.values[:uy].func : nothing
                Kraken._fill_bc_arrays!(ua, va, ux_fn, uy_fn, bc.face, dx, dy, Nx, Ny, dom, T; t=t)
                set_face_profile!(lat, bc.face, ua, va)
            elseif kind === :pressure && td
                set_face!(lat, bc.face, face(:zouhe_pressure; rho=Float64(evaluate(bc.values[:rho]; t=t))))
            end
        end
    end
    refresh!(0.0, false)
    time_dep = any(h -> h[3], handlers)
    time_dep || _tune!(lat, setup.max_steps)
    step = 0
    while step < setup.max_steps
        if time_dep
            refresh!(Float64(step + 1), true)
            n = 1
        else
            n = setup.max_steps - step
            callback !== nothing && (n = min(n, callback_every - step % callback_every))
        end
        step!(lat, n)
        step += n
        if callback !== nothing && step % callback_every == 0
            m = moments(lat)
            callback(step, (; rho=m.ρ, ux=m.ux, uy=m.uy))
        end
    end
    m = moments(lat)
    _destroy!(lat)
    return (ρ=m.ρ, ux=m.ux, uy=m.uy, setup=setup)
end

# ---- batched cavities (kmi_batch_*) -------------------------------------------------------------
"""
    Batch(b, n, ν; storage=:f32s, collision=:mrt, u_lid=0.1, zou_he_lid=true)

Hundreds of square D2Q9 lid-driven cavities of side n[c] and viscosity ν[c], advanced by one
launch per step (or per several steps), each run to its own residual: the ensemble form of
run_cavity_2d / the MRT loop of test/test_mrt.jl:19-25.
"""
mutable struct Batch
    h::Ptr{Cvoid}
    backend::MI355XBackend
    n::Vector{Int32}
end

function Batch(b::MI355XBackend, n::AbstractVector{<:Integer}, ν::AbstractVector{<:Real};
               storage::Symbol=:f32s, collision::Symbol=:mrt, u_lid::Real=0.1, zou_he_lid::Bool=true)
    d = KmiBatchDesc(STORAGE[storage], COLLISION[collision],
                     zou_he_lid ? FACE_KIND[:zouhe_velocity] : FACE_KIND[:moving_wall], 0,
                     u_lid, 1.4, 1.4, 1.2, 3 / 16)
    nn, νν = Int32.(collect(n)), Float64.(collect(ν))
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:kmi_batch_create, libkmi), Cint,
                (Ptr{Cvoid}, Ref{KmiBatchDesc}, Int32, Ptr{Int32}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
                b.ctx, d, length(nn), nn, νν, r), b.ctx)
    _retain_ctx!(b)
    bt = Batch(r[], b, nn)
    finalizer(_destroy!, bt)
    return bt
end

function _destroy!(bt::Batch)
    if bt.h != C_NULL
        ccall((:kmi_batch_destroy, libkmi), Cint, (Ptr{Cvoid},), bt.h)
        bt.h = C_NULL
        _release_ctx!(bt.backend)
    end
    return nothing
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

# Warm start of cavity c (1-based): populations set to feq of the given fields (n_c × n_c).
init_equilibrium!(bt::Batch, c::Integer, ρ, ux, uy) =
    check(ccall((:kmi_batch_init_equilibrium, libkmi), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                bt.h, c - 1, _f64(ρ), _f64(ux), _f64(uy)), bt.backend.ctx)

"""Step every running cavity to r < tol (checked every `check_every` steps, r over the last
`window` steps), divergence or `max_steps`; returns the device milliseconds."""
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

# ---- learned warm start (kmi_nn_*) ---------------------------------------------------------------
"""
    WarmStartNet(b; hidden=256, latent_hw=16, max_batch=32, seed=1234)

The conv decoder that predicts a cavity's steady (u, v, ρ) field from (Re, lid amplitude),
trained and run on the device (BF16 MFMA, FP32 master weights); its prediction seeds
`init_equilibrium!` of a batch, the role of f_init in gnf_periodic_steady.jl:114-150.
"""
mutable struct WarmStartNet
    h::Ptr{Cvoid}
    backend::MI355XBackend
    desc::KmiNNDesc
end

function WarmStartNet(b::MI355XBackend; hidden::Integer=256, latent_hw::Integer=16,
                      max_batch::Integer=32, seed::Integer=1234)
    r0 = Ref{KmiNNDesc}()
    ccall((:kmi_nn_desc_init, libkmi), Cvoid, (Ref{KmiNNDesc},), r0)
    d0 = r0[]
    d = KmiNNDesc(d0.in_dim, hidden, latent_hw, d0.latent_c, d0.c1, d0.c2, d0.out_c, max_batch, seed)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:kmi_nn_create, libkmi), Cint, (Ptr{Cvoid}, Ref{KmiNNDesc}, Ref{Ptr{Cvoid}}), b.ctx, d, r),
          b.ctx)
    _retain_ctx!(b)
    net = WarmStartNet(r[], b, d)
    finalizer(_destroy!, net)
    return net
end

function _destroy!(net::WarmStartNet)
    if net.h != C_NULL
        ccall((:kmi_nn_destroy, libkmi), Cint, (Ptr{Cvoid},), net.h)
        net.h = C_NULL
        _release_ctx!(net.backend)
    end
    return nothing
end

# x: in_dim × batch (column-major, one sample per column, the C layout batch × in_dim);
# returns out_c × W × W × batch, the C layout batch × W × W × out_c (NHWC).
function forward(net::WarmStartNet, x::AbstractMatrix)
    B = size(x, 2)
    W = 8 * Int(net.desc.latent_hw)
    y = Array{Float32}(undef, Int(net.desc.out_c), W, W, B)
    check(ccall((:kmi_nn_forward, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{Float32}, Ptr{Float32}),
                net.h, B, convert(Matrix{Float32}, x), y), net.backend.ctx)
    return y
end

set_dataset!(net::WarmStartNet, x::AbstractMatrix, y::AbstractArray) =
    check(ccall((:kmi_nn_set_dataset, libkmi), Cint, (Ptr{Cvoid}, Int32, Ptr{Float32}, Ptr{Float32}),
                net.h, size(x, 2), convert(Matrix{Float32}, x), convert(Array{Float32}, y)), net.backend.ctx)

function train_epochs!(net::WarmStartNet, epochs::Integer, batch::Integer, lr::Real; seed::Integer=0)
    loss = Ref{Float64}(0.0)
    check(ccall((:kmi_nn_train_epochs, libkmi), Cint, (Ptr{Cvoid}, Int32, Int32, Float32, UInt64, Ref{Float64}),
                net.h, epochs, batch, lr, seed, loss), net.backend.ctx)
    return loss[]
end

end # module
