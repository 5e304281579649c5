# SymCodeHIP.jl -- Julia binding of libsymcode_hip.so for SymmetryCode.jl.
#
# Adds methods of the reference's own functions for a new backend tag, so the
# existing call sites run on an MI355X through the C ABI in
# include/symcode.h.  Fields are `HIPField`s: device buffers in the library's
# device layout, created by the allocators below (the Julia GC owns them
# through a finalizer).  A field iterates over its components (views that
# share its memory), as the reference's NamedTuples of arrays do, so the
# reference's own create_dns march loop, turbulence_statistics, sfs! and
# solve_les! run unchanged on them; their broadcasts reduce to `dest .= a`,
# `dest .= a .± b` and `dest .= c .* a` on spectral fields.  Artefact saving
# (`u |> cpu_device()`, JLD2) is not covered: `download` returns the host
# arrays.  Julia is not installed where this binding was written, so it has
# not been run.
#
# Usage:
#     using SymmetryCode, SymCodeHIP
#     g = SymmetryCode.Grid{3}(; l = 2π, n = 810, backend = SymCodeHIP.HIPBackend())
#     u = SymmetryCode.randomfield(SymmetryCode.peak_profile, g; kpeak = 4)
#     cache = SymmetryCode.getcache(g)
#     SymmetryCode.wray3!(SymmetryCode.convectiondiffusion!, u, Δt, g, cache; visc)
module SymCodeHIP

using SymmetryCode: SymmetryCode as S

const libsymcode = joinpath(@__DIR__, "..", "symcode", "libsymcode_hip.so")

"Backend tag for `Grid(; backend = HIPBackend())`."
struct HIPBackend end

# ------------------------------------------------------------------ context
mutable struct Context
    ptr::Ptr{Cvoid}
    D::Int
    n::Int
    nh::Int
    nhp::Int
    spec_elems::Int
    phys_elems::Int
    prec::Int
end

const CONTEXTS = Dict{Tuple{Int,Int,Float64,DataType},Context}()

function check(rc, ctx)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:sc_last_error, libsymcode), Cstring, (Ptr{Cvoid},), ctx))
    error("symcode error $rc: $msg")
end

"One library context per (D, n, l, T), created on first use."
function context(g::S.Grid{D,T,HIPBackend}) where {D,T}
    get!(CONTEXTS, (D, g.n, Float64(g.l), T)) do
        out = Ref{Ptr{Cvoid}}(C_NULL)
        prec = T === Float64 ? 64 : 32
        check(ccall((:sc_create, libsymcode), Cint, (Cint, Cint, Cdouble, Cint, Cint, Ref{Ptr{Cvoid}}),
                    D, g.n, g.l, prec, 0, out), C_NULL)
        info = zeros(Int64, 8)
        check(ccall((:sc_info, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Int64}), out[], info), out[])
        ctx = Context(out[], D, g.n, info[3], info[4], info[6], info[7], prec)
        finalizer(c -> ccall((:sc_destroy, libsymcode), Cint, (Ptr{Cvoid},), c.ptr), ctx)
        ctx
    end
end

# -------------------------------------------------------------------- fields
"""
A device field of `ncomp` spectral (complex) or physical (real) scalars.  A
field owns its memory (`parent === nothing`) or is a view of one component of
another field (`parent` keeps that field alive).
"""
mutable struct HIPField{T}
    ptr::Ptr{Cvoid}
    ncomp::Int
    spectral::Bool
    ctx::Context
    parent::Any
end

function HIPField{T}(ctx::Context, ncomp, spectral) where {T}
    elems = spectral ? ctx.spec_elems : ctx.phys_elems
    bytes = ncomp * elems * (spectral ? 2 : 1) * sizeof(T)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    # hipMalloc through the HIP runtime the library links against
    rc = ccall((:hipMalloc, "libamdhip64"), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, bytes)
    rc == 0 || error("hipMalloc failed ($rc)")
    ccall((:hipMemset, "libamdhip64"), Cint, (Ptr{Cvoid}, Cint, Csize_t), p[], 0, bytes)
    f = HIPField{T}(p[], ncomp, spectral, ctx, nothing)
    finalizer(x -> ccall((:hipFree, "libamdhip64"), Cint, (Ptr{Cvoid},), x.ptr), f)
    f
end

# components: f[i] is a one-component view (component stride = spec_elems or
# phys_elems scalars); iteration yields the components in the reference's
# NamedTuple order (x, y, z; xx, yy, zz, xy, yz, zx)
component_bytes(f::HIPField{T}) where {T} =
    (f.spectral ? 2 * f.ctx.spec_elems : f.ctx.phys_elems) * sizeof(T)
function Base.getindex(f::HIPField{T}, i::Integer) where {T}
    1 <= i <= f.ncomp || throw(BoundsError(f, i))
    f.ncomp == 1 && return f
    HIPField{T}(f.ptr + (i - 1) * component_bytes(f), 1, f.spectral, f.ctx, f)
end
Base.length(f::HIPField) = f.ncomp
Base.iterate(f::HIPField, i = 1) = i > f.ncomp ? nothing : (f[i], i + 1)

# Broadcasting on spectral fields, for the reference's `dest .= a`,
# `dest .= a .+ b`, `dest .= a .- b`, `dest .= c .* a` and `dest .*= c`
# (sc_lincomb: dest = α x + β y); anything else raises.
struct HIPStyle <: Base.Broadcast.BroadcastStyle end
Base.BroadcastStyle(::Type{<:HIPField}) = HIPStyle()
Base.BroadcastStyle(s::HIPStyle, ::Base.Broadcast.DefaultArrayStyle{0}) = s
Base.broadcastable(f::HIPField) = f
Base.axes(f::HIPField) = (Base.OneTo(f.ncomp),)
Base.Broadcast.instantiate(bc::Base.Broadcast.Broadcasted{HIPStyle}) = bc
function lincomb!(dest::HIPField, x::HIPField, α, y::HIPField, β)
    (dest.spectral && x.spectral && y.spectral) || error("broadcasts are supported on spectral fields only")
    check(ccall((:sc_lincomb, libsymcode), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Cdouble, Cint),
                dest.ctx.ptr, dest.ptr, x.ptr, α, y.ptr, β, dest.ncomp), dest.ctx.ptr)
    dest
end
# α x for a field or a product of a number and a field
scaled(x::HIPField) = (1.0, x)
function scaled(bc::Base.Broadcast.Broadcasted)
    if bc.f === (*) && length(bc.args) == 2
        a, b = bc.args
        a isa Number && b isa HIPField && return (Float64(a), b)
        b isa Number && a isa HIPField && return (Float64(b), a)
    end
    error("unsupported broadcast on HIPField: $(bc.f)")
end
Base.copyto!(dest::HIPField, x::HIPField) = lincomb!(dest, x, 1.0, x, 0.0)
function Base.copyto!(dest::HIPField, bc::Base.Broadcast.Broadcasted{HIPStyle})
    if bc.f === identity && length(bc.args) == 1
        return copyto!(dest, bc.args[1])
    elseif (bc.f === (+) || bc.f === (-)) && length(bc.args) == 2
        α, x = scaled(bc.args[1])
        β, y = scaled(bc.args[2])
        return lincomb!(dest, x, α, y, bc.f === (-) ? -β : β)
    end
    α, x = scaled(bc)
    lincomb!(dest, x, α, x, 0.0)
end

const HIPGrid{D,T} = S.Grid{D,T,HIPBackend}

# the reference's allocators (src/solver.jl:246-317) return one HIPField per
# vector/tensor field instead of a NamedTuple of arrays
S.scalarfield(g::HIPGrid{D,T}) where {D,T} = HIPField{T}(context(g), 1, true)
S.vectorfield(g::HIPGrid{D,T}) where {D,T} = HIPField{T}(context(g), D, true)
S.tensorfield(g::HIPGrid{D,T}) where {D,T} = HIPField{T}(context(g), D == 2 ? 3 : 6, true)
S.tensorfield_nonsym(g::HIPGrid{D,T}) where {D,T} = HIPField{T}(context(g), D * D, true)
S.spacescalarfield(g::HIPGrid{D,T}) where {D,T} = HIPField{T}(context(g), 1, false)
S.spacevectorfield(g::HIPGrid{D,T}) where {D,T} = HIPField{T}(context(g), D, false)
S.spacetensorfield(g::HIPGrid{D,T}) where {D,T} = HIPField{T}(context(g), D == 2 ? 3 : 6, false)

"Copy a NamedTuple of host RFFT arrays (the reference layout) to the device."
function upload!(f::HIPField{T}, u::NamedTuple) where {T}
    host = stack(map(Array, values(u)))
    check(ccall((:sc_upload_spectral, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                f.ctx.ptr, f.ptr, host, f.ncomp), f.ctx.ptr)
    f
end
"Device field -> host array of size (n/2+1, n, n, ncomp)."
function download(f::HIPField{T}) where {T}
    c = f.ctx
    host = Array{Complex{T}}(undef, c.nh, c.n, c.D == 3 ? c.n : 1, f.ncomp)
    check(ccall((:sc_download_spectral, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                c.ptr, host, f.ptr, f.ncomp), c.ptr)
    host
end

# ------------------------------------------------------------------ solver
# cache: the library owns the solver's scratch; getcache binds the context and
# allocates the stress σ that sfs! reads from the DNS cache (the reference's
# getcache holds it too); the other entries exist for the call sites that
# name them
S.getcache(g::HIPGrid) = (; ctx = context(g), σ = S.tensorfield(g), vi_vj = nothing, v = nothing, plan = nothing,
                          ustart = nothing, du = nothing)
S.statscache(g::HIPGrid) = (; ctx = context(g), field = nothing)

S.apply!(::typeof(S.twothirds!), g::HIPGrid, (u, _)) =
    check(ccall((:sc_twothirds, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), u.ctx.ptr, u.ptr, u.ncomp), u.ctx.ptr)
S.apply!(::typeof(S.project!), g::HIPGrid, (u, _)) =
    check(ccall((:sc_project, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), u.ctx.ptr, u.ptr), u.ctx.ptr)
S.apply!(::typeof(S.gaussianfilter!), g::HIPGrid, (u, Δ, _)) =
    check(ccall((:sc_gaussianfilter, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble),
                u.ctx.ptr, u.ptr, u.ncomp, Δ), u.ctx.ptr)
S.apply!(::typeof(S.cutoff!), g::HIPGrid, (ubar, u)) =
    check(ccall((:sc_cutoff, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                ubar.ctx.ptr, u.ctx.ptr, ubar.ptr, u.ptr, u.ncomp), ubar.ctx.ptr)

S.to_phys!(phys::HIPField, spec::HIPField, plan, g::HIPGrid) =
    (check(ccall((:sc_to_phys, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                 spec.ctx.ptr, phys.ptr, spec.ptr, spec.ncomp), spec.ctx.ptr); phys)
S.to_spec!(spec::HIPField, phys::HIPField, plan, g::HIPGrid) =
    (check(ccall((:sc_to_spec, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                 spec.ctx.ptr, spec.ptr, phys.ptr, phys.ncomp), spec.ctx.ptr); spec)

S.nonlinearity!(σ::HIPField, vi_vj, v, u::HIPField, plan, g::HIPGrid) =
    check(ccall((:sc_nonlinearity, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                u.ctx.ptr, σ.ptr, u.ptr), u.ctx.ptr)

function S.convectiondiffusion!(du::HIPField, u::HIPField, grid::HIPGrid, cache; visc)
    check(ccall((:sc_convectiondiffusion, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                u.ctx.ptr, du.ptr, u.ptr, visc), u.ctx.ptr)
end

function S.wray3!(::typeof(S.convectiondiffusion!), u::HIPField, Δt, grid::HIPGrid, cache; visc)
    check(ccall((:sc_wray3, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble),
                u.ctx.ptr, u.ptr, Δt, visc), u.ctx.ptr)
    u
end

function S.propose_timestep(u::HIPField, grid::HIPGrid, visc, cache)
    dt = Ref{Cdouble}(0)
    check(ccall((:sc_propose_timestep, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ref{Cdouble}),
                u.ctx.ptr, u.ptr, visc, dt), u.ctx.ptr)
    dt[]
end

function S.energy(u::HIPField)
    e = Ref{Cdouble}(0)
    check(ccall((:sc_energy, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}),
                u.ctx.ptr, u.ptr, u.ncomp, e), u.ctx.ptr)
    e[]
end

# getenergy (src/solver.jl:701) of one spectral array (a component)
function S.getenergy(u::HIPField)
    e = Ref{Cdouble}(0)
    check(ccall((:sc_energy, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{Cdouble}),
                u.ctx.ptr, u.ptr, u.ncomp, e), u.ctx.ptr)
    2 * e[]
end

function S.get_palinstrophy!(p, u::HIPField, visc, g::HIPGrid)
    e = Ref{Cdouble}(0)
    check(ccall((:sc_palinstrophy, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ref{Cdouble}),
                u.ctx.ptr, u.ptr, visc, e), u.ctx.ptr)
    e[]
end

function S.make_tracefree!(τ::HIPField, g::HIPGrid)
    f = τ.spectral ? :sc_make_tracefree_spectral : :sc_make_tracefree_physical
    check(ccall((f, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), τ.ctx.ptr, τ.ptr), τ.ctx.ptr)
    τ
end

function S.filter_reynolds(ubar::HIPField, g::HIPGrid, visc, Δ)
    r = Ref{Cdouble}(0)
    check(ccall((:sc_filter_reynolds, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Ref{Cdouble}),
                ubar.ctx.ptr, ubar.ptr, visc, Δ, r), ubar.ctx.ptr)
    r[]
end

function S.get_dissipation!(diss, u::HIPField, visc, g::HIPGrid)
    e = Ref{Cdouble}(0)
    check(ccall((:sc_dissipation, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ref{Cdouble}),
                u.ctx.ptr, u.ptr, visc, e), u.ctx.ptr)
    e[]
end

function S.spectrum(u::HIPField, grid::HIPGrid, stuff = nothing)
    kcut = div(2 * div(grid.n, 2), 3)
    s = zeros(kcut)
    check(ccall((:sc_spectrum, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}),
                u.ctx.ptr, u.ptr, u.ncomp, kcut, s), u.ctx.ptr)
    (; k = 2π / grid.l * (1:kcut), s)
end

# shell clamp (src/solver.jl:1135-1152): shells are 1..m
function S.energy_shells(grid::HIPGrid, shells, u::HIPField)
    m = length(shells)
    e = zeros(m + 1)
    check(ccall((:sc_shell_energies, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cdouble}),
                u.ctx.ptr, u.ptr, u.ncomp, m + 1, e), u.ctx.ptr)
    map(i -> (; shell = i, eref = e[i+1]), shells)
end
function S.maintain_shell_energy!(u::HIPField, shells)
    eref = [s.eref for s in shells]
    check(ccall((:sc_maintain_shell_energy, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cdouble}),
                u.ctx.ptr, u.ptr, length(eref), eref), u.ctx.ptr)
    u
end

# The device draws its normals from a counter-based stream keyed by a seed
# (Julia's Xoshiro stream cannot be reproduced on the device).  The
# reference's callers pass `rng` (create_dns: `rng = Xoshiro(dns.seed)`,
# src/data.jl:39-46): the seed is then drawn from it, so distinct rngs give
# distinct fields and equal rngs equal ones; without rng, `seed` is used.
function S.randomfield(profile, grid::HIPGrid; totalenergy = 1, rng = nothing, seed = 0, kwargs...)
    seed = rng === nothing ? UInt64(seed) : rand(rng, UInt64)
    kdiag = floor(Int, sqrt(S.dim(grid)) * div(grid.n, 2))
    prof = [profile(k; kwargs...) for k in 0:kdiag]
    E0 = totalenergy .* prof ./ sum(prof)
    u = S.vectorfield(grid)
    check(ccall((:sc_randomfield, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, UInt64, Cint, Ptr{Cdouble}),
                u.ctx.ptr, u.ptr, seed, kdiag + 1, E0), u.ctx.ptr)
    u
end

# ------------------------------------------------------------------ closures
struct BuiltinClosure
    kind::Cint
    p1::Cdouble
    p2::Cdouble
end
S.create_clark(Δ, g::HIPGrid) = BuiltinClosure(1, Δ, 0)
S.create_smagorinsky(CS, Δ, g::HIPGrid) = BuiltinClosure(2, CS, Δ)
S.create_dynamic_smagorinsky(Δ, g::HIPGrid) = BuiltinClosure(3, Δ, 0)
function S.create_verstappen(C, Δ, g::HIPGrid{D}) where {D}
    D == 3 || throw(AssertionError("Q-R model is only defined in 3D"))
    BuiltinClosure(5, C, Δ)
end
S.create_bardina(s, Δ, g::HIPGrid) = BuiltinClosure(6, s, Δ)


# ------------------------------------------------------------ learned closure
"The BF16-MFMA 3x3x3 circular CNN (sc_cnn_*); parameters stay on the device."
mutable struct CNN
    ptr::Ptr{Cvoid}
    n::Int
end
function CNN(n; batch = 1, cin = 4, width = 64, cout = 6, seed = 0, tracefree = cout == 6)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:sc_cnn_create, libsymcode), Cint, (Cint, Cint, Cint, Cint, Cint, UInt64, Ref{Ptr{Cvoid}}),
               n, batch, cin, width, cout, seed, out)
    rc == 0 || error("sc_cnn_create failed ($rc)")
    net = CNN(out[], n)
    finalizer(x -> ccall((:sc_cnn_destroy, libsymcode), Cint, (Ptr{Cvoid},), x.ptr), net)
    # the `symm` trace-removal layer of the reference's nets (src/nets.jl:275-293)
    if tracefree
        ccall((:sc_cnn_set_tracefree, libsymcode), Cint, (Ptr{Cvoid}, Cint), net.ptr, 1) == 0 ||
            error("sc_cnn_set_tracefree failed")
    end
    net
end

# One step of the reference's optimiser, OptimiserChain(ClipNorm(clip),
# AdamW(lr)) (src/training.jl:50), on the gradient accumulated on the device;
# the caller supplies lr from the warm-up + cosine schedule (:64-70).
adamw!(net::CNN, lr; clip = 1.0, β = (0.9, 0.999), ϵ = 1e-8, λ = 0.0) =
    ccall((:sc_cnn_adamw, libsymcode), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble),
          net.ptr, lr, β[1], β[2], ϵ, λ, clip) == 0 || error("sc_cnn_adamw failed")

"""
The network as an LES model, the counterpart of `fullchain`
(src/nets.jl:311-334): tau = s_out * net(u_bar / s_in, z(Re_Delta)).
`les!` and the RK4 steps take it like a built-in closure (kind 4).
"""
struct CNNClosure
    net::CNN
    s_in::Cdouble
    s_out::Cdouble
    extra::Cdouble
end
kind(m::BuiltinClosure) = m.kind
kind(::CNNClosure) = Cint(4)
kind(::Nothing) = Cint(0)
attach(m, ctx) = nothing
attach(m::CNNClosure, ctx) = check(ccall((:sc_set_cnn_closure, libsymcode), Cint,
                                         (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble),
                                         ctx.ptr, m.net.ptr, m.s_in, m.s_out, m.extra), ctx.ptr)


# ------------------------------------------------------------ pointwise MLP closure
"The reference's pointwise `mlp` (src/nets.jl:247-298) on the device (sc_pmlp_*)."
mutable struct PointwiseMLP
    ptr::Ptr{Cvoid}
    D::Int
end
# tbnn = true: the tensor-basis network (tbnn_net, src/nets.jl:535-556);
# equi = true: equivariant_net (src/nets.jl:102-238, compact trainables)
function PointwiseMLP(D, nchan; use_redelta = false, same_as_equi = false, seed = 0, nvox_max = 1 << 21,
                      tbnn = false, equi = false)
    w = same_as_equi ? nchan .* (D == 3 ? 48 : 8) : nchan
    out = Ref{Ptr{Cvoid}}(C_NULL)
    rc = if tbnn
        ccall((:sc_tbnn_create, libsymcode), Cint, (Cint, Cint, Cint, Cint, Cint, UInt64, Clong, Ref{Ptr{Cvoid}}),
              D, use_redelta, w[1], w[2], w[3], seed, nvox_max, out)
    elseif equi
        ccall((:sc_equi_create, libsymcode), Cint, (Cint, Cint, Cint, Cint, Cint, UInt64, Clong, Ref{Ptr{Cvoid}}),
              D, use_redelta, w[1], w[2], w[3], seed, nvox_max, out)
    else
        ccall((:sc_pmlp_create, libsymcode), Cint, (Cint, Cint, Cint, Cint, Cint, UInt64, Clong, Ref{Ptr{Cvoid}}),
              D, use_redelta, w[1], w[2], w[3], seed, nvox_max, out)
    end
    rc == 0 || error("network creation failed ($rc)")
    net = PointwiseMLP(out[], D)
    finalizer(x -> ccall((:sc_pmlp_destroy, libsymcode), Cint, (Ptr{Cvoid},), x.ptr), net)
    net
end

# Lux parameters of `mlp` (layers lift, mid_1, mid_2, sink; each weight
# (1,…,1, c_in, c_out)) in the library's flat order [W1, b1, W2, b2, W3, b3, W4]
function setparams!(net::PointwiseMLP, ps)
    # every leaf array in Lux's order, column-major: mlp / tbnn_net weights
    # and biases; equivariant_net's compact lift (weight, bias[, weight_re]),
    # mid_1, mid_2 and sink arrays
    h = Float32[]
    for layer in values(ps), leaf in values(layer)
        append!(h, vec(leaf))
    end
    ccall((:sc_pmlp_set_params, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Float32}), net.ptr, h) == 0 ||
        error("sc_pmlp_set_params failed (parameter count?)")
    net
end

adamw!(net::PointwiseMLP, lr; clip = 1.0, β = (0.9, 0.999), ϵ = 1e-8, λ = 0.0) =
    ccall((:sc_pmlp_adamw, libsymcode), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble, Cdouble),
          net.ptr, lr, β[1], β[2], ϵ, λ, clip) == 0 || error("sc_pmlp_adamw failed")

"""
`fullchain(setup, net, ps, st, Δ; use_redelta, redelta_norm)` (src/nets.jl:311-334)
for a `PointwiseMLP`: les! runs it as closure kind 7 (FP64 inference, live Re_Δ).
"""
struct MLPClosure
    net::PointwiseMLP
    Δ::Cdouble
    visc::Cdouble
    μ::Cdouble
    σ::Cdouble
    symmetrize::Bool  # symmetrize_pointwise (src/nets.jl:350-389) of an mlp
end
MLPClosure(net::PointwiseMLP, Δ; visc = 0.0, redelta_norm = (; μ = 0.0, σ = 1.0), symmetrize = false) =
    MLPClosure(net, Δ, visc, redelta_norm.μ, redelta_norm.σ, symmetrize)
kind(::MLPClosure) = Cint(7)
const LESModel = Union{BuiltinClosure,CNNClosure,MLPClosure}
function attach(m::MLPClosure, ctx)
    check(ccall((:sc_set_pmlp_closure, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Cdouble),
                ctx.ptr, m.net.ptr, m.Δ, m.visc, m.μ, m.σ), ctx.ptr)
    m.symmetrize && check(ccall((:sc_set_pmlp_symmetrize, libsymcode), Cint, (Ptr{Cvoid}, Cint), ctx.ptr, 1), ctx.ptr)
end

# les! and wray3!(les!, …) (src/les.jl:4-33, src/solver.jl:968-1012): one
# method each (keyword types do not take part in dispatch), the model's kind
# and parameters passed to the library, a network attached first
params(m::BuiltinClosure) = (m.p1, m.p2)
params(m) = (0.0, 0.0)
function S.les!(du::HIPField, u::HIPField, grid::HIPGrid, cache; model, visc)
    attach(model, u.ctx)
    p1, p2 = params(model)
    check(ccall((:sc_les_rhs, libsymcode), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Cdouble, Cdouble),
                u.ctx.ptr, du.ptr, u.ptr, visc, kind(model), p1, p2), u.ctx.ptr)
end

function S.wray3!(::typeof(S.les!), u::HIPField, Δt, grid::HIPGrid, cache; model, visc)
    attach(model, u.ctx)
    p1, p2 = params(model)
    check(ccall((:sc_les_wray3, libsymcode), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cint, Cdouble, Cdouble),
                u.ctx.ptr, u.ptr, Δt, visc, kind(model), p1, p2), u.ctx.ptr)
    u
end

# ------------------------------------------------------------ adjoint
# Differentiable RK4 step for the training loop (src/training.jl:45-129):
# out-of-place, fixed dt, optional closure (none or a CNNClosure).  The
# pullback is the hand-written adjoint (sc_rk4_vjp / sc_les_rk4_vjp) under
# the RFFT-weighted inner product of spectraldot (src/solver.jl:708); with a
# CNNClosure the network's parameter gradient accumulates on the device
# (sc_cnn_zero_grad / sc_cnn_get_grads, then adamw!).
using ChainRulesCore: ChainRulesCore, NoTangent, unthunk

function copyfield(u::HIPField{T}) where {T}
    v = HIPField{T}(u.ctx, u.ncomp, u.spectral)
    bytes = u.ncomp * (u.spectral ? 2 * u.ctx.spec_elems : u.ctx.phys_elems) * sizeof(T)
    ccall((:hipMemcpy, "libamdhip64"), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint), v.ptr, u.ptr, bytes, 3)
    v
end

function rk4_step(u::HIPField, Δt, grid::HIPGrid; visc, model = nothing)
    v = copyfield(u)
    if model === nothing
        check(ccall((:sc_rk4, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble),
                    u.ctx.ptr, v.ptr, Δt, visc), u.ctx.ptr)
    else
        attach(model, u.ctx)
        p1, p2 = params(model)
        check(ccall((:sc_les_rk4, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cint, Cdouble, Cdouble),
                    u.ctx.ptr, v.ptr, Δt, visc, kind(model), p1, p2), u.ctx.ptr)
    end
    v
end

function ChainRulesCore.rrule(::typeof(rk4_step), u::HIPField, Δt, grid::HIPGrid; visc, model = nothing)
    v = rk4_step(u, Δt, grid; visc, model)
    function rk4_step_pullback(v̄)
        w = unthunk(v̄)
        ū = HIPField{eltype(u)}(u.ctx, u.ncomp, true)
        if model === nothing
            check(ccall((:sc_rk4_vjp, libsymcode), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble),
                        u.ctx.ptr, ū.ptr, u.ptr, w.ptr, Δt, visc), u.ctx.ptr)
        else
            attach(model, u.ctx)
            check(ccall((:sc_les_rk4_vjp, libsymcode), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cint),
                        u.ctx.ptr, ū.ptr, u.ptr, w.ptr, Δt, visc, kind(model)), u.ctx.ptr)
        end
        return NoTangent(), ū, NoTangent(), NoTangent()
    end
    v, rk4_step_pullback
end
Base.eltype(::HIPField{T}) where {T} = T

end # module
