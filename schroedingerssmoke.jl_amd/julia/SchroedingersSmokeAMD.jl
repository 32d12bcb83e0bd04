# SchroedingersSmokeAMD.jl -- Julia binding of libisf (include/isf.h), the MI355X (gfx950)
# implementation of the SchroedingersSmoke.jl ISF step.
#
# It keeps the names and call forms of the reference's exports (src/parallel.jl:264-266,
# src/particles.jl:97) so the driver of test/runtests.jl (and examples/benchmark.jl) runs
# unchanged apart from `ArrayType = ISFArray`, loaded under the reference's package name
# through SchroedingersSmoke.jl beside this file.  Every call goes through `ccall` into
# libisf; device arrays are `ISFArray`s (device pointer + shape, freed by a finalizer).
#
# Julia is not installed in the build container, so this file is not exercised by the
# test suite; the same C ABI is exercised through the Python mirror (isf_amd).
module SchroedingersSmokeAMD

export ISF, ISFArray, normalize_psi!, pressure_project!, velocity_one_form!, schroedinger_flow!
export staggered_advect!, add_particles!, restrict_velocity!, isjet, synchronize
export ISFBatch, ISFRollout, rollout_loss, rollout_gradient, normalize_psi, set_spawn_seed!, simloop!
export Complex64, GPUArrays

const libisf = joinpath(@__DIR__, "..", "lib", "libisf.so")

# The Julia 0.6 name the reference drivers use for the element type of psi
# (test/runtests.jl:19, examples/benchmark.jl:20).
const Complex64 = ComplexF32

function check(rc::Cint, what::AbstractString)
    if rc != 0
        msg = unsafe_string(ccall((:isf_last_error, libisf), Cstring, ()))
        error("$what: libisf error $rc: $msg")
    end
    nothing
end

# ---------------------------------------------------------------- device arrays
"""
    ISFArray{T,N}

Device array in the reference's column-major layout (x fastest), the counterpart of
`ArrayType(hostarray)` / `Array(devarray)` (test/runtests.jl:19,33, examples/jet.jl:83).
"""
mutable struct ISFArray{T,N} <: AbstractArray{T,N}
    ptr::Ptr{Cvoid}
    dims::NTuple{N,Int}
    function ISFArray{T,N}(dims::NTuple{N,Int}) where {T,N}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:isf_malloc, libisf), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, prod(dims) * sizeof(T)), "isf_malloc")
        a = new{T,N}(p[], dims)
        finalizer(x -> ccall((:isf_free, libisf), Cint, (Ptr{Cvoid},), x.ptr), a)
        a
    end
end
Base.size(a::ISFArray) = a.dims
Base.getindex(::ISFArray, ::Int...) = error("scalar indexing of an ISFArray; copy it to the host with Array(a)")
function ISFArray(h::Array{T,N}) where {T,N}
    a = ISFArray{T,N}(size(h))
    check(ccall((:isf_memcpy_h2d, libisf), Cint, (Ptr{Cvoid}, Ptr{T}, Csize_t), a.ptr, h, sizeof(h)), "isf_memcpy_h2d")
    a
end
function Base.Array(a::ISFArray{T,N}) where {T,N}
    h = Array{T,N}(undef, a.dims)
    check(ccall((:isf_memcpy_d2h, libisf), Cint, (Ptr{T}, Ptr{Cvoid}, Csize_t), h, a.ptr, sizeof(h)), "isf_memcpy_d2h")
    h
end
devptr(a::ISFArray) = a.ptr
devptr(::Nothing) = C_NULL

# Broadcasts over ISFArrays that libisf has no kernel for (everything but the reference's
# normalize_psi, isjet and StaggeredSharp forms below) run on the host: the device
# operands are copied there, the broadcast is evaluated by Base, and the result comes
# back as an ISFArray.  Logical-mask indexing works the same way, so host-style lines
# such as `psi1[isJet] = amp1[isJet] .* exp.(1.0im*phase[isJet])`
# (examples/matlabport/simple.jl:107-110) also accept device arrays.
struct ISFStyle <: Base.Broadcast.AbstractArrayStyle{Any} end
ISFStyle(::Val) = ISFStyle()
Base.BroadcastStyle(::Type{<:ISFArray}) = ISFStyle()
Base.BroadcastStyle(::ISFStyle, ::Base.Broadcast.DefaultArrayStyle) = ISFStyle()
Base.BroadcastStyle(::ISFStyle, ::Base.Broadcast.Style{Tuple}) = ISFStyle()
hostify(x::ISFArray) = Array(x)
hostify(bc::Base.Broadcast.Broadcasted) = Base.Broadcast.broadcasted(bc.f, map(hostify, bc.args)...)
hostify(x) = x
Base.Broadcast.materialize(bc::Base.Broadcast.Broadcasted{ISFStyle}) =
    ISFArray(Base.Broadcast.materialize(hostify(bc)))
function Base.Broadcast.materialize!(dest::ISFArray, bc::Base.Broadcast.Broadcasted)
    h = Array(dest)
    Base.Broadcast.materialize!(h, hostify(bc))
    upload!(dest, h)
end
function upload!(dest::ISFArray{T}, h::Array{T}) where {T}
    check(ccall((:isf_memcpy_h2d, libisf), Cint, (Ptr{Cvoid}, Ptr{T}, Csize_t), dest.ptr, h, sizeof(h)),
          "isf_memcpy_h2d")
    dest
end
Base.getindex(a::ISFArray, mask::AbstractArray{Bool}) = Array(a)[hostify(mask)]
function Base.setindex!(a::ISFArray, v, mask::AbstractArray{Bool})
    h = Array(a)
    h[hostify(mask)] = hostify(v)
    upload!(a, h)
end

# ---------------------------------------------------------------- ISF
"""
    ISF{ISFArray, UInt32, Float32}(physical_size, dims, hbar, dt)

The reference constructor (src/parallel.jl:41-98).  Fields read by the drivers:
`d`, `dt`, `hbar`, `grid_res`, `velocity` (test/runtests.jl:25,44,59-63).
"""
mutable struct ISF{A,I,F}
    h::Ptr{Cvoid}
    grid_res::NTuple{3,Int}
    physical_size::NTuple{3,Int}
    d::NTuple{3,Float32}
    hbar::Float32
    dt::Float32
    function ISF{A,I,F}(physical_size, dims, hbar, dt) where {A,I,F}
        h = Ref{Ptr{Cvoid}}(C_NULL)
        res = Cint.(collect(dims))
        box = Cint.(collect(physical_size))
        check(ccall((:isf_create, libisf), Cint, (Ref{Ptr{Cvoid}}, Ptr{Cint}, Ptr{Cint}, Cfloat, Cfloat),
                    h, res, box, Float32(hbar), Float32(dt)), "isf_create")
        d = Float32.(Tuple(collect(physical_size) ./ collect(dims)))
        isf = new{A,I,F}(h[], Tuple(dims), Tuple(physical_size), d, Float32(hbar), Float32(dt))
        finalizer(x -> ccall((:isf_destroy, libisf), Cint, (Ptr{Cvoid},), x.h), isf)
        CURRENT[] = isf  # the grid the reference's psi broadcasts refer to
        isf
    end
end

const CURRENT = Ref{Any}(nothing)

"`isf.velocity` (src/parallel.jl:14): the handle's NTuple{3,Float32} field (borrowed)."
struct ISFVelocity
    isf::Any
end
"`isf.positions` (src/parallel.jl:75-77): cell positions, only meaningful inside `isjet.(...)`."
struct ISFPositions
    isf::Any
end
function devptr(v::ISFVelocity)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:isf_velocity_ptr, libisf), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), v.isf.h, p), "isf_velocity_ptr")
    p[]
end
function Base.Array(v::ISFVelocity)
    h = Array{NTuple{3,Float32},3}(undef, v.isf.grid_res)
    check(ccall((:isf_memcpy_d2h, libisf), Cint, (Ptr{NTuple{3,Float32}}, Ptr{Cvoid}, Csize_t), h, devptr(v),
                sizeof(h)), "isf_memcpy_d2h")
    h
end

function Base.getproperty(isf::ISF, s::Symbol)
    s === :velocity && return ISFVelocity(isf)
    s === :positions && return ISFPositions(isf)
    getfield(isf, s)
end

synchronize(isf::ISF) = check(ccall((:isf_synchronize, libisf), Cint, (Ptr{Cvoid},), isf.h), "isf_synchronize")

# ---------------------------------------------------------------- reference operators
# psi is an ISFArray{NTuple{2,ComplexF32},3}; particles an ISFArray{NTuple{3,Float32},1}.
schroedinger_flow!(isf::ISF, psi) =
    (check(ccall((:isf_schroedinger_flow, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), isf.h, devptr(psi)),
           "schroedinger_flow!"); psi)

"`psi .= normalize_psi.(psi)` (src/parallel.jl:162-165, test/runtests.jl:49)."
normalize_psi!(isf::ISF, psi) =
    (check(ccall((:isf_normalize, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), isf.h, devptr(psi)), "normalize_psi"); psi)

pressure_project!(isf::ISF, psi) =
    (check(ccall((:isf_pressure_project, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), isf.h, devptr(psi)),
           "pressure_project!"); psi)

"`velocity_one_form!(isf, psi, hbar)` writes isf.velocity; `sharp=true` also applies the driver's
in-place StaggeredSharp (test/runtests.jl:61-62)."
velocity_one_form!(isf::ISF, psi, hbar = 1f0; sharp::Bool = false) =
    check(ccall((:isf_velocity_one_form, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cfloat, Cint, Ptr{Cvoid}),
                isf.h, devptr(psi), Float32(hbar), Cint(sharp), C_NULL), "velocity_one_form!")

"`isjet.(isf.positions, cen, len, rad)` (src/particles.jl:6-10, test/runtests.jl:25)."
function isjet(isf::ISF, nozzle_cen, nozzle_len, nozzle_rad)
    out = ISFArray{UInt8,3}(isf.grid_res)
    check(ccall((:isf_isjet, libisf), Cint, (Ptr{Cvoid}, Ptr{Cfloat}, Cfloat, Cfloat, Ptr{Cvoid}),
                isf.h, Float32.(collect(nozzle_cen)), Float32(nozzle_len), Float32(nozzle_rad), out.ptr), "isjet")
    out
end

"`restrict_velocity!(isf, psi, kvec, isjet, omgterm)` (src/particles.jl:25-27); `phase_mode` 1 (default,
the reference's own call) reproduces its `tuple_dot` sum, 0 the k.x phase of the MatlabPort driver."
restrict_velocity!(isf::ISF, psi, kvec, isjetarr, omgterm = 0f0; phase_mode::Integer = 1) =
    (check(ccall((:isf_restrict_velocity, libisf), Cint,
                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cfloat}, Cfloat, Cint),
                 isf.h, devptr(psi), devptr(isjetarr), Float32.(collect(kvec)), Float32(omgterm), Cint(phase_mode)),
           "restrict_velocity!"); psi)

"`add_particles!(isf, particles, range, cen, rad; seed, step)` (src/particles.jl:34-94): the
draws of `range` from Philox4x32-10 with counter (j, step) and key `seed`."
add_particles!(isf::ISF, particles::ISFArray, range::UnitRange, nozzle_cen, nozzle_rad; seed = 0, step = 0) =
    check(ccall((:isf_add_particles, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Clonglong, Clonglong, Clonglong, Ptr{Cfloat}, Cfloat, Culonglong, Cuint),
                isf.h, particles.ptr, length(particles), first(range) - 1, length(range),
                Float32.(collect(nozzle_cen)), Float32(nozzle_rad), UInt64(seed), UInt32(step)), "add_particles!")

# The reference's own call form `add_particles!(particles, start:stop, nozzle_cen, nozzle_rad)`
# (src/particles.jl:60, test/runtests.jl:54) on the ISF created last; its unseeded rand
# becomes Philox draws with the seed of set_spawn_seed! and the call number as the step.
const SPAWN = Ref((seed = UInt64(20261203), calls = UInt32(0)))
"Seed of the reference-form `add_particles!` (restarts its call counter)."
set_spawn_seed!(seed::Integer) = (SPAWN[] = (seed = UInt64(seed), calls = UInt32(0)); nothing)
function add_particles!(particles::ISFArray, range::UnitRange, nozzle_cen, nozzle_rad)
    s = SPAWN[]
    SPAWN[] = (seed = s.seed, calls = s.calls + UInt32(1))
    add_particles!(CURRENT[], particles, range, nozzle_cen, nozzle_rad; seed = s.seed, step = s.calls + UInt32(1))
    particles
end

"`staggered_advect!(particles, isf)` (src/parallel.jl:248-261), using isf.velocity."
staggered_advect!(particles::ISFArray, isf::ISF) =
    (check(ccall((:isf_staggered_advect, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Clonglong, Ptr{Cvoid}),
                 isf.h, particles.ptr, length(particles), C_NULL), "staggered_advect!"); particles)

# ---------------------------------------------------------------- fused driver step
"Configure the jet of the driver (test/runtests.jl:7-12,23-25) for `step!`."
jet_configure!(isf::ISF, kvec, omega, nozzle_cen, nozzle_len, nozzle_rad, n_spawn; seed = 0, phase_mode = 0) =
    check(ccall((:isf_jet_configure, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cfloat}, Cfloat, Ptr{Cfloat}, Cfloat, Cfloat, Clonglong, Culonglong, Cint),
                isf.h, Float32.(collect(kvec)), Float32(omega), Float32.(collect(nozzle_cen)), Float32(nozzle_len),
                Float32(nozzle_rad), n_spawn, UInt64(seed), Cint(phase_mode)), "isf_jet_configure")

"One iteration of the simloop body (test/runtests.jl:45-64)."
step!(isf::ISF, psi, particles::ISFArray, iter::Integer) =
    check(ccall((:isf_step, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Clonglong, Cint),
                isf.h, devptr(psi), particles.ptr, length(particles), Cint(iter)), "isf_step")
"""
    simloop!(isf, psi, particles, first_iter, count)

The driver's `simloop(N, …)` (test/runtests.jl:40-66): `count` iterations from `first_iter`
in one call (`isf_simloop`), equal to `count` `step!` calls.
"""
simloop!(isf::ISF, psi, particles::ISFArray, first_iter::Integer, count::Integer) =
    check(ccall((:isf_simloop, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Clonglong, Cint, Cint),
                isf.h, devptr(psi), particles.ptr, length(particles), Cint(first_iter), Cint(count)), "isf_simloop")

# ---------------------------------------------------------------- the reference's broadcast call forms
# The reference drivers write (test/runtests.jl:21,25,49,61-62)
#     psi .= normalize_psi.(psi)
#     jet = isjet.(isf.positions, nozzle_cen, nozzle_len, nozzle_rad)
#     broadcast!((x, y) -> x .* y, isf.velocity, isf.velocity, (inv.(isf.d),))
# Each lowers to Base.broadcasted / materialize(!) / broadcast!, which are specialised
# here so that the whole-array form runs as one libisf call: isf_normalize, isf_isjet,
# isf_staggered_sharp.  The psi broadcast uses the grid of the most recently created ISF.

"`normalize_psi` of the reference (src/parallel.jl:162-165); here only its broadcast form."
function normalize_psi end
struct NormalizeBC{A}
    psi::A
end
Base.Broadcast.broadcasted(::typeof(normalize_psi), psi::ISFArray) = NormalizeBC(psi)
function Base.Broadcast.materialize!(dest::ISFArray, bc::NormalizeBC)
    if dest !== bc.psi
        check(ccall((:isf_memcpy_d2d, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), dest.ptr, bc.psi.ptr,
                    prod(size(dest)) * sizeof(eltype(dest))), "isf_memcpy_d2d")
    end
    normalize_psi!(CURRENT[], dest)
    dest
end

struct IsjetBC
    isf::Any
    args::Tuple
end
# The reference passes the scalars wrapped as 1-tuples so they broadcast as scalars,
# `isjet.(isf.positions, (nozzle_cen,), (nozzle_len,), (nozzle_rad,))` (test/runtests.jl:25);
# Ref(x) is the modern spelling of the same thing.
unwrap1(x) = x
unwrap1(x::Tuple{Any}) = x[1]
unwrap1(x::Base.RefValue) = x[]
Base.Broadcast.broadcasted(::typeof(isjet), p::ISFPositions, cen, len, rad) =
    IsjetBC(p.isf, (unwrap1(cen), unwrap1(len), unwrap1(rad)))
Base.Broadcast.materialize(bc::IsjetBC) = isjet(bc.isf, bc.args...)

# the driver's in-place StaggeredSharp: velocity .* inv.(isf.d) per component
function Base.broadcast!(f, dest::ISFVelocity, src::ISFVelocity, scale::Tuple{<:Tuple})
    dest.isf === src.isf || error("broadcast! between the velocities of two ISF instances")
    all(Float32.(scale[1]) .== inv.(dest.isf.d)) ||
        error("broadcast! on isf.velocity: only the StaggeredSharp scaling inv.(isf.d) is supported")
    check(ccall((:isf_staggered_sharp, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), dest.isf.h, C_NULL),
          "isf_staggered_sharp")
    dest
end

"GPUArrays.synchronize(array) of the drivers (test/runtests.jl:65)."
synchronize(::ISFArray) = synchronize(CURRENT[])

"The `GPUArrays.synchronize(particle)` call of the drivers (test/runtests.jl:65) resolves here."
module GPUArrays
import ..SchroedingersSmokeAMD: synchronize
end

# ---------------------------------------------------------------- batched uneven ensembles
"Mirror of `isf_member_desc` (include/isf.h)."
struct MemberDesc
    res::NTuple{3,Cint}
    box::NTuple{3,Cint}
    hbar::Cfloat
    dt::Cfloat
    kvec::NTuple{3,Cfloat}
    omega::Cfloat
    cen::NTuple{3,Cfloat}
    len::Cfloat
    rad::Cfloat
    n_spawn::Clonglong
    seed::Culonglong
    phase_mode::Cint
end

"""
    ISFBatch(members::Vector{MemberDesc})

Independent jets of mixed grid sizes stepped together (`isf_batch_*`): one hipGraph
replay per ensemble step.  `member(b, i)` gives the i-th member's handle pointer for
set-up; `bind!` attaches its psi and particle arrays.
"""
mutable struct ISFBatch
    h::Ptr{Cvoid}
    n::Int
    bound::Vector{Any}
    function ISFBatch(members::Vector{MemberDesc})
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:isf_batch_create, libisf), Cint, (Ref{Ptr{Cvoid}}, Cint, Ptr{MemberDesc}), h,
                    length(members), members), "isf_batch_create")
        b = new(h[], length(members), Any[nothing for _ in members])
        finalizer(x -> ccall((:isf_batch_destroy, libisf), Cint, (Ptr{Cvoid},), x.h), b)
        b
    end
end
function member(b::ISFBatch, i::Integer)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:isf_batch_member, libisf), Cint, (Ptr{Cvoid}, Cint, Ref{Ptr{Cvoid}}), b.h, i - 1, p),
          "isf_batch_member")
    p[]
end
function bind!(b::ISFBatch, i::Integer, psi::ISFArray, particles::ISFArray)
    check(ccall((:isf_batch_bind, libisf), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Clonglong), b.h, i - 1,
                psi.ptr, particles.ptr, length(particles)), "isf_batch_bind")
    b.bound[i] = (psi, particles)
    b
end
step!(b::ISFBatch, iter::Integer) =
    check(ccall((:isf_batch_step, libisf), Cint, (Ptr{Cvoid}, Cint), b.h, iter), "isf_batch_step")
"`count` ensemble steps from `first_iter` in one call (`isf_batch_simloop`), equal to `count` `step!` calls."
simloop!(b::ISFBatch, first_iter::Integer, count::Integer) =
    check(ccall((:isf_batch_simloop, libisf), Cint, (Ptr{Cvoid}, Cint, Cint), b.h, first_iter, count),
          "isf_batch_simloop")
tune!(b::ISFBatch, cache::AbstractString = "") =
    check(ccall((:isf_batch_tune, libisf), Cint, (Ptr{Cvoid}, Cstring), b.h, isempty(cache) ? C_NULL : cache),
          "isf_batch_tune")
synchronize(b::ISFBatch) = check(ccall((:isf_batch_synchronize, libisf), Cint, (Ptr{Cvoid},), b.h), "isf_batch_synchronize")
"The FFT schedule in use for grid `res`: x-pass tile and plane-pass elements per thread."
function schedule(b::ISFBatch, res::NTuple{3,Integer})
    r = Cint[res...]
    ty = Ref{Cint}(0)
    pe = Ref{Cint}(0)
    check(ccall((:isf_batch_tile, libisf), Cint, (Ptr{Cvoid}, Ptr{Cint}, Ref{Cint}), b.h, r, ty), "isf_batch_tile")
    check(ccall((:isf_batch_plane_elems, libisf), Cint, (Ptr{Cvoid}, Ptr{Cint}, Ref{Cint}), b.h, r, pe),
          "isf_batch_plane_elems")
    return (tile = Int(ty[]), plane_elements = Int(pe[]))
end

# ---------------------------------------------------------------- vector-Jacobian products
# One per operator (isf_vjp_*): the cotangent of the output becomes that of the input, in place.
vjp_schroedinger_flow!(isf::ISF, psibar) =
    check(ccall((:isf_vjp_schroedinger_flow, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), isf.h, devptr(psibar)), "vjp")
vjp_normalize!(isf::ISF, psi_in, psibar) =
    check(ccall((:isf_vjp_normalize, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), isf.h, devptr(psi_in),
                devptr(psibar)), "vjp")
vjp_pressure_project!(isf::ISF, psi_in, psibar) =
    check(ccall((:isf_vjp_pressure_project, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), isf.h,
                devptr(psi_in), devptr(psibar)), "vjp")

# ---------------------------------------------------------------- differentiable rollout
"Mirror of `isf_rollout_desc` (include/isf.h)."
struct RolloutDesc
    steps::Cint
    checkpoint_every::Cint
    constraint_iters::Cint
    phase_res::NTuple{3,Cint}
    splat_res::NTuple{3,Cint}
    splat_sigma::Cfloat
    splat_cutoff::Cfloat
    nslots::Clonglong
    n_init::Clonglong
    fill_seed::Culonglong
end

"""
    ISFRollout(isf, desc, target)

The driver's set-up and `steps` simloop bodies as a function of the nozzle velocity and
a coarse initial phase field, with a splat-density loss against `target`
(an `ISFArray{Float32,3}` of `desc.splat_res`).
"""
mutable struct ISFRollout
    h::Ptr{Cvoid}
    isf::ISF
    desc::RolloutDesc
    function ISFRollout(isf::ISF, desc::RolloutDesc, target::ISFArray)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:isf_rollout_create, libisf), Cint, (Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Ref{RolloutDesc}), h, isf.h,
                    desc), "isf_rollout_create")
        r = new(h[], isf, desc)
        finalizer(x -> ccall((:isf_rollout_destroy, libisf), Cint, (Ptr{Cvoid},), x.h), r)
        check(ccall((:isf_rollout_set_target, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), r.h, target.ptr),
              "isf_rollout_set_target")
        r
    end
end

"Loss of one rollout for nozzle velocity `v` and phase field `theta` (an ISFArray or nothing)."
function rollout_loss(r::ISFRollout, v::NTuple{3,<:Real}, theta)
    loss = Ref{Cdouble}(0)
    check(ccall((:isf_rollout_forward, libisf), Cint, (Ptr{Cvoid}, Ptr{Cfloat}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Cdouble}),
                r.h, Cfloat[v...], devptr(theta), C_NULL, loss), "isf_rollout_forward")
    loss[]
end

"Gradients of the last `rollout_loss`: (dv::NTuple{3,Float64}, dtheta::ISFArray or nothing)."
function rollout_gradient(r::ISFRollout, theta)
    gv = zeros(Cdouble, 3)
    gth = theta === nothing ? nothing : ISFArray{Float32,3}(size(theta))
    check(ccall((:isf_rollout_backward, libisf), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                r.h, gv, devptr(gth), C_NULL, C_NULL), "isf_rollout_backward")
    (Tuple(gv), gth)
end

# ChainRulesCore.rrule of rollout_loss, the Julia counterpart of a torch.autograd.Function:
# defined when ChainRulesCore is loaded in the session.
function __init__()
    if isdefined(Main, :ChainRulesCore)
        CRC = getfield(Main, :ChainRulesCore)
        @eval function $(CRC).rrule(::typeof(rollout_loss), r::ISFRollout, v, theta)
            y = rollout_loss(r, v, theta)
            function rollout_pullback(ybar)
                gv, gth = rollout_gradient(r, theta)
                # the theta cotangent stays on the device; it is returned for ybar == 1 only
                gth === nothing || ybar == 1 || error("rollout_loss: device theta cotangent needs ybar == 1")
                (($CRC).NoTangent(), ($CRC).NoTangent(), ybar .* gv,
                 gth === nothing ? ($CRC).NoTangent() : gth)
            end
            y, rollout_pullback
        end
    end
end

# ---------------------------------------------------------------- MatlabPort names
"""
    SchroedingersSmokeAMD.MatlabPort

The CamelCase API of the reference's MatlabPort module (src/MatlabPort.jl) over libisf's
`isf_mp_*` entry points, with that module's layout: psi1, psi2 separate complex arrays,
velocity as separate vx, vy, vz.  Device arguments are `ISFArray{ComplexF32,3}` /
`ISFArray{Float32,3}` and give device results; host `Array`s are copied to the device,
computed in Float32 and returned as host arrays, so the driver of
examples/matlabport/simple.jl keeps its call forms.
"""
module MatlabPort

import ..SchroedingersSmokeAMD: libisf, check, ISFArray, devptr
export ISF, land, SchroedingerFlow, Normalize, VelocityOneForm, StaggeredSharp, Div, PoissonSolve, GaugeTransform
export PressureProject, StaggeredVelocity, StaggeredAdvect, Particles, Keep

"land(a, b), src/MatlabPort.jl:5-8."
land(a, b) = a .& b

"""
    ISF(vol_size, vol_res, hbar, dt)

src/MatlabPort.jl:38-73.  Grid fields as in the reference (`sizex`.., `resx`.., `dx`..,
`hbar`, `dt`, host `px`, `py`, `pz`); the Schroedinger mask of BuildSchroedinger (:75-83)
lives in the device tables of the handle.
"""
mutable struct ISF
    h::Ptr{Cvoid}
    sizex::Int; sizey::Int; sizez::Int
    resx::Int; resy::Int; resz::Int
    dx::Float64; dy::Float64; dz::Float64
    hbar::Float64
    dt::Float64
    px::Array{Float64,3}; py::Array{Float64,3}; pz::Array{Float64,3}
    function ISF(vol_size::NTuple{3}, vol_res::NTuple{3}, hbar, dt)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        res = Cint[round(Int, v) for v in vol_res]
        box = Cint.(collect(vol_size))
        check(ccall((:isf_create, libisf), Cint, (Ref{Ptr{Cvoid}}, Ptr{Cint}, Ptr{Cint}, Cfloat, Cfloat),
                    h, res, box, Float32(hbar), Float32(dt)), "isf_create")
        d = collect(vol_size) ./ res
        px = [(i - 1) * d[1] for i in 1:res[1], j in 1:res[2], k in 1:res[3]]
        py = [(j - 1) * d[2] for i in 1:res[1], j in 1:res[2], k in 1:res[3]]
        pz = [(k - 1) * d[3] for i in 1:res[1], j in 1:res[2], k in 1:res[3]]
        obj = new(h[], vol_size..., res..., d..., Float64(hbar), Float64(dt), px, py, pz)
        finalizer(x -> ccall((:isf_destroy, libisf), Cint, (Ptr{Cvoid},), x.h), obj)
        CURRENT[] = obj  # Normalize and GaugeTransform take no ISF argument
        obj
    end
end
const CURRENT = Ref{Any}(nothing)
dims(obj::ISF) = (obj.resx, obj.resy, obj.resz)

# device view of an argument; host arrays are uploaded (converted to Float32 types)
dev(a::ISFArray, ::Type{T}) where {T} = (eltype(a) === T || error("expected an ISFArray{$T}"); a)
dev(a::AbstractArray, ::Type{T}) where {T} = ISFArray(convert(Array{T}, a))
ondevice(args...) = all(a -> a isa ISFArray, args)
result(ondev::Bool, outs...) = ondev ? outs : map(Array, outs)
newc(obj) = ISFArray{ComplexF32,3}(dims(obj))
newf(obj) = ISFArray{Float32,3}(dims(obj))

"SchroedingerFlow(obj, psi1, psi2), src/MatlabPort.jl:87-93."
function SchroedingerFlow(obj::ISF, psi1, psi2)
    a, b = dev(psi1, ComplexF32), dev(psi2, ComplexF32)
    o1, o2 = newc(obj), newc(obj)
    check(ccall((:isf_mp_schroedinger_flow, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                obj.h, a.ptr, b.ptr, o1.ptr, o2.ptr), "SchroedingerFlow")
    result(ondevice(psi1, psi2), o1, o2)
end

"Normalize(psi1, psi2), src/MatlabPort.jl:155-160 (on the ISF constructed last)."
function Normalize(psi1, psi2)
    obj = CURRENT[]
    a, b = dev(psi1, ComplexF32), dev(psi2, ComplexF32)
    o1, o2 = newc(obj), newc(obj)
    check(ccall((:isf_mp_normalize, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                obj.h, a.ptr, b.ptr, o1.ptr, o2.ptr), "Normalize")
    result(ondevice(psi1, psi2), o1, o2)
end

"VelocityOneForm(obj, psi1, psi2, hbar=1.0), src/MatlabPort.jl:106-126."
function VelocityOneForm(obj::ISF, psi1, psi2, hbar = 1.0)
    a, b = dev(psi1, ComplexF32), dev(psi2, ComplexF32)
    vx, vy, vz = newf(obj), newf(obj), newf(obj)
    check(ccall((:isf_mp_velocity_one_form, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cfloat, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                obj.h, a.ptr, b.ptr, Float32(hbar), vx.ptr, vy.ptr, vz.ptr), "VelocityOneForm")
    result(ondevice(psi1, psi2), vx, vy, vz)
end

"StaggeredSharp(obj, vx, vy, vz), src/MatlabPort.jl:99-104."
function StaggeredSharp(obj::ISF, vx, vy, vz)
    a, b, c = dev(vx, Float32), dev(vy, Float32), dev(vz, Float32)
    ux, uy, uz = newf(obj), newf(obj), newf(obj)
    check(ccall((:isf_mp_staggered_sharp, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                obj.h, a.ptr, b.ptr, c.ptr, ux.ptr, uy.ptr, uz.ptr), "StaggeredSharp")
    result(ondevice(vx, vy, vz), ux, uy, uz)
end

"Div(obj, vx, vy, vz), src/MatlabPort.jl:162-170."
function Div(obj::ISF, vx, vy, vz)
    a, b, c = dev(vx, Float32), dev(vy, Float32), dev(vz, Float32)
    f = newf(obj)
    check(ccall((:isf_mp_div, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                obj.h, a.ptr, b.ptr, c.ptr, f.ptr), "Div")
    result(ondevice(vx, vy, vz), f)[1]
end

"PoissonSolve(obj, f), src/MatlabPort.jl:128-138 (complex result, imaginary part 0)."
function PoissonSolve(obj::ISF, f)
    fd = dev(f, Float32)
    q = newc(obj)
    check(ccall((:isf_mp_poisson_solve, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), obj.h, fd.ptr, q.ptr),
          "PoissonSolve")
    result(ondevice(f), q)[1]
end

"GaugeTransform(psi1, psi2, q) = psi .* exp.(im .* q), src/MatlabPort.jl:140-145."
function GaugeTransform(psi1, psi2, q)
    obj = CURRENT[]
    a, b, qd = dev(psi1, ComplexF32), dev(psi2, ComplexF32), dev(q, ComplexF32)
    o1, o2 = newc(obj), newc(obj)
    check(ccall((:isf_mp_gauge_transform, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                obj.h, a.ptr, b.ptr, qd.ptr, o1.ptr, o2.ptr), "GaugeTransform")
    result(ondevice(psi1, psi2, q), o1, o2)
end

"PressureProject(obj, psi1, psi2), src/MatlabPort.jl:147-153."
function PressureProject(obj::ISF, psi1, psi2)
    a, b = dev(psi1, ComplexF32), dev(psi2, ComplexF32)
    o1, o2 = newc(obj), newc(obj)
    check(ccall((:isf_mp_pressure_project, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                obj.h, a.ptr, b.ptr, o1.ptr, o2.ptr), "PressureProject")
    result(ondevice(psi1, psi2), o1, o2)
end

"StaggeredVelocity(px, py, pz, torus, vx, vy, vz), src/MatlabPort.jl:237-273."
function StaggeredVelocity(px, py, pz, torus::ISF, vx, vy, vz)
    n = length(px)
    p = (dev(px, Float32), dev(py, Float32), dev(pz, Float32))
    v = (dev(vx, Float32), dev(vy, Float32), dev(vz, Float32))
    u = (ISFArray{Float32,1}((n,)), ISFArray{Float32,1}((n,)), ISFArray{Float32,1}((n,)))
    check(ccall((:isf_mp_staggered_velocity, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Clonglong, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                torus.h, p[1].ptr, p[2].ptr, p[3].ptr, n, v[1].ptr, v[2].ptr, v[3].ptr, u[1].ptr, u[2].ptr, u[3].ptr),
          "StaggeredVelocity")
    result(ondevice(px, py, pz), u...)
end

"""
    Particles(x, y, z, length, active)

src/MatlabPort.jl:178-184 on the device (`isf_particles_*`): capacity `length(x)`.  The
driver's start `Particles(zeros(Float32, n), zeros(Float32, n), zeros(Float32, n), 0, Int[])`
(examples/matlabport/simple.jl:74-78), or the first `length` slots filled and active.
`p.x`, `p.y`, `p.z` and `p.active` (1-based) read back host copies.
"""
mutable struct Particles
    h::Ptr{Cvoid}
    isf::ISF
    capacity::Int
    snap::Any  # host copies of x, y, z while filter! runs over the active list
    function Particles(x::AbstractVector, y::AbstractVector, z::AbstractVector, len::Integer, active::AbstractVector)
        obj = CURRENT[]
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:isf_particles_create, libisf), Cint, (Ref{Ptr{Cvoid}}, Clonglong), h, length(x)),
              "isf_particles_create")
        p = new(h[], obj, length(x), nothing)
        finalizer(q -> ccall((:isf_particles_destroy, libisf), Cint, (Ptr{Cvoid},), q.h), p)
        if len > 0
            active == 1:len || error("Particles: active must be 1:length")
            append!(p, x[1:len], y[1:len], z[1:len])
        end
        p
    end
end

function state(p::Particles)
    len, cnt = Ref{Clonglong}(0), Ref{Clonglong}(0)
    px, py, pz, pa = Ref{Ptr{Cvoid}}(C_NULL), Ref{Ptr{Cvoid}}(C_NULL), Ref{Ptr{Cvoid}}(C_NULL), Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:isf_particles_state, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Clonglong}, Ref{Clonglong}, Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}},
                 Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}),
                p.isf.h, p.h, len, cnt, px, py, pz, pa), "isf_particles_state")
    (len[], cnt[], px[], py[], pz[], pa[])
end
function download(ptr::Ptr{Cvoid}, ::Type{T}, n::Integer) where {T}
    h = Vector{T}(undef, n)
    n > 0 && check(ccall((:isf_memcpy_d2h, libisf), Cint, (Ptr{T}, Ptr{Cvoid}, Csize_t), h, ptr, n * sizeof(T)),
                   "isf_memcpy_d2h")
    h
end
Base.length(p::Particles) = Int(state(p)[1])
"""
    ActiveList

`p.active`: the 1-based active slot list read from the device.  `filter!(f, p.active)` (the
driver's own removethedead!, examples/matlabport/simple.jl:63-70) evaluates `f` on the host,
with `p.x`, `p.y`, `p.z` served from one snapshot of the device arrays, and writes the
surviving list back to the device store (isf_particles_set_active).
"""
struct ActiveList <: AbstractVector{Int}
    p::Particles
    idx::Vector{Int}
end
Base.size(a::ActiveList) = size(a.idx)
Base.getindex(a::ActiveList, i::Int) = a.idx[i]
function Base.filter!(f, a::ActiveList)
    p = a.p
    st = state(p)
    setfield!(p, :snap, (download(st[3], Float32, p.capacity), download(st[4], Float32, p.capacity),
                         download(st[5], Float32, p.capacity)))
    kept = try
        filter(f, a.idx)
    finally
        setfield!(p, :snap, nothing)
    end
    zb = Cint.(kept .- 1)
    check(ccall((:isf_particles_set_active, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cint}, Clonglong),
                p.isf.h, p.h, zb, length(zb)), "filter!(p.active)")
    empty!(a.idx)
    append!(a.idx, kept)
    a
end
function Base.getproperty(p::Particles, s::Symbol)
    s in (:h, :isf, :capacity, :snap) && return getfield(p, s)
    snap = getfield(p, :snap)
    snap !== nothing && s === :x && return snap[1]
    snap !== nothing && s === :y && return snap[2]
    snap !== nothing && s === :z && return snap[3]
    st = state(p)
    s === :length && return Int(st[1])
    s === :x && return download(st[3], Float32, p.capacity)
    s === :y && return download(st[4], Float32, p.capacity)
    s === :z && return download(st[5], Float32, p.capacity)
    s === :active && return ActiveList(p, Int.(download(st[6], Cint, st[2])) .+ 1)
    error("Particles has no field $s")
end

"append!(p, x, y, z), src/MatlabPort.jl:186-195 (refuses unless length + n < capacity)."
function Base.append!(p::Particles, x, y, z)
    a, b, c = dev(x, Float32), dev(y, Float32), dev(z, Float32)
    check(ccall((:isf_particles_append, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Clonglong),
                p.isf.h, p.h, a.ptr, b.ptr, c.ptr, length(x)), "append!")
    check(ccall((:isf_synchronize, libisf), Cint, (Ptr{Cvoid},), p.isf.h), "isf_synchronize")
    nothing
end

"removethedead!(p, vol_size) of examples/matlabport/simple.jl:63-70 (strictly inside the box) in one
device pass; not exported, since that driver defines its own over `filter!(p.active)`."
removethedead!(p::Particles, vol_size = nothing) =
    check(ccall((:isf_particles_remove_dead, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), p.isf.h, p.h), "removethedead!")

"StaggeredAdvect(p, torus, vx, vy, vz, dt), src/MatlabPort.jl:201-222, over the active slots."
function StaggeredAdvect(p::Particles, torus::ISF, vx, vy, vz, dt)
    v = (dev(vx, Float32), dev(vy, Float32), dev(vz, Float32))
    check(ccall((:isf_mp_staggered_advect, libisf), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cfloat),
                torus.h, p.h, v[1].ptr, v[2].ptr, v[3].ptr, Float32(dt)), "StaggeredAdvect")
    ondevice(vx, vy, vz) ||
        check(ccall((:isf_synchronize, libisf), Cint, (Ptr{Cvoid},), torus.h), "isf_synchronize")
    nothing
end

"Keep(particle, ind), src/MatlabPort.jl:226-230 (1-based ind): x, y, z become x[ind], y[ind], z[ind];
length and the active list are left as they are."
function Keep(p::Particles, ind)
    d = ISFArray(Cint.(collect(ind)) .- Cint(1))
    check(ccall((:isf_particles_keep, libisf), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Clonglong), p.isf.h, p.h,
                d.ptr, length(d)), "Keep")
    check(ccall((:isf_synchronize, libisf), Cint, (Ptr{Cvoid},), p.isf.h), "isf_synchronize")
    nothing
end

end # module MatlabPort

end # module
