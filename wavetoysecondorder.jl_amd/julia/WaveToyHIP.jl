# WaveToyHIP.jl -- ccall binding of libwt2.so (include/wt2.h) for
# WaveToySecondOrder.jl callers.  Not exercised in this repository (no Julia
# toolchain here); it is the binding a maintainer adds next to
# src/wave.jl:573-611 so that `rhs_wave3d!` and the RK4 step dispatch to the
# MI355X kernels when the arrays live on an AMDGPU (ROCArray) device.
module WaveToyHIP

const libwt2 = get(ENV, "WT2_LIB", joinpath(@__DIR__, "..", "wavetoy_hip", "libwt2.so"))

const WT2_HOST_PTRS = Cint(1)

struct MeshDesc
    N::Int32
    M::NTuple{3,Int32}
    x0::NTuple{3,Float64}
    x1::NTuple{3,Float64}
    periodic::NTuple{3,Int32}
    warp_A::Float64
    tau::Float64
    eps_ko::Float64
    bdry_values::NTuple{6,Float64}
    precision::Int32      # 0 = FP64, 1 = mixed
    flags::Int32
    face_kind::NTuple{6,Int32}   # 0 = Dirichlet, 7 = SOMMERFELD_BDRY_TAG
    sommerfeld_inv_R::Float64    # 1/R; 0 for R = Inf
end

mutable struct Context
    ptr::Ptr{Cvoid}
    npoints::Int
end

function check(rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:wt2_last_error, libwt2), Cstring, ()))
    error("libwt2 error $rc: $msg")
end

"""
    Context(N, M; x0=0.0, x1=1.0, periodic=false, warp_A=0.0, τ, ε_KO=0.0,
            bdry_values=ntuple(_->0.0, 6), precision=:fp64,
            sommerfeld_faces=(), sommerfeld_R=Inf)

The mesh + operators + geometry + workspace bundle (make_uniform_hex +
make_geometry + make_workspace + to_device of src/evolve.jl:910-936).
`sommerfeld_faces` lists the faces (1..6) tagged SOMMERFELD_BDRY_TAG.
"""
function Context(N::Integer, M::Integer; x0 = 0.0, x1 = 1.0, periodic = false,
                 warp_A = 0.0, τ, ε_KO = 0.0, bdry_values = ntuple(_ -> 0.0, 6),
                 precision = :fp64, sommerfeld_faces = (), sommerfeld_R = Inf)
    d = MeshDesc(Int32(N), ntuple(_ -> Int32(M), 3), ntuple(_ -> Float64(x0), 3),
                 ntuple(_ -> Float64(x1), 3), ntuple(_ -> Int32(periodic), 3),
                 Float64(warp_A), Float64(τ), Float64(ε_KO), NTuple{6,Float64}(bdry_values),
                 Int32(precision === :mixed ? 1 : 0), Int32(0),
                 ntuple(f -> Int32(f in sommerfeld_faces ? 7 : 0), 6),
                 isinf(sommerfeld_R) ? 0.0 : 1.0 / Float64(sommerfeld_R))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:wt2_ctx_create, libwt2), Cint, (Ref{MeshDesc}, Ref{Ptr{Cvoid}}), d, out))
    ctx = Context(out[], N^3 * M^3)
    finalizer(c -> ccall((:wt2_ctx_destroy, libwt2), Cvoid, (Ptr{Cvoid},), c.ptr), ctx)
    return ctx
end

"""
    Context(geom, conn; τ, bdry_values=ntuple(_->0.0, 6), precision=:fp64, sommerfeld_R=Inf)

The same bundle on a multi-patch mesh (make_cubed_cube_mesh /
make_inflated_cube_mesh + make_geometry + to_device, src/evolve.jl:916-936):
`geom.coords` (3,N,N,N,Ne) and `conn` = mesh.conn (`neighbour`,
`neighbour_face`, `bdry`; (6,Ne)).  The D4 orientation of every shared face is
found by the library from the face nodes, so HexMeshes' `orientation` codes
are not passed.  Outer faces with bdry 7 are Sommerfeld faces, those with
bdry 8 (EXCISION_BDRY_TAG) excision faces that carry no SAT.
"""
function Context(geom, conn; τ, bdry_values = ntuple(_ -> 0.0, 6), precision = :fp64,
                 sommerfeld_R = Inf)
    N = size(geom.coords, 2)
    Ne = size(geom.coords, 5)
    d = MeshDesc(Int32(N), (Int32(1), Int32(1), Int32(1)), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0),
                 (Int32(0), Int32(0), Int32(0)), 0.0, Float64(τ), 0.0, NTuple{6,Float64}(bdry_values),
                 Int32(precision === :mixed ? 1 : 0), Int32(0), ntuple(_ -> Int32(0), 6),
                 isinf(sommerfeld_R) ? 0.0 : 1.0 / Float64(sommerfeld_R))
    coords = Array{Float64}(geom.coords)
    nbr = Array{Int32}(conn.neighbour); nface = Array{Int32}(conn.neighbour_face)
    bdry = Array{Int32}(conn.bdry)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:wt2_ctx_create_mesh, libwt2), Cint,
                (Ref{MeshDesc}, Int64, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
                 Ref{Ptr{Cvoid}}),
                d, Ne, coords, nbr, nface, C_NULL, bdry, out))
    ctx = Context(out[], N^3 * Ne)
    finalizer(c -> ccall((:wt2_ctx_destroy, libwt2), Cvoid, (Ptr{Cvoid},), c.ptr), ctx)
    return ctx
end

"""
    rhs_wave3d!(ü, u, u̇, bdry_values; ctx) -> ü

Same arguments and result as WaveToySecondOrder.rhs_wave3d! (src/wave.jl:583-611);
host arrays are copied through, device pointers (ROCArray) are used in place.
"""
function rhs_wave3d!(ü::Array{Float64,4}, u::Array{Float64,4}, u̇::Array{Float64,4},
                     bdry_values::NTuple{6,Float64}; ctx::Context)
    @assert size(ü) == size(u) == size(u̇) && length(u) == ctx.npoints
    bv = Ref(bdry_values)
    check(ccall((:wt2_rhs3d, libwt2), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{NTuple{6,Float64}}, Cint),
                ctx.ptr, ü, u, u̇, bv, WT2_HOST_PTRS))
    return ü
end

"""
    rk4_steps!(u, u̇, dt, nsteps; ctx, energy_every=0) -> energies or nothing

nsteps classical RK4 steps in place (the step! loop of src/evolve.jl:1028-1031).
"""
function rk4_steps!(u::Array{Float64,4}, u̇::Array{Float64,4}, dt::Float64, nsteps::Integer;
                    ctx::Context, energy_every::Integer = 0)
    ne = energy_every > 0 ? 1 + nsteps ÷ energy_every : 0
    E = zeros(max(ne, 1))
    check(ccall((:wt2_rk4_steps, libwt2), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cint, Cint, Ptr{Float64}, Cint),
                ctx.ptr, u, u̇, dt, nsteps, energy_every, ne > 0 ? E : C_NULL, WT2_HOST_PTRS))
    return ne > 0 ? E[1:ne] : nothing
end

# ---- first-order (Φ, Π) path (wt2_fo_*)

const _BC = Dict(:dirichlet => Int32(1), :sommerfeld => Int32(2), :excision => Int32(3),
                 :full_dirichlet => Int32(4))

struct BC3d
    kinds::NTuple{6,Int32}
    excision_tag::Int32      # general meshes: outer faces with this bdry tag get no SAT
    σ::Float64
    gΦ::Ptr{Float64}
    gΠ::Ptr{Float64}
    gD::Ptr{Float64}         # general meshes, :dirichlet: the ∇Φ target (3 fields)
end

const EXCISION_BDRY_TAG = Int32(8)

"""
    make_bc3d(kinds; σ = 1, gΦ = nothing, gΠ = nothing, gD = nothing, excision_tag = 0)

make_bc3d of src/boundaries3d.jl:44-54.  `kinds` is one kind per face on an
affine mesh, or a single kind for every outer face of a general mesh
(_apply_bc3d_curv!, src/boundaries3d.jl:131-165), whose faces tagged
`excision_tag` get no SAT.  Targets are host arrays (kept alive across the
call) or device pointers (`Ptr{Float64}`), to match the call's flags.
"""
function make_bc3d(kinds; σ = 1.0, gΦ = nothing, gΠ = nothing, gD = nothing, excision_tag = 0)
    ks = kinds isa Symbol || kinds isa Integer ? ntuple(_ -> kinds, 6) : kinds
    ptr(a) = a === nothing ? Ptr{Float64}(C_NULL) : a isa Ptr ? Ptr{Float64}(a) : pointer(a)
    BC3d(ntuple(i -> ks[i] isa Symbol ? _BC[ks[i]] : Int32(ks[i]), 6), Int32(excision_tag), Float64(σ),
         ptr(gΦ), ptr(gΠ), ptr(gD))
end

"""
    fo_coef(bg10; ctx) -> coef (11 fields of N^3 Ne values)

make_coef3d + sample_background3d! (src/wave3d_curved.jl:75-124) from 10
background values per node.
"""
function fo_coef(bg10::Vector{Float64}; ctx::Context)
    coef = zeros(11 * ctx.npoints)
    check(ccall((:wt2_fo_coef, libwt2), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint),
                ctx.ptr, bg10, coef, WT2_HOST_PTRS))
    return coef
end

"""
    wave3d_curved_rhs!(Φ̇, Π̇, Φ, Π, coef; ctx, ε_KO = 0.1, bc3d = nothing) -> (Φ̇, Π̇)

Same arguments and result as WaveToySecondOrder.wave3d_curved_rhs! on its affine
path (src/wave3d_curved.jl:205-258).
"""
function wave3d_curved_rhs!(Φ̇::Array{Float64,4}, Π̇::Array{Float64,4}, Φ::Array{Float64,4},
                            Π::Array{Float64,4}, coef::Vector{Float64}; ctx::Context,
                            ε_KO::Real = 0.1, bc3d = nothing)
    bc = bc3d === nothing ? Ptr{BC3d}(C_NULL) : Ref(bc3d)
    check(ccall((:wt2_fo_rhs, libwt2), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble,
                 Ptr{BC3d}, Cint),
                ctx.ptr, Φ̇, Π̇, Φ, Π, coef, Float64(ε_KO), bc, WT2_HOST_PTRS))
    return Φ̇, Π̇
end

"""
    wave3d_energy(Φ, Π, coef; ctx) -> E   (src/wave3d_curved.jl:268-287)
"""
function wave3d_energy(Φ::Array{Float64,4}, Π::Array{Float64,4}, coef::Vector{Float64}; ctx::Context)
    E = Ref(0.0)
    check(ccall((:wt2_fo_energy, libwt2), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint),
                ctx.ptr, Φ, Π, coef, E, WT2_HOST_PTRS))
    return E[]
end

"""
    fo_rk4_stage!(stage, Φ, Π, coef, dt; ctx, ε_KO = 0.1, bc3d = nothing)

One classical RK4 stage (1..4) of the first-order system in place on device
arrays (`Ptr{Float64}` to 16-byte aligned ROCArray storage): the stages of
OrdinaryDiffEq RK4() in src/evolve.jl:1309-1313, each the right-hand side
fused with its update.  Boundary targets may change between stages.
"""
function fo_rk4_stage!(stage::Integer, Φ::Ptr{Float64}, Π::Ptr{Float64}, coef::Ptr{Float64}, dt::Real;
                       ctx::Context, ε_KO::Real = 0.1, bc3d = nothing)
    bc = bc3d === nothing ? Ptr{BC3d}(C_NULL) : Ref(bc3d)
    check(ccall((:wt2_fo_rk4_stage, libwt2), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cdouble, Ptr{BC3d}),
                ctx.ptr, Int32(stage), Φ, Π, coef, Float64(dt), Float64(ε_KO), bc))
    return nothing
end

"""
    rk4_steps_noalloc!(u, u̇, dt, nsteps; ctx, energy_every=0) -> energies or nothing

rk4_steps! on device arrays (`Ptr{Float64}`) for callers that cannot allocate
device memory; no counterpart in the reference.  At most 65536 energy
samples; calls that overlap in time on one GPU race.
"""
function rk4_steps_noalloc!(u::Ptr{Float64}, u̇::Ptr{Float64}, dt::Float64, nsteps::Integer;
                            ctx::Context, energy_every::Integer = 0)
    ne = energy_every > 0 ? 1 + nsteps ÷ energy_every : 0
    E = zeros(max(ne, 1))
    check(ccall((:wt2_rk4_steps_noalloc, libwt2), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cint, Cint, Ptr{Float64}),
                ctx.ptr, u, u̇, dt, nsteps, energy_every, ne > 0 ? E : C_NULL))
    return ne > 0 ? E[1:ne] : nothing
end

# ---- batches: many independent evolutions (an ensemble of step! loops,
# src/evolve.jl:1028-1031) advanced together on one device

mutable struct Batch
    ptr::Ptr{Cvoid}
    members::Vector{Any}     # keeps contexts and arrays alive while they are in the batch
end

function Batch()
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:wt2_batch_create, libwt2), Cint, (Ref{Ptr{Cvoid}},), out))
    b = Batch(out[], Any[])
    finalizer(x -> ccall((:wt2_batch_destroy, libwt2), Cvoid, (Ptr{Cvoid},), x.ptr), b)
    return b
end

"""
    add!(b, ctx, u, u̇, dt, nsteps; energy_every = 0) -> slot

Adds one evolution on device arrays (`Ptr{Float64}`); returns its slot.
"""
function add!(b::Batch, ctx::Context, u::Ptr{Float64}, u̇::Ptr{Float64}, dt::Real, nsteps::Integer;
              energy_every::Integer = 0)
    slot = Ref{Int32}(0)
    check(ccall((:wt2_batch_add, libwt2), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Int64, Cint, Ref{Int32}),
                b.ptr, ctx.ptr, u, u̇, Float64(dt), Int64(nsteps), Int32(energy_every), slot))
    push!(b.members, ctx)
    return Int(slot[])
end

"""
    step!(b, max_steps) -> number of evolutions that finished
"""
function step!(b::Batch, max_steps::Integer)
    fin = Ref{Int32}(0)
    check(ccall((:wt2_batch_step, libwt2), Cint, (Ptr{Cvoid}, Cint, Ref{Int32}), b.ptr, Int32(max_steps), fin))
    return Int(fin[])
end

"""
    slot_state(b, slot, max_energies) -> (steps_done, energies)
"""
function slot_state(b::Batch, slot::Integer, max_energies::Integer = 0)
    done = Ref{Int64}(0)
    nE = Ref{Int32}(0)
    E = zeros(max(max_energies, 1))
    check(ccall((:wt2_batch_slot, libwt2), Cint,
                (Ptr{Cvoid}, Cint, Ref{Int64}, Ptr{Float64}, Cint, Ref{Int32}),
                b.ptr, Int32(slot), done, max_energies > 0 ? E : C_NULL, Int32(max_energies), nE))
    return Int(done[]), E[1:min(Int(nE[]), max_energies)]
end

remove!(b::Batch, slot::Integer) =
    check(ccall((:wt2_batch_remove, libwt2), Cint, (Ptr{Cvoid}, Cint), b.ptr, Int32(slot)))

end # module
