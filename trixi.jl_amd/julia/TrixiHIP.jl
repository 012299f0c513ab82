# TrixiHIP.jl -- Julia binding of libtrixi_hip.so (include/trixi_hip.h).
#
# Makes the MI355X right-hand side a drop-in for Trixi.jl's
#   rhs_hyperbolic!(du_ode, u_ode, semi, t)   (src/semidiscretization/semidiscretization_hyperbolic.jl:578-597)
# for TreeMesh{2,3} (uniform or refined with L2 mortars) +
# CompressibleEulerEquations{2,3}D + DGSEM with periodic boundaries,
# BoundaryConditionDirichlet(initial_condition) or boundary_condition_slip_wall. The ODEProblem keeps `semi` as its parameter, so the 2N
# integrators and the Stepsize/Analysis callbacks run unchanged; only the RHS
# function is replaced by a callable that ccalls the library.
#
# Julia is not installed in the build image, so this file is shipped but not
# executed there; the C++/Python harness of the repository makes exactly these
# calls.
module TrixiHIP

using Trixi
using Trixi: TreeMesh, DGSEM, CompressibleEulerEquations2D, CompressibleEulerEquations3D,
             VolumeIntegralWeakForm, VolumeIntegralFluxDifferencing, VolumeIntegralShockCapturingHGType,
             VolumeIntegralPureLGLFiniteVolume, FluxLaxFriedrichs, flux_ranocha, flux_central, flux_hllc,
             max_abs_speed, max_abs_speed_naive, source_terms_convergence_test, SolutionAnalyzer,
             mesh_equations_solver_cache, nelements, ninterfaces, nboundaries, nmortars,
             BoundaryConditionDirichlet, boundary_condition_slip_wall, initial_condition_constant,
             initial_condition_convergence_test,
             initial_condition_weak_blast_wave
using SciMLBase: ODEProblem

const libtrixi_hip = get(ENV, "TRIXI_HIP_LIB", joinpath(@__DIR__, "..", "libtrixi_hip.so"))

# ---- C structs (field order and types of include/trixi_hip.h) ----
struct HipConfig
    ndims::Int32
    polydeg::Int32
    nelements::Int32
    ninterfaces::Int32
    gamma::Float64
    precision::Int32
    volume_integral::Int32
    volume_flux::Int32
    surface_flux::Int32
    fv_flux::Int32
    alpha_max::Float64
    alpha_min::Float64
    alpha_smooth::Int32
    source_terms::Int32
    elements_per_block::Int32
    threads_per_block::Int32
    nboundaries::Int32
    boundary_condition::Int32
    nmortars::Int32
    hg_variable::Int32
end

struct HipMesh
    neighbor_ids::Ptr{Int32}
    orientations::Ptr{Int32}
    inverse_jacobian::Ptr{Float64}
    element_centers::Ptr{Float64}
    boundary_neighbor_ids::Ptr{Int32}
    boundary_orientations::Ptr{Int32}
    boundary_neighbor_sides::Ptr{Int32}
    mortar_neighbor_ids::Ptr{Int32}
    mortar_large_sides::Ptr{Int32}
    mortar_orientations::Ptr{Int32}
end

struct HipBasis
    nodes::Ptr{Float64}
    weights::Ptr{Float64}
    derivative_split::Ptr{Float64}
    derivative_hat::Ptr{Float64}
    inverse_vandermonde_legendre::Ptr{Float64}
    analysis_nnodes::Int32
    analysis_weights::Ptr{Float64}
    analysis_vandermonde::Ptr{Float64}
    mortar_forward_upper::Ptr{Float64}
    mortar_forward_lower::Ptr{Float64}
    mortar_reverse_upper::Ptr{Float64}
    mortar_reverse_lower::Ptr{Float64}
end

const HOST = Int32(0)
const DEVICE = Int32(1)

# ---- numerical-flux and volume-integral codes ----
flux_code(f::FluxLaxFriedrichs) = f.dissipation.max_abs_speed === max_abs_speed_naive ? Int32(1) : Int32(0)
flux_code(::typeof(flux_ranocha)) = Int32(2)
flux_code(::typeof(flux_central)) = Int32(3)
flux_code(::typeof(flux_hllc)) = Int32(4)

vi_codes(::VolumeIntegralWeakForm) = (Int32(0), Int32(2), Int32(0), nothing)
vi_codes(v::VolumeIntegralFluxDifferencing) = (Int32(1), flux_code(v.volume_flux), Int32(0), nothing)
# The device implements the classic first-order HG blend only
# (VolumeIntegralShockCapturingHG, dg.jl:208-221): flux differencing everywhere
# outside the blended region (volume_integral_default, used for dg_only
# elements, calc_volume_integral.jl:250-255) and as the high-order part, and
# the first-order LGL finite volume as the low-order part. Other combinations
# (e.g. VolumeIntegralShockCapturingRRG, whose low-order part is
# VolumeIntegralPureLGLFiniteVolumeO2) are refused.
function vi_codes(v::VolumeIntegralShockCapturingHGType)
    hi, lo, default = v.volume_integral_blend_high_order, v.volume_integral_blend_low_order,
                      v.volume_integral_default
    hi isa VolumeIntegralFluxDifferencing && lo isa VolumeIntegralPureLGLFiniteVolume &&
        default isa VolumeIntegralFluxDifferencing && default.volume_flux === hi.volume_flux ||
        throw(ArgumentError("TrixiHIP supports VolumeIntegralShockCapturingHG (flux differencing blended " *
                            "with the first-order LGL finite volume) only"))
    (Int32(2), flux_code(hi.volume_flux), flux_code(lo.volume_flux_fv), v.indicator)
end
vi_codes(v::VolumeIntegralPureLGLFiniteVolume) = (Int32(3), Int32(2), flux_code(v.volume_flux_fv), nothing)

# TRIXI_HIP_VAR_* of the indicator variable
indicator_variable_code(::typeof(Trixi.density_pressure)) = Int32(0)
indicator_variable_code(::typeof(Trixi.density)) = Int32(1)
indicator_variable_code(::typeof(Trixi.pressure)) = Int32(2)

source_code(::Nothing) = Int32(0)
source_code(::typeof(source_terms_convergence_test)) = Int32(1)

# initial conditions the device can evaluate (TRIXI_HIP_IC_*), for Dirichlet boundary values
ic_code(::typeof(initial_condition_constant)) = Int32(0)
ic_code(::typeof(initial_condition_convergence_test)) = Int32(1)
ic_code(::typeof(initial_condition_weak_blast_wave)) = Int32(2)
ic_code(f) = throw(ArgumentError("TrixiHIP cannot evaluate $f on the device"))

# Every boundary must carry the same condition: BoundaryConditionDirichlet(semi.initial_condition)
# (the device evaluates the initial condition as the outer state, code 1) or
# boundary_condition_slip_wall (code 2). Returns the TRIXI_HIP_BC_* code.
function boundary_condition_code(semi)
    codes = Int32[]
    for bc in values(semi.boundary_conditions)
        bc isa Trixi.BoundaryConditionPeriodic && continue
        if bc === boundary_condition_slip_wall
            push!(codes, Int32(2))
        elseif bc isa BoundaryConditionDirichlet && bc.boundary_value_function === semi.initial_condition
            push!(codes, Int32(1))
        else
            throw(ArgumentError("TrixiHIP supports BoundaryConditionDirichlet(initial_condition) and " *
                                "boundary_condition_slip_wall only"))
        end
    end
    all(==(first(codes)), codes) ||
        throw(ArgumentError("TrixiHIP needs the same boundary condition on every boundary"))
    return first(codes)
end

check(rc, ctx) = rc == 0 || error("libtrixi_hip: ", unsafe_string(ccall((:trixi_hip_last_error, libtrixi_hip),
                                                                        Cstring, (Ptr{Cvoid},), ctx)))

# Row-major copy of a Julia (column-major) matrix, as the C ABI expects.
rowmajor(M) = vec(permutedims(Matrix{Float64}(M)))

"""
    HipRHS(semi; precision = :fp64)

Callable `f(du_ode, u_ode, semi, t)` evaluating `rhs_hyperbolic!` on the GPU.
Owns one library context (device copies of mesh and basis).
"""
mutable struct HipRHS
    ctx::Ptr{Cvoid}
end

function HipRHS(semi; precision::Symbol = :fp64)
    mesh, equations, solver, cache = mesh_equations_solver_cache(semi)
    mesh isa TreeMesh || throw(ArgumentError("TrixiHIP supports TreeMesh only"))
    D = ndims(mesh)
    E = nelements(solver, cache)
    I = ninterfaces(solver, cache)
    B = nboundaries(solver, cache)
    bc = B > 0 ? boundary_condition_code(semi) : Int32(0)
    M = nmortars(solver, cache)
    vi, vflux, fvflux, indicator = vi_codes(solver.volume_integral)
    amax, amin, asmooth = indicator === nothing ? (0.5, 0.001, true) :
                          (indicator.alpha_max, indicator.alpha_min, indicator.alpha_smooth)
    hgvar = indicator === nothing ? Int32(0) : indicator_variable_code(indicator.variable)
    cfg = HipConfig(D, Trixi.polydeg(solver), E, I, equations.gamma, precision === :fp32 ? 1 : 0,
                    vi, vflux, flux_code(solver.surface_integral.surface_flux), fvflux, amax, amin,
                    asmooth ? 1 : 0, source_code(semi.source_terms), 0, 0, B, bc, M, hgvar)
    # interface container: 1-based ids and orientations -> 0-based
    nbr = Int32.(cache.interfaces.neighbor_ids .- 1)            # [2, I]
    ori = Int32.(cache.interfaces.orientations .- 1)            # [I]
    invj = Vector{Float64}(cache.elements.inverse_jacobian)     # [E]
    centers = Matrix{Float64}(mesh.tree.coordinates[:, cache.elements.cell_ids])  # [D, E]
    # boundaries container (containers_2d.jl:475-543): 0-based ids/orientations, sides 1/2
    bel = Int32.(cache.boundaries.neighbor_ids[1:B] .- 1)
    bor = Int32.(cache.boundaries.orientations[1:B] .- 1)
    bsd = Int32.(cache.boundaries.neighbor_sides[1:B])
    # L2 mortars container (containers_3d.jl:652-788): ids [2^(ndims-1) + 1, M], 0-based
    mid = Int32.(cache.mortars.neighbor_ids[:, 1:M] .- 1)
    msd = Int32.(cache.mortars.large_sides[1:M])
    mor = Int32.(cache.mortars.orientations[1:M] .- 1)
    mortar = solver.mortar  # LobattoLegendreMortarL2
    mfu, mfl = rowmajor(mortar.forward_upper), rowmajor(mortar.forward_lower)
    mru, mrl = rowmajor(mortar.reverse_upper), rowmajor(mortar.reverse_lower)
    basis = solver.basis
    analyzer = SolutionAnalyzer(basis)
    nodes, weights = Vector{Float64}(basis.nodes), Vector{Float64}(basis.weights)
    ds, dh = rowmajor(basis.derivative_split), rowmajor(basis.derivative_hat)
    vinv = rowmajor(basis.inverse_vandermonde_legendre)
    aw, av = Vector{Float64}(analyzer.weights), rowmajor(analyzer.vandermonde)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve nbr ori invj centers bel bor bsd mid msd mor nodes weights ds dh vinv aw av mfu mfl mru mrl begin
        hm = Ref(HipMesh(pointer(nbr), pointer(ori), pointer(invj), pointer(centers),
                         pointer(bel), pointer(bor), pointer(bsd), pointer(mid), pointer(msd), pointer(mor)))
        hb = Ref(HipBasis(pointer(nodes), pointer(weights), pointer(ds), pointer(dh), pointer(vinv),
                          length(aw), pointer(aw), pointer(av), pointer(mfu), pointer(mfl), pointer(mru),
                          pointer(mrl)))
        rc = ccall((:trixi_hip_create, libtrixi_hip), Cint,
                   (Ref{HipConfig}, Ref{HipMesh}, Ref{HipBasis}, Ref{Ptr{Cvoid}}), Ref(cfg), hm, hb, out)
        rc == 0 || error("trixi_hip_create failed with code $rc")
    end
    if B > 0  # the boundary value function
        rc = ccall((:trixi_hip_set_exact_solution, libtrixi_hip), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Ptr{Float64}), out[], ic_code(semi.initial_condition),
                   C_NULL, 0, C_NULL)
        check(rc, out[])
    end
    f = HipRHS(out[])
    finalizer(f) do g
        ccall((:trixi_hip_destroy, libtrixi_hip), Cint, (Ptr{Cvoid},), g.ctx)
    end
    return f
end

# Host vectors (Vector{Float64}): the library stages them through device memory.
function (f::HipRHS)(du_ode::Vector{Float64}, u_ode::Vector{Float64}, semi, t)
    GC.@preserve du_ode u_ode begin
        check(ccall((:trixi_hip_rhs, libtrixi_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cint),
                    f.ctx, du_ode, u_ode, t, HOST), f.ctx)
    end
    return nothing
end

# Device vectors (e.g. AMDGPU.ROCVector{Float64}): pass the device pointers.
function (f::HipRHS)(du_ode::AbstractVector{Float64}, u_ode::AbstractVector{Float64}, semi, t)
    check(ccall((:trixi_hip_rhs, libtrixi_hip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cdouble, Cint),
                f.ctx, pointer(du_ode), pointer(u_ode), t, DEVICE), f.ctx)
    return nothing
end

"""
    semidiscretize_hip(semi, tspan; precision = :fp64)

Like `Trixi.semidiscretize(semi, tspan)` (semidiscretization.jl:102-153) but with
the RHS evaluated by libtrixi_hip. `p` stays `semi`, so callbacks are unchanged.
"""
function semidiscretize_hip(semi, tspan; precision::Symbol = :fp64)
    ode = Trixi.semidiscretize(semi, tspan)
    return ODEProblem{true}(HipRHS(semi; precision = precision), ode.u0, ode.tspan, semi)
end

# --- AMR with a device-resident state ---------------------------------------
# Trixi's AMRCallback keeps its tree logic (amr.jl:244-392, abstract_tree.jl);
# the per-node work runs on the device: IndicatorMax for the controller, and
# the solution transfer into a context created on the adapted cache.

"""
    hip_indicator_max(f::HipRHS, variable = 1)

`IndicatorMax(semi; variable)` (indicators_3d.jl:239-259) of the resident state,
for `variable` = the 1-based index of a conserved variable (`first` = 1).
"""
function hip_indicator_max(f::HipRHS, nelements::Integer, variable::Integer = 1)
    alpha = Vector{Float64}(undef, nelements)
    check(ccall((:trixi_hip_indicator_max, libtrixi_hip), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
                f.ctx, variable - 1, alpha), f.ctx)
    return alpha
end

"""
    transfer_map(n_old, elements_to_refine, elements_to_remove, ndims)

The (kind, src) arrays of `trixi_hip_transfer` for one adaptation, from the
element lists the solver's `refine!` and `coarsen!` receive (amr.jl:288-360):
`elements_to_refine` are old element ids, `elements_to_remove` are ids after the
refinement (the 2^ndims consecutive children of every coarsened parent).
"""
function transfer_map(n_old::Integer, elements_to_refine, elements_to_remove, ndims::Integer)
    nch = 2^ndims
    refine = falses(n_old)
    refine[elements_to_refine] .= true
    kind, src = Int32[], Int32[]
    for e in 1:n_old
        if refine[e]
            for q in 0:(nch - 1)
                push!(kind, Int32(1 + q)); push!(src, Int32(e - 1))
            end
        else
            push!(kind, Int32(0)); push!(src, Int32(e - 1))
        end
    end
    isempty(elements_to_remove) && return kind, src
    removed = falses(length(kind))
    removed[elements_to_remove] .= true
    kind2, src2 = Int32[], Int32[]
    e = 1
    while e <= length(kind)
        if removed[e]  # first child of a coarsened parent; its children are unrefined old elements
            push!(kind2, Int32(-1)); push!(src2, src[e])
            e += nch
        else
            push!(kind2, kind[e]); push!(src2, src[e])
            e += 1
        end
    end
    return kind2, src2
end

"""
    hip_transfer!(new::HipRHS, old::HipRHS, kind, src)

Project the resident state of `old` into `new` (created with `HipRHS(semi)` on
the adapted cache): refine!/coarsen! of the DG solver (amr_dg3d.jl:138-430).
"""
function hip_transfer!(new::HipRHS, old::HipRHS, kind::Vector{Int32}, src::Vector{Int32})
    GC.@preserve kind src begin
        check(ccall((:trixi_hip_transfer, libtrixi_hip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}),
                    new.ctx, old.ctx, kind, src), new.ctx)
    end
    return new
end

export HipRHS, semidiscretize_hip, hip_indicator_max, transfer_map, hip_transfer!

end # module
