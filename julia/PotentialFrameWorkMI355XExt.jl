# PotentialFrameWork.jl extension that runs the force-evaluation hot path on an
# MI355X through libpfw_mi355x (include/pfw_mi355x.h). A maintainer adds it as
# ext/PotentialFrameWorkMI355XExt.jl; it keeps the reference's calling
# conventions:
#   * evaluate!(w, sys, p; request, parameters) and the AtomsCalculators
#     methods with `neighbourlist = w` for the pair models
#     (src/submodules/PairPotentials/evaluation.jl:31-52, calculators.jl);
#   * AC.potential_energy / forces / virial / energy_forces_virial / calculate
#     for a device ACE calculator (src/submodules/ACE/calculators.jl:93-129);
#   * rrules for the pair path (ext/PotentialFrameWorkChainRulesCoreExt.jl:26-52)
#     and for AC.calculate (:197-236) through the device adjoints;
#   * the linear design (src/submodules/ACE/fitting.jl:18-130) and the
#     equivariant features with their pullback (src/submodules/ACE/features.jl:174-255).
# Units: positions Å, energies eV, forces eV/Å; positions 3×N column-major in
# original order, lattice vectors in columns. Errors map to ArgumentError,
# DimensionMismatch, DomainError and ErrorException as in the reference. One
# workspace (one HIP stream) serves one caller at a time, the reference's own
# rule (README.md:217-219).
#
# This file has not been run here: the build image has no Julia. The Python
# host layer (potentialframework.jl_amd/) drives the same ABI and is what the
# tests exercise; each function below names its Python counterpart.
module PotentialFrameWorkMI355XExt

using PotentialFrameWork
using AtomsBase
using Unitful
using StaticArrays: SVector, SMatrix
using SparseArrays: findnz
import AtomsCalculators as AC
import ChainRulesCore: rrule, RuleConfig, HasReverseMode, NoTangent, ZeroTangent, AbstractZero, unthunk,
                       @not_implemented
const PP = PotentialFrameWork.PairPotentials
const ACEm = PotentialFrameWork.ACE
const NIm = PotentialFrameWork.NeighbourInterface

const LIB = get(ENV, "PFW_MI355X_LIB", "libpfw_mi355x.so")
const PFW_ENERGY, PFW_FORCES, PFW_VIRIAL = Cuint(1), Cuint(2), Cuint(4)
const PFW_FP64, PFW_MIXED = Cint(0), Cint(1)
const PFW_SMOOTH, PFW_LJ = Cint(0), Cint(1)
const ENVELOPE_KIND = Dict(:PolyEnvelope2sX => Int32(0), :PolyEnvelope1sR => Int32(1),
                           :ACE1_PolyEnvelope1sR => Int32(2))

# ---------------------------------------------------------------- errors (capi.hip, _capi.py: check)
function check(ctx::Ptr{Cvoid}, rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:pfw_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx))
    rc == 1 && throw(ArgumentError(msg))
    rc == 2 && throw(DimensionMismatch(msg))
    rc == 3 && throw(DomainError(nothing, msg))
    throw(ErrorException(msg))                       # 4 stale, 5 capacity, 6 device
end

# ---------------------------------------------------------------- workspace (neighbourlist.py: Workspace)
"""MI355XWorkspace(; precision=:float64|:mixed, device=0, skin=0): the device
counterpart of `NeighbourList.Workspace` / `GPUWorkspace`
(src/submodules/NeighbourList/storage.jl:59-65, gpu.jl:61-116). One context is one
HIP stream. `skin > 0` is `GPUConfig(skin=...)` (gpu.jl:504-549)."""
mutable struct MI355XWorkspace
    ctx::Ptr{Cvoid}
    mode::Cint
    key::Any                                         # (n, H, pbc, channels, cutoff) of the uploaded system
    model_key::Any                                   # (kind, nz, parameters, cutoff) of the uploaded pair model
    function MI355XWorkspace(; precision=:float64, device=0, skin=0.0)
        precision in (:float64, :mixed) || throw(ArgumentError("precision must be :float64 or :mixed"))
        r = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:pfw_ctx_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), device, r)
        rc == 0 || throw(ErrorException("pfw_ctx_create failed ($rc)"))
        w = new(r[], precision === :mixed ? PFW_MIXED : PFW_FP64, nothing, nothing)
        finalizer(x -> ccall((:pfw_ctx_destroy, LIB), Cvoid, (Ptr{Cvoid},), x.ctx), w)
        check(w.ctx, ccall((:pfw_set_skin, LIB), Cint, (Ptr{Cvoid}, Cdouble), w.ctx, skin))
        w
    end
end

"""The `rebuilds` counter of a GPUWorkspace (gpu.jl:100,508)."""
function rebuilds(w::MI355XWorkspace)
    info = zeros(Int64, 5)
    check(w.ctx, ccall((:pfw_verlet_info, LIB), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Float64}), w.ctx, info, C_NULL))
    info[1]
end

"""Launch knobs of GPUConfig's strategy/workgroup/tile (gpu.jl:61-73) and the ACE
strategy (ACE/gpu.jl:301-303); 0 keeps the built-in choice: ACE sites per
wavefront, ACE radial tile, ACE product wavefronts, pair lanes per atom, ACE
contractions (1 VALU, 2 densities on matrix cores, 3 densities and pullback)
(pfw_set_tuning; tune.py: autotune sweeps them)."""
function set_tuning!(w::MI355XWorkspace; ace_sites_per_wave=0, ace_radial_chunk=0, ace_product_waves=0,
                     pair_lanes_per_atom=0, ace_density_contraction=0)
    k = Int32[ace_sites_per_wave, ace_radial_chunk, ace_product_waves, pair_lanes_per_atom, ace_density_contraction]
    check(w.ctx, ccall((:pfw_set_tuning, LIB), Cint, (Ptr{Cvoid}, Ptr{Int32}, Cint), w.ctx, k, 5))
end

"""Replay fresh steps from a captured HIP graph while only positions change (pfw_set_graphs)."""
set_graphs!(w::MI355XWorkspace, enabled::Bool) =
    check(w.ctx, ccall((:pfw_set_graphs, LIB), Cint, (Ptr{Cvoid}, Cint), w.ctx, enabled))

"""Positions already on the device (an integrator's ROCArray, 3×N Float64), as
`copyto!(executor.positions, x)` (ACE/gpu.jl:329-343, NeighbourList/gpu.jl:363-372)."""
update_positions!(w::MI355XWorkspace, X::DenseMatrix{Float64}) =
    check(w.ctx, ccall((:pfw_update_positions_device, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), w.ctx, pointer(X)))

request_mask(::PP.EnergyOnly) = PFW_ENERGY
request_mask(::PP.EnergyForcesVirial) = PFW_ENERGY | PFW_FORCES | PFW_VIRIAL
property_bit(p) = p isa AC.Energy ? PFW_ENERGY : p isa AC.Forces ? PFW_FORCES : p isa AC.Virial ? PFW_VIRIAL :
                  throw(ArgumentError("unsupported property $(typeof(p))"))
request_mask(t::Tuple) = reduce(|, (property_bit(p) for p in t); init=Cuint(0))
request_mask(p::Union{AC.Energy,AC.Forces,AC.Virial}) = property_bit(p)

"""Owning numeric extraction (storage.jl:135-170, ACE/environments.jl:21-27): X 3×N
in Å, H 3×3 with lattice vectors in columns (open axes zeroed), pbc flags."""
function geometry(sys::AbstractSystem{3})
    n = length(sys)
    X = Matrix{Float64}(undef, 3, n)
    for i in 1:n, a in 1:3
        X[a, i] = ustrip(u"Å", position(sys, i)[a])
    end
    pb = periodicity(sys)
    cv = cell_vectors(sys)
    H = [pb[b] ? ustrip(u"Å", cv[b][a]) : 0.0 for a in 1:3, b in 1:3]
    X, H, UInt8[p ? 1 : 0 for p in pb]
end

"""Upload a system (numeric form: X 3×N Å, H 3×3 columns, pbc flags, 0-based
channels), or only its positions when nothing else changed, which keeps a
Verlet list, ACE skin rows and a captured step alive (neighbourlist.py: prepare)."""
function set_system!(w::MI355XWorkspace, X::AbstractMatrix, H::AbstractMatrix, pbc::AbstractVector,
                     channels::Vector{Int32}, cutoff::Real)
    size(X, 1) == 3 || throw(DimensionMismatch("positions must be 3×N"))
    size(H) == (3, 3) || throw(DimensionMismatch("cell must be 3×3"))
    length(channels) == size(X, 2) || throw(DimensionMismatch("one channel per atom"))
    Xd = Matrix{Float64}(X); Hd = Matrix{Float64}(H); pb = UInt8[p != 0 ? 1 : 0 for p in pbc]
    key = (size(Xd, 2), Hd, pb, channels, Float64(cutoff))
    if key == w.key
        check(w.ctx, ccall((:pfw_update_positions, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}), w.ctx, Xd))
        return nothing
    end
    check(w.ctx, ccall((:pfw_set_system, LIB), Cint,
        (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Int32}, Ptr{Float64}, Ptr{UInt8}, Cdouble),
        w.ctx, size(Xd, 2), Xd, channels, Hd, pb, cutoff))
    w.key = key
    w.model_key = nothing                            # a new system takes the pair model again
    nothing
end
function set_system!(w::MI355XWorkspace, sys::AbstractSystem{3}, channels::Vector{Int32}, cutoff::Real)
    X, H, pb = geometry(sys)
    set_system!(w, X, H, pb, channels, cutoff)
end
# open lattice columns of a numeric cell are zero (ACE/environments.jl:21-27)
periodic_flags(H::AbstractMatrix) = UInt8[any(!iszero, view(H, :, b)) ? 1 : 0 for b in 1:3]

function read_outputs(w::MI355XWorkspace, n::Integer, entry::Symbol, mask::Cuint)
    E = Ref(0.0); F = zeros(3, n); W = zeros(3, 3)
    check(w.ctx, ccall((entry, LIB), Cint, (Ptr{Cvoid}, Cuint, Cint, Ref{Float64}, Ptr{Float64}, Ptr{Float64}),
        w.ctx, mask, w.mode, E, F, W))
    E[], F, W
end
force_vectors(F) = [SVector{3}(F[:, i]) * u"eV/Å" for i in axes(F, 2)]
static_virial(W) = SMatrix{3,3}(W) * u"eV"

# ---------------------------------------------------------------- pair models (pair.py)
# models.jl:138-176 (SmoothPair: amplitudes over its labels), lennard_jones.jl:5-22
device_model(p::PP.SmoothPair) = (PFW_SMOOTH, Cint(length(p.labels)), Matrix{Float64}(p.amplitudes))
device_model(p::PP.RadialPotential{<:PP.LennardJonesRadial}) =
    (PFW_LJ, Cint(1), Float64[p.radial.epsilon, p.radial.sigma])
device_model(p) = throw(ArgumentError("$(typeof(p)) has no device kernel"))
# bind_species (models.jl): channel of each atom in the model's labels
function pair_channels(sys::AbstractSystem{3}, p::PP.SmoothPair)
    ch = Vector{Int32}(undef, length(sys))
    for i in 1:length(sys)
        k = findfirst(isequal(species(sys, i)), p.labels)
        isnothing(k) && throw(ArgumentError("species $(species(sys, i)) is not in the pair model"))
        ch[i] = k - 1
    end
    ch
end
pair_channels(sys::AbstractSystem{3}, p) = zeros(Int32, length(sys))
pair_channels(labels::AbstractVector, p::PP.SmoothPair) =
    Int32[something(findfirst(isequal(s), p.labels), 0) - 1 for s in labels]
pair_channels(labels::AbstractVector, p) = zeros(Int32, length(labels))

function set_pair_model!(w::MI355XWorkspace, p)
    kind, nz, params = device_model(p)
    rc = Float64(PP.cutoff(p))
    key = (kind, nz, params, rc)
    key == w.model_key && return rc                  # unchanged: keeps a captured step alive
    check(w.ctx, ccall((:pfw_set_pair_model, LIB), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Cdouble),
        w.ctx, kind, nz, params, rc))
    w.model_key = key
    rc
end

"""`evaluate!(w, sys, p; request, parameters)` (evaluation.jl:31-52) on the device."""
function PP.evaluate!(w::MI355XWorkspace, sys::AbstractSystem{3}, p::PP.AbstractPairPotential;
                      request=PP.EnergyForcesVirial(), parameters=nothing)
    isnothing(parameters) || (p = PP.with_parameters(p, parameters))
    ch = pair_channels(sys, p)
    rc = Float64(PP.cutoff(p))
    set_system!(w, sys, ch, rc)
    set_pair_model!(w, p)
    E, F, W = read_outputs(w, length(sys), :pfw_pair_evaluate, request_mask(request))
    (energy=E * u"eV", forces=F * u"eV/Å", virial=W * u"eV")
end

# AtomsCalculators with neighbourlist = w (PairPotentials/calculators.jl)
function pair_result(sys, p, w; parameters=nothing)
    PP.evaluate!(w, sys, p; request=PP.EnergyForcesVirial(), parameters)
end
AC.potential_energy(sys::AbstractSystem{3}, p::PP.AbstractPairPotential; neighbourlist::MI355XWorkspace, kwargs...) =
    PP.evaluate!(neighbourlist, sys, p; request=PP.EnergyOnly()).energy
AC.forces(sys::AbstractSystem{3}, p::PP.AbstractPairPotential; neighbourlist::MI355XWorkspace, kwargs...) =
    force_vectors(ustrip.(u"eV/Å", pair_result(sys, p, neighbourlist).forces))
AC.virial(sys::AbstractSystem{3}, p::PP.AbstractPairPotential; neighbourlist::MI355XWorkspace, kwargs...) =
    static_virial(ustrip.(u"eV", pair_result(sys, p, neighbourlist).virial))
function AC.energy_forces_virial(sys::AbstractSystem{3}, p::PP.AbstractPairPotential;
                                 neighbourlist::MI355XWorkspace, kwargs...)
    r = pair_result(sys, p, neighbourlist)
    (energy=r.energy, forces=force_vectors(ustrip.(u"eV/Å", r.forces)), virial=static_virial(ustrip.(u"eV", r.virial)))
end

"""`pair_evaluate(X, p, ps, env, H)` (PairPotentials/numerical.jl:60-75) with the
workspace first: the device builds the pair set itself (bit-identical to the
reference's directed snapshot; env supplies the species labels and the cutoff)."""
function PP.pair_evaluate(w::MI355XWorkspace, X::AbstractMatrix, p::PP.AbstractPairPotential, ps::NamedTuple,
                          env::PP.PairEnvironment, H::AbstractMatrix; request=PP.EnergyForcesVirial(),
                          state=NamedTuple())
    q = PP.with_parameters(p, ps)
    Float64(PP.cutoff(q)) == Float64(env.radius) || throw(ArgumentError("environment cutoff mismatch"))
    set_system!(w, X, H, periodic_flags(H), pair_channels(env.labels, q), env.radius)
    set_pair_model!(w, q)
    E, F, W = read_outputs(w, size(X, 2), :pfw_pair_evaluate, request_mask(request))
    (energy=E, forces=F, virial=W)
end

"""The pair rrule (ext/PotentialFrameWorkChainRulesCoreExt.jl:26-52) through
pfw_pair_adjoint: SmoothPair amplitude gradients (the upper triangle in column
order, parameters.jl:34-37); position gradients are not provided by the device."""
function rrule(::typeof(PP.pair_evaluate), w::MI355XWorkspace, X::AbstractMatrix, p::PP.SmoothPair,
               ps::NamedTuple, env::PP.PairEnvironment, H::AbstractMatrix; request=PP.EnergyForcesVirial(),
               state=NamedTuple())
    y = PP.pair_evaluate(w, X, p, ps, env, H; request, state)
    saved = deepcopy((X, p, ps, env, H))
    function pair_pullback(delta)
        delta = unthunk(delta)
        delta isa AbstractZero && return (NoTangent(), NoTangent(), ZeroTangent(), NoTangent(), ZeroTangent(),
                                          NoTangent(), NoTangent())
        cot(name, zero_value) = hasproperty(delta, name) ? (v = unthunk(getproperty(delta, name));
                                                             v isa AbstractZero ? zero_value : v) : zero_value
        X0, p0, ps0, env0, H0 = saved
        PP.pair_evaluate(w, X0, p0, ps0, env0, H0)          # the saved system and parameters on the device again
        dE = Float64[cot(:energy, 0.0)]
        dF = Matrix{Float64}(cot(:forces, zeros(size(X0))))
        dW = Matrix{Float64}(cot(:virial, zeros(3, 3)))
        nz = length(p.labels)
        gA = zeros(nz, nz)
        check(w.ctx, ccall((:pfw_pair_adjoint, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), w.ctx, dE, dF, dW, gA))
        g = gA + transpose(gA) - diagonal_part(gA)
        (NoTangent(), NoTangent(), @not_implemented("position gradients of the device pair path"), NoTangent(),
         (amplitudes=[g[i, j] for j in 1:nz for i in 1:j],), NoTangent(),
         @not_implemented("cell derivatives are outside the fixed-cell pair AD contract"))
    end
    y, pair_pullback
end
diagonal_part(A) = [i == j ? A[i, j] : 0.0 for i in axes(A, 1), j in axes(A, 2)]

# ---------------------------------------------------------------- ACE (ace.py: ResolvedModel._struct)
struct PfwRadial                                    # pfw_ace_radial
    nr::Int32
    nnodes::Int32
    transform::Ptr{Float64}
    envelope_kind::Ptr{Int32}
    envelope::Ptr{Float64}
    nodes::Ptr{Float64}
    coefficients::Ptr{Float64}
end
struct PfwModel                                     # pfw_ace_model
    nz::Int32
    lmax::Int32
    ndens::Int32
    nprod::Int32
    nbasis::Int32
    reference::Ptr{Float64}
    radial::PfwRadial
    pair::PfwRadial
    density::Ptr{Int32}
    product_ptr::Ptr{Int32}
    product_idx::Ptr{Int32}
    coupling_nnz::Int64
    coupling_rows::Ptr{Int32}
    coupling_cols::Ptr{Int32}
    coupling_vals::Ptr{Float64}
    WB::Ptr{Float64}
    Wpair::Ptr{Float64}
    angular::Int32
    selected::Ptr{Int32}
    order_lo::Int32
    order_hi::Int32
end

"""Pack one radial family (a centre × neighbour Matrix of RadialChannel,
resolved.jl:1-17) into the ABI arrays, channel index centre + nz·neighbour."""
function pack_radial(family::AbstractMatrix, keep::Vector{Any})
    chans = vec(family)                              # centre fastest (column-major)
    isempty(chans) && return PfwRadial(0, 2, C_NULL, C_NULL, C_NULL, C_NULL, C_NULL)
    nr = size(first(chans).coefficients, 2); nnodes = length(first(chans).nodes)
    all(c -> size(c.coefficients, 2) == nr && length(c.nodes) == nnodes, chans) ||
        throw(DimensionMismatch("radial channels must share the radial count and the spline nodes' count"))
    tf = Float64[]; ek = Int32[]; ep = Float64[]; nodes = Float64[]; coef = Float64[]
    for c in chans
        t = c.transform
        append!(tf, (t.p, t.q, t.a, t.rin, t.r0, t.yin, t.ycut, t.rcut))
        haskey(ENVELOPE_KIND, c.envelope) || throw(ArgumentError("unsupported radial envelope $(c.envelope)"))
        push!(ek, ENVELOPE_KIND[c.envelope])
        e = zeros(5); e[1:length(c.envelope_parameters)] .= c.envelope_parameters
        append!(ep, e); append!(nodes, c.nodes); append!(coef, vec(c.coefficients))
    end
    push!(keep, tf, ek, ep, nodes, coef)
    PfwRadial(nr, nnodes, pointer(tf), pointer(ek), pointer(ep), pointer(nodes), pointer(coef))
end
pack_radial(family, keep) = throw(ArgumentError("device ACE needs piecewise-cubic Agnesi radials (RadialChannel)"))

"""Pack a ResolvedModel (resolved.jl:18-38) with its basis policy
(ACE/interface.jl:2-18, 115-125) and readout into the ABI struct: 0-based density,
product and coupling indices; effective_WB (mask applied) and Wpair (pair flag
applied), both nbasis|nq × nz column-major; reference_value per species;
selected species flags and the restricted orders. `keep` holds the arrays."""
function pack_model(m, ps)
    keep = Any[]
    nz = length(m.atomic_numbers)
    pol = m.policy
    ref = Float64[ACEm.reference_value(m, z) for z in 1:nz]
    dens = Int32[]
    for (r, h) in m.density
        push!(dens, r - 1, h - 1)
    end
    pptr = Int32[0]; pidx = Int32[]
    for p in m.products
        append!(pidx, Int32.(p .- 1)); push!(pptr, length(pidx))
    end
    rows, cols, vals = findnz(m.coupling)
    crow = Int32.(rows .- 1); ccol = Int32.(cols .- 1); cval = Float64.(vals)
    WB = Matrix{Float64}(ACEm.effective_WB(m, ps))
    Wp = Matrix{Float64}(ps.Wpair) .* (pol.pair ? 1.0 : 0.0)
    sel = zeros(Int32, nz); for z in pol.species; sel[z] = 1; end
    lo, hi = isnothing(pol.orders) ? (-1, -1) : (first(pol.orders), last(pol.orders))
    push!(keep, ref, dens, pptr, pidx, crow, ccol, cval, WB, Wp, sel)
    st = PfwModel(nz, m.lmax, length(m.density), length(m.products), size(m.coupling, 1), pointer(ref),
                  pack_radial(m.radial, keep), pack_radial(m.pair_radial, keep), pointer(dens), pointer(pptr),
                  pointer(pidx), length(cval), pointer(crow), pointer(ccol), pointer(cval), pointer(WB), pointer(Wp),
                  m.angular_kind === :real_solid ? Int32(1) : Int32(0), pointer(sel), lo, hi)
    st, keep
end

"""Device ACE calculator (ace.py: ACECalculator; calculators.jl:16-29)."""
mutable struct MI355XACECalculator
    w::MI355XWorkspace
    model::Any
    parameters::Any
end
function MI355XACECalculator(model, ps; precision=:float64, device=0, skin=0.0)
    w = MI355XWorkspace(; precision, device, skin)
    st, keep = pack_model(model, ps)
    GC.@preserve keep check(w.ctx, ccall((:pfw_set_ace_model, LIB), Cint, (Ptr{Cvoid}, Ref{PfwModel}), w.ctx, st))
    MI355XACECalculator(w, model, deepcopy(ps))
end
AC.get_parameters(c::MI355XACECalculator) = deepcopy(c.parameters)
function AC.set_parameters!(c::MI355XACECalculator, ps)
    size(ps.WB) == size(c.parameters.WB) && size(ps.Wpair) == size(c.parameters.Wpair) ||
        throw(DimensionMismatch("ACE readout shapes"))
    WB = Matrix{Float64}(ACEm.effective_WB(c.model, ps))
    Wp = Matrix{Float64}(ps.Wpair) .* (c.model.policy.pair ? 1.0 : 0.0)
    check(c.w.ctx, ccall((:pfw_set_ace_parameters, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
        c.w.ctx, WB, Wp))
    c.parameters = deepcopy(ps)
    c
end
AC.energy_unit(::MI355XACECalculator) = u"eV"
AC.length_unit(::MI355XACECalculator) = u"Å"

# element channels of the model (ACE/environments.jl:1-10), 0-based
ace_channels(sys::AbstractSystem{3}, m) = Int32[ACEm.element_channel(m, species(sys, i)) - 1 for i in 1:length(sys)]

function ace_result(sys::AbstractSystem{3}, c::MI355XACECalculator; parameters=nothing, mask=PFW_ENERGY | PFW_FORCES | PFW_VIRIAL)
    isnothing(parameters) || parameters === c.parameters || AC.set_parameters!(c, parameters)
    set_system!(c.w, sys, ace_channels(sys, c.model), NIm.search_cutoff(c.model))
    E, F, W = read_outputs(c.w, length(sys), :pfw_ace_evaluate, mask)
    (energy=E * u"eV", forces=force_vectors(F), virial=static_virial(W))
end
# calculators.jl:93-129
AC.potential_energy(sys::AbstractSystem{3}, c::MI355XACECalculator; kwargs...) =
    ace_result(sys, c; mask=PFW_ENERGY).energy
AC.forces(sys::AbstractSystem{3}, c::MI355XACECalculator; kwargs...) = ace_result(sys, c).forces
AC.virial(sys::AbstractSystem{3}, c::MI355XACECalculator; kwargs...) = ace_result(sys, c).virial
AC.energy_forces_virial(sys::AbstractSystem{3}, c::MI355XACECalculator; kwargs...) = ace_result(sys, c)
function AC.energy_forces(sys::AbstractSystem{3}, c::MI355XACECalculator; kwargs...)
    r = ace_result(sys, c); (energy=r.energy, forces=r.forces)
end
function calculate_device(props, sys, c::MI355XACECalculator, parameters, state)
    ps = props isa Tuple ? props : (props,)
    isempty(ps) && return (; state)
    r = ace_result(sys, c; parameters, mask=request_mask(ps))
    key(::AC.Energy) = :energy; key(::AC.Forces) = :forces; key(::AC.Virial) = :virial
    merge((NamedTuple{(key(p),)}((r[key(p)],)) for p in ps)..., (; state))
end
for PropertyType in (Union{AC.Energy,AC.Forces,AC.Virial}, Tuple)
    @eval begin
        AC.calculate(props::$PropertyType, sys::AbstractSystem{3}, c::MI355XACECalculator; kwargs...) =
            calculate_device(props, sys, c, nothing, nothing)
        AC.calculate(props::$PropertyType, sys::AbstractSystem{3}, c::MI355XACECalculator, parameters;
                     state=nothing, kwargs...) = calculate_device(props, sys, c, parameters, state)
        AC.calculate(props::$PropertyType, sys::AbstractSystem{3}, c::MI355XACECalculator, parameters, state;
                     kwargs...) = calculate_device(props, sys, c, parameters, state)
    end
end

strip_bar(x::AbstractZero, u, shape) = zeros(shape)
strip_bar(x::Number, u, shape) = fill(Float64(ustrip(u, x)), shape)
strip_bar(x::AbstractArray, u, shape) = reshape(Float64[ustrip(u, v) for v in Iterators.flatten(x)], shape)

"""AC.calculate's parameter rrule (ext/PotentialFrameWorkChainRulesCoreExt.jl:197-236)
through pfw_ace_adjoint: the gradient with respect to (WB, Wpair); the pullback
recomputes the environments of the saved system on the device."""
function rrule(config::RuleConfig{>:HasReverseMode}, ::typeof(AC.calculate), props, sys::AbstractSystem{3},
               c::MI355XACECalculator, ps, state; kwargs...)
    y = AC.calculate(props, sys, c, ps, state)
    saved = deepcopy(sys)
    n = length(sys)
    function pullback(d)
        d = unthunk(d)
        d isa AbstractZero && return (NoTangent(), NoTangent(), NoTangent(), NoTangent(), ZeroTangent(), NoTangent())
        field(k) = hasproperty(d, k) ? unthunk(getproperty(d, k)) : ZeroTangent()
        dE = strip_bar(field(:energy), u"eV", (1,))
        dF = strip_bar(field(:forces), u"eV/Å", (3, n))
        dW = strip_bar(field(:virial), u"eV", (3, 3))
        AC.set_parameters!(c, ps)
        set_system!(c.w, saved, ace_channels(saved, c.model), NIm.search_cutoff(c.model))
        gWB = zeros(size(ps.WB)); gWp = zeros(size(ps.Wpair))
        check(c.w.ctx, ccall((:pfw_ace_adjoint, LIB), Cint,
            (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
            c.w.ctx, c.w.mode, dE, dF, dW, gWB, gWp))
        (NoTangent(), NoTangent(), NoTangent(), NoTangent(), (WB=gWB, Wpair=gWp), NoTangent())
    end
    y, pullback
end
function rrule(config::RuleConfig{>:HasReverseMode}, ::typeof(AC.calculate), props, sys::AbstractSystem{3},
               c::MI355XACECalculator, ps; state=nothing, kwargs...)
    y, back = rrule(config, AC.calculate, props, sys, c, ps, state)
    y, d -> back(d)[1:5]
end

# ---------------------------------------------------------------- linear design (fitting.py: design_matrix)
"""`design_matrix(X, model, ps, env, H; fit_pair)` (ACE/fitting.jl:18-130) for one
configuration on the device (fitting.py: design_matrix): rows [E; vec(F); vec(W)],
columns vec(WB) then vec(Wpair); `matrix` excludes the baseline (the zero
readout's outputs: the reference energies and, without fit_pair, the current pair
term). For an unprojected basis the vec(WB) columns come from one pass of
per-site Jacobians (pfw_ace_design); the Wpair columns and the baseline, and
every column of a projected basis, from unit-readout parameter sets."""
function design_matrix(X::AbstractMatrix, model, ps, env, H::AbstractMatrix; fit_pair=true,
                       c::MI355XACECalculator)
    set_system!(c.w, X, H, periodic_flags(H), Int32.(env.species .- 1), NIm.search_cutoff(model))
    nb, nz = size(ps.WB); nq = size(ps.Wpair, 1)
    n = length(env.species); rows = 1 + 3n + 9
    mask = Float64.(collect(Bool, model.policy.mask))
    jac = isnothing(model.policy.orders)
    nsb = jac ? 0 : nb * nz                          # basis columns taken from parameter sets
    npair = fit_pair ? nq * nz : 0
    nsets = nsb + npair + 1
    WB = zeros(nb, nz, nsets); WP = zeros(nq, nz, nsets)
    for k in 1:nsb                                   # vec(WB) columns, masked as effective_WB
        WB[mod1(k, nb), cld(k, nb), k] = mask[mod1(k, nb)]
    end
    if fit_pair && model.policy.pair
        for k in 1:npair
            WP[mod1(k, nq), cld(k, nq), nsb + k] = 1.0
        end
    elseif !fit_pair
        # the current pair term (off when the policy disables it), cancelling in the columns
        WP .= reshape(Matrix{Float64}(ps.Wpair) .* (model.policy.pair ? 1.0 : 0.0), nq, nz, 1)
    end
    out = Matrix{Float64}(undef, rows, nsets)
    check(c.w.ctx, ccall((:pfw_ace_parameter_sets, LIB), Cint,
        (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        c.w.ctx, c.w.mode, nsets, WB, WP, out))
    baseline = out[:, end]
    A = Matrix{Float64}(undef, rows, nb * nz + npair)
    if jac
        At = Matrix{Float64}(undef, nb * nz, rows)     # the row-major block of pfw_ace_design, transposed view
        check(c.w.ctx, ccall((:pfw_ace_design, LIB), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Int64),
            c.w.ctx, c.w.mode, mask, At, nb * nz))
        A[:, 1:nb*nz] .= transpose(At)
    else
        A[:, 1:nb*nz] .= out[:, 1:nb*nz] .- baseline
    end
    npair > 0 && (A[:, nb*nz+1:end] .= out[:, nsb+1:nsb+npair] .- baseline)
    (; matrix=A, baseline)
end

# ---------------------------------------------------------------- features (features.py)
struct PfwFeatureLayer                              # pfw_feature_layer
    nz::Int32
    nr::Int32
    lmax::Int32
    ndens::Int32
    nprod::Int32
    nblocks::Int32
    dens_z::Ptr{Int32}
    dens_n::Ptr{Int32}
    dens_h::Ptr{Int32}
    prod_ptr::Ptr{Int32}
    prod_idx::Ptr{Int32}
    block_L::Ptr{Int32}
    block_channels::Ptr{Int32}
    block_nbasis::Ptr{Int32}
    row_ptr::Ptr{Int32}
    col::Ptr{Int32}
    val::Ptr{Float64}
    mixing::Ptr{Float64}
    cutoff::Float64
    radial_cutoff::Float64
    radial_kind::Int32
    powers::Ptr{Int32}
    rate::Float64
    amplitude::Float64
    scale::Ptr{Float64}
end

"""Flatten an ACEFeatureLayer (features.jl:9-19) and its parameters
(features.jl:106-116) for the device: densities (neighbour channel 0 = all, then
0-based radial and harmonic), products, the coupling rows of every block in
turn as CSR, mixing blocks column-major in turn, the smooth-exponential radial's
powers, cutoff, rate, amplitude and scale (features.py: _tables, _struct)."""
function pack_feature_layer(l, ps)
    rad = first(l.radial)
    rad.evaluator isa ACEm.SmoothExponential ||
        throw(ArgumentError("the device feature kernels take the SmoothExponential radial family"))
    keep = Any[]
    dz = Int32[d[1] for d in l.density]; dn = Int32[d[2] - 1 for d in l.density]; dh = Int32[d[3] - 1 for d in l.density]
    pptr = Int32[0]; pidx = Int32[]
    for p in l.products
        append!(pidx, Int32.(p .- 1)); push!(pptr, length(pidx))
    end
    rowp = Int32[0]; col = Int32[]; val = Float64[]
    for C in l.coupling, r in axes(C, 1)
        for k in axes(C, 2)
            v = C[r, k]
            iszero(v) && continue
            push!(col, k - 1); push!(val, v)
        end
        push!(rowp, length(col))
    end
    L = Int32[s.L for s in l.irreps]; ch = Int32[s.channels for s in l.irreps]
    nb = Int32[size(l.coupling[b], 1) ÷ (2l.irreps[b].L + 1) for b in eachindex(l.irreps)]
    mix = reduce(vcat, [vec(Float64.(getproperty(ps.mixing, Symbol(:block_, b)))) for b in eachindex(l.irreps)])
    powers = Int32.(rad.config.powers)
    scale = vec(Float64.(ps.radial.scale))
    push!(keep, dz, dn, dh, pptr, pidx, rowp, col, val, L, ch, nb, mix, powers, scale)
    st = PfwFeatureLayer(length(l.atomic_numbers), length(l.radial_spec), l.lmax, length(l.density),
                         length(l.products), length(l.irreps), pointer(dz), pointer(dn), pointer(dh), pointer(pptr),
                         pointer(pidx), pointer(L), pointer(ch), pointer(nb), pointer(rowp), pointer(col), pointer(val),
                         pointer(mix), l.cutoff, rad.config.cutoff, Int32(1), pointer(powers),
                         Float64(ps.custom_radial.rate[1]), Float64(ps.custom_radial.amplitude[1]), pointer(scale))
    st, keep
end

"""Split the device's flat parameter gradients into the structure of `ps`."""
function unpack_feature_parameters(l, ps, dmix, dscale, drad)
    off = 0
    blocks = map(eachindex(l.irreps)) do b
        W = getproperty(ps.mixing, Symbol(:block_, b))
        g = reshape(dmix[off+1:off+length(W)], size(W)); off += length(W)
        g
    end
    (mixing=NamedTuple{Tuple(Symbol(:block_, b) for b in eachindex(l.irreps))}(Tuple(blocks)),
     radial=(scale=reshape(dscale, size(ps.radial.scale)),),
     custom_radial=(rate=[drad[1]], amplitude=[drad[2]]))
end

"""`ace_features(X, l, ps, env, H)` (features.jl:174-200) on the device."""
function ace_features(X::AbstractMatrix, l, ps, env, H::AbstractMatrix; w::MI355XWorkspace)
    set_system!(w, X, H, periodic_flags(H), Int32.(env.species .- 1), l.cutoff)
    st, keep = pack_feature_layer(l, ps)
    n = size(X, 2)
    sizes = [(2s.L + 1) * s.channels * n for s in l.irreps]
    out = zeros(sum(sizes))
    GC.@preserve keep check(w.ctx, ccall((:pfw_features, LIB), Cint,
        (Ptr{Cvoid}, Ref{PfwFeatureLayer}, Ptr{Float64}), w.ctx, st, out))
    offs = cumsum([0; sizes])
    Tuple(reshape(out[offs[b]+1:offs[b+1]], 2s.L + 1, s.channels, n) for (b, s) in enumerate(l.irreps))
end

"""`feature_vjp(X, l, ps, env, H, cotangents)` (features.jl:206-255) on the device:
the position gradient (3×N) and the parameter gradient in the structure of ps."""
function feature_vjp(X::AbstractMatrix, l, ps, env, H::AbstractMatrix, cotangents; w::MI355XWorkspace)
    set_system!(w, X, H, periodic_flags(H), Int32.(env.species .- 1), l.cutoff)
    st, keep = pack_feature_layer(l, ps)
    cot = reduce(vcat, [vec(Float64.(c)) for c in cotangents])     # blocks in turn, (2L+1, channels, n) each
    dx = zeros(3, size(X, 2))
    dmix = zeros(sum(length(getproperty(ps.mixing, Symbol(:block_, b))) for b in eachindex(l.irreps)))
    dscale = zeros(length(ps.radial.scale)); drad = zeros(2)
    GC.@preserve keep check(w.ctx, ccall((:pfw_features_vjp, LIB), Cint,
        (Ptr{Cvoid}, Ref{PfwFeatureLayer}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        w.ctx, st, cot, dx, dmix, dscale, drad))
    (positions=dx, parameters=unpack_feature_parameters(l, ps, dmix, dscale, drad))
end

end # module
