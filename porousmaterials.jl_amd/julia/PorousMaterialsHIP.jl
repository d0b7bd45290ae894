"""
    PorousMaterialsHIP

`ccall` binding of libpmhip.so (include/pmhip.h) with the call surface of
PorousMaterials.jl v0.4.2: `henry_coefficient`, `energy_grid`, `vdw_energy`,
`electrostatic_potential_energy` and `total_electrostatic_potential_energy`
take and return the reference's own types (`Crystal`, `Molecule`,
`LJForceField`, `EwaldParams`, `Eikr`, `EwaldSum`, `Grid`, the 16-key results
`Dict`).  Device handles are cached per object and released by finalizers.

Usage: `using PorousMaterials, PorousMaterialsHIP`, then call
`PorousMaterialsHIP.henry_coefficient(...)` exactly as the reference's, or
`PorousMaterialsHIP.enable!()` to route the reference's own methods through
the device.
"""
module PorousMaterialsHIP

using PorousMaterials
using ChainRulesCore
using LinearAlgebra
using Optim
using Printf
using Statistics

const libpmhip = normpath(joinpath(@__DIR__, "..", "libpmhip.so"))

const PM_FP64 = Int32(0)
const PM_MIXED = Int32(1)

struct PMError <: Exception
    code::Int
    msg::String
end
Base.showerror(io::IO, e::PMError) = print(io, "libpmhip error ", e.code, ": ", e.msg)

function check(rc::Integer)
    if rc != 0
        msg = unsafe_string(ccall((:pm_last_error, libpmhip), Cstring, ()))
        throw(PMError(rc, msg))
    end
    return nothing
end

precision_code(p::Symbol) = p === :fp64 ? PM_FP64 : p === :mixed ? PM_MIXED :
    error("precision must be :mixed or :fp64")

function __init__()
    check(ccall((:pm_init, libpmhip), Cint, (Int32,), Int32(0)))
end

# ------------------------------------------------------------ handles ----
mutable struct Handle
    ptr::Ptr{Cvoid}
    destroy::Symbol
    function Handle(ptr::Ptr{Cvoid}, destroy::Symbol)
        h = new(ptr, destroy)
        finalizer(h) do x
            if x.ptr != C_NULL
                if x.destroy === :pm_framework_destroy
                    ccall((:pm_framework_destroy, libpmhip), Cint, (Ptr{Cvoid},), x.ptr)
                elseif x.destroy === :pm_forcefield_destroy
                    ccall((:pm_forcefield_destroy, libpmhip), Cint, (Ptr{Cvoid},), x.ptr)
                elseif x.destroy === :pm_ewald_destroy
                    ccall((:pm_ewald_destroy, libpmhip), Cint, (Ptr{Cvoid},), x.ptr)
                elseif x.destroy === :pm_access_destroy
                    ccall((:pm_access_destroy, libpmhip), Cint, (Ptr{Cvoid},), x.ptr)
                elseif x.destroy === :pm_gcmc_destroy
                    ccall((:pm_gcmc_destroy, libpmhip), Cint, (Ptr{Cvoid},), x.ptr)
                else
                    ccall((:pm_probe_destroy, libpmhip), Cint, (Ptr{Cvoid},), x.ptr)
                end
                x.ptr = C_NULL
            end
        end
        return h
    end
end

# species order of the dense tables: sorted symbols of the force field, plus a
# zero-epsilon slot for framework charges that sit on no LJ atom
ff_species(ljff::LJForceField) = sort(collect(keys(ljff.pure_σ)))

function ff_tables(ljff::LJForceField)
    sp = ff_species(ljff)
    n = length(sp) + 1
    eps = zeros(n, n)
    sig2 = ones(n, n)
    for (i, a) in enumerate(sp), (j, b) in enumerate(sp)
        eps[i, j] = ljff.ϵ[a][b]      # row-major [i*n + j] == column-major transpose;
        sig2[i, j] = ljff.σ²[a][b]    # the tables are symmetric
    end
    return sp, eps, sig2
end

# ------------------------------------------------------------ caches ----
# Device handles of reference objects are cached under a fingerprint of the
# object's contents.  `LJForceField` (src/forcefield.jl:12) and Xtals'
# `Crystal` are immutable structs: Julia refuses finalizers on them, so they
# can key neither a WeakKeyDict nor a finalizer.  A content fingerprint also
# changes when a caller edits a crystal's arrays in place, so a stale device
# copy is never used.  Each cache keeps at most CACHE_LIMIT handles (it is
# emptied when full); `release_handles!()` empties both.  A handle removed from
# a cache is freed by its finalizer once no call holds it.
const CACHE_LIMIT = 64
const FF_CACHE = Dict{UInt64, Handle}()
const FW_CACHE = Dict{UInt64, Handle}()

fingerprint(ljff::LJForceField) =
    hash((ljff.name, ljff.pure_σ, ljff.pure_ϵ, ljff.σ², ljff.ϵ, ljff.r²_cutoff))

fingerprint(crystal::Crystal) =
    hash((crystal.box.f_to_c, crystal.atoms.species, crystal.atoms.coords.xf, crystal.charges.q,
          crystal.charges.coords.xf))

function cached!(make::Function, cache::Dict{UInt64, Handle}, key::UInt64)
    h = get(cache, key, nothing)
    h === nothing || return h
    length(cache) >= CACHE_LIMIT && empty!(cache)
    h = make()
    cache[key] = h
    return h
end

"""
    release_handles!()

Drop every cached device handle (force fields, frameworks, Ewald sums); their
device memory is freed by the handles' finalizers once no call holds them.
"""
function release_handles!()
    empty!(FF_CACHE)
    empty!(FW_CACHE)
    empty!(EW_CACHE)
    return nothing
end

function forcefield_handle(ljff::LJForceField)
    cached!(FF_CACHE, fingerprint(ljff)) do
        sp, eps, sig2 = ff_tables(ljff)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:pm_forcefield_create, libpmhip), Cint,
                    (Int32, Ptr{Float64}, Ptr{Float64}, Float64, Ref{Ptr{Cvoid}}),
                    Int32(size(eps, 1)), eps, sig2, ljff.r²_cutoff, out))
        Handle(out[], :pm_forcefield_destroy)
    end
end

# Framework atoms for the device: LJ atoms carrying the charge at their
# position, plus charge-only sites on the zero-epsilon species.
function framework_arrays(crystal::Crystal, ljff::LJForceField)
    sp = ff_species(ljff)
    idx = Dict(s => Int32(i - 1) for (i, s) in enumerate(sp))
    ghost = Int32(length(sp))
    n = crystal.atoms.n
    xf = copy(crystal.atoms.coords.xf)
    species = Int32[idx[s] for s in crystal.atoms.species]
    q = zeros(n)
    extra_x = Vector{Vector{Float64}}()
    extra_q = Float64[]
    for c in 1:crystal.charges.n
        x = crystal.charges.coords.xf[:, c]
        k = findfirst(i -> xf[:, i] == x, 1:n)
        if k === nothing
            push!(extra_x, x)
            push!(extra_q, crystal.charges.q[c])
        else
            q[k] += crystal.charges.q[c]
        end
    end
    if !isempty(extra_q)
        xf = hcat(xf, reduce(hcat, extra_x))
        species = vcat(species, fill(ghost, length(extra_q)))
        q = vcat(q, extra_q)
    end
    return xf, species, q
end

# species-only force field (every epsilon 0) for Ewald-only device frameworks
function es_only_forcefield(crystal::Crystal, r_cutoff::Float64)
    species = unique(crystal.atoms.species)
    pure_σ = Dict{Symbol, Float64}(s => 1.0 for s in species)
    pure_ϵ = Dict{Symbol, Float64}(s => 0.0 for s in species)
    σ² = Dict{Symbol, Dict{Symbol, Float64}}(a => Dict{Symbol, Float64}(b => 1.0 for b in species)
                                             for a in species)
    ϵ = Dict{Symbol, Dict{Symbol, Float64}}(a => Dict{Symbol, Float64}(b => 0.0 for b in species)
                                            for a in species)
    return LJForceField("electrostatics-only", pure_σ, pure_ϵ, σ², ϵ, r_cutoff^2)
end

# `crystal` is the unit cell; `rep` the replication factors the reference uses
function framework_handle(crystal::Crystal, ljff::LJForceField, rep::Tuple{Int, Int, Int})
    cached!(FW_CACHE, hash((fingerprint(crystal), fingerprint(ljff), rep))) do
        xf, species, q = framework_arrays(crystal, ljff)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:pm_framework_create, libpmhip), Cint,
                    (Ptr{Float64}, Ptr{Int32}, Int64, Ptr{Float64}, Ptr{Int32}, Ptr{Float64},
                     Ref{Ptr{Cvoid}}),
                    Matrix{Float64}(crystal.box.f_to_c), Int32[rep...], Int64(size(xf, 2)), xf,
                    species, q, out))
        Handle(out[], :pm_framework_destroy)
    end
end

fingerprint(eparams::EwaldParams) =
    hash((eparams.α, eparams.sr_cutoff_r, [(k.ka, k.kb, k.kc, k.wt) for k in eparams.kvecs]))

# Ewald handles per (framework handle, Ewald parameters); an entry holds its
# framework handle, so that handle stays alive (and its address unreused)
# while the entry exists
const EW_CACHE = Dict{UInt64, Tuple{Handle, Handle}}()

function ewald_handle(fw::Handle, eparams::EwaldParams)
    key = hash((objectid(fw), fingerprint(eparams)))
    hit = get(EW_CACHE, key, nothing)
    if hit !== nothing && hit[1] === fw
        return hit[2]
    end
    kidx = Int32[v for kv in eparams.kvecs for v in (kv.ka, kv.kb, kv.kc)]
    wt = Float64[kv.wt for kv in eparams.kvecs]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pm_ewald_create, libpmhip), Cint,
                (Ptr{Cvoid}, Float64, Float64, Int64, Ptr{Int32}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
                fw.ptr, eparams.α, eparams.sr_cutoff_r, Int64(length(wt)), kidx, wt, out))
    ew = Handle(out[], :pm_ewald_destroy)
    length(EW_CACHE) >= CACHE_LIMIT && empty!(EW_CACHE)
    EW_CACHE[key] = (fw, ew)
    return ew
end

function probe_handle(molecule::Molecule{Cart}, ljff::LJForceField)
    sp = ff_species(ljff)
    idx = Dict(s => Int32(i - 1) for (i, s) in enumerate(sp))
    com = molecule.com.x
    lj = molecule.atoms.coords.x .- com
    qx = molecule.charges.coords.x .- com
    species = Int32[idx[s] for s in molecule.atoms.species]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pm_probe_create, libpmhip), Cint,
                (Int32, Ptr{Float64}, Ptr{Int32}, Int32, Ptr{Float64}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
                Int32(molecule.atoms.n), lj, species, Int32(molecule.charges.n), qx,
                molecule.charges.q, out))
    return Handle(out[], :pm_probe_destroy)
end

# ------------------------------------------------- reference surface ----
"""
    vdw_energy(crystal, molecule, ljff; precision=:fp64)

`src/vdw_energetics.jl:62-64` on the device: the crystal is the (already
replicated) framework, the molecule's fractional coordinates refer to its box.
"""
function vdw_energy(crystal::Crystal, molecule::Molecule, ljff::LJForceField;
                    precision::Symbol=:fp64)
    m = isa(molecule, Molecule{Cart}) ? Frac(molecule, crystal.box) : molecule
    m.atoms.n == 0 && return 0.0
    fw = framework_handle(crystal, ljff, (1, 1, 1))
    ff = forcefield_handle(ljff)
    sp = ff_species(ljff)
    idx = Dict(s => Int32(i - 1) for (i, s) in enumerate(sp))
    species = Int32[idx[s] for s in m.atoms.species]
    out = zeros(1)
    check(ccall((:pm_vdw_energy, libpmhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Int32}, Int64, Ptr{Float64}, Int32, Ptr{Float64}),
                fw.ptr, ff.ptr, Int32(m.atoms.n), species, Int64(1),
                Matrix{Float64}(m.atoms.coords.xf), precision_code(precision), out))
    return out[1]
end

"""
    electrostatic_potential_energy(crystal, molecule, eparams, eikr; precision=:fp64)

`src/electrostatics.jl:331-342` on the device; returns an `EwaldSum`.
"""
function electrostatic_potential_energy(crystal::Crystal, molecule::Molecule{Frac},
                                        eparams::EwaldParams, eikr::Eikr;
                                        precision::Symbol=:fp64)
    molecule.charges.n == 0 && return EwaldSum(0.0, 0.0)
    fw = framework_handle(crystal, es_only_forcefield(crystal, eparams.sr_cutoff_r), (1, 1, 1))
    ew = ewald_handle(fw, eparams)
    sr = zeros(1)
    lr = zeros(1)
    check(ccall((:pm_electrostatic_energy, libpmhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Float64}, Int32,
                 Ptr{Float64}, Ptr{Float64}),
                fw.ptr, ew.ptr, Int32(molecule.charges.n), molecule.charges.q, Int64(1),
                Matrix{Float64}(molecule.charges.coords.xf), precision_code(precision), sr, lr))
    return EwaldSum(sr[1], lr[1])
end

function total_electrostatic_potential_energy(crystal::Crystal, molecules::Array{Molecule{Frac}, 1},
                                              eparams::EwaldParams, eikr::Eikr;
                                              precision::Symbol=:fp64)
    total = EwaldSum(0.0, 0.0)
    for m in molecules
        total += electrostatic_potential_energy(crystal, m, eparams, eikr; precision=precision)
    end
    return total
end

const UNIV_GAS_CONST = 8.3144598e-5
const K_TO_KJ_PER_MOL = 8.3144598e-3
const N_BLOCKS = 5

"""
    widom_results(sums, nb_per_block, temperature, ρ, crystal)

Block statistics and the results `Dict` of `src/henry.jl:112-159` from the
device's per-block sums (3 x N_BLOCKS: Σw, Σw E_vdw, Σw E_es).
"""
function widom_results(sums::Matrix{Float64}, nb_per_block::Int, temperature::Float64, ρ::Float64,
                       crystal::Crystal)
    boltzmann_factor_sums = sums[1, :]
    average_energies = [PotentialEnergy(sums[2, b] / sums[1, b], sums[3, b] / sums[1, b])
                        for b in 1:N_BLOCKS]
    henry_coefficients = boltzmann_factor_sums / (UNIV_GAS_CONST * temperature * nb_per_block)
    err_kh = 2.0 * std(henry_coefficients) / sqrt(N_BLOCKS)
    err_vdw = 2.0 * std([e.vdw for e in average_energies]) / sqrt(N_BLOCKS)
    err_es = 2.0 * std([e.es for e in average_energies]) / sqrt(N_BLOCKS)
    results = Dict{String, Any}()
    results["henry coefficient [mol/(m³-bar)]"] = mean(henry_coefficients)
    results["xtal"] = crystal.name
    results["henry coefficient [mmol/(g-bar)]"] = results["henry coefficient [mol/(m³-bar)]"] / ρ
    results["err henry coefficient [mmol/(g-bar)]"] = err_kh / ρ
    results["henry coefficient [mol/(kg-Pa)]"] = results["henry coefficient [mmol/(g-bar)]"] / 100000.0
    results["⟨U, vdw⟩ (K)"] = mean([e.vdw for e in average_energies])
    results["⟨U, es⟩ (K)"] = mean([e.es for e in average_energies])
    results["⟨U⟩ (K)"] = results["⟨U, vdw⟩ (K)"] + results["⟨U, es⟩ (K)"]
    results["⟨U⟩ (kJ/mol)"] = results["⟨U⟩ (K)"] * K_TO_KJ_PER_MOL
    results["⟨U, vdw⟩ (kJ/mol)"] = results["⟨U, vdw⟩ (K)"] * K_TO_KJ_PER_MOL
    results["err ⟨U, vdw⟩ (kJ/mol)"] = err_vdw * K_TO_KJ_PER_MOL
    results["⟨U, es⟩ (kJ/mol)"] = results["⟨U, es⟩ (K)"] * K_TO_KJ_PER_MOL
    results["err ⟨U, es⟩ (kJ/mol)"] = err_es * K_TO_KJ_PER_MOL
    results["Qst (kJ/mol)"] = -results["⟨U⟩ (kJ/mol)"] + temperature * K_TO_KJ_PER_MOL
    results["err Qst (kJ/mol)"] = (err_vdw + err_es) * K_TO_KJ_PER_MOL
    return results
end

"""
    henry_coefficient(crystal, molecule, temperature, ljff; insertions_per_volume=200,
                      verbose=true, ewald_precision=1e-6, autosave=true,
                      filename_comment="", seed=0, precision=:fp64)

`src/henry.jl:34-181` with the insertions on the device: same insertion count,
blocks, replication, statistics and 16-key results `Dict`.  Insertions come
from the Philox stream keyed by `seed`, each with a fresh Arvo rotation.
"""
function henry_coefficient(crystal::Crystal, molecule::Molecule, temperature::Float64,
                           ljff::LJForceField; insertions_per_volume::Union{Int, Float64}=200,
                           verbose::Bool=true, ewald_precision::Float64=1e-6,
                           autosave::Bool=true, filename_comment::AbstractString="",
                           accessibility_grid::Union{Nothing, Grid{Bool}}=nothing, seed::Integer=0,
                           precision::Symbol=:fp64)
    assert_P1_symmetry(crystal)
    acc = access_handle(crystal, accessibility_grid, verbose)
    time_start = time()
    nb_insertions = max(N_BLOCKS, ceil(Int, insertions_per_volume * crystal.box.Ω))
    nb_per_block = ceil(Int, nb_insertions / N_BLOCKS)
    repfactors = replication_factors(crystal.box, ljff)
    xtal_rep = replicate(crystal, repfactors)
    charged_system = has_charges(crystal) & has_charges(molecule)
    ρ = crystal_density(xtal_rep)
    fw = framework_handle(crystal, ljff, repfactors)
    ff = forcefield_handle(ljff)
    ew = C_NULL
    ew_h = nothing
    if charged_system
        eparams = setup_Ewald_sum(xtal_rep.box, sqrt(ljff.r²_cutoff), ϵ=ewald_precision,
                                  verbose=verbose)
        ew_h = ewald_handle(fw, eparams)
        ew = ew_h.ptr
    end
    pr = probe_handle(molecule, ljff)
    sums = zeros(3, N_BLOCKS)
    GC.@preserve acc begin
        check(ccall((:pm_widom, libpmhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, UInt64, Int32, Int64,
                     Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                    fw.ptr, ff.ptr, ew, pr.ptr, temperature, UInt64(seed), Int32(N_BLOCKS),
                    Int64(nb_per_block), precision_code(precision), sums, C_NULL, handle_ptr(acc)))
    end
    results = widom_results(sums, nb_per_block, temperature, ρ, crystal)
    results["elapsed time (min)"] = (time() - time_start) / 60
    if autosave
        path = rc[:paths][:simulations]
        isdir(path) || mkdir(path)
        savename = joinpath(path, henry_result_savename(crystal, molecule, temperature, ljff,
                                                        insertions_per_volume,
                                                        comment=filename_comment))
        PorousMaterials.JLD2.@save savename results
    end
    return results
end

"""
Mirror of `pm_widom_run` (include/pmhip.h): one run of `pm_widom_batch`.
"""
struct WidomRun
    framework::Ptr{Cvoid}
    forcefield::Ptr{Cvoid}
    ewald::Ptr{Cvoid}
    probe::Ptr{Cvoid}
    temperature::Float64
    seed::UInt64
    n_per_block::Int64
    n_blocks::Int32
    reserved::Int32
    access::Ptr{Cvoid}
end

"""
    henry_coefficients(pairs, temperature, ljff; insertions_per_volume=200,
                       ewald_precision=1e-6, seeds=zeros(Int, length(pairs)),
                       precision=:fp64)

Screening: `henry_coefficient` for every `(crystal, molecule)` in `pairs`, all
runs in one `pm_widom_batch` call.  Each results `Dict` equals the one
`henry_coefficient` returns for that pair and seed.
"""
function henry_coefficients(pairs::Vector{<:Tuple{Crystal, Molecule}}, temperature::Float64,
                            ljff::LJForceField; insertions_per_volume::Union{Int, Float64}=200,
                            ewald_precision::Float64=1e-6, seeds=zeros(Int, length(pairs)),
                            precision::Symbol=:fp64,
                            accessibility_grids=fill(nothing, length(pairs)))
    time_start = time()
    handles = Any[]   # keeps every handle alive until the call returns
    runs = WidomRun[]
    setups = Tuple{Int, Float64}[]
    ff = forcefield_handle(ljff)
    for ((crystal, molecule), seed, agrid) in zip(pairs, seeds, accessibility_grids)
        assert_P1_symmetry(crystal)
        acc = access_handle(crystal, agrid, false)
        isnothing(acc) || push!(handles, acc)
        nb_insertions = max(N_BLOCKS, ceil(Int, insertions_per_volume * crystal.box.Ω))
        nb_per_block = ceil(Int, nb_insertions / N_BLOCKS)
        repfactors = replication_factors(crystal.box, ljff)
        xtal_rep = replicate(crystal, repfactors)
        fw = framework_handle(crystal, ljff, repfactors)
        ew = C_NULL
        if has_charges(crystal) & has_charges(molecule)
            eparams = setup_Ewald_sum(xtal_rep.box, sqrt(ljff.r²_cutoff), ϵ=ewald_precision, verbose=false)
            ew_h = ewald_handle(fw, eparams)
            push!(handles, ew_h)
            ew = ew_h.ptr
        end
        pr = probe_handle(molecule, ljff)
        push!(handles, fw, pr)
        push!(runs, WidomRun(fw.ptr, ff.ptr, ew, pr.ptr, temperature, UInt64(seed), Int64(nb_per_block),
                             Int32(N_BLOCKS), Int32(0), handle_ptr(acc)))
        push!(setups, (nb_per_block, crystal_density(xtal_rep)))
    end
    sums = zeros(3, N_BLOCKS * length(runs))
    GC.@preserve handles ff begin
        check(ccall((:pm_widom_batch, libpmhip), Cint, (Int32, Ptr{WidomRun}, Int32, Ptr{Float64}),
                    Int32(length(runs)), runs, precision_code(precision), sums))
    end
    out = Dict{String, Any}[]
    for (i, (crystal, molecule)) in enumerate(pairs)
        nb_per_block, ρ = setups[i]
        results = widom_results(sums[:, (i - 1) * N_BLOCKS + 1:i * N_BLOCKS], nb_per_block, temperature, ρ,
                                crystal)
        results["elapsed time (min)"] = (time() - time_start) / 60
        push!(out, results)
    end
    return out
end

"""
    energy_grid(crystal, molecule, ljff; resolution=1.0, n_rotations=1000,
                temperature=NaN, units=:kJ_mol, center=false, verbose=true,
                precision=:fp64, electrostatics=false)

`src/grid.jl:291-362` for single-site probes on the device.  As in the
reference, a point's energy is the probe's `vdw_energy` alone
(src/grid.jl:337-338), so a bare point charge (the only charged single-site
probe) gives a grid of zeros.  `electrostatics=true` (no counterpart in the
reference) adds the probe charge times the framework's Ewald potential, the
point-charge energy of `electrostatic_potential_energy(crystal, ion(q, xf), ...)`.
"""
function energy_grid(crystal::Crystal, molecule::Molecule{Cart}, ljforcefield::LJForceField;
                     resolution::Union{Float64, Tuple{Int, Int, Int}}=1.0, n_rotations::Int=1000,
                     temperature::Float64=NaN, units::Symbol=:kJ_mol, center::Bool=false,
                     verbose::Bool=true, precision::Symbol=:fp64, ewald_precision::Float64=1e-6,
                     electrostatics::Bool=false)
    assert_P1_symmetry(crystal)
    units in (:kJ_mol, :K) || error("Pass :kJ_mol or :K for units of kJ/mol or K, respectively.")
    needs_rotations(molecule) &&
        error("energy grids of multi-site probes are not supported by the device path")
    n_pts = isa(resolution, Tuple{Int, Int, Int}) ? resolution : required_n_pts(crystal.box, resolution)
    origin = center ? crystal.box.f_to_c * [-0.5, -0.5, -0.5] : [0.0, 0.0, 0.0]
    charged = electrostatics && has_charges(crystal) && has_charges(molecule)
    # a single-site probe without an LJ sphere is a bare point charge: the
    # reference's grid is its vdw_energy, zero everywhere (src/grid.jl:337-338)
    if molecule.atoms.n == 0 && !charged
        return Grid(crystal.box, n_pts, zeros(Float64, n_pts...), units, origin)
    end
    repfactors = replication_factors(crystal.box, ljforcefield)
    fw = framework_handle(crystal, ljforcefield, repfactors)
    ff = forcefield_handle(ljforcefield)
    ew = C_NULL
    ew_h = nothing
    if charged
        xtal_rep = replicate(crystal, repfactors)
        eparams = setup_Ewald_sum(xtal_rep.box, sqrt(ljforcefield.r²_cutoff), ϵ=ewald_precision)
        ew_h = ewald_handle(fw, eparams)
        ew = ew_h.ptr
    end
    pr = probe_handle(molecule, ljforcefield)
    data = zeros(Float64, n_pts...)
    GC.@preserve ew_h begin
        check(ccall((:pm_energy_grid, libpmhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int32}, Int32, Ptr{Float64}),
                    fw.ptr, ff.ptr, ew, pr.ptr, Int32[n_pts...], precision_code(precision), data))
    end
    if units == :kJ_mol
        data .= data .* 8.314 ./ 1000.0   # the reference's order of operations (src/grid.jl:358)
    end
    return Grid(crystal.box, n_pts, data, units, origin)
end

# ------------------------------------------------------ accessibility ----
handle_ptr(h::Nothing) = C_NULL
handle_ptr(h::Handle) = h.ptr

# Device copy of an accessibility grid (`Grid{Bool}`, i fastest like the
# reference's Array{Bool,3}); the box check is the reference's, which it
# makes only when verbose (src/henry.jl:57-63).
function access_handle(crystal::Crystal, grid, verbose::Bool)
    isnothing(grid) && return nothing
    if verbose
        println("Using provided accessibility grid to block inaccessible pockets")
        isapprox(grid.box, crystal.box) ||
            error(@sprintf("accessibility grid box does not match box of %s.\n", crystal.name))
    end
    data = UInt8.(vec(grid.data))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pm_access_create, libpmhip), Cint, (Ptr{Int32}, Ptr{UInt8}, Ref{Ptr{Cvoid}}),
                Int32[grid.n_pts...], data, out))
    return Handle(out[], :pm_access_destroy)
end

"""
    accessible, segments, (n_segments, n_blocked, n_cycles) = block_pockets(grid, energy_tol;
                                                                            block_pockets=true)

The steps of `compute_accessibility_grid` after the energy grid
(src/grid.jl:383-646, 704-764) in libpmhip.so's host code (pm_block_pockets).
"""
function block_pockets(grid::Grid, energy_tol::Float64; block_pockets::Bool=true, max_cycles::Integer=0)
    n = length(grid.data)
    acc = zeros(UInt8, grid.n_pts...)
    seg = zeros(Int32, grid.n_pts...)
    counts = zeros(Int64, 3)
    check(ccall((:pm_block_pockets, libpmhip), Cint,
                (Ptr{Int32}, Ptr{Float64}, Float64, Int32, Int64, Ptr{UInt8}, Ptr{Int32}, Ptr{Int64}),
                Int32[grid.n_pts...], convert(Array{Float64, 3}, grid.data), energy_tol,
                Int32(block_pockets), Int64(max_cycles), acc, seg, counts))
    return acc .== 1, seg, (counts[1], counts[2], counts[3])
end

"""
    compute_accessibility_grid(crystal, probe, forcefield; resolution=1.0, energy_tol=10.0,
                               energy_units=:kJ_mol, verbose=true, write_b4_after_grids=true,
                               block_inaccessible_pockets=true, precision=:fp64)

`src/grid.jl:689-774` with the energy grid on the device (FP64 by default)
and the flood fill / segment graph / cycles in `block_pockets`; same return
values (grid, segments blocked, porosity).
"""
function compute_accessibility_grid(crystal::Crystal, probe::Molecule, forcefield::LJForceField;
                                    resolution::Union{Float64, Tuple{Int, Int, Int}}=1.0,
                                    energy_tol::Float64=10.0, energy_units::Symbol=:kJ_mol,
                                    verbose::Bool=true, write_b4_after_grids::Bool=true,
                                    block_inaccessible_pockets::Bool=true, precision::Symbol=:fp64)
    assert_P1_symmetry(crystal)
    grid = energy_grid(crystal, probe, forcefield, resolution=resolution, verbose=verbose,
                       units=energy_units, precision=precision)
    if !block_inaccessible_pockets
        acc, _, _ = block_pockets(grid, energy_tol, block_pockets=false)
        return Grid{Bool}(grid.box, grid.n_pts, acc, :accessibility, grid.origin), 0, sum(acc) / length(acc)
    end
    acc, seg, (n_seg, n_blocked, n_cycles) = block_pockets(grid, energy_tol)
    stem = @sprintf("%s_in_%s_%s", probe.species, replace(replace(crystal.name, ".cif" => ""), ".cssr" => ""),
                    replace(forcefield.name, ".csv" => ""))
    if write_b4_after_grids
        write_cube(Grid(grid.box, grid.n_pts, map(l -> l == -1 ? -1 : 1, seg), :Segment_No, grid.origin),
                   stem * "_b4_pocket_blocking.cube")
        write_cube(Grid(grid.box, grid.n_pts, map(a -> a ? 1 : -1, acc), :Segment_No, grid.origin),
                   stem * "_after_pocket_blocking.cube")
    end
    any(acc) || @warn @sprintf("%s cannot enter the pores of %s with %f K energy tolerance.",
                               probe.species, crystal.name, energy_tol)
    porosity = Dict(:b4_blocking => sum(grid.data .< energy_tol) / length(grid.data),
                    :after_blocking => sum(acc) / length(acc))
    if n_blocked != 0
        printstyled(@sprintf("\t%d pockets in %s were found to be inaccessible to %s and blocked.\n",
                             n_blocked, crystal.name, probe.species), color=:yellow)
    end
    return Grid{Bool}(grid.box, grid.n_pts, acc, :accessibility, grid.origin), n_blocked, porosity
end

# ------------------------------------------------------------ adjoint ----
# (epsilon_ij, sigma_ij) of the reference's mixing rules (src/forcefield.jl:75-92),
# generic in the number type so that a complex step differentiates it exactly
function mix(rule::AbstractString, ea, sa, eb, sb)
    r = lowercase(rule)
    if r == "lorentz-berthelot"
        return sqrt(ea * eb), (sa + sb) / 2
    elseif r == "kong"
        es6 = sqrt(ea * sa^6 * eb * sb^6)
        es12 = (((ea * sa^12)^(1 / 13) + (eb * sb^12)^(1 / 13)) / 2)^13
        return es6^2 / es12, (es12 / es6)^(1 / 6)
    else
        return sqrt(ea * eb), sqrt(sa * sb)
    end
end

"""
    mixing_jacobian(ljff, mixing_rules) -> jac (4 x n x n)

d(eps_ij, sigma_ij)/d(eps_i, sigma_i) for the dense tables' species order (the
zero-epsilon slot last, all zeros), by complex step: jac[:, j, i] = (d eps/d eps_i,
d eps/d sigma_i, d sigma/d eps_i, d sigma/d sigma_i) for framework species i and
probe species j, the layout pm_widom_gradient reads as jac[(i*n + j)*4 + k].
"""
function mixing_jacobian(ljff::LJForceField, mixing_rules::AbstractString)
    sp = ff_species(ljff)
    n = length(sp) + 1
    jac = zeros(4, n, n)
    h = 1e-30
    for (i, a) in enumerate(sp), (j, b) in enumerate(sp)
        ea, sa, eb, sb = ljff.pure_ϵ[a], ljff.pure_σ[a], ljff.pure_ϵ[b], ljff.pure_σ[b]
        e1, s1 = mix(mixing_rules, complex(ea, h), sa, eb, sb)
        e2, s2 = mix(mixing_rules, ea, complex(sa, h), eb, sb)
        jac[:, j, i] .= (imag(e1) / h, imag(e2) / h, imag(s1) / h, imag(s2) / h)
    end
    return jac
end

"""
    henry_coefficient_gradient(crystal, molecule, temperature, ljff;
                               mixing_rules="Lorentz-Berthelot", kwargs...) -> (results, grad)

One adjoint device pass (pm_widom_gradient): the `henry_coefficient` results
`Dict` and d ln K_H / d theta for every framework species, theta = its pure
epsilon, its pure sigma (the probe's held fixed) and a uniform shift of the
charges of its atoms: `grad[species] = (epsilon=..., sigma=..., charge=...)`.
"""
function henry_coefficient_gradient(crystal::Crystal, molecule::Molecule, temperature::Float64,
                                    ljff::LJForceField; mixing_rules::AbstractString="Lorentz-Berthelot",
                                    insertions_per_volume::Union{Int, Float64}=200,
                                    ewald_precision::Float64=1e-6, seed::Integer=0, precision::Symbol=:fp64,
                                    accessibility_grid::Union{Nothing, Grid{Bool}}=nothing)
    acc = access_handle(crystal, accessibility_grid, false)
    nb_insertions = max(N_BLOCKS, ceil(Int, insertions_per_volume * crystal.box.Ω))
    nb_per_block = ceil(Int, nb_insertions / N_BLOCKS)
    repfactors = replication_factors(crystal.box, ljff)
    xtal_rep = replicate(crystal, repfactors)
    fw = framework_handle(crystal, ljff, repfactors)
    ff = forcefield_handle(ljff)
    ew = C_NULL
    ew_h = nothing
    if has_charges(crystal) & has_charges(molecule)
        eparams = setup_Ewald_sum(xtal_rep.box, sqrt(ljff.r²_cutoff), ϵ=ewald_precision, verbose=false)
        ew_h = ewald_handle(fw, eparams)
        ew = ew_h.ptr
    end
    pr = probe_handle(molecule, ljff)
    jac = mixing_jacobian(ljff, mixing_rules)
    n = size(jac, 2)
    sums = zeros(3, N_BLOCKS)
    grads = zeros(3, n, N_BLOCKS)
    GC.@preserve acc begin
        check(ccall((:pm_widom_gradient, libpmhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Float64, UInt64, Int32, Int64, Int32,
                     Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}),
                    fw.ptr, ff.ptr, ew, pr.ptr, temperature, UInt64(seed), Int32(N_BLOCKS), Int64(nb_per_block),
                    precision_code(precision), jac, sums, grads, handle_ptr(acc)))
    end
    results = widom_results(sums, nb_per_block, temperature, crystal_density(xtal_rep), crystal)
    S = sum(sums[1, :])
    G = dropdims(sum(grads, dims=3), dims=3)
    sp = ff_species(ljff)
    grad = Dict{Symbol, NamedTuple}()
    for s in unique(crystal.atoms.species)
        k = findfirst(==(s), sp)
        d = -G[:, k] / (temperature * S)
        grad[s] = (epsilon=d[1], sigma=d[2], charge=d[3])
    end
    return results, grad
end

"""
    henry_kh(theta, crystal, molecule, temperature, ljff; species, mixing_rules, kwargs...)

K_H (mol/(kg Pa)) as a function of theta = [eps_t; sigma_t; dq_t] for the
framework species `species` (pure parameters, and a uniform shift of the
charges of their atoms): the differentiable form of `henry_coefficient`,
with the ChainRulesCore rrule below as its backward pass.
"""
function henry_kh(theta::AbstractVector, crystal::Crystal, molecule::Molecule, temperature::Float64,
                  ljff::LJForceField; species::Vector{Symbol}, mixing_rules::AbstractString="Lorentz-Berthelot",
                  kwargs...)
    xtal, ff = apply_parameters(theta, crystal, ljff, species, mixing_rules)
    return henry_coefficient(xtal, molecule, temperature, ff; verbose=false, autosave=false,
                             kwargs...)["henry coefficient [mol/(kg-Pa)]"]
end

# The crystal and force field at parameters theta: pure (eps, sigma) of
# `species` replaced, cross terms re-mixed, and the charges of the atoms of
# each species shifted by dq_t.
function apply_parameters(theta, crystal::Crystal, ljff::LJForceField, species::Vector{Symbol},
                          mixing_rules::AbstractString)
    nt = length(species)
    pure_ϵ = copy(ljff.pure_ϵ)
    pure_σ = copy(ljff.pure_σ)
    for (t, s) in enumerate(species)
        pure_ϵ[s] = theta[t]
        pure_σ[s] = theta[nt + t]
    end
    ϵ = Dict{Symbol, Dict{Symbol, Float64}}()
    σ² = Dict{Symbol, Dict{Symbol, Float64}}()
    for a in keys(pure_σ)
        ϵ[a] = Dict{Symbol, Float64}()
        σ²[a] = Dict{Symbol, Float64}()
        for b in keys(pure_σ)
            e, sg = mix(mixing_rules, pure_ϵ[a], pure_σ[a], pure_ϵ[b], pure_σ[b])
            ϵ[a][b] = e
            σ²[a][b] = sg^2
        end
    end
    ff = LJForceField(ljff.name, pure_σ, pure_ϵ, σ², ϵ, ljff.r²_cutoff)
    q = copy(crystal.charges.q)
    for c in 1:crystal.charges.n
        x = crystal.charges.coords.xf[:, c]
        k = findfirst(i -> crystal.atoms.coords.xf[:, i] == x, 1:crystal.atoms.n)
        k === nothing && continue
        t = findfirst(==(crystal.atoms.species[k]), species)
        t === nothing || (q[c] += theta[2 * nt + t])
    end
    charges = Charges(q, Frac(crystal.charges.coords.xf))
    xtal = Crystal(crystal.name, crystal.box, crystal.atoms, charges, crystal.bonds, crystal.symmetry)
    return xtal, ff
end

function ChainRulesCore.rrule(::typeof(henry_kh), theta::AbstractVector, crystal::Crystal, molecule::Molecule,
                              temperature::Float64, ljff::LJForceField; species::Vector{Symbol},
                              mixing_rules::AbstractString="Lorentz-Berthelot", kwargs...)
    xtal, ff = apply_parameters(theta, crystal, ljff, species, mixing_rules)
    results, grad = henry_coefficient_gradient(xtal, molecule, temperature, ff; mixing_rules=mixing_rules,
                                               kwargs...)
    kh = results["henry coefficient [mol/(kg-Pa)]"]
    dlnk = vcat([grad[s].epsilon for s in species], [grad[s].sigma for s in species],
                [grad[s].charge for s in species])
    henry_kh_pullback(ȳ) = (NoTangent(), ȳ * kh .* dlnk, NoTangent(), NoTangent(), NoTangent(), NoTangent())
    return kh, henry_kh_pullback
end

# ------------------------------------------------------ forces, minimum ----
"""
    site_forces(crystal, molecule, ljff; eparams=nothing, precision=:fp64) -> (E, dE)

Energy (K) and dE/d(site Cartesian xyz) (3 x n_sites, K/A) of a molecule pose
in the (replicated) crystal through pm_pose_gradient: LJ spheres first, then
charges.
"""
function site_forces(crystal::Crystal, molecule::Molecule{Frac}, ljff::LJForceField; eparams=nothing,
                     precision::Symbol=:fp64, rep::Tuple{Int, Int, Int}=(1, 1, 1))
    sp = ff_species(ljff)
    idx = Dict(s => Int32(i - 1) for (i, s) in enumerate(sp))
    charged = eparams !== nothing && has_charges(molecule)
    species = vcat(Int32[idx[s] for s in molecule.atoms.species], fill(Int32(-1), charged ? molecule.charges.n : 0))
    q = vcat(zeros(molecule.atoms.n), charged ? molecule.charges.q : Float64[])
    xf = charged ? hcat(molecule.atoms.coords.xf, molecule.charges.coords.xf) : molecule.atoms.coords.xf
    fw = framework_handle(crystal, ljff, rep)
    ew_h = charged ? ewald_handle(fw, eparams) : nothing
    e = zeros(3)
    g = zeros(3, length(species))
    check(ccall((:pm_pose_gradient, libpmhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Float64}, Int64, Ptr{Float64}, Int32,
                 Ptr{Float64}, Ptr{Float64}),
                fw.ptr, forcefield_handle(ljff).ptr, charged ? ew_h.ptr : C_NULL, Int32(length(species)), species,
                q, Int64(1), Matrix{Float64}(xf), precision_code(precision), e, g))
    return sum(e), g
end

"""
    find_energy_minimum(xtal, molecule, ljff) -> (molecule, E_min kJ/mol)

`src/energy_min.jl:19-58` with Optim's L-BFGS on device forces in place of
Nelder-Mead on `vdw_energy`: centre of mass only, as the reference.
"""
function find_energy_minimum(xtal::Crystal, molecule::Molecule, ljff::LJForceField)
    rep = replication_factors(xtal.box, ljff)
    xtal_rep = replicate(xtal, rep)
    m = isa(molecule, Molecule{Cart}) ? Frac(molecule, xtal_rep.box) :
        (translate_to!(molecule, Frac(molecule.com.xf ./ rep)); molecule)
    F = xtal_rep.box.f_to_c
    function fg!(f, g, xf)
        translate_to!(m, Frac(mod.(xf, 1.0)))
        e, dE = site_forces(xtal, m, ljff; rep=rep)
        g === nothing || (g .= transpose(F) * vec(sum(dE, dims=2)))  # d/d xf = F^T d/dx
        return e * 8.314 / 1000
    end
    res = optimize(Optim.only_fg!((f, g, x) -> (g === nothing ? fg!(f, nothing, x) :
                                                (v = fg!(f, g, x); g .*= 8.314 / 1000; v))),
                   m.com.xf, LBFGS())
    xf_min = mod.(mod.(Optim.minimizer(res), 1.0) .* rep, 1.0)
    translate_to!(m, Frac(xf_min ./ rep))
    return Frac(Cart(m, xtal_rep.box), xtal.box), Optim.minimum(res)
end

# ------------------------------------------------------------- GCMC ----
const PM_MAX_GCMC_SPECIES = 4
const PM_GCMC_BLOCKS = 5
const KB_PA_A3 = 1.38064852e7  # src/PorousMaterials.jl:14

"""Mirror of `pm_gcmc_settings` (include/pmhip.h), same C layout."""
struct GcmcSettings
    n_burn_cycles::Int64
    n_sample_cycles::Int64
    sample_frequency::Int32
    snapshot_frequency::Int32
    capacity::Int32
    density_species::Int32
    density_site_mask::UInt32
    density_n_pts::NTuple{3, Int32}
    density_sim_box::Int32
    delta::Float64
    seed::UInt64
end

"""Mirror of `pm_gcmc_state` (include/pmhip.h); `block_n[b * 4 + s + 1]` is
block b (0-based), species s."""
struct GcmcState
    cycles_done::Int64
    markov_time::Int64
    num_snapshots::Int64
    n::NTuple{4, Int32}
    block::Int32
    error::Int32
    delta::Float64
    energy_gh::Float64
    energy_gg::Float64
    n_proposed::NTuple{6, Int64}
    n_accepted::NTuple{6, Int64}
    block_samples::NTuple{5, Int64}
    block_n::NTuple{20, Int64}
    block_n2::NTuple{20, Int64}
    block_gh::NTuple{5, Float64}
    block_gg::NTuple{5, Float64}
    block_gh2::NTuple{5, Float64}
    block_gg2::NTuple{5, Float64}
    block_un::NTuple{5, Float64}
end

function gcmc_state(gc::Handle, chain::Integer)
    st = Ref{GcmcState}()
    check(ccall((:pm_gcmc_get_state, libpmhip), Cint, (Ptr{Cvoid}, Int32, Ref{GcmcState}), gc.ptr, Int32(chain), st))
    return st[]
end

# the reference's per-block statistics (src/gcmc.jl:30-62) from the device's block sums
function gcmc_blocks(st::GcmcState, nb_species::Int)
    E(gh, gg) = PorousMaterials.SystemPotentialEnergy(PorousMaterials.PotentialEnergy(gh, 0.0),
                                                      PorousMaterials.PotentialEnergy(gg, 0.0))
    return [PorousMaterials.GCMCstats(Int(st.block_samples[b]),
                                      [Int(st.block_n[(b - 1) * PM_MAX_GCMC_SPECIES + s]) for s in 1:nb_species],
                                      [Int(st.block_n2[(b - 1) * PM_MAX_GCMC_SPECIES + s]) for s in 1:nb_species],
                                      E(st.block_gh[b], st.block_gg[b]), E(st.block_gh2[b], st.block_gg2[b]),
                                      st.block_un[b]) for b in 1:PM_GCMC_BLOCKS]
end

# a device molecule as the reference returns it: Cartesian in the simulation
# box, charges carried by the rigid motion that takes the template's LJ
# spheres onto the device's (Kabsch)
function device_molecule(template::Molecule{Cart}, sites_xf::Matrix{Float64}, com_xf::Vector{Float64}, box::Box)
    x = box.f_to_c * sites_xf
    c = box.f_to_c * com_xf
    P = template.atoms.coords.x .- template.com.x
    F = svd((x .- c) * P')
    R = F.U * Diagonal([1.0, 1.0, sign(det(F.U * F.Vt))]) * F.Vt
    qx = c .+ R * (template.charges.coords.x .- template.com.x)
    return Molecule(template.species, Atoms(template.atoms.species, Cart(x)),
                    Charges(template.charges.q, Cart(qx)), Cart(c))
end

# chains for every pressure set in one pm_gcmc handle (src/gcmc.jl:175-817 per chain)
function gcmc_sims(xtal::Crystal, molecule_templates::Array{Molecule{Cart}, 1}, temperature::Float64,
                   pressure_sets::Vector{Vector{Float64}}, ljff::LJForceField, molecule_sets::Vector;
                   n_cycles_per_volume::Int=200, fraction_burn_cycles::Float64=0.5, sample_frequency::Int=1,
                   verbose::Bool=true, ewald_precision::Float64=1e-6, eos::Symbol=:ideal, autosave::Bool=true,
                   show_progress_bar::Bool=false, write_adsorbate_snapshots::Bool=false,
                   snapshot_frequency::Int=1, calculate_density_grid::Bool=false, density_grid_dx::Float64=1.0,
                   density_grid_molecular_species::Union{Nothing, Symbol}=nothing,
                   density_grid_sim_box::Bool=true, results_filename_comment::String="",
                   n_burn_cycles::Union{Nothing, Int}=nothing, n_sample_cycles::Union{Nothing, Int}=nothing,
                   seed::Integer=0, capacity::Union{Nothing, Int}=nothing)
    assert_P1_symmetry(xtal)
    write_adsorbate_snapshots && error("xyz snapshots of the adsorbates are not written by the device chains")
    start_time = time()
    nb_species = length(molecule_templates)
    molecular_species = [mt.species for mt in molecule_templates]
    all(p -> length(p) == nb_species, pressure_sets) || error("# molecules in simulation: $nb_species; " *
                                                              "each pressure set needs as many partial pressures")
    if has_charges(xtal) && any(has_charges, molecule_templates)
        error("sorry, electrostatics not supported")  # src/gcmc.jl:378-382
    end
    nb_cycles = max(N_BLOCKS, ceil(Int, n_cycles_per_volume * xtal.box.Ω))
    @assert (0.0 < fraction_burn_cycles) && (fraction_burn_cycles < 1.0)
    n_burn = isnothing(n_burn_cycles) ? ceil(Int, nb_cycles * fraction_burn_cycles) : n_burn_cycles
    n_sample = isnothing(n_sample_cycles) ? ceil(Int, nb_cycles * (1 - fraction_burn_cycles)) : n_sample_cycles
    # fugacities (Pa) from an equation of state (src/gcmc.jl:228-243)
    fugacity_sets = map(pressure_sets) do pressures
        if eos == :ideal
            pressures * 100000.0
        elseif eos == :PengRobinson
            nb_species == 1 || error("Peng Robinson EOS not implemented for >1 species")
            props = calculate_properties(PengRobinsonFluid(molecule_templates[1].species), temperature,
                                         pressures[1], verbose=false)
            [props["fugacity (bar)"] * 100000.0]
        else
            error("eos=:ideal and eos=:PengRobinson are the only valid options for an equation of state.")
        end
    end
    repfactors = replication_factors(xtal.box, ljff)
    xtal_rep = replicate(xtal, repfactors)
    box = xtal_rep.box
    # density grid (src/gcmc.jl:286-325)
    dens_species, dens_mask, n_pts = Int32(-1), UInt32(0), (Int32(1), Int32(1), Int32(1))
    if calculate_density_grid
        if isnothing(density_grid_molecular_species)
            nb_species == 1 || error("Pass `density_grid_molecular_species` to specify which adsorbate to make a grid for.")
            density_grid_molecular_species = molecular_species[1]
        end
        sid = findfirst(==(density_grid_molecular_species), molecular_species)
        dens_species = Int32(sid - 1)
        dens_mask = UInt32(sum(UInt32(1) << (k - 1) for (k, a) in enumerate(molecule_templates[sid].atoms.species)
                               if a == density_grid_molecular_species; init=UInt32(0)))
        n_pts = Int32.(required_n_pts(density_grid_sim_box ? box : xtal.box, density_grid_dx))
    end
    cap = isnothing(capacity) ? Int(min(1 << 18, max(256, ceil(Int, 0.1 * box.Ω)))) : capacity
    settings = GcmcSettings(n_burn, n_sample, Int32(sample_frequency), Int32(snapshot_frequency), Int32(cap),
                            dens_species, dens_mask, n_pts, Int32(density_grid_sim_box), 2.0, UInt64(seed))
    fw = framework_handle(xtal, ljff, repfactors)
    ff = forcefield_handle(ljff)
    probes = [probe_handle(m, ljff) for m in molecule_templates]
    n_chains = length(pressure_sets)
    fug = Float64[f for fs in fugacity_sets for f in fs]   # [chain * n_species + s]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pm_gcmc_create, libpmhip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Int32, Ptr{Float64}, Ptr{Float64},
                 Ref{GcmcSettings}, Ref{Ptr{Cvoid}}),
                fw.ptr, ff.ptr, Int32(nb_species), [p.ptr for p in probes], Int32(n_chains),
                fill(temperature, n_chains), fug, settings, out))
    gc = Handle(out[], :pm_gcmc_destroy)
    # starting configurations (src/gcmc.jl:386-402)
    for (c, mols) in enumerate(molecule_sets)
        isnothing(mols) && continue
        for (s, ms) in enumerate(mols)
            isempty(ms) && continue
            all(m -> m.species == molecular_species[s], ms) || error("initializing with wrong molecule species")
            sites = reduce(vcat, [vec(box.c_to_f * m.atoms.coords.x) for m in ms])
            coms = reduce(vcat, [box.c_to_f * m.com.x for m in ms])
            check(ccall((:pm_gcmc_set_molecules, libpmhip), Cint,
                        (Ptr{Cvoid}, Int32, Int32, Int32, Ptr{Float64}, Ptr{Float64}),
                        gc.ptr, Int32(c - 1), Int32(s - 1), Int32(length(ms)), sites, coms))
        end
    end
    verbose && @printf("μVT simulation of %s in %s at %f K (%d chains), %d burn + %d sample cycles\n",
                       join(molecular_species, ", "), xtal.name, temperature, n_chains, n_burn, n_sample)
    # launches grow towards ~1 s of device time; the chains stay on the device between them
    done, step = 0, 1
    while done < n_burn + n_sample
        k = min(step, n_burn + n_sample - done)
        t = @elapsed check(ccall((:pm_gcmc_run, libpmhip), Cint, (Ptr{Cvoid}, Int64), gc.ptr, Int64(k)))
        done += k
        step = max(1, Int(floor(k * clamp(1.0 / max(t, 1e-3), 0.5, 4.0))))
        show_progress_bar && @printf("\tcycle %d / %d\n", done, n_burn + n_sample)
    end
    elapsed_time = time() - start_time
    nr = repfactors[1] * repfactors[2] * repfactors[3]
    outputs = Tuple{Dict{String, Any}, Vector{Vector{Molecule{Cart}}}}[]
    for c in 1:n_chains
        st = gcmc_state(gc, c - 1)
        st.error == 0 || error("chain $c reached the molecule capacity $cap; pass a larger capacity")
        gh, gg = Ref(0.0), Ref(0.0)
        check(ccall((:pm_gcmc_total_energy, libpmhip), Cint, (Ptr{Cvoid}, Int32, Ref{Float64}, Ref{Float64}),
                    gc.ptr, Int32(c - 1), gh, gg))
        if !(isapprox(gh[], st.energy_gh, atol=0.01) && isapprox(gg[], st.energy_gg, atol=0.01))
            error("energy incremented improperly during simulation...")  # src/gcmc.jl:729-741
        end
        stats = gcmc_blocks(st, nb_species)
        avg_n, err_n, avg_U, err_U = PorousMaterials.mean_stderr_n_U(stats)
        tot = sum(stats)
        results = Dict{String, Any}()
        results["xtal"] = xtal.name
        results["adsorbate"] = molecular_species
        results["forcefield"] = ljff.name
        results["pressure (bar)"] = pressure_sets[c]
        results["fugacity (bar)"] = fugacity_sets[c] / 100000.0
        results["temperature (K)"] = temperature
        results["repfactors"] = repfactors
        results["# sample cycles"] = n_sample
        results["# burn cycles"] = n_burn
        results["# samples"] = tot.n_samples
        results["elapsed time (min)"] = elapsed_time / 60
        results["⟨N⟩ (molecules)"] = avg_n
        results["⟨U_gh, vdw⟩ (K)"] = avg_U.gh.vdw
        results["⟨U_gh, electro⟩ (K)"] = avg_U.gh.es
        results["⟨U_gg, vdw⟩ (K)"] = avg_U.gg.vdw
        results["⟨U_gg, electro⟩ (K)"] = avg_U.gg.es
        results["⟨U⟩ (K)"] = sum(avg_U)
        results["var(N)"] = (tot.n² / tot.n_samples) - (results["⟨N⟩ (molecules)"] .^ 2)
        results["Q_st (K)"] = temperature .- (tot.Un / tot.n_samples .- results["⟨U⟩ (K)"] * results["⟨N⟩ (molecules)"]) ./
                              results["var(N)"]
        results["err ⟨N⟩ (molecules)"] = err_n
        results["err ⟨U_gh, vdw⟩ (K)"] = err_U.gh.vdw
        results["err ⟨U_gh, electro⟩ (K)"] = err_U.gh.es
        results["err ⟨U_gg, vdw⟩ (K)"] = err_U.gg.vdw
        results["err ⟨U_gg, electro⟩ (K)"] = err_U.gg.es
        results["err ⟨U⟩ (K)"] = sum(err_U)
        results["⟨N⟩ (molecules/unit cell)"] = avg_n / nr
        results["err ⟨N⟩ (molecules/unit cell)"] = err_n / nr
        to_mmol_g = 1000 / (6.022140857e23 * molecular_weight(xtal_rep) * 1.66054e-24) * nr
        results["⟨N⟩ (mmol/g)"] = results["⟨N⟩ (molecules/unit cell)"] * to_mmol_g
        results["err ⟨N⟩ (mmol/g)"] = results["err ⟨N⟩ (molecules/unit cell)"] * to_mmol_g
        for (id, desc) in PorousMaterials.PROPOSAL_ENCODINGS
            results[@sprintf("Total # %s proposals", desc)] = Int(st.n_proposed[id])
            results[@sprintf("Fraction of %s proposals accepted", desc)] = st.n_accepted[id] / st.n_proposed[id]
        end
        dbox = density_grid_sim_box ? box : xtal.box
        density = Grid(dbox, Tuple(Int.(n_pts)), zeros(Int.(n_pts)...), :inverse_A3, [0.0, 0.0, 0.0])
        if calculate_density_grid
            counts = zeros(prod(Int.(n_pts)))
            check(ccall((:pm_gcmc_get_density, libpmhip), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                        gc.ptr, Int32(c - 1), counts))
            density.data .= reshape(counts, Int.(n_pts)...) ./ max(st.num_snapshots, 1)
        end
        results["density grid"] = density
        results["num snapshots"] = Int(st.num_snapshots)
        verbose && PorousMaterials.print_results(results, print_title=false)
        molecules = Vector{Vector{Molecule{Cart}}}()
        for s in 1:nb_species
            n = Ref{Int32}(0)
            check(ccall((:pm_gcmc_get_molecules, libpmhip), Cint,
                        (Ptr{Cvoid}, Int32, Int32, Int32, Ref{Int32}, Ptr{Float64}, Ptr{Float64}),
                        gc.ptr, Int32(c - 1), Int32(s - 1), Int32(0), n, C_NULL, C_NULL))
            ns = molecule_templates[s].atoms.n
            sites = zeros(3, ns, max(n[], 1))
            coms = zeros(3, max(n[], 1))
            check(ccall((:pm_gcmc_get_molecules, libpmhip), Cint,
                        (Ptr{Cvoid}, Int32, Int32, Int32, Ref{Int32}, Ptr{Float64}, Ptr{Float64}),
                        gc.ptr, Int32(c - 1), Int32(s - 1), n[], n, sites, coms))
            push!(molecules, [device_molecule(molecule_templates[s], sites[:, :, i], coms[:, i], box)
                              for i in 1:n[]])
        end
        if autosave
            path = rc[:paths][:simulations]
            isdir(path) || mkdir(path)
            savename = joinpath(path, μVT_output_filename(xtal_rep, molecule_templates, temperature,
                                                          pressure_sets[c], ljff, n_burn, n_sample,
                                                          comment=results_filename_comment))
            PorousMaterials.JLD2.@save savename results
            verbose && println("\tresults dictionary saved in ", savename)
        end
        push!(outputs, (results, molecules))
    end
    return outputs
end

"""
    results, molecules = μVT_sim(xtal, molecule_templates, temperature, pressures, ljff; molecules=nothing,
                                 <the reference's keywords>, seed=0, capacity=nothing)

`src/gcmc.jl:175-817` with the Markov chain on the device (`pm_gcmc_*`): the
same cycle counts, fugacities, replication, moves, acceptance rules, block
statistics (the reference's own `GCMCstats` and `mean_stderr_n_U`), energy
check and results `Dict`, saved as JLD2.  The chain draws from the Philox
stream keyed by `seed`.
"""
function μVT_sim(xtal::Crystal, molecule_templates::Array{Molecule{Cart}, 1}, temperature::Float64,
                 pressures::Array{Float64, 1}, ljff::LJForceField;
                 molecules::Union{Nothing, Array{Array{Molecule{Cart}, 1}, 1}}=nothing, kwargs...)
    return gcmc_sims(xtal, molecule_templates, temperature, [pressures], ljff, [molecules]; kwargs...)[1]
end

μVT_sim(xtal::Crystal, molecule_template::Molecule{Cart}, temperature::Float64, pressure::Float64,
        ljff::LJForceField; kwargs...) = μVT_sim(xtal, [molecule_template], temperature, [pressure], ljff; kwargs...)

"""
    results = adsorption_isotherm(xtal, molecule_templates, temperature, pressures, ljff; kwargs...)

`src/gcmc.jl:996-1030`: every pressure set is a chain of one device handle
(the reference's `pmap` over pressures), all advanced by the same launches.
"""
function adsorption_isotherm(xtal::Crystal, molecule_templates::Array{Molecule{Cart}, 1}, temperature::Float64,
                             pressures::Array{Array{Float64, 1}, 1}, ljff::LJForceField; kwargs...)
    return [r for (r, _) in gcmc_sims(xtal, molecule_templates, temperature, pressures, ljff,
                                      fill(nothing, length(pressures)); kwargs...)]
end

adsorption_isotherm(xtal::Crystal, molecule_template::Molecule{Cart}, temperature::Float64,
                    pressures::Array{Float64, 1}, ljff::LJForceField; kwargs...) =
    adsorption_isotherm(xtal, [molecule_template], temperature, [[p] for p in pressures], ljff; kwargs...)

"""
    results = stepwise_adsorption_isotherm(xtal, molecule_templates, temperature, pressures, ljff; kwargs...)

`src/gcmc.jl:940-975`: one device chain per pressure in turn, each starting
from the previous one's molecules; pressure i draws from stream `seed + i - 1`.
"""
function stepwise_adsorption_isotherm(xtal::Crystal, molecule_templates::Array{Molecule{Cart}, 1},
                                      temperature::Float64, pressures::Array{Array{Float64, 1}, 1},
                                      ljff::LJForceField; seed::Integer=0, kwargs...)
    length(molecule_templates) == 1 || error("We only calculate pure gas isotherms")
    results = Dict{String, Any}[]
    molecules = nothing
    for (i, pressure) in enumerate(pressures)
        result, molecules = μVT_sim(xtal, molecule_templates, temperature, pressure, ljff; molecules=molecules,
                                    seed=seed + i - 1, kwargs...)
        push!(results, result)
    end
    return results
end

stepwise_adsorption_isotherm(xtal::Crystal, molecule_template::Molecule{Cart}, temperature::Float64,
                             pressures::Array{Float64, 1}, ljff::LJForceField; kwargs...) =
    stepwise_adsorption_isotherm(xtal, [molecule_template], temperature, [[p] for p in pressures], ljff; kwargs...)

"""
    enable!()

Route PorousMaterials' own `henry_coefficient`, `energy_grid`, `vdw_energy`,
`electrostatic_potential_energy`, `total_electrostatic_potential_energy` and
`μVT_sim` through the device (methods replacing the reference's for the same
argument types).  Every routed call computes in FP64, as the reference does;
`precision=:mixed` stays an explicit opt-in of this module's own methods.
The reference's guest-guest `vdw_energy` methods (molecule vs molecules) are
left as they are.
"""
function enable!()
    @eval PorousMaterials.henry_coefficient(crystal::Crystal, molecule::Molecule, temperature::Float64,
                                            ljff::LJForceField; kwargs...) =
        PorousMaterialsHIP.henry_coefficient(crystal, molecule, temperature, ljff; kwargs...)
    @eval PorousMaterials.energy_grid(crystal::Crystal, molecule::Molecule{Cart}, ljff::LJForceField;
                                      kwargs...) =
        PorousMaterialsHIP.energy_grid(crystal, molecule, ljff; kwargs...)
    @eval PorousMaterials.vdw_energy(crystal::Crystal, molecule::Molecule, ljff::LJForceField) =
        PorousMaterialsHIP.vdw_energy(crystal, molecule, ljff)
    @eval PorousMaterials.electrostatic_potential_energy(crystal::Crystal, molecule::Molecule{Frac},
                                                         eparams::EwaldParams, eikr::Eikr) =
        PorousMaterialsHIP.electrostatic_potential_energy(crystal, molecule, eparams, eikr)
    @eval PorousMaterials.total_electrostatic_potential_energy(crystal::Crystal,
                                                               molecules::Array{Molecule{Frac}, 1},
                                                               eparams::EwaldParams, eikr::Eikr) =
        PorousMaterialsHIP.total_electrostatic_potential_energy(crystal, molecules, eparams, eikr)
    # the reference's isotherm drivers call μVT_sim, so they reach the device too
    @eval PorousMaterials.μVT_sim(xtal::Crystal, molecule_templates::Array{Molecule{Cart}, 1},
                                  temperature::Float64, pressures::Array{Float64, 1}, ljff::LJForceField;
                                  kwargs...) =
        PorousMaterialsHIP.μVT_sim(xtal, molecule_templates, temperature, pressures, ljff; kwargs...)
    return nothing
end

end # module
