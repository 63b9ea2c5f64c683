#=
FLOWVPMHIP -- Julia binding of libflowvpm_hip.so (include/flowvpm_hip.h) for
FLOWVPM.jl.  Thin ccall layer: every numerical call runs on the MI355X.

    using FLOWVPM, FLOWVPMHIP
    pfield = FLOWVPM.ParticleField(n; UJ=FLOWVPMHIP.UJ_hip, kernel=FLOWVPM.gaussianerf)

`UJ_hip` has the keyword signature of FLOWVPM.UJ_direct (src/FLOWVPM_UJ.jl:21-37)
and mutates pfield.particles in place; `direct_hip!` is the FastMultipole
leaf-pair hook (src/FLOWVPM_fmm.jl:102-168); `rungekutta3_hip` runs one full
rungekutta3 step on device-resident state (src/FLOWVPM_timeintegration.jl:388-461);
the rrule of `uj_hip` replaces the ChainRules pullback of src/FLOWVPM_rrules.jl.
=#
module FLOWVPMHIP

import FLOWVPM
import ChainRulesCore

const vpm = FLOWVPM
const libvpm = get(ENV, "FLOWVPM_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libflowvpm_hip.so"))

const VPM_OK = Cint(0)
const VPM_UJ_RESET, VPM_UJ_RESET_SFS, VPM_UJ_SFS = UInt32(1), UInt32(2), UInt32(4)
const VPM_UJ_F64_FAR = UInt32(16)

struct VpmConfig
    kernel::Int32
    precision::Int32      # 0 FP32 pairs + compensated sums, 1 FP64
    transposed::Int32
    reserved::Int32
end

mutable struct VpmStepConfig
    uj::VpmConfig
    f::Float64; g::Float64; Uinf::NTuple{3,Float64}
    sfs::Int32; sfs_force_positive::Int32; sfs_clip::Int32; relax::Int32
    sfs_alpha::Float64; sfs_rlxf::Float64; sfs_minC::Float64; sfs_maxC::Float64
    rlxf::Float64
    viscous::Int32; uj_flags::Int32     # 0 or VPM_UJ_F64_FAR
    nu::Float64; sgm0::Float64; beta::Float64; t_sgm::Float64
    rbf_itmax::Int32; rbf_iterror::Int32; rbf_tol::Float64   # CoreSpreading itmax, iterror, tol
end

check(rc) = rc == VPM_OK ? nothing :
    error("libflowvpm_hip: ", unsafe_string(ccall((:vpm_last_error, libvpm), Cstring, ())))

function kernel_id(k)
    k === vpm.kernel_singular    && return Int32(0)
    k === vpm.kernel_gaussian    && return Int32(1)
    k === vpm.kernel_gaussianerf && return Int32(2)
    k === vpm.kernel_winckelmans && return Int32(3)
    error("kernel not provided by libflowvpm_hip")
end

config(pfield; precision=0) = VpmConfig(kernel_id(pfield.kernel), Int32(precision), Int32(pfield.transposed), 0)

# One device field per ParticleField, kept across calls (no re-allocation).
# The keys are held weakly, so a ParticleField that is no longer referenced
# is collected, and its finalizer releases the device field (records, stream,
# scratch buffers).  The finalizer owns the handle it frees; it does not look
# the field up again.
const FIELDS = WeakKeyDict{Any,Ptr{Cvoid}}()
const FIELDS_LOCK = ReentrantLock()

function device_field(pfield)
    lock(FIELDS_LOCK) do
        h = get(FIELDS, pfield, C_NULL)
        if h == C_NULL
            r = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:vpm_create, libvpm), Cint, (Cint, Int64, Ref{Ptr{Cvoid}}),
                        pfield.useGPU > 0 ? pfield.useGPU - 1 : 0, pfield.maxparticles, r))
            h = r[]
            FIELDS[pfield] = h
            handle = h
            finalizer(_ -> ccall((:vpm_destroy, libvpm), Cint, (Ptr{Cvoid},), handle), pfield)
        end
        return h
    end
end

upload(pfield) = GC.@preserve pfield check(ccall((:vpm_set_particles, libvpm), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Int64), device_field(pfield), pfield.particles, pfield.np))

download(pfield, row0, row1) = GC.@preserve pfield check(ccall((:vpm_get_particles, libvpm), Cint,
    (Ptr{Cvoid}, Ptr{Float64}, Int64, Int32, Int32), device_field(pfield), pfield.particles, pfield.np, row0, row1))

"""
    UJ_hip(pfield; rbf=false, sfs=false, reset=true, reset_sfs=false, precision=0, far64=false)

UJ_direct semantics on the GPU: reset U, vorticity, J, PSE of non-static
particles; SFS reset if `reset_sfs`; all-pairs U and J; E_str if `sfs`.
`far64=true` (with `precision=1`) takes the FP64 far-form pair loop
(VPM_UJ_F64_FAR): one rsqrt per pair, g = 1 past q = 10 (Gaussian-erf) or
q^2 = 13 (Gaussian), the reference arithmetic to FP64 rounding.
"""
function UJ_hip(pfield::vpm.ParticleField; rbf::Bool=false, sfs::Bool=false,
                reset=true, reset_sfs=false, precision=0, far64::Bool=false, optargs...)
    pfield.np == 0 && return nothing
    upload(pfield)
    flags = (reset ? VPM_UJ_RESET : 0x0) | (reset_sfs ? VPM_UJ_RESET_SFS : 0x0) | (sfs ? VPM_UJ_SFS : 0x0) |
            (far64 ? VPM_UJ_F64_FAR : 0x0)
    cfg = Ref(config(pfield; precision))
    check(ccall((:vpm_uj_direct, libvpm), Cint, (Ptr{Cvoid}, Ref{VpmConfig}, UInt32),
                device_field(pfield), cfg, flags))
    download(pfield, Int32(9), Int32(27))                # U, vorticity, J, PSE (0-based rows)
    (sfs || reset_sfs) && download(pfield, Int32(39), Int32(42))
    return nothing
end

"""
    direct_hip!(pfield, target_ranges, source_ranges; precision=0)

FastMultipole leaf-pair P2P: for each k, the particles in source_ranges[k]
act on the particles in target_ranges[k] (1-based UnitRanges), accumulating
U and J as fmm.direct! + buffer_to_target_system! do.  The ranges are
positions in `pfield.particles` as it is at the call: a tree's
`bodies_index` ranges index the tree's sorted order, so on a field that is
not stored in that order use `plan_direct_list` (which maps them through
`tree.sort_index_list`, as Estr_fmm! does) instead.
"""
function direct_hip!(pfield::vpm.ParticleField, target_ranges, source_ranges; precision=0)
    n = length(target_ranges)
    tb = Int64[first(r) - 1 for r in target_ranges]; te = Int64[last(r) for r in target_ranges]
    sb = Int64[first(r) - 1 for r in source_ranges]; se = Int64[last(r) for r in source_ranges]
    upload(pfield)
    cfg = Ref(config(pfield; precision))
    check(ccall((:vpm_p2p_ranges, libvpm), Cint,
                (Ptr{Cvoid}, Ref{VpmConfig}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64),
                device_field(pfield), cfg, tb, te, sb, se, n))
    download(pfield, Int32(9), Int32(24))
    return nothing
end

"""
    estr_hip!(pfield, target_ranges, source_ranges; precision=0)

The pair loop of Estr_fmm! (src/FLOWVPM_subfilterscale_models.jl:157-188)
over explicit ranges: E_str of every (target range, source range) pair onto
SFS, with no reset and no static skip.  As for `direct_hip!`, the ranges are
positions in `pfield.particles`; Estr_fmm! itself takes a tree and maps its
branches through `sort_index_list` -- that is `estr_fmm_hip!`.
"""
function estr_hip!(pfield::vpm.ParticleField, target_ranges, source_ranges; precision=0)
    n = length(target_ranges)
    tb = Int64[first(r) - 1 for r in target_ranges]; te = Int64[last(r) for r in target_ranges]
    sb = Int64[first(r) - 1 for r in source_ranges]; se = Int64[last(r) for r in source_ranges]
    upload(pfield)
    cfg = Ref(config(pfield; precision))
    check(ccall((:vpm_estr_ranges, libvpm), Cint,
                (Ptr{Cvoid}, Ref{VpmConfig}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64),
                device_field(pfield), cfg, tb, te, sb, se, n))
    download(pfield, Int32(39), Int32(42))
    return nothing
end

"""
    FastMultipole.nearfield_device!(target, target_indices, switch, source, source_indices)

FastMultipole's device near-field hook: `fmm.fmm!` calls it instead of its
CPU `direct!` loop when `nearfield_device=true`, which UJ_fmm sets for
`useGPU > 0` (src/FLOWVPM_UJ.jl:97); the call shape is the reference's own
(src/FLOWVPM_gpu.jl:643).  `target_indices[k]` and `source_indices[k]` are
the k-th leaf pair (UnitRanges of particles); all pairs go to the device in
one `vpm_p2p_ranges` call, which accumulates U and J as `fmm.direct!` +
`buffer_to_target_system!` do (src/FLOWVPM_fmm.jl:102-176).  The ranges are
taken as positions in `target.particles`, the order the field is in when
FastMultipole calls the hook; FastMultipole 2.x's own indexing contract for
this hook is external to FLOWVPM and not restated here.  Target and source
must be the same field (UJ_fmm's case).
"""
function vpm.fmm.nearfield_device!(target::vpm.ParticleField, target_indices, derivatives_switch,
                                   source::vpm.ParticleField, source_indices)
    target === source || error("nearfield_device!: libflowvpm_hip serves a field acting on itself")
    length(target_indices) == length(source_indices) ||
        error("nearfield_device!: $(length(target_indices)) target and $(length(source_indices)) source ranges")
    isempty(target_indices) && return nothing
    direct_hip!(target, target_indices, source_indices)
    return nothing
end

# (target, source) bodies_index ranges of every direct_list entry as 0-based
# half-open positions in the tree's sorted order
function list_ranges(target_tree, source_tree, direct_list)
    m = length(direct_list)
    tb = Vector{Int64}(undef, m); te = similar(tb); sb = similar(tb); se = similar(tb)
    for (k, (i_target, i_source)) in enumerate(direct_list)
        t = target_tree.branches[i_target].bodies_index[1]
        s = source_tree.branches[i_source].bodies_index[1]
        tb[k] = first(t) - 1; te[k] = last(t); sb[k] = first(s) - 1; se[k] = last(s)
    end
    return tb, te, sb, se
end

"""
    DirectListPlan

One FastMultipole direct list of a field, planned once on the device
(`vpm_plan_ranges`) and served to the near-field U/J (`nearfield_plan!`) and
then to E_str (`estr_plan!`), as UJ_fmm hands one `direct_list` to the near
field and then to `Estr_fmm!` (src/FLOWVPM_UJ.jl:90-125).  The particles are
taken in tree order (`tree.sort_index_list[1]`) into a device field of the
plan's own, so the tree's `bodies_index` ranges are used as they are, and
results are scattered back to `pfield.particles` in storage order.
"""
mutable struct DirectListPlan
    field::Ptr{Cvoid}
    plan::Ptr{Cvoid}
    perm::Vector{Int}
    n::Int
end

function plan_direct_list(pfield::vpm.ParticleField, tree, direct_list)
    n = pfield.np
    perm = collect(tree.sort_index_list[1][1:n])
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:vpm_create, libvpm), Cint, (Cint, Int64, Ref{Ptr{Cvoid}}),
                pfield.useGPU > 0 ? pfield.useGPU - 1 : 0, max(n, 1), r))
    p = DirectListPlan(r[], C_NULL, perm, n)
    finalizer(p) do q
        q.plan != C_NULL && ccall((:vpm_plan_destroy, libvpm), Cint, (Ptr{Cvoid},), q.plan)
        q.field != C_NULL && ccall((:vpm_destroy, libvpm), Cint, (Ptr{Cvoid},), q.field)
    end
    Psorted = pfield.particles[:, perm]
    check(ccall((:vpm_set_particles, libvpm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), p.field, Psorted, n))
    tb, te, sb, se = list_ranges(tree, tree, direct_list)
    pr = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:vpm_plan_ranges, libvpm), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64, Ref{Ptr{Cvoid}}),
                p.field, tb, te, sb, se, length(tb), pr))
    p.plan = pr[]
    return p
end

# upload pfield in the plan's tree order, run `entry` on the plan, scatter
# rows [row0, row1) (0-based, half-open) back
function run_plan!(pfield, p::DirectListPlan, entry::Symbol, row0, row1; precision=0)
    pfield.np == p.n || error("the direct list was planned for $(p.n) particles, the field holds $(pfield.np)")
    Psorted = pfield.particles[:, p.perm]
    check(ccall((:vpm_set_particles, libvpm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), p.field, Psorted, p.n))
    cfg = Ref(config(pfield; precision))
    if entry === :p2p
        check(ccall((:vpm_p2p_plan, libvpm), Cint, (Ptr{Cvoid}, Ref{VpmConfig}, Ptr{Cvoid}), p.field, cfg, p.plan))
    else
        check(ccall((:vpm_estr_plan, libvpm), Cint, (Ptr{Cvoid}, Ref{VpmConfig}, Ptr{Cvoid}), p.field, cfg, p.plan))
    end
    check(ccall((:vpm_get_particles, libvpm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64, Int32, Int32),
                p.field, Psorted, p.n, Int32(row0), Int32(row1)))
    pfield.particles[row0+1:row1, p.perm] .= Psorted[row0+1:row1, :]
    return nothing
end

"""
    nearfield_plan!(pfield, plan; precision=0)

The near field of UJ_fmm over the planned direct list: U and J of every
leaf pair accumulated onto `pfield.particles` (fmm.direct! per pair).
"""
nearfield_plan!(pfield::vpm.ParticleField, p::DirectListPlan; precision=0) =
    run_plan!(pfield, p, :p2p, 9, 24; precision)

"""
    estr_plan!(pfield, plan; precision=0)

Estr_fmm! over the planned direct list (src/FLOWVPM_subfilterscale_models.jl:
94-188): E_str of every leaf pair onto SFS, using the field's complete J
(near and far field), with no reset and no static skip.
"""
estr_plan!(pfield::vpm.ParticleField, p::DirectListPlan; precision=0) =
    run_plan!(pfield, p, :estr, 39, 42; precision)

"""
    estr_fmm_hip!(target_pfield, source_pfield, target_tree, source_tree, direct_list; precision=0)

Estr_fmm! (src/FLOWVPM_subfilterscale_models.jl:94-188) on the device, with
the reference's signature: the branches' `bodies_index` ranges are mapped
through `sort_index_list`, as the reference's loop does.  One field acting
on itself (UJ_fmm's call, src/FLOWVPM_UJ.jl:125).
"""
function estr_fmm_hip!(target_pfield::vpm.ParticleField, source_pfield::vpm.ParticleField, target_tree,
                       source_tree, direct_list; precision=0)
    (target_pfield === source_pfield && target_tree === source_tree) ||
        error("estr_fmm_hip!: one field and one tree (UJ_fmm's call)")
    isempty(direct_list) && return nothing
    p = plan_direct_list(target_pfield, target_tree, direct_list)
    try
        estr_plan!(target_pfield, p; precision)
    finally
        finalize(p)
    end
    return nothing
end

"""
    zeta_hip(pfield; precision=0)

zeta_direct(pfield) (src/FLOWVPM_viscous.jl:486-515) on the device: J[1:3]
of every particle <- the basis-function field.  Pass it as CoreSpreading's
`zeta` argument.
"""
function zeta_hip(pfield::vpm.ParticleField; precision=0)
    upload(pfield)
    cfg = Ref(config(pfield; precision))
    check(ccall((:vpm_zeta_direct, libvpm), Cint, (Ptr{Cvoid}, Ref{VpmConfig}), device_field(pfield), cfg))
    download(pfield, Int32(15), Int32(18))
    return nothing
end

"""
    zeta_fmm_hip(pfield; precision=0)

zeta_fmm(pfield) (src/FLOWVPM_viscous.jl:523-558) with its pair loop on the
device: the same FastMultipole tree and direct list; the particles are taken
in tree order (sort_index) into a scratch device field, and for each
(target branch, source branch) of the list the target branch's strengths are
added onto J[1:3] of the source branch's particles, as the reference's loop
does.  J is not reset.  Pass it as CoreSpreading's `zeta` argument.
"""
function zeta_fmm_hip(pfield::vpm.ParticleField; precision=0)
    n = pfield.np
    n == 0 && return nothing
    fmm = pfield.fmm
    FMM = vpm.FastMultipole
    tree = FMM.Tree((pfield,); leaf_size=fmm.ncrit, shrink_recenter=fmm.shrink_recenter)
    _, direct_list = FMM.build_interaction_lists(tree.branches, tree.branches, fmm.ncrit, fmm.theta,
                                                 false, true, true)
    perm = tree.sort_index_list[1]
    Psorted = pfield.particles[:, perm[1:n]]
    tb = Int64[]; te = Int64[]; sb = Int64[]; se = Int64[]
    for (i_target, i_source) in direct_list
        writes = tree.branches[i_source].bodies_index[1]     # particles written (the reference's Pi)
        reads = tree.branches[i_target].bodies_index[1]      # particles read (the reference's Pj)
        push!(tb, first(writes) - 1); push!(te, last(writes))
        push!(sb, first(reads) - 1); push!(se, last(reads))
    end
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:vpm_create, libvpm), Cint, (Cint, Int64, Ref{Ptr{Cvoid}}),
                pfield.useGPU > 0 ? pfield.useGPU - 1 : 0, n, r))
    h = r[]
    try
        cfg = Ref(config(pfield; precision))
        check(ccall((:vpm_set_particles, libvpm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), h, Psorted, n))
        check(ccall((:vpm_zeta_ranges, libvpm), Cint,
                    (Ptr{Cvoid}, Ref{VpmConfig}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Int64),
                    h, cfg, tb, te, sb, se, length(tb)))
        check(ccall((:vpm_get_particles, libvpm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64, Int32, Int32),
                    h, Psorted, n, Int32(15), Int32(18)))
    finally
        ccall((:vpm_destroy, libvpm), Cint, (Ptr{Cvoid},), h)
    end
    pfield.particles[16:18, perm[1:n]] .= Psorted[16:18, :]
    return nothing
end

"""
    rbf_hip(pfield, cs::CoreSpreading; precision=0)

rbf_conjugategradient(pfield, cs) (src/FLOWVPM_viscous.jl:314-478) on the
device, with zeta = zeta_direct; pass it as CoreSpreading's `rbf` keyword.
Errors like the reference when cs.iterror and itmax is reached unconverged.
"""
function rbf_hip(pfield::vpm.ParticleField, cs; precision=0)
    upload(pfield)
    cfg = Ref(config(pfield; precision))
    it = Ref{Int32}(0); res = zeros(3)
    rc = ccall((:vpm_rbf_conjugategradient, libvpm), Cint,
               (Ptr{Cvoid}, Ref{VpmConfig}, Int32, Float64, Int32, Ref{Int32}, Ptr{Float64}),
               device_field(pfield), cfg, Int32(cs.itmax), Float64(cs.tol), Int32(cs.iterror), it, res)
    download(pfield, Int32(0), Int32(46))
    check(rc)
    return nothing
end

"""
    rungekutta3_hip(pfield, dt; relax=false, precision=0, far64=false)

One rungekutta3 step of ReformulatedVPM with UJ = UJ_direct, entirely on the
device (pfield.particles is uploaded once and downloaded once).  With
`precision=1, far64=true` the step's U/J evaluations take the FP64 far form
(VPM_UJ_F64_FAR); everything else runs in FP64 as with `precision=1`.
"""
"""
    sfs_settings(SFS) -> (sfs, force_positive, clip)

The SFS models the device step runs: NoSFS, and DynamicSFS with the
pseudo-3-level procedure (src/FLOWVPM_subfilterscale.jl:447-673), whose
afterUJ half forces C >= 0 only as pseudo3level_positive_afterUJ
(src/FLOWVPM.jl:202), no controls, and either no clipping or
clipping_backscatter alone.  Any other model (ConstantSFS such as
SFS_Cs_nobackscatter, sensor-function procedures, controls) raises.
"""
function sfs_settings(SFS)
    SFS isa vpm.NoSFS && return (false, false, false)
    SFS isa vpm.DynamicSFS ||
        error("rungekutta3_hip runs NoSFS and the pseudo-3-level DynamicSFS only; got $(typeof(SFS))")
    SFS.procedure_beforeUJ === vpm.dynamicprocedure_pseudo3level_beforeUJ ||
        error("rungekutta3_hip: unsupported SFS procedure_beforeUJ $(SFS.procedure_beforeUJ)")
    pa = SFS.procedure_afterUJ
    force_positive = pa === vpm.pseudo3level_positive_afterUJ ? true :
                     pa === vpm.dynamicprocedure_pseudo3level_afterUJ ? false :
                     error("rungekutta3_hip: unsupported SFS procedure_afterUJ $pa")
    isempty(SFS.controls) || error("rungekutta3_hip: SFS controls are not supported")
    clip = if isempty(SFS.clippings)
        false
    elseif length(SFS.clippings) == 1 && SFS.clippings[1] === vpm.clipping_backscatter
        true
    else
        error("rungekutta3_hip: only clippings=(clipping_backscatter,) is supported")
    end
    return (true, force_positive, clip)
end

function rungekutta3_hip(pfield::vpm.ParticleField, dt::Real; relax::Bool=false, precision=0, far64::Bool=false,
                         optargs...)
    sfs, force_positive, clip = sfs_settings(pfield.SFS)
    upload(pfield)
    visc = pfield.viscous isa vpm.CoreSpreading
    rl = pfield.relaxation
    method = rl.relax === vpm.relax_pedrizzetti ? 1 : rl.relax === vpm.relax_correctedpedrizzetti ? 2 : 0
    s = Ref(VpmStepConfig(config(pfield; precision), pfield.formulation.f, pfield.formulation.g,
            Tuple(pfield.Uinf(pfield.t)),
            Int32(sfs), Int32(force_positive), Int32(clip), Int32(method),
            sfs ? pfield.SFS.alpha : 0.0, sfs ? pfield.SFS.rlxf : 0.0,
            sfs ? pfield.SFS.minC : 0.0, sfs ? pfield.SFS.maxC : 1.0, rl.rlxf,
            Int32(visc), Int32(far64 ? VPM_UJ_F64_FAR : 0), visc ? pfield.viscous.nu : 0.0,
            visc ? pfield.viscous.sgm0 : 1.0,
            visc ? pfield.viscous.beta : 1.5, visc ? pfield.viscous.t_sgm : 0.0,
            Int32(visc ? pfield.viscous.itmax : 0), Int32(visc && pfield.viscous.iterror),
            visc ? Float64(pfield.viscous.tol) : 0.0))
    rc = ccall((:vpm_rk3_step, libvpm), Cint, (Ptr{Cvoid}, Ref{VpmStepConfig}, Float64, Int32),
               device_field(pfield), s, dt, Int32(relax))
    visc && (pfield.viscous.t_sgm = s[].t_sgm)
    download(pfield, Int32(0), Int32(46))
    check(rc)
    return nothing
end

"""
    uj_hip(pfield; precision=0) -> (U::Matrix{Float64}(3,np), J::Matrix{Float64}(9,np))

Pure function of (X, Gamma, sigma) for AD; its rrule calls the hand-written
adjoint kernel instead of differentiating through the pair loop (precision 0:
fused FP32 pair kernel with FP64 partial sums, 1: FP64 pairs).
"""
function uj_hip(pfield::vpm.ParticleField; precision=0)
    UJ_hip(pfield; precision)
    return pfield.particles[10:12, 1:pfield.np], pfield.particles[16:24, 1:pfield.np]
end

function ChainRulesCore.rrule(::typeof(uj_hip), pfield::vpm.ParticleField; precision=0)
    U, J = uj_hip(pfield; precision)
    # the forward point: the pullback differentiates at these records even if
    # pfield (host or device) changes between the forward and the pullback
    n = pfield.np
    P0 = pfield.particles[:, 1:n]
    # a cotangent that is structurally zero (a loss on U alone leaves Jbar as
    # ZeroTangent or nothing) becomes zeros; all-zero Jbar takes the library's
    # kernel without the Jbar terms
    dense(x, rows, n) = (x === nothing || x isa ChainRulesCore.AbstractZero) ? zeros(rows, n) :
                        Matrix{Float64}(ChainRulesCore.unthunk(x))
    function uj_hip_pullback((Ubar, Jbar))
        Ub = dense(Ubar, 3, n); Jb = dense(Jbar, 9, n)
        Xb = zeros(3, n); Gb = zeros(3, n); Sb = zeros(n)
        cfg = Ref(config(pfield; precision))
        h = device_field(pfield)
        GC.@preserve P0 check(ccall((:vpm_set_particles, libvpm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64),
                                    h, P0, n))
        check(ccall((:vpm_uj_adjoint, libvpm), Cint,
                    (Ptr{Cvoid}, Ref{VpmConfig}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                    h, cfg, Ub, Jb, Xb, Gb, Sb))
        dparticles = zeros(size(pfield.particles))
        dparticles[1:3, 1:n] .= Xb; dparticles[4:6, 1:n] .= Gb; dparticles[7, 1:n] .= Sb
        return ChainRulesCore.NoTangent(), ChainRulesCore.Tangent{typeof(pfield)}(particles=dparticles)
    end
    return (U, J), uj_hip_pullback
end

"""
    rungekutta3_taped!(pfield, dt; precision=0) -> tape

One rungekutta3 step (ReformulatedVPM, no SFS, inviscid, no relaxation) that
also keeps the three stage states on the device (per-stage checkpoints of X,
Gamma, sigma with their U and J).  The tape feeds `rungekutta3_pullback!`.
"""
mutable struct Tape
    h::Ptr{Cvoid}
    pfield::Any
end

function rungekutta3_taped!(pfield::vpm.ParticleField, dt::Real; precision=0)
    upload(pfield)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:vpm_tape_create, libvpm), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), device_field(pfield), r))
    t = Tape(r[], pfield)
    finalizer(t -> ccall((:vpm_tape_destroy, libvpm), Cint, (Ptr{Cvoid},), t.h), t)
    s = Ref(VpmStepConfig(config(pfield; precision), pfield.formulation.f, pfield.formulation.g,
            Tuple(pfield.Uinf(pfield.t)), Int32(0), Int32(0), Int32(0), Int32(0), 0.0, 0.0, 0.0, 1.0, 0.0,
            Int32(0), 0, 0.0, 1.0, 1.5, 0.0, Int32(0), Int32(0), 0.0))
    check(ccall((:vpm_rk3_step_taped, libvpm), Cint, (Ptr{Cvoid}, Ref{VpmStepConfig}, Float64, Ptr{Cvoid}),
                device_field(pfield), s, dt, t.h))
    download(pfield, Int32(0), Int32(46))
    return t
end

"""
    rungekutta3_pullback!(tape, bar)

bar :: Matrix{Float64}(46, np), the cotangent of pfield.particles after the
taped step (rows X, Gamma, sigma, M); replaced by the cotangent before it.
"""
rungekutta3_pullback!(t::Tape, bar::Matrix{Float64}) =
    GC.@preserve bar check(ccall((:vpm_rk3_step_adjoint, libvpm), Cint, (Ptr{Cvoid}, Ptr{Float64}), t.h, bar))

# rrule of one step on the particle matrix: the forward records the tape, the
# pullback runs the reverse sweep over its three stages on the device.
function rungekutta3_hip_step(pfield::vpm.ParticleField, dt::Real; precision=0)
    rungekutta3_taped!(pfield, dt; precision)
    return copy(pfield.particles)
end

function ChainRulesCore.rrule(::typeof(rungekutta3_hip_step), pfield::vpm.ParticleField, dt::Real; precision=0)
    tape = rungekutta3_taped!(pfield, dt; precision)
    out = copy(pfield.particles)
    function step_pullback(outbar)
        bar = Matrix{Float64}(ChainRulesCore.unthunk(outbar))
        rungekutta3_pullback!(tape, bar)
        return ChainRulesCore.NoTangent(), ChainRulesCore.Tangent{typeof(pfield)}(particles=bar),
               ChainRulesCore.NoTangent()
    end
    return out, step_pullback
end

end # module
