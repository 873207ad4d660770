"""
    SeismicWaves_MI355XBackendExt

Package extension of SeismicWaves.jl (loaded with `SWXMI355X`) adding the
backend `parall = :mi355x`: the finite-difference forward and adjoint
gradient of a batch of shots run in libswx.so (HIP kernels for gfx950)
through the C ABI of include/swx.h.

It plugs in where the reference's GPU extensions do
(ext/SeismicWaves_AMDGPUBackendExt.jl:15-26: `select_backend` methods for
`Val{parall}`), with one difference: the reference drives its backend once
per time step and shot (run_swforward!, forward.jl:71-115;
run_swgradient!, gradient.jl:93-136), libswx runs the whole time loop of a
batch of shots per call.  The backend module therefore gives the wave
simulation its own array type (`MI355X.SWXArray`, host memory), and
`run_swforward!`, `run_swmisfit!` and `run_swgradient!` get methods for wave
simulations of that type.  Everything on the host side is the reference's
own code: checks (init_shot!), C-PML coefficients (init_bdc!), source
scaling and sinc spreading (possrcrec_scaletf), material factors
(precompute_fact!, precomp_elaprop!), misfits (calcmisfit, ∂χ_∂u) and the
gradient post-processing (back_interp, mutearoundmultiplepoints!, the chain
rules of acou_gradient.jl:84-95, :176-202 and ela_gradient.jl:155-185).

`ChainRulesCore.rrule` of `swmisfit!` lives in
`SeismicWaves_MI355XChainRulesExt` (file SeismicWaves_MI355XChainRulesExt.jl).
"""
module SeismicWaves_MI355XBackendExt

using SeismicWaves, SWXMI355X
using SeismicWaves: CPMLBoundaryCondition, LocalGrid, WaveSimulation, MaterialProperties, Shot, ScalarShot,
                    MomentTensorShot, ExternalForceShot, AbstractMisfit, L2Misfit, AcousticCDCPMLWaveSimulation,
                    AcousticVDStaggeredCPMLWaveSimulation, ElasticIsoCPMLWaveSimulation

const lib = SWXMI355X.libswx

# ---------------------------------------------------------------------------
# C structs of include/swx.h (same field order and types)
# ---------------------------------------------------------------------------
struct SwxPlanDesc
    physics::Int32
    dtype::Int32
    ndim::Int32
    halo::Int32
    n::NTuple{3, Int64}
    h::NTuple{3, Float64}
    dt::Float64
    freetop::Int32
    max_shots::Int32
    nt_max::Int32
    check_freq::Int32
    gradient::Int32
    device::Int32
    mem_budget::Float64
end

struct SwxShotDesc
    nt::Int32
    nsrc::Int32
    src_idx::Ptr{Int64}
    srctf::Ptr{Cvoid}
    nrec::Int32
    rec_idx::Ptr{Int64}
    cpml::NTuple{12, Ptr{Cvoid}}      # [axis][a, a_h, b, b_h], axis-major
end

struct SwxTaps
    ptr::Ptr{Int32}
    ij::Ptr{Int64}
    coef::Ptr{Cvoid}
end

struct SwxElasticShotDesc
    nt::Int32
    source_kind::Int32
    nsrc::Int32
    src::NTuple{2, SwxTaps}
    moments::Ptr{Cvoid}
    srctf::Ptr{Cvoid}
    nrec::Int32
    rec::NTuple{2, SwxTaps}
    cpml::NTuple{12, Ptr{Cvoid}}
end

const SWX_OK, SWX_EINVAL, SWX_ENOMEM = 0, 1, 2
const SWX_ACOUSTIC_CD, SWX_ACOUSTIC_VD, SWX_ELASTIC_ISO = 1, 2, 3
const SWX_F32, SWX_F64 = 1, 2
const SWX_SRC_MOMENT_TENSOR, SWX_SRC_EXTERNAL_FORCE = 1, 2

function check(rc::Cint)
    rc == SWX_OK && return nothing
    msg = unsafe_string(ccall((:swx_last_error, lib), Cstring, ()))
    rc == SWX_EINVAL && throw(AssertionError(msg))       # the reference's @assert failures
    rc == SWX_ENOMEM && throw(OutOfMemoryError())
    error("libswx error $rc: $msg")
end

# ---------------------------------------------------------------------------
# Backend module: host arrays of a type of their own, so that a wave
# simulation built with parall = :mi355x has a distinct type parameter A
# (acou_models.jl:112-113 takes A = backend.Data.Array{T, N}).
# ---------------------------------------------------------------------------
module MI355X
    struct SWXArray{T, N} <: DenseArray{T, N}
        a::Array{T, N}
    end
    SWXArray(a::Array{T, N}) where {T, N} = SWXArray{T, N}(a)
    SWXArray{T, N}(a::AbstractArray) where {T, N} = SWXArray{T, N}(Array{T, N}(a))
    Base.size(x::SWXArray) = size(x.a)
    Base.IndexStyle(::Type{<:SWXArray}) = IndexLinear()
    Base.@propagate_inbounds Base.getindex(x::SWXArray, i::Int) = x.a[i]
    Base.@propagate_inbounds Base.setindex!(x::SWXArray, v, i::Int) = (x.a[i] = v)
    Base.similar(x::SWXArray, ::Type{S}, dims::Dims) where {S} = SWXArray(similar(x.a, S, dims))
    Base.unsafe_convert(::Type{Ptr{T}}, x::SWXArray{T}) where {T} = Base.unsafe_convert(Ptr{T}, x.a)
    Base.strides(x::SWXArray) = strides(x.a)
    Base.elsize(::Type{SWXArray{T, N}}) where {T, N} = sizeof(T)
    Base.Array(x::SWXArray) = copy(x.a)
    module Data
        import ..SWXArray
        const Array = SWXArray
    end
    zeros(T, dims...) = SWXArray(Base.zeros(T, dims...))
    ones(T, dims...) = SWXArray(Base.ones(T, dims...))
end

for W in (AcousticCDCPMLWaveSimulation, AcousticVDStaggeredCPMLWaveSimulation, ElasticIsoCPMLWaveSimulation)
    @eval SeismicWaves.select_backend(::CPMLBoundaryCondition, ::LocalGrid, ::Type{<:$W}, ::Type{Val{:mi355x}}) = MI355X
end

const SwxAcousticCD{T, N} = AcousticCDCPMLWaveSimulation{T, N, <:MI355X.SWXArray}
const SwxAcousticVD{T, N} = AcousticVDStaggeredCPMLWaveSimulation{T, N, <:MI355X.SWXArray}
const SwxElastic{T, N} = ElasticIsoCPMLWaveSimulation{T, N, <:MI355X.SWXArray}
const SwxSim{T, N} = Union{SwxAcousticCD{T, N}, SwxAcousticVD{T, N}, SwxElastic{T, N}}

host(x::MI355X.SWXArray) = x.a
host(x::Array) = x

# ---------------------------------------------------------------------------
# Plans: one per wave simulation and batch size, destroyed with the wave
# simulation (finalizer) or when a larger batch needs a new one.
# ---------------------------------------------------------------------------
mutable struct Plan
    ptr::Ptr{Cvoid}
    nshots::Int
end

const PLANS = WeakKeyDict{Any, Plan}()

physics(::SwxAcousticCD) = SWX_ACOUSTIC_CD
physics(::SwxAcousticVD) = SWX_ACOUSTIC_VD
physics(::SwxElastic) = SWX_ELASTIC_ISO

function plan_desc(ws::SwxSim{T, N}, nshots::Int) where {T, N}
    n = ntuple(k -> k <= N ? Int64(ws.params.gridsize[k]) : Int64(1), 3)
    h = ntuple(k -> k <= N ? Float64(ws.params.gridspacing[k]) : 1.0, 3)
    gp = ws.gradparams
    return SwxPlanDesc(physics(ws), T === Float32 ? SWX_F32 : SWX_F64, N, ws.cpmlparams.halo, n, h, Float64(ws.dt),
                       ws.cpmlparams.freeboundtop, nshots, ws.nt, gp === nothing ? 1 : gp.check_freq,
                       gp !== nothing, 0, 0.0)
end

"swx_plan_create for `nshots` shots of this wave simulation (the reference's constructors allocate per shot)."
function plan_for(ws::SwxSim, nshots::Int)
    p = get(PLANS, ws, nothing)
    p !== nothing && p.nshots >= nshots && return p
    p === nothing || plan_destroy!(p)
    ref = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:swx_plan_create, lib), Cint, (Ref{SwxPlanDesc}, Ref{Ptr{Cvoid}}), plan_desc(ws, nshots), ref))
    p = Plan(ref[], nshots)
    PLANS[ws] = p
    finalizer(plan_destroy!, p)
    return p
end

function plan_destroy!(p::Plan)
    p.ptr == C_NULL && return nothing
    ccall((:swx_plan_destroy, lib), Cint, (Ptr{Cvoid},), p.ptr)
    p.ptr = C_NULL
    return nothing
end

# ---------------------------------------------------------------------------
# Model fields: the reference's own precomputed factors (precompute_fact!,
# acou_models.jl:60 and :292-300; precomp_elaprop!, ela_models.jl:156-173)
# ---------------------------------------------------------------------------
model_fields(ws::SwxAcousticCD) = [host(ws.grid.fields["fact"].value)]
model_fields(ws::SwxAcousticVD) =
    [host(ws.grid.fields["fact_m0"].value); [host(v) for v in ws.grid.fields["fact_m1_stag"].value]]
model_fields(ws::SwxElastic) = [host(ws.grid.fields[k].value) for k in ("λ", "μ", "ρ_ihalf", "ρ_jhalf", "μ_ihalf_jhalf")]

function set_model!(p::Plan, ws::SwxSim)
    fields = model_fields(ws)
    ptrs = [Ptr{Cvoid}(pointer(f)) for f in fields]
    GC.@preserve fields check(ccall((:swx_set_model, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}),
                                    p.ptr, length(fields), ptrs))
end

# C-PML coefficients of the shot just initialised (init_bdc! fills
# ws.cpmlcoeffs per shot, acou_init_bc.jl:7-40): copies, since the next
# shot's init_shot! overwrites them
function cpml_copies(ws::SwxSim{T, N}, keep) where {T, N}
    ptrs = Ptr{Cvoid}[]
    for ax in 1:3, v in (:a, :a_h, :b, :b_h)
        if ax <= N
            c = collect(T, host(getfield(ws.cpmlcoeffs[ax], v)))
            push!(keep, c)
            push!(ptrs, isempty(c) ? C_NULL : Ptr{Cvoid}(pointer(c)))
        else
            push!(ptrs, C_NULL)
        end
    end
    return ntuple(k -> ptrs[k], 12)
end

"Shot descriptors of a batch (init_shot! and possrcrec_scaletf per shot, acou_forward.jl:6-20, :67-81)."
function shot_descs(ws::Union{SwxAcousticCD{T}, SwxAcousticVD{T}}, shots) where {T}
    keep = Any[]
    descs = SwxShotDesc[]
    for shot in shots
        SeismicWaves.init_shot!(ws, shot)
        possrcs, posrecs, scal_srctf = SeismicWaves.possrcrec_scaletf(ws, shot)
        si = Matrix{Int64}(possrcs)
        ri = Matrix{Int64}(posrecs)
        tf = Matrix{T}(scal_srctf)
        cp = cpml_copies(ws, keep)
        push!(keep, si, ri, tf)
        push!(descs, SwxShotDesc(ws.nt, size(si, 1), pointer(si), Ptr{Cvoid}(pointer(tf)), size(ri, 1), pointer(ri), cp))
    end
    return descs, keep
end

# one family of spread points (srccoeij_*/srccoeval_* of ela_models.jl:6-90):
# point p owns rows ptr[p]+1 .. ptr[p+1] of the (i, j) tap list
function taps(ij::Vector, coef::Vector, ::Type{T}, keep) where {T}
    ptr = Int32[0]
    for c in coef
        push!(ptr, ptr[end] + length(c))
    end
    ijs = Int64[]
    for m in ij
        for r in axes(m, 1)
            push!(ijs, m[r, 1], m[r, 2])
        end
    end
    cs = T[c for v in coef for c in v]
    push!(keep, ptr, ijs, cs)
    return SwxTaps(pointer(ptr), pointer(ijs), Ptr{Cvoid}(pointer(cs)))
end

function shot_descs(ws::SwxElastic{T}, shots) where {T}
    keep = Any[]
    descs = SwxElasticShotDesc[]
    for shot in shots
        SeismicWaves.init_shot!(ws, shot)
        s1ij, s1v, s2ij, s2v, rxij, rxv, rzij, rzv, scal_srctf =
            SeismicWaves.possrcrec_scaletf(ws, shot; sincinterp=ws.sincinterp)
        mt = shot isa MomentTensorShot
        src = (taps(s1ij, s1v, T, keep), taps(s2ij, s2v, T, keep))
        rec = (taps(rxij, rxv, T, keep), taps(rzij, rzv, T, keep))
        tf = Array{T}(scal_srctf)
        push!(keep, tf)
        mom = C_NULL
        if mt
            M = T[getfield(m, f) for m in shot.srcs.momtens for f in (:Mxx, :Mzz, :Mxz)]
            push!(keep, M)
            mom = Ptr{Cvoid}(pointer(M))
        end
        cp = cpml_copies(ws, keep)
        push!(descs, SwxElasticShotDesc(ws.nt, mt ? SWX_SRC_MOMENT_TENSOR : SWX_SRC_EXTERNAL_FORCE,
                                        size(shot.srcs.positions, 1), src, mom, Ptr{Cvoid}(pointer(tf)),
                                        size(shot.recs.positions, 1), rec, cp))
    end
    return descs, keep
end

set_shots!(p::Plan, ws::Union{SwxAcousticCD, SwxAcousticVD}, d::Vector{SwxShotDesc}) =
    check(ccall((:swx_set_shots, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{SwxShotDesc}), p.ptr, length(d), d))
set_shots!(p::Plan, ws::SwxElastic, d::Vector{SwxElasticShotDesc}) =
    check(ccall((:swx_set_shots_elastic, lib), Cint, (Ptr{Cvoid}, Cint, Ptr{SwxElasticShotDesc}),
                p.ptr, length(d), d))

forward!(p::Plan, traces, save::Bool) =
    check(ccall((:swx_forward, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), p.ptr, traces, save))
adjoint!(p::Plan, adjsrc, grad, per_shot::Bool) =
    check(ccall((:swx_adjoint, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint), p.ptr, adjsrc, grad, per_shot))
set_observed!(p::Plan, obs, w) =
    check(ccall((:swx_set_observed, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cdouble}), p.ptr, obs,
                w === nothing ? C_NULL : w))
adjoint_l2!(p::Plan, grad, per_shot::Bool, misfit::Vector{Float64}) =
    check(ccall((:swx_adjoint_l2, lib), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cdouble}), p.ptr, grad, per_shot,
                misfit))

"Traces of a batch (swx_forward's layout) back into each shot's seismograms (copyto!, acou_forward.jl:61)."
function scatter!(shots, traces::Vector{T}) where {T}
    off = 0
    for shot in shots
        n = length(shot.recs.seismograms)
        copyto!(shot.recs.seismograms, reshape(view(traces, off+1:off+n), size(shot.recs.seismograms)))
        off += n
    end
end

# sizes of the raw gradient fields swx_adjoint returns (include/swx.h)
function grad_shapes(ws::SwxSim{T, N}) where {T, N}
    n = ws.params.gridsize
    ws isa SwxAcousticCD && return [n]
    ws isa SwxAcousticVD && return [n; [ntuple(k -> n[k] - (k == d ? 1 : 0), N) for d in 1:N]]
    return [(n[1] - 1, n[2]), (n[1], n[2] - 1), n, n, (n[1] - 1, n[2] - 1)]
end
raw_length(ws::SwxSim) = sum(prod, grad_shapes(ws))

function split_raw(ws::SwxSim{T}, raw::Vector{T}, k::Int) where {T}
    out = Array{T}[]
    off = (k - 1) * raw_length(ws)
    for sh in grad_shapes(ws)
        n = prod(sh)
        push!(out, reshape(raw[off+1:off+n], sh))
        off += n
    end
    return out
end

# ---------------------------------------------------------------------------
# Gradient post-processing of one shot (or of the batch sum when nothing is
# muted): the reference's own helpers
# ---------------------------------------------------------------------------
function mute!(ws::SwxSim, g, shot)
    SeismicWaves.mutearoundmultiplepoints!(g, shot.srcs.positions, ws.grid, ws.gradparams.mute_radius_src)
    SeismicWaves.mutearoundmultiplepoints!(g, shot.recs.positions, ws.grid, ws.gradparams.mute_radius_rec)
    return g
end

function finish(ws::SwxAcousticCD{T}, f, shot) where {T}          # acou_gradient.jl:84-94
    g = f[1]
    shot === nothing || mute!(ws, g, shot)
    return Dict("vp" => (convert(T, 2.0) ./ (ws.matprop.vp .^ 3)) .* g)
end

function finish(ws::SwxAcousticVD{T, N}, f, shot) where {T, N}     # acou_gradient.jl:176-202
    g_m0 = f[1]
    g_m1 = zeros(T, ws.grid.size...)
    for i in 1:N
        g_m1 .+= SeismicWaves.back_interp(ws.matprop.interp_method, 1 ./ ws.matprop.rho, f[1+i], i)
    end
    if shot !== nothing
        mute!(ws, g_m0, shot)
        mute!(ws, g_m1, shot)
    end
    return Dict("vp" => .-convert(T, 2.0) .* g_m0 ./ (ws.matprop.vp .^ 3 .* ws.matprop.rho),
                "rho" => .-g_m0 ./ (ws.matprop.vp .^ 2 .* ws.matprop.rho .^ 2) .- g_m1 ./ ws.matprop.rho)
end

function finish(ws::SwxElastic{T}, f, shot) where {T}              # ela_gradient.jl:155-185
    g_rih, g_rjh, g_λ, g_μ, g_μhh = f
    mp = ws.matprop
    g_ρ = SeismicWaves.back_interp(mp.interp_method_ρ, mp.ρ, g_rih, 1) .+
          SeismicWaves.back_interp(mp.interp_method_ρ, mp.ρ, g_rjh, 2)
    g_μ = g_μ .+ SeismicWaves.back_interp(mp.interp_method_μ, mp.μ, g_μhh, [1, 2])
    if shot !== nothing
        foreach(g -> mute!(ws, g, shot), (g_ρ, g_λ, g_μ))
    end
    return Dict("rho" => g_ρ, "lambda" => g_λ, "mu" => g_μ)
end

# L2 misfits the device evaluates: identity or diagonal inverse covariance,
# any windows (L2Misfit.jl:24-95)
function device_l2(m)
    m isa L2Misfit || return false
    C = m.invcov
    return isempty(C) || isdiag(C)
end
isdiag(C) = all(C[i, j] == 0 for j in axes(C, 2), i in axes(C, 1) if i != j)

# observed data in the traces layout, and float64 weights mask x diag(C^-1)
# per sample (nothing when all are 1)
function l2_observed(misfits, shots, ::Type{T}) where {T}
    obs = T[]
    weighted = any(m -> !isempty(m.windows) || !isempty(m.invcov), misfits)
    w = Float64[]
    for (m, shot) in zip(misfits, shots)
        o = m.observed
        @assert size(o) == size(shot.recs.seismograms) "observed data and seismograms must have the same shape"
        append!(obs, vec(T.(o)))
        if weighted
            nt = size(o, 1)
            mask = isempty(m.windows) ? ones(nt) : zeros(nt)
            for wnd in m.windows
                mask[wnd.first:wnd.second] .= 1
            end
            c = isempty(m.invcov) ? ones(nt) : [Float64(m.invcov[i, i]) for i in 1:nt]
            wt = mask .* c
            append!(w, vec(repeat(wt, 1, div(length(o), nt))))
        end
    end
    return obs, weighted ? w : nothing
end

# ---------------------------------------------------------------------------
# Batched drivers
# ---------------------------------------------------------------------------
"run_swforward! (forward.jl:71-115) for the whole batch of shots in one plan."
function SeismicWaves.run_swforward!(ws::SwxSim{T, N}, matprop::MaterialProperties{T, N},
                                     shots::Vector{<:Shot{T}}) where {T, N}
    SeismicWaves.check_sim_consistency(ws, matprop, shots)
    SeismicWaves.set_wavesim_matprop!(ws, matprop)
    p = plan_for(ws, length(shots))
    set_model!(p, ws)
    descs, keep = shot_descs(ws, shots)
    traces = Vector{T}(undef, sum(length(s.recs.seismograms) for s in shots))
    GC.@preserve keep begin
        set_shots!(p, ws, descs)
        forward!(p, traces, false)
    end
    scatter!(shots, traces)
    return nothing
end

"run_swmisfit! (misfit.jl:85-104), summed over every shot (the reference's loop counts shot 1 only, misfit.jl:97-99)."
function SeismicWaves.run_swmisfit!(ws::SwxSim{T, N}, matprop::MaterialProperties{T, N}, shots::Vector{<:Shot{T}},
                                    misfit::Vector{<:AbstractMisfit{T}}) where {T, N}
    SeismicWaves.run_swforward!(ws, matprop, shots)
    return sum(SeismicWaves.calcmisfit(m, s.recs) for (m, s) in zip(misfit, shots))
end

"run_swgradient! (gradient.jl:93-136) for the whole batch: forward with device checkpoints, adjoint source, adjoint."
function SeismicWaves.run_swgradient!(ws::SwxSim{T, N}, matprop::MaterialProperties{T, N}, shots::Vector{<:Shot{T}},
                                      misfit::Vector{<:AbstractMisfit{T}}) where {T, N}
    SeismicWaves.check_sim_consistency(ws, matprop, shots)
    SeismicWaves.set_wavesim_matprop!(ws, matprop)
    gp = ws.gradparams
    nsh = length(shots)
    p = plan_for(ws, nsh)
    set_model!(p, ws)
    descs, keep = shot_descs(ws, shots)
    per_shot = gp.mute_radius_src > 0 || gp.mute_radius_rec > 0
    traces = Vector{T}(undef, sum(length(s.recs.seismograms) for s in shots))
    raw = Vector{T}(undef, (per_shot ? nsh : 1) * raw_length(ws))
    totmisfit = zero(T)
    GC.@preserve keep begin
        set_shots!(p, ws, descs)
        forward!(p, traces, true)
        scatter!(shots, traces)
        if all(device_l2, misfit)
            # adjoint source and misfit on the device (acou_gradient.jl:45-51)
            obs, w = l2_observed(misfit, shots, T)
            mis = zeros(Float64, nsh)
            GC.@preserve obs w begin
                set_observed!(p, obs, w)
                adjoint_l2!(p, raw, per_shot, mis)
            end
            totmisfit = convert(T, sum(mis))
        else
            adj = reduce(vcat, [vec(.-SeismicWaves.∂χ_∂u(m, s.recs)) for (m, s) in zip(misfit, shots)])
            adj = Vector{T}(adj)
            adjoint!(p, adj, raw, per_shot)
            if gp.compute_misfit
                totmisfit = sum(SeismicWaves.calcmisfit(m, s.recs) for (m, s) in zip(misfit, shots))
            end
        end
    end
    if per_shot
        totgrad = nothing
        for (k, shot) in enumerate(shots)
            g = finish(ws, split_raw(ws, raw, k), shot)
            totgrad === nothing ? (totgrad = g) : foreach(key -> totgrad[key] .+= g[key], keys(g))
        end
    else
        totgrad = finish(ws, split_raw(ws, raw, 1), nothing)
    end
    return gp.compute_misfit ? (totgrad, totmisfit) : totgrad
end

end # module
