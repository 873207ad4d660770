"""
    SeismicWaves_MI355XChainRulesExt

ChainRulesCore rule for the misfit as a function of the material
properties, J(m) = sum_s chi_s(d_s(m)), on the `:mi355x` backend
(loaded with SWXMI355X and ChainRulesCore).  The pullback scales the batched
adjoint gradient of `run_swgradient!`.  Signs: the reference's returned
acoustic gradients are -dJ/dm (its adjoint is driven by -fact C^-1 (syn -
obs) with the same operator, acou_gradient.jl:45-51), the elastic ones
+dJ/dm (ela_gradient.jl:45-60); both were measured by central differences
of the discrete misfit.
"""
module SeismicWaves_MI355XChainRulesExt

using SeismicWaves, SWXMI355X, ChainRulesCore
using SeismicWaves: ElasticIsoCPMLWaveSimulation, VpAcousticCDMaterialProperties, VpRhoAcousticVDMaterialProperties,
                    ElasticIsoMaterialProperties

const Ext = Base.get_extension(SeismicWaves, :SeismicWaves_MI355XBackendExt)

# material property field names of each gradient key
fieldname_of(::VpAcousticCDMaterialProperties, k) = Dict("vp" => :vp)[k]
fieldname_of(::VpRhoAcousticVDMaterialProperties, k) = Dict("vp" => :vp, "rho" => :rho)[k]
fieldname_of(::ElasticIsoMaterialProperties, k) = Dict("rho" => :ρ, "lambda" => :λ, "mu" => :μ)[k]

function ChainRulesCore.rrule(::typeof(SeismicWaves.run_swmisfit!), ws::Ext.SwxSim, matprop, shots, misfits)
    compute = ws.gradparams.compute_misfit
    compute || error("the rrule of run_swmisfit! needs GradParameters(compute_misfit=true)")
    g, J = SeismicWaves.run_swgradient!(ws, matprop, shots, misfits)
    sgn = ws isa ElasticIsoCPMLWaveSimulation ? 1 : -1
    function run_swmisfit_pullback(ΔJ)
        parts = (fieldname_of(matprop, k) => sgn .* unthunk(ΔJ) .* v for (k, v) in g)
        Δm = Tangent{typeof(matprop)}(; parts...)
        return NoTangent(), NoTangent(), Δm, NoTangent(), NoTangent()
    end
    return J, run_swmisfit_pullback
end

end # module
