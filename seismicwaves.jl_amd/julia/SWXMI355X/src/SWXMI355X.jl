"""
    SWXMI355X

Locates `libswx.so`, the MI355X (gfx950) backend of SeismicWaves.jl's
finite-difference forward/adjoint path.  Loading this package next to
SeismicWaves activates the extension `SeismicWaves_MI355XBackendExt`, which
adds the backend symbol `parall = :mi355x`.

The library path is taken from the environment variable `SWX_LIBRARY`, or
else from `libswx.so` of the Python package beside this one
(`seismicwaves.jl_amd/seismicwaves/libswx.so`).
"""
module SWXMI355X

const libswx = let p = get(ENV, "SWX_LIBRARY", "")
    isempty(p) ? normpath(joinpath(@__DIR__, "..", "..", "..", "seismicwaves", "libswx.so")) : p
end

"Version of the C ABI the library was built with (`SWX_ABI_VERSION` of include/swx.h)."
abi_version() = ccall((:swx_abi_version, libswx), Cint, ())

end # module
