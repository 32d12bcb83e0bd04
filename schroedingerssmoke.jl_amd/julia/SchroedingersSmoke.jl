# SchroedingersSmoke.jl -- the libisf binding under the reference's package name, so that
# `using SchroedingersSmoke` in the reference drivers (test/runtests.jl:1,
# examples/benchmark.jl:1) loads it; with `ArrayType = ISFArray` the drivers then run on
# the MI355X path.  Everything is defined in SchroedingersSmokeAMD.jl and re-exported.
module SchroedingersSmoke

include("SchroedingersSmokeAMD.jl")
using .SchroedingersSmokeAMD

for name in names(SchroedingersSmokeAMD)
    name === :SchroedingersSmokeAMD && continue
    @eval const $name = SchroedingersSmokeAMD.$name
    @eval export $name
end
const MatlabPort = SchroedingersSmokeAMD.MatlabPort

end # module
