"""rang (MI355X build): rang-core's exact w-gridder path — `rang._core`'s
gridder functions and `rang.core_wgridder` — on hand-written HIP kernels for
gfx950 behind the reference's C ABI (include/rang_core_wgridder.h)."""
