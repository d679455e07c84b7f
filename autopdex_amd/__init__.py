"""autopdex_amd -- MI355X-native assemble-and-solve backend for AutoPDEx.

Drop-in 'hip' solver backend behind solver.solver(dofs, settings,
static_settings): hand-written HIP kernels (gfx950) for the 'user potential'
element loop on Q1 quads and Hex8 bricks, deterministic assembly into the
reduced CSR system, a device-resident Newton loop with Jacobi-PCG, FP32/FP64
mixed precision and a learned preconditioner for flexible CG.

Modules mirror the reference's: mesher, geometry, seeder, utility, spaces,
models, assembler, solver.  The compute library is lib/libautopdex_amd.so
(C ABI: include/autopdex_amd.h), built by autopdex_amd.build.
"""
__version__ = "0.1.0"
