"""delfem2_amd -- DelFEM2's implicit tet-FEM solve path on MI355X (gfx950).

tet assembly -> 3x3 block-CSR CMatrixSparse -> CPreconditionerILU ILU(0)
-> Solve_PCG, as hand-written HIP kernels behind a C ABI
(include/dfm2_hip.h, built into ``libdfm2_hip.so``).

Submodules:
* ``_native``  -- ctypes bindings of the C ABI (loads without a GPU);
* ``solver``   -- the reference-shaped Python API (needs torch + a GPU);
* ``workload`` -- synthetic meshes, orderings, seeded materials and loads;
* ``problem``  -- end-to-end drivers (assemble, factor, solve) used by the
  benchmark and the tests;
* ``build``    -- compiles the library for gfx950.
"""

__all__ = ["_native", "workload"]
