"""diver_amd: DIVeR's render-and-train hot path on AMD Instinct MI355X (gfx950).

Layout:
- ``csrc/``     HIP kernels and the C ABI (built into ``libdiver.so``; see include/diver.h)
- ``calls``     tensor-level wrappers of the C ABI (device tensors, torch's current stream)
- ``decoder``   canonical decoder parameters and their device handle
- ``renderer``  fused render path and its autograd adjoint
- ``server``    continuously batched multi-request renderer
- ``utils/``, ``model/``  the reference's module surface (utils.ray_voxel_intersection,
  utils.ray_march, utils.mlp_evaluation{,64}, utils.integrator, model.diver, model.mlps)
- ``synthetic`` seeded scenes, decoders and cameras used by tests and benchmarks
"""
__version__ = '0.1.0'
