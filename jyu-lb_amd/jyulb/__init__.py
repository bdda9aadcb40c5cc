"""jyulb: MI355X-native drop-in for the evolve path of the JYU-LB isothermal
D3Q19 lattice-Boltzmann solver.

Same module layout and public names as the reference package
(jyulb/__init__.py:1): isothermal (Solver), D3Q19 (velocity set and moment
constants), defs (enumerations) and flow_field (analytic profiles). The
reference's bcs module has no Python-visible surface (bcs.pxd declares cdef
functions only); its Zou-He faces are device code here (csrc/lb_physics.h).
Additions: adjoint (gradients through evolve), geometry (synthetic porous
media) and raw_io (the raw-format lattice files of the reference's C++
front-end). Solver.evolve runs hand-written HIP kernels
for gfx950 through the C ABI of libjyulb_hip.so (include/jyulb_gpu.h).
"""
__all__ = ['isothermal', 'D3Q19', 'defs', 'flow_field', 'adjoint', 'geometry', 'raw_io']
__version__ = '0.2.0'
