"""Zou-He fixed-density / fixed-velocity boundary conditions for D3Q19.

In the reference (jyulb/bcs.pyx:31-716) these are cdef functions selected per
face by the Solver (isothermal.pyx:431-471) and called per boundary site.
Here they are device functions of the fused step kernel
(csrc/lb_physics.h: zh_fix_den, zh_fix_vel, assign_face), selected by the
face's bc id. This module keeps the naming of the per-face procedures and the
populations each one assigns, as data.
"""
from . import D3Q19 as _q
from .defs import FACE_1X, FACE_X1, FACE_1Y, FACE_Y1, FACE_1Z, FACE_Z1, FIXED_DEN, FIXED_VEL

# populations assigned at each face (the 5 unknowns entering the domain)
ASSIGNED = {
    FACE_1X: (_q.L_I1_1J_K0, _q.L_I1_J0_1K, _q.L_I1_J0_K0, _q.L_I1_J0_K1, _q.L_I1_J1_K0),
    FACE_X1: (_q.L_1I_1J_K0, _q.L_1I_J0_1K, _q.L_1I_J0_K0, _q.L_1I_J0_K1, _q.L_1I_J1_K0),
    FACE_1Y: (_q.L_1I_J1_K0, _q.L_I0_J1_1K, _q.L_I0_J1_K0, _q.L_I0_J1_K1, _q.L_I1_J1_K0),
    FACE_Y1: (_q.L_1I_1J_K0, _q.L_I0_1J_1K, _q.L_I0_1J_K0, _q.L_I0_1J_K1, _q.L_I1_1J_K0),
    FACE_1Z: (_q.L_1I_J0_K1, _q.L_I0_1J_K1, _q.L_I0_J0_K1, _q.L_I0_J1_K1, _q.L_I1_J0_K1),
    FACE_Z1: (_q.L_1I_J0_1K, _q.L_I0_1J_1K, _q.L_I0_J0_1K, _q.L_I0_J1_1K, _q.L_I1_J0_1K),
}

# names of the reference's per-face procedures, by (bc id, face)
PROCEDURES = {
    (FIXED_DEN, FACE_1X): 'fix_den_face_1x_D3Q19', (FIXED_DEN, FACE_X1): 'fix_den_face_x1_D3Q19',
    (FIXED_DEN, FACE_1Y): 'fix_den_face_1y_D3Q19', (FIXED_DEN, FACE_Y1): 'fix_den_face_y1_D3Q19',
    (FIXED_DEN, FACE_1Z): 'fix_den_face_1z_D3Q19', (FIXED_DEN, FACE_Z1): 'fix_den_face_z1_D3Q19',
    (FIXED_VEL, FACE_1X): 'fix_vel_face_1x_D3Q19', (FIXED_VEL, FACE_X1): 'fix_vel_face_x1_D3Q19',
    (FIXED_VEL, FACE_1Y): 'fix_vel_face_1y_D3Q19', (FIXED_VEL, FACE_Y1): 'fix_vel_face_y1_D3Q19',
    (FIXED_VEL, FACE_1Z): 'fix_vel_face_1z_D3Q19', (FIXED_VEL, FACE_Z1): 'fix_vel_face_z1_D3Q19',
}
