"""Enumerations of coordinate axes, faces, edges, corners, phases, flow
regimes, collision operators and boundary conditions.

Values are identical to the reference's cpdef enums (jyulb/defs.pxd:17-103),
so scripts written against the reference work unchanged.
"""
# Cartesian axes (defs.pxd:17-20)
X = 0
Y = 1
Z = 2

# Cuboid faces: 1X has outward normal (-1,0,0), X1 has (1,0,0), ... (defs.pxd:27-33)
FACE_1X = 0
FACE_X1 = 1
FACE_1Y = 2
FACE_Y1 = 3
FACE_1Z = 4
FACE_Z1 = 5

# Cuboid edges (defs.pxd:39-51)
EDGE_1X_1Y = 0
EDGE_1X_Y1 = 1
EDGE_X1_1Y = 2
EDGE_X1_Y1 = 3
EDGE_1X_1Z = 4
EDGE_1X_Z1 = 5
EDGE_X1_1Z = 6
EDGE_X1_Z1 = 7
EDGE_1Y_1Z = 8
EDGE_1Y_Z1 = 9
EDGE_Y1_1Z = 10
EDGE_Y1_Z1 = 11

# Cuboid corners (defs.pxd:58-66)
CORNER_1X_1Y_1Z = 0
CORNER_1X_1Y_Z1 = 1
CORNER_1X_Y1_1Z = 2
CORNER_1X_Y1_Z1 = 3
CORNER_X1_1Y_1Z = 4
CORNER_X1_1Y_Z1 = 5
CORNER_X1_Y1_1Z = 6
CORNER_X1_Y1_Z1 = 7

# Phases (defs.pxd:71-75); any nonzero phase is treated as fluid
SOLID = 0
FLUID = 1
LIQUID = 2
VAPOUR = 3

# Flow regimes (defs.pxd:80-83)
STOKES = 0
LAMINAR = 1
TURBULENT = 2

# Collision operators (defs.pxd:88-92)
BGK = 0
TRT = 1
REGULARIZATION = 2
RECURRENCE_REL = 3

# Boundary conditions (defs.pxd:97-102)
WALL = 0       # immobile and adiabatic
PERIODIC = 1
FIXED_DEN = 2  # defines implicitly the boundary normal velocity
FIXED_VEL = 3  # defines implicitly the boundary density
SYMMETRIC = 4  # specular reflection
