"""Physical constants and dtypes of the FDTD path (values as core/constants.py:6-9 of the reference)."""
import numpy as np

DTYPE = np.float32
C0 = 299792458.0
EPS0 = 8.8541878128e-12
MU0 = 1.2566370614e-6
ETA0 = float(np.sqrt(MU0 / EPS0))
FACES = ('x_lo', 'x_hi', 'y_lo', 'y_hi')
FACES_3D = FACES + ('z_lo', 'z_hi')
