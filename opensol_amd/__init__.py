"""opensol_amd: the FDTD (Yee) hot path of OpenSol rebuilt for AMD MI355X (gfx950).

2D E_z / H_z and 3D Yee steppers with CPML, calibrated sheet / dipole sources and FP64 Kahan running-DFT flux monitors,
stepped by hand-written HIP kernels in libyee.so (C ABI, include/yee.h) and driven from Python through ctypes.
PyTorch-ROCm provides device memory and streams only. There is no CPU fallback: stepping without the built library
or without a GPU raises.
"""
from .constants import C0, DTYPE, EPS0, ETA0, MU0
from .waveforms import ContinuousWave, GaussianPulse, RickerWavelet
from .boundaries import BoundaryConfig
from .dispersion import DrudePole, LorentzPole
from .monitors import DFTMonitor, DFTPlane3D, FieldProbe, FluxMonitor, FluxPlane
from .sources import LineSource, PointDipole
from .simulation import Simulation, Simulation3D

__all__ = ['C0', 'DTYPE', 'EPS0', 'ETA0', 'MU0', 'ContinuousWave', 'GaussianPulse', 'RickerWavelet',
           'BoundaryConfig', 'DrudePole', 'LorentzPole', 'DFTMonitor', 'DFTPlane3D', 'FieldProbe', 'FluxMonitor', 'FluxPlane', 'LineSource', 'PointDipole', 'Simulation', 'Simulation3D']
