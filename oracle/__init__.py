"""ORACLE — test infrastructure, not product code.

CPU restatements of the reference's FDTD (Yee) hot path, used only as the checker:
  * yee2d: 2D E_z / H_z Yee stepper with CPML, periodic wraps, calibrated sheet sources and flux monitors
    (numpy; float64 by default, float32 for the numpy-twin CPU baseline);
  * yee3d: 3D Yee stepper with fused CPML, material-index coefficients, point dipole and plane flux monitors.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package, and there only to
check or to time a CPU baseline, never as the thing measured or shipped. The product package (opensol_amd) does not
import it and fails loudly when its HIP library is missing.

Pinning (what each part is checked against, see tests/test_oracle_*.py):
  * end-to-end 2D composition: golden vectors written by the reference's own numpy path
    (tests/golden/thin_film_{ez,hz}.npz, precision 'double');
  * bulk Yee update: the exact discrete dispersion of a PEC/PMC cavity mode;
  * CPML: reflection against a larger-domain run; profile closed forms;
  * sheet calibration: plane-wave amplitude in a periodic strip;
  * DFT / flux: the transform of an analytic signal, Poynting flux of a plane wave;
  * 3D: discrete box resonances, the 2D-invariant reduction to the 2D stepper, dipole power.
"""
