"""tdep_amd: MI355X engine for the three-phonon scattering and lattice thermal
conductivity path of TDEP's `thermal_conductivity` program.

The compute lives in libtc3.so (HIP for gfx950, C ABI in include/tc3.h); this
package holds its ctypes binding (`lib`), input readers (`io`), the Python
mirror of the reference driver (`engine`) and the synthetic Morse inputs
(`inputs`).
"""
from .lib import load, Tc3Error  # noqa: F401
