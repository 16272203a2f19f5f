"""MI355X split-step wave-propagation library (host side).

Thin Python layer over librsim.so, mirroring big-wave's entry points:
  multisim.run_single_simulation  -> ravesim.multisim.run_single_simulation
  multisim.setup_simulation       -> ravesim.simdir.setup_simulation
  propagation / optical_element   -> ravesim.core.Field operators
and a batched runner (ravesim.core.run_batch) for many energies / sources.
"""
from ._lib import RS_C8, RS_C16, RS_MODE_FUSED, RS_MODE_PLAIN, RsError  # noqa: F401
from .core import BatchSim, Context, DBAGrating, Field, Grating, Sample, energy_to_wavelength, run_batch, run_mixed  # noqa: F401
