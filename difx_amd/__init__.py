"""MI355X-native FX hot path of the DiFX software correlator (mpifxcorr Mode/Core).

Modules:
  config   job description (freq/datastream/baseline tables, delay models) and its C ABI form
  core     FxCore: Core::processdata / Mode getters over libfxg.so (include/fxg.h)
  synth    seeded synthetic VDIF/Mark5B slot inputs
  build    hipcc build of libfxg.so for gfx950
"""
from .config import Config, Datastream, Freq, Baseline, standard_config  # noqa: F401

__all__ = ["Config", "Datastream", "Freq", "Baseline", "standard_config"]
