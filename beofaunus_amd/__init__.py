"""beofaunus_amd -- MI355X-native Metropolis Delta-U engine for the Faunus
hot path (Energy::Nonbonded / NonbondedEwald, Move::AtomicTranslation /
TranslateRotate / ParallelTempering).

The compute lives in libfgx.so (HIP, gfx950) behind the C ABI in
include/fgx.h; `beofaunus_amd.fgx` is its ctypes binding and
`beofaunus_amd.faunus` mirrors the reference's class names on top of it.
"""
from . import fgx  # noqa: F401

__all__ = ["fgx"]
