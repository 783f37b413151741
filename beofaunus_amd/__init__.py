"""beofaunus_amd -- MI355X-native Metropolis Delta-U engine for the Faunus
hot path (Energy::Nonbonded / NonbondedEwald, Move::AtomicTranslation /
TranslateRotate / ParallelTempering).

The compute lives in libfgx.so (HIP, gfx950) behind the C ABI in
include/fgx.h; `beofaunus_amd.fgx` is its ctypes binding,
`beofaunus_amd.faunus_io` reads and writes the reference's JSON inputs and
Space::save state files, `beofaunus_amd.workloads` builds the benchmark
systems and `beofaunus_amd.analysis` samples g(r) and <U> on the device.
The C++ binding that plugs into the reference's own classes is
include/faunus_amd/reference_adapter.h.
"""
from . import fgx  # noqa: F401

__all__ = ["fgx"]
