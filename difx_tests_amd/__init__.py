"""MI355X FX correlator engine.

The compute path is libdifxfx.so (HIP kernels for gfx950 behind a C ABI, see
include/difxfx.h and include/kernel.h); this package holds its Python
bindings (engine), the config grammar (config), the synthetic-data
generator (synth), the workload definitions (workloads) and the build
script (build).
"""
from .engine import FxBatch, FxEngine, FxError, load_library  # noqa: F401
