"""zombie_amd: an MI355X (gfx950) Walk-on-Spheres / Walk-on-Stars engine.

The GPU work happens in libzombie_amd.so (hand-written HIP for gfx950) behind the C-ABI of
include/zombie_amd.h.  `engine` is the array-level host API; `zombie` mirrors the reference's
Python surface (zombie.Core / zombie.Solvers / zombie.Utils) over it and `demo` runs model
problems in the reference's JSON format; `problems` builds the manufactured benchmark problems;
`learned_cv`, `bvc` and `autotune` hold the learned control variate, boundary value caching and
the launch-shape autotuner.
"""
from . import _abi
from ._abi import ZombieError, ZombieLibraryError

__version__ = "0.1.0"


def library_path():
    return _abi.LIB_PATH


def load_library():
    """Loads libzombie_amd.so; raises ZombieLibraryError when it is missing."""
    return _abi.lib()
