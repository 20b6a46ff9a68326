"""MI355X-native Mille-feuille: tiled sparse format and single-kernel CG/BiCGSTAB.

Import name ``mille_feuille_amd`` (the directory is loaded by path, see
``load_package`` in ``__graft_entry__.py``).  Compute runs in ``libmf.so``
(HIP kernels for gfx950 behind the C-ABI of ``include/mf.h``).
"""
from . import workloads  # noqa: F401
from ._lib import LIB_PATH, MfError, exported_symbols, lib  # noqa: F401
from . import autograd  # noqa: F401
from .api import (Batch, TiledMatrix, Tiles, decode, device_name, dyn_classify, encode, read_matrix,  # noqa: F401
                  solve_options, tile_classes, tile_create, value_copies, write_cbd)

__all__ = ["Batch", "TiledMatrix", "Tiles", "tile_create", "tile_classes", "value_copies", "encode", "decode",
           "read_matrix", "write_cbd", "device_name", "solve_options", "workloads", "MfError", "lib"]
