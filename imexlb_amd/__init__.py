"""MI355X-native D3Q27 BGK time step of imexlb behind a C-ABI.

The compute path is libimexlb.so (hand-written HIP for gfx950, built in-tree by
``python -m imexlb_amd.build``); this package is its Python binding, with the
reference's LBM / System / CommHelper names.
"""
from .lbm import (BANDWIDTH_PROBES, LBM, Ensemble, System, bandwidth_probe, check, comm_layout,  # noqa: F401
                  copy_bandwidth, tecplot_write)
from ._lib import ImexlbError, LIB_PATH, lib  # noqa: F401
