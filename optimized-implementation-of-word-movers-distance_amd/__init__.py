"""MI355X-native Sinkhorn-Knopp Word Mover's Distance.

The compute path is libwmd.so (hand-written HIP for gfx950 behind the C ABI
of include/wmd.h).  This package is its Python host mirror: the reference's
``sinkhorn_wmd`` interface (main.py:37), the resident-corpus engine, the
synthetic workload generators and the reference's on-disk formats.
"""
from . import datagen, formats, prep  # noqa: F401
from ._native import (VARIANT_FUSED, VARIANT_NONFUSED, VARIANT_RESIDENT,  # noqa: F401
                      WMDError, load)
from .api import (WMDEngine, mlp_init, mlp_layout, mlp_loss_grad,  # noqa: F401
                  mlp_param_count, shapes, sinkhorn_wmd, sinkhorn_wmd_fused,
                  sinkhorn_wmd_nonfused)

PACKAGE = __name__
