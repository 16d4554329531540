"""rwig_amd: MI355X-native entropic optimal transport behind rwig's
sinkhorn() / barycenter() / wdl() / tsvd() / wig() entry points.

The solves run in librwig_hip.so (hand-written HIP kernels for gfx950),
reached through a C ABI (include/rwig_hip.h).
"""
from .api import (barycenter, barycenter_log_ck, cuda_available, doc2dist,  # noqa: F401
                  euclidean, sinkhorn, wdl)
from .wig import tsvd, wig, wig_index, wig_scores  # noqa: F401
from ._lib import RwigError  # noqa: F401
