"""costal_amd: CosTaL's exact cosine kNN + Tanimoto edge weights on MI355X (gfx950).

The public functions keep the signatures of ike12th/CosTaL ``python/costal`` (knn, reweight,
truncate, build_graph, find_communities); ``costal_amd._costal`` mirrors the reference's low-level
``costal._costal`` module.  The compute runs in ``libcostal_hip.so`` (include/costal_hip.h).
"""

from . import _costal
from ._api import CostalWarning, KnnGraph, build_graph, find_communities, knn, reweight, truncate
from ._costal import default_matmul

__version__ = _costal.__version__

__all__ = ["CostalWarning", "KnnGraph", "build_graph", "default_matmul", "find_communities", "knn",
           "reweight", "truncate"]
