"""ennemi_amd: ennemi's nearest-neighbour MI / entropy estimators on MI355X.

Drop-in for the estimate_mi / estimate_entropy / pairwise_mi call surface of
ennemi 1.5.0; the k-th-neighbour radii, ball counts, digamma sums and means
run in hand-written HIP kernels for gfx950 (libennemi_hip.so).
"""

from ._driver import (estimate_corr, estimate_entropy, estimate_mi, normalize_mi,
                      pairwise_corr, pairwise_mi)

__all__ = ["estimate_entropy", "estimate_mi", "normalize_mi", "pairwise_mi",
           "estimate_corr", "pairwise_corr"]
__version__ = "0.1.0"
