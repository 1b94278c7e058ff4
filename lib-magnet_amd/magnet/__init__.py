"""MAGNET's full-batch loss="tsne" pass on MI355X (HIP, gfx950).

The reference package is kerighan/lib-magnet v3.0.2 (magnet/__init__.py, magnet/tsne.py); this
package keeps its t-SNE interface -- optimize_tsne, FFTRepulsion2D, MAGNET(loss="tsne") with fit_transform and
transform -- over a HIP implementation, plus a continuous-batching transform service (TransformServer).
"""
from .tsne import FFTRepulsion2D, optimize_tsne, grid_params
from .affinity import perplexity_affinities
from .estimator import MAGNET
from .serving import TransformServer

__version__ = "3.0.2+mi355x"
__all__ = ["FFTRepulsion2D", "optimize_tsne", "grid_params", "perplexity_affinities", "MAGNET", "TransformServer",
           "__version__"]
