"""MAGNET's full-batch loss="tsne" pass on MI355X (HIP, gfx950).

The reference package is kerighan/lib-magnet v3.0.2 (magnet/__init__.py, magnet/tsne.py); this
package keeps its t-SNE interface -- optimize_tsne, FFTRepulsion2D -- over a HIP implementation.
"""
from .tsne import FFTRepulsion2D, optimize_tsne, grid_params
from .affinity import perplexity_affinities, exact_knn, calibrate_perplexity

__version__ = "3.0.2+mi355x"
__all__ = ["FFTRepulsion2D", "optimize_tsne", "grid_params", "perplexity_affinities", "exact_knn",
           "calibrate_perplexity", "__version__"]
