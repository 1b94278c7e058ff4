"""The compiled reference's nn_descent (oracle/_ref, src/nndescent.cpp:636-645) on a saved array, run by
tests/test_gpu_nndescent.py in a process of its own: `python tests/ref_nndescent_graph.py X.npy out.npz k n_trees seed`
writes the indices, squared distances and seconds."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

if __name__ == "__main__":
    src, dst, k, n_trees, seed = sys.argv[1], sys.argv[2], int(sys.argv[3]), int(sys.argv[4]), int(sys.argv[5])
    R = O.ref_module()
    X = np.load(src)
    t = time.time()
    ind, dist = R.nn_descent(X, k, 50, 10, 0.001, n_trees, 30, seed, 16, False)
    np.savez(dst, ind=np.asarray(ind), dist=np.asarray(dist), seconds=time.time() - t)
