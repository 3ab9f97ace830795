"""Loader of the committed input/output vectors under tests/golden/ (written by the scripts there)."""
import functools
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@functools.lru_cache(maxsize=None)
def df_vectors():
    """The df-function vectors of tests/golden/make_df_vectors.py, as a dict of arrays."""
    with np.load(os.path.join(GOLDEN, "df_vectors.npz")) as z:
        return {k: z[k] for k in z.files}
