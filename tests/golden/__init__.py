"""Golden vectors for the FFT path (inputs and closed-form expected outputs), written by
make_fixtures.py in this directory."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_fixture(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
