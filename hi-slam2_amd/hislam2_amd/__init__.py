"""MI355X-native dense-tracking hot path of HI-SLAM2.

``droid_backends`` (the compiled extension next to this package) is the drop-in
for the reference module of the same name; this package holds the Python-side
mirror of the reference's callers of it (``corr`` for hislam2/modules/corr.py) and
the synthetic workloads (``synth``).  The extension is required: importing
anything that computes raises ImportError when it has not been built
(``python hi-slam2_amd/build.py``).
"""
import importlib
import os
import sys

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def backend():
    """The in-tree droid_backends extension module (built by hi-slam2_amd/build.py)."""
    if _HERE not in sys.path:
        sys.path.insert(0, _HERE)
    try:
        return importlib.import_module('droid_backends')
    except ImportError as e:      # no silent fallback: the HIP path is the product
        raise ImportError('droid_backends is not built; run `python hi-slam2_amd/build.py` '
                          f'({e})') from e
