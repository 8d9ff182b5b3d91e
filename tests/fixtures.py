"""Known answers of the reference's own tests, kept as data in
tests/golden/reference_vectors.json (each entry names the file:line it was
transcribed from)."""
import json
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.json")
_VECTORS = None


def golden(name):
    global _VECTORS
    if _VECTORS is None:
        with open(_PATH) as f:
            _VECTORS = json.load(f)
    return _VECTORS[name]


def hexint(s):
    return int(s, 16)
