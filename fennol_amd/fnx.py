"""Reading and writing FeNNol ``.fnx`` model files.

A ``.fnx`` file (FENNIX.save / FENNIX.load, models/fennix.py:647-688) is the
msgpack encoding, in Flax's serialization format, of

    {"cutoff", "modules": [[name, params], ...], "preprocessing": [[name, params], ...],
     "energy_unit", "energy_terms", "variables"}

Flax encodes NumPy arrays as msgpack extension type 1 whose payload is the
msgpack of (shape, dtype name, C-order bytes), and NumPy scalars as type 3
with the same payload.  The format is kept because it is the file format
users already have; arrays larger than Flax's chunking threshold (2**30
bytes) are not handled (no FeNNol ANI model comes near it).
"""

from __future__ import annotations

from collections import OrderedDict

import msgpack
import numpy as np

_EXT_NDARRAY = 1
_EXT_COMPLEX = 2
_EXT_NPSCALAR = 3


def _array_payload(arr: np.ndarray) -> bytes:
    arr = np.asarray(arr)
    if arr.dtype.hasobject:
        raise ValueError("object arrays cannot be serialised")
    return msgpack.packb((list(arr.shape), arr.dtype.name, arr.tobytes("C")), use_bin_type=True)


def _array_from_payload(data: bytes) -> np.ndarray:
    shape, dtype_name, buf = msgpack.unpackb(data, raw=False)
    return np.frombuffer(buf, dtype=np.dtype(dtype_name)).reshape(shape).copy()


def _default(x):
    if isinstance(x, np.ndarray):
        return msgpack.ExtType(_EXT_NDARRAY, _array_payload(x))
    if isinstance(x, np.generic):
        return msgpack.ExtType(_EXT_NPSCALAR, _array_payload(np.asarray(x)))
    if isinstance(x, complex):
        return msgpack.ExtType(_EXT_COMPLEX, msgpack.packb((x.real, x.imag)))
    if isinstance(x, tuple):
        return list(x)
    raise TypeError(f"cannot serialise {type(x)}")


def _ext_hook(code, data):
    if code == _EXT_NDARRAY:
        return _array_from_payload(data)
    if code == _EXT_NPSCALAR:
        return _array_from_payload(data)[()]
    if code == _EXT_COMPLEX:
        re, im = msgpack.unpackb(data)
        return complex(re, im)
    return msgpack.ExtType(code, data)


def dumps(state: dict) -> bytes:
    return msgpack.packb(state, default=_default, strict_types=True, use_bin_type=True)


def loads(data: bytes) -> dict:
    return msgpack.unpackb(data, ext_hook=_ext_hook, raw=False, strict_map_key=False)


def save(model, filename) -> None:
    """Write a FENNIX model (fennol_amd.fennix.FENNIX) as a .fnx file."""
    state = {
        "cutoff": model._input_args["cutoff"],
        "modules": [[k, dict(v)] for k, v in model._input_args["modules"].items()],
        "preprocessing": [[k, dict(v)] for k, v in model._input_args["preprocessing"].items()],
        "energy_unit": model._input_args["energy_unit"],
        "energy_terms": model.energy_terms,
        "variables": model.variables,
    }
    with open(filename, "wb") as f:
        f.write(dumps(state))


def load_state(filename) -> dict:
    with open(filename, "rb") as f:
        state = loads(f.read())
    state["preprocessing"] = OrderedDict((k, v) for k, v in state["preprocessing"])
    state["modules"] = OrderedDict((k, v) for k, v in state["modules"])
    return state
