"""JSON codec for numpy arrays, the on-disk format of the reference's to_json /
from_json (pyANNonGPU/json_numpy.py:5-38): arrays become
{"type": "ndarray", "dtype", "complex", "real"/"imag" or "data"}."""
import json

import numpy as np


class NumpyEncoder(json.JSONEncoder):
    def default(self, obj):
        if isinstance(obj, np.ndarray):
            out = {"type": "ndarray", "dtype": str(obj.dtype), "complex": bool(np.iscomplexobj(obj))}
            if out["complex"]:
                out["real"] = obj.real.tolist()
                out["imag"] = obj.imag.tolist()
            else:
                out["data"] = obj.tolist()
            return out
        if isinstance(obj, np.generic):
            return obj.item() if not np.iscomplexobj(obj) else {"real": float(obj.real), "imag": float(obj.imag)}
        if isinstance(obj, complex):
            return {"real": obj.real, "imag": obj.imag, "type": "complex"}
        return super().default(obj)


class NumpyDecoder(json.JSONDecoder):
    def __init__(self, *args, **kwargs):
        super().__init__(object_hook=self.object_hook, *args, **kwargs)

    @staticmethod
    def object_hook(obj):
        if obj.get("type") == "ndarray":
            dtype = np.dtype(obj["dtype"])
            if obj["complex"]:
                return np.array(obj["real"], dtype=dtype) + 1j * np.array(obj["imag"], dtype=dtype)
            return np.array(obj["data"], dtype=dtype)
        if obj.get("type") == "complex":
            return complex(obj["real"], obj["imag"])
        return obj
