"""Records the courant.h calls a Python scene builder makes through the ctypes binding, so
that tests/napi/harness.js can make the same calls through the Node addon.

Each call is [name, args, result, outputs after the call (base64)]; an argument is one of
  {"i": int} / {"d": float}             scalar (ints and floats as Python passed them)
  {"sim": n}                            the n-th simulation handle cem_create returned
  {"null": true}                        a null pointer
  {"in": base64, "dtype": str}          an array or struct passed by pointer (its bytes)
  {"out": bytes, "init": base64}        an output buffer (ctypes byref / out array), with the
                                        bytes it held before the call
and the result is the call's return value (null for pointers other than sim handles).
Only calls made while `active` are recorded; every call goes through either way.
"""
import base64
import ctypes as C

import numpy as np

# pointer arguments a function writes (output buffers), by position
OUTPUTS = {
    "cem_fit_djordjevic_sarkar": (6, 7, 8),
}


def _b64(raw):
    return base64.b64encode(bytes(raw)).decode()


class Recorder:
    def __init__(self, lib):
        self._lib = lib
        self.calls = []
        self.handles = {}
        self.active = True

    def __getattr__(self, name):
        f = getattr(self._lib, name)
        if not name.startswith("cem_"):
            return f

        def call(*args):
            if not self.active:
                return f(*args)
            outs = OUTPUTS.get(name, ())
            rec = [self._arg(a, k in outs) for k, a in enumerate(args)]
            r = f(*args)
            after = [_b64(self._raw(args[k])) for k in outs]
            if name == "cem_create" and r:
                self.handles[r] = len(self.handles)
                res = {"sim": self.handles[r]}
            elif isinstance(r, (int, float)) and not isinstance(r, bool):
                res = r
            else:
                res = None
            self.calls.append([name, rec, res, after])
            return r

        return call

    @staticmethod
    def _raw(a):
        arr = getattr(a, "_arr", None)
        if arr is not None:
            return np.ascontiguousarray(arr).tobytes()
        return bytes(memoryview(a._obj).cast("B"))

    def _arg(self, a, is_out):
        if a is None:
            return {"null": True}
        if isinstance(a, (bool, np.bool_)):
            return {"i": int(a)}
        if isinstance(a, np.integer):
            a = int(a)
        if isinstance(a, np.floating):
            a = float(a)
        if isinstance(a, int):
            if a in self.handles:
                return {"sim": self.handles[a]}
            return {"i": a}
        if isinstance(a, float):
            return {"d": a}
        arr = getattr(a, "_arr", None)          # numpy .ctypes.data_as(...)
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            raw = arr.tobytes()
            if is_out:
                return {"out": len(raw), "init": _b64(raw)}
            return {"in": _b64(raw), "dtype": str(arr.dtype)}
        obj = getattr(a, "_obj", None)          # ctypes byref(x)
        if obj is not None:
            raw = bytes(memoryview(obj).cast("B"))
            return {"out": len(raw), "init": _b64(raw)} if is_out else {"in": _b64(raw), "dtype": "raw"}
        objs = getattr(a, "_objects", None)     # ctypes cast of a pointer to a struct
        if isinstance(objs, dict):
            for v in objs.values():
                if isinstance(v, C.Structure):
                    return {"in": _b64(bytes(v)), "dtype": "struct"}
        raise TypeError(f"cannot record argument {a!r}")


def decode_out(arg, dtype=np.float64):
    """The initial contents of a recorded output buffer, as an array."""
    return np.frombuffer(base64.b64decode(arg["init"]), dtype=dtype)
