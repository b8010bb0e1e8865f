"""Runs oracle/_ref/ref_golden, the reference's own CPU sources linked with a small driver.

TEST INFRASTRUCTURE ONLY (golden-fixture generation and the reference CPU baseline). The
binary is built by `make -C oracle ref` from the reference checkout; it is absent where that
checkout is absent, and callers must treat it as optional.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "_ref", "ref_golden")


def available() -> bool:
    return os.path.exists(BIN) and os.access(BIN, os.X_OK)


def _fmt(v):
    if isinstance(v, (list, tuple)):
        return ",".join(_fmt(x) for x in v)
    if isinstance(v, float):
        return repr(float(v))
    return str(v)


def parse_meta(text: str) -> dict:
    out = {}
    for line in text.splitlines():
        if ": " not in line:
            continue
        k, v = line.split(": ", 1)
        if k == "residual_history":
            out[k] = np.array([float(x) for x in v.split(",") if x])
            continue
        try:
            out[k] = int(v)
        except ValueError:
            try:
                out[k] = float(v)
            except ValueError:
                out[k] = v
    return out


def run(command: str, inputs=(), out_shapes=(), timeout=600, **params):
    """Runs one command; returns (list of output arrays reshaped to out_shapes, meta dict)."""
    if not available():
        raise RuntimeError("reference driver not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory() as td:
        args = [BIN, command]
        if inputs:
            fin = os.path.join(td, "in.f64")
            np.concatenate([np.ascontiguousarray(a, dtype=np.float64).ravel() for a in inputs]).tofile(fin)
            args.append(f"in={fin}")
        fout = os.path.join(td, "out.f64")
        fmeta = os.path.join(td, "meta.txt")
        args += [f"out={fout}", f"meta={fmeta}"]
        args += [f"{k}={_fmt(v)}" for k, v in params.items()]
        subprocess.run(args, check=True, timeout=timeout)
        meta = parse_meta(open(fmeta).read()) if os.path.exists(fmeta) else {}
        outs = []
        if out_shapes:
            flat = np.fromfile(fout, dtype=np.float64)
            pos = 0
            for shp in out_shapes:
                n = int(np.prod(shp))
                outs.append(flat[pos:pos + n].reshape(shp).copy())
                pos += n
            if pos != flat.size:
                raise RuntimeError(f"{command}: output size {flat.size} != expected {pos}")
        return outs, meta
