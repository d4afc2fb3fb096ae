"""Two-body MLP weights: loaders and the seeded initialiser.

The pair function is the TorchScript ``Net`` of python/convert.py:6-20 around
``Sequential(Linear(3,128), Tanh, Linear(128,512), Tanh, Linear(512,9))``
(nn/two_body_unbounded.pkl).  Weights travel as one flat float32 vector in
the order ``W1(128,3) | b1(128) | W2(512,128) | b2(512) | W3(9,512) | b3(9)``
(71,177 floats), row-major, exactly the tensor layout of the pickle.

``load_pkl`` reads the pickle as what it is on disk -- a zip archive whose
``<root>/data/<k>`` members hold the raw little-endian float32 storages --
without unpickling anything.  The storage keys 0..5 map to
W1, b1, W2, b2, W3, b3 (the order in which ``torch.save`` wrote the
Sequential's parameters).
"""
from __future__ import annotations

import os
import zipfile

import numpy as np

H1, H2, NOUT = 128, 512, 9
SHAPES = [("W1", (H1, 3)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)),
          ("W3", (NOUT, H2)), ("b3", (NOUT,))]
TOTAL = sum(int(np.prod(s)) for _, s in SHAPES)  # 71,177


def pack(parts: dict) -> np.ndarray:
    """Concatenate W1,b1,W2,b2,W3,b3 into the flat float32 layout."""
    out = []
    for name, shape in SHAPES:
        a = np.asarray(parts[name], dtype=np.float32)
        if a.shape != shape:
            raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        out.append(a.reshape(-1))
    return np.ascontiguousarray(np.concatenate(out))


def unpack(flat: np.ndarray) -> dict:
    flat = np.asarray(flat, dtype=np.float32).reshape(-1)
    if flat.size != TOTAL:
        raise ValueError(f"expected {TOTAL} floats, got {flat.size}")
    parts, off = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        parts[name] = flat[off:off + n].reshape(shape)
        off += n
    return parts


def xavier(seed: int = 0) -> np.ndarray:
    """Seeded initialisation of the same layer shapes.

    Weights: Xavier/Glorot uniform, U(-a, a) with a = sqrt(6/(fan_in+fan_out)).
    Biases: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) (the default bias range of a
    torch Linear layer).  Generator: numpy PCG64 seeded with ``seed``; the
    draws go W1, b1, W2, b2, W3, b3.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = {}
    for (wn, (fo, fi)), (bn, _) in zip(SHAPES[0::2], SHAPES[1::2]):
        a = np.sqrt(6.0 / (fi + fo))
        parts[wn] = rng.uniform(-a, a, size=(fo, fi)).astype(np.float32)
        bb = 1.0 / np.sqrt(fi)
        parts[bn] = rng.uniform(-bb, bb, size=(fo,)).astype(np.float32)
    return pack(parts)


def load_pkl(path: str) -> np.ndarray:
    """Raw weights from a torch-saved Sequential (zip of float32 storages)."""
    with zipfile.ZipFile(path) as z:
        names = z.namelist()
        blobs = {}
        for n in names:
            head, _, tail = n.rpartition("/data/")
            if head and tail.isdigit():
                blobs[int(tail)] = z.read(n)
    if sorted(blobs) != list(range(6)):
        raise ValueError(f"{path}: expected storages 0..5, found {sorted(blobs)}")
    parts = {}
    for k, (name, shape) in enumerate(SHAPES):
        a = np.frombuffer(blobs[k], dtype="<f4")
        if a.size != int(np.prod(shape)):
            raise ValueError(f"{path}: storage {k} has {a.size} floats, {name} needs {shape}")
        parts[name] = a.reshape(shape)
    return pack(parts)


def resolve_model_path(base: str) -> str:
    """load_two_body_model(base) resolution (src/HignnModel.cpp:404-411 loads
    ``base.pt`` / ``base_<dev>.pt``): here ``base.pkl`` (raw weights) or a
    ``.npy`` flat vector; TorchScript archives are not parsed."""
    for cand in (base, base + ".pkl", base + ".npy"):
        if os.path.isfile(cand):
            return cand
    raise FileNotFoundError(f"no two-body model at {base}(.pkl|.npy)")


def load(base: str) -> np.ndarray:
    path = resolve_model_path(base)
    if path.endswith(".npy"):
        return pack(unpack(np.load(path, allow_pickle=False)))
    return load_pkl(path)
