"""Two-body MLP weights: loaders and the seeded initialiser.

The pair function is the TorchScript ``Net`` of python/convert.py:6-20 around
``Sequential(Linear(3,128), Tanh, Linear(128,512), Tanh, Linear(512,9))``
(nn/two_body_unbounded.pkl).  Weights travel as one flat float32 vector in
the order ``W1(128,3) | b1(128) | W2(512,128) | b2(512) | W3(9,512) | b3(9)``
(71,177 floats), row-major, exactly the tensor layout of the archives.

Both model files of the reference are zip archives: ``torch.save`` of the
Sequential (``.pkl``) and ``torch.jit.save`` of ``Net`` (``.pt``,
``<base>_<device>.pt`` for GPU runs, python/convert.py:41-57,
src/HignnModel.cpp:404-411).  Each holds ``<root>/data.pkl`` describing the
objects and ``<root>/data/<key>`` raw little-endian storages.  The loaders
here never unpickle: ``data.pkl`` is read opcode by opcode
(``pickletools.genops``) by a small stack machine that only records which
names, dictionaries and tensor references the stream builds -- no global is
imported and no callable is invoked -- and the tensors are then read from
the raw storages.
"""
from __future__ import annotations

import os
import zipfile

import numpy as np

H1, H2, NOUT = 128, 512, 9
SHAPES = [("W1", (H1, 3)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)),
          ("W3", (NOUT, H2)), ("b3", (NOUT,))]
TOTAL = sum(int(np.prod(s)) for _, s in SHAPES)  # 71,177


def net_shapes(din: int):
    """Layer shapes of Three_body_net (din 6) and Two_body_self_net (din 3):
    Linear(din,256) Tanh Linear(256,1024) Tanh Linear(1024,512) Tanh
    Linear(512,9) (python/HIGNN/model_structure.py:41-68; the nets of
    python/Saved_Model/<run>/HIGNN_nn_3body.pkl and HIGNN_nn_self.pkl)."""
    return [("W1", (256, din)), ("b1", (256,)), ("W2", (1024, 256)), ("b2", (1024,)),
            ("W3", (512, 1024)), ("b3", (512,)), ("W4", (9, 512)), ("b4", (9,))]


def pack(parts: dict, shapes=None) -> np.ndarray:
    """Concatenate the weights (W1,b1,W2,b2,... in order) into the flat float32 layout."""
    out = []
    for name, shape in shapes or SHAPES:
        a = np.asarray(parts[name], dtype=np.float32)
        if a.shape != shape:
            raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        out.append(a.reshape(-1))
    return np.ascontiguousarray(np.concatenate(out))


def unpack(flat: np.ndarray, shapes=None) -> dict:
    shapes = shapes or SHAPES
    total = sum(int(np.prod(s)) for _, s in shapes)
    flat = np.asarray(flat, dtype=np.float32).reshape(-1)
    if flat.size != total:
        raise ValueError(f"expected {total} floats, got {flat.size}")
    parts, off = {}, 0
    for name, shape in shapes:
        n = int(np.prod(shape))
        parts[name] = flat[off:off + n].reshape(shape)
        off += n
    return parts


def xavier(seed: int = 0, shapes=None) -> np.ndarray:
    """Seeded initialisation of the same layer shapes.

    Weights: Xavier/Glorot uniform, U(-a, a) with a = sqrt(6/(fan_in+fan_out)).
    Biases: U(-1/sqrt(fan_in), 1/sqrt(fan_in)) (the default bias range of a
    torch Linear layer).  Generator: numpy PCG64 seeded with ``seed``; the
    draws go W1, b1, W2, b2, W3, b3.
    """
    shapes = shapes or SHAPES
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = {}
    for (wn, (fo, fi)), (bn, _) in zip(shapes[0::2], shapes[1::2]):
        a = np.sqrt(6.0 / (fi + fo))
        parts[wn] = rng.uniform(-a, a, size=(fo, fi)).astype(np.float32)
        bb = 1.0 / np.sqrt(fi)
        parts[bn] = rng.uniform(-bb, bb, size=(fo,)).astype(np.float32)
    return pack(parts, shapes)


def round_bf16(a: np.ndarray) -> np.ndarray:
    """float32 values rounded to the nearest BF16 value (ties to even)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def smooth(seed: int = 0, w1_scale: float = 0.01, bf16_exact: bool = True) -> np.ndarray:
    """A smooth pair function of the same shapes: ``xavier(seed)`` with the
    layer-1 weights W1 multiplied by ``w1_scale`` (biases unchanged) and,
    with ``bf16_exact``, W2 and W3 rounded to BF16 values (the weights of a
    BF16 checkpoint; every path, FP32 and FP64 included, then computes with
    the same numbers).

    With Xavier's W1 (|W1 r| ~ 0.2 |r|) the first tanh layer saturates a few
    radii from the origin, so the pair function is close to a piecewise
    constant function of the direction of r, cut by the 128 planes of W1,
    divided by |r|: far blocks are far from low rank and no small net follows
    it.  At w1_scale = 0.01 the first layer varies on the scale of a few
    hundred radii (the size of a 10^6-particle suspension), so (y + I)/|r| is
    a smooth function divided by |r|: far blocks keep the low ranks of an
    Oseen-like kernel (0.5-0.9x those of the RPY tensor at the same
    separation in the tolerance range 1e-4..1e-3), and the ACA tolerance test
    ends the iterations.  Not a physical mobility: a seeded stand-in for a
    trained net, whose far field is smooth in the same way.

    Why BF16-exact W2 and W3: the BF16 MFMA path rounds its operands to BF16.
    Rounding FP32 weights moves every block entry by a fixed relative amount
    (measured on this net for |r| in [2, 30]: 9e-4 from W2, 7e-4 from W3,
    the same sign for every pair), which sums of many pairs do not average
    out; the activation rounding (about 8e-4 per layer) is random and does.
    """
    w = xavier(seed)
    p = unpack(w)
    p = {k: v.copy() for k, v in p.items()}
    p["W1"] = (p["W1"].astype(np.float64) * w1_scale).astype(np.float32)
    if bf16_exact:
        p["W2"] = round_bf16(p["W2"])
        p["W3"] = round_bf16(p["W3"])
    return pack(p)


class _Ref:
    """A global (module, name) named by the opcode stream; never imported."""

    def __init__(self, module, name):
        self.module, self.name = module, name


class _Obj:
    """An object the stream would build: constructor, arguments, state."""

    def __init__(self, fn, args):
        self.fn, self.args, self.state, self.items = fn, args, None, {}


class _Pers:
    def __init__(self, pid):
        self.pid = pid


def _walk_pickle(data: bytes):
    """Replays the opcodes of a pickle symbolically; returns the root."""
    import pickletools

    stack, memo, marks = [], {}, []

    def pop_mark():
        i = marks.pop()
        items = stack[i:]
        del stack[i:]
        return items

    for op, arg, _ in pickletools.genops(data):
        n = op.name
        if n in ("PROTO", "FRAME", "STOP"):
            continue
        if n == "GLOBAL":
            mod, _, name = arg.partition(" ")
            stack.append(_Ref(mod, name))
        elif n == "STACK_GLOBAL":
            name = stack.pop()
            mod = stack.pop()
            stack.append(_Ref(mod, name))
        elif n in ("BINPUT", "LONG_BINPUT", "PUT"):
            memo[arg] = stack[-1]
        elif n == "MEMOIZE":
            memo[len(memo)] = stack[-1]
        elif n in ("BINGET", "LONG_BINGET", "GET"):
            stack.append(memo[arg])
        elif n == "MARK":
            marks.append(len(stack))
        elif n == "EMPTY_TUPLE":
            stack.append(())
        elif n in ("TUPLE1", "TUPLE2", "TUPLE3"):
            k = int(n[-1])
            t = tuple(stack[-k:])
            del stack[-k:]
            stack.append(t)
        elif n == "TUPLE":
            stack.append(tuple(pop_mark()))
        elif n == "EMPTY_LIST":
            stack.append([])
        elif n == "LIST":
            stack.append(list(pop_mark()))
        elif n == "APPEND":
            v = stack.pop()
            stack[-1].append(v)
        elif n == "APPENDS":
            vs = pop_mark()
            stack[-1].extend(vs)
        elif n == "EMPTY_DICT":
            stack.append({})
        elif n == "DICT":
            vs = pop_mark()
            stack.append(dict(zip(vs[0::2], vs[1::2])))
        elif n in ("SETITEM", "SETITEMS"):
            vs = [stack.pop(), stack.pop()][::-1] if n == "SETITEM" else pop_mark()
            tgt = stack[-1]
            d = tgt.items if isinstance(tgt, _Obj) else tgt
            for k, v in zip(vs[0::2], vs[1::2]):
                d[k] = v
        elif n in ("EMPTY_SET", "FROZENSET"):
            stack.append(set() if n == "EMPTY_SET" else frozenset(pop_mark()))
        elif n == "ADDITEMS":
            pop_mark()
        elif n in ("REDUCE", "NEWOBJ"):
            args = stack.pop()
            fn = stack.pop()
            stack.append(_Obj(fn, args))
        elif n == "BUILD":
            state = stack.pop()
            if isinstance(stack[-1], _Obj):
                stack[-1].state = state
        elif n == "BINPERSID":
            stack.append(_Pers(stack.pop()))
        elif n == "NONE":
            stack.append(None)
        elif n in ("NEWTRUE", "NEWFALSE"):
            stack.append(n == "NEWTRUE")
        elif n in ("BININT", "BININT1", "BININT2", "LONG1", "LONG4", "INT", "LONG", "BINFLOAT",
                   "FLOAT", "BINUNICODE", "SHORT_BINUNICODE", "BINUNICODE8", "UNICODE",
                   "BINSTRING", "SHORT_BINSTRING", "STRING", "BINBYTES", "SHORT_BINBYTES",
                   "BINBYTES8"):
            stack.append(arg)
        elif n == "POP":
            stack.pop()
        elif n == "POP_MARK":
            pop_mark()
        elif n == "DUP":
            stack.append(stack[-1])
        else:
            raise ValueError(f"unsupported pickle opcode {n}")
    return stack[-1] if stack else None


def _tensors(node, path="", out=None, seen=None):
    """(path, storage key, dtype name, offset, shape, stride) of every tensor
    reference below node, in stream order."""
    if out is None:
        out, seen = [], set()
    if id(node) in seen:
        return out
    if isinstance(node, (_Obj, dict, list, tuple)):
        seen.add(id(node))
    if isinstance(node, _Obj):
        fn = node.fn
        if isinstance(fn, _Ref) and fn.module == "torch._utils" and fn.name == "_rebuild_tensor_v2":
            pers, offset, size, stride = node.args[:4]
            pid = pers.pid
            stype = pid[1]
            dtype = stype.name if isinstance(stype, _Ref) else str(stype)
            out.append((path, str(pid[2]), dtype, int(offset), tuple(size), tuple(stride)))
            return out
        if isinstance(fn, _Ref) and fn.module == "torch._utils" and fn.name == "_rebuild_parameter":
            return _tensors(node.args[0], path, out, seen)
        for k, v in node.items.items():
            _tensors(v, f"{path}.{k}", out, seen)
        if node.state is not None:
            _tensors(node.state, path, out, seen)
        if isinstance(node.args, tuple):
            for v in node.args:
                _tensors(v, path, out, seen)
    elif isinstance(node, dict):
        for k, v in node.items():
            _tensors(v, f"{path}.{k}", out, seen)
    elif isinstance(node, (list, tuple)):
        for v in node:
            _tensors(v, path, out, seen)
    return out


def load_torch_zip(path: str, shapes=None) -> np.ndarray:
    """Raw weights from a torch-saved module (.pkl) or a TorchScript Net
    (.pt) archive, without unpickling: the tensors of the given layer shapes
    (default: the pair function's W1, b1, W2, b2, W3, b3), taken in the
    stream's order."""
    shapes = shapes or SHAPES
    with zipfile.ZipFile(path) as z:
        names = z.namelist()
        pk = [n for n in names if n.endswith("/data.pkl")]
        if len(pk) != 1:
            raise ValueError(f"{path}: expected one data.pkl, found {pk}")
        root = pk[0][: -len("data.pkl")]
        refs = _tensors(_walk_pickle(z.read(pk[0])))
        want = [shape for _, shape in shapes]
        parts, k = {}, 0
        for p_, key, dtype, offset, size, stride in refs:
            if k < len(want) and size == want[k]:
                if dtype != "FloatStorage":
                    raise ValueError(f"{path}: tensor {p_} is {dtype}, expected float32")
                n = int(np.prod(size))
                expect = tuple(int(np.prod(size[i + 1:])) for i in range(len(size)))
                if stride != expect:
                    raise ValueError(f"{path}: tensor {p_} is not contiguous")
                raw = np.frombuffer(z.read(f"{root}data/{key}"), dtype="<f4")
                parts[shapes[k][0]] = raw[offset:offset + n].reshape(size)
                k += 1
        if k != len(want):
            raise ValueError(f"{path}: found {k} of the {len(want)} MLP tensors "
                             f"({[r[4] for r in refs]})")
    return pack(parts, shapes)


def load_pkl(path: str) -> np.ndarray:
    """Raw weights from a torch-saved Sequential (zip of float32 storages)."""
    return load_torch_zip(path)


def resolve_model_path(base: str, device: int | None = None) -> str:
    """load_two_body_model(base) resolution: the reference loads
    ``base_<device>.pt`` on GPU builds and ``base.pt`` otherwise
    (src/HignnModel.cpp:404-411).  Tried here in this order: ``base`` itself,
    ``base_<device>.pt`` (device from LOCAL_RANK, default 0), ``base.pt``,
    ``base.pkl`` (the torch-saved Sequential) and ``base.npy`` (a flat float32
    vector)."""
    if device is None:
        device = int(os.environ.get("LOCAL_RANK", "0"))
    for cand in (base, f"{base}_{device}.pt", base + ".pt", base + ".pkl", base + ".npy"):
        if os.path.isfile(cand):
            return cand
    raise FileNotFoundError(f"no two-body model at {base}(_<device>.pt|.pt|.pkl|.npy)")


def load(base: str) -> np.ndarray:
    """Weights of load_two_body_model(base); also "xavier:<seed>" and
    "smooth:<seed>" for the seeded weight sets above."""
    if base.startswith("xavier:"):
        return xavier(int(base.split(":", 1)[1]))
    if base.startswith("smooth:"):
        return smooth(int(base.split(":", 1)[1]))
    path = resolve_model_path(base)
    if path.endswith(".npy"):
        return pack(unpack(np.load(path, allow_pickle=False)))
    return load_torch_zip(path)


def load_net(path: str, din: int) -> np.ndarray:
    """Weights of a three-body (din 6) or self (din 3) net from a torch-saved
    module archive (.pkl), a .npy flat vector, or "xavier:<seed>"."""
    shapes = net_shapes(din)
    if path.startswith("xavier:"):
        return xavier(int(path.split(":", 1)[1]), shapes)
    if path.endswith(".npy"):
        return pack(unpack(np.load(path, allow_pickle=False), shapes), shapes)
    return load_torch_zip(path, shapes)
