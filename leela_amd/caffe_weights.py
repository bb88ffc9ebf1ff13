"""Caffe weight files: the reference's USE_CAFFE import path and its
WRITE_WEIGHTS export (Network.cpp:251-341), without Caffe.

The reference loads a net with Caffe's C++ API (`Net::CopyTrainedLayersFrom`
on a .caffemodel) and, with WRITE_WEIGHTS defined, prints every layer's
blobs as C++ `std::array<float, N>` definitions (the form of the NN*.cpp
weight files the nets are compiled from). Here the .caffemodel is read
straight from its protobuf wire format (varints, length-delimited
messages, packed repeated fields), using the public Caffe schema's field
numbers:

    NetParameter   name = 1, layer = 100 (LayerParameter)
    LayerParameter name = 1, type = 2, blobs = 7 (BlobProto),
                   batch_norm_param = 139 (BatchNormParameter)
    BlobProto      num/channels/height/width = 1-4 (legacy), data = 5
                   (float, packed or not), shape = 7 (BlobShape)
    BlobShape      dim = 1 (int64, packed or not)
    BatchNormParameter eps = 3 (float, default 1e-5)

`write_weights` reproduces the WRITE_WEIGHTS text byte for byte: the
`// <shape_string>` comment per blob (Caffe's Blob::shape_string, dims
then the count in parentheses), the array names (convolutions and inner
products number the arrays, `val_conv<n>_w/_b`, `val_ip<n>_w/_b`;
BatchNorm and other layers' blobs take the current number,
`val_bn<n>_w<k>`, `val_sc<n>_w<k>`) and the values as an ostream prints a
float (6 significant digits, %g). `batchnorm_affine` gives the per-channel
scale and shift of a BatchNorm layer at inference (Caffe's BatchNorm with
global statistics: mean and variance divided by the stored scale factor).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import BinaryIO, Iterator, List, Optional, Tuple, Union

import numpy as np

_VARINT, _I64, _LEN, _I32 = 0, 1, 2, 5


def _varint(b: bytes, i: int) -> Tuple[int, int]:
    r = s = 0
    while True:
        if i >= len(b):
            raise ValueError("truncated varint")
        c = b[i]
        i += 1
        r |= (c & 0x7F) << s
        s += 7
        if c < 0x80:
            return r, i


def _fields(b: bytes) -> Iterator[Tuple[int, int, object]]:
    """(field number, wire type, value) of one message; length-delimited
    values are bytes, fixed-width ones raw bytes, varints ints."""
    i = 0
    while i < len(b):
        key, i = _varint(b, i)
        f, wt = key >> 3, key & 7
        if wt == _VARINT:
            v, i = _varint(b, i)
        elif wt == _I64:
            v, i = b[i:i + 8], i + 8
        elif wt == _I32:
            v, i = b[i:i + 4], i + 4
        elif wt == _LEN:
            n, i = _varint(b, i)
            v, i = b[i:i + n], i + n
        else:
            raise ValueError(f"wire type {wt} (groups are not used by Caffe)")
        if i > len(b):
            raise ValueError("truncated message")
        yield f, wt, v


def _ints(wt: int, v) -> List[int]:
    if wt == _VARINT:
        return [v]
    out, j = [], 0
    while j < len(v):
        x, j = _varint(v, j)
        out.append(x)
    return out


def _floats(wt: int, v) -> np.ndarray:
    if wt == _I32:
        return np.frombuffer(v, "<f4")
    if wt == _LEN:
        if len(v) % 4:
            raise ValueError("packed float field of odd length")
        return np.frombuffer(v, "<f4")
    raise ValueError(f"float field with wire type {wt}")


@dataclass
class Layer:
    name: str
    type: str
    blobs: List[np.ndarray] = field(default_factory=list)
    bn_eps: float = 1e-5


@dataclass
class CaffeNet:
    name: str
    layers: List[Layer]

    def weighted(self) -> List[Layer]:
        return [la for la in self.layers if la.blobs]

    def total_weights(self) -> int:
        return sum(b.size for la in self.layers for b in la.blobs)


def _blob(msg: bytes) -> np.ndarray:
    shape: Optional[List[int]] = None
    legacy = [None, None, None, None]
    data: List[np.ndarray] = []
    for f, wt, v in _fields(msg):
        if f == 7 and wt == _LEN:
            shape = []
            for g, wt2, u in _fields(v):
                if g == 1:
                    shape += _ints(wt2, u)
        elif f in (1, 2, 3, 4) and wt == _VARINT:
            legacy[f - 1] = v
        elif f == 5:
            data.append(_floats(wt, v))
    values = (np.concatenate(data) if data else np.zeros(0, np.float32))
    if shape is None:
        shape = [d for d in legacy if d is not None] or [values.size]
    count = int(np.prod(shape)) if shape else 0
    if count != values.size:
        raise ValueError(f"blob of shape {shape} holds {values.size} values")
    return values.astype(np.float32).reshape(shape)


def _layer(msg: bytes) -> Layer:
    name = typ = ""
    blobs: List[np.ndarray] = []
    eps = 1e-5
    for f, wt, v in _fields(msg):
        if f == 1 and wt == _LEN:
            name = v.decode()
        elif f == 2 and wt == _LEN:
            typ = v.decode()
        elif f == 7 and wt == _LEN:
            blobs.append(_blob(v))
        elif f == 139 and wt == _LEN:
            for g, wt2, u in _fields(v):
                if g == 3 and wt2 == _I32:
                    eps = struct.unpack("<f", u)[0]
    return Layer(name, typ, blobs, eps)


def read_caffemodel(src: Union[str, bytes]) -> CaffeNet:
    """The layers of a .caffemodel (path or bytes), in file order."""
    b = src if isinstance(src, (bytes, bytearray)) else open(src, "rb").read()
    name, layers = "", []
    for f, wt, v in _fields(bytes(b)):
        if f == 1 and wt == _LEN:
            name = v.decode()
        elif f == 100 and wt == _LEN:
            layers.append(_layer(v))
    return CaffeNet(name, layers)


def shape_string(blob: np.ndarray) -> str:
    """Caffe's Blob::shape_string: each dim and a space, then (count)."""
    return "".join(f"{d} " for d in blob.shape) + f"({blob.size})"


def _float_text(x: np.float32) -> str:
    return "%g" % float(x)


def write_weights(net: CaffeNet, out: Union[str, BinaryIO, None] = None) -> str:
    """The WRITE_WEIGHTS text of Network.cpp:275-339 for `net`; written to
    `out` (a path or a binary file) when given, and returned."""
    text = ["#include <array>\n\n"]
    count = 0
    for la in net.layers:
        for k, blob in enumerate(la.blobs):
            text.append(f"// {shape_string(blob)}\n")
            n = blob.size
            if la.type == "Convolution":
                if k == 0:
                    count += 1
                    text.append(f"std::array<float, {n}> val_conv{count}_w = {{{{\n")
                else:
                    text.append(f"std::array<float, {n}> val_conv{count}_b = {{{{\n")
            elif la.type == "BatchNorm":
                text.append(f"std::array<float, {n}> val_bn{count}_w{k + 1} = {{{{\n")
            elif la.type == "InnerProduct":
                if k == 0:
                    count += 1
                    text.append(f"std::array<float, {n}> val_ip{count}_w = {{{{\n")
                else:
                    text.append(f"std::array<float, {n}> val_ip{count}_b = {{{{\n")
            else:
                text.append(f"std::array<float, {n}> val_sc{count}_w{k + 1} = {{{{\n")
            vals = blob.reshape(-1)
            text.append(", ".join(_float_text(x) for x in vals))
            text.append(" }};\n\n")
    s = "".join(text)
    if isinstance(out, str):
        with open(out, "w") as fh:
            fh.write(s)
    elif out is not None:
        out.write(s.encode())
    return s


def batchnorm_affine(layer: Layer) -> Tuple[np.ndarray, np.ndarray]:
    """(scale, shift) per channel with y = scale * x + shift equal to
    Caffe's BatchNorm at inference: mean and variance are the stored sums
    times 1 / blobs[2][0] (0 when that factor is 0), then
    y = (x - mean) / sqrt(var + eps)."""
    if layer.type != "BatchNorm" or len(layer.blobs) < 3:
        raise ValueError(f"{layer.name}: not a BatchNorm layer with 3 blobs")
    f = float(layer.blobs[2].reshape(-1)[0])
    k = 0.0 if f == 0.0 else 1.0 / f
    mean = layer.blobs[0].astype(np.float64) * k
    var = layer.blobs[1].astype(np.float64) * k
    inv = 1.0 / np.sqrt(var + layer.bn_eps)
    return inv, -mean * inv


# ---- writing .caffemodel bytes (test fixtures) ---------------------------

def _key(f: int, wt: int) -> bytes:
    return _enc_varint((f << 3) | wt)


def _enc_varint(x: int) -> bytes:
    out = bytearray()
    while True:
        b = x & 0x7F
        x >>= 7
        if x:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def _len_field(f: int, payload: bytes) -> bytes:
    return _key(f, _LEN) + _enc_varint(len(payload)) + payload


def encode_caffemodel(net: CaffeNet, legacy_dims: bool = False,
                      packed: bool = True) -> bytes:
    """A .caffemodel holding `net` (shape messages, or the legacy
    num/channels/height/width fields; packed or one-per-value floats)."""
    out = bytearray(_len_field(1, net.name.encode()))
    for la in net.layers:
        m = bytearray(_len_field(1, la.name.encode()))
        m += _len_field(2, la.type.encode())
        for blob in la.blobs:
            bm = bytearray()
            if legacy_dims:
                dims = list(blob.shape)
                for i, d in enumerate(dims[:4]):
                    bm += _key(i + 1, _VARINT) + _enc_varint(int(d))
            else:
                dims = b"".join(_enc_varint(int(d)) for d in blob.shape)
                bm += _len_field(7, _len_field(1, dims))
            vals = np.ascontiguousarray(blob, "<f4").reshape(-1)
            if packed:
                bm += _len_field(5, vals.tobytes())
            else:
                for x in vals:
                    bm += _key(5, _I32) + struct.pack("<f", float(x))
            m += _len_field(7, bytes(bm))
        if la.type == "BatchNorm":
            m += _len_field(139, _key(3, _I32) + struct.pack("<f", la.bn_eps))
        out += _len_field(100, bytes(m))
    return bytes(out)


def summary(net: CaffeNet) -> str:
    """The layer listing the reference prints while loading (Network.cpp:
    287-301): one line per layer with its blobs' shape strings, then the
    total."""
    out = [f"{len(net.layers)} layers:"]
    for i, la in enumerate(net.layers, 1):
        line = f"layer {i} ({la.type})"
        if la.blobs:
            line += " = " + " + ".join(shape_string(b) for b in la.blobs) + " "
        out.append(line)
    out.append(f"{net.total_weights()} total DCNN weights")
    return "\n".join(out)


# ---- running a Caffe net on the library's kernels -------------------------

ACT_ELU, ACT_RELU, ACT_NONE = 0, 1, 2  # LZ_ACT_* (include/leela_amd.h)
_ACTS = {"ReLU": ACT_RELU, "ELU": ACT_ELU}


def layer_table(net: CaffeNet) -> List[dict]:
    """The net as the library's layer list, in the order Caffe runs it:
    one dict per Convolution / InnerProduct (kind 'conv' / 'fc', filter,
    cin, cout, w, b), with the activation layer that follows a conv in
    place ('act': ACT_*, ACT_NONE when none follows) and the BatchNorm
    after it ('bn': (means, variances, scale factor) or None). Data,
    Flatten, loss and Silence layers carry nothing to run."""
    out: List[dict] = []
    for la in net.layers:
        if la.type == "Convolution":
            w = la.blobs[0]
            if w.ndim != 4 or w.shape[2] != w.shape[3]:
                raise ValueError(f"{la.name}: weights of shape {w.shape}")
            b = (la.blobs[1].reshape(-1) if len(la.blobs) > 1 else
                 np.zeros(w.shape[0], np.float32))
            out.append(dict(kind="conv", name=la.name, filter=w.shape[2],
                            cin=w.shape[1], cout=w.shape[0], w=w, b=b,
                            act=ACT_NONE, bn=None))
        elif la.type == "InnerProduct":
            w = la.blobs[0].reshape(la.blobs[0].shape[0], -1)
            b = (la.blobs[1].reshape(-1) if len(la.blobs) > 1 else
                 np.zeros(w.shape[0], np.float32))
            out.append(dict(kind="fc", name=la.name, filter=0,
                            cin=w.shape[1], cout=w.shape[0], w=w, b=b,
                            act=ACT_NONE, bn=None))
        elif la.type in _ACTS:
            if not out or out[-1]["kind"] != "conv" or out[-1]["bn"]:
                raise ValueError(f"{la.name}: activation not after a conv")
            out[-1]["act"] = _ACTS[la.type]
        elif la.type == "BatchNorm":
            if not out or out[-1]["kind"] != "conv" or out[-1]["bn"]:
                raise ValueError(f"{la.name}: BatchNorm not after a conv")
            if abs(la.bn_eps - 1e-5) > 1e-12:
                raise ValueError(f"{la.name}: eps {la.bn_eps} (the reference "
                                 "normalises with 1e-5)")
            out[-1]["bn"] = (la.blobs[0].reshape(-1), la.blobs[1].reshape(-1),
                             float(la.blobs[2].reshape(-1)[0]))
        elif la.type in ("Data", "Input", "Flatten", "Softmax",
                         "SoftmaxWithLoss", "Silence", "Slice"):
            continue
        else:
            raise ValueError(f"{la.name}: layer type {la.type} is not run")
    return out


def push_layers(backend, which: int, table: List[dict]) -> None:
    """Pushes a layer table (layer_table) to net `which` of an unfinalized
    leela_amd.backend.Backend: each conv, its activation and its batch
    normalisation, in order."""
    for L in table:
        if L["kind"] == "conv":
            backend.push_convolve(which, L["filter"], L["cin"], L["cout"],
                                  L["w"], L["b"])
            backend.push_activation(which, L["act"])
            if L["bn"] is not None:
                m, v, f = L["bn"]
                backend.push_batchnorm(which, m, v, f)
        else:
            backend.push_innerproduct(which, L["cin"], L["cout"], L["w"],
                                      L["b"])


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="Read a .caffemodel and write "
                                 "its WRITE_WEIGHTS arrays")
    ap.add_argument("model")
    ap.add_argument("out", nargs="?", default="weights.txt")
    a = ap.parse_args(argv)
    net = read_caffemodel(a.model)
    print(summary(net))
    write_weights(net, a.out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
