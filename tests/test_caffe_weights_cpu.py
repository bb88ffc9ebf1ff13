"""The .caffemodel reader and the WRITE_WEIGHTS export (leela_amd/
caffe_weights.py; Network.cpp:251-341).

Pins: round trips through a .caffemodel writer covering the encodings
Caffe produces (shape messages or legacy dims, packed or unpacked floats);
the WRITE_WEIGHTS text written out by hand from the reference's stream
statements for a small net; BatchNorm's inference affine against its
definition; and, where the reference tree is present, its two shipped
models (layer shapes recorded from them)."""
import os
from pathlib import Path

import numpy as np
import pytest

from leela_amd import caffe_weights as cw

REF = Path(os.environ.get("LEELA_REFERENCE", "/root/reference"))


def small_net():
    rng = np.random.default_rng(3)
    return cw.CaffeNet("LeelaNet", [
        cw.Layer("data", "Data"),
        cw.Layer("conv1", "Convolution",
                 [rng.standard_normal((4, 2, 3, 3)).astype(np.float32),
                  rng.standard_normal(4).astype(np.float32)]),
        cw.Layer("relu1", "ReLU"),
        cw.Layer("bn1", "BatchNorm",
                 [rng.standard_normal(4).astype(np.float32),
                  rng.uniform(1, 2, 4).astype(np.float32),
                  np.array([0.999], np.float32)], bn_eps=1e-3),
        cw.Layer("ip2", "InnerProduct",
                 [rng.standard_normal((3, 4)).astype(np.float32),
                  rng.standard_normal(3).astype(np.float32)]),
    ])


@pytest.mark.parametrize("legacy,packed", [(False, True), (False, False),
                                           (True, True)])
def test_round_trip(legacy, packed):
    net = small_net()
    back = cw.read_caffemodel(cw.encode_caffemodel(net, legacy, packed))
    assert back.name == net.name
    assert [(a.name, a.type) for a in back.layers] == \
        [(a.name, a.type) for a in net.layers]
    for a, b in zip(net.layers, back.layers):
        assert len(a.blobs) == len(b.blobs)
        for x, y in zip(a.blobs, b.blobs):
            assert x.shape == y.shape and np.array_equal(x, y)
    assert back.layers[3].bn_eps == np.float32(1e-3)
    assert back.total_weights() == 72 + 4 + 9 + 12 + 3


def test_write_weights_text():
    net = cw.CaffeNet("n", [
        cw.Layer("c", "Convolution", [np.array([[[[0.5, -1.25e-5]]]],
                                               np.float32),
                                      np.array([3.0], np.float32)]),
        cw.Layer("bn", "BatchNorm", [np.array([1e7], np.float32),
                                     np.array([123.456789], np.float32),
                                     np.array([1.0], np.float32)]),
        cw.Layer("s", "Scale", [np.array([-0.0], np.float32)]),
        cw.Layer("ip", "InnerProduct", [np.array([[2.0, 0.1]], np.float32),
                                        np.array([1.0], np.float32)]),
    ])
    want = ("#include <array>\n\n"
            "// 1 1 1 2 (2)\n"
            "std::array<float, 2> val_conv1_w = {{\n0.5, -1.25e-05 }};\n\n"
            "// 1 (1)\n"
            "std::array<float, 1> val_conv1_b = {{\n3 }};\n\n"
            "// 1 (1)\n"
            "std::array<float, 1> val_bn1_w1 = {{\n1e+07 }};\n\n"
            "// 1 (1)\n"
            "std::array<float, 1> val_bn1_w2 = {{\n123.457 }};\n\n"
            "// 1 (1)\n"
            "std::array<float, 1> val_bn1_w3 = {{\n1 }};\n\n"
            "// 1 (1)\n"
            "std::array<float, 1> val_sc1_w1 = {{\n-0 }};\n\n"
            "// 1 2 (2)\n"
            "std::array<float, 2> val_ip2_w = {{\n2, 0.1 }};\n\n"
            "// 1 (1)\n"
            "std::array<float, 1> val_ip2_b = {{\n1 }};\n\n")
    assert cw.write_weights(net) == want


def test_summary_lines():
    s = cw.summary(small_net()).splitlines()
    assert s[0] == "5 layers:"
    assert s[1] == "layer 1 (Data)"
    assert s[2] == "layer 2 (Convolution) = 4 2 3 3 (72) + 4 (4) "
    assert s[-1] == "100 total DCNN weights"


def test_batchnorm_affine():
    la = small_net().layers[3]
    scale, shift = cw.batchnorm_affine(la)
    x = np.linspace(-3, 3, 7)[:, None]
    f = 1.0 / float(la.blobs[2][0])  # the stored float32 factor
    mean = la.blobs[0].astype(np.float64) * f
    var = la.blobs[1].astype(np.float64) * f
    want = (x - mean) / np.sqrt(var + la.bn_eps)
    assert np.allclose(scale * x + shift, want, rtol=1e-12, atol=1e-12)
    # A zero scale factor zeroes the statistics (Caffe's convention).
    la.blobs[2] = np.array([0.0], np.float32)
    scale, shift = cw.batchnorm_affine(la)
    assert np.allclose(scale, 1 / np.sqrt(la.bn_eps)) and np.all(shift == 0)


def test_rejects_bad_blobs():
    net = small_net()
    b = bytearray(cw.encode_caffemodel(net))
    with pytest.raises(ValueError):
        cw.read_caffemodel(bytes(b[:-3]))


@pytest.mark.skipif(not (REF / "model_4750.caffemodel").exists(),
                    reason="reference models not present")
def test_reference_models():
    net = cw.read_caffemodel(str(REF / "model_4750.caffemodel"))
    convs = [la for la in net.layers if la.type == "Convolution"]
    shapes = [tuple(la.blobs[0].shape) for la in convs]
    assert shapes == [(96, 22, 5, 5), (160, 96, 3, 3), (128, 160, 3, 3),
                      (128, 128, 3, 3), (96, 128, 3, 3), (96, 96, 3, 3),
                      (96, 96, 3, 3), (64, 96, 3, 3), (64, 64, 3, 3),
                      (1, 64, 3, 3)]
    assert sum(la.type == "BatchNorm" for la in net.layers) == 9
    text = cw.write_weights(net)
    assert "val_conv10_b" in text and "val_bn9_w3" in text
    # Values survive the 6-digit print to within its rounding.
    w = convs[0].blobs[0].reshape(-1)
    body = text.split("val_conv1_w = {{\n")[1].split(" }};")[0]
    printed = np.array([float(t) for t in body.split(", ")])
    assert printed.size == w.size
    assert np.all(np.abs(printed - w) <= 5e-6 * np.maximum(np.abs(w), 1e-30))
    small = cw.read_caffemodel(str(REF / "model_4258.caffemodel"))
    assert [tuple(la.blobs[0].shape) for la in small.layers
            if la.type == "Convolution"] == \
        [(160, 22, 3, 3), (96, 160, 3, 3), (32, 96, 3, 3), (32, 32, 3, 3),
         (1, 32, 3, 3)]
