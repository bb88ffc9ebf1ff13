"""The trained nets the reference ships (model_4258 / model_4750), as layer
tables read back from tests/golden/caffe_models.npz (written by
tests/golden/make_caffe_fixture.py from the .caffemodel files)."""
from pathlib import Path

import numpy as np

FIXTURE = Path(__file__).resolve().parent / "golden" / "caffe_models.npz"
MODELS = ("model_4258", "model_4750")


def load_tables():
    """{model: [layer dict]} in the form of caffe_weights.layer_table."""
    z = np.load(FIXTURE)
    out = {}
    for m in MODELS:
        table = []
        for i in range(int(z[f"{m}/n"])):
            w = z[f"{m}/{i}/w"]
            bn = None
            if f"{m}/{i}/bn_mean" in z:
                bn = (z[f"{m}/{i}/bn_mean"], z[f"{m}/{i}/bn_var"],
                      float(z[f"{m}/{i}/bn_scale"]))
            table.append(dict(kind="conv", name=f"conv{i + 1}",
                              filter=w.shape[2], cin=w.shape[1],
                              cout=w.shape[0], w=w, b=z[f"{m}/{i}/b"],
                              act=int(z[f"{m}/{i}/act"]), bn=bn))
        out[m] = table
    return out


def planes22(n: int, seed0: int = 1, max_moves: int = 250) -> np.ndarray:
    """Inputs for the 22-plane nets: the first 22 of the 32 policy planes of
    generator positions (stones, liberties and liberties after the move;
    the older 22-plane feature set itself is not in the reference)."""
    from leela_amd.positions import make_batch
    return np.ascontiguousarray(make_batch(n, seed0=seed0,
                                           max_moves=max_moves)[:, :22])
