"""Host side of the learned AO surrogate (no GPU): the RAOW weights file layout that
raw_ao_net_load reads, the parameter initialisation and the learning-rate schedule."""
import numpy as np

import raw_ao.learned as L
import unet_oracle as U


def test_weights_file_layout(tmp_path):
    p = L.init_params(1)
    f = str(tmp_path / "w.raow")
    L.save_weights(f, p, {"mode": "surrogate", "steps": 3})
    raw = open(f, "rb").read()
    assert raw[:4] == b"RAOW"
    assert np.frombuffer(raw, "<u4", 1, 4)[0] == 1
    assert np.frombuffer(raw, "<u8", 1, 8)[0] == p.size
    mlen = np.frombuffer(raw, "<u4", 1, 16)[0]
    assert len(raw) == 20 + 4 * p.size + mlen
    assert np.array_equal(np.frombuffer(raw, "<f4", p.size, 20), p)
    q, meta = L.load_weights(f)
    assert np.array_equal(q, p) and meta == {"mode": "surrogate", "steps": 3}


def test_layer_table_and_init_match_the_network():
    assert [(c0, c1, co) for c0, c1, co, _ in U.LAYERS] == L.LAYERS
    p = L.init_params(2, head_bias=0.7)
    assert p.size == U.param_count()
    layers, (hw, hb) = U.unpack(p)
    assert hb == np.float32(0.7)
    for (w, b), (c0, c1, co) in zip(layers, L.LAYERS):
        assert not b.any()
        assert abs(w.std() / np.sqrt(2.0 / (9 * (c0 + c1))) - 1) < 0.1


def test_lr_schedule():
    assert L.lr_at(0.0, 1e-3) == 0.0
    assert np.isclose(L.lr_at(0.02, 1e-3), 1e-3)
    assert np.isclose(L.lr_at(1.0, 1e-3), 2e-5)
    xs = [L.lr_at(f, 1e-3) for f in np.linspace(0.02, 1, 50)]
    assert all(a >= b for a, b in zip(xs, xs[1:]))
