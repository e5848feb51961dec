"""The committed learned-estimator weights match the network's architecture (CPU only)."""

import numpy as np


def test_committed_weights_match_architecture():
    from hermespy_amd.configs import mimo2x2_512_pilots
    from hermespy_amd.learned import SKIP_TAPS, TAPS, WEIGHTS_PATH, layer_shapes, skip_shape

    cfg = mimo2x2_512_pilots()
    with np.load(WEIGHTS_PATH, allow_pickle=False) as f:
        for i, (cin, cout, _) in enumerate(layer_shapes(cfg)):
            assert f[f"w{i}"].shape == (cout, TAPS, cin) and f[f"b{i}"].shape == (cout,)
        cin, cout = skip_shape(cfg)
        assert f["w4"].shape == (cout, SKIP_TAPS, cin) and f["b4"].shape == (cout,)
        assert all(np.isfinite(f[k]).all() for k in f.files)
