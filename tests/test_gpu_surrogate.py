"""The learned warm start end to end on the GPU: models solved by this library's COCG, the
encoder-decoder trained on them with the library's BF16-MFMA kernels, and warm-started solves of
held-out models compared with cold (X = 0) solves of the same systems.

What this pins: training lowers the loss; every warm-started solve still reaches the same FP64
true residual; and the fallback rule (use the prediction only when its minimal-residual scaled
relative residual is at most `accept`) keeps a warm-started solve from costing more iterations
than the cold one. The measured gain is printed: on this operator it is about zero -- see the
surrogate section of the design notes for why."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_train_and_warm_start():
    from actfemtide_amd import nn, surrogate as S
    grid = (16, 16, 8)
    w = 2 * math.pi / 44714.0
    d = S.Dataset(grid, w)
    for i in range(8):
        smp = d.add(300 + i, tol=1e-5)
        print(f"training model {i}: {smp.iters} iterations to 1e-5")
    net = nn.Surrogate(base=16, cin=8, seed=0)
    log = S.train(net, d, range(8), steps=200, batch=4, lr=2e-3, log_every=50)
    print("loss", log.losses, f"{1e3 * log.seconds / log.steps:.1f} ms/step")
    first, last = log.losses[0][1], log.losses[-1][1]
    assert last < 0.5 * first
    tot_c = tot_w = 0
    for k in range(2):
        case = S.grf_case(grid, 9000 + k)
        r = S.warm_solve(net, d, case, tol=1e-8, accept=0.5)
        print({key: (round(v, 6) if isinstance(v, float) else v) for key, v in r.items()})
        assert r["true_cold"] <= 1e-8 and r["true_warm"] <= 1e-8
        if r["fallback"]:
            assert r["iters_warm"] == r["iters_cold"]
        tot_c += r["iters_cold"]
        tot_w += r["iters_warm"]
    print(f"iterations cold {tot_c} warm {tot_w}: gain {1 - tot_w / tot_c:+.3f}")
    assert tot_w <= tot_c * 1.02 + 5
