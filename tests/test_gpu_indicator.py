"""Learned refinement indicator on the coastline workload (39,280 elements,
N=3): features from the GPU modal kernel, labels from the GPU estimator
(bit-exact to the reference's GradientErrorEstimator restatement), training
by the project's own BF16 forward/backward kernels.

What the trained network gains over the method without it: the untrained
network (the same kernels, no training) and the best constant prediction
both misclassify far more elements than the trained one; the trained
network's misclassification against the estimator's decisions on unseen
seeds is reported and bounded."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_indicator_learns_estimator_decisions():
    import torch

    from drifter_amd import indicator, mlp
    from drifter_amd.mesh import coastline_mesh

    mesh = coastline_mesh(band=7000.0)
    fields = indicator.CoastlineFields(mesh)
    print(f"[indicator] E={fields.E}", flush=True)
    Xtr, ytr, atr = indicator.build_dataset(fields, range(100, 106))
    Xte, yte, ate = indicator.build_dataset(fields, range(900, 902))
    counts = np.bincount(ate.cpu().numpy(), minlength=3)
    print(f"  test decisions none/refine/coarsen: {counts.tolist()}", flush=True)
    model = mlp.IndicatorMLP(seed=11)
    before = indicator.misclassification(model, Xte, ate)
    loss = indicator.train(model, Xtr, ytr, epochs=30, batch=2048, lr=2e-3)
    after = indicator.misclassification(model, Xte, ate)
    train_err = indicator.misclassification(model, Xtr, atr)
    const = 1.0 - counts.max() / counts.sum()
    print(f"  misclassification: untrained {before:.4f}, constant {const:.4f}, "
          f"trained {after:.4f} (train set {train_err:.4f}, train loss {loss:.4f})", flush=True)
    assert after < 0.5 * min(before, const)
    assert after < 0.15
