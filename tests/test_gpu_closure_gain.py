"""What the learned closure gains (the closure workflow of
powerllel_channel_sycl_amd/closure_pipeline.py at Re_tau~395): a fine DNS
384x256x192 (FP32-mixed) from a perturbed laminar start, 200 box-filtered
snapshots (factors 4x4x3) feeding the training set and the filtered-DNS
statistics, Adam training of the perceptron on the GPU, then two coarse
96x64x64 runs from the filtered DNS state -- without a model and with the
closure in the stage kernels -- compared with the filtered DNS mean profile,
Reynolds stresses and wall shear.  The runs are deterministic (fixed seeds,
fixed-order reductions), so the comparison is reproducible on one GPU."""
import json

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(300)
def test_closure_gains_over_no_model():
    from powerllel_channel_sycl_amd import closure_pipeline as cp
    print("CASE closure gain Re_tau~395", flush=True)
    res = cp.run(cp.Config())
    summary = {k: res[k] for k in ("dns", "labels", "training", "coarse")}
    print(json.dumps(summary, indent=1), flush=True)
    dns, tr, co = res["dns"], res["training"], res["coarse"]
    # the fine DNS is turbulent at the target friction Reynolds number
    assert 360 < dns["re_tau"] < 420
    # training on the GPU generalises to held-out snapshots
    assert tr["val_loss"] < 0.6 * tr["val_loss_before"]
    assert tr["val_loss"] < tr["label_variance"]
    nm, cl = co["no_model"], co["closure"]
    assert nm["finite"] and cl["finite"]
    assert 0.05 < cl["nut_over_nu_mean"] < 5.0
    # the gain: the closure brings the coarse statistics closer to the filtered
    # DNS than the same coarse run without it
    for k in ("U", "re_tau", "uu", "vv", "ww", "uw"):
        assert cl["errors"][k] < nm["errors"][k], (k, cl["errors"][k], nm["errors"][k])
    assert cl["errors"]["U"] < 0.5 * nm["errors"]["U"]
    assert cl["errors"]["re_tau"] < 0.5 * nm["errors"]["re_tau"]
