"""The accumulated image after 4096 samples per pixel against the oracle's samples accumulated in float64.

The GPU folds every sample into a float32 running mean (data.cu:28-40: m = m (n-1)/n + v/n); the
oracle renders the same 4096 samples (each pass bit-exact with the GPU's) and sums them in float64.
The only difference left is the running mean's float32 rounding over 4096 folds, so the stated
bound is tight: relative RMSE at most 1e-5 and no pixel off by more than 1e-4 relative (the oracle's
own float32 running mean of the same samples stays inside the same bounds)."""
import numpy as np
import pytest

from nectarrender_amd import abi, scenes

pytestmark = pytest.mark.gpu

N = 4096


@pytest.mark.parametrize("res", [24])
def test_4096_sample_accumulation_against_fp64(ctx, oracle, res):
    s, cam, rp = scenes.mixed_scene(spp=N + 1, res=res)     # passes 1..N index an (N + 1)-sample pattern
    desc = abi.render_desc(1, N + 1, rp["depth"], rp["seed"])
    ctx.upload(s)
    ctx.set_camera(0, cam, rp["seed"])
    ctx.reset(0)
    ctx.render(0, desc)
    layers = (abi.LAYER_BEAUTY, abi.LAYER_DIFFUSE, abi.LAYER_WORLD_NORMAL)
    gpu = [ctx.read(0, L).astype(np.float64) for L in layers]

    osc, ocam = oracle.OracleScene(s), oracle.OracleCamera(cam, rp["seed"])
    acc = [np.zeros((3, res, res), np.float64) for _ in layers]
    for k in range(1, N + 1):
        for a, img in zip(acc, oracle.render_pass(osc, ocam, desc, k)):
            a += img
    ref = [a / N for a in acc]
    run32 = [x.astype(np.float64) for x in oracle.render(osc, ocam, desc)]

    for name, g, r, o32 in zip(("beauty", "diffuse", "normal"), gpu, ref, run32):
        assert np.array_equal(np.isnan(g), np.isnan(r)), name
        ok = ~np.isnan(r)
        g, r = g[ok], r[ok]
        scale = max(float(np.abs(r).mean()), 1e-6)
        rel_rmse = float(np.sqrt(np.mean((g - r) ** 2))) / scale
        rel_max = float(np.max(np.abs(g - r) / np.maximum(np.abs(r), scale)))
        print(f"{name}: relative RMSE {rel_rmse:.2e}, max {rel_max:.2e} against the float64 sum")
        assert rel_rmse <= 1e-5 and rel_max <= 1e-4, name
        assert np.array_equal(g, o32[ok]), name      # the float32 running mean is the reference's, bit for bit
