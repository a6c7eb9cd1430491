"""What the trained denoiser gains over the plain path tracer: at the same 4 samples per pixel, the
U-Net frame's error against a 1024-sample render (display space: white balance, ACES, gamma) is a
fraction of the noisy frame's, on a scene of each family the network was trained on."""
import os

import numpy as np
import pytest

from nectarrender_amd import abi, scenes
from nectarrender_amd.quality import display_desc, with_spp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def unet():
    from nectarrender_amd.denoise import WEIGHTS, UNet
    if not os.path.exists(WEIGHTS):
        pytest.skip("no trained weights")
    u = UNet(0).load()
    yield u
    u.close()


def _scene(kind):
    if kind == "spheres":                     # three spheres + 60 media/dielectric spheres, reference estimator
        return scenes.config5(n_spheres=60, res=(128, 128), spp=1024)
    # displaced icosphere under an HDRI with a sun (a seed outside the training set), MIS estimator
    s, cam, rp = scenes.training_scene_hdri_mesh(90_001, 128, 1024)
    rp["mode"] = abi.MODE_FAST
    return s, cam, rp


@pytest.mark.parametrize("kind", ["spheres", "hdri_mesh"])
def test_denoiser_beats_plain_at_equal_samples(ctx, unet, kind):
    s, cam, rp = _scene(kind)
    ctx.upload(s)
    H, W = int(cam.resolution[0]), int(cam.resolution[1])
    mode = rp.get("mode", abi.MODE_PARITY)
    p0 = 0 if mode == abi.MODE_FAST else 1
    post = display_desc(abi)

    def display(ptr):
        d = torch.empty((3, H, W), dtype=torch.float32, device="cuda")
        ctx.post_device(ptr, H, W, post, 0, d.data_ptr())
        return d.double()

    ctx.set_camera(0, with_spp(cam, 1024 + p0), 91)
    ctx.reset(0)
    ctx.render(0, abi.render_desc(p0, 1024 + p0, rp["depth"], 91, mode=mode))
    ref = display(ctx.aov_device_ptr(0, abi.LAYER_BEAUTY))
    ctx.set_camera(1, with_spp(cam, 4 + p0), 5)
    ctx.reset(1)
    ctx.render(1, abi.render_desc(p0, 4 + p0, rp["depth"], 5, mode=mode))
    plain = display(ctx.aov_device_ptr(1, abi.LAYER_BEAUTY))
    out = torch.empty((3, H, W), dtype=torch.float32, device="cuda")
    unet.denoise_view(ctx, 1, out)
    torch.cuda.synchronize()
    den = display(out.data_ptr())
    e_plain = float(((plain - ref) ** 2).mean().sqrt())
    e_den = float(((den - ref) ** 2).mean().sqrt())
    print(f"{kind}: 4-sample RMSE plain {e_plain:.4f}, denoised {e_den:.4f} ({e_plain / e_den:.1f}x lower)")
    assert np.isfinite(e_den)
    assert e_den < 0.5 * e_plain
