"""The whole post path on the GPU (process_stream_kernel, csrc/data/src/stream.cu:7-159) against the
oracle: every tonemap method, the Reinhard fix, the tm_alpha blend, two white points, white-balance
temperatures on both branches of the Kelvin fit and at the clamps, tints of 0, 1 and -40, gamma on and
off.  Bytes and floats must be equal (same IEEE fp32 operations; the Kelvin ratios come from the host's
libm on both sides)."""
import itertools

import numpy as np
import pytest

from nectarrender_amd import abi

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _image(H=48, W=64):
    rng = np.random.default_rng(5)
    img = (rng.lognormal(-0.5, 1.2, (3, H, W))).astype(np.float32)
    img[:, 0, :8] = np.float32([-0.5, 0.0, 1e-9, 0.5, 1.0, 7.0, 1e6, np.inf])     # edge values
    img[0, 1, 0] = np.nan
    img[1, 1, 1] = -np.inf
    return img


@pytest.mark.parametrize("tm_method,fix", [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1)])
def test_post_grid_equals_oracle(ctx, oracle, tm_method, fix):
    img = _image()
    H, W = img.shape[1:]
    d_img = torch.from_numpy(img).cuda()
    d_u8 = torch.empty((3, H, W), dtype=torch.uint8, device="cuda")
    d_f = torch.empty((3, H, W), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    n = 0
    for alpha, white, kelvin, tint, gamma, wb, tm in itertools.product(
            (0.0, 0.5, 1.0), (1.0, 4.0), (500.0, 1000.0, 3000.0, 6500.0, 6600.0, 12000.0, 40000.0),
            (0.0, 1.0, -40.0), (True, False), (True, False), (True,)):
        d = abi.post_desc(linear_to_gamma=gamma, apply_white_balance=wb, wb_temperature=kelvin, wb_tint=tint,
                          apply_tonemapping=tm, tm_method=tm_method, tm_white_point=white, tm_alpha=alpha,
                          fix_reinhard=bool(fix))
        ctx.post_device(d_img.data_ptr(), H, W, d, d_u8.data_ptr(), d_f.data_ptr())
        ref_u8, ref_f = oracle.post(img, d)
        case = (tm_method, fix, alpha, white, kelvin, tint, gamma, wb)
        np.testing.assert_array_equal(d_u8.cpu().numpy(), ref_u8, err_msg=str(case))
        got_f = d_f.cpu().numpy()
        assert np.array_equal(got_f, ref_f), case
        n += 1
    assert n == 3 * 2 * 7 * 3 * 2 * 2


def test_post_of_a_view_layer_without_tonemapping(ctx, oracle):
    """nr_post on a rendered layer with tonemapping off (white balance and gamma only)."""
    from nectarrender_amd import scenes
    s, cam, rp = scenes.mixed_scene(spp=4, res=40)
    ctx.upload(s)
    ctx.set_camera(0, cam, rp["seed"])
    ctx.reset(0)
    ctx.render(0, abi.render_desc(1, 4, rp["depth"], rp["seed"]))
    for kelvin, tint in ((3000.0, -40.0), (12000.0, 0.0)):
        d = abi.post_desc(apply_tonemapping=False, wb_temperature=kelvin, wb_tint=tint)
        u8, f = ctx.post(0, d)
        ref_u8, ref_f = oracle.post(ctx.read(0), d)
        np.testing.assert_array_equal(u8, ref_u8)
        assert np.array_equal(f, ref_f)
