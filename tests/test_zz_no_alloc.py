"""nr_trace_closest_device_noalloc: the entry point that allocates nothing.

It traces with its traversal-stack spill and counters in the library's per-GPU
static array (16 MiB), in chunks of as many rays as that array holds spill
columns for.  Its hits must be the bytes nr_trace_closest_device writes for the
same rays, on two calls in a row, the second with fewer rays (and a count that
is not a multiple of the 128-ray workgroup), and it must not write past n hits.
"""
import numpy as np
import pytest

from nectarrender_amd import abi, scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HIT_BYTES = 64


def _scene():
    v, t = scenes.displaced_icosphere(freq=96)                 # 184,320 triangles: a deep mesh tree
    s = abi.SceneDesc()
    m = s.add_mesh(v, t)
    s.mesh(m, (0.1, 0, 0), (10, 20, 30), (1, 1.2, 0.9), abi.Material.lambertian())
    s.sphere((1.6, 0.2, 0), 0.4, abi.Material.dielectric())
    s.quad((0, -1.3, 0), (0, 0, 0), (6, 1, 6), abi.Material.lambertian())
    return s


def test_noalloc_trace_equals_device_trace(ctx):
    ctx.upload(_scene())
    _, _, depth = ctx.mesh_tree(0)
    assert depth > 16, "the mesh tree must be deep enough for the stack to spill past LDS"
    for call, n in enumerate((600_000, 70_001)):
        rays = scenes.random_rays(n, 21 + call, spread=2.4)
        print(f"no-alloc trace: call {call}, {n} rays, mesh depth {depth}", flush=True)
        d_rays = torch.from_numpy(rays).cuda()
        ref = torch.empty((n, HIT_BYTES), dtype=torch.uint8, device="cuda")
        got = torch.full((n + 300, HIT_BYTES), 0xAB, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        ctx.trace_closest_device(d_rays.data_ptr(), n, ref.data_ptr())
        ctx.trace_closest_device_noalloc(d_rays.data_ptr(), n, got.data_ptr())
        ctx.sync()
        assert torch.equal(got[:n], ref), f"call {call}: hits differ from nr_trace_closest_device"
        assert bool((got[n:] == 0xAB).all()), f"call {call}: wrote past the last hit"
        hits = ref.cpu().numpy().view(abi.HIT_DTYPE).reshape(-1)
        assert (hits["object_id"] >= 0).mean() > 0.2


def test_noalloc_trace_rejects_bad_arguments(ctx):
    ctx.upload(_scene())
    with pytest.raises(Exception):
        ctx.trace_closest_device_noalloc(0, 10, 0)
    ctx.trace_closest_device_noalloc(0, 0, 0)                   # nothing to trace: nothing read or written
