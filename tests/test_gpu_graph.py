"""Streaming iterations captured as HIP graphs (tuning field graph_iterations).

A render in graph mode replays a per-stream-group graph of several iterations whose
grids cover the group's whole path capacity; a later render with the same launch
arguments replays the same cached graph.  Every replay must give the AOVs of the
host-launched loop bit for bit: renders in graph mode are interleaved with host-
launched renders and synchronisations, several times over, and each is compared with
the host-launched result (and that with the oracle).
"""
import numpy as np
import pytest

from nectarrender_amd import abi, scenes

pytestmark = pytest.mark.gpu

LAYERS = (abi.LAYER_BEAUTY, abi.LAYER_DIFFUSE, abi.LAYER_WORLD_NORMAL, abi.LAYER_DEPTH)


def _same(a, b):
    return bool(((a == b) | (np.isnan(a) & np.isnan(b))).all())


def _render(ctx, view, desc):
    ctx.reset(view)
    ctx.render(view, desc)
    ctx.sync()
    return [ctx.read(view, L) for L in LAYERS]


@pytest.mark.parametrize("mode", [abi.MODE_PARITY, abi.MODE_FAST])
def test_graph_replays_equal_host_loop(ctx, oracle, mode):
    s, cam, rp = scenes.mixed_scene(spp=9, res=80)
    ctx.upload(s)
    ctx.set_camera(0, cam, rp["seed"])
    p0 = 1 if mode == abi.MODE_PARITY else 0
    desc = abi.render_desc(p0, 9, rp["depth"], rp["seed"], mode=mode)
    try:
        ctx.set_tuning(4, 2, 8, 0, 0)
        eager = _render(ctx, 0, desc)
        if mode == abi.MODE_PARITY:
            ref = oracle.render(oracle.OracleScene(s), oracle.OracleCamera(cam, rp["seed"]), desc, depth=True)
            for k in range(4):
                assert _same(eager[k], ref[k]), f"host loop vs oracle, layer {k}"
        for rep in range(3):
            for iters in (2, 8):
                ctx.set_tuning(4, 2, 8, 0, iters)
                got = _render(ctx, 0, desc)          # the second round replays the cached graphs
                for k in range(4):
                    assert _same(got[k], eager[k]), f"replay {rep}, {iters} iterations per graph, layer {k}"
                ctx.set_tuning(4, 2, 8, 0, 0)        # host-launched work between the replays
                again = _render(ctx, 0, desc)
                for k in range(4):
                    assert _same(again[k], eager[k]), f"host loop after replay {rep}, layer {k}"
    finally:
        ctx.set_tuning(4, 2, 8)


def test_graph_mode_with_views_tiles_and_edits(ctx, oracle):
    """Several views in one call, per-tile sample limits, and a scene edit between two renders with
    otherwise equal arguments (the cached graph is replayed against the edited scene tables)."""
    s, cam, rp = scenes.mixed_scene(spp=6, res=64)
    ctx.upload(s)
    cam2 = abi.camera_desc((40, 56), (0.5, 1.0, 4.0), (-8.0, 5.0, 0.0), 6, 2.0, 4.0, 0.0, 2.0, 1.0)
    ctx.set_camera(0, cam, rp["seed"])
    ctx.set_camera(1, cam2, rp["seed"])
    desc = abi.render_desc(1, 6, rp["depth"], rp["seed"])
    tiles = np.array([[6, 3], [1, 6]], np.uint32)
    try:
        results = {}
        for iters in (0, 4, 4):
            ctx.set_tuning(4, 2, 8, 262144, iters)
            for v in (0, 1):
                ctx.reset(v)
            ctx.render_views([0, 1], desc)
            ctx.sync()
            views = [ctx.read(v) for v in (0, 1)]
            ctx.reset(0)
            ctx.render_tiles(0, desc, tiles, 32)
            ctx.sync()
            results.setdefault("views", []).append(views)
            results.setdefault("tiles", []).append(ctx.read(0))
        for k in (1, 2):
            assert all(_same(a, b) for a, b in zip(results["views"][k], results["views"][0])), f"views, run {k}"
            assert _same(results["tiles"][k], results["tiles"][0]), f"tiles, run {k}"
        # an edit between two graph-mode renders with the same arguments
        ctx.set_tuning(4, 2, 8, 262144, 4)
        before = _render(ctx, 0, desc)
        ctx.set_object_transform(0, (0.3, 0.0, 0.0), (0.0, 20.0, 0.0), (1.0, 1.0, 1.0))
        after = _render(ctx, 0, desc)
        ctx.set_tuning(4, 2, 8, 262144, 0)
        eager_after = _render(ctx, 0, desc)
        for k in range(4):
            assert _same(after[k], eager_after[k]), f"edited scene, layer {k}"
        assert not _same(before[0], after[0])
    finally:
        ctx.set_tuning(4, 2, 8)


def test_graph_iterations_validated(ctx):
    for bad in (1, 3, 18, -2):
        with pytest.raises(RuntimeError):
            ctx.set_tuning(4, 2, 8, 262144, bad)
    assert ctx.get_tuning()["graph_iterations"] == 0
