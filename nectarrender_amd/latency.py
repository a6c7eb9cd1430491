"""Latency of short renders of the 1080p benchmark workload.

Uploads `scenes.config2()` (1920 x 1080, 1,003,520-triangle icosphere, HDRI with a 1e4 sun,
depth 12), sets a K-sample pattern and renders passes 0..K-1 with the MIS estimator under
each (stream groups, tail threshold) tuning with `--slots` path slots (8), nine times, timing from `nr_render` to the end of
a stream synchronisation; prints the median and the minimum, then renders three more times
under the last tuning (the dispatches a kernel trace of this command ends with).

    python -m nectarrender_amd.latency [--samples 1]
"""
from __future__ import annotations

import argparse
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=1)
    ap.add_argument("--reps", type=int, default=9)
    ap.add_argument("--slots", type=int, default=8, help="path slots per pixel of every tuning")
    args = ap.parse_args(argv)
    import torch
    torch.cuda.init()
    from . import abi, native, scenes
    from .quality import with_spp
    s, cam, rp = scenes.config2()
    ctx = native.Context(0)
    ctx.upload(s)
    ctx.set_camera(0, with_spp(cam, args.samples), 11)
    d = abi.render_desc(0, args.samples, rp["depth"], 11, mode=abi.MODE_FAST)

    def render():
        ctx.reset(0)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.render(0, d)
        ctx.sync()
        return time.perf_counter() - t0

    for groups in (1, 2):
        for tail in (262144, 4194304):
            ctx.set_tuning(args.slots, groups, 16, tail, 0)
            ts = sorted(render() for _ in range(args.reps))
            print(f"samples {args.samples}, slots {args.slots}, groups {groups}, tail {tail}: median {ts[len(ts) // 2] * 1e3:.3f} ms, "
                  f"min {ts[0] * 1e3:.3f} ms", flush=True)
    ctx.set_tuning(args.slots, 2, 16, 262144, 0)
    for _ in range(3):
        render()
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
