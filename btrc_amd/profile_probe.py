"""Workloads for rocprofv3 counter calibration and capture.

    python btrc_amd/profile_probe.py copy [--mib 1024]
        one device-to-device copy of a known size (torch's copy kernel), to
        scale FETCH_SIZE / WRITE_SIZE readings to bytes
    python btrc_amd/profile_probe.py frame [--spp 16]
        renders the 1080p frame of the 1M-triangle scene once and prints the
        number of extension rays, so counters summed over the extension-trace
        dispatches divide by a known ray count
    python btrc_amd/profile_probe.py batch [--rays 4194304] [--repeats 3]
        captures the extension rays of the first two bounces of the 1080p
        frame of the 1M-triangle scene, then traces them as one resident
        batch per repeat (k_trace_batch), so counters per dispatch divide by
        a known ray count
"""
import argparse
import os
import sys

if __name__ == "__main__" and __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def copy(mib):
    import torch
    n = mib * 1024 * 1024 // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy: {mib} MiB read + {mib} MiB written", flush=True)


def frame(spp):
    import btrc_amd
    from btrc_amd import _abi, scenes
    scene, cam = scenes.sphere_field()
    ctx = btrc_amd.Context(0)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 1920, 1080)
    p = _abi.RenderParams()
    p.width, p.height, p.spp = 1920, 1080, spp
    p.min_depth, p.max_depth = 3, 8
    p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
    p.state_count = 1 << 21
    st = ctx.render(p)
    print(f"frame: extension_rays {st.extension_rays} shadow_rays {st.shadow_rays} iterations {st.iterations}",
          flush=True)


def batch(rays, repeats):
    import numpy as np
    import btrc_amd
    from btrc_amd import _abi, scenes
    scene, cam = scenes.sphere_field()
    ctx = btrc_amd.Context(0)
    ctx.upload_scene(scene)
    ctx.set_camera(cam, 1920, 1080)
    p = _abi.RenderParams()
    p.width, p.height, p.spp = 1920, 1080, 64
    p.min_depth, p.max_depth = 3, 8
    p.rr_threshold, p.rr_cont_prob = 0.1, 0.5
    p.state_count = 1 << 21
    r = ctx.capture_rays(p, 2, rays)
    r = r[np.random.default_rng(0).permutation(len(r))]
    ctx.batch_upload(r)
    for k in range(repeats):
        ms = ctx.batch_trace(_abi.TRI_WATERTIGHT, False, 1)
        print(f"batch: {len(r)} rays in {ms:.3f} ms ({len(r) / ms / 1e3:.1f} Mrays/s)", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["copy", "batch", "frame"])
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--rays", type=int, default=1 << 22)
    ap.add_argument("--repeats", type=int, default=3)
    a = ap.parse_args()
    if a.mode == "copy":
        copy(a.mib)
    elif a.mode == "frame":
        frame(a.spp)
    else:
        batch(a.rays, a.repeats)


if __name__ == "__main__":
    sys.exit(main())
