"""Writes the traversal golden vectors of tests/golden/ from the reference's
own CPU traversal (lib/bvh, compiled into oracle/_ref by oracle/build_ref.sh).

    python tests/golden/make_golden.py

Fixtures (inputs and expected outputs only):
  bvhref_cornell.npz  rays against lib/bvh's test scene cornell_box.obj and
                      lib/bvh's closest hit (prim, t, u, v) plus per-ray
                      traversal statistics (steps, intersections)
  bvhref_field.npz    the same for a 64-instance field of the seeded
                      256-triangle sphere (world-space flattened triangles)

cornell_box.obj and cornell_box_reference.png are lib/bvh's own test data
(lib/bvh/test/scene/), copied as they are.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from btrc_amd import _abi, scenes  # noqa: E402
from oracle import oracle  # noqa: E402


def random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, dtype=_abi.ray_dtype())
    rays["o"] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["d"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = np.float32(3.4028234663852886e38)
    return rays


def flatten(scene):
    out = []
    for inst in scene.instances:
        p = scene.geometries[inst["geometry"]]["positions"].reshape(-1, 3).astype(np.float32)
        m = inst["mat"].astype(np.float32)
        # CTransform3D::apply_to_point order (utils/cmath/ctransform.cpp:31-37)
        w = ((m[:, 0] * p[:, :1] + m[:, 1] * p[:, 1:2]) + m[:, 2] * p[:, 2:3]) + m[:, 3]
        out.append(w[:, :3].reshape(-1, 9))
    return np.concatenate(out).astype(np.float32)


def field_scene():
    return scenes.sphere_field(instances=64, lights=0, seed=7, extent=20.0)[0]


def main():
    tris = scenes.load_obj_triangles(os.path.join(HERE, "cornell_box.obj"))
    rays = random_rays(4096, [-1, 0, -1], [1, 2, 1], 11)
    prim, t, u, v, st, isc = oracle.RefBvh(tris).trace(rays)
    np.savez_compressed(os.path.join(HERE, "bvhref_cornell.npz"), rays=rays, prim=prim, t=t, u=u, v=v,
                        steps=st, isects=isc)

    world = flatten(field_scene())
    rays = random_rays(4096, [0, 0, 0], [20, 20, 20], 12)
    prim, t, u, v, st, isc = oracle.RefBvh(world).trace(rays)
    np.savez_compressed(os.path.join(HERE, "bvhref_field.npz"), rays=rays, prim=prim, t=t, u=u, v=v,
                        steps=st, isects=isc, world=world)


if __name__ == "__main__":
    main()
