"""A small run of every oracle path, for the sanitizer builds of the oracle
(`make -C oracle asan tsan`): run as `python -m tests.sanitizer_workload
<path to libslo_*.so>` with the matching sanitizer runtime preloaded
(tests/test_oracle_sanitizers.py does this).  Multithreaded where the oracle
has a thread pool."""
from __future__ import annotations

import ctypes
import sys

from softlight_amd import api as A
from softlight_amd import scenes


def main(path: str) -> None:
    lib = A.Library(ctypes.CDLL(path), "slo_")

    def ctx(threads=4):
        c = A.SLContext(lib)
        c.num_threads(threads)
        return c

    c = ctx()
    s = scenes.build_random_triangles(c, seed=7, ntri=300, W=203, H=157, debug=True)
    s.render(); s.read()
    c = ctx()
    s = scenes.build_sphere_scene(c, seed=2, n_instances=4, stacks=16, sectors=16, W=160, H=96, tex_size=32, debug=True)
    s.render(); s.read()
    c = ctx()
    s = scenes.build_random_points(c, seed=2, npoints=3000, W=120, H=90, debug=True)
    s.render(); s.read()
    c = ctx()
    s = scenes.build_random_lines(c, seed=2, nlines=600, W=120, H=90, debug=True)
    s.render(); s.read()
    c = ctx()
    s = scenes.build_batched_scene(c, seed=5, n_meshes=20, median_tris=100, W=160, H=90, tex_size=32, debug=True)
    s.render(); s.read()
    # differentiable: forward, loss gradient, fixed-visibility adjoint
    c = ctx()
    d = scenes.DiffScene(c, level=2, views=2, size=48, views_per_fbo=2, tex_size=16, edge_views=1)
    d.make_targets()
    d.forward(); d.loss_grad_images(); d.backward()
    c.finish()
    # smoothed coverage with position gradients, clipped winners included
    c = ctx()
    sm = scenes.SmoothScene(c, seed=6, meshes=2, min_tris=200, max_tris=600, views=2, size=48, views_per_fbo=2,
                            tex_size=16)
    sm.make_targets(1e-3)
    sm.forward(1e-3); sm.backward(1e-3)
    c.finish()
    print("sanitizer workload: ok", flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
