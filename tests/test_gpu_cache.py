"""The collider cache on the GPU (lsg_collider_create_cached / save / load): a tree loaded from
its file equals the tree that was built, node for node, with the same leaf order and the same
hits; a file of another mesh, another build setting or a damaged file is not used (the tree is
rebuilt and the file rewritten); small trees and "*" write no file."""
import os

import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

pytestmark = pytest.mark.gpu


def _files(d):
    return sorted(f for f in os.listdir(d) if f.endswith(".lsgbvh"))


def _same_tree(a, b, rays):
    ia, ib = a.info(), b.info()
    for f in ("num_triangles", "num_valid", "num_nodes", "root_ref", "depth", "leaf_size", "num_nodes4", "stack_need"):
        assert getattr(ia, f) == getattr(ib, f), f
    ta, tb = a.download_tree(), b.download_tree()
    np.testing.assert_array_equal(ta["nodes"].view(np.uint32), tb["nodes"].view(np.uint32))
    np.testing.assert_array_equal(ta["prim"], tb["prim"])
    ha, fa = a.intersect_batch(rays)
    hb, fb = b.intersect_batch(rays)
    np.testing.assert_array_equal(ha.view(np.uint8), hb.view(np.uint8))
    np.testing.assert_array_equal(fa, fb)


def test_gpu_cache_round_trip(gpu, tmp_path):
    from lightsprintsdk_amd.collider import RRCollider
    V, I = S.icosphere(S.ICOSPHERE_100K_FREQUENCY, seed=1)
    R = S.random_rays(V, 200_000, seed=5)
    loc = str(tmp_path) + "/"
    a = RRCollider.create(V, I, cacheLocation=loc)
    assert not a.from_cache and len(_files(tmp_path)) == 1
    b = RRCollider.create(V, I, cacheLocation=loc)
    assert b.from_cache
    _same_tree(a, b, R)
    print(f"build {a.info().build_ms:.1f} ms, load {b.info().build_ms:.1f} ms", flush=True)
    # explicit save / load
    path = str(tmp_path / "tree.bin")
    a.save(path)
    c = RRCollider.load(V, I, path)
    _same_tree(a, c, R)
    # LBVH-only and another leaf size are other trees: rebuilt and written again
    d = RRCollider.create(V, I, cacheLocation=loc, lbvh_only=True)
    assert not d.from_cache
    e = RRCollider.create(V, I, cacheLocation=loc, lbvh_only=True)
    assert e.from_cache
    _same_tree(d, e, R[:20_000])
    f = RRCollider.create(V, I, cacheLocation=loc, leaf_size=4)
    assert not f.from_cache
    g = RRCollider.create(V, I, cacheLocation=loc, leaf_size=4)
    assert g.from_cache and g.info().leaf_size == 4
    _same_tree(f, g, R[:20_000])
    for col in (a, b, c, d, e, f, g):
        col.close()


def test_gpu_cache_refuses_other_mesh_and_damage(gpu, tmp_path):
    from lightsprintsdk_amd import native as N
    from lightsprintsdk_amd.collider import RRCollider
    V, I = S.icosphere(20, seed=2)
    R = S.random_rays(V, 20_000, seed=6)
    loc = str(tmp_path) + "/"
    a = RRCollider.create(V, I, cacheLocation=loc)
    path = tmp_path / _files(tmp_path)[0]
    # the same triangle count with one vertex moved: another mesh, another file name
    V2 = V.copy()
    V2[7] += 0.01
    b = RRCollider.create(V2, I, cacheLocation=loc)
    assert not b.from_cache and len(_files(tmp_path)) == 2
    with pytest.raises(N.LsgError):
        RRCollider.load(V2, I, str(path))          # the first mesh's file
    # one flipped payload byte: refused, rebuilt, rewritten
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0x40
    path.write_bytes(bytes(raw))
    with pytest.raises(N.LsgError):
        RRCollider.load(V, I, str(path))
    c = RRCollider.create(V, I, cacheLocation=loc)
    assert not c.from_cache
    d = RRCollider.create(V, I, cacheLocation=loc)
    assert d.from_cache
    _same_tree(a, d, R)
    # a truncated file
    path.write_bytes(path.read_bytes()[:-100])
    with pytest.raises(N.LsgError):
        RRCollider.load(V, I, str(path))
    for col in (a, b, c, d):
        col.close()


def test_gpu_cache_small_trees_and_star(gpu, tmp_path):
    from lightsprintsdk_amd.collider import RRCollider
    v = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    i = np.uint32([[0, 1, 2], [0, 2, 3]])
    a = RRCollider.create(v, i, cacheLocation=str(tmp_path) + "/")
    assert not _files(tmp_path)                    # under 1024 bytes: not written
    V, I = S.icosphere(20, seed=2)
    b = RRCollider.create(V, I, cacheLocation="*")
    assert not b.from_cache and not _files(tmp_path)
    a.close()
    b.close()


def test_gpu_cache_large_mesh_load_time(gpu, tmp_path):
    """The 10M-triangle city: the cached tree skips the GPU build (LBVH, SAH rebuild, collapse)."""
    import time
    from lightsprintsdk_amd.collider import RRCollider
    sc = S.city_terrain(seed=5, sun_direction=S.sun_sweep(0, 60))
    loc = str(tmp_path) + "/"
    t0 = time.perf_counter()
    a = RRCollider.create(sc.verts, sc.idx, cacheLocation=loc)
    t1 = time.perf_counter()
    b = RRCollider.create(sc.verts, sc.idx, cacheLocation=loc)
    t2 = time.perf_counter()
    assert not a.from_cache and b.from_cache
    size = os.path.getsize(tmp_path / _files(tmp_path)[0])
    print(f"city {sc.ntri} triangles: build + write {t1 - t0:.2f} s (tree {a.info().build_ms:.0f} ms), "
          f"load {t2 - t1:.2f} s (tree {b.info().build_ms:.0f} ms), file {size / 1e6:.0f} MB", flush=True)
    R = S.random_rays(sc.verts, 100_000, seed=7)
    _same_tree(a, b, R)
    a.close()
    b.close()
