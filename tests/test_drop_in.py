"""The IT_HIP_LBVH technique driven through the reference's own RRCollider API
(oracle/_ref/drop_in_test: compiled reference collider + the RRColliderHIP
adapter + liblsgpu.so)."""
import os

import numpy as np
import pytest

import oracle_py as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
IT_LINEAR, IT_BSP_FASTER, IT_HIP_LBVH = 0, 3, 8

needs_bin = pytest.mark.skipif(not O.have_drop_in(), reason="oracle/_ref/drop_in_test is built only where /root/reference exists")


def _fixture():
    z = np.load(os.path.join(GOLDEN, "collider_planes.npz"))
    return z["verts"], z["idx"], z["rays"], z


def _same(a, b, fields=("tri", "t", "u", "v", "p", "plane", "front")):
    np.testing.assert_array_equal(a["tri"], b["tri"])
    m = b["tri"] >= 0
    for f in fields[1:]:
        np.testing.assert_array_equal(a[f][m], b[f][m], err_msg=f)


@needs_bin
def test_drop_in_falls_back_without_device():
    """No gfx950 device: the builder returns nullptr and RRCollider::create retries
    with IT_LINEAR (RRCollider.cpp:236-239); results equal the IT_LINEAR golden."""
    from lightsprintsdk_amd import native
    if native.device_count() > 0:
        pytest.skip("a device is present; the GPU test covers this binary")
    v, i, r, z = _fixture()
    single, batch, handler, info = O.drop_in(v, i, r, IT_HIP_LBVH)
    assert info["technique"] == IT_HIP_LBVH and info["technique_used"] == IT_LINEAR
    _same(single, z["linear"])
    _same(batch, z["linear"])


@needs_bin
@pytest.mark.gpu
def test_drop_in_gpu_technique(gpu, oracle):
    """With a device: RRCollider::create(..., IT_HIP_LBVH) returns the GPU collider;
    single-ray intersect (host mirror of the tree), intersectBatch (GPU) and the
    compiled IT_BSP_FASTER agree; with a front-side-only collision handler the
    handler sees hits nearest first, as the reference's BSP offers them."""
    v, i, r, z = _fixture()
    single, batch, handler, info = O.drop_in(v, i, r, IT_HIP_LBVH)
    assert info["technique_used"] == IT_HIP_LBVH, info
    _same(single, batch)
    _same(batch, z["bsp_faster"])
    ref_single, _, ref_handler, rinfo = O.drop_in(v, i, r, IT_BSP_FASTER)
    assert rinfo["technique_used"] == IT_BSP_FASTER
    _same(ref_single, z["bsp_faster"])
    np.testing.assert_array_equal(handler["tri"], ref_handler["tri"])
    m = handler["tri"] >= 0
    np.testing.assert_array_equal(handler["t"][m], ref_handler["t"][m])
    assert np.all(handler["front"][m] == 1)
