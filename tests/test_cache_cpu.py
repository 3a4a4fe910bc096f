"""Mesh hash and cache file names of the collider cache (host code of liblsgpu):

- lsg_mesh_hash equals Python's hashlib SHA-1 of RRMesh::getHash's byte stream (every vertex
  as three binary32, then every triangle as three uint32) for meshes whose streams end on
  and around every padding boundary of the 64-byte blocks;
- on the benchmark meshes it equals the reference's own RRMesh::getHash, and
  lsg_cache_file_name equals RRHash::getFileName for several versions, cache directories and
  extensions (both compiled from the reference's sources into oracle/_ref/hash_ref)."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from lightsprintsdk_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HASH_BIN = os.path.join(ROOT, "oracle", "_ref", "hash_ref")


def _mesh(nv, nt, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(nv, 3)).astype(np.float32)
    i = rng.integers(0, max(nv, 1), (nt, 3)).astype(np.uint32) if nv else np.zeros((nt, 3), np.uint32)
    return v, i


def test_mesh_hash_equals_hashlib():
    from lightsprintsdk_amd.collider import mesh_hash
    # 12-byte records: streams of 0..26 records cross the 55/56/64-byte padding cases
    for nv in range(0, 14):
        for nt in range(0, 14 if nv else 1):
            v, i = _mesh(nv, nt, 100 * nv + nt)
            assert mesh_hash(v, i) == hashlib.sha1(v.tobytes() + i.tobytes()).digest(), (nv, nt)
    v, i = S.icosphere(40, seed=3)
    v, i = v.astype(np.float32), i.astype(np.uint32)
    assert mesh_hash(v, i) == hashlib.sha1(v.tobytes() + i.tobytes()).digest()


def _reference(v, i, version, location, ext, tmp_path):
    f = tmp_path / "mesh.bin"
    with open(f, "wb") as fh:
        fh.write(np.array([len(v), len(i)], np.uint32).tobytes())
        fh.write(np.ascontiguousarray(v, np.float32).tobytes())
        fh.write(np.ascontiguousarray(i, np.uint32).tobytes())
    out = subprocess.run([HASH_BIN, str(f), str(version), location, ext], check=True, capture_output=True, text=True).stdout
    hexdigest, name = out.splitlines()[:2]
    return bytes.fromhex(hexdigest), name


@pytest.mark.skipif(not os.path.exists(HASH_BIN), reason="oracle/_ref/hash_ref is built only where the reference sources exist")
def test_hash_and_file_name_equal_reference(tmp_path):
    from lightsprintsdk_amd.collider import cache_file_name, mesh_hash
    sc = S.box_room(seed=3, n_boxes=6, target_tris=8000, lightmap_size=32)
    v0, i0 = S.icosphere(S.ICOSPHERE_100K_FREQUENCY, seed=1)
    for v, i in ((sc.verts, sc.idx), (v0, i0), _mesh(5, 3, 9)):
        v, i = np.ascontiguousarray(v, np.float32), np.ascontiguousarray(i, np.uint32)
        h = mesh_hash(v, i)
        for version, loc, ext in ((4, "", ".faster"), (1, "/tmp/cache/", ".lsgbvh"), (0xFFFFFFFF, "c/", ""), (77, "x", ".fast")):
            ref_hash, ref_name = _reference(v, i, version, loc, ext, tmp_path)
            assert h == ref_hash
            assert cache_file_name(h, version, loc, ext) == ref_name, (version, loc, ext)
