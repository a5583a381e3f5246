"""glTF ingestion (the caller side of create_render_mesh / submit_render_mesh
for asset files): attribute, index, material and node-transform decoding on
synthetic files, and, when the reference's DamagedHelmet asset is present,
the library's BLAS builder against the oracle's on a real mesh."""
import math
import os

import numpy as np
import pytest

from wavefrontpathtracing_amd import gltf, scenes
from wavefrontpathtracing_amd import renderer as R

import gltf_fixtures as F
import oracle_api as O

HELMET = "/root/reference/assets/meshes/damaged_helmet/DamagedHelmet.gltf"


@pytest.mark.parametrize("interleave,glb", [(True, False), (False, False), (True, True)])
def test_synthetic_file_decodes(tmp_path, interleave, glb):
    path = tmp_path / ("s.glb" if glb else "s.gltf")
    path.write_bytes(F.build(interleave=interleave, glb=glb))
    sc = gltf.load(str(path))
    rv, ri = scenes.box_room()
    qp, qn, quv, qi = F.quad_mesh()
    assert len(sc.meshes) == 2
    assert np.array_equal(sc.meshes[0][0]["position"], rv["position"]) and np.array_equal(sc.meshes[0][1], ri)
    v, i = sc.meshes[1]
    assert np.array_equal(v["position"], qp) and np.array_equal(v["normal"], qn) and np.array_equal(v["uv"], quv)
    assert i.dtype == np.uint32 and np.array_equal(i, qi.astype(np.uint32))
    assert len(sc.instances) == 2
    room, q = sc.instances
    assert np.allclose(room.transform.reshape(4, 4).T, np.diag([2, 2, 2, 1]))
    # child of (rotate 90 deg about +y, translate (0,1,0)): quad origin at (0,1,0) + R*(0,0,0.5) = (0.5, 1, 0)
    m = q.transform.reshape(4, 4).T
    assert np.allclose(m[:3, 3], [0.5, 1.0, 0.0], atol=1e-6)
    assert np.allclose(m[:3, :3] @ [0, 0, 1], [1, 0, 0], atol=1e-6)
    assert room.base_color == pytest.approx((0.7, 0.6, 0.5)) and room.emissive_strength == 0.0
    assert q.emissive == (1.0, 0.5, 0.25) and q.emissive_strength == 4.0


def test_texture_decodes_to_linear(tmp_path):
    rgb8 = np.array([[[0, 128, 255], [10, 200, 40]]], np.uint8)
    path = tmp_path / "t.gltf"
    path.write_bytes(F.build(texture=rgb8))
    sc = gltf.load(str(path))
    assert len(sc.textures) == 1
    w, h, rgb = sc.textures[0]
    assert (w, h) == (2, 1)
    c = rgb8.astype(np.float64) / 255.0
    want = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4).astype(np.float32)
    assert np.array_equal(rgb, want)
    assert [ins.texture for ins in sc.instances] == [-1, 0, -1]
    assert sc.instances[2].emissive_strength == 6.0


def test_emissive_texture_decodes_to_linear(tmp_path):
    """emissiveTexture is sRGB like baseColorTexture (asset_loader.cpp:404-408,
    460-467) and lands in the light instance's emissive_texture slot."""
    rgb8 = np.array([[[0, 128, 255], [10, 200, 40]]], np.uint8)
    em8 = np.array([[[255, 64, 0]], [[30, 30, 220]]], np.uint8)
    path = tmp_path / "e.gltf"
    path.write_bytes(F.build(texture=rgb8, emissive_texture=em8))
    sc = gltf.load(str(path))
    assert len(sc.textures) == 2
    w, h, rgb = sc.textures[1]
    assert (w, h) == (1, 2)
    c = em8.astype(np.float64) / 255.0
    want = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4).astype(np.float32)
    assert np.array_equal(rgb, want)
    assert [ins.emissive_texture for ins in sc.instances] == [-1, -1, 1]
    assert [ins.texture for ins in sc.instances] == [-1, 0, -1]


def test_emission_needs_a_pbr_block(tmp_path):
    """The reference reads emissiveFactor, its strength and emissiveTexture only
    inside `has_pbr_metallic_roughness` (asset_loader.cpp:426-468): a material
    without that block emits nothing."""
    import json
    doc = json.loads(F.build())
    del doc["materials"][1]["pbrMetallicRoughness"]
    path = tmp_path / "nopbr.gltf"
    path.write_text(json.dumps(doc))
    sc = gltf.load(str(path))
    assert sc.instances[1].emissive_strength == 0.0 and sc.instances[1].emissive == (0.0, 0.0, 0.0)
    assert sc.instances[1].base_color == (1.0, 1.0, 1.0)


def test_missing_normals_rejected(tmp_path):
    import json
    doc = json.loads(F.build())
    del doc["meshes"][1]["primitives"][0]["attributes"]["NORMAL"]
    path = tmp_path / "bad.gltf"
    path.write_text(json.dumps(doc))
    with pytest.raises(gltf.GltfError):
        gltf.load(str(path))


@pytest.mark.skipif(not os.path.exists(HELMET), reason="the reference's assets are not on this machine")
def test_helmet_blas_matches_oracle():
    sc = gltf.load(HELMET, textures=False)
    assert len(sc.meshes) == 1 and len(sc.instances) == 1
    v, i = sc.meshes[0]
    assert len(i) == 46356 and len(v) == 14556    # 15,452 triangles
    tri = v["position"][i.reshape(-1, 3)].reshape(-1, 9).astype(np.float32)
    for bins in (8, 16):
        assert R.bvh_build(tri, interval_count=bins) == O.blas_blob(tri, bins)


@pytest.mark.skipif(not os.path.exists(HELMET), reason="the reference's assets are not on this machine")
def test_helmet_material_textures():
    """DamagedHelmet's material: base colour and emissive maps decoded as
    sRGB, normal and metallic-roughness maps as linear (asset_loader.cpp:
    404-408), each in its own texture slot."""
    sc = gltf.load(HELMET)
    ins = sc.instances[0]
    slots = [ins.texture, ins.normal_texture, ins.metallic_roughness_texture, ins.emissive_texture]
    assert min(slots) >= 0 and len(set(slots)) == 4 and len(sc.textures) == 4
    assert ins.emissive_strength > 0
    for k in slots:
        w, h, rgb = sc.textures[k]
        assert rgb.shape == (h, w, 3) and rgb.dtype == np.float32 and 0.0 <= rgb.min() and rgb.max() <= 1.0
    # a linear map keeps byte / 255; the normal map's mean blue is close to 1 (tangent-space z)
    n = sc.textures[ins.normal_texture][2]
    assert np.allclose(n * 255.0, np.round(n * 255.0), atol=1e-3) and n[..., 2].mean() > 0.7
