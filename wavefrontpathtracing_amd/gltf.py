"""glTF 2.0 ingestion: the caller side of create_render_mesh /
submit_render_mesh for asset files (the reference's load_scene_gltf,
core/assets/asset_loader.cpp:367-560, which parses with cgltf).

What is read, following the reference loader:
- every primitive of every mesh becomes one render mesh (triangles only):
  POSITION and NORMAL are required, TEXCOORD_0 is read (zeros when absent),
  indices as u8/u16/u32 (a non-indexed primitive gets 0..n-1);
- materials: pbrMetallicRoughness.baseColorFactor, baseColorTexture and
  metallicRoughnessTexture, normalTexture, emissiveFactor, emissiveTexture
  and KHR_materials_emissive_strength
  (default 1, as cgltf); as in the reference, the emissive fields are read
  only for materials that carry a pbrMetallicRoughness block;
- every node with a mesh submits each of its primitives with the node's
  world transform (matrix, or T * R * S, composed down the hierarchy).

Base-color and emissive images (both sRGB in glTF 2.0) are decoded with
Pillow and converted from sRGB to linear RGB floats here, once, so that the library and any checker receive the same
texel values.  Unlike the reference, a node's primitive is found through its
own mesh (the reference indexes its global primitive list by the primitive's
index within the mesh, which is correct only for files with one mesh).
"""
from __future__ import annotations

import base64
import json
import os
import struct

import numpy as np

from . import scenes
from ._lib import VERTEX

_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


class GltfError(ValueError):
    pass


def _srgb_table() -> np.ndarray:
    c = np.arange(256, dtype=np.float64) / 255.0
    lin = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    return lin.astype(np.float32)


def _read_buffers(doc, base_dir, glb_bin):
    out = []
    for k, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            if glb_bin is None or k != 0:
                raise GltfError("buffer without uri outside a .glb")
            data = glb_bin
        elif uri.startswith("data:"):
            data = base64.b64decode(uri.split(",", 1)[1])
        else:
            with open(os.path.join(base_dir, uri), "rb") as f:
                data = f.read()
        if len(data) < b["byteLength"]:
            raise GltfError(f"buffer {k} holds {len(data)} bytes, expected {b['byteLength']}")
        out.append(data)
    return out


def _accessor(doc, buffers, idx) -> np.ndarray:
    a = doc["accessors"][idx]
    dt = np.dtype(_COMPONENT[a["componentType"]])
    width = _WIDTH[a["type"]]
    count = a["count"]
    if "bufferView" not in a:
        return np.zeros((count, width), dtype=dt)
    bv = doc["bufferViews"][a["bufferView"]]
    data = buffers[bv["buffer"]]
    start = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
    elem = dt.itemsize * width
    stride = bv.get("byteStride", elem)
    if count and start + stride * (count - 1) + elem > bv.get("byteOffset", 0) + bv["byteLength"]:
        raise GltfError(f"accessor {idx} reads past its buffer view")
    raw = np.frombuffer(data, dtype=np.uint8, count=stride * (count - 1) + elem if count else 0, offset=start)
    if stride == elem:
        arr = raw.view(dt).reshape(count, width)
    else:
        rows = np.lib.stride_tricks.as_strided(raw, shape=(count, elem), strides=(stride, 1))
        arr = np.ascontiguousarray(rows).view(dt).reshape(count, width)
    if a.get("normalized") and dt.kind in "iu":
        arr = arr.astype(np.float32) / float(np.iinfo(dt).max)
    return arr


def _node_matrix(node) -> np.ndarray:
    if "matrix" in node:
        return np.array(node["matrix"], dtype=np.float64).reshape(4, 4).T   # column-major in the file
    t = np.array(node.get("translation", [0, 0, 0]), dtype=np.float64)
    x, y, z, w = node.get("rotation", [0, 0, 0, 1])
    s = np.array(node.get("scale", [1, 1, 1]), dtype=np.float64)
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    m = np.eye(4)
    m[:3, :3] = r * s[None, :]
    m[:3, 3] = t
    return m


def _decode_image(doc, buffers, base_dir, img_idx, srgb=True):
    from PIL import Image
    import io
    img = doc["images"][img_idx]
    if "uri" in img:
        uri = img["uri"]
        if uri.startswith("data:"):
            blob = base64.b64decode(uri.split(",", 1)[1])
        else:
            with open(os.path.join(base_dir, uri), "rb") as f:
                blob = f.read()
    else:
        bv = doc["bufferViews"][img["bufferView"]]
        o = bv.get("byteOffset", 0)
        blob = buffers[bv["buffer"]][o:o + bv["byteLength"]]
    im = Image.open(io.BytesIO(blob)).convert("RGB")
    rgb8 = np.asarray(im, dtype=np.uint8)
    if not srgb:
        return im.width, im.height, rgb8.astype(np.float32) / np.float32(255.0)
    return im.width, im.height, _srgb_table()[rgb8]


def load(path: str, width: int = 1280, height: int = 720, vfov_deg: float = 60.0, textures: bool = True) -> scenes.Scene:
    """Parse a .gltf (external, embedded or data-URI buffers) or .glb file into a
    scenes.Scene: one mesh per primitive, one instance per (node, primitive),
    base-color textures as linear RGB floats.  The camera looks at the scene's
    bounds from the -z side."""
    base_dir = os.path.dirname(os.path.abspath(path))
    with open(path, "rb") as f:
        blob = f.read()
    glb_bin = None
    if blob[:4] == b"glTF":
        _, _, total = struct.unpack_from("<III", blob, 0)
        off, doc = 12, None
        while off < total:
            clen, ctype = struct.unpack_from("<II", blob, off)
            chunk = blob[off + 8: off + 8 + clen]
            if ctype == 0x4E4F534A:
                doc = json.loads(chunk.decode("utf-8"))
            elif ctype == 0x004E4942:
                glb_bin = chunk
            off += 8 + clen
        if doc is None:
            raise GltfError("glb without a JSON chunk")
    else:
        doc = json.loads(blob.decode("utf-8"))
    buffers = _read_buffers(doc, base_dir, glb_bin)

    # materials -> (base color, base-color texture, emissive, strength, emissive texture)
    # (asset_loader.cpp:409-470)
    tex_of_image, scene_textures = {}, []

    def texture_of(ref, srgb=True):
        if not textures or ref is None:
            return -1
        src = doc["textures"][ref["index"]].get("source")
        if src is None:
            return -1
        if (src, srgb) not in tex_of_image:
            tex_of_image[(src, srgb)] = len(scene_textures)
            scene_textures.append(_decode_image(doc, buffers, base_dir, src, srgb))
        return tex_of_image[(src, srgb)]

    mats = []
    for m in doc.get("materials", []):
        pbr = m.get("pbrMetallicRoughness")
        if pbr is None:
            mats.append(((1.0, 1.0, 1.0), -1, (0.0, 0.0, 0.0), 0.0, -1, -1, -1))
            continue
        bc = pbr.get("baseColorFactor", [1.0, 1.0, 1.0, 1.0])
        tex = texture_of(pbr.get("baseColorTexture"))
        strength = m.get("extensions", {}).get("KHR_materials_emissive_strength", {}).get("emissiveStrength", 1.0)
        em = tuple(float(x) for x in m.get("emissiveFactor", [0, 0, 0]))
        etex = texture_of(m.get("emissiveTexture")) if any(em) else -1
        # normal and metallic-roughness maps are linear (asset_loader.cpp:404-408); the debug views read them
        ntex = texture_of(m.get("normalTexture"), srgb=False)
        mrtex = texture_of(pbr.get("metallicRoughnessTexture"), srgb=False)
        mats.append((tuple(float(x) for x in bc[:3]), tex, em, float(strength), etex, ntex, mrtex))
    default_mat = ((1.0, 1.0, 1.0), -1, (0.0, 0.0, 0.0), 1.0, -1, -1, -1)

    # primitives -> meshes
    meshes, prim_index, prim_material = [], {}, {}
    for mi, mesh in enumerate(doc.get("meshes", [])):
        for pi, prim in enumerate(mesh["primitives"]):
            if prim.get("mode", 4) != 4:
                continue   # only triangle lists are rendered
            attr = prim["attributes"]
            if "POSITION" not in attr or "NORMAL" not in attr:
                raise GltfError(f"mesh {mi} primitive {pi} lacks POSITION or NORMAL")
            pos = _accessor(doc, buffers, attr["POSITION"]).astype(np.float32)
            nrm = _accessor(doc, buffers, attr["NORMAL"]).astype(np.float32)
            v = np.zeros(len(pos), dtype=VERTEX)
            v["position"] = pos
            v["normal"] = nrm
            if "TEXCOORD_0" in attr:
                v["uv"] = _accessor(doc, buffers, attr["TEXCOORD_0"]).astype(np.float32)
            if "indices" in prim:
                idx = _accessor(doc, buffers, prim["indices"]).reshape(-1).astype(np.uint32)
            else:
                idx = np.arange(len(pos), dtype=np.uint32)
            if len(idx) % 3 or (len(idx) and idx.max() >= len(pos)):
                raise GltfError(f"mesh {mi} primitive {pi}: bad index list")
            prim_index[(mi, pi)] = len(meshes)
            prim_material[(mi, pi)] = prim.get("material")
            meshes.append((v, np.ascontiguousarray(idx)))

    # nodes -> instances (world transforms down the hierarchy)
    nodes = doc.get("nodes", [])
    roots = doc["scenes"][doc.get("scene", 0)]["nodes"] if doc.get("scenes") else list(range(len(nodes)))
    instances = []
    stack = [(n, np.eye(4)) for n in reversed(roots)]
    while stack:
        n, parent = stack.pop()
        node = nodes[n]
        world = parent @ _node_matrix(node)
        if "mesh" in node:
            for pi, _ in enumerate(doc["meshes"][node["mesh"]]["primitives"]):
                key = (node["mesh"], pi)
                if key not in prim_index:
                    continue
                mat_i = prim_material[key]
                bc, tex, em, strength, etex, ntex, mrtex = mats[mat_i] if mat_i is not None else default_mat
                instances.append(scenes.Instance(prim_index[key], world.T.astype(np.float32).reshape(16).copy(),
                                                 base_color=bc, emissive=em,
                                                 emissive_strength=strength if any(em) else 0.0, texture=tex,
                                                 emissive_texture=etex, normal_texture=ntex,
                                                 metallic_roughness_texture=mrtex))
        for c in reversed(node.get("children", [])):
            stack.append((c, world))

    # camera: look at the bounds' centre from -z, far enough to see them
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for ins in instances:
        m = ins.transform.reshape(4, 4).T.astype(np.float64)
        p = meshes[ins.mesh][0]["position"].astype(np.float64)
        w = p @ m[:3, :3].T + m[:3, 3]
        lo, hi = np.minimum(lo, w.min(axis=0)), np.maximum(hi, w.max(axis=0))
    if not instances:
        raise GltfError("the scene has no triangle meshes")
    centre = (lo + hi) / 2
    radius = float(np.linalg.norm(hi - lo)) / 2
    dist = radius / np.tan(np.radians(vfov_deg) / 2) * 1.1
    eye = (float(centre[0]), float(centre[1]), float(centre[2] - dist))
    return scenes.Scene(meshes=meshes, instances=instances, eye=eye, target=tuple(float(x) for x in centre),
                        vfov_deg=vfov_deg, width=width, height=height, textures=scene_textures,
                        name=os.path.basename(path))
