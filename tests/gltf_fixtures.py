"""Synthetic glTF 2.0 files for the loader tests (written on the fly)."""
import base64
import io
import json
import math
import struct

import numpy as np

from wavefrontpathtracing_amd import scenes


def _pad4(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 4)


def quad_mesh():
    """Unit quad in the xy plane facing -z (positions, normals, uvs, u16 indices)."""
    p = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    n = np.tile(np.array([0, 0, -1], np.float32), (4, 1))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    i = np.array([0, 2, 1, 0, 3, 2], np.uint16)
    return p, n, uv, i


def png_bytes(rgb8: np.ndarray) -> bytes:
    from PIL import Image
    buf = io.BytesIO()
    Image.fromarray(rgb8, "RGB").save(buf, format="PNG")
    return buf.getvalue()


def build(interleave: bool = True, glb: bool = False, texture: np.ndarray = None, emissive_texture: np.ndarray = None):
    """Two meshes: the room of scenes.box_room() (u32 indices) and a quad
    (u16 indices, interleaved position/normal/uv when `interleave`).  Node 0
    (scale 2) holds the room; node 1 rotates 90 degrees about +y, translates
    by (0, 1, 0) and has child node 2 (translation (0, 0, 0.5), the quad).  The
    quad's material is emissive (factor (1, 0.5, 0.25), strength 4) or, with
    `texture`, textured, and then a third node holds a downward emissive quad
    (factor 1, strength 6, with `emissive_texture` as its emissiveTexture).
    Returns the file bytes (JSON text or GLB)."""
    rv, ri = scenes.box_room()
    qp, qn, quv, qi = quad_mesh()
    views, accessors, blob = [], [], b""

    def add_view(data: bytes, stride=None):
        nonlocal blob
        off = len(blob)
        blob += _pad4(data)
        v = {"buffer": 0, "byteOffset": off, "byteLength": len(data)}
        if stride:
            v["byteStride"] = stride
        views.append(v)
        return len(views) - 1

    def add_acc(view, ctype, count, typ, offset=0, mn=None, mx=None):
        a = {"bufferView": view, "componentType": ctype, "count": count, "type": typ, "byteOffset": offset}
        if mn is not None:
            a["min"], a["max"] = mn, mx
        accessors.append(a)
        return len(accessors) - 1

    # room: separate views
    rpos = np.ascontiguousarray(rv["position"], np.float32)
    a_rp = add_acc(add_view(rpos.tobytes()), 5126, len(rv), "VEC3", mn=rpos.min(0).tolist(), mx=rpos.max(0).tolist())
    a_rn = add_acc(add_view(np.ascontiguousarray(rv["normal"], np.float32).tobytes()), 5126, len(rv), "VEC3")
    a_ruv = add_acc(add_view(np.ascontiguousarray(rv["uv"], np.float32).tobytes()), 5126, len(rv), "VEC2")
    a_ri = add_acc(add_view(ri.astype(np.uint32).tobytes()), 5125, len(ri), "SCALAR")
    # quad
    if interleave:
        inter = np.concatenate([qp, qn, quv], axis=1).astype(np.float32)       # 32-byte rows
        vq = add_view(inter.tobytes(), stride=32)
        a_qp = add_acc(vq, 5126, 4, "VEC3", 0, qp.min(0).tolist(), qp.max(0).tolist())
        a_qn = add_acc(vq, 5126, 4, "VEC3", 12)
        a_quv = add_acc(vq, 5126, 4, "VEC2", 24)
    else:
        a_qp = add_acc(add_view(qp.tobytes()), 5126, 4, "VEC3", mn=qp.min(0).tolist(), mx=qp.max(0).tolist())
        a_qn = add_acc(add_view(qn.tobytes()), 5126, 4, "VEC3")
        a_quv = add_acc(add_view(quv.tobytes()), 5126, 4, "VEC2")
    a_qi = add_acc(add_view(qi.tobytes()), 5123, 6, "SCALAR")
    doc = {
        "asset": {"version": "2.0"},
        "scene": 0,
        "scenes": [{"nodes": [0, 1]}],
        "nodes": [
            {"mesh": 0, "scale": [2.0, 2.0, 2.0]},
            {"rotation": [0.0, math.sin(math.pi / 4), 0.0, math.cos(math.pi / 4)], "translation": [0.0, 1.0, 0.0],
             "children": [2]},
            {"mesh": 1, "translation": [0.0, 0.0, 0.5]},
        ],
        "meshes": [
            {"primitives": [{"attributes": {"POSITION": a_rp, "NORMAL": a_rn, "TEXCOORD_0": a_ruv}, "indices": a_ri,
                             "material": 0}]},
            {"primitives": [{"attributes": {"POSITION": a_qp, "NORMAL": a_qn, "TEXCOORD_0": a_quv}, "indices": a_qi,
                             "material": 1}]},
        ],
        "materials": [
            {"pbrMetallicRoughness": {"baseColorFactor": [0.7, 0.6, 0.5, 1.0]}},
            {"pbrMetallicRoughness": {"baseColorFactor": [1.0, 1.0, 1.0, 1.0]}, "emissiveFactor": [1.0, 0.5, 0.25],
             "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0}}},
        ],
        "accessors": accessors,
        "bufferViews": views,
    }
    if texture is not None:
        img_view = add_view(png_bytes(texture))
        doc["bufferViews"] = views
        doc["images"] = [{"bufferView": img_view, "mimeType": "image/png"}]
        doc["textures"] = [{"source": 0}]
        doc["materials"][1] = {"pbrMetallicRoughness": {"baseColorFactor": [1.0, 1.0, 1.0, 1.0],
                                                        "baseColorTexture": {"index": 0}}}
        # light: the room's material becomes emissive only on the quad's side -> use a third node
        doc["nodes"].append({"mesh": 1, "translation": [0.0, 1.9, 0.0],
                             "rotation": [math.sin(-math.pi / 4), 0.0, 0.0, math.cos(-math.pi / 4)]})
        doc["meshes"].append({"primitives": [{"attributes": {"POSITION": a_qp, "NORMAL": a_qn, "TEXCOORD_0": a_quv},
                                              "indices": a_qi, "material": 2}]})
        doc["nodes"][-1]["mesh"] = 2
        doc["scenes"][0]["nodes"].append(3)
        light = {"pbrMetallicRoughness": {"baseColorFactor": [1.0, 1.0, 1.0, 1.0]}, "emissiveFactor": [1.0, 1.0, 1.0],
                 "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 6.0}}}
        if emissive_texture is not None:
            doc["images"].append({"bufferView": add_view(png_bytes(emissive_texture)), "mimeType": "image/png"})
            doc["textures"].append({"source": 1})
            light["emissiveTexture"] = {"index": 1}
        doc["materials"].append(light)
    if glb:
        doc["buffers"] = [{"byteLength": len(blob)}]
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
        return struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body
    doc["buffers"] = [{"byteLength": len(blob),
                       "uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode()}]
    return json.dumps(doc).encode()
