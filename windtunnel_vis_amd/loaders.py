"""Model files into merged triangle meshes (src/voxelize/loaders.ts): STL (binary and ASCII), OBJ
and glTF 2.0 (.gltf with its buffers, or .glb).

`load_model(path)` returns (Mesh, up) where `up` is the file format's usual up axis ('z' for STL,
'y' otherwise), to be passed to `normalize_model` (mesh.ts: normalizeMesh with `up` / `flip`).
Node transforms of glTF scenes are applied; every primitive becomes one part, re-oriented to a
positive enclosed volume as the reference's mergeObject does. Materials and textures are not read
(they only feed the reference's renderer). Compressed glTF geometry (Draco, meshopt) is refused
with a clear error: the decoders are not available here.
"""
import base64
import json
import math
import os
import struct

import numpy as np

from .geometry import Mesh, merge_meshes, normalize_mesh, rot_y
from .presets import orient

MODEL_EXTS = ("glb", "gltf", "obj", "stl")


def _ext(path):
    return os.path.splitext(path)[1].lower().lstrip(".")


# ---- STL -----------------------------------------------------------------------------------------
def load_stl(path):
    data = open(path, "rb").read()
    tris = None
    if len(data) >= 84:
        n = struct.unpack_from("<I", data, 80)[0]
        if 84 + 50 * n == len(data):  # binary: 80-byte header, count, 50 bytes per triangle
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]), count=n, offset=84)
            tris = rec["v"].astype(np.float32)
    if tris is None:
        verts = []
        for line in data.decode("utf-8", errors="replace").splitlines():
            w = line.split()
            if len(w) == 4 and w[0].lower() == "vertex":
                verts.append([float(w[1]), float(w[2]), float(w[3])])
        if len(verts) % 3:
            raise ValueError("%s: ASCII STL with %d vertices (not a multiple of 3)" % (path, len(verts)))
        tris = np.asarray(verts, dtype=np.float32).reshape(-1, 3, 3)
    if len(tris) == 0:
        raise ValueError("%s: no triangles" % path)
    pos = tris.reshape(-1, 3)
    return Mesh(pos, np.arange(len(pos), dtype=np.uint32).reshape(-1, 3))


# ---- OBJ -----------------------------------------------------------------------------------------
def load_obj(path):
    """Vertices `v` and faces `f` (polygons fanned, negative indices relative); one part per
    object / group (`o`, `g`)."""
    verts = []
    parts = [[]]
    for line in open(path, "r", errors="replace"):
        w = line.split()
        if not w:
            continue
        if w[0] == "v" and len(w) >= 4:
            verts.append([float(w[1]), float(w[2]), float(w[3])])
        elif w[0] in ("o", "g"):
            if parts[-1]:
                parts.append([])
        elif w[0] == "f" and len(w) >= 4:
            ids = []
            for tok in w[1:]:
                i = int(tok.split("/")[0])
                ids.append(i - 1 if i > 0 else len(verts) + i)
            for k in range(1, len(ids) - 1):
                parts[-1].append((ids[0], ids[k], ids[k + 1]))
    pos = np.asarray(verts, dtype=np.float32).reshape(-1, 3)
    meshes = []
    for tri in parts:
        if not tri:
            continue
        t = np.asarray(tri, dtype=np.int64)
        if t.min() < 0 or t.max() >= len(pos):
            raise ValueError("%s: face index out of range" % path)
        used, inv = np.unique(t.reshape(-1), return_inverse=True)
        meshes.append(orient(Mesh(pos[used], inv.reshape(-1, 3))))
    if not meshes:
        raise ValueError("%s: no faces" % path)
    return merge_meshes(meshes)


# ---- glTF 2.0 ------------------------------------------------------------------------------------
_COMP = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


def _node_matrix(node):
    if "matrix" in node:
        return np.asarray(node["matrix"], dtype=np.float64).reshape(4, 4).T  # column-major
    T = np.eye(4)
    T[:3, 3] = node.get("translation", [0, 0, 0])
    x, y, z, w = node.get("rotation", [0, 0, 0, 1])
    R = np.eye(4)
    R[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                 [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                 [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    S = np.diag(list(node.get("scale", [1, 1, 1])) + [1.0])
    return T @ R @ S


def load_gltf(path):
    raw = open(path, "rb").read()
    bin_chunk = None
    if raw[:4] == b"glTF":
        _, _, length = struct.unpack_from("<III", raw, 0)
        off = 12
        doc = None
        while off < length:
            clen, ctype = struct.unpack_from("<II", raw, off)
            chunk = raw[off + 8: off + 8 + clen]
            if ctype == 0x4E4F534A:
                doc = json.loads(chunk.decode("utf-8"))
            elif ctype == 0x004E4942:
                bin_chunk = chunk
            off += 8 + clen
        if doc is None:
            raise ValueError("%s: GLB without a JSON chunk" % path)
    else:
        doc = json.loads(raw.decode("utf-8"))
    for ext in doc.get("extensionsRequired", []):
        if ext in ("KHR_draco_mesh_compression", "EXT_meshopt_compression", "KHR_meshopt_compression"):
            raise ValueError("%s: compressed geometry (%s) is not supported; export it uncompressed" % (path, ext))
    base = os.path.dirname(os.path.abspath(path))
    buffers = []
    for i, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            if bin_chunk is None:
                raise ValueError("%s: buffer %d has no data" % (path, i))
            buffers.append(bin_chunk)
        elif uri.startswith("data:"):
            buffers.append(base64.b64decode(uri.split(",", 1)[1]))
        else:
            p = os.path.join(base, uri)
            if not os.path.exists(p):
                raise ValueError("%s: missing buffer file %s (select it together with the .gltf)" % (path, uri))
            buffers.append(open(p, "rb").read())

    def accessor(k):
        a = doc["accessors"][k]
        if "bufferView" not in a:
            return np.zeros((a["count"], _NCOMP[a["type"]]), dtype=_COMP[a["componentType"]])
        bv = doc["bufferViews"][a["bufferView"]]
        dt = np.dtype(_COMP[a["componentType"]])
        nc = _NCOMP[a["type"]]
        buf = buffers[bv["buffer"]]
        start = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
        stride = bv.get("byteStride", 0) or dt.itemsize * nc
        cnt = a["count"]
        out = np.empty((cnt, nc), dtype=dt)
        for j in range(nc):
            out[:, j] = np.ndarray((cnt,), dtype=dt, buffer=buf, offset=start + j * dt.itemsize, strides=(stride,))
        return out

    meshes = []

    def visit(ni, parent):
        node = doc["nodes"][ni]
        M = parent @ _node_matrix(node)
        if "mesh" in node:
            for prim in doc["meshes"][node["mesh"]]["primitives"]:
                if prim.get("mode", 4) != 4 or "POSITION" not in prim.get("attributes", {}):
                    continue
                if "extensions" in prim and any(e in prim["extensions"] for e in ("KHR_draco_mesh_compression",)):
                    raise ValueError("%s: Draco-compressed primitive" % path)
                p = accessor(prim["attributes"]["POSITION"]).astype(np.float64)
                idx = accessor(prim["indices"]).reshape(-1).astype(np.int64) if "indices" in prim else np.arange(len(p))
                idx = idx[: len(idx) - len(idx) % 3].reshape(-1, 3)
                pw = (np.c_[p, np.ones(len(p))] @ M.T)[:, :3]
                if np.linalg.det(M[:3, :3]) < 0:
                    idx = idx[:, ::-1]
                meshes.append(orient(Mesh(pw, idx)))
        for c in node.get("children", []):
            visit(c, M)

    scene = doc.get("scenes", [{}])[doc.get("scene", 0)] if doc.get("scenes") else {"nodes": list(range(len(doc.get("nodes", []))))}
    for ni in scene.get("nodes", []):
        visit(ni, np.eye(4))
    if not meshes:
        raise ValueError("%s: no triangle meshes found in the file" % path)
    return merge_meshes(meshes)


def load_model(path):
    """(Mesh, up axis) of a .stl / .obj / .gltf / .glb file (loaders.ts: loadModelFiles)."""
    e = _ext(path)
    if e == "stl":
        return load_stl(path), "z"
    if e == "obj":
        return load_obj(path), "y"
    if e in ("gltf", "glb"):
        return load_gltf(path), "y"
    raise ValueError("no model among the files — use .glb, .gltf, .obj or .stl (got %r)" % path)


def normalize_model(mesh, up="y", flip=False, align_longest=True):
    """mesh.ts normalizeMesh with the `up` and `flip` options: a z-up model is turned y-up (rotation
    of -90 degrees about x), `flip` turns it 180 degrees about y (the other driving direction)."""
    p = mesh.positions.astype(np.float64)
    if up == "z":
        p = np.c_[p[:, 0], p[:, 2], -p[:, 1]]
    if flip:
        p = p @ rot_y(180).T
    return normalize_mesh(Mesh(p, mesh.indices, mesh.wheels), align_longest=align_longest)


def save_stl(path, mesh):
    """Binary STL of a mesh (used to hand meshes to other tools and in tests)."""
    p = mesh.positions.astype(np.float32)
    t = mesh.indices.astype(np.int64)
    a, b, c = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]
    nrm = np.cross(b - a, c - a)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    rec = np.zeros(len(t), dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    rec["n"] = nrm
    rec["v"] = np.stack([a, b, c], axis=1)
    with open(path, "wb") as fh:
        fh.write(b"windtunnel_vis_amd binary STL".ljust(80, b" "))
        fh.write(struct.pack("<I", len(t)))
        fh.write(rec.tobytes())
