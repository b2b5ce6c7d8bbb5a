"""Wavefront OBJ import (M/geometry/shape/TriangleMesh3F.java:534-960 and its
DefaultObjectModel / IndexedObjectModel, :1200-1700).

`read_wavefront_object(path)` returns the triangle meshes of a file in the
reference's order and grouping:

* a face with n vertices becomes the fan (v0, v[i+1], v[i+2]), i < n - 2
  (:1344-1358);
* indices are 1-based; a vertex is the triple (position, texture
  coordinates, normal) and equal triples share one index (:1219-1263);
* each face vertex records the group, material and object names current
  where it is read, and a vertex triple keeps the names of its FIRST
  occurrence in the file (the indexed model's computeIfAbsent, :1250-1263);
  a triangle takes the names of its vertex A (:651-655), so a face whose
  vertex A appeared earlier under another `usemtl` or `g` belongs to that
  earlier group and material;
* a new mesh starts where a triangle's group or material name differs from
  the previous triangle's; an object name alone does not split, and a mesh
  carries the object name of its last triangle (:657-683, :705-709);
* a file without any `vn` gets per-position normals: the normalised sum of
  the unit normals cross(B - A, C - A) of the faces around each position
  (:1567-1583), shared by every vertex with that position;
* `scale` multiplies positions; normals go through the inverse transpose,
  so they keep their direction for a positive scale and flip for a
  negative one; a zero scale (not invertible) gives no meshes (:620-625);
  `flip_texture_y` replaces v by 1 - v.

The reference indexes its lists with `index - 1` and so fails on negative
(relative) indices; so does this reader (ValueError).  Files that give
normals for some face vertices and not for others are rejected: the
reference's indexed model mis-associates normals there (it appends normals
only for the vertices that carry one).
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np


@dataclasses.dataclass
class TriangleMesh:
    """One mesh of an OBJ file: world-space arrays plus the reference's names."""
    group_name: str
    material_name: str
    object_name: str
    positions: np.ndarray   # (N, 3) float32
    normals: np.ndarray     # (N, 3) float32, unit length
    uvs: np.ndarray         # (N, 2) float32
    indices: np.ndarray     # (M, 3) int32, into this mesh's arrays


def _index(token, count, what, line_no):
    i = int(token) - 1
    if i < 0 or i >= count:
        raise ValueError(f"line {line_no}: {what} index {token} out of range 1..{count}")
    return i


def _parse(lines, flip_texture_y):
    pos, tex, nrm = [], [], []
    verts, groups, mats, objs = [], [], [], []
    group = material = obj = ""
    for line_no, line in enumerate(lines, 1):
        el = line.split()
        if not el or el[0].startswith("#"):
            continue
        k = el[0]
        if k == "v":
            pos.append((float(el[1]), float(el[2]), float(el[3])))
        elif k == "vt":
            u, v = float(el[1]), float(el[2]) if len(el) > 2 else 0.0
            tex.append((u, 1.0 - v if flip_texture_y else v))
        elif k == "vn":
            nrm.append((float(el[1]), float(el[2]), float(el[3])))
        elif k == "g":
            group = el[1] if len(el) > 1 else ""
        elif k == "o":
            obj = el[1] if len(el) > 1 else ""
        elif k == "usemtl":
            material = el[1] if len(el) > 1 else ""
        elif k == "f":
            face = []
            for tok in el[1:]:
                parts = tok.split("/")
                p = _index(parts[0], len(pos), "position", line_no)
                t = _index(parts[1], len(tex), "texture", line_no) if len(parts) > 1 and parts[1] else -1
                n = _index(parts[2], len(nrm), "normal", line_no) if len(parts) > 2 and parts[2] else -1
                face.append((p, t, n))
            for i in range(len(face) - 2):
                for vtx in (face[0], face[i + 1], face[i + 2]):
                    verts.append(vtx)
                    groups.append(group)
                    mats.append(material)
                    objs.append(obj)
    return (np.asarray(pos, np.float64).reshape(-1, 3), np.asarray(tex, np.float64).reshape(-1, 2),
            np.asarray(nrm, np.float64).reshape(-1, 3), verts, groups, mats, objs)


def _position_normals(pos, verts):
    """Per-position normals: normalised sum of unit face normals."""
    acc = np.zeros_like(pos)
    tri = np.asarray([v[0] for v in verts], np.int64).reshape(-1, 3)
    a, b, c = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    fn = np.cross(b - a, c - a)
    ln = np.linalg.norm(fn, axis=1, keepdims=True)
    fn = np.divide(fn, ln, out=np.zeros_like(fn), where=ln > 0)
    for k in range(3):
        np.add.at(acc, tri[:, k], fn)
    ln = np.linalg.norm(acc, axis=1, keepdims=True)
    return np.divide(acc, ln, out=np.zeros_like(acc), where=ln > 0)


def parse_wavefront_object(text, flip_texture_y=False, scale=1.0):
    """The meshes of OBJ source text (see the module docstring)."""
    if scale == 0.0:
        return []
    pos, tex, nrm, verts, groups, mats, objs = _parse(text.splitlines(), flip_texture_y)
    nsign = 1.0 if scale > 0.0 else -1.0
    has_n = [v[2] >= 0 for v in verts]
    if any(has_n) and not all(has_n):
        raise ValueError("some face vertices have normals and others do not")
    pos_n = None if any(has_n) else _position_normals(pos, verts)
    first = {}                                   # vertex triple -> names where it first occurs
    for vtx, g, mt, o in zip(verts, groups, mats, objs):
        first.setdefault(vtx, (g, mt, o))
    meshes = []
    cur, key = None, None
    for i in range(0, len(verts), 3):
        g, mt, o = first[verts[i]]
        if cur is None or (g, mt) != key:
            cur = {"names": [g, mt, o], "map": {}, "p": [], "n": [], "t": [], "idx": []}
            meshes.append(cur)
            key = (g, mt)
        cur["names"][2] = o
        for vtx in verts[i:i + 3]:
            j = cur["map"].get(vtx)
            if j is None:
                j = len(cur["p"])
                cur["map"][vtx] = j
                cur["p"].append(pos[vtx[0]] * scale)
                n = nrm[vtx[2]] if vtx[2] >= 0 else pos_n[vtx[0]]
                ln = np.linalg.norm(n)
                cur["n"].append(nsign * n / ln if ln > 0 else n)
                cur["t"].append(tex[vtx[1]] if vtx[1] >= 0 else (0.0, 0.0))
            cur["idx"].append(j)
    out = []
    for m in meshes:
        g, mt, o = m["names"]
        out.append(TriangleMesh(g, mt, o, np.asarray(m["p"], np.float32).reshape(-1, 3),
                                np.asarray(m["n"], np.float32).reshape(-1, 3),
                                np.asarray(m["t"], np.float32).reshape(-1, 2),
                                np.asarray(m["idx"], np.int32).reshape(-1, 3)))
    return out


def read_wavefront_object(path, flip_texture_y=False, scale=1.0):
    """TriangleMesh3F.readWavefrontObject(file, isFlippingTextureCoordinateY, scale)."""
    with open(os.fspath(path), "r") as f:
        return parse_wavefront_object(f.read(), flip_texture_y, scale)


def add_to_scene(scene, meshes, materials):
    """Adds each mesh to `scene` with the material index `materials[name]`
    (or `materials[""]` for names without an entry); returns the primitive
    indices."""
    prims = []
    for m in meshes:
        mat = materials.get(m.material_name, materials.get(""))
        if mat is None:
            raise KeyError(f"no material for '{m.material_name}'")
        prims.append(scene.add_mesh(m.positions, m.indices, mat, normals=m.normals))
    return prims
