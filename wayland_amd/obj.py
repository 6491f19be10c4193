"""Wavefront OBJ/MTL loader for the examples/PathTracer drop-in.

Replaces the assimp import of examples/Common/SimpleModel/Model.h:95-99
(aiProcess_Triangulate | aiProcess_GenSmoothNormals |
aiProcess_JoinIdenticalVertices) for the files the example uses: one mesh
per object, polygons fan-triangulated, identical (position, normal) corners
joined in first-occurrence order, flat normals generated for faces without
normals.  Materials keep the diffuse colour Kd and the emission Ke that
GenerateHitData reads (examples/PathTracer/main.cpp:12-33).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np


@dataclass
class ObjMesh:
    name: str
    vertices: np.ndarray   # (V, 3) float32
    normals: np.ndarray    # (V, 3) float32
    indices: np.ndarray    # (T, 3) uint32
    material: str = ""


@dataclass
class ObjMaterial:
    kd: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    ke: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))


def load_mtl(path: str) -> Dict[str, ObjMaterial]:
    mats: Dict[str, ObjMaterial] = {}
    cur = None
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "newmtl":
                cur = mats.setdefault(t[1], ObjMaterial())
            elif cur is not None and t[0] == "Kd":
                cur.kd = np.array([float(x) for x in t[1:4]], np.float32)
            elif cur is not None and t[0] == "Ke":
                cur.ke = np.array([float(x) for x in t[1:4]], np.float32)
    return mats


def load_obj(path: str):
    """Returns (meshes, materials)."""
    pos: List[List[float]] = []
    nrm: List[List[float]] = []
    meshes: List[ObjMesh] = []
    mats: Dict[str, ObjMaterial] = {}
    cur = {"name": "default", "mtl": "", "faces": []}

    def flush():
        if not cur["faces"]:
            return
        key_to_idx = {}
        V, N, I = [], [], []
        for face in cur["faces"]:
            corners = []
            for (vi, ni) in face:
                if ni is None:
                    corners.append(None)
                    continue
                k = (vi, ni)
                if k not in key_to_idx:
                    key_to_idx[k] = len(V)
                    V.append(pos[vi])
                    N.append(nrm[ni])
                corners.append(key_to_idx[k])
            if any(c is None for c in corners):
                # no normals: flat normal of the face, one vertex per corner
                p = [np.array(pos[vi], np.float32) for vi, _ in face]
                n = np.cross(p[1] - p[0], p[2] - p[0])
                n = n / max(np.linalg.norm(n), 1e-30)
                corners = []
                for (vi, _) in face:
                    corners.append(len(V))
                    V.append(pos[vi])
                    N.append(list(n))
            for k in range(1, len(corners) - 1):
                I.append([corners[0], corners[k], corners[k + 1]])
        meshes.append(ObjMesh(cur["name"], np.array(V, np.float32).reshape(-1, 3),
                              np.array(N, np.float32).reshape(-1, 3),
                              np.array(I, np.uint32).reshape(-1, 3), cur["mtl"]))
        cur["faces"] = []

    base = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            if t[0] == "v":
                pos.append([float(x) for x in t[1:4]])
            elif t[0] == "vn":
                nrm.append([float(x) for x in t[1:4]])
            elif t[0] == "o" or t[0] == "g":
                flush()
                cur["name"] = t[1] if len(t) > 1 else "default"
            elif t[0] == "usemtl":
                if cur["faces"] and cur["mtl"] != t[1]:
                    flush()
                cur["mtl"] = t[1]
            elif t[0] == "mtllib":
                mats.update(load_mtl(os.path.join(base, t[1])))
            elif t[0] == "f":
                face = []
                for c in t[1:]:
                    parts = c.split("/")
                    vi = int(parts[0])
                    vi = vi - 1 if vi > 0 else len(pos) + vi
                    ni = None
                    if len(parts) >= 3 and parts[2]:
                        ni = int(parts[2])
                        ni = ni - 1 if ni > 0 else len(nrm) + ni
                    face.append((vi, ni))
                cur["faces"].append(face)
    flush()
    return meshes, mats
