"""Oracle for scene ingestion -- TEST INFRASTRUCTURE ONLY.

A pure-Python restatement of the reference's model loading, for files of a
few thousand lines: the checker of libvisionaray_amd's vr_mesh_* entry
points, never part of the product path.

    model::load            src/common/model.cpp:144-200
    load_obj               src/common/obj_loader.cpp:306-520
      remap_index            :40-44    store_triangle   :51-74
      store_faces            :81-132   parse_mtl        :182-242
      add_material           :248-263  mtl defaults     :163-175
    obj_grammar            src/common/obj_grammar.cpp:35-64
    load_ply               src/common/ply_loader.cpp:64-211
    obj_material defaults  src/common/sg/material.h:41-72
    map_material           src/viewer/viewer.cpp:745-795

Boost (Spirit, filesystem, iostreams) and tinyply are not in this image, so
the reference loaders cannot be built here.  Their behaviour is restated from
the grammar and loader sources above, and from the PLY 1.0 format that
tinyply (submodule src/3rdparty/tinyply, ddiakopoulos/tinyply, commit not
recorded in the snapshot) implements.  Pinned by the known-answer fixtures
under tests/golden/meshes (expected values derived by hand from those
sources) and by round trips whose answers follow from float arithmetic.

Reals: qi::float_ accumulates the digits exactly and scales once by a power
of ten, which is correctly rounded for mantissas below 2^24 and power-of-ten
scales up to 10^10; this oracle rounds float(token) (correctly rounded to
double) to float32, which equals the correctly rounded float32 for such
short decimals.  The fixtures stay inside that domain.
"""
from __future__ import annotations

import math
import os
import re
import struct

import numpy as np

F32 = np.float32

_REAL = re.compile(r"[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|[nN][aA][nN]|[iI][nN][fF](?:[iI][nN][iI][tT][yY])?)")
_INT = re.compile(r"[+-]?\d+")


def _f(tok: str) -> np.float32:
    return F32(float(tok))


def _lines(data: bytes):
    """Lines ending in \\r\\n, \\n or \\r; a final line without an end matches no
    rule of the grammar (each ends in qi::eol) and is dropped."""
    text = data.decode("latin-1")
    out = []
    i = 0
    n = len(text)
    while i < n:
        j = i
        while j < n and text[j] not in "\r\n":
            j += 1
        if j == n:
            break
        out.append(text[i:j])
        i = j + 2 if text[j] == "\r" and j + 1 < n and text[j + 1] == "\n" else j + 1
    return out


def _reals(s: str):
    """Every real on the rest of the line, or None when anything else is there."""
    vals = []
    pos = 0
    while True:
        while pos < len(s) and s[pos] in " \t":
            pos += 1
        if pos == len(s):
            return vals
        m = _REAL.match(s, pos)
        if not m:
            return None
        vals.append(_f(m.group(0)))
        pos = m.end()


def _corner(s: str, pos: int):
    """int ['/'] [int] ['/'] [int] from pos (a lexeme): ((v, t, n), end) or None."""
    m = _INT.match(s, pos)
    if not m:
        return None
    v = int(m.group(0))
    pos = m.end()
    t = n = None
    if pos < len(s) and s[pos] == "/":
        pos += 1
    m = _INT.match(s, pos)
    if m:
        t = int(m.group(0))
        pos = m.end()
    if pos < len(s) and s[pos] == "/":
        pos += 1
    m = _INT.match(s, pos)
    if m:
        n = int(m.group(0))
        pos = m.end()
    for x in (v, t, n):
        if x is not None and not (-2 ** 31 <= x < 2 ** 31):
            return None
    return (v, t, n), pos


def _face(s: str):
    corners = []
    pos = 0
    while True:
        while pos < len(s) and s[pos] in " \t":
            pos += 1
        if pos == len(s):
            break
        r = _corner(s, pos)
        if r is None:
            return None
        corners.append(r[0])
        pos = r[1]
    return corners if len(corners) >= 3 else None


def default_material():
    return dict(ca=[F32(0.2)] * 3, cd=[F32(0.8)] * 3, cs=[F32(0.1)] * 3, ce=[F32(0.0)] * 3, ior=[F32(1.0)] * 3,
                transmission=F32(0.0), specular_exp=F32(32.0), illum=2, name="")


def _parse_mtl(path: str, lib: dict):
    cur = None
    with open(path, "rb") as fh:
        lines = _lines(fh.read())
    for line in lines:
        s = line.lstrip(" \t")
        if s.startswith("newmtl"):
            name = s[6:].strip(" \t\n\r\f\v")
            if name not in lib:
                lib[name] = dict(ka=[F32(0.2)] * 3, kd=[F32(0.8)] * 3, ke=[F32(0.0)] * 3, ks=[F32(0.1)] * 3,
                                 tr=F32(0.0), d=F32(1.0), ns=F32(32.0), ni=F32(1.0), map_kd="", illum=2)
            cur = lib[name]
            continue
        if cur is None:
            continue
        done = False
        for kw, key in (("Ka", "ka"), ("Kd", "kd"), ("Ke", "ke"), ("Ks", "ks")):
            if not done and s.startswith(kw):
                r = _reals(s[2:])
                if r is not None and len(r) == 3:
                    cur[key] = list(r)
                    done = True
        for kw, key in (("Tr", "tr"), ("d", "d"), ("Ns", "ns"), ("Ni", "ni")):
            if not done and s.startswith(kw):
                r = _reals(s[len(kw):])
                if r is not None and len(r) == 1:
                    cur[key] = r[0]
                    done = True
        if not done and s.startswith("map_Kd"):
            cur["map_kd"] = s[6:].lstrip(" \t")
        elif not done and s.startswith("illum"):
            m = re.fullmatch(r"[ \t]*([+-]?\d+)[ \t]*", s[5:])
            if m:
                cur["illum"] = int(m.group(1))


def _add_material(m: dict, name: str) -> dict:
    tr = m["tr"] if m["tr"] > F32(0.0) else F32(0.0)
    one_minus_d = F32(F32(1.0) - m["d"])
    tr = one_minus_d if one_minus_d > F32(0.0) else tr
    return dict(ca=list(m["ka"]), cd=list(m["kd"]), cs=list(m["ks"]), ce=list(m["ke"]), ior=[m["ni"]] * 3,
                transmission=tr, specular_exp=m["ns"], illum=m["illum"], name=name)


class Model:
    def __init__(self):
        self.prims = []       # (geom_id, prim_id, v1, e1, e2) with float32 3-lists
        self.geometric_normals = []
        self.shading_normals = []
        self.tex_coords = []
        self.materials = []
        self.lo = [F32(np.finfo(np.float32).max)] * 3
        self.hi = [F32(-np.finfo(np.float32).max)] * 3
        self.degenerate = 0
        self.missing = 0

    def _insert(self, v):
        self.lo = [v[k] if v[k] < self.lo[k] else self.lo[k] for k in range(3)]
        self.hi = [v[k] if self.hi[k] < v[k] else self.hi[k] for k in range(3)]


def _sub(a, b):
    return [F32(a[k] - b[k]) for k in range(3)]


def _add(a, b):
    return [F32(a[k] + b[k]) for k in range(3)]


def _cross(u, v):
    return [F32(F32(u[1] * v[2]) - F32(u[2] * v[1])), F32(F32(u[2] * v[0]) - F32(u[0] * v[2])),
            F32(F32(u[0] * v[1]) - F32(u[1] * v[0]))]


def _dot(u, v):
    return F32(F32(F32(u[0] * v[0]) + F32(u[1] * v[1])) + F32(u[2] * v[2]))


def _normalize(v):
    r = F32(F32(1.0) / np.sqrt(_dot(v, v)))
    return [F32(v[k] * r) for k in range(3)]


def _remap(i: int, size: int) -> int:
    r = size + i if i < 0 else i - 1
    if not (0 <= r < size):
        raise ValueError(f"index {i} out of range (size {size})")
    return r


def load_obj(mod: Model, paths):
    lib, parsed, geom_id = {}, [], 0
    for path in paths:
        with open(path, "rb") as fh:
            lines = _lines(fh.read())
        verts, tcs, nrms = [], [], []
        for line in lines:
            s = line.lstrip(" \t")
            if s.startswith("#"):
                continue
            if s.startswith("mtllib"):
                name = s[6:].lstrip(" \t")
                if name in parsed:
                    continue
                d = os.path.dirname(path)
                lp = name if d == "" else d + "/" + name
                if os.path.isfile(lp):
                    _parse_mtl(lp, lib)
                else:
                    mod.missing += 1
                parsed.append(name)
                continue
            if s.startswith("usemtl"):
                name = s[6:].strip(" \t\n\r\f\v")
                if name in lib:
                    mod.materials.append(_add_material(lib[name], name))
                else:
                    mod.missing += 1
                geom_id = max(len(mod.materials) - 1, 0)
                continue
            if s.startswith("v"):
                r = _reals(s[1:])
                if r is not None and len(r) in (3, 4, 6):
                    verts.append(r[:3])
                    continue
                r = _reals(s[2:]) if s.startswith("vt") else None
                if r is not None and len(r) in (2, 3):
                    tcs.append(r[:2])
                    continue
                r = _reals(s[2:]) if s.startswith("vn") else None
                if r is not None and len(r) == 3:
                    nrms.append(r)
                continue
            if s.startswith("f"):
                corners = _face(s[1:])
                if corners is None:
                    continue
                i1 = _remap(corners[0][0], len(verts))
                for last in range(2, len(corners)):
                    c2, c3 = corners[last - 1], corners[last]
                    i2, i3 = _remap(c2[0], len(verts)), _remap(c3[0], len(verts))
                    v1 = verts[i1]
                    e1, e2 = _sub(verts[i2], v1), _sub(verts[i3], v1)
                    c = _cross(e1, e2)
                    if np.sqrt(_dot(c, c)) == F32(0.0):
                        mod.degenerate += 1
                        continue
                    gid = len(mod.materials) - 1 if mod.materials else 0
                    mod.prims.append((gid, len(mod.prims), v1, e1, e2))
                    mod.geometric_normals.append(("obj", e1, e2))
                    if corners[0][1] is not None and c2[1] is not None and c3[1] is not None:
                        for t in (corners[0][1], c2[1], c3[1]):
                            mod.tex_coords.append(tcs[_remap(t, len(tcs))])
                    if corners[0][2] is not None and c2[2] is not None and c3[2] is not None:
                        for n in (corners[0][2], c2[2], c3[2]):
                            mod.shading_normals.append(nrms[_remap(n, len(nrms))])
        while len(mod.materials) <= geom_id:
            mod.materials.append(default_material())
    while len(mod.tex_coords) < 3 * len(mod.prims):
        mod.tex_coords.append([F32(0.0), F32(0.0)])
    for (_, _, v1, e1, e2) in mod.prims:
        mod._insert(v1)
        mod._insert(_add(v1, e1))
        mod._insert(_add(v1, e2))


_PLY_TYPES = {"char": "b", "int8": "b", "uchar": "B", "uint8": "B", "short": "h", "int16": "h",
              "ushort": "H", "uint16": "H", "int": "i", "int32": "i", "uint": "I", "uint32": "I",
              "float": "f", "float32": "f", "double": "d", "float64": "d"}


def load_ply(mod: Model, path: str):
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.index(b"end_header")
    body = data[data.index(b"\n", end) + 1:]
    elems, fmt = [], None
    for raw in data[:end].decode("latin-1").split("\n")[1:]:
        tok = raw.split()
        if not tok:
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elems.append((tok[1], int(tok[2]), []))
        elif tok[0] == "property":
            if tok[1] == "list":
                elems[-1][2].append((tok[4], _PLY_TYPES[tok[3]], _PLY_TYPES[tok[2]]))
            else:
                elems[-1][2].append((tok[2], _PLY_TYPES[tok[1]], None))
    if fmt == "ascii":
        toks = body.split()
        pos = [0]

        def read(code):
            t = toks[pos[0]].decode()
            pos[0] += 1
            return _f(t) if code in "fd" else int(t)
    else:
        end_ch = "<" if fmt == "binary_little_endian" else ">"
        pos = [0]

        def read(code):
            v = struct.unpack_from(end_ch + code, body, pos[0])[0]
            pos[0] += struct.calcsize(code)
            return F32(v) if code in "fd" else v
    values = {}
    for name, count, props in elems:
        recs = []
        for _ in range(count):
            rec = {}
            for pname, code, ccode in props:
                if ccode is not None:
                    k = read(ccode)
                    rec[pname] = [read(code) for _ in range(k)]
                else:
                    rec[pname] = read(code)
            recs.append(rec)
        values.setdefault(name, recs)
    verts = values["vertex"]
    has_n = all(k in verts[0] for k in ("nx", "ny", "nz")) if verts else False
    has_uv = has_n and all(k in verts[0] for k in ("u", "v"))
    gid = len(mod.materials)
    for face in values["face"]:
        ids = face["vertex_indices"]
        assert len(ids) == 3
        vs = [[verts[i]["x"], verts[i]["y"], verts[i]["z"]] for i in ids]
        e1, e2 = _sub(vs[1], vs[0]), _sub(vs[2], vs[0])
        mod.prims.append((gid, len(mod.prims), vs[0], e1, e2))
        mod.geometric_normals.append(("ply", e1, e2))
        for i in ids:
            if has_n:
                mod.shading_normals.append([verts[i]["nx"], verts[i]["ny"], verts[i]["nz"]])
            mod.tex_coords.append([verts[i]["u"], verts[i]["v"]] if has_uv else [F32(0.0), F32(0.0)])
        for v in vs:
            mod._insert(v)
    mod.materials.append(default_material())


def load(paths) -> Model:
    """model::load over a list of file names."""
    mod = Model()
    exts = [os.path.splitext(p)[1] for p in paths]
    same = all(e == exts[0] for e in exts)
    if same and exts[0] in (".obj", ".OBJ"):
        load_obj(mod, paths)
    else:
        for p, e in zip(paths, exts):
            if e in (".obj", ".OBJ"):
                load_obj(mod, [p])
            elif e in (".ply", ".PLY"):
                load_ply(mod, p)
            else:
                raise ValueError(f"unsupported format {e}")
    return mod


def triangles(mod: Model, dtype) -> np.ndarray:
    out = np.zeros(len(mod.prims), dtype)
    for i, (g, p, v1, e1, e2) in enumerate(mod.prims):
        out[i] = (g, p, v1, e1, e2)
    return out


def geometric_normals(mod: Model) -> np.ndarray:
    out = np.zeros((len(mod.geometric_normals), 3), np.float32)
    for i, (kind, e1, e2) in enumerate(mod.geometric_normals):
        c = _cross(e1, e2)
        out[i] = _normalize(c) if kind == "obj" else c
    return out


def pt_material(m: dict, diffuse_fallback: bool):
    """map_material restricted to emissive / matte: (type, cd, kd, ce, ls)."""
    if math.sqrt(float(_dot(m["ce"], m["ce"]))) > 0.0:
        return (1, [0.0] * 3, 0.0, list(m["ce"]), 1.0)
    if m["illum"] == 1 or diffuse_fallback:
        return (0, list(m["cd"]), 1.0, [0.0] * 3, 0.0)
    raise ValueError("unsupported material")
