"""Readers for the reference's on-disk character data and the embedded-graph
build, so the released character data (not only synthetic scenes) feeds the
GPU path.

* camera calibration: ``read_calibration`` (Projects/AdditionalUtils/CameraReader.py:16-80)
* meshes: ``read_off`` / ``write_off`` (Skeletool/Mesh/trimesh.cpp:610-723 loadOff,
  writeOff / writeCOff) and ``read_obj`` (trimesh.cpp:337-520 loadObj)
* skeleton: ``read_skeleton`` for "Skeletool Skeleton Definition V1.0"
  (Skeletool/Character/skeleton.cpp:205-545 loadSkeleton10b / loadSkeleton10) and
  "Skeleton v0.9" / "v0.10" / "v0.3" / "v0.1" (skeleton.cpp:533-1354), whose DOFs
  ``read_dof`` takes from a .dof file (skeleton.cpp:1679-2095 loadDof / loadDof05),
  Pinocchio .pskel (:1357-1455) and BVH hierarchies (:1459-1676)
* skinning: ``read_skin`` (Skeletool/Character/skinnedcharacter.cpp:154-206 loadSkinningData),
  ``read_pskin`` (:210-268 loadPinocchioSkinning), ``read_maya_skin`` (:272-362 loadMayaSkinningData)
* motion: ``read_motion`` (Skeletool/Parameters/parameter_container.cpp:129-190,330-404)
* character: ``read_character`` / ``load_character`` (skinnedcharacter.cpp:425-557 loadCharacter)
* embedded graph: ``build_embedded_graph`` (Skeletool/Mesh/EmbeddedGraph.cpp:5-460)

All of this is host-side set-up that runs once per character; the arrays it
returns are what the GPU entry points take.
"""
from __future__ import annotations

import dataclasses
import os
from dataclasses import dataclass, field

import numpy as np

from . import scenes


# ----------------------------------------------------------------------------
# Camera calibration
# ----------------------------------------------------------------------------

@dataclass
class Calibration:
    extrinsics: np.ndarray   # [C, 12] float32, 3x4 [R|t] row-major per camera
    intrinsics: np.ndarray   # [C, 9] float32, 3x3 row-major per camera (rescaled to width x height)
    width: int
    height: int
    original_size: np.ndarray  # [C, 2] (U, V) of the calibration


def read_calibration(path: str, render_u: int = -1, render_v: int = -1) -> Calibration:
    """Cameras of a calibration file (CameraReader.py:16-80): per camera an
    ``intrinsic`` line of 16 numbers (4x4 row-major; the top-left 3x3 is
    kept), an ``extrinsic`` line of 16 numbers (the first 12, 3x4, are kept)
    and a ``size U V`` line.  With a render resolution the focal lengths and
    principal point are rescaled by render / size of the first camera, as the
    reference does."""
    ext, intr, su, sv = [], [], [], []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if t[0] == "intrinsic":
                intr.extend(float(t[i]) for i in range(1, len(t)) if i % 4 != 0 and i <= 11)
            elif t[0] == "extrinsic":
                ext.extend(float(t[i]) for i in range(1, len(t)) if i <= 12)
            elif t[0] == "size":
                su.append(float(t[1]))
                sv.append(float(t[2]))
    if not ext or len(ext) % 12 or len(intr) != len(ext) // 12 * 9 or not su:
        raise ValueError(f"{path}: not a camera calibration file")
    C = len(ext) // 12
    e = np.asarray(ext, np.float32).reshape(C, 12)
    k = np.asarray(intr, np.float32).reshape(C, 3, 3)
    if render_u > 0 and render_v > 0:
        u0, v0 = np.float32(su[0]), np.float32(sv[0])
        for c in range(C):
            k[c, 0, 0] = (k[c, 0, 0] / u0) * np.float32(render_u)
            k[c, 1, 1] = (k[c, 1, 1] / v0) * np.float32(render_v)
            k[c, 0, 2] = (k[c, 0, 2] / u0) * np.float32(render_u)
            k[c, 1, 2] = (k[c, 1, 2] / v0) * np.float32(render_v)
        w, h = int(render_u), int(render_v)
    else:
        w, h = int(su[0]), int(sv[0])
    return Calibration(e, k.reshape(C, 9), w, h, np.stack([su, sv], 1).astype(np.float32))


def write_calibration(path: str, extrinsics: np.ndarray, intrinsics: np.ndarray, sizes) -> None:
    """Writes cameras in the layout read_calibration reads (4x4 rows padded
    with the homogeneous row/column)."""
    E = np.asarray(extrinsics, np.float64).reshape(-1, 3, 4)
    K = np.asarray(intrinsics, np.float64).reshape(-1, 3, 3)
    with open(path, "w") as fh:
        for c in range(E.shape[0]):
            k4 = np.eye(4)
            k4[:3, :3] = K[c]
            e4 = np.eye(4)
            e4[:3] = E[c]
            fh.write(f"name cam{c}\n")
            fh.write("intrinsic " + " ".join(repr(float(x)) for x in k4.reshape(-1)) + "\n")
            fh.write("extrinsic " + " ".join(repr(float(x)) for x in e4.reshape(-1)) + "\n")
            fh.write(f"size {sizes[c][0]} {sizes[c][1]}\n")


# ----------------------------------------------------------------------------
# Meshes
# ----------------------------------------------------------------------------

@dataclass
class Mesh:
    verts: np.ndarray                   # [N, 3] float32
    faces: np.ndarray                   # [F, 3] int32
    colors: np.ndarray | None = None    # [N, 3] in [0, 1]
    texcoords: np.ndarray | None = None  # OBJ: [T, 2] table; STOFF: [N, 2] per vertex
    face_tindex: np.ndarray | None = None  # OBJ: [F, 3] texcoord index per face corner (-1 if none)

    def face_texcoords(self) -> np.ndarray:
        """[F, 3, 2] per-corner texture coordinates, the renderer's
        ``texture_coordinates`` attribute layout."""
        if self.texcoords is None:
            raise ValueError("mesh has no texture coordinates")
        if self.face_tindex is not None:
            if (self.face_tindex < 0).any():
                raise ValueError("some faces have no texture coordinates")
            return self.texcoords[self.face_tindex].astype(np.float32)
        return self.texcoords[self.faces].astype(np.float32)


def read_off(path: str, flip: bool = False) -> Mesh:
    """OFF / COFF (colours 0-255, divided by 255) / STOFF (u v per vertex)
    triangle meshes (trimesh.cpp:610-688).  Faces are ``3 a b c`` lines;
    ``flip`` swaps the last two indices as the reference's loader option."""
    with open(path) as fh:
        lines = [ln.split() for ln in fh]
    lines = [t for t in lines if t]
    head = lines[0]
    if len(head) != 1 or head[0] not in ("OFF", "COFF", "STOFF"):
        raise ValueError(f"{path}: expected an OFF header")
    coff, stoff = head[0] == "COFF", head[0] == "STOFF"
    nv, nf = int(lines[1][0]), int(lines[1][1])
    vrows = lines[2:2 + nv]
    frows = lines[2 + nv:2 + nv + nf]
    if len(vrows) != nv or len(frows) != nf:
        raise ValueError(f"{path}: truncated")
    verts = np.array([[float(x) for x in t[:3]] for t in vrows], np.float32)
    colors = texcoords = None
    if coff:
        colors = np.array([[np.float32(float(x)) / np.float32(255.0) for x in t[3:6]] for t in vrows], np.float32)
    if stoff:
        texcoords = np.array([[float(x) for x in t[3:5]] for t in vrows], np.float32)
    for t in frows:
        if len(t) != 4:
            raise ValueError(f"{path}: only triangles are supported ({len(t)} tokens in a face line)")
    faces = np.array([[int(x) for x in t[1:4]] for t in frows], np.int32).reshape(-1, 3)
    if flip:
        faces = faces[:, [0, 2, 1]]
    return Mesh(verts, faces, colors, texcoords)


def write_off(path: str, verts: np.ndarray, faces: np.ndarray, colors: np.ndarray | None = None) -> None:
    """OFF as trimesh::writeOff; with colours, COFF as trimesh::writeCOff,
    which writes colours as int(c * 255) and each face as (a, c, b)
    (trimesh.cpp:697-743)."""
    with open(path, "w") as fh:
        fh.write("COFF\n" if colors is not None else "OFF\n")
        fh.write(f"{len(verts)} {len(faces)} 0\n")
        for i, v in enumerate(np.asarray(verts, np.float32)):
            line = f"{repr(float(v[0]))} {repr(float(v[1]))} {repr(float(v[2]))}"
            if colors is not None:
                c = np.asarray(colors[i], np.float32)
                line += " " + " ".join(str(int(np.float32(x) * np.float32(255.0))) for x in c)
            fh.write(line + "\n")
        for f in np.asarray(faces):
            a, b, c = (int(f[0]), int(f[2]), int(f[1])) if colors is not None else (int(f[0]), int(f[1]), int(f[2]))
            fh.write(f"3 {a} {b} {c}\n")


def read_obj(path: str) -> Mesh:
    """Wavefront OBJ as trimesh::loadObj (trimesh.cpp:337-520): ``v x y z
    [r g b]``, ``vc r g b``, ``vt u v``, ``f`` with 1-based ``v[/vt[/vn]]``
    corners; polygons are split into fans (0, d-1, d).  Texture indices are
    kept when every corner of a face has one."""
    verts, cols, vt, faces, tind = [], [], [], [], []
    with open(path) as fh:
        for line in fh:
            t = line.rstrip("\r\n").split()
            if not t:
                continue
            if t[0] == "v":
                verts.append([float(x) for x in t[1:4]])
                if len(t) > 5:
                    cols.append([float(x) for x in t[4:7]])
            elif t[0] == "vc":
                cols.append([float(x) for x in t[1:4]])
            elif t[0] == "vt":
                vt.append([float(x) for x in t[1:3]])
            elif t[0] == "f":
                vi, ti = [], []
                for tok in t[1:]:
                    sub = tok.split("/")
                    vi.append(int(sub[0]) - 1)
                    if len(sub) > 1 and sub[1]:
                        ti.append(int(sub[1]) - 1)
                for d in range(2, len(vi)):
                    faces.append([vi[0], vi[d - 1], vi[d]])
                    tind.append([ti[0], ti[d - 1], ti[d]] if len(ti) == len(vi) else [-1, -1, -1])
    V = np.asarray(verts, np.float32).reshape(-1, 3)
    Fc = np.asarray(faces, np.int32).reshape(-1, 3)
    if Fc.size and (Fc.min() < 0 or Fc.max() >= len(V)):
        raise ValueError(f"{path}: vertex index out of range")
    colors = np.asarray(cols, np.float32).reshape(-1, 3) if len(cols) == len(V) and cols else None
    texc = np.asarray(vt, np.float32).reshape(-1, 2) if vt else None
    ti = np.asarray(tind, np.int32).reshape(-1, 3) if vt else None
    return Mesh(V, Fc, colors, texc, ti)


# ----------------------------------------------------------------------------
# Skeleton (V1.0; older layouts + .dof), skinning, character
# ----------------------------------------------------------------------------

@dataclass
class SkeletonFile:
    names: list[str]
    types: list[str]             # "revolute" | "prismatic"
    parents: list[int]           # -1 for the root
    offsets: np.ndarray          # [J, 3] as written
    axes: np.ndarray             # [J, 3] (revolute axes normalised on load)
    scales: np.ndarray           # [J]
    markers: list[tuple]         # (name, parent joint, type, offset [3])
    scaling: list[tuple]         # (name, [joint ids])
    dofs: list[tuple]            # (name, limits or None, [(joint id, weight)])
    joint_limits: np.ndarray | None = None  # pre-V1.0: [J, 2] per-joint limits (lo == hi: none)
    dof_file_required: bool = False         # pre-V1.0: the DOFs come from a .dof file (read_dof)


def _lines(path):
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if t:
                yield line.strip(), t


def read_skeleton(path: str) -> SkeletonFile:
    """A "Skeletool Skeleton Definition V1.0" file (skeleton.cpp:246-545):
    ``joints: J`` then ``<name> <revolute|prismatic> <parent> ox oy oz ax ay az
    scale`` lines (a parent name that is no joint makes the root), ``markers:
    P`` lines, ``scaling joints: S`` groups and ``dofs: D`` blocks of
    ``<name> <count>``, ``limits lo hi`` or ``nolimits``, then ``<joint>
    <weight>`` lines.  Joint references in markers, scaling groups and DOFs
    match the first joint whose name starts with the given name (the
    reference's getJointByName prefix lookup).  "Skeleton v0.9" / "v0.10" /
    "v0.3" / "v0.1" files go to the older layouts' readers (no DOFs: see
    read_dof); .pskel (Pinocchio) and .bvh files are told by their extension
    first, as skeleton.cpp:2097-2130 does."""
    ext = os.path.splitext(path)[1]
    if ext == ".pskel":
        return _read_skeleton_pinocchio(path)
    if ext == ".bvh":
        return _read_skeleton_bvh(path)
    it = _lines(path)
    header, _ = next(it)
    if header in ("Skeleton v0.9", "Skeleton v0.10"):
        return _read_skeleton09(path)
    if header in ("Skeleton v0.3", "Skeleton v0.1"):
        return _read_skeleton_v01_v03(path, header.split()[1])
    if header != "Skeletool Skeleton Definition V1.0":
        raise ValueError(f"{path}: '{header}' skeletons are not supported "
                         "(V1.0, and v0.9 / v0.10 / v0.3 / v0.1 with a .dof file, are)")
    _, t = next(it)
    if len(t) != 2 or t[0] != "joints:":
        raise ValueError(f"{path}: expected 'joints:'")
    nj = int(t[1])
    names, types, pnames, off, ax, sc = [], [], [], [], [], []
    for _ in range(nj):
        _, t = next(it)
        if len(t) != 10 or t[1] not in ("revolute", "prismatic"):
            raise ValueError(f"{path}: bad joint line {t}")
        names.append(t[0])
        types.append(t[1])
        pnames.append(t[2])
        off.append([float(x) for x in t[3:6]])
        a = np.asarray([float(x) for x in t[6:9]], np.float32)
        if t[1] == "revolute":
            a = a / np.float32(np.sqrt(np.float32(a @ a)))
        ax.append(a)
        sc.append(float(t[9]))

    def exact(n):
        return names.index(n) if n in names else -1

    def prefix(n):
        for i, jn in enumerate(names):
            if jn.startswith(n):
                return i
        return -1

    parents = [exact(p) for p in pnames]
    if sum(p < 0 for p in parents) != 1:
        raise ValueError(f"{path}: the skeleton needs exactly one root joint")
    _, t = next(it)
    if len(t) != 2 or t[0] != "markers:":
        raise ValueError(f"{path}: expected 'markers:'")
    markers = []
    for _ in range(int(t[1])):
        _, t = next(it)
        markers.append((t[0], prefix(t[1]), t[2], np.asarray([float(x) for x in t[3:6]], np.float32)))
    _, t = next(it)
    if len(t) != 3 or t[0] != "scaling" or t[1] != "joints:":
        raise ValueError(f"{path}: expected 'scaling joints:'")
    scaling = []
    for _ in range(int(t[2])):
        _, t = next(it)
        ids = []
        for _ in range(int(t[1])):
            _, r = next(it)
            j = prefix(r[0])
            if j >= 0:
                ids.append(j)
        scaling.append((t[0], ids))
    _, t = next(it)
    if len(t) != 2 or t[0] != "dofs:":
        raise ValueError(f"{path}: expected 'dofs:'")
    dofs = []
    for _ in range(int(t[1])):
        _, t = next(it)
        name, cnt = t[0], int(t[1])
        _, lim = next(it)
        limits = (float(lim[1]), float(lim[2])) if lim[0] == "limits" and len(lim) == 3 else None
        infl = []
        for _ in range(cnt):
            _, r = next(it)
            j = prefix(r[0])
            if j < 0:
                raise ValueError(f"{path}: unknown joint '{r[0]}' in dof '{name}'")
            infl.append((j, float(r[1])))
        dofs.append((name, limits, infl))
    return SkeletonFile(names, types, parents, np.asarray(off, np.float32), np.asarray(ax, np.float32),
                        np.asarray(sc, np.float32), markers, scaling, dofs)


def write_skeleton(path: str, desc: dict) -> None:
    """A V1.0 skeleton file (the layout read_skeleton reads) from a kinematic
    description with joint ``names``: one DOF per driven joint, weight 1,
    named after it, no limits, no markers or scaling groups."""
    names = desc["names"]
    J = len(names)
    with open(path, "w") as fh:
        fh.write("Skeletool Skeleton Definition V1.0\n")
        fh.write(f"joints: {J}\n")
        for j in range(J):
            p = int(desc["parent"][j])
            t = "revolute" if int(desc["type"][j]) == scenes.JOINT_REVOLUTE else "prismatic"
            o, a = desc["offset"][j], desc["axis"][j]
            fh.write(f"{names[j]} {t} {names[p] if p >= 0 else 'none'} "
                     + " ".join(repr(float(x)) for x in (*o, *a)) + " 1.0\n")
        fh.write("markers: 0\nscaling joints: 0\n")
        driven = [j for j in range(J) if int(desc["dof"][j]) >= 0]
        driven.sort(key=lambda j: int(desc["dof"][j]))
        fh.write(f"dofs: {len(driven)}\n")
        for j in driven:
            fh.write(f"dof_{names[j]} 1\nnolimits\n{names[j]} 1.0\n")


def _token_rows(path):
    with open(path) as fh:
        return [ln.split() for ln in fh]


def _first_prefix(names, n):
    for i, jn in enumerate(names):
        if jn.startswith(n):
            return i
    return -1


def _read_skeleton09(path: str) -> SkeletonFile:
    """A "Skeleton v0.9" / "Skeleton v0.10" file (skeleton.cpp:533-872), the
    layout whose DOFs live in a separate .dof file (read_dof).
    ``BEGIN joints global`` is followed by joint lines ``<name> <parent>
    <r|t|s> ox oy oz ax ay az ...`` (at least 10 tokens; revolute, prismatic,
    scaling) whose offsets are global: each is stored relative to the global
    offset of the earlier joint named exactly as its parent, and a parent
    name that is no earlier joint makes the root.  ``BEGIN limits`` lines
    ``<joint> <weight> lo hi`` set per-joint limits (exact names; unknown ones
    skipped).  ``BEGIN proxies global level 0`` lines ``<name> <parent> ox oy
    oz size`` are markers, their parent found by prefix and their global
    offset made relative like the joints'.  The marker colour block after
    them is not read."""
    rows = _token_rows(path)
    n, i = len(rows), 0
    while i < n and not rows[i]:
        i += 1
    i += 1
    while i < n and len(rows[i]) < 2:
        i += 1
    if i >= n or rows[i] != ["BEGIN", "joints", "global"]:
        raise ValueError(f"{path}: expected 'BEGIN joints global'")
    i += 1
    kinds = {"r": "revolute", "t": "prismatic", "s": "scaling"}
    names, types, parents, glob, off, ax = [], [], [], [], [], []
    while i < n and len(rows[i]) >= 10:
        t = rows[i]
        i += 1
        if t[2] not in kinds:
            raise ValueError(f"{path}: unknown joint type '{t[2]}' of joint '{t[0]}'")
        g = np.asarray([float(x) for x in t[3:6]], np.float32)
        a = np.asarray([float(x) for x in t[6:9]], np.float32)
        if t[2] == "r":
            a = a / np.float32(np.sqrt(np.float32(a @ a)))
        p = names.index(t[1]) if t[1] in names else -1
        names.append(t[0])
        types.append(kinds[t[2]])
        parents.append(p)
        glob.append(g)
        off.append(g - glob[p] if p >= 0 else g)
        ax.append(a)
    J = len(names)
    if sum(p < 0 for p in parents) != 1:
        raise ValueError(f"{path}: the skeleton needs exactly one root joint")
    limits = np.zeros((J, 2), np.float32)
    while i < n and len(rows[i]) < 2:
        i += 1
    if i < n and rows[i] == ["BEGIN", "limits"]:
        i += 1
        while i < n and (not rows[i] or rows[i][0] != "BEGIN"):
            t = rows[i]
            i += 1
            if len(t) >= 4 and t[0] in names:
                limits[names.index(t[0])] = (float(t[2]), float(t[3]))
    markers = []
    while i < n and len(rows[i]) < 5:
        i += 1
    if i < n and rows[i] == ["BEGIN", "proxies", "global", "level", "0"]:
        i += 1
        while i < n and len(rows[i]) == 6:
            t = rows[i]
            i += 1
            o = np.asarray([float(x) for x in t[2:5]], np.float32)
            if t[1] in names:
                o = o - glob[names.index(t[1])]
            markers.append((t[0], _first_prefix(names, t[1]), "proxy", o))
    return SkeletonFile(names, types, parents, np.asarray(off, np.float32).reshape(J, 3),
                        np.asarray(ax, np.float32).reshape(J, 3), np.ones(J, np.float32), markers, [], [],
                        joint_limits=limits, dof_file_required=True)


def _read_skeleton_v01_v03(path: str, version: str) -> SkeletonFile:
    """"Skeleton v0.3" (skeleton.cpp:874-1019) and "Skeleton v0.1"
    (:1022-1354) files, whose DOFs also come from a .dof file.  ``joints: J``
    then J joint lines ``<name> <parent> <r|t|s> ox oy oz ax ay az`` (parent:
    an earlier joint of that exact name, else the root; v0.3 has no scaling
    joints, a v0.1 scaling joint has zero offset and no axis).

    - v0.3: offsets are parent-relative; ``proxies: P`` lines ``<parent> ox oy
      oz size -1 r g b`` (parent by prefix, offset as written, named by index).
    - v0.1 with limits inline (the first joint line has 11 values): offsets
      relative, ``lo hi`` after the axis; ``proxies: P`` lines ``<name>
      <joint index> ox oy oz size -1 h s v`` (the first proxy line's length
      picks the form for all of them), offsets as written.
    - v0.1 without: offsets global (made relative to the exact-named parent's
      global offset; a scaling joint's is zero, placing it at the origin), then ``BEGIN
      limits`` lines ``<joint> <weight> lo hi`` up to ``End::``; 6-value
      proxy lines ``<tag> <parent> ox oy oz size`` (exact parent name, global
      offsets made relative, named by index)."""
    rows = _token_rows(path)
    n, i = len(rows), 0
    while i < n and not rows[i]:
        i += 1
    i += 1
    while i < n and len(rows[i]) < 2:
        i += 1
    if i >= n or len(rows[i]) != 2 or rows[i][0] != "joints:":
        raise ValueError(f"{path}: expected 'joints:'")
    nj = int(rows[i][1])
    jrows = rows[i + 1:i + 1 + nj]
    i += 1 + nj
    if len(jrows) != nj or nj == 0:
        raise ValueError(f"{path}: expected {nj} joint lines")
    inline = version == "v0.1" and len(jrows[0]) == 11
    need = 11 if inline else 9
    kinds = {"r": "revolute", "t": "prismatic"} if version == "v0.3" else \
        {"r": "revolute", "t": "prismatic", "s": "scaling"}
    global_offsets = version == "v0.1" and not inline
    names, types, parents, glob, off, ax = [], [], [], [], [], []
    limits = np.zeros((nj, 2), np.float32)
    for j, t in enumerate(jrows):
        if len(t) != need:
            raise ValueError(f"{path}: joint line {t} has {len(t)} values, expected {need}")
        if t[2] not in kinds:
            raise ValueError(f"{path}: unknown joint type '{t[2]}' of joint '{t[0]}'")
        o = np.asarray([float(x) for x in t[3:6]], np.float32)
        a = np.asarray([float(x) for x in t[6:9]], np.float32)
        if t[2] == "r":
            a = a / np.float32(np.sqrt(np.float32(a @ a)))
        elif t[2] == "s":
            o = np.zeros(3, np.float32)
            a = np.zeros(3, np.float32)
        p = names.index(t[1]) if t[1] in names else -1
        names.append(t[0])
        types.append(kinds[t[2]])
        parents.append(p)
        glob.append(o)
        off.append(o - glob[p] if global_offsets and p >= 0 else o)
        ax.append(a)
        if inline:
            limits[j] = (float(t[9]), float(t[10]))
    if sum(p < 0 for p in parents) != 1:
        raise ValueError(f"{path}: the skeleton needs exactly one root joint")
    if global_offsets:
        while i < n and len(rows[i]) < 2:
            i += 1
        if i >= n or rows[i] != ["BEGIN", "limits"]:
            raise ValueError(f"{path}: expected 'BEGIN limits'")
        i += 1
        while i < n and (not rows[i] or rows[i][0] != "End::"):
            t = rows[i]
            i += 1
            if len(t) >= 4 and t[0] in names:
                limits[names.index(t[0])] = (float(t[2]), float(t[3]))
        i += 1
    while i < n and len(rows[i]) < 2:
        i += 1
    if i >= n or len(rows[i]) != 2 or rows[i][0] != "proxies:":
        raise ValueError(f"{path}: expected 'proxies:'")
    npx = int(rows[i][1])
    prows = rows[i + 1:i + 1 + npx]
    if len(prows) != npx:
        raise ValueError(f"{path}: expected {npx} proxy lines")
    markers = []
    coloured = bool(prows) and len(prows[0]) == 10
    for k, t in enumerate(prows):
        want = 9 if version == "v0.3" else (10 if coloured else 6)
        if len(t) != want:
            raise ValueError(f"{path}: proxy line {t} has {len(t)} values, expected {want}")
        if version == "v0.3":
            o = np.asarray([float(x) for x in t[1:4]], np.float32)
            markers.append((str(k), _first_prefix(names, t[0]), "proxy", o))
        elif coloured:
            pj = int(t[1])
            markers.append((t[0], pj if 0 <= pj < nj else -1, "proxy",
                            np.asarray([float(x) for x in t[2:5]], np.float32)))
        else:
            pj = names.index(t[1]) if t[1] in names else -1
            o = np.asarray([float(x) for x in t[2:5]], np.float32)
            markers.append((str(k), pj, "proxy", o - glob[pj] if pj >= 0 else o))
    return SkeletonFile(names, types, parents, np.asarray(off, np.float32).reshape(nj, 3),
                        np.asarray(ax, np.float32).reshape(nj, 3), np.ones(nj, np.float32), markers, [], [],
                        joint_limits=limits, dof_file_required=True)


def _read_skeleton_pinocchio(path: str) -> SkeletonFile:
    """A Pinocchio skeleton (.pskel, skeleton.cpp:1357-1455): one line per
    joint, ``<name> ax ay az x y z <parent>``, all revolute (axes
    normalised).  The parent is the earlier joint of that exact name (none:
    the root); the position is global with its last two components swapped,
    (x, z, y), and is made relative to the position of the joint whose index
    the parent field reads as.  Each joint gets its own DOF (weight 1, no
    limits)."""
    rows = [t for t in _token_rows(path) if t]
    names, parents, off, ax, glob, dofs = [], [], [], [], [], []
    for t in rows:
        if len(t) != 8:
            raise ValueError(f"{path}: joint line {t} has {len(t)} values, expected 8")
        a = np.asarray([float(x) for x in t[1:4]], np.float32)
        a = a / np.float32(np.sqrt(np.float32(a @ a)))
        g = np.asarray([float(t[4]), float(t[6]), float(t[5])], np.float32)
        try:
            ip = int(t[7])
        except ValueError:
            raise ValueError(f"{path}: parent field '{t[7]}' is not a joint index") from None
        p = names.index(t[7]) if t[7] in names else -1
        glob.append(g)
        off.append(g - glob[ip] if 0 <= ip < len(glob) else g)
        dofs.append((t[0], None, [(len(names), 1.0)]))
        names.append(t[0])
        parents.append(p)
        ax.append(a)
    J = len(names)
    if J == 0 or sum(p < 0 for p in parents) != 1:
        raise ValueError(f"{path}: the skeleton needs exactly one root joint")
    return SkeletonFile(names, ["revolute"] * J, parents, np.asarray(off, np.float32).reshape(J, 3),
                        np.asarray(ax, np.float32).reshape(J, 3), np.ones(J, np.float32), [], [], dofs)


def _bone_name(joint_name: str) -> str:
    k = joint_name.rfind("_")
    return joint_name if k < 0 else joint_name[:k]


_BVH_CHANNELS = (("Xposition", "_tx", "prismatic", (1, 0, 0)), ("Yposition", "_ty", "prismatic", (0, 1, 0)),
                 ("Zposition", "_tz", "prismatic", (0, 0, 1)), ("Xrotation", "_rx", "revolute", (1, 0, 0)),
                 ("Yrotation", "_ry", "revolute", (0, 1, 0)), ("Zrotation", "_rz", "revolute", (0, 0, 1)))


def _read_skeleton_bvh(path: str) -> SkeletonFile:
    """The HIERARCHY part of a BVH file (skeleton.cpp:1459-1676).  Every
    channel of a ROOT (bone name "Root") or JOINT becomes a joint named
    ``<bone>_tx|ty|tz|rx|ry|rz`` (prismatic / revolute about the unit axis),
    chained in channel order under the previous joint, with a DOF of its own
    (weight 1, no limits).  A channel joint's offset is the bone's OFFSET
    minus the offset of its parent joint, or, when that parent's own parents
    are channels of the same bone, of the first of them; the root's first
    channel keeps the OFFSET.  An End Site is a revolute joint
    ``<parent bone>_end`` about x with the OFFSET as written and no DOF.
    Each closing brace pops the innermost bone's channel chain.  Bone names
    are joint names up to their last '_'.  The MOTION part is not read."""
    rows = [t for t in _token_rows(path) if t]
    if not rows or rows[0] != ["HIERARCHY"]:
        raise ValueError(f"{path}: expected 'HIERARCHY'")
    names, types, parents, off, ax, dofs = [], [], [], [], [], []
    stack = []
    i, n = 1, len(rows)

    def row(k, what):
        if k >= n:
            raise ValueError(f"{path}: ended before {what}")
        return rows[k]

    t = row(i, "MOTION")
    while "MOTION" not in t[0]:
        is_root, is_joint, is_end = "ROOT" in t[0], "JOINT" in t[0], "End" in t[0]
        if not (is_root or is_joint or is_end) or (is_joint and len(t) < 2):
            raise ValueError(f"{path}: expected ROOT, JOINT or End Site, got {t}")
        bone = "Root" if is_root else (t[1] if is_joint else "End")
        i += 1
        if "{" not in row(i, "'{'")[0]:
            raise ValueError(f"{path}: expected '{{' after {t}")
        i += 1
        o = row(i, "OFFSET")
        if "OFFSET" not in o[0] or len(o) < 4:
            raise ValueError(f"{path}: expected 'OFFSET x y z', got {o}")
        offset = np.asarray([float(x) for x in o[1:4]], np.float32)
        if is_end:
            if not stack:
                raise ValueError(f"{path}: End Site outside a joint")
            p = stack[-1]
            names.append(_bone_name(names[p]) + "_end")
            types.append("revolute")
            parents.append(p)
            off.append(offset)
            ax.append(np.asarray((1, 0, 0), np.float32))
            i += 1
            if "}" not in row(i, "'}'")[0]:
                raise ValueError(f"{path}: expected '}}' closing an End Site")
        else:
            i += 1
            c = row(i, "CHANNELS")
            if "CHANNELS" not in c[0] or len(c) < 2 or len(c) < int(c[1]) + 2:
                raise ValueError(f"{path}: expected 'CHANNELS n ...', got {c}")
            for ch in c[2:2 + int(c[1])]:
                spec = next((s for s in _BVH_CHANNELS if s[0] in ch), None)
                if spec is None:
                    raise ValueError(f"{path}: unknown channel '{ch}'")
                jn = bone + spec[1]
                if not stack:
                    p, rel = -1, offset
                else:
                    p = stack[-1]
                    father, fj = off[p], parents[p]
                    while fj >= 0 and _bone_name(names[fj]) == _bone_name(jn):
                        father, fj = off[fj], parents[fj]
                    rel = offset - father
                dofs.append((jn, None, [(len(names), 1.0)]))
                stack.append(len(names))
                names.append(jn)
                types.append(spec[2])
                parents.append(p)
                off.append(rel)
                ax.append(np.asarray(spec[3], np.float32))
        i += 1
        while i < n and "}" in rows[i][0]:
            if stack:
                last = _bone_name(names[stack.pop()])
                while stack and _bone_name(names[stack[-1]]) == last:
                    stack.pop()
            i += 1
        t = row(i, "MOTION")
    J = len(names)
    if J == 0 or sum(p < 0 for p in parents) != 1:
        raise ValueError(f"{path}: the skeleton needs exactly one root joint")
    return SkeletonFile(names, types, parents, np.asarray(off, np.float32).reshape(J, 3),
                        np.asarray(ax, np.float32).reshape(J, 3), np.ones(J, np.float32), [], [], dofs)


def read_dof(path: str, skel: SkeletonFile) -> SkeletonFile:
    """The DOFs of a pre-V1.0 skeleton from its .dof file (skeleton.cpp:1679-2095),
    as a copy of ``skel`` with ``dofs`` filled in.

    - "Dofs v0.3": ``dofs: D`` then ``<name> <value> limits lo hi`` or
      ``<name> <value> nolimits`` lines.
    - "Dofs v0.1": ``dofs: D`` then ``[<id>] <name> <value>`` lines; limits
      come from the skeleton's per-joint limits.
    - "Dofs v0.5": ``BEGIN dofs`` (``<name> <value>`` lines), ``BEGIN
      scalingpose`` (initial values of scaling DOFs, not kept: a scaling joint
      driven by a DOF is rejected by skeleton_desc), ``BEGIN map`` up to the
      first empty line; limits from the skeleton, then DOFs that drive no
      joint are dropped.

    The joint map (after ``joints: J`` in v0.1 / v0.3) has lines ``<joint> =
    0 + w1 <dof1> + w2 <dof2> ...``; the joint is found by prefix, each DOF by
    exact name, and unknown joints or DOF names are skipped.  A per-joint
    limit (lo != hi) goes to the first DOF that drives the joint, later
    joints overriding earlier ones."""
    if not skel.names:
        raise ValueError("read_dof needs a loaded skeleton")
    rows = _token_rows(path)
    n, i = len(rows), 0
    while i < n and not rows[i]:
        i += 1
    header = " ".join(rows[i]) if i < n else ""
    if header not in ("Dofs v0.5", "Dofs v0.3", "Dofs v0.1"):
        raise ValueError(f"{path}: expected a 'Dofs v0.1', 'Dofs v0.3' or 'Dofs v0.5' header")
    version = header.split()[1]
    i += 1
    names, lims = [], []
    if version == "v0.5":
        while i < n and len(rows[i]) < 2:
            i += 1
        if i >= n or rows[i] != ["BEGIN", "dofs"]:
            raise ValueError(f"{path}: expected 'BEGIN dofs'")
        i += 1
        while i < n and len(rows[i]) >= 2 and rows[i][0] != "BEGIN":
            names.append(rows[i][0])
            lims.append(None)
            i += 1
        while i < n and rows[i] != ["BEGIN", "scalingpose"]:
            i += 1
        i += 1
        while i < n and len(rows[i]) >= 2 and rows[i][0] != "BEGIN":
            i += 1
        while i < n and rows[i] != ["BEGIN", "map"]:
            i += 1
        i += 1
        map_rows = []
        while i < n and rows[i]:
            map_rows.append(rows[i])
            i += 1
    else:
        while i < n and len(rows[i]) < 2:
            i += 1
        if i >= n or len(rows[i]) != 2 or rows[i][0] != "dofs:":
            raise ValueError(f"{path}: expected 'dofs:'")
        nd = int(rows[i][1])
        i += 1
        for _ in range(nd):
            t = rows[i] if i < n else []
            i += 1
            if version == "v0.3":
                if len(t) == 5 and t[2] == "limits":
                    lims.append((float(t[3]), float(t[4])))
                elif len(t) == 3 and t[2] == "nolimits":
                    lims.append(None)
                else:
                    raise ValueError(f"{path}: bad dof line {t}")
                names.append(t[0])
            else:
                if len(t) < 2:
                    raise ValueError(f"{path}: bad dof line {t}")
                names.append(t[1] if len(t) == 3 else t[0])
                lims.append(None)
        while i < n and len(rows[i]) < 2:
            i += 1
        if i >= n or len(rows[i]) != 2 or rows[i][0] != "joints:":
            raise ValueError(f"{path}: expected 'joints:'")
        nj = int(rows[i][1])
        map_rows = rows[i + 1:i + 1 + nj]
        if len(map_rows) != nj or any(len(t) < 3 for t in map_rows):
            raise ValueError(f"{path}: expected {nj} joint lines of at least 3 values")
    infl = [[] for _ in names]
    for t in map_rows:
        j = _first_prefix(skel.names, t[0])
        if j < 0:
            continue
        for d in range(3, len(t), 3):
            if t[d] != "+":
                break
            if d + 2 >= len(t):
                raise ValueError(f"{path}: truncated dof influence in {t}")
            if t[d + 2] in names:
                infl[names.index(t[d + 2])].append((j, float(t[d + 1])))
    if version in ("v0.1", "v0.5") and skel.joint_limits is not None:
        for j, (lo, hi) in enumerate(skel.joint_limits):
            if lo != hi:
                for q in range(len(names)):
                    if any(jj == j for jj, _ in infl[q]):
                        lims[q] = (float(lo), float(hi))
                        break
    dofs = [(nm, lm, inf) for nm, lm, inf in zip(names, lims, infl)]
    if version == "v0.5":
        dofs = [d for d in dofs if d[2]]
    return dataclasses.replace(skel, dofs=dofs, dof_file_required=False)


_DEG2RAD = np.float32(np.float32(np.pi) / np.float32(180.0))


def read_motion(path: str, skel: SkeletonFile):
    """Per-frame DOF vectors of a motion file (parameter_container.cpp:129-190,
    330-404): (frame numbers [F], DOFs [F, D] float32) for a skeleton with D
    DOFs.  "Skeletool Motion File V1.0" has lines ``<frame> p_0 ... p_D-1``
    (extra values ignored); "V0.1" the same with the sign of every DOF after
    the first three flipped; a BVH file (header HIERARCHY) has ``Frames: F``,
    ``Frame Time: t`` and F lines of D values after its MOTION line, in
    degrees for every DOF whose first joint is not prismatic (converted with
    the single-precision pi / 180).  .proj motion files are not read."""
    if os.path.splitext(path)[1] == ".proj":
        raise ValueError(f"{path}: .proj motion files are not supported")
    D = len(skel.dofs)
    rows = [t for t in _token_rows(path) if t]
    if not rows:
        raise ValueError(f"{path}: empty motion file")
    header = " ".join(rows[0])
    frames, params = [], []
    if header in ("Skeletool Motion File V1.0", "Skeletool Motion File V0.1"):
        for t in rows[1:]:
            if len(t) < D + 1:
                raise ValueError(f"{path}: {len(t) - 1} parameters in a frame line, the skeleton has {D}")
            p = np.asarray([float(x) for x in t[1:D + 1]], np.float32)
            if header.endswith("V0.1"):
                p[3:] = -p[3:]
            frames.append(int(t[0]))
            params.append(p)
    elif header == "HIERARCHY":
        i = next((k for k, t in enumerate(rows) if "MOTION" in t[0]), None)
        if i is None or i + 2 >= len(rows) or rows[i + 1][0] != "Frames:" or rows[i + 2][:2] != ["Frame", "Time:"]:
            raise ValueError(f"{path}: expected MOTION, 'Frames: F' and 'Frame Time: t'")
        nf = int(rows[i + 1][1])
        deg = np.array([skel.types[d[2][0][0]] != "prismatic" if d[2] else True for d in skel.dofs])
        lines = rows[i + 3:i + 3 + nf]
        if len(lines) != nf:
            raise ValueError(f"{path}: {len(lines)} frame lines, expected {nf}")
        for f, t in enumerate(lines):
            if len(t) != D:
                raise ValueError(f"{path}: {len(t)} values in a frame line, the skeleton has {D} DOFs")
            p = np.asarray([float(x) for x in t], np.float32)
            p[deg] = p[deg] * _DEG2RAD
            frames.append(f)
            params.append(p)
    else:
        raise ValueError(f"{path}: not a skeletool motion file or BVH")
    return np.asarray(frames, np.int64), np.asarray(params, np.float32).reshape(len(params), D)


def last_joint_by_name(skel: SkeletonFile, bone: str) -> int:
    """The last joint whose name starts with ``bone`` (skeleton.cpp:3667-3675),
    the joint a skinning bone binds to."""
    for i in range(len(skel.names) - 1, -1, -1):
        if skel.names[i].startswith(bone):
            return i
    raise ValueError(f"no joint matches bone '{bone}'")


def _global_transforms(parents, types, offset, axis, params) -> np.ndarray:
    """[J, 4, 4] joint transforms at joint parameters ``params`` [J]: G_j =
    G_parent translate(offset_j) M_j(p_j), M a rotation by p about the unit
    axis (Rodrigues) or a translation by p along it."""
    J = len(parents)
    G = np.zeros((J, 4, 4))
    for j in _topological(list(parents)):
        L = np.eye(4)
        L[:3, 3] = offset[j]
        M = np.eye(4)
        a, t = np.asarray(axis[j], np.float64), float(params[j])
        if int(types[j]) == scenes.JOINT_REVOLUTE:
            K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
            M[:3, :3] = np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * (K @ K)
        else:
            M[:3, 3] = a * t
        G[j] = (G[parents[j]] if parents[j] >= 0 else np.eye(4)) @ L @ M
    return G


def skeleton_desc(skel: SkeletonFile, bones: list[str], joint_ids: list[int] | None = None,
                  bind_dofs: np.ndarray | None = None) -> dict:
    """The kinematic description skinning.Skeleton takes.  Joint i's local
    transform is translate(offset_i * scale_i) then the rotation about (or
    translation along) its axis by its parameter, and the parameter is the
    weighted sum of the DOFs that drive it (dof.cpp:11-13); ``dof_map``
    [J, D] holds those weights.  When every joint is driven by at most one DOF
    with weight 1, ``dof`` indexes the DOF vector directly; otherwise the FK
    runs on the J joint parameters ``dofs @ dof_map.T`` (``dof`` = joint id).
    Skinning joint b is the joint its bone name resolves to (or
    ``joint_ids[b]``, the Maya export's joint index), bound at its global
    transform with all DOFs zero, or at the DOF vector ``bind_dofs`` (a
    character's initial pose, which the reference applies before it reads
    the skinning, skinnedcharacter.cpp:519-547)."""
    J, D = len(skel.names), len(skel.dofs)
    dof_map = np.zeros((J, D), np.float32)
    for d, (_, _, infl) in enumerate(skel.dofs):
        for j, w in infl:
            dof_map[j, d] += np.float32(w)
    simple = all(np.count_nonzero(dof_map[j]) <= 1 and dof_map[j].max(initial=0.0) in (0.0, 1.0)
                 and dof_map[j].min(initial=0.0) >= 0.0 for j in range(J))
    if simple:
        dof = np.array([int(np.argmax(dof_map[j])) if dof_map[j].any() else -1 for j in range(J)], np.int32)
        num_dofs = D
    else:
        dof = np.arange(J, dtype=np.int32)
        num_dofs = J
    # a scaling joint (type "s" of the pre-V1.0 layouts, a 2^p scaling) is the identity at p = 0:
    # undriven, it is a prismatic joint with no axis; driven, the FK cannot
    # represent it
    for j, t in enumerate(skel.types):
        if t == "scaling" and dof_map[j].any():
            raise ValueError(f"scaling joint '{skel.names[j]}' is driven by a DOF; the FK has no scaling joints")
    typ = np.array([scenes.JOINT_REVOLUTE if t == "revolute" else scenes.JOINT_PRISMATIC for t in skel.types],
                   np.int32)
    axes = skel.axes.astype(np.float64)
    axes[[t == "scaling" for t in skel.types]] = 0.0
    offset = skel.offsets.astype(np.float64) * skel.scales[:, None].astype(np.float64)
    # global positions at zero parameters: rotations are identity, so the
    # transform of each joint is the translation summed along its chain
    order = _topological(skel.parents)
    gpos = np.zeros((J, 3))
    for j in order:
        p = skel.parents[j]
        gpos[j] = (gpos[p] if p >= 0 else 0.0) + offset[j]
    G = None
    if bind_dofs is not None:
        bind_dofs = np.asarray(bind_dofs, np.float64).reshape(-1)
        if bind_dofs.shape[0] != D:
            raise ValueError(f"bind pose has {bind_dofs.shape[0]} DOFs, the skeleton {D}")
        G = _global_transforms(skel.parents, typ, offset, axes, dof_map.astype(np.float64) @ bind_dofs)
        gpos = G[:, :3, 3].copy()
    if joint_ids is not None:
        if len(joint_ids) != len(bones) or not all(0 <= j < J for j in joint_ids):
            raise ValueError(f"skinning joint indices {joint_ids} do not fit {J} joints")
        skin_joint = np.asarray(joint_ids, np.int32)
    else:
        skin_joint = np.array([last_joint_by_name(skel, b) for b in bones], np.int32)
    bind = np.zeros((len(bones), 12))
    for b, j in enumerate(skin_joint):
        if G is None:
            bind[b] = [1, 0, 0, -gpos[j, 0], 0, 1, 0, -gpos[j, 1], 0, 0, 1, -gpos[j, 2]]
        else:
            bind[b] = np.linalg.inv(G[j])[:3].reshape(-1)
    return dict(parent=np.asarray(skel.parents, np.int32), type=typ, dof=dof, offset=offset,
                axis=axes, skin_joint=skin_joint, skin_scale=np.ones(len(bones)),
                skin_bind=bind, num_dofs=num_dofs, names=list(skel.names), dof_map=dof_map,
                dof_names=[d[0] for d in skel.dofs], dof_limits=[d[1] for d in skel.dofs], joint_positions=gpos,
                joint_params=not simple)


def joint_parameters(desc: dict, dofs: np.ndarray) -> np.ndarray:
    """The FK input for a skeleton_desc: the DOF vector itself when each joint
    has at most one unit-weight DOF, else the joint parameters
    ``dofs @ dof_map.T``."""
    dofs = np.asarray(dofs)
    if not desc.get("joint_params", False):
        return dofs
    return dofs @ desc["dof_map"].T


def _topological(parents):
    J = len(parents)
    done, order = [False] * J, []

    def visit(j):
        chain = []
        while j >= 0 and not done[j]:
            chain.append(j)
            j = parents[j]
        for k in reversed(chain):
            done[k] = True
            order.append(k)

    for j in range(J):
        visit(j)
    return order


@dataclass
class SkinFile:
    bones: list[str]
    influences: list[list[tuple[int, float]]]  # per vertex: (bone index, weight)
    joint_ids: list[int] | None = None         # Maya export: each bone's skeleton joint index

    def top_k(self, K: int = 4, renormalise: bool = True):
        """[N, K] bone indices and weights: each vertex's K largest weights
        (ties keep the file order), padded with (0, 0).  Vertices with more
        than K influences are renormalised when ``renormalise``."""
        N = len(self.influences)
        idx = np.zeros((N, K), np.int32)
        w = np.zeros((N, K), np.float32)
        for v, inf in enumerate(self.influences):
            srt = sorted(range(len(inf)), key=lambda i: -inf[i][1])[:K]
            for k, i in enumerate(srt):
                idx[v, k], w[v, k] = inf[i][0], inf[i][1]
            if renormalise and len(inf) > K and w[v].sum() > 0:
                w[v] /= w[v].sum()
        return idx, w


def read_skin(path: str) -> SkinFile:
    """A "Skeletool character skinning file V1.0" (skinnedcharacter.cpp:154-206):
    ``bones:`` then one line of bone names, ``vertex weights:`` then one line
    per vertex, ``<vertex> <bone> <weight> <bone> <weight> ...`` (the first
    token is not used)."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    lines = [ln for ln in lines if ln]
    if lines[0] != "Skeletool character skinning file V1.0" or lines[1] != "bones:":
        raise ValueError(f"{path}: not a skeletool skin file")
    bones = lines[2].split()
    if lines[3] != "vertex weights:":
        raise ValueError(f"{path}: expected 'vertex weights:'")
    infl = []
    for ln in lines[4:]:
        t = ln.split()
        infl.append([(int(t[j]), float(t[j + 1])) for j in range(1, len(t) - 1, 2)])
    return SkinFile(bones, infl)


def read_pskin(path: str) -> SkinFile:
    """A Pinocchio skinning file (.pskin, skinnedcharacter.cpp:210-268): the
    .skin header, ``bones:`` and its line of bone names, ``vertex weights:``,
    then one line per vertex holding a weight for every bone in bone order;
    weights > 0 are kept.  Empty lines are skipped."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    lines = [ln for ln in lines if ln]
    if len(lines) < 4 or lines[0] != "Skeletool character skinning file V1.0" or lines[1] != "bones:":
        raise ValueError(f"{path}: not a skeletool skin file")
    bones = lines[2].split()
    if lines[3] != "vertex weights:":
        raise ValueError(f"{path}: expected 'vertex weights:'")
    infl = []
    for ln in lines[4:]:
        w = [float(x) for x in ln.split()]
        if len(w) > len(bones):
            raise ValueError(f"{path}: {len(w)} weights for {len(bones)} bones")
        infl.append([(j, x) for j, x in enumerate(w) if x > 0.0])
    return SkinFile(bones, infl)


def read_maya_skin(path: str, num_vertices: int) -> SkinFile:
    """The Maya skinning export (any skin extension other than .skin and
    .pskin, skinnedcharacter.cpp:272-362): a ``<label> B`` line, B lines ``<joint
    index> <bone name>`` (each bone binds to the skeleton joint of that index),
    a ``:``-separated header whose fields after the first name the bone of
    each weight column, then ``num_vertices`` ``:``-separated lines whose
    fields after the first are the weights; weights > 0 are kept.  A column
    naming no bone is an error here (the reference reports it and then
    misaligns the later columns)."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    lines = [ln for ln in lines if ln]
    t = lines[0].split()
    if len(t) != 2:
        raise ValueError(f"{path}: expected the number of bones")
    nb = int(t[1])
    bones, ids = [], []
    for ln in lines[1:1 + nb]:
        t = ln.split()
        if len(t) != 2:
            raise ValueError(f"{path}: expected '<joint index> <bone name>', got {t}")
        ids.append(int(t[0]))
        bones.append(t[1])
    cols = [c.strip() for c in lines[1 + nb].split(":")][1:]
    for c in cols:
        if c not in bones:
            raise ValueError(f"{path}: weight column '{c}' names no bone")
    col_bone = [bones.index(c) for c in cols]
    rows = lines[2 + nb:2 + nb + num_vertices]
    if len(rows) != num_vertices:
        raise ValueError(f"{path}: {len(rows)} weight lines for {num_vertices} vertices")
    infl = []
    for ln in rows:
        w = [float(x) for x in ln.split(":")[1:]]
        if len(w) > len(cols):
            raise ValueError(f"{path}: {len(w)} weights for {len(cols)} columns")
        infl.append([(col_bone[j], x) for j, x in enumerate(w) if x > 0.0])
    return SkinFile(bones, infl, ids)


def write_skin(path: str, bones: list[str], idx: np.ndarray, w: np.ndarray) -> None:
    with open(path, "w") as fh:
        fh.write("Skeletool character skinning file V1.0\nbones:\n")
        fh.write(" ".join(bones) + "\nvertex weights:\n")
        for v in range(len(idx)):
            pairs = " ".join(f"{int(i)} {repr(float(x))}" for i, x in zip(idx[v], w[v]) if x != 0)
            fh.write(f"{v} {pairs}\n")


def read_character(path: str) -> dict:
    """Paths of a "skeletool character file" (v1.0, v0.3, v0.1;
    skinnedcharacter.cpp:425-505): ``skeleton`` (and, before v1.0, a .dof
    line after it), ``mesh``, ``skin`` and optional ``pose`` sections, each
    followed by a path relative to the character file."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh]
    lines = [ln for ln in lines if ln]
    versions = {"skeletool character file v1.0": "v1.0", "skeletool character file v0.3": "v0.3",
                "skeletool character file v0.1": "v0.1"}
    if lines[0] not in versions:
        raise ValueError(f"{path}: not a skeletool character file")
    version = versions[lines[0]]
    base = os.path.dirname(path)
    out = {"version": version}
    i = 1
    while i < len(lines):
        key = lines[i].split()[0]
        if key == "skeleton":
            out["skeleton"] = os.path.join(base, lines[i + 1].split()[0])
            i += 2
            if version in ("v0.1", "v0.3"):
                out["dof"] = os.path.join(base, lines[i].split()[0])
                i += 1
        elif key in ("mesh", "skin", "pose"):
            out[key] = os.path.join(base, lines[i + 1].split()[0])
            i += 2
        else:
            i += 1
    if "mesh" not in out or "skin" not in out:
        raise ValueError(f"{path}: no mesh or skinning definition")
    return out


def load_character(path: str, K: int = 4, ed_graph_mesh: str | None = None):
    """A character from its files: (scenes.Character, skeleton description).
    Skeletons are V1.0, or v0.9 / v0.10 completed by the character's .dof file
    (skinnedcharacter.cpp:515-518); skinning is chosen by extension as the
    reference does (:542-547): .skin, .pskin, else the Maya export; the mesh
    is OFF or OBJ.  A ``pose`` file's frame 0 is the pose the mesh is bound
    in.
    Skinning joints are the skin file's bones in file order.  With
    ``ed_graph_mesh`` (the decimated graph mesh) the embedded graph is built
    as the reference builds it."""
    paths = read_character(path)
    if "skeleton" not in paths:
        raise ValueError(f"{path}: no skeleton")
    skel = read_skeleton(paths["skeleton"])
    if skel.dof_file_required:
        if "dof" not in paths:
            raise ValueError(f"{path}: the skeleton needs a .dof file, which only v0.1 / v0.3 characters name")
        skel = read_dof(paths["dof"], skel)
    mesh = read_obj(paths["mesh"]) if paths["mesh"].lower().endswith(".obj") else read_off(paths["mesh"])
    ext = os.path.splitext(paths["skin"])[1]
    if ext == ".skin":
        skin = read_skin(paths["skin"])
    elif ext == ".pskin":
        skin = read_pskin(paths["skin"])
    else:
        skin = read_maya_skin(paths["skin"], len(mesh.verts))
    if len(skin.influences) != len(mesh.verts):
        raise ValueError(f"{path}: {len(skin.influences)} skinned vertices for {len(mesh.verts)} mesh vertices")
    bind = None
    if "pose" in paths:
        frames, poses = read_motion(paths["pose"], skel)
        if not (frames == 0).any():
            raise ValueError(f"{paths['pose']}: no frame 0 for the initial pose")
        bind = poses[int(np.flatnonzero(frames == 0)[-1])]
    desc = skeleton_desc(skel, skin.bones, skin.joint_ids, bind)
    ji, jw = skin.top_k(K)
    verts = mesh.verts.astype(np.float64)
    colors = mesh.colors if mesh.colors is not None else np.full((len(verts), 3), 0.5)
    ch = scenes.Character(verts, mesh.faces, scenes.vertex_normals(verts, mesh.faces),
                          desc["joint_positions"][desc["skin_joint"]], ji, jw, colors)
    if ed_graph_mesh is not None:
        g = read_obj(ed_graph_mesh) if ed_graph_mesh.lower().endswith(".obj") else read_off(ed_graph_mesh)
        built = build_embedded_graph(mesh.verts, mesh.faces, g.verts, g.faces)
        ch.graph = built.graph
        ch.meta["ed_build"] = built
    return ch, desc


# ----------------------------------------------------------------------------
# Embedded graph construction
# ----------------------------------------------------------------------------

_FAR = 1000000  # the reference's "unreached" hop distance


@dataclass
class EDGraphBuild:
    graph: scenes.EmbeddedGraph          # node -> base vertex, vertex -> node CSR and weights
    radius: np.ndarray                   # [M] hop radius per node
    node_neighbours: list[list[int]]     # node -> node (the graph mesh's edges), ascending
    n2v_offsets: np.ndarray              # [M + 1] node -> vertex CSR (vertices ascending)
    n2v_vertices: np.ndarray
    n2v_weights: np.ndarray
    rigidity: np.ndarray                 # [M] mean segmentation weight of the node's vertices
    unconnected: int = 0                 # vertices outside every radius (tied to the nearest node)
    meta: dict = field(default_factory=dict)


def _adjacency(n: int, faces: np.ndarray):
    """Vertex adjacency from the faces' edges (trimesh::computeNeighbors), as
    a CSR matrix and the sorted neighbour lists."""
    from scipy.sparse import csr_matrix
    f = np.asarray(faces, np.int64)
    a = np.concatenate([f[:, [0, 1]], f[:, [0, 2]], f[:, [1, 0]], f[:, [1, 2]], f[:, [2, 0]], f[:, [2, 1]]])
    a = np.unique(a, axis=0)
    m = csr_matrix((np.ones(len(a), np.int8), (a[:, 0], a[:, 1])), shape=(n, n))
    m.sort_indices()
    return m


def _hops(adj, sources) -> np.ndarray:
    """Hop distances from each source over the adjacency (trimesh::
    computeGeodesicDistance: breadth-first, unreached = 1000000)."""
    from scipy.sparse.csgraph import shortest_path
    d = shortest_path(adj, method="D", unweighted=True, directed=False, indices=np.atleast_1d(sources))
    d[~np.isfinite(d)] = _FAR
    return d.astype(np.int64)


def _extreme_y(verts):
    y = np.asarray(verts, np.float32)[:, 1]
    hi = lo = -1
    ymax, ymin = np.float32(-10000.0), np.float32(100000.0)
    for i, v in enumerate(y):  # strict comparisons: the first vertex reaching the extreme
        if v > ymax:
            ymax, hi = v, i
        if v < ymin:
            ymin, lo = v, i
    return hi, lo


def build_embedded_graph(base_verts, base_faces, graph_verts, graph_faces, seg_weights=None,
                         refinement: bool = False) -> EDGraphBuild:
    """The reference's embedded graph (EmbeddedGraph.cpp:34-460) from the
    character mesh and its decimated graph mesh:

    1. hop distances from the highest and lowest (y) vertex of each mesh mark
       which of those two each vertex is connected to;
    2. each graph-mesh vertex, in order, becomes a node at the nearest
       (float32 Euclidean, first on ties) base vertex that is not taken yet
       and has the same connectivity marks; its graph-mesh neighbours are its
       node neighbours;
    3. node radius = half (integer) the largest hop distance on the base mesh
       to its neighbour nodes' vertices, at least 3;
    4. vertex v joins every node i with d = hops(i, v) / radius_i <= 1 at
       weight exp(-d^2 / 2) (nodes ascending); a vertex no node reaches joins
       the node with the smallest hop distance (the last one on ties) at
       weight 1; weights are normalised by their float sum;
    5. rigidity of a node = mean segmentation weight of its vertices (with
       ``refinement``, weights below 20 count a tenth).
    """
    bv = np.asarray(base_verts, np.float32)
    gv = np.asarray(graph_verts, np.float32)
    N, M = len(bv), len(gv)
    adj_b = _adjacency(N, base_faces)
    adj_g = _adjacency(M, graph_faces)
    hb, lb = _extreme_y(bv)
    hg, lg = _extreme_y(gv)
    db = _hops(adj_b, [hb, lb])
    dg = _hops(adj_g, [hg, lg])
    mark_b = np.stack([db[0] != _FAR, db[1] != _FAR], 1)
    mark_g = np.stack([dg[0] != _FAR, dg[1] != _FAR], 1)
    used = np.zeros(N, bool)
    node_vtx = np.zeros(M, np.int32)
    for vl in range(M):
        diff = gv[vl][None, :] - bv
        dist = np.sqrt((diff[:, 0] * diff[:, 0] + diff[:, 1] * diff[:, 1]) + diff[:, 2] * diff[:, 2])
        ok = ~used & (mark_b[:, 0] == mark_g[vl, 0]) & (mark_b[:, 1] == mark_g[vl, 1])
        if not ok.any():
            raise ValueError(f"graph vertex {vl} has no base vertex left to attach to")
        cand = np.where(ok, dist, np.float32(np.inf))
        j = int(np.argmin(cand))  # first minimum: the reference's strict '<' scan
        used[j] = True
        node_vtx[vl] = j
    neigh = [list(adj_g.indices[adj_g.indptr[i]:adj_g.indptr[i + 1]]) for i in range(M)]
    if any(len(n) == 0 for n in neigh):
        raise ValueError("every graph-mesh vertex needs at least one edge")
    dist = _hops(adj_b, node_vtx)                       # [M, N]
    radius = np.zeros(M, np.int64)
    for n in range(M):
        r = max(int(dist[n, node_vtx[j]]) for j in neigh[n])
        radius[n] = max(r // 2, 3)
    d = dist.astype(np.float32) / radius[:, None].astype(np.float32)   # [M, N] float32 like the reference
    inside = d <= np.float32(1.0)
    w = np.where(inside, np.exp(-0.5 * d.astype(np.float64) * d.astype(np.float64)).astype(np.float32),
                 np.float32(0.0))
    v2n_off = [0]
    v2n_idx, v2n_w = [], []
    unconnected = 0
    for v in range(N):
        nodes = np.nonzero(inside[:, v])[0]
        if len(nodes):
            ws = w[nodes, v]
        else:
            unconnected += 1
            col = dist[:, v]
            nodes = np.array([int(np.flatnonzero(col == col.min())[-1])])  # '<=': the last minimum
            ws = np.array([1.0], np.float32)
        tot = np.float32(0.0)
        for x in ws:  # the reference's sequential float sum
            tot = np.float32(tot + x)
        v2n_idx.extend(int(n) for n in nodes)
        v2n_w.extend(np.float32(x) / tot for x in ws)
        v2n_off.append(len(v2n_idx))
    v2n_off = np.asarray(v2n_off, np.int32)
    v2n_idx = np.asarray(v2n_idx, np.int32)
    v2n_w = np.asarray(v2n_w, np.float32)
    # node -> vertex CSR in vertex order (EmbeddedGraph.cpp:364-396)
    order = np.argsort(v2n_idx, kind="stable")
    vert_of_entry = np.repeat(np.arange(N, dtype=np.int32), np.diff(v2n_off))
    n2v_vtx = vert_of_entry[order]
    n2v_w = v2n_w[order]
    n2v_off = np.concatenate([[0], np.cumsum(np.bincount(v2n_idx, minlength=M))]).astype(np.int32)
    sw = np.ones(N, np.float32) if seg_weights is None else np.asarray(seg_weights, np.float32)
    rig = np.zeros(M, np.float32)
    for n in range(M):
        vs = n2v_vtx[n2v_off[n]:n2v_off[n + 1]]
        acc = np.float32(0.0)
        for v in vs:
            s = sw[v]
            acc = np.float32(acc + (s / np.float32(10.0) if refinement and s < 20.0 else s))
        rig[n] = acc / np.float32(len(vs)) if len(vs) else np.float32(0.0)
    graph = scenes.EmbeddedGraph(node_vtx, v2n_off, v2n_idx, v2n_w.astype(np.float64))
    return EDGraphBuild(graph, radius.astype(np.int32), neigh, n2v_off, n2v_vtx, n2v_w, rig, unconnected)
