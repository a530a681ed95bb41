"""Seeded synthetic characters and camera rigs for tests, smoke and bench.

Everything is numpy on the host (float64 geometry, cast to float32 at upload).
Units are metres.  The generators follow the workload descriptions:

* ``icosphere``: subdivided icosahedron (subdiv 5 -> 10,242 verts / 20,480 tris)
* ``chain_skeleton``: joints evenly spaced along an axis
* ``capsule_body``: a body made of capped cylinders (torso, head, limbs)
* ``skin_weights``: softmax(-dist/0.05) truncated to the top-K joints
* ``random_joint_dqs``: per-joint rotation about the joint centre (random axis,
  |theta| <= max_angle) plus a N(0, sigma) translation, as dual quaternions
* ``embedded_graph``: farthest-point-sampled nodes, K nearest nodes per vertex
  with Gaussian weights
* ``ring_cameras``: pinhole cameras on a horizontal ring looking at the origin
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


# ----------------------------------------------------------------------------
# Meshes
# ----------------------------------------------------------------------------

def icosphere(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = f
    for _ in range(subdiv):
        cache: dict[tuple[int, int], int] = {}

        def mid(a: int, b: int) -> int:
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, dtype=np.int32)


def capped_cylinder(p0, p1, radius: float, n_around: int, n_along: int,
                    n_cap: int = 2) -> tuple[np.ndarray, np.ndarray]:
    """Closed tube from p0 to p1 with hemispherical-ish caps (rings shrinking)."""
    p0 = np.asarray(p0, float)
    p1 = np.asarray(p1, float)
    axis = p1 - p0
    L = np.linalg.norm(axis)
    a = axis / L
    tmp = np.array([1.0, 0, 0]) if abs(a[0]) < 0.9 else np.array([0, 1.0, 0])
    u = np.cross(a, tmp)
    u /= np.linalg.norm(u)
    w = np.cross(a, u)
    rings = []
    # caps: rings at shrinking radius beyond the ends
    for i in range(n_cap, 0, -1):
        ang = (np.pi / 2) * i / (n_cap + 1)
        rings.append((-radius * np.sin(ang), radius * np.cos(ang)))
    for i in range(n_along + 1):
        rings.append((L * i / n_along, radius))
    for i in range(1, n_cap + 1):
        ang = (np.pi / 2) * i / (n_cap + 1)
        rings.append((L + radius * np.sin(ang), radius * np.cos(ang)))
    th = 2 * np.pi * np.arange(n_around) / n_around
    verts = []
    for s, r in rings:
        for t_ in th:
            verts.append(p0 + a * s + r * (np.cos(t_) * u + np.sin(t_) * w))
    nr = len(rings)
    verts.append(p0 - a * radius)  # bottom pole
    verts.append(p1 + a * radius)  # top pole
    bot, top = len(verts) - 2, len(verts) - 1
    faces = []
    for r_ in range(nr - 1):
        for k in range(n_around):
            a0 = r_ * n_around + k
            a1 = r_ * n_around + (k + 1) % n_around
            b0 = a0 + n_around
            b1 = a1 + n_around
            faces += [(a0, a1, b1), (a0, b1, b0)]
    for k in range(n_around):
        faces.append((bot, (k + 1) % n_around, k))
        o = (nr - 1) * n_around
        faces.append((top, o + k, o + (k + 1) % n_around))
    return np.array(verts), np.array(faces, dtype=np.int32)


def vertex_normals(verts: np.ndarray, faces: np.ndarray) -> np.ndarray:
    fn = np.cross(verts[faces[:, 1]] - verts[faces[:, 0]], verts[faces[:, 2]] - verts[faces[:, 0]])
    vn = np.zeros_like(verts)
    for i in range(3):
        np.add.at(vn, faces[:, i], fn)
    n = np.linalg.norm(vn, axis=1, keepdims=True)
    n[n == 0] = 1.0
    return vn / n


# ----------------------------------------------------------------------------
# Skeleton / skinning
# ----------------------------------------------------------------------------

def chain_skeleton(n_joints: int, axis: int = 0, lo: float = -0.5, hi: float = 0.5) -> np.ndarray:
    c = np.zeros((n_joints, 3))
    c[:, axis] = np.linspace(lo, hi, n_joints)
    return c


def skin_weights(verts: np.ndarray, centres: np.ndarray, K: int = 4, tau: float = 0.05,
                 allowed: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
    """softmax(-|v - c_j| / tau) over joints, truncated to top-K, renormalised.

    ``allowed`` (optional bool [N,J]) restricts which joints may skin a vertex.
    Returns joint indices [N,K] int32 (descending weight) and weights [N,K].
    """
    d = np.linalg.norm(verts[:, None, :] - centres[None, :, :], axis=2)
    logits = -d / tau
    if allowed is not None:
        logits = np.where(allowed, logits, -np.inf)
    K = min(K, centres.shape[0])
    idx = np.argsort(-logits, axis=1, kind="stable")[:, :K]
    lg = np.take_along_axis(logits, idx, axis=1)
    lg = lg - lg[:, :1]
    w = np.exp(lg)
    w /= w.sum(axis=1, keepdims=True)
    return idx.astype(np.int32), w


def axis_angle_to_quat(axis: np.ndarray, angle: np.ndarray) -> np.ndarray:
    """(x, y, z, w) quaternions; axis [...,3] unit, angle [...]."""
    h = 0.5 * angle
    return np.concatenate([axis * np.sin(h)[..., None], np.cos(h)[..., None]], axis=-1)


def quat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    av, aw = a[..., :3], a[..., 3:]
    bv, bw = b[..., :3], b[..., 3:]
    v = aw * bv + bw * av + np.cross(av, bv)
    w = aw * bw - np.sum(av * bv, axis=-1, keepdims=True)
    return np.concatenate([v, w], axis=-1)


def quat_to_mat(q: np.ndarray) -> np.ndarray:
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.stack([
        1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * z * w, 2 * x * z + 2 * y * w,
        2 * x * y + 2 * z * w, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * x * w,
        2 * x * z - 2 * y * w, 2 * y * z + 2 * x * w, 1 - 2 * x * x - 2 * y * y], axis=-1)
    return R.reshape(q.shape[:-1] + (3, 3))


def dq_from_rt(q_r: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Dual quaternion [..., 8] of the transform x -> R(q_r) x + t."""
    tq = np.concatenate([t, np.zeros(t.shape[:-1] + (1,))], axis=-1)
    q_d = 0.5 * quat_mul(tq, q_r)
    return np.concatenate([q_r, q_d], axis=-1)


def random_unit(rng: np.random.Generator, shape) -> np.ndarray:
    a = rng.normal(size=tuple(shape) + (3,))
    return a / np.linalg.norm(a, axis=-1, keepdims=True)


def random_joint_dqs(rng: np.random.Generator, centres: np.ndarray, B: int,
                     max_angle: float = 0.5, trans_sigma: float = 0.02) -> np.ndarray:
    """[B,J,8] DQs: rotation about each joint centre + N(0, sigma) translation."""
    J = centres.shape[0]
    axis = random_unit(rng, (B, J))
    ang = rng.uniform(0.0, max_angle, size=(B, J))
    q = axis_angle_to_quat(axis, ang)
    R = quat_to_mat(q)
    t = rng.normal(0.0, trans_sigma, size=(B, J, 3))
    c = centres[None]
    trans = c - np.einsum("bjik,bjk->bji", R, np.broadcast_to(c, (B, J, 3))) + t
    return dq_from_rt(q, trans)


def euler_from_axis_angle(axis: np.ndarray, angle: np.ndarray) -> np.ndarray:
    """Euler (a, b, g) with R = Rz(g) Ry(b) Rx(a) for the given rotation."""
    R = quat_to_mat(axis_angle_to_quat(axis, angle))
    sy = np.sqrt(R[..., 0, 0] ** 2 + R[..., 1, 0] ** 2)
    a = np.arctan2(R[..., 2, 1], R[..., 2, 2])
    b = np.arctan2(-R[..., 2, 0], sy)
    g = np.arctan2(R[..., 1, 0], R[..., 0, 0])
    return np.stack([a, b, g], axis=-1)


# ----------------------------------------------------------------------------
# Embedded graph
# ----------------------------------------------------------------------------

def farthest_point_sample(verts: np.ndarray, M: int, seed_index: int = 0) -> np.ndarray:
    N = verts.shape[0]
    M = min(M, N)
    sel = np.empty(M, dtype=np.int64)
    sel[0] = seed_index
    x = np.ascontiguousarray(verts, dtype=np.float64)
    d = np.einsum("ij,ij->i", x - x[seed_index], x - x[seed_index])
    diff = np.empty_like(x)
    dn = np.empty(N)
    for i in range(1, M):
        sel[i] = int(np.argmax(d))
        np.subtract(x, x[sel[i]], out=diff)
        np.einsum("ij,ij->i", diff, diff, out=dn)
        np.minimum(d, dn, out=d)
    return sel


@dataclass
class EmbeddedGraph:
    node_vertex: np.ndarray      # [M] int32 base-mesh vertex of each node
    v2n_offsets: np.ndarray      # [N+1] int32 CSR vertex -> nodes
    v2n_nodes: np.ndarray        # [E] int32
    v2n_weights: np.ndarray      # [E] float

    @property
    def num_nodes(self) -> int:
        return int(self.node_vertex.shape[0])

    @property
    def num_vertices(self) -> int:
        return int(self.v2n_offsets.shape[0] - 1)


def embedded_graph(verts: np.ndarray, M: int, K: int = 4, sigma: float = 0.05) -> EmbeddedGraph:
    nodes = farthest_point_sample(verts, M)
    gpos = verts[nodes]
    N = verts.shape[0]
    K = min(K, len(nodes))
    from scipy.spatial import cKDTree
    dist, idx = cKDTree(gpos).query(verts, k=K)
    dist, idx = dist.reshape(N, K), idx.reshape(N, K)
    w = np.exp(-dist ** 2 / (2 * sigma ** 2))
    tot = w.sum(axis=1, keepdims=True)
    bad = tot[:, 0] <= 1e-12
    w[bad] = 0.0
    w[bad, 0] = 1.0
    w /= w.sum(axis=1, keepdims=True)
    offsets = np.arange(0, N * K + 1, K, dtype=np.int32)
    return EmbeddedGraph(nodes.astype(np.int32), offsets, idx.reshape(-1).astype(np.int32),
                         w.reshape(-1))


# ----------------------------------------------------------------------------
# Cameras
# ----------------------------------------------------------------------------

def node_neighbours(verts: np.ndarray, node_vertex: np.ndarray, k: int = 6):
    """Symmetric node -> node adjacency of an embedded graph: each node's k
    nearest nodes (by base position), symmetrised.  Returns (offsets [M],
    sizes [M], indices [E]) int32, neighbours in increasing id order."""
    from scipy.spatial import cKDTree
    P = np.asarray(verts, np.float64)[node_vertex]
    M = len(node_vertex)
    _, nn = cKDTree(P).query(P, k=min(k + 1, M))
    adj = [set() for _ in range(M)]
    for i in range(M):
        for j in np.atleast_1d(nn[i])[1:]:
            adj[i].add(int(j))
            adj[int(j)].add(i)
    sizes = np.array([len(a) for a in adj], np.int32)
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    indices = np.concatenate([sorted(a) for a in adj]).astype(np.int32)
    return offsets, sizes, indices


def look_at_extrinsic(centre: np.ndarray, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """3x4 world->camera [R|t]; camera x right, y down, z forward."""
    c = np.asarray(centre, float)
    z = np.asarray(target, float) - c
    z /= np.linalg.norm(z)
    x = np.cross(z, np.asarray(up, float))
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    return np.concatenate([R, (-R @ c)[:, None]], axis=1)


def ring_cameras(C: int, radius: float, f: float, W: int, H: int, height: float = 0.0,
                 f_range: tuple[float, float] | None = None, rng=None,
                 phase: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
    """Extrinsics [C,12] and intrinsics [C,9] (row-major), as the reference lays them out."""
    ext, intr = [], []
    for i in range(C):
        phi = phase + 2 * np.pi * i / C
        centre = np.array([radius * np.sin(phi), height, radius * np.cos(phi)])
        ext.append(look_at_extrinsic(centre).reshape(-1))
        fi = f if f_range is None else rng.uniform(*f_range)
        intr.append(np.array([fi, 0.0, W / 2.0, 0.0, fi, H / 2.0, 0.0, 0.0, 1.0]))
    return np.array(ext), np.array(intr)


# ----------------------------------------------------------------------------
# Character bundle
# ----------------------------------------------------------------------------

@dataclass
class Character:
    verts: np.ndarray            # [N,3] rest positions
    faces: np.ndarray            # [F,3] int32
    normals: np.ndarray          # [N,3] rest normals
    joint_centres: np.ndarray    # [J,3]
    joint_idx: np.ndarray        # [N,K] int32
    joint_w: np.ndarray          # [N,K]
    colors: np.ndarray           # [N,3] in [0,1]
    graph: EmbeddedGraph | None = None
    texcoords: np.ndarray | None = None  # [F,3,2]
    meta: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.verts.shape[0])

    @property
    def F(self) -> int:
        return int(self.faces.shape[0])

    @property
    def J(self) -> int:
        return int(self.joint_centres.shape[0])


def uv_atlas(faces: np.ndarray) -> np.ndarray:
    """Uniform-grid atlas: every face gets its own cell of a square grid."""
    F = faces.shape[0]
    g = int(np.ceil(np.sqrt(F)))
    k = np.arange(F)
    cu, cv = (k % g) / g, (k // g) / g
    s = 1.0 / g
    tc = np.stack([np.stack([cu + 0.1 * s, cv + 0.1 * s], -1),
                   np.stack([cu + 0.9 * s, cv + 0.1 * s], -1),
                   np.stack([cu + 0.1 * s, cv + 0.9 * s], -1)], axis=1)
    return tc


def ellipsoid_character(subdiv: int, scale=(1.0, 0.4, 0.3), n_joints: int = 24, K: int = 4,
                        seed: int = 0, ed_nodes: int = 0) -> Character:
    """Icosphere scaled to an ellipsoid, skinned to a chain along x."""
    rng = np.random.default_rng(seed)
    v, f = icosphere(subdiv)
    v = v * np.asarray(scale)[None]
    centres = chain_skeleton(n_joints, 0, -0.5 * scale[0], 0.5 * scale[0])
    ji, jw = skin_weights(v, centres, K)
    col = rng.uniform(0.0, 1.0, size=v.shape)
    ch = Character(v, f, vertex_normals(v, f), centres, ji, jw, col)
    if ed_nodes:
        ch.graph = embedded_graph(v, ed_nodes)
    ch.texcoords = uv_atlas(f)
    return ch


_BODY_PARTS = [
    # name, p0, p1, radius, joint chain (indices into joint list)
    ("torso", (0.0, 0.95, 0.0), (0.0, 1.45, 0.0), 0.16, [0, 1, 2]),
    ("head", (0.0, 1.55, 0.0), (0.0, 1.72, 0.0), 0.10, [3]),
    ("l_arm", (0.20, 1.40, 0.0), (0.75, 1.40, 0.0), 0.05, [4, 5, 6]),
    ("r_arm", (-0.20, 1.40, 0.0), (-0.75, 1.40, 0.0), 0.05, [7, 8, 9]),
    ("l_leg", (0.10, 0.90, 0.0), (0.10, 0.08, 0.0), 0.07, [10, 11, 12]),
    ("r_leg", (-0.10, 0.90, 0.0), (-0.10, 0.08, 0.0), 0.07, [13, 14, 15]),
    ("l_foot", (0.10, 0.04, 0.0), (0.10, 0.04, 0.18), 0.04, [16]),
    ("r_foot", (-0.10, 0.04, 0.0), (-0.10, 0.04, 0.18), 0.04, [17]),
]

_BODY_JOINTS = np.array([
    (0.0, 0.95, 0.0), (0.0, 1.15, 0.0), (0.0, 1.35, 0.0), (0.0, 1.55, 0.0),
    (0.20, 1.40, 0.0), (0.45, 1.40, 0.0), (0.72, 1.40, 0.0),
    (-0.20, 1.40, 0.0), (-0.45, 1.40, 0.0), (-0.72, 1.40, 0.0),
    (0.10, 0.90, 0.0), (0.10, 0.50, 0.0), (0.10, 0.10, 0.0),
    (-0.10, 0.90, 0.0), (-0.10, 0.50, 0.0), (-0.10, 0.10, 0.0),
    (0.10, 0.04, 0.08), (-0.10, 0.04, 0.08)])


def capsule_body(target_verts: int, K: int = 4, seed: int = 0, ed_nodes: int = 500,
                 ed_sigma: float = 0.05) -> Character:
    """Cylinder-union body (8 capped tubes, 18 joints) with ~target_verts vertices."""
    rng = np.random.default_rng(seed)
    lengths = [np.linalg.norm(np.subtract(p1, p0)) + 2 * r for _, p0, p1, r, _ in _BODY_PARTS]
    circ = [2 * np.pi * r for _, _, _, r, _ in _BODY_PARTS]
    areas = np.array(lengths) * np.array(circ)
    share = areas / areas.sum()
    vs, fs, allowed_rows = [], [], []
    off = 0
    J = _BODY_JOINTS.shape[0]
    for (name, p0, p1, r, chain), s in zip(_BODY_PARTS, share):
        nv = max(64, int(target_verts * s))
        L = np.linalg.norm(np.subtract(p1, p0)) + 2 * r
        ratio = L / (2 * np.pi * r)
        n_around = max(8, int(np.sqrt(nv / ratio)))
        n_along = max(2, nv // n_around - 5)
        v, f = capped_cylinder(p0, p1, r, n_around, n_along)
        vs.append(v)
        fs.append(f + off)
        off += v.shape[0]
        allow = np.zeros(J, bool)
        allow[chain] = True
        # neighbouring joints (parents) may also influence the part
        near = np.argsort(np.linalg.norm(_BODY_JOINTS - np.asarray(p0)[None], axis=1))[:2]
        allow[near] = True
        allowed_rows.append(np.broadcast_to(allow, (v.shape[0], J)))
    v = np.concatenate(vs)
    f = np.concatenate(fs).astype(np.int32)
    allowed = np.concatenate(allowed_rows)
    ji, jw = skin_weights(v, _BODY_JOINTS, K, tau=0.05, allowed=allowed)
    col = rng.uniform(0.0, 1.0, size=v.shape)
    ch = Character(v, f, vertex_normals(v, f), _BODY_JOINTS.copy(), ji, jw, col)
    if ed_nodes:
        ch.graph = embedded_graph(v, ed_nodes, 4, ed_sigma)
    ch.texcoords = uv_atlas(f)
    ch.meta["parts"] = [p[0] for p in _BODY_PARTS]
    return ch


# ----------------------------------------------------------------------------
# Kinematic skeleton (DOFs) for the capsule body
# ----------------------------------------------------------------------------

_BODY_PARENTS = [-1, 0, 1, 2, 2, 4, 5, 2, 7, 8, 0, 10, 11, 0, 13, 14, 12, 15]
JOINT_REVOLUTE, JOINT_PRISMATIC = 0, 1


def body_skeleton(centres: np.ndarray | None = None) -> dict:
    """57-DOF kinematic tree over the 18 capsule-body joints.

    Root: three prismatic (x, y, z) then three revolute (x, y, z) joints at the
    pelvis; every other body joint: a ball of three revolute joints (x, y, z)
    offset from its parent body joint.  One DOF per kinematic joint (57).
    Skinning joint i is the last kinematic joint of body joint i with bind
    pose T(-centre_i), so zero DOFs give identity transforms.
    Returns dict(parent, type, dof, offset, axis, skin_joint, skin_scale,
    skin_bind, num_dofs, names).
    """
    c = np.asarray(_BODY_JOINTS if centres is None else centres, np.float64)
    eye = np.eye(3)
    parent, typ, dof, off, axis, names = [], [], [], [], [], []
    last = {}

    def add(p, t, o, a, name):
        parent.append(p)
        typ.append(t)
        dof.append(len(dof))
        off.append(o)
        axis.append(a)
        names.append(name)
        return len(parent) - 1

    j = -1
    for k in range(3):
        j = add(j, JOINT_PRISMATIC, c[0] if k == 0 else np.zeros(3), eye[k], f"root_t{'xyz'[k]}")
    for k in range(3):
        j = add(j, JOINT_REVOLUTE, np.zeros(3), eye[k], f"root_r{'xyz'[k]}")
    last[0] = j
    for b in range(1, len(c)):
        p = last[_BODY_PARENTS[b]]
        j = p
        for k in range(3):
            j = add(j, JOINT_REVOLUTE, (c[b] - c[_BODY_PARENTS[b]]) if k == 0 else np.zeros(3), eye[k],
                    f"j{b}_r{'xyz'[k]}")
        last[b] = j
    bind = np.zeros((len(c), 12))
    for b in range(len(c)):
        bind[b] = [1, 0, 0, -c[b, 0], 0, 1, 0, -c[b, 1], 0, 0, 1, -c[b, 2]]
    return dict(parent=np.array(parent, np.int32), type=np.array(typ, np.int32), dof=np.array(dof, np.int32),
                offset=np.array(off), axis=np.array(axis), skin_joint=np.array([last[b] for b in range(len(c))],
                                                                              np.int32),
                skin_scale=np.ones(len(c)), skin_bind=bind, num_dofs=len(dof), names=names)


def random_dofs(rng: np.random.Generator, skel: dict, B: int, max_angle: float = 0.3,
                trans_sigma: float = 0.02) -> np.ndarray:
    """[B,D] DOFs: revolute angles U(-max_angle, max_angle), prismatic N(0, sigma)."""
    D = int(skel["num_dofs"])
    out = np.zeros((B, D))
    for j, (t, d) in enumerate(zip(skel["type"], skel["dof"])):
        if d < 0:
            continue
        out[:, d] = rng.uniform(-max_angle, max_angle, B) if t == JOINT_REVOLUTE else rng.normal(0, trans_sigma, B)
    return out
