"""The five synthetic workloads the throughput and parity figures are quoted
on, built from the procedural scenes (metres, seeded):

* plain_forward_scene(): icosphere subdivision 5 (10,242 vertices, 20,480
  triangles) scaled to 1.0 x 0.4 x 0.3 m, a 24-joint chain, joint rotations
  |theta| <= 0.5 rad and N(0, 0.02 m) translations, top-4 softmax skin
  weights, LBS, four 512^2 pinhole cameras (f = 1000 px) on a 3 m ring;
  tri-id, barycentric and vertex-colour buffers; seed 0;
* single_character_work(): a 25k-vertex body posed by the 57-DOF skeleton
  (forward kinematics to joint DQs), DQ skinning top-4, a 500-node embedded
  graph, four 1024^2 views, colour + silhouette losses;
* batched_works(): 16 bodies with vertices uniform in [8k, 60k], 500-1500
  graph nodes, DQ skinning top-4, 4 views each at 1024^2, 256^2 DT images,
  seed 2;
* temporal_works(): one 50k-vertex body over 64 frames of a random walk of
  the joint rotations (axis-angle steps N(0, 0.02)), 8 views per frame at
  1285x940 with focal lengths in [900, 1200] px;
* stress_works(): 64 bodies x 4 views, vertices log-uniform in [5k, 100k],
  resolutions from {512^2, 1024^2, 2048^2}, every fifth body with limbs bent
  up to 1.5 rad.

Targets (colour images and DT masks from a perturbed pose) are rendered by
`render_fn`, by default the GPU pose renderer.
"""
from __future__ import annotations

import numpy as np

from . import pipeline, scenes


def fast_graph(ch: scenes.Character, M: int, rng) -> scenes.EmbeddedGraph:
    """An embedded graph on a random vertex subset (4 nearest nodes, Gaussian
    weights, sigma 0.05 m): node placement does not matter for a throughput
    workload, and farthest-point sampling of many large meshes would dominate
    the set-up time."""
    from scipy.spatial import cKDTree
    nodes = np.sort(rng.choice(ch.N, M, replace=False)).astype(np.int32)
    d, idx = cKDTree(ch.verts[nodes]).query(ch.verts, k=4)
    w = np.exp(-d ** 2 / (2 * 0.05 ** 2))
    bad = w.sum(1) <= 1e-12
    w[bad] = 0.0
    w[bad, 0] = 1.0
    w /= w.sum(1, keepdims=True)
    return scenes.EmbeddedGraph(nodes, np.arange(0, 4 * ch.N + 1, 4, dtype=np.int32),
                                idx.reshape(-1).astype(np.int32), w.reshape(-1))


def _centred_ring(C, f, W, H, phase=0.0, rng=None, f_range=None):
    """C cameras on a 3 m ring looking at (0, 0.9, 0), the body's centre."""
    kw = {} if f_range is None else dict(f_range=f_range, rng=rng)
    ext, intr = scenes.ring_cameras(C, 3.0, f, W, H, phase=phase, **kw)
    E = ext.reshape(C, 3, 4).copy()
    E[:, :, 3] -= E[:, :, :3] @ np.array([0.0, 0.9, 0.0])
    return E.reshape(C, 12), intr


def plain_forward_scene():
    """(character, LBS transforms [1,J,12], skin weights [1,N,4], zero
    displacements [1,N,3], extrinsics [4,12], intrinsics [4,9], W, H)."""
    rng = np.random.default_rng(0)
    ch = scenes.ellipsoid_character(5, scale=(1.0, 0.4, 0.3), n_joints=24, K=4, seed=0)
    dq = scenes.random_joint_dqs(rng, ch.joint_centres, 1, 0.5, 0.02)[0]
    qr, qd = dq[:, :4], dq[:, 4:]
    R = scenes.quat_to_mat(qr)
    t = 2 * scenes.quat_mul(qd, qr * np.array([-1, -1, -1, 1]))[:, :3]
    T = np.concatenate([R.reshape(-1, 9), t], 1)[None].astype(np.float32)
    w = ch.joint_w[None].astype(np.float32)
    disp = np.zeros((1, ch.N, 3), np.float32)
    ext, intr = scenes.ring_cameras(4, 3.0, 1000.0, 512, 512)
    return ch, T, w, disp, ext, intr, 512, 512


def single_character_work(render_fn=None, device="cuda"):
    """One CharacterWork; its joint DQs come from 57 random DOFs through the
    GPU forward kinematics."""
    import torch

    from . import skinning
    rng = np.random.default_rng(1)
    ch = scenes.capsule_body(25000, seed=1, ed_nodes=500)
    skel = scenes.body_skeleton()
    dofs = scenes.random_dofs(rng, skel, 1, 0.3, 0.02)
    dq, _, _ = skinning.skeleton_fk(skinning.Skeleton(skel),
                                    torch.as_tensor(dofs, dtype=torch.float32, device=device))
    _, dT, dA = pipeline.pose_params(rng, ch)
    E, intr = _centred_ring(4, 2000.0, 1024, 1024, phase=0.4)
    wk = pipeline.CharacterWork(ch, dq[0].cpu().numpy(), dT, dA, E, intr, 1024, 1024, 256)
    pipeline.make_targets([wk], render_fn or pipeline.gpu_pose_renderer(device), np.random.default_rng(2))
    return wk, skel, dofs


def batched_works(render_fn=None, n_chars=16):
    works = pipeline.build_works(n_chars, seed=2)
    pipeline.make_targets(works, render_fn or pipeline.gpu_pose_renderer(), np.random.default_rng(2))
    return works


def temporal_works(render_fn=None, W=1285, H=940, C=8, T=64):
    rng = np.random.default_rng(33)
    ch = scenes.capsule_body(50000, seed=3300, ed_nodes=0)
    ch.graph = fast_graph(ch, 1000, rng)
    E, intr = _centred_ring(C, 1000.0, W, H, phase=0.3, rng=rng, f_range=(900.0, 1200.0))
    J = ch.J
    axis = scenes.random_unit(rng, (J,))
    ang = rng.uniform(0, 0.2, J)
    _, dT, dA = pipeline.pose_params(rng, ch)
    ws = []
    for _ in range(T):
        # random walk: compose the previous rotation with a small step
        step = rng.normal(0, 0.02, (J, 3))
        rv = axis * ang[:, None] + step
        ang = np.linalg.norm(rv, axis=1)
        axis = rv / np.maximum(ang, 1e-12)[:, None]
        q = scenes.axis_angle_to_quat(axis, ang)
        # rotate each joint about its own centre: t = c - R c
        R = scenes.quat_to_mat(q)
        t = ch.joint_centres - np.einsum("jab,jb->ja", R, ch.joint_centres)
        ws.append(pipeline.CharacterWork(ch, scenes.dq_from_rt(q, t), dT, dA, E, intr, W, H, 256))
    pipeline.make_targets(ws, render_fn or pipeline.gpu_pose_renderer(), np.random.default_rng(6))
    return ws


def stress_works(render_fn=None, n_chars=64):
    rng = np.random.default_rng(44)
    nv = np.exp(rng.uniform(np.log(5000), np.log(100000), n_chars)).astype(int)
    res = rng.choice([512, 1024, 2048], n_chars)
    ws = []
    for b in range(n_chars):
        ch = scenes.capsule_body(int(nv[b]), seed=900 + b, ed_nodes=0)
        ch.graph = fast_graph(ch, int(rng.uniform(500, 1500)), rng)
        dq, dT, dA = pipeline.pose_params(rng, ch, max_angle=1.5 if b % 5 == 0 else 0.3)
        W = H = int(res[b])
        # the body spans about a third of the image at 3 m
        ext, intr = scenes.ring_cameras(4, 3.0, 2.0 * W, W, H, phase=rng.uniform(0, 2 * np.pi))
        E = ext.reshape(4, 3, 4).copy()
        E[:, :, 3] -= E[:, :, :3] @ np.array([0.0, 0.9, 0.0])
        ws.append(pipeline.CharacterWork(ch, dq, dT, dA, E.reshape(4, 12), intr, W, H, 256))
    pipeline.make_targets(ws, render_fn or pipeline.gpu_pose_renderer(), np.random.default_rng(5))
    return ws
