"""TensorFlow gradient registrations for the op shim (dcx_tf_ops.cc).

Import after tf.load_op_library(".../libdcx_tf_ops.so"): `load(path)` does
both.  The wiring follows the reference's Python:

* CudaRendererGpu (CudaRendererGPU.py:194-237): CudaRendererGradGpu for the
  vertexColor and textured albedos with (grad of the render buffer, the seven
  inputs, outputs 0, 1 and 3); zeros otherwise; target image and cameras get
  zeros;
* EmbeddedGraphGpu (EmbeddedGraphGpu.py:82-94): EmbeddedGraphGpuGrad with
  (grad of the vertices, grad of the markers, outputs 4 and 5);
* CameraProjectionGpu (CameraProjectionGPU.py:60-74): CameraProjectionGpuGrad
  with the forward's is_point_cam_proj (the shim's kernel returns zeros for
  directions); zeros for the directions input and the cameras;
* Pose2EmbeddedGraphCpu (Pose2EmbeddedGraphCpu.py:55-60): zeros;

and registers the three the reference leaves unregistered although their
gradient ops exist, pairing each with the forward outputs marked "output for
gradient operator" (DQSkinningGPUOp.cpp:8-13, LBSkinningGPUOp.cpp:9-14,
MultiViewSilhouetteLossGPUOp.cpp:13-14):

* DqSkinningGpu -> DqSkinningGpuGrad(grad of positions, outputs 0, 2, 3, 4, 5);
* LbSkinningGpu -> LbSkinningGpuGrad(grad of positions, outputs 0, 2, 3, 4, 5);
* MultiViewSilhouetteLossGpu -> MultiViewSilhouetteLossGpuGrad(grads of
  outputs 0 and 1, outputs 2 and 3, the crops input);

plus EmbeddedGraphArapGpu -> EmbeddedGraphArapGpuGrad(grad of the ARAP
residual, output 2) and ProjectedMeshBoundaryGpu as not differentiable.
"""
from __future__ import annotations

_ops = None


def load(library_path: str):
    """Loads the op library and registers the gradients; returns the ops module."""
    global _ops
    import tensorflow as tf
    _ops = tf.load_op_library(library_path)
    _register(tf)
    return _ops


def _register(tf):
    from tensorflow.python.framework import ops

    ops.NotDifferentiable("ProjectedMeshBoundaryGpu")

    @ops.RegisterGradient("CudaRendererGpu")
    def _cuda_renderer_grad(op, g_bary, g_face, g_render, g_normal, g_normal_map):
        albedo = op.get_attr("albedo_mode").decode("utf-8")
        if albedo in ("vertexColor", "textured"):
            grads = _ops.cuda_renderer_grad_gpu(
                render_buffer_grad=g_render,
                vertex_pos=op.inputs[0], vertex_color=op.inputs[1], texture=op.inputs[2], sh_coeff=op.inputs[3],
                target_image=op.inputs[4], vertex_normal=op.outputs[3], barycentric_buffer=op.outputs[0],
                face_buffer=op.outputs[1], extrinsics=op.inputs[5], intrinsics=op.inputs[6],
                faces=op.get_attr("faces"), texture_coordinates=op.get_attr("texture_coordinates"),
                number_of_vertices=op.get_attr("number_of_vertices"),
                number_of_cameras=op.get_attr("number_of_cameras"),
                render_resolution_u=op.get_attr("render_resolution_u"),
                render_resolution_v=op.get_attr("render_resolution_v"),
                albedo_mode=op.get_attr("albedo_mode"), shading_mode=op.get_attr("shading_mode"),
                image_filter_size=op.get_attr("image_filter_size"),
                texture_filter_size=op.get_attr("texture_filter_size"))
        else:
            grads = [tf.zeros(tf.shape(op.inputs[k])) for k in range(4)]
        return (grads[0], grads[1], grads[2], grads[3], tf.zeros(tf.shape(op.inputs[4])),
                tf.zeros(tf.shape(op.inputs[5])), tf.zeros(tf.shape(op.inputs[6])))

    @ops.RegisterGradient("DqSkinningGpu")
    def _dq_skinning_grad(op, g_pos, g_nrm, g_jp, g_ja, g_dq, g_w):
        return _ops.dq_skinning_gpu_grad(
            g_pos, op.outputs[0], op.outputs[2], op.outputs[3], op.outputs[4], op.outputs[5],
            character_file_path_skinning_grad=op.get_attr("character_file_path_skinning"),
            mini_batch_size_skinning_grad=op.get_attr("mini_batch_size_skinning"))

    @ops.RegisterGradient("LbSkinningGpu")
    def _lb_skinning_grad(op, g_pos, g_nrm, g_jp, g_ja, g_tr, g_w):
        return _ops.lb_skinning_gpu_grad(
            g_pos, op.outputs[0], op.outputs[2], op.outputs[3], op.outputs[4], op.outputs[5],
            character_file_path_skinning_grad=op.get_attr("character_file_path_skinning"),
            mini_batch_size_skinning_grad=op.get_attr("mini_batch_size_skinning"))

    @ops.RegisterGradient("EmbeddedGraphGpu")
    def _embedded_graph_grad(op, g_vertices, g_normals, g_markers, g_graph, g_da, g_sa):
        return _ops.embedded_graph_gpu_grad(
            g_vertices, g_markers, op.outputs[4], op.outputs[5],
            character_file_path_eg_grad=op.get_attr("character_file_path_eg"),
            graph_file_path_grad=op.get_attr("graph_file_path"),
            refinement_grad=op.get_attr("refinement"))

    @ops.RegisterGradient("EmbeddedGraphArapGpu")
    def _arap_grad(op, g_res, g_cw, g_rot):
        return _ops.embedded_graph_arap_gpu_grad(
            g_res, op.outputs[2],
            character_file_path_eg_grad=op.get_attr("character_file_path_eg"),
            graph_file_path_grad=op.get_attr("graph_file_path"),
            number_of_batches_eg_grad=op.get_attr("number_of_batches_eg"),
            max_number_of_node_connections_grad=op.get_attr("max_number_of_node_connections"),
            refinement_grad=op.get_attr("refinement"))

    @ops.RegisterGradient("MultiViewSilhouetteLossGpu")
    def _silhouette_grad(op, g_res, g_res1, g_dtg, g_cid):
        g_pts = _ops.multi_view_silhouette_loss_gpu_grad(g_res, g_res1, op.outputs[2], op.outputs[3], op.inputs[4])
        return (g_pts, tf.zeros(tf.shape(op.inputs[1])), None, None, tf.zeros(tf.shape(op.inputs[4])))

    @ops.RegisterGradient("CameraProjectionGpu")
    def _camera_projection_grad(op, g):
        gp = _ops.camera_projection_gpu_grad(points_image_space=g, points_global_space=op.inputs[0],
                                             extrinsics=op.inputs[2], intrinsics=op.inputs[3],
                                             is_point_cam_proj_grad=op.get_attr("is_point_cam_proj"))
        return gp, tf.zeros(tf.shape(op.inputs[1])), tf.zeros(tf.shape(op.inputs[2])), tf.zeros(tf.shape(op.inputs[3]))

    @ops.RegisterGradient("Pose2EmbeddedGraphCpu")
    def _pose2eg_grad(op, g_t, g_a):
        return tf.zeros(tf.shape(op.inputs[0]))
