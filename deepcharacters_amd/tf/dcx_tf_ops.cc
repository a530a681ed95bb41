// TensorFlow custom-op shim over libdcx: the reference's GPU op surface
// (REGISTER_OP text identical to the reference's: op names, input and output
// names, order and dtypes, attribute names, types and defaults), with every
// kernel a thin call into the C ABI of include/dcx.h on TensorFlow's own
// stream.
//
// Build (needs the TensorFlow headers and library of a ROCm TensorFlow, not
// present in this repository's build image):
//   TF_CFLAGS=$(python -c 'import tensorflow as tf; print(" ".join(tf.sysconfig.get_compile_flags()))')
//   TF_LFLAGS=$(python -c 'import tensorflow as tf; print(" ".join(tf.sysconfig.get_link_flags()))')
//   g++ -std=c++17 -O2 -shared -fPIC -D__HIP_PLATFORM_AMD__ -DEIGEN_USE_HIP -DTENSORFLOW_USE_ROCM \
//       -I../../include -I$(python3-config --includes) $TF_CFLAGS dcx_tf_ops.cc dcx_tf_host.cc \
//       -o libdcx_tf_ops.so -L.. -ldcx $TF_LFLAGS
// and in Python: tf.load_op_library(".../libdcx_tf_ops.so"), then import
// deepcharacters_amd.tf.gradients to register the gradients.
//
// Differences from the reference, each a decision, not an omission:
//  * file-reading constructors (character, graph and mesh files) read them
//    with the package's readers through the embedding Python interpreter
//    (dcx_tf_host.cc), so the op library must be loaded from Python;
//  * the DOF -> dual-quaternion / transform forward kinematics run on the GPU
//    (no host round trip per call);
//  * the skinning-weight gradients are the exact adjoint; setting the
//    environment variable DCX_TF_REFERENCE_WEIGHT_GRADIENT=1 reproduces the
//    reference's formulas (dcx_dqs_bwd / dcx_lbs_bwd with DCX_FLAG_REF_COMPAT);
//  * outputs the reference leaves unwritten (the normal map when
//    compute_normal_map is 'none', the markers of a graph without markers)
//    are zero-filled;
//  * Pose2EmbeddedGraphCpu keeps its name but runs on the GPU;
//  * as in the reference, the `refinement` attributes are read but the graph
//    is always built without refinement (EmbeddedGraphGPUOp.cpp:53), and
//    `use_gap_detection` is read but unused.
#include <algorithm>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "tensorflow/core/framework/op.h"
#include "tensorflow/core/framework/op_kernel.h"
#include "tensorflow/core/framework/tensor.h"

#include "../../include/dcx.h"
#include "dcx_tf_host.h"

using namespace tensorflow;

namespace {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

void* stream_of(OpKernelContext* ctx) { return (void*)ctx->eigen_gpu_device().stream(); }

#define DCX_TF_CALL(ctx, call)                                                              \
  do {                                                                                      \
    const dcx_status st_ = (call);                                                          \
    OP_REQUIRES(ctx, st_ == DCX_OK, errors::Internal(#call, " failed: ", dcx_last_error())); \
  } while (0)

Status workspace(OpKernelContext* ctx, size_t bytes, Tensor* t) {
  return ctx->allocate_temp(DT_UINT8, TensorShape({(int64_t)std::max<size_t>(bytes, 256)}), t);
}

template <class T>
T* out_ptr(Tensor* t) { return t->flat<T>().data(); }
template <class T>
const T* in_ptr(const Tensor& t) { return t.flat<T>().data(); }
void* ws_ptr(Tensor& t) { return (void*)t.flat<uint8_t>().data(); }

void zero_fill(OpKernelContext* ctx, Tensor* t) {
  if (t->TotalBytes()) (void)hipMemsetAsync((void*)t->tensor_data().data(), 0, t->TotalBytes(),
                                            (hipStream_t)stream_of(ctx));
}

bool reference_weight_gradient() {
  const char* v = std::getenv("DCX_TF_REFERENCE_WEIGHT_GRADIENT");
  return v && std::string(v) == "1";
}

Status load_tables(const std::string& fn, const std::vector<std::string>& args, dcx_tf::HostTables* t) {
  const std::string err = dcx_tf::CallCharacterIO(fn, args, t);
  if (!err.empty()) return errors::InvalidArgument(fn, ": ", err);
  return Status::OK();
}

const dcx_tf::HostArray& table(const dcx_tf::HostTables& t, const char* name) { return t.at(name); }

// Device copy of a host table, kept by the kernel for its lifetime.
struct DeviceArray {
  void* p = nullptr;
  ~DeviceArray() { if (p) (void)hipFree(p); }
  Status upload(const dcx_tf::HostArray& a) {
    if (a.bytes.empty()) return Status::OK();
    if (hipMalloc(&p, a.bytes.size()) != hipSuccess) return errors::ResourceExhausted("hipMalloc");
    if (hipMemcpy(p, a.bytes.data(), a.bytes.size(), hipMemcpyHostToDevice) != hipSuccess)
      return errors::Internal("hipMemcpy");
    return Status::OK();
  }
  const float* f() const { return (const float*)p; }
  const int32_t* i() const { return (const int32_t*)p; }
};

int albedo_code(const std::string& s) {
  static const char* names[] = {"vertexColor", "textured", "normal", "lighting", "foregroundMask", "position", "depth", "uv"};
  for (int k = 0; k < 8; ++k)
    if (s == names[k]) return k;
  return -1;
}

// The CudaRenderer attributes (CudaRenderer.cpp:35-101): the topology from
// `faces` / `texture_coordinates`, and the raster configuration.
struct RendererAttrs {
  std::vector<int> faces;
  std::vector<float> texcoords;
  int N = 0, C = 0, U = 0, V = 0, img_fs = 2, tex_fs = 2;
  std::string albedo, shading, normal_map = "none";
  dcx_topo* topo = nullptr;
  ~RendererAttrs() { if (topo) dcx_topo_destroy(topo); }

  Status read(OpKernelConstruction* ctx, bool with_normal_map) {
    TF_RETURN_IF_ERROR(ctx->GetAttr("faces", &faces));
    TF_RETURN_IF_ERROR(ctx->GetAttr("texture_coordinates", &texcoords));
    TF_RETURN_IF_ERROR(ctx->GetAttr("number_of_vertices", &N));
    TF_RETURN_IF_ERROR(ctx->GetAttr("number_of_cameras", &C));
    TF_RETURN_IF_ERROR(ctx->GetAttr("render_resolution_u", &U));
    TF_RETURN_IF_ERROR(ctx->GetAttr("render_resolution_v", &V));
    TF_RETURN_IF_ERROR(ctx->GetAttr("albedo_mode", &albedo));
    TF_RETURN_IF_ERROR(ctx->GetAttr("shading_mode", &shading));
    TF_RETURN_IF_ERROR(ctx->GetAttr("image_filter_size", &img_fs));
    TF_RETURN_IF_ERROR(ctx->GetAttr("texture_filter_size", &tex_fs));
    if (with_normal_map) TF_RETURN_IF_ERROR(ctx->GetAttr("compute_normal_map", &normal_map));
    if (N <= 0) return errors::InvalidArgument("number_of_vertices not set!");
    if (C <= 0) return errors::InvalidArgument("number_of_cameras not set!");
    if (U <= 0 || V <= 0) return errors::InvalidArgument("render_resolution_u/v not set!");
    if (albedo_code(albedo) < 0) return errors::InvalidArgument("INVALID ALBEDO MODE ", albedo);
    if (shading != "shaded" && shading != "shadeless") return errors::InvalidArgument("INVALID SHADING MODE ", shading);
    if (albedo != "vertexColor" && albedo != "textured") shading = "shadeless";  // CudaRenderer.cpp:71-74
    if (normal_map != "none" && normal_map != "normal" && normal_map != "position" && normal_map != "face")
      return errors::InvalidArgument("INVALID Normal MODE ", normal_map);
    if (faces.size() % 3) return errors::InvalidArgument("faces: length not a multiple of 3");
    const int F = (int)faces.size() / 3;
    const float* tc = texcoords.size() == (size_t)F * 6 ? texcoords.data() : nullptr;
    if (dcx_topo_create(faces.data(), F, N, tc, &topo) != DCX_OK) return errors::InvalidArgument(dcx_last_error());
    return Status::OK();
  }

  dcx_raster_cfg cfg() const {
    dcx_raster_cfg c{};
    c.num_cameras = C;
    c.width = U;
    c.height = V;
    c.albedo_mode = albedo_code(albedo);
    c.shading_mode = shading == "shaded" ? DCX_SHADING_SHADED : DCX_SHADING_SHADELESS;
    c.image_filter_size = img_fs;
    c.texture_filter_size = tex_fs;
    c.rule = DCX_RULE_NATIVE;
    c.path = DCX_PATH_BINNED;
    return c;
  }
};

}  // namespace

// ===========================================================================
// CudaRendererGpu / CudaRendererGradGpu (CudaRenderer.cpp, CudaRendererGrad.cpp)
// ===========================================================================

REGISTER_OP("CudaRendererGpu")
.Input("vertex_pos: float")
.Input("vertex_color: float")
.Input("texture: float")
.Input("sh_coeff: float")
.Input("target_image: float")
.Input("extrinsics: float")
.Input("intrinsics: float")
.Output("barycentric_buffer: float")
.Output("face_buffer: int32")
.Output("render_buffer: float")
.Output("vertex_normal: float")
.Output("normal_map: float")
.Attr("faces: list(int)")
.Attr("texture_coordinates: list(float)")
.Attr("number_of_vertices: int")
.Attr("number_of_cameras: int")
.Attr("render_resolution_u: int = 512")
.Attr("render_resolution_v: int = 512")
.Attr("albedo_mode: string")
.Attr("shading_mode: string")
.Attr("image_filter_size: int = 2")
.Attr("texture_filter_size: int = 2")
.Attr("compute_normal_map: string");

class CudaRendererGpuOp : public OpKernel {
 public:
  explicit CudaRendererGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) { OP_REQUIRES_OK(ctx, a_.read(ctx, true)); }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &vp = ctx->input(0), &vc = ctx->input(1), &tex = ctx->input(2), &sh = ctx->input(3);
    const Tensor &ext = ctx->input(5), &intr = ctx->input(6);
    // the batch size is the texture's leading dimension (CudaRenderer.cpp:233-235)
    const int B = (int)tex.dim_size(0), TH = (int)tex.dim_size(1), TW = (int)tex.dim_size(2);
    const int C = a_.C, H = a_.V, W = a_.U, N = a_.N;
    OP_REQUIRES(ctx, vp.NumElements() == (int64_t)B * N * 3, errors::InvalidArgument("vertex_pos: expected [B,N,3]"));
    Tensor *bary = nullptr, *face = nullptr, *render = nullptr, *vn = nullptr, *nmap = nullptr;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, C, H, W, 2}), &bary));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, C, H, W}), &face));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(2, TensorShape({B, C, H, W, 3}), &render));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(3, TensorShape({B, C, N, 3}), &vn));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(4, TensorShape({B, TH, TW, 3}), &nmap));
    const dcx_raster_cfg cfg = a_.cfg();
    void* s = stream_of(ctx);
    Tensor ws;
    if (a_.normal_map == "none") {
      OP_REQUIRES_OK(ctx, workspace(ctx, dcx_render_workspace_size(a_.topo, &cfg, B), &ws));
      DCX_TF_CALL(ctx, dcx_render_fwd(a_.topo, &cfg, B, in_ptr<float>(vp), in_ptr<float>(vc), in_ptr<float>(tex), TH,
                                      TW, in_ptr<float>(sh), in_ptr<float>(ext), in_ptr<float>(intr),
                                      out_ptr<float>(bary), out_ptr<int32_t>(face), out_ptr<float>(render),
                                      out_ptr<float>(vn), ws_ptr(ws), ws.NumElements(), s));
      zero_fill(ctx, nmap);
    } else {
      const int mode = a_.normal_map == "normal" ? DCX_NORMAL_MAP_NORMAL
                       : a_.normal_map == "position" ? DCX_NORMAL_MAP_POSITION : DCX_NORMAL_MAP_FACE;
      DCX_TF_CALL(ctx, dcx_topo_texel_table(a_.topo, TH, TW, s));
      OP_REQUIRES_OK(ctx, workspace(ctx, dcx_normal_map_workspace_size(a_.topo, B), &ws));
      DCX_TF_CALL(ctx, dcx_render_normal_map(a_.topo, &cfg, B, mode, in_ptr<float>(vp), TH, TW, out_ptr<float>(bary),
                                             out_ptr<int32_t>(face), out_ptr<float>(render), out_ptr<float>(vn),
                                             out_ptr<float>(nmap), ws_ptr(ws), ws.NumElements(), s));
    }
  }

 private:
  RendererAttrs a_;
};
REGISTER_KERNEL_BUILDER(Name("CudaRendererGpu").Device(DEVICE_GPU), CudaRendererGpuOp);

REGISTER_OP("CudaRendererGradGpu")
.Input("render_buffer_grad: float")
.Input("vertex_pos: float")
.Input("vertex_color: float")
.Input("texture: float")
.Input("sh_coeff: float")
.Input("target_image: float")
.Input("vertex_normal: float")
.Input("barycentric_buffer: float")
.Input("face_buffer: int32")
.Input("extrinsics: float")
.Input("intrinsics: float")
.Output("vertex_pos_grad: float")
.Output("vertex_color_grad: float")
.Output("texture_grad: float")
.Output("sh_coeff_grad: float")
.Attr("faces: list(int)")
.Attr("texture_coordinates: list(float)")
.Attr("number_of_vertices: int")
.Attr("number_of_cameras: int")
.Attr("render_resolution_u: int = 512")
.Attr("render_resolution_v: int = 512")
.Attr("albedo_mode: string")
.Attr("shading_mode: string")
.Attr("image_filter_size: int = 2")
.Attr("texture_filter_size: int = 2");

class CudaRendererGradGpuOp : public OpKernel {
 public:
  explicit CudaRendererGradGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) { OP_REQUIRES_OK(ctx, a_.read(ctx, false)); }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &g = ctx->input(0), &vp = ctx->input(1), &vc = ctx->input(2), &tex = ctx->input(3);
    const Tensor &sh = ctx->input(4), &tgt = ctx->input(5), &vn = ctx->input(6), &bary = ctx->input(7);
    const Tensor &face = ctx->input(8), &ext = ctx->input(9), &intr = ctx->input(10);
    // the batch size is vertex_pos's leading dimension (CudaRendererGrad.cpp:184)
    const int B = (int)vp.dim_size(0), TH = (int)tex.dim_size(1), TW = (int)tex.dim_size(2);
    Tensor *gvp = nullptr, *gvc = nullptr, *gtex = nullptr, *gsh = nullptr;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, a_.N, 3}), &gvp));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, a_.N, 3}), &gvc));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(2, TensorShape({B, TH, TW, 3}), &gtex));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(3, TensorShape({B, a_.C, 27}), &gsh));
    const dcx_raster_cfg cfg = a_.cfg();
    Tensor ws;
    OP_REQUIRES_OK(ctx, workspace(ctx, dcx_render_bwd_workspace_size(a_.topo, &cfg, B), &ws));
    DCX_TF_CALL(ctx, dcx_render_bwd(a_.topo, &cfg, B, in_ptr<float>(g), in_ptr<float>(vp), in_ptr<float>(vc),
                                    in_ptr<float>(tex), TH, TW, in_ptr<float>(sh), in_ptr<float>(tgt),
                                    in_ptr<float>(vn), in_ptr<float>(bary), in_ptr<int32_t>(face),
                                    in_ptr<float>(ext), in_ptr<float>(intr), out_ptr<float>(gvp), out_ptr<float>(gvc),
                                    out_ptr<float>(gtex), out_ptr<float>(gsh), ws_ptr(ws), ws.NumElements(),
                                    stream_of(ctx)));
  }

 private:
  RendererAttrs a_;
};
REGISTER_KERNEL_BUILDER(Name("CudaRendererGradGpu").Device(DEVICE_GPU), CudaRendererGradGpuOp);

// ===========================================================================
// DqSkinningGpu / LbSkinningGpu and their gradients (DQSkinningGPUOp.cpp,
// DQSkinningGPUOpGrad.cpp, LBSkinningGPUOp.cpp, LBSkinningGPUOpGrad.cpp)
// ===========================================================================

namespace {

// The skinned character of the skinning ops: rest mesh and skinning table on
// the device, the kinematic tree as a dcx_skeleton, and (for the gradient
// ops) the DOF influence table.
struct SkinnedCharacter {
  dcx_tf::HostTables t;
  DeviceArray rest_pos, rest_nrm, joint_idx, influence;
  dcx_skeleton* skel = nullptr;
  int N = 0, K = 0, J = 0, Js = 0, D = 0, max_entries = 0;
  ~SkinnedCharacter() { if (skel) dcx_skeleton_destroy(skel); }

  Status load(const std::string& path, bool with_influence) {
    if (path == "None") return errors::InvalidArgument("character_file_path_skinning not set!");
    TF_RETURN_IF_ERROR(load_tables("skinning_character", {path}, &t));
    const auto &P = table(t, "rest_pos"), &JI = table(t, "joint_idx"), &parent = table(t, "parent");
    N = (int)P.shape[0];
    K = (int)JI.shape[1];
    J = (int)parent.shape[0];
    Js = (int)table(t, "skin_joint").shape[0];
    D = table(t, "num_dofs").i()[0];
    if (dcx_skeleton_create(J, D, parent.i(), table(t, "type").i(), table(t, "dof").i(), table(t, "offset").f(),
                            table(t, "axis").f(), Js, table(t, "skin_joint").i(), table(t, "skin_scale").f(),
                            table(t, "skin_bind").f(), &skel) != DCX_OK)
      return errors::InvalidArgument(dcx_last_error());
    TF_RETURN_IF_ERROR(rest_pos.upload(P));
    TF_RETURN_IF_ERROR(rest_nrm.upload(table(t, "rest_nrm")));
    TF_RETURN_IF_ERROR(joint_idx.upload(JI));
    if (with_influence) {
      // the DOF influence table of DQSkinningGPUOpGrad.cpp:90-217, for the
      // static skinning weights of the character file
      int need = 0;
      if (dcx_dof_influence(skel, N, K, JI.i(), table(t, "joint_w").f(), 1, nullptr, &need) != DCX_OK)
        return errors::Internal(dcx_last_error());
      max_entries = std::max(1, need);
      dcx_tf::HostArray tab;
      tab.kind = 'f';
      tab.bytes.resize(sizeof(float) * 4 * (size_t)D * max_entries);
      if (dcx_dof_influence(skel, N, K, JI.i(), table(t, "joint_w").f(), max_entries, (float*)tab.bytes.data(),
                            &need) != DCX_OK)
        return errors::Internal(dcx_last_error());
      TF_RETURN_IF_ERROR(influence.upload(tab));
    }
    return Status::OK();
  }
};

}  // namespace

REGISTER_OP("DqSkinningGpu")
.Input("dofs: float")
.Input("skinning_weights: float")
.Output("skinned_vertex_positions: float")
.Output("skinned_vertex_normals: float")
.Output("global_joint_position: float")
.Output("global_joint_axis: float")
.Output("dual_quaternions: float")
.Output("output_skinning_weights: float")
.Attr("character_file_path_skinning: string = 'None'")
.Attr("mini_batch_size_skinning: int = 0");

class DqSkinningGpuOp : public OpKernel {
 public:
  explicit DqSkinningGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string path;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_skinning", &path));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("mini_batch_size_skinning", &B_));
    OP_REQUIRES_OK(ctx, c_.load(path, false));
  }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &dofs = ctx->input(0), &w = ctx->input(1);
    const int B = B_, N = c_.N, K = c_.K;
    OP_REQUIRES(ctx, dofs.NumElements() == (int64_t)B * c_.D, errors::InvalidArgument("dofs: expected [B, ", c_.D, "]"));
    OP_REQUIRES(ctx, w.NumElements() == (int64_t)B * N * K, errors::InvalidArgument("skinning_weights: expected [B,N,K]"));
    Tensor *pos, *nrm, *jp, *ja, *dq, *wo;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, N, 3}), &pos));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, N, 3}), &nrm));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(2, TensorShape({B, c_.J, 3}), &jp));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(3, TensorShape({B, c_.J, 3}), &ja));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(4, TensorShape({B, c_.Js, 8}), &dq));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(5, TensorShape({B, N, K}), &wo));
    void* s = stream_of(ctx);
    DCX_TF_CALL(ctx, dcx_fk(c_.skel, B, in_ptr<float>(dofs), out_ptr<float>(dq), out_ptr<float>(jp), out_ptr<float>(ja), s));
    DCX_TF_CALL(ctx, dcx_dqs_fwd(out_ptr<float>(dq), in_ptr<float>(w), c_.joint_idx.i(), c_.rest_pos.f(), c_.rest_nrm.f(),
                                 B, N, K, c_.Js, out_ptr<float>(pos), out_ptr<float>(nrm), nullptr, out_ptr<float>(wo), s));
  }

 private:
  SkinnedCharacter c_;
  int B_ = 0;
};
REGISTER_KERNEL_BUILDER(Name("DqSkinningGpu").Device(DEVICE_GPU), DqSkinningGpuOp);

REGISTER_OP("DqSkinningGpuGrad")
.Input("skinned_vertex_positions_grad: float")
.Input("skinned_vertex_positions: float")
.Input("global_joint_position: float")
.Input("global_joint_axis: float")
.Input("dual_quaternions: float")
.Input("skinning_weights: float")
.Output("dofs_grad: float")
.Output("skinning_weights_grad: float")
.Attr("character_file_path_skinning_grad: string = 'None'")
.Attr("mini_batch_size_skinning_grad: int = 0");

class DqSkinningGpuGradOp : public OpKernel {
 public:
  explicit DqSkinningGpuGradOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string path;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_skinning_grad", &path));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("mini_batch_size_skinning_grad", &B_));
    OP_REQUIRES_OK(ctx, c_.load(path, true));
  }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &g = ctx->input(0), &pos = ctx->input(1), &jp = ctx->input(2), &ja = ctx->input(3);
    const Tensor &dq = ctx->input(4), &w = ctx->input(5);
    const int B = B_, N = c_.N, K = c_.K;
    Tensor *gd, *gw;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, c_.D}), &gd));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, N, K}), &gw));
    void* s = stream_of(ctx);
    DCX_TF_CALL(ctx, dcx_dof_grad(in_ptr<float>(g), in_ptr<float>(pos), in_ptr<float>(jp), in_ptr<float>(ja),
                                  c_.influence.f(), B, N, c_.J, c_.D, c_.max_entries, out_ptr<float>(gd), s));
    DCX_TF_CALL(ctx, dcx_dqs_bwd(in_ptr<float>(g), nullptr, in_ptr<float>(dq), in_ptr<float>(w), c_.joint_idx.i(),
                                 c_.rest_pos.f(), c_.rest_nrm.f(), B, N, K, c_.Js, nullptr, out_ptr<float>(gw),
                                 nullptr, reference_weight_gradient() ? DCX_FLAG_REF_COMPAT : 0, s));
  }

 private:
  SkinnedCharacter c_;
  int B_ = 0;
};
REGISTER_KERNEL_BUILDER(Name("DqSkinningGpuGrad").Device(DEVICE_GPU), DqSkinningGpuGradOp);

REGISTER_OP("LbSkinningGpu")
.Input("dofs: float")
.Input("skinning_weights: float")
.Input("displacement: float")
.Output("skinned_vertex_positions: float")
.Output("skinned_vertex_normals: float")
.Output("global_joint_position: float")
.Output("global_joint_axis: float")
.Output("transformation: float")
.Output("output_skinning_weights: float")
.Attr("character_file_path_skinning: string = 'None'")
.Attr("mini_batch_size_skinning: int = 0");

class LbSkinningGpuOp : public OpKernel {
 public:
  explicit LbSkinningGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string path;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_skinning", &path));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("mini_batch_size_skinning", &B_));
    OP_REQUIRES_OK(ctx, c_.load(path, false));
  }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &dofs = ctx->input(0), &w = ctx->input(1), &disp = ctx->input(2);
    const int B = B_, N = c_.N, K = c_.K;
    OP_REQUIRES(ctx, dofs.NumElements() == (int64_t)B * c_.D, errors::InvalidArgument("dofs: expected [B, ", c_.D, "]"));
    Tensor *pos, *nrm, *jp, *ja, *tr, *wo;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, N, 3}), &pos));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, N, 3}), &nrm));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(2, TensorShape({B, c_.J, 3}), &jp));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(3, TensorShape({B, c_.J, 3}), &ja));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(4, TensorShape({B, c_.Js, 12}), &tr));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(5, TensorShape({B, N, K}), &wo));
    void* s = stream_of(ctx);
    DCX_TF_CALL(ctx, dcx_fk_transforms(c_.skel, B, in_ptr<float>(dofs), out_ptr<float>(tr), out_ptr<float>(jp),
                                       out_ptr<float>(ja), s));
    DCX_TF_CALL(ctx, dcx_lbs_fwd(out_ptr<float>(tr), in_ptr<float>(w), c_.joint_idx.i(), c_.rest_pos.f(), c_.rest_nrm.f(),
                                 in_ptr<float>(disp), B, N, K, c_.Js, out_ptr<float>(pos), out_ptr<float>(nrm), nullptr,
                                 out_ptr<float>(wo), s));
  }

 private:
  SkinnedCharacter c_;
  int B_ = 0;
};
REGISTER_KERNEL_BUILDER(Name("LbSkinningGpu").Device(DEVICE_GPU), LbSkinningGpuOp);

REGISTER_OP("LbSkinningGpuGrad")
.Input("skinned_vertex_positions_grad: float")
.Input("skinned_vertex_positions: float")
.Input("global_joint_position: float")
.Input("global_joint_axis: float")
.Input("transformation: float")
.Input("skinning_weights: float")
.Output("dofs_grad: float")
.Output("skinning_weights_grad: float")
.Output("displacement_grad: float")
.Attr("character_file_path_skinning_grad: string = 'None'")
.Attr("mini_batch_size_skinning_grad: int = 0");

class LbSkinningGpuGradOp : public OpKernel {
 public:
  explicit LbSkinningGpuGradOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string path;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_skinning_grad", &path));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("mini_batch_size_skinning_grad", &B_));
    OP_REQUIRES_OK(ctx, c_.load(path, true));
  }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &g = ctx->input(0), &pos = ctx->input(1), &jp = ctx->input(2), &ja = ctx->input(3);
    const Tensor &tr = ctx->input(4), &w = ctx->input(5);
    const int B = B_, N = c_.N, K = c_.K;
    Tensor *gd, *gw, *gdisp;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, c_.D}), &gd));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, N, K}), &gw));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(2, TensorShape({B, N, 3}), &gdisp));
    Tensor zero_disp;
    OP_REQUIRES_OK(ctx, ctx->allocate_temp(DT_FLOAT, TensorShape({B, N, 3}), &zero_disp));
    zero_fill(ctx, &zero_disp);  // the displacements the forward used are not an input of this op
    void* s = stream_of(ctx);
    DCX_TF_CALL(ctx, dcx_dof_grad(in_ptr<float>(g), in_ptr<float>(pos), in_ptr<float>(jp), in_ptr<float>(ja),
                                  c_.influence.f(), B, N, c_.J, c_.D, c_.max_entries, out_ptr<float>(gd), s));
    DCX_TF_CALL(ctx, dcx_lbs_bwd(in_ptr<float>(g), in_ptr<float>(tr), in_ptr<float>(w), c_.joint_idx.i(), c_.rest_pos.f(),
                                 zero_disp.flat<float>().data(), B, N, K, c_.Js, out_ptr<float>(gw),
                                 out_ptr<float>(gdisp), nullptr, reference_weight_gradient() ? DCX_FLAG_REF_COMPAT : 0, s));
  }

 private:
  SkinnedCharacter c_;
  int B_ = 0;
};
REGISTER_KERNEL_BUILDER(Name("LbSkinningGpuGrad").Device(DEVICE_GPU), LbSkinningGpuGradOp);

// ===========================================================================
// EmbeddedGraphGpu / EmbeddedGraphGpuGrad, EmbeddedGraphArapGpu(+Grad),
// Pose2EmbeddedGraphCpu (EmbeddedGraphGPUOp.cpp, EmbeddedGraphArapGPUOp.cpp,
// Pose2EmbeddedGraphCPUOp.cpp)
// ===========================================================================

namespace {

// The character and graph tables of the graph ops.
struct GraphCharacter {
  dcx_tf::HostTables t;
  DeviceArray rest_pos, node_vertex, nn_off, nn_size, nn_idx, rigidity, joint_idx, joint_w;
  dcx_topo* topo = nullptr;
  dcx_edgraph* graph = nullptr;
  int N = 0, M = 0, Mk = 0, K = 0;
  ~GraphCharacter() {
    if (graph) dcx_edgraph_destroy(graph);
    if (topo) dcx_topo_destroy(topo);
  }

  Status load(const std::string& cpath, const std::string& gpath, bool refinement, bool device_graph) {
    if (cpath == "None") return errors::InvalidArgument("character_file_path_eg not set!");
    if (gpath == "None") return errors::InvalidArgument("graph_file_path not set!");
    TF_RETURN_IF_ERROR(load_tables("graph_character", {cpath, gpath, refinement ? "1" : "0"}, &t));
    const auto &P = table(t, "rest_pos"), &F = table(t, "faces"), &nv = table(t, "node_vertex");
    N = (int)P.shape[0];
    M = (int)nv.shape[0];
    Mk = (int)table(t, "markers").shape[0];
    K = (int)table(t, "joint_idx").shape[1];
    TF_RETURN_IF_ERROR(rest_pos.upload(P));
    TF_RETURN_IF_ERROR(node_vertex.upload(nv));
    TF_RETURN_IF_ERROR(nn_off.upload(table(t, "nn_offsets")));
    TF_RETURN_IF_ERROR(nn_size.upload(table(t, "nn_sizes")));
    TF_RETURN_IF_ERROR(nn_idx.upload(table(t, "nn_indices")));
    TF_RETURN_IF_ERROR(rigidity.upload(table(t, "rigidity")));
    TF_RETURN_IF_ERROR(joint_idx.upload(table(t, "joint_idx")));
    TF_RETURN_IF_ERROR(joint_w.upload(table(t, "joint_w")));
    if (device_graph) {
      if (dcx_topo_create(F.i(), (int)F.shape[0], N, nullptr, &topo) != DCX_OK)
        return errors::InvalidArgument(dcx_last_error());
      if (dcx_edgraph_create(P.f(), N, M, nv.i(), table(t, "v2n_offsets").i(), table(t, "v2n_nodes").i(),
                             table(t, "v2n_weights").f(), Mk, Mk ? table(t, "markers").f() : nullptr,
                             Mk ? table(t, "marker_node").i() : nullptr, topo, &graph) != DCX_OK)
        return errors::InvalidArgument(dcx_last_error());
    }
    return Status::OK();
  }
};

}  // namespace

REGISTER_OP("EmbeddedGraphGpu")
.Input("nodes_delta_translation: float")
.Input("nodes_delta_rotation: float")
.Input("nodes_skinned_translation: float")
.Input("nodes_skinned_rotation: float")
.Input("vertex_displacements: float")
.Output("deformed_vertices: float")
.Output("deformed_normals: float")
.Output("deformed_markers: float")
.Output("deformed_graph: float")
.Output("d_delta_rotation: float")
.Output("d_skinned_rotation: float")
.Attr("character_file_path_eg: string = 'None'")
.Attr("graph_file_path: string = 'None'")
.Attr("number_of_batches_eg: int = 0")
.Attr("refinement: bool = false");

class EmbeddedGraphGpuOp : public OpKernel {
 public:
  explicit EmbeddedGraphGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string cpath, gpath;
    bool refinement = false;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_eg", &cpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("graph_file_path", &gpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("number_of_batches_eg", &B_));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("refinement", &refinement));
    OP_REQUIRES_OK(ctx, g_.load(cpath, gpath, false, true));  // refinement is ignored (EmbeddedGraphGPUOp.cpp:53)
  }

  void Compute(OpKernelContext* ctx) override {
    const int B = B_, N = g_.N, M = g_.M;
    Tensor* o[6];
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, N, 3}), &o[0]));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, N, 3}), &o[1]));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(2, TensorShape({B, g_.Mk, 3}), &o[2]));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(3, TensorShape({B, M, 3}), &o[3]));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(4, TensorShape({B, M, 3}), &o[4]));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(5, TensorShape({B, M, 3}), &o[5]));
    Tensor ws;
    OP_REQUIRES_OK(ctx, workspace(ctx, dcx_ed_workspace_size(g_.graph, B), &ws));
    DCX_TF_CALL(ctx, dcx_ed_fwd(g_.graph, B, in_ptr<float>(ctx->input(0)), in_ptr<float>(ctx->input(1)),
                                in_ptr<float>(ctx->input(2)), in_ptr<float>(ctx->input(3)), in_ptr<float>(ctx->input(4)),
                                out_ptr<float>(o[0]), out_ptr<float>(o[1]), g_.Mk ? out_ptr<float>(o[2]) : nullptr,
                                out_ptr<float>(o[3]), out_ptr<float>(o[4]), out_ptr<float>(o[5]), ws_ptr(ws),
                                ws.NumElements(), stream_of(ctx)));
  }

 private:
  GraphCharacter g_;
  int B_ = 0;
};
REGISTER_KERNEL_BUILDER(Name("EmbeddedGraphGpu").Device(DEVICE_GPU), EmbeddedGraphGpuOp);

REGISTER_OP("EmbeddedGraphGpuGrad")
.Input("deformed_vertices_grad: float")
.Input("deformed_markers_grad: float")
.Input("d_delta_a: float")
.Input("d_skinned_a: float")
.Output("nodes_t_grad: float")
.Output("nodes_r_grad: float")
.Output("nodes_skinned_t_grad: float")
.Output("nodes_skinned_r_grad: float")
.Output("vertex_displacements_grad: float")
.Attr("character_file_path_eg_grad: string = 'None'")
.Attr("graph_file_path_grad: string = 'None'")
.Attr("refinement_grad: bool = false");

class EmbeddedGraphGpuGradOp : public OpKernel {
 public:
  explicit EmbeddedGraphGpuGradOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string cpath, gpath;
    bool refinement = false;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_eg_grad", &cpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("graph_file_path_grad", &gpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("refinement_grad", &refinement));
    OP_REQUIRES_OK(ctx, g_.load(cpath, gpath, false, true));  // refinement is ignored (EmbeddedGraphGPUOp.cpp:53)
  }

  void Compute(OpKernelContext* ctx) override {
    const Tensor& gv = ctx->input(0);
    const int B = (int)gv.dim_size(0), N = g_.N, M = g_.M;  // EmbeddedGraphGPUOpGrad.cpp:112
    Tensor* o[5];
    for (int k = 0; k < 4; ++k) OP_REQUIRES_OK(ctx, ctx->allocate_output(k, TensorShape({B, M, 3}), &o[k]));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(4, TensorShape({B, N, 3}), &o[4]));
    Tensor ws;
    OP_REQUIRES_OK(ctx, workspace(ctx, dcx_ed_workspace_size(g_.graph, B), &ws));
    DCX_TF_CALL(ctx, dcx_ed_bwd(g_.graph, B, in_ptr<float>(gv), g_.Mk ? in_ptr<float>(ctx->input(1)) : nullptr,
                                in_ptr<float>(ctx->input(2)), in_ptr<float>(ctx->input(3)), out_ptr<float>(o[0]),
                                out_ptr<float>(o[1]), out_ptr<float>(o[2]), out_ptr<float>(o[3]), out_ptr<float>(o[4]),
                                ws_ptr(ws), ws.NumElements(), stream_of(ctx)));
  }

 private:
  GraphCharacter g_;
};
REGISTER_KERNEL_BUILDER(Name("EmbeddedGraphGpuGrad").Device(DEVICE_GPU), EmbeddedGraphGpuGradOp);

REGISTER_OP("EmbeddedGraphArapGpu")
.Input("nodes_translation: float")
.Input("nodes_rotation: float")
.Output("nodes_arap_loss: float")
.Output("d_connection_weights: float")
.Output("d_rotation: float")
.Attr("character_file_path_eg: string = 'None'")
.Attr("graph_file_path: string = 'None'")
.Attr("number_of_batches_eg: int = 0")
.Attr("max_number_of_node_connections: int = 0")
.Attr("refinement: bool = false");

class EmbeddedGraphArapGpuOp : public OpKernel {
 public:
  explicit EmbeddedGraphArapGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string cpath, gpath;
    bool refinement = false;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_eg", &cpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("graph_file_path", &gpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("number_of_batches_eg", &B_));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("max_number_of_node_connections", &C_));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("refinement", &refinement));
    OP_REQUIRES(ctx, C_ > 0, errors::InvalidArgument("max_number_of_node_connections not set!"));
    OP_REQUIRES_OK(ctx, g_.load(cpath, gpath, false, false));  // refinement is ignored, as above
  }

  void Compute(OpKernelContext* ctx) override {
    const int B = B_, M = g_.M;
    Tensor *res, *cw, *rot;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, M, C_, 3}), &res));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, M, C_}), &cw));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(2, TensorShape({B, M, 3}), &rot));
    DCX_TF_CALL(ctx, dcx_arap_fwd(g_.rest_pos.f(), g_.node_vertex.i(), g_.nn_off.i(), g_.nn_size.i(), g_.nn_idx.i(),
                                  g_.rigidity.f(), B, M, C_, in_ptr<float>(ctx->input(0)), in_ptr<float>(ctx->input(1)),
                                  out_ptr<float>(res), out_ptr<float>(cw), out_ptr<float>(rot), stream_of(ctx)));
  }

 private:
  GraphCharacter g_;
  int B_ = 0, C_ = 0;
};
REGISTER_KERNEL_BUILDER(Name("EmbeddedGraphArapGpu").Device(DEVICE_GPU), EmbeddedGraphArapGpuOp);

REGISTER_OP("EmbeddedGraphArapGpuGrad")
.Input("nodes_arap_loss: float")
.Input("d_a: float")
.Output("nodes_t_grad: float")
.Output("nodes_r_grad: float")
.Attr("character_file_path_eg_grad: string = 'None'")
.Attr("graph_file_path_grad: string = 'None'")
.Attr("number_of_batches_eg_grad: int = 0")
.Attr("max_number_of_node_connections_grad: int = 0")
.Attr("refinement_grad: bool = false");

class EmbeddedGraphArapGpuGradOp : public OpKernel {
 public:
  explicit EmbeddedGraphArapGpuGradOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string cpath, gpath;
    bool refinement = false;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_eg_grad", &cpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("graph_file_path_grad", &gpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("number_of_batches_eg_grad", &B_));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("max_number_of_node_connections_grad", &C_));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("refinement_grad", &refinement));
    OP_REQUIRES(ctx, C_ > 0, errors::InvalidArgument("max_number_of_node_connections_grad not set!"));
    OP_REQUIRES_OK(ctx, g_.load(cpath, gpath, false, false));  // refinement is ignored, as above
  }

  void Compute(OpKernelContext* ctx) override {
    const int B = B_, M = g_.M;
    Tensor *gt, *ga;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, M, 3}), &gt));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, M, 3}), &ga));
    DCX_TF_CALL(ctx, dcx_arap_bwd(g_.rest_pos.f(), g_.node_vertex.i(), g_.nn_off.i(), g_.nn_size.i(), g_.nn_idx.i(), B, M,
                                  C_, in_ptr<float>(ctx->input(0)), in_ptr<float>(ctx->input(1)), out_ptr<float>(gt),
                                  out_ptr<float>(ga), stream_of(ctx)));
  }

 private:
  GraphCharacter g_;
  int B_ = 0, C_ = 0;
};
REGISTER_KERNEL_BUILDER(Name("EmbeddedGraphArapGpuGrad").Device(DEVICE_GPU), EmbeddedGraphArapGpuGradOp);

REGISTER_OP("Pose2EmbeddedGraphCpu")
.Input("dofs: float")
.Output("d_skinned_translation: float")
.Output("d_skinned_rotation: float")
.Attr("character_file_path_eg: string = 'None'")
.Attr("graph_file_path: string = 'None'");

class Pose2EmbeddedGraphOp : public OpKernel {
 public:
  explicit Pose2EmbeddedGraphOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string cpath, gpath;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("character_file_path_eg", &cpath));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("graph_file_path", &gpath));
    OP_REQUIRES_OK(ctx, g_.load(cpath, gpath, false, false));
    OP_REQUIRES_OK(ctx, s_.load(cpath, false));
  }

  void Compute(OpKernelContext* ctx) override {
    const Tensor& dofs = ctx->input(0);
    const int B = (int)dofs.dim_size(0), M = g_.M;  // Pose2EmbeddedGraphCPUOp.cpp:101
    Tensor *t, *a;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, M, 3}), &t));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, M, 3}), &a));
    Tensor dq, jp, ja;
    OP_REQUIRES_OK(ctx, ctx->allocate_temp(DT_FLOAT, TensorShape({B, s_.Js, 8}), &dq));
    OP_REQUIRES_OK(ctx, ctx->allocate_temp(DT_FLOAT, TensorShape({B, s_.J, 3}), &jp));
    OP_REQUIRES_OK(ctx, ctx->allocate_temp(DT_FLOAT, TensorShape({B, s_.J, 3}), &ja));
    void* st = stream_of(ctx);
    DCX_TF_CALL(ctx, dcx_fk(s_.skel, B, in_ptr<float>(dofs), dq.flat<float>().data(), jp.flat<float>().data(),
                            ja.flat<float>().data(), st));
    DCX_TF_CALL(ctx, dcx_pose_to_graph(dq.flat<float>().data(), g_.joint_idx.i(), g_.joint_w.f(), g_.node_vertex.i(), B,
                                       g_.N, g_.K, s_.Js, M, out_ptr<float>(t), out_ptr<float>(a), st));
  }

 private:
  GraphCharacter g_;
  SkinnedCharacter s_;
};
REGISTER_KERNEL_BUILDER(Name("Pose2EmbeddedGraphCpu").Device(DEVICE_GPU), Pose2EmbeddedGraphOp);

// ===========================================================================
// MultiViewSilhouetteLossGpu(+Grad), CameraProjectionGpu(+Grad),
// ProjectedMeshBoundaryGpu (MultiViewSilhouetteLossGPUOp.cpp,
// CameraProjectionGPUOp.cpp, ProjectedMeshBoundaryGPUOp.cpp)
// ===========================================================================

REGISTER_OP("MultiViewSilhouetteLossGpu")
.Input("points_image_space: float")
.Input("normals_image_space: float")
.Input("is_boundary: bool")
.Input("distance_transform_image: uint8")
.Input("multi_view_crops: float")
.Output("mv_dt_residual: float")
.Output("mv_dt_residual_1: float")
.Output("dt_image_gradient: float")
.Output("closest_vertex_id: float");

class MultiViewSilhouetteLossGpuOp : public OpKernel {
 public:
  explicit MultiViewSilhouetteLossGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) {}

  void Compute(OpKernelContext* ctx) override {
    const Tensor &pts = ctx->input(0), &nrm = ctx->input(1), &isb = ctx->input(2), &dt = ctx->input(3);
    const Tensor& crops = ctx->input(4);
    // MultiViewSilhouetteLossGPUOp.cpp:136-140
    const int B = (int)pts.dim_size(0), C = (int)pts.dim_size(1), N = (int)pts.dim_size(2);
    const int U = (int)dt.dim_size(2), V = (int)dt.dim_size(3);
    Tensor *r0, *r1, *dtg, *cid;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, C, N}), &r0));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(1, TensorShape({B, C, U, V, 2}), &r1));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(2, TensorShape({B, C, N, 2}), &dtg));
    OP_REQUIRES_OK(ctx, ctx->allocate_output(3, TensorShape({B, C, U, V}), &cid));
    Tensor ws;
    OP_REQUIRES_OK(ctx, workspace(ctx, dcx_mvsil_workspace_size(B, C, N, U, V), &ws));
    DCX_TF_CALL(ctx, dcx_mvsil_fwd(in_ptr<float>(pts), in_ptr<float>(nrm), (const uint8_t*)in_ptr<bool>(isb),
                                   in_ptr<uint8_t>(dt), in_ptr<float>(crops), B, C, N, U, V, DCX_SIL_BINNED,
                                   out_ptr<float>(r0), out_ptr<float>(r1), out_ptr<float>(dtg), out_ptr<float>(cid),
                                   ws_ptr(ws), ws.NumElements(), stream_of(ctx)));
  }
};
REGISTER_KERNEL_BUILDER(Name("MultiViewSilhouetteLossGpu").Device(DEVICE_GPU), MultiViewSilhouetteLossGpuOp);

REGISTER_OP("MultiViewSilhouetteLossGpuGrad")
.Input("mv_dt_grad: float")
.Input("mv_dt_grad_1: float")
.Input("dt_image_grad: float")
.Input("closest_vertex_id_grad: float")
.Input("multi_view_crops: float")
.Output("points_image_space_grad: float");

class MultiViewSilhouetteLossGpuGradOp : public OpKernel {
 public:
  explicit MultiViewSilhouetteLossGpuGradOp(OpKernelConstruction* ctx) : OpKernel(ctx) {}

  void Compute(OpKernelContext* ctx) override {
    const Tensor &g0 = ctx->input(0), &g1 = ctx->input(1), &dtg = ctx->input(2), &cid = ctx->input(3);
    const Tensor& crops = ctx->input(4);
    // MultiViewSilhouetteLossGPUOpGrad.cpp:65-70
    const int B = (int)g0.dim_size(0), C = (int)g0.dim_size(1), N = (int)g0.dim_size(2);
    const int U = (int)cid.dim_size(2), V = (int)cid.dim_size(3);
    Tensor* gp;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, C, N, 2}), &gp));
    DCX_TF_CALL(ctx, dcx_mvsil_bwd(in_ptr<float>(g0), in_ptr<float>(g1), in_ptr<float>(dtg), in_ptr<float>(cid),
                                   in_ptr<float>(crops), B, C, N, U, V, out_ptr<float>(gp), stream_of(ctx)));
  }
};
REGISTER_KERNEL_BUILDER(Name("MultiViewSilhouetteLossGpuGrad").Device(DEVICE_GPU), MultiViewSilhouetteLossGpuGradOp);

REGISTER_OP("CameraProjectionGpu")
.Input("points_global_space: float")
.Input("vectors_global_space: float")
.Input("extrinsics: float")
.Input("intrinsics: float")
.Output("points_image_space: float")
.Attr("is_point_cam_proj: bool = false");

class CameraProjectionGpuOp : public OpKernel {
 public:
  explicit CameraProjectionGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    OP_REQUIRES_OK(ctx, ctx->GetAttr("is_point_cam_proj", &is_point_));
  }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &p = ctx->input(0), &v = ctx->input(1), &ext = ctx->input(2), &intr = ctx->input(3);
    // CameraProjectionGPUOp.cpp:106-108
    const int C = (int)intr.dim_size(1) / 9, B = (int)p.dim_size(0), N = (int)p.dim_size(1);
    Tensor* out;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, C, N, 2}), &out));
    DCX_TF_CALL(ctx, dcx_camproj_fwd(in_ptr<float>(p), is_point_ ? nullptr : in_ptr<float>(v), in_ptr<float>(ext),
                                     in_ptr<float>(intr), B, C, N, is_point_ ? 1 : 0, out_ptr<float>(out), stream_of(ctx)));
  }

 private:
  bool is_point_ = true;
};
REGISTER_KERNEL_BUILDER(Name("CameraProjectionGpu").Device(DEVICE_GPU), CameraProjectionGpuOp);

REGISTER_OP("CameraProjectionGpuGrad")
.Input("points_image_space: float")
.Input("points_global_space: float")
.Input("extrinsics: float")
.Input("intrinsics: float")
.Output("points_global_space_grad: float")
.Attr("is_point_cam_proj_grad: bool = false");

class CameraProjectionGpuGradOp : public OpKernel {
 public:
  explicit CameraProjectionGpuGradOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    OP_REQUIRES_OK(ctx, ctx->GetAttr("is_point_cam_proj_grad", &is_point_));
  }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &g = ctx->input(0), &p = ctx->input(1), &ext = ctx->input(2), &intr = ctx->input(3);
    // CameraProjectionGPUOpGrad.cpp:58-60
    const int B = (int)g.dim_size(0), C = (int)g.dim_size(1), N = (int)g.dim_size(2);
    Tensor* out;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B, N, 3}), &out));
    if (!is_point_) {  // the projection of directions has no position gradient (CameraProjectionGPU.py:60-74)
      zero_fill(ctx, out);
      return;
    }
    DCX_TF_CALL(ctx, dcx_camproj_bwd(in_ptr<float>(g), in_ptr<float>(p), in_ptr<float>(ext), in_ptr<float>(intr), B, C, N,
                                     out_ptr<float>(out), stream_of(ctx)));
  }

 private:
  bool is_point_ = true;
};
REGISTER_KERNEL_BUILDER(Name("CameraProjectionGpuGrad").Device(DEVICE_GPU), CameraProjectionGpuGradOp);

REGISTER_OP("ProjectedMeshBoundaryGpu")
.Input("points_global_space: float")
.Input("extrinsics: float")
.Input("intrinsics: float")
.Output("is_boundary: bool")
.Attr("mesh_file_path_boundary_check: string = 'None'")
.Attr("use_gap_detection: bool")
.Attr("number_batches: int")
.Attr("number_cameras: int")
.Attr("render_u: int")
.Attr("render_v: int");

class ProjectedMeshBoundaryGpuOp : public OpKernel {
 public:
  explicit ProjectedMeshBoundaryGpuOp(OpKernelConstruction* ctx) : OpKernel(ctx) {
    std::string path;
    bool gaps = false;
    OP_REQUIRES_OK(ctx, ctx->GetAttr("mesh_file_path_boundary_check", &path));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("use_gap_detection", &gaps));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("number_batches", &B_));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("number_cameras", &C_));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("render_u", &U_));
    OP_REQUIRES_OK(ctx, ctx->GetAttr("render_v", &V_));
    OP_REQUIRES(ctx, path != "None", errors::InvalidArgument("mesh_file_path_boundary_check not set!"));
    dcx_tf::HostTables t;
    OP_REQUIRES_OK(ctx, load_tables("boundary_mesh", {path}, &t));
    const auto& F = table(t, "faces");
    N_ = table(t, "num_vertices").i()[0];
    OP_REQUIRES(ctx, dcx_topo_create(F.i(), (int)F.shape[0], N_, nullptr, &topo_) == DCX_OK,
                errors::InvalidArgument(dcx_last_error()));
  }
  ~ProjectedMeshBoundaryGpuOp() override { if (topo_) dcx_topo_destroy(topo_); }

  void Compute(OpKernelContext* ctx) override {
    const Tensor &p = ctx->input(0), &ext = ctx->input(1), &intr = ctx->input(2);
    Tensor* out;
    OP_REQUIRES_OK(ctx, ctx->allocate_output(0, TensorShape({B_, C_, N_}), &out));
    Tensor ws, cov;
    OP_REQUIRES_OK(ctx, workspace(ctx, dcx_projected_mesh_boundary_workspace_size(B_, C_, N_, V_, U_), &ws));
    OP_REQUIRES_OK(ctx, ctx->allocate_temp(DT_UINT8, TensorShape({B_, C_, V_, U_}), &cov));
    DCX_TF_CALL(ctx, dcx_projected_mesh_boundary(topo_, B_, C_, V_, U_, in_ptr<float>(p), in_ptr<float>(ext),
                                                 in_ptr<float>(intr), (uint8_t*)out_ptr<bool>(out),
                                                 cov.flat<uint8_t>().data(), ws_ptr(ws), ws.NumElements(),
                                                 stream_of(ctx)));
  }

 private:
  dcx_topo* topo_ = nullptr;
  int B_ = 0, C_ = 0, U_ = 0, V_ = 0, N_ = 0;
};
REGISTER_KERNEL_BUILDER(Name("ProjectedMeshBoundaryGpu").Device(DEVICE_GPU), ProjectedMeshBoundaryGpuOp);
