// Embedded-graph deformation (EmbeddedGraphGpu / EmbeddedGraphGpuGrad) on gfx950.
//
// Forward: a node pass evaluates each node's two Euler rotations once
// (sin/cos per node, not per (vertex, node) as in EmbeddedGraphGPUOp.cu:50-125)
// into a 27-float node table; the vertex pass then gathers its K nodes' rows.
// Backward: a vertex pass forms u_v = (sum_s w_s R_s)^T g_v, which is exactly
// the displacement gradient, and a node pass (one wavefront per node) gathers
// w * u_v and w * (dR/d(angle) (v - g))·u_v over the node's vertices from the
// node -> vertex adjacency with a deterministic wave reduction — no atomics.
#include <vector>

#include "raster_internal.hpp"

struct dcx_edgraph {
  int N = 0, M = 0, Mk = 0;
  float* base = nullptr;         // [N,3]
  int* node_vertex = nullptr;    // [M]
  int* v2n_off = nullptr;        // [N+1]
  int* v2n_node = nullptr;       // [E]
  float* v2n_w = nullptr;        // [E]
  int* n2v_off = nullptr;        // [M+1]
  int* n2v_vtx = nullptr;        // [E]
  float* n2v_w = nullptr;        // [E]
  float* base_markers = nullptr; // [Mk,3]
  int* marker_node = nullptr;    // [Mk]
  int* node_marker = nullptr;    // [M]
  const dcx_topo* topo = nullptr;
};

namespace dcx {
namespace {

constexpr int kB = 256;
constexpr int kFwdRow = 27;  // Rd(9) Rs(9) g(3) td(3) ts(3)
constexpr int kBwdRow = 48;  // Rs(9) dRa(9) dRb(9) dRg(9) g(3), padded

__global__ __launch_bounds__(kB) void ed_node_fwd_kernel(const dcx_edgraph g, int B, const float* __restrict__ dT,
                                                          const float* __restrict__ dA,
                                                          const float* __restrict__ sT,
                                                          const float* __restrict__ sA, float* __restrict__ table,
                                                          float* __restrict__ graph_out, float* __restrict__ dA_out,
                                                          float* __restrict__ sA_out) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= B * g.M) return;
  const int n = i % g.M;
  const f3 a = ld3(dA + 3 * (size_t)i), s = ld3(sA + 3 * (size_t)i);
  const m3 Rd = euler_mat(euler_sc(a)), Rs = euler_mat(euler_sc(s));
  const f3 gn = ld3(g.base + 3 * g.node_vertex[n]);
  const f3 td = ld3(dT + 3 * (size_t)i), ts = ld3(sT + 3 * (size_t)i);
  float* row = table + (size_t)i * kFwdRow;
  for (int k = 0; k < 9; ++k) { row[k] = Rd.m[k]; row[9 + k] = Rs.m[k]; }
  row[18] = gn.x; row[19] = gn.y; row[20] = gn.z;
  row[21] = td.x; row[22] = td.y; row[23] = td.z;
  row[24] = ts.x; row[25] = ts.y; row[26] = ts.z;
  if (graph_out) {
    const f3 o = mul(Rs, gn + td) + ts;
    graph_out[3 * (size_t)i] = o.x; graph_out[3 * (size_t)i + 1] = o.y; graph_out[3 * (size_t)i + 2] = o.z;
  }
  if (dA_out) { dA_out[3 * (size_t)i] = a.x; dA_out[3 * (size_t)i + 1] = a.y; dA_out[3 * (size_t)i + 2] = a.z; }
  if (sA_out) { sA_out[3 * (size_t)i] = s.x; sA_out[3 * (size_t)i + 1] = s.y; sA_out[3 * (size_t)i + 2] = s.z; }
}

__device__ __forceinline__ m3 row_mat(const float* r) {
  m3 R;
  for (int k = 0; k < 9; ++k) R.m[k] = r[k];
  return R;
}

__global__ __launch_bounds__(kB) void ed_vertex_fwd_kernel(const dcx_edgraph g, int B, const float* __restrict__ table,
                                                            const float* __restrict__ disp, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * g.N) return;
  const int b = (int)(i / g.N), v = (int)(i - (long long)b * g.N);
  const float* tb = table + (size_t)b * g.M * kFwdRow;
  const f3 p = ld3(g.base + 3 * v);
  f3 vnr = mk3(0.f, 0.f, 0.f);
  const int e0 = g.v2n_off[v], e1 = g.v2n_off[v + 1];
  for (int e = e0; e < e1; ++e) {
    const float* r = tb + (size_t)g.v2n_node[e] * kFwdRow;
    const float w = g.v2n_w[e];
    const f3 gn = ld3(r + 18);
    vnr += w * (mul(row_mat(r), p - gn) + gn + ld3(r + 21));
  }
  vnr += ld3(disp + 3 * (size_t)i);
  f3 o = mk3(0.f, 0.f, 0.f);
  for (int e = e0; e < e1; ++e) {
    const float* r = tb + (size_t)g.v2n_node[e] * kFwdRow;
    o += g.v2n_w[e] * (mul(row_mat(r + 9), vnr) + ld3(r + 24));
  }
  out[3 * i] = o.x; out[3 * i + 1] = o.y; out[3 * i + 2] = o.z;
}

// Normalised sum of the deformed mesh's face normals around each vertex
// (EmbeddedGraphGPUOp.cu:180-225), faces in increasing id order.
__global__ __launch_bounds__(kB) void ed_normal_kernel(const int* __restrict__ faces, const int* __restrict__ off,
                                                        const int* __restrict__ lst, int N, int B,
                                                        const float* __restrict__ verts, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * N) return;
  const int b = (int)(i / N), v = (int)(i - (long long)b * N);
  const float* vb = verts + (size_t)b * N * 3;
  f3 s = mk3(0.f, 0.f, 0.f);
  for (int k = off[v]; k < off[v + 1]; ++k) {
    const int f = lst[k];
    const f3 a = ld3(vb + 3 * faces[3 * f]), c1 = ld3(vb + 3 * faces[3 * f + 1]), c2 = ld3(vb + 3 * faces[3 * f + 2]);
    s += cross(c1 - a, c2 - a);
  }
  const float l = sqrtf(dot(s, s));
  out[3 * i] = s.x / l; out[3 * i + 1] = s.y / l; out[3 * i + 2] = s.z / l;
}

__global__ __launch_bounds__(kB) void ed_marker_kernel(const dcx_edgraph g, int B, const float* __restrict__ table,
                                                        float* __restrict__ out) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= B * g.Mk) return;
  const int b = i / g.Mk, k = i % g.Mk;
  const float* r = table + ((size_t)b * g.M + g.marker_node[k]) * kFwdRow;
  const f3 gn = ld3(r + 18);
  const f3 mnr = mul(row_mat(r), ld3(g.base_markers + 3 * k) - gn) + gn + ld3(r + 21);
  const f3 o = mul(row_mat(r + 9), mnr) + ld3(r + 24);
  out[3 * (size_t)i] = o.x; out[3 * (size_t)i + 1] = o.y; out[3 * (size_t)i + 2] = o.z;
}

__global__ __launch_bounds__(kB) void ed_node_bwd_table_kernel(const dcx_edgraph g, int B, const float* __restrict__ dA,
                                                                const float* __restrict__ sA, float* __restrict__ table) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= B * g.M) return;
  const EulerSC sd = euler_sc(ld3(dA + 3 * (size_t)i));
  const m3 Rs = euler_mat(euler_sc(ld3(sA + 3 * (size_t)i)));
  const m3 da = euler_dA(sd), db = euler_dB(sd), dg = euler_dG(sd);
  float* row = table + (size_t)i * kBwdRow;
  for (int k = 0; k < 9; ++k) { row[k] = Rs.m[k]; row[9 + k] = da.m[k]; row[18 + k] = db.m[k]; row[27 + k] = dg.m[k]; }
  const f3 gn = ld3(g.base + 3 * g.node_vertex[i % g.M]);
  row[36] = gn.x; row[37] = gn.y; row[38] = gn.z;
}

// u_v = (sum_s w_s Rs_s)^T g_v  == displacement gradient (EmbeddedGraphGPUOpGrad.cu:149-201).
__global__ __launch_bounds__(kB) void ed_vertex_bwd_kernel(const dcx_edgraph g, int B, const float* __restrict__ table,
                                                            const float* __restrict__ gV, float* __restrict__ u_out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * g.N) return;
  const int b = (int)(i / g.N), v = (int)(i - (long long)b * g.N);
  const float* tb = table + (size_t)b * g.M * kBwdRow;
  m3 W = m3_zero();
  for (int e = g.v2n_off[v]; e < g.v2n_off[v + 1]; ++e) {
    const float* r = tb + (size_t)g.v2n_node[e] * kBwdRow;
    const float w = g.v2n_w[e];
    for (int k = 0; k < 9; ++k) W.m[k] += r[k] * w;
  }
  const f3 u = mulT(W, ld3(gV + 3 * i));
  u_out[3 * i] = u.x; u_out[3 * i + 1] = u.y; u_out[3 * i + 2] = u.z;
}

// One wavefront per (batch, node): gathers over the node's vertices.
__global__ __launch_bounds__(kB) void ed_node_bwd_kernel(const dcx_edgraph g, int B, const float* __restrict__ table,
                                                          const float* __restrict__ u, const float* __restrict__ gM,
                                                          float* __restrict__ gT, float* __restrict__ gR,
                                                          float* __restrict__ gsT, float* __restrict__ gsR) {
  const int wave = (blockIdx.x * kB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= B * g.M) return;
  const int b = wave / g.M, n = wave % g.M;
  const float* row = table + (size_t)wave * kBwdRow;
  const m3 da = row_mat(row + 9), db = row_mat(row + 18), dgm = row_mat(row + 27);
  const f3 gn = ld3(row + 36);
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const float* ub = u + (size_t)b * g.N * 3;
  for (int e = g.n2v_off[n] + lane; e < g.n2v_off[n + 1]; e += 64) {
    const int v = g.n2v_vtx[e];
    const float w = g.n2v_w[e];
    const f3 uv = ld3(ub + 3 * v);
    const f3 d = ld3(g.base + 3 * v) - gn;
    acc[0] += w * uv.x; acc[1] += w * uv.y; acc[2] += w * uv.z;
    acc[3] += w * dot(mul(da, d), uv);
    acc[4] += w * dot(mul(db, d), uv);
    acc[5] += w * dot(mul(dgm, d), uv);
  }
  for (int k = 0; k < 6; ++k) acc[k] = wave_sum(acc[k]);
  if (lane == 0) {
    const int mk = g.node_marker ? g.node_marker[n] : -1;
    if (gM && mk >= 0) {
      const f3 um = mulT(row_mat(row), ld3(gM + ((size_t)b * g.Mk + mk) * 3));
      const f3 d = ld3(g.base_markers + 3 * mk) - gn;
      acc[0] += um.x; acc[1] += um.y; acc[2] += um.z;
      acc[3] += dot(mul(da, d), um);
      acc[4] += dot(mul(db, d), um);
      acc[5] += dot(mul(dgm, d), um);
    }
    const size_t o = (size_t)wave * 3;
    gT[o] = acc[0]; gT[o + 1] = acc[1]; gT[o + 2] = acc[2];
    gR[o] = acc[3]; gR[o + 1] = acc[4]; gR[o + 2] = acc[5];
    if (gsT) { gsT[o] = 0.f; gsT[o + 1] = 0.f; gsT[o + 2] = 0.f; }
    if (gsR) { gsR[o] = 0.f; gsR[o + 1] = 0.f; gsR[o + 2] = 0.f; }
  }
}

template <class T>
hipError_t upload(T** dst, const T* src, size_t n) {
  hipError_t e = hipMalloc(dst, sizeof(T) * (n ? n : 1));
  if (e != hipSuccess || n == 0) return e;
  return hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice);
}

void free_graph(dcx_edgraph* g) {
  hipFree(g->base); hipFree(g->node_vertex); hipFree(g->v2n_off); hipFree(g->v2n_node); hipFree(g->v2n_w);
  hipFree(g->n2v_off); hipFree(g->n2v_vtx); hipFree(g->n2v_w); hipFree(g->base_markers);
  hipFree(g->marker_node); hipFree(g->node_marker);
  delete g;
}

// As-rigid-as-possible residuals of the node graph (EmbeddedGraphArapGPUOp.cu:
// 10-105): one lane per (pose, node j, connection slot c).  For c below the
// node's neighbour count, with k its c-th neighbour,
//   r_jc = R_j (g_k - g_j) + g_j + t_j - (g_k + t_k),  w_jc = (rig_j + rig_k) / 2,
// and zeros past the neighbour count (neighbours past max_conn are dropped).
__global__ __launch_bounds__(kB) void arap_fwd_kernel(const float* __restrict__ base, const int* __restrict__ node_vertex,
                                                      const int* __restrict__ nn_off, const int* __restrict__ nn_size,
                                                      const int* __restrict__ nn_idx, const float* __restrict__ rig,
                                                      int B, int M, int C, const float* __restrict__ dT,
                                                      const float* __restrict__ dA, float* __restrict__ res,
                                                      float* __restrict__ cw) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * M * C) return;
  const int c = (int)(i % C);
  const long long bj = i / C;
  const int j = (int)(bj % M), b = (int)(bj / M);
  f3 r = mk3(0.f, 0.f, 0.f);
  float w = 0.f;
  if (c < nn_size[j]) {
    const int k = nn_idx[nn_off[j] + c];
    const m3 R = euler_mat(euler_sc(ld3(dA + ((size_t)b * M + j) * 3)));
    const f3 gj = ld3(base + 3 * (size_t)node_vertex[j]), gk = ld3(base + 3 * (size_t)node_vertex[k]);
    const f3 tj = ld3(dT + ((size_t)b * M + j) * 3), tk = ld3(dT + ((size_t)b * M + k) * 3);
    r = mul(R, gk - gj) + gj + tj - (gk + tk);
    w = (rig[j] + rig[k]) / 2.f;
  }
  res[3 * i] = r.x; res[3 * i + 1] = r.y; res[3 * i + 2] = r.z;
  cw[i] = w;
}

// Its gradient (EmbeddedGraphArapGPUOpGrad.cu:11-100), one lane per (pose,
// node i): dL/dA_i = sum over i's slots c of (dR/dA (g_k - g_i))^T G_ic and
// dL/dt_i = sum over i's neighbours k of (G_ik - G_ki), G_ki read from the
// slot of k's list that holds i (the last such slot; zero if k does not list
// i) -- the exact adjoint for a symmetric node graph.
__global__ __launch_bounds__(kB) void arap_bwd_kernel(const float* __restrict__ base, const int* __restrict__ node_vertex,
                                                      const int* __restrict__ nn_off, const int* __restrict__ nn_size,
                                                      const int* __restrict__ nn_idx, int B, int M, int C,
                                                      const float* __restrict__ g_res, const float* __restrict__ dA,
                                                      float* __restrict__ g_t, float* __restrict__ g_a) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * M) return;
  const int n = (int)(i % M), b = (int)(i / M);
  const EulerSC sc = euler_sc(ld3(dA + 3 * (size_t)i));
  const m3 Ra = euler_dA(sc), Rb = euler_dB(sc), Rg = euler_dG(sc);
  const f3 gi = ld3(base + 3 * (size_t)node_vertex[n]);
  const float* G = g_res + (size_t)b * M * C * 3;
  f3 gt = mk3(0.f, 0.f, 0.f), ga = mk3(0.f, 0.f, 0.f);
  const int cn = min(nn_size[n], C);  // slots past max_conn hold no residual
  for (int c = 0; c < cn; ++c) {
    const int k = nn_idx[nn_off[n] + c];
    const f3 Gik = ld3(G + ((size_t)n * C + c) * 3);
    const f3 d = ld3(base + 3 * (size_t)node_vertex[k]) - gi;
    ga = ga + mk3(dot(mul(Ra, d), Gik), dot(mul(Rb, d), Gik), dot(mul(Rg, d), Gik));
    f3 Gki = mk3(0.f, 0.f, 0.f);
    const int ck = min(nn_size[k], C);
    for (int c2 = 0; c2 < ck; ++c2)
      if (nn_idx[nn_off[k] + c2] == n) Gki = ld3(G + ((size_t)k * C + c2) * 3);
    gt = gt + (Gik - Gki);
  }
  g_t[3 * i] = gt.x; g_t[3 * i + 1] = gt.y; g_t[3 * i + 2] = gt.z;
  g_a[3 * i] = ga.x; g_a[3 * i + 1] = ga.y; g_a[3 * i + 2] = ga.z;
}

__global__ __launch_bounds__(kB) void copy_f_kernel(const float* __restrict__ src, float* __restrict__ dst, long long n) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" dcx_status dcx_edgraph_create(const float* base_pos, int N, int M, const int32_t* node_vertex,
                                         const int32_t* v2n_offsets, const int32_t* v2n_nodes,
                                         const float* v2n_weights, int num_markers, const float* base_markers,
                                         const int32_t* marker_node, const dcx_topo* topo, dcx_edgraph** out) {
  DCX_REQUIRE(out && base_pos && node_vertex && v2n_offsets && v2n_nodes && v2n_weights,
              "dcx_edgraph_create: null input");
  DCX_REQUIRE(N > 0 && M > 0 && num_markers >= 0, "dcx_edgraph_create: bad sizes N=%d M=%d", N, M);
  DCX_REQUIRE(num_markers == 0 || (base_markers && marker_node), "dcx_edgraph_create: markers missing");
  DCX_REQUIRE(!topo || topo->N == N, "dcx_edgraph_create: topology has %d vertices, graph %d", topo ? topo->N : 0, N);
  DCX_REQUIRE(v2n_offsets[0] == 0, "dcx_edgraph_create: offsets must start at 0");
  const int E = v2n_offsets[N];
  for (int v = 0; v < N; ++v)
    DCX_REQUIRE(v2n_offsets[v + 1] >= v2n_offsets[v], "dcx_edgraph_create: offsets not monotone at %d", v);
  for (int e = 0; e < E; ++e)
    DCX_REQUIRE(v2n_nodes[e] >= 0 && v2n_nodes[e] < M, "dcx_edgraph_create: node index %d out of range", v2n_nodes[e]);
  for (int n = 0; n < M; ++n)
    DCX_REQUIRE(node_vertex[n] >= 0 && node_vertex[n] < N, "dcx_edgraph_create: node vertex %d out of range", node_vertex[n]);
  std::vector<int> node_marker(M, -1);
  for (int k = 0; k < num_markers; ++k) {
    DCX_REQUIRE(marker_node[k] >= 0 && marker_node[k] < M, "dcx_edgraph_create: marker node out of range");
    node_marker[marker_node[k]] = k;  // the last marker of a node wins, as in the reference's map
  }
  // node -> vertex adjacency, vertices in increasing order per node
  std::vector<int> n2v_off(M + 1, 0), n2v_vtx(E);
  std::vector<float> n2v_w(E);
  for (int e = 0; e < E; ++e) n2v_off[v2n_nodes[e] + 1]++;
  for (int n = 0; n < M; ++n) n2v_off[n + 1] += n2v_off[n];
  std::vector<int> fill(n2v_off.begin(), n2v_off.end() - 1);
  for (int v = 0; v < N; ++v)
    for (int e = v2n_offsets[v]; e < v2n_offsets[v + 1]; ++e) {
      const int k = fill[v2n_nodes[e]]++;
      n2v_vtx[k] = v;
      n2v_w[k] = v2n_weights[e];
    }
  dcx_edgraph* g = new dcx_edgraph();
  g->N = N; g->M = M; g->Mk = num_markers; g->topo = topo;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = upload(&g->base, base_pos, (size_t)3 * N);
  if (e == hipSuccess) e = upload(&g->node_vertex, node_vertex, (size_t)M);
  if (e == hipSuccess) e = upload(&g->v2n_off, v2n_offsets, (size_t)N + 1);
  if (e == hipSuccess) e = upload(&g->v2n_node, v2n_nodes, (size_t)E);
  if (e == hipSuccess) e = upload(&g->v2n_w, v2n_weights, (size_t)E);
  if (e == hipSuccess) e = upload(&g->n2v_off, n2v_off.data(), (size_t)M + 1);
  if (e == hipSuccess) e = upload(&g->n2v_vtx, n2v_vtx.data(), (size_t)E);
  if (e == hipSuccess) e = upload(&g->n2v_w, n2v_w.data(), (size_t)E);
  if (e == hipSuccess) e = upload(&g->node_marker, node_marker.data(), (size_t)M);
  if (e == hipSuccess && num_markers) e = upload(&g->base_markers, base_markers, (size_t)3 * num_markers);
  if (e == hipSuccess && num_markers) e = upload(&g->marker_node, marker_node, (size_t)num_markers);
  if (e != hipSuccess) {
    free_graph(g);
    return set_error(DCX_ERR_OOM, "dcx_edgraph_create: %s", hipGetErrorString(e));
  }
  *out = g;
  return DCX_OK;
}

extern "C" dcx_status dcx_edgraph_destroy(dcx_edgraph* g) {
  if (g) free_graph(g);
  return DCX_OK;
}

extern "C" size_t dcx_ed_workspace_size(const dcx_edgraph* g, int B) {
  if (!g || B <= 0) return 0;
  return sizeof(float) * (size_t)B * g->M * kBwdRow;
}

extern "C" dcx_status dcx_ed_fwd(const dcx_edgraph* g, int B, const float* delta_t, const float* delta_a,
                                 const float* skinned_t, const float* skinned_a, const float* displacements,
                                 float* deformed_vertices, float* deformed_normals, float* deformed_markers,
                                 float* deformed_graph, float* delta_a_out, float* skinned_a_out, void* workspace,
                                 size_t ws_bytes, void* stream) {
  DCX_REQUIRE(g && B > 0, "dcx_ed_fwd: null graph or B=%d", B);
  DCX_REQUIRE(delta_t && delta_a && skinned_t && skinned_a && displacements && deformed_vertices,
              "dcx_ed_fwd: null input");
  DCX_REQUIRE(!deformed_normals || g->topo, "dcx_ed_fwd: normals need a topology");
  DCX_REQUIRE(!deformed_markers || g->Mk > 0, "dcx_ed_fwd: graph has no markers");
  DCX_REQUIRE(workspace && ws_bytes >= dcx_ed_workspace_size(g, B), "dcx_ed_fwd: workspace too small");
  hipStream_t s = as_stream(stream);
  float* table = (float*)workspace;
  hipLaunchKernelGGL(ed_node_fwd_kernel, dim3(ceil_div((long long)B * g->M, kB)), dim3(kB), 0, s, *g, B, delta_t,
                     delta_a, skinned_t, skinned_a, table, deformed_graph, delta_a_out, skinned_a_out);
  hipLaunchKernelGGL(ed_vertex_fwd_kernel, dim3(ceil_div((long long)B * g->N, kB)), dim3(kB), 0, s, *g, B, table,
                     displacements, deformed_vertices);
  if (deformed_normals)
    hipLaunchKernelGGL(ed_normal_kernel, dim3(ceil_div((long long)B * g->N, kB)), dim3(kB), 0, s, g->topo->faces,
                       g->topo->vf_off, g->topo->vf_list, g->N, B, deformed_vertices, deformed_normals);
  if (deformed_markers)
    hipLaunchKernelGGL(ed_marker_kernel, dim3(ceil_div((long long)B * g->Mk, kB)), dim3(kB), 0, s, *g, B, table,
                       deformed_markers);
  return check_launch("dcx_ed_fwd");
}

extern "C" dcx_status dcx_ed_bwd(const dcx_edgraph* g, int B, const float* g_vertices, const float* g_markers,
                                 const float* delta_a, const float* skinned_a, float* g_delta_t, float* g_delta_a,
                                 float* g_skinned_t, float* g_skinned_a, float* g_displacements, void* workspace,
                                 size_t ws_bytes, void* stream) {
  DCX_REQUIRE(g && B > 0, "dcx_ed_bwd: null graph or B=%d", B);
  DCX_REQUIRE(g_vertices && delta_a && skinned_a && g_delta_t && g_delta_a && g_displacements,
              "dcx_ed_bwd: null input/output");
  DCX_REQUIRE(workspace && ws_bytes >= dcx_ed_workspace_size(g, B), "dcx_ed_bwd: workspace too small");
  hipStream_t s = as_stream(stream);
  float* table = (float*)workspace;
  hipLaunchKernelGGL(ed_node_bwd_table_kernel, dim3(ceil_div((long long)B * g->M, kB)), dim3(kB), 0, s, *g, B,
                     delta_a, skinned_a, table);
  hipLaunchKernelGGL(ed_vertex_bwd_kernel, dim3(ceil_div((long long)B * g->N, kB)), dim3(kB), 0, s, *g, B, table,
                     g_vertices, g_displacements);
  hipLaunchKernelGGL(ed_node_bwd_kernel, dim3(ceil_div((long long)B * g->M * 64, kB)), dim3(kB), 0, s, *g, B, table,
                     g_displacements, g->Mk ? g_markers : nullptr, g_delta_t, g_delta_a, g_skinned_t, g_skinned_a);
  return check_launch("dcx_ed_bwd");
}

extern "C" dcx_status dcx_arap_fwd(const float* base_pos, const int32_t* node_vertex, const int32_t* nn_offsets,
                                   const int32_t* nn_sizes, const int32_t* nn_indices, const float* rigidity, int B,
                                   int M, int max_conn, const float* delta_t, const float* delta_a, float* residual,
                                   float* connection_weights, float* rotation_out, void* stream) {
  DCX_REQUIRE(base_pos && node_vertex && nn_offsets && nn_sizes && nn_indices && rigidity && delta_t && delta_a &&
                  residual && connection_weights,
              "dcx_arap_fwd: null argument");
  DCX_REQUIRE(B > 0 && M > 0 && max_conn > 0, "dcx_arap_fwd: bad sizes");
  hipStream_t s = as_stream(stream);
  hipLaunchKernelGGL(arap_fwd_kernel, dim3(ceil_div((long long)B * M * max_conn, kB)), dim3(kB), 0, s, base_pos,
                     node_vertex, nn_offsets, nn_sizes, nn_indices, rigidity, B, M, max_conn, delta_t, delta_a,
                     residual, connection_weights);
  if (rotation_out)
    hipLaunchKernelGGL(copy_f_kernel, dim3(ceil_div(3ll * B * M, kB)), dim3(kB), 0, s, delta_a, rotation_out,
                       3ll * B * M);
  return check_launch("dcx_arap_fwd");
}

extern "C" dcx_status dcx_arap_bwd(const float* base_pos, const int32_t* node_vertex, const int32_t* nn_offsets,
                                   const int32_t* nn_sizes, const int32_t* nn_indices, int B, int M, int max_conn,
                                   const float* residual_grad, const float* delta_a, float* delta_t_grad,
                                   float* delta_a_grad, void* stream) {
  DCX_REQUIRE(base_pos && node_vertex && nn_offsets && nn_sizes && nn_indices && residual_grad && delta_a &&
                  delta_t_grad && delta_a_grad,
              "dcx_arap_bwd: null argument");
  DCX_REQUIRE(B > 0 && M > 0 && max_conn > 0, "dcx_arap_bwd: bad sizes");
  hipLaunchKernelGGL(arap_bwd_kernel, dim3(ceil_div((long long)B * M, kB)), dim3(kB), 0, as_stream(stream), base_pos,
                     node_vertex, nn_offsets, nn_sizes, nn_indices, B, M, max_conn, residual_grad, delta_a,
                     delta_t_grad, delta_a_grad);
  return check_launch("dcx_arap_bwd");
}
