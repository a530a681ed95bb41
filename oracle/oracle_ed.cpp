// ORACLE — test infrastructure only.  CPU restatement of the embedded-graph
// deformation op used as the checker for the HIP kernels.
//
//   ed_fwd  EmbeddedGraphGPUOp.cu:11-225 (nodes, vertices, normals, markers)
//           with the Euler rotation R = Rz(g) Ry(b) Rx(a) of
//           EmbeddedGraphRotationHelper.h:53-77
//   ed_bwd  EmbeddedGraphGPUOpGrad.cu:9-201: node translation / rotation
//           gradients (node-centric sums), the marker term, displacement
//           gradient; skinned-node gradients are zero as in the reference
//   arap    EmbeddedGraphArapGPUOp.cu:10-105 / ...Grad.cu:11-100: node-graph
//           as-rigid-as-possible residuals and weights; the gradient is the
//           exact adjoint (each residual's +t_j / -t_k / rotation terms),
//           which is the reference's neighbour-list formula on a symmetric
//           graph
//
// Pinning (tests/test_oracle_ed.py): identity at zero parameters, pure
// translation, single-node rigid rotation about the node, the Euler
// convention (R(a,0,0) = Rx(a) ...), marker/vertex consistency, unit normals,
// and FP64 central finite differences of every gradient.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_common.hpp"

namespace orc {

template <class T> static void euler(const T* a, T R[9]) {
  const T ca = std::cos(a[0]), cb = std::cos(a[1]), cg = std::cos(a[2]);
  const T sa = std::sin(a[0]), sb = std::sin(a[1]), sg = std::sin(a[2]);
  R[0] = cb * cg; R[1] = sa * sb * cg - ca * sg; R[2] = ca * sb * cg + sa * sg;
  R[3] = cb * sg; R[4] = sa * sb * sg + ca * cg; R[5] = ca * sb * sg - sa * cg;
  R[6] = -sb;     R[7] = sa * cb;                R[8] = ca * cb;
}

// dR/d(angle k), k = 0 (alpha), 1 (beta), 2 (gamma)
template <class T> static void euler_d(const T* a, int k, T R[9]) {
  const T ca = std::cos(a[0]), cb = std::cos(a[1]), cg = std::cos(a[2]);
  const T sa = std::sin(a[0]), sb = std::sin(a[1]), sg = std::sin(a[2]);
  if (k == 0) {
    R[0] = 0; R[1] = ca * sb * cg + sa * sg; R[2] = -sa * sb * cg + ca * sg;
    R[3] = 0; R[4] = ca * sb * sg - sa * cg; R[5] = -sa * sb * sg - ca * cg;
    R[6] = 0; R[7] = ca * cb;                R[8] = -sa * cb;
  } else if (k == 1) {
    R[0] = -sb * cg; R[1] = sa * cb * cg; R[2] = ca * cb * cg;
    R[3] = -sb * sg; R[4] = sa * cb * sg; R[5] = ca * cb * sg;
    R[6] = -cb;      R[7] = -sa * sb;     R[8] = -ca * sb;
  } else {
    R[0] = -cb * sg; R[1] = -sa * sb * sg - ca * cg; R[2] = -ca * sb * sg + sa * cg;
    R[3] = cb * cg;  R[4] = sa * sb * cg - ca * sg;  R[5] = ca * sb * cg + sa * sg;
    R[6] = 0;        R[7] = 0;                       R[8] = 0;
  }
}

template <class T> static void mv(const T R[9], const T* v, T* o) {
  for (int i = 0; i < 3; ++i) o[i] = R[3 * i] * v[0] + R[3 * i + 1] * v[1] + R[3 * i + 2] * v[2];
}
template <class T> static void mtv(const T R[9], const T* v, T* o) {
  for (int i = 0; i < 3; ++i) o[i] = R[i] * v[0] + R[3 + i] * v[1] + R[6 + i] * v[2];
}

struct EdGraph {
  int N, M, Mk;
  const int32_t* node_vertex;   // [M]
  const int32_t* v2n_off;       // [N+1]
  const int32_t* v2n_node;      // [E]
  const double* v2n_w64;        // weights (double or float)
  const float* v2n_w32;
  const int32_t* marker_node;   // [Mk]
  const int32_t* node_marker;   // [M], -1 if none
};

template <class T> static T wgt(const EdGraph& g, int e) { return g.v2n_w64 ? T(g.v2n_w64[e]) : T(g.v2n_w32[e]); }

template <class T>
void ed_fwd(const EdGraph& g, const T* base, const T* base_markers, const int32_t* faces, int F,
            const T* dT, const T* dA, const T* sT, const T* sA, const T* disp, int B, T* verts, T* nrms,
            T* markers, T* graph) {
  const int N = g.N, M = g.M;
  for (int b = 0; b < B; ++b) {
    std::vector<T> Rd((size_t)M * 9), Rs((size_t)M * 9);
    for (int n = 0; n < M; ++n) {
      euler(dA + ((size_t)b * M + n) * 3, &Rd[9 * n]);
      euler(sA + ((size_t)b * M + n) * 3, &Rs[9 * n]);
      const T* gn = base + 3 * g.node_vertex[n];
      const T* t = dT + ((size_t)b * M + n) * 3;
      T v[3] = {gn[0] + t[0], gn[1] + t[1], gn[2] + t[2]}, o[3];
      mv(&Rs[9 * n], v, o);
      const T* ts = sT + ((size_t)b * M + n) * 3;
      if (graph)
        for (int i = 0; i < 3; ++i) graph[((size_t)b * M + n) * 3 + i] = o[i] + ts[i];
    }
    T* vo = verts + (size_t)b * N * 3;
    for (int v = 0; v < N; ++v) {
      const T* p = base + 3 * v;
      T vnr[3] = {0, 0, 0};
      for (int e = g.v2n_off[v]; e < g.v2n_off[v + 1]; ++e) {
        const int n = g.v2n_node[e];
        const T w = wgt<T>(g, e);
        const T* gn = base + 3 * g.node_vertex[n];
        const T d[3] = {p[0] - gn[0], p[1] - gn[1], p[2] - gn[2]};
        T r[3];
        mv(&Rd[9 * n], d, r);
        const T* t = dT + ((size_t)b * M + n) * 3;
        for (int i = 0; i < 3; ++i) vnr[i] += w * (r[i] + gn[i] + t[i]);
      }
      const T* dp = disp + ((size_t)b * N + v) * 3;
      for (int i = 0; i < 3; ++i) vnr[i] += dp[i];
      T out[3] = {0, 0, 0};
      for (int e = g.v2n_off[v]; e < g.v2n_off[v + 1]; ++e) {
        const int s = g.v2n_node[e];
        const T w = wgt<T>(g, e);
        T r[3];
        mv(&Rs[9 * s], vnr, r);
        const T* ts = sT + ((size_t)b * M + s) * 3;
        for (int i = 0; i < 3; ++i) out[i] += w * (r[i] + ts[i]);
      }
      for (int i = 0; i < 3; ++i) vo[3 * v + i] = out[i];
    }
    if (nrms) {
      std::vector<T> acc((size_t)N * 3, T(0));
      std::vector<std::vector<int>> vf(N);
      for (int f = 0; f < F; ++f)
        for (int c = 0; c < 3; ++c) vf[faces[3 * f + c]].push_back(f);
      for (int v = 0; v < N; ++v) {
        T s[3] = {0, 0, 0};
        for (int f : vf[v]) {
          const T* a = vo + 3 * faces[3 * f];
          const T* b1 = vo + 3 * faces[3 * f + 1];
          const T* c1 = vo + 3 * faces[3 * f + 2];
          const T u[3] = {b1[0] - a[0], b1[1] - a[1], b1[2] - a[2]};
          const T w[3] = {c1[0] - a[0], c1[1] - a[1], c1[2] - a[2]};
          s[0] += u[1] * w[2] - u[2] * w[1];
          s[1] += u[2] * w[0] - u[0] * w[2];
          s[2] += u[0] * w[1] - u[1] * w[0];
        }
        const T l = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        for (int i = 0; i < 3; ++i) nrms[((size_t)b * N + v) * 3 + i] = s[i] / l;
      }
    }
    if (markers)
      for (int k = 0; k < g.Mk; ++k) {
        const int n = g.marker_node[k];
        const T* m = base_markers + 3 * k;
        const T* gn = base + 3 * g.node_vertex[n];
        const T d[3] = {m[0] - gn[0], m[1] - gn[1], m[2] - gn[2]};
        T r[3], o[3];
        mv(&Rd[9 * n], d, r);
        const T* t = dT + ((size_t)b * M + n) * 3;
        T mnr[3] = {r[0] + gn[0] + t[0], r[1] + gn[1] + t[1], r[2] + gn[2] + t[2]};
        mv(&Rs[9 * n], mnr, o);
        const T* ts = sT + ((size_t)b * M + n) * 3;
        for (int i = 0; i < 3; ++i) markers[((size_t)b * g.Mk + k) * 3 + i] = o[i] + ts[i];
      }
  }
}

template <class T>
void ed_bwd(const EdGraph& g, const T* base, const T* base_markers, const T* gV, const T* gM, const T* dA,
            const T* sA, int B, T* gT, T* gR, T* gsT, T* gsR, T* gdisp) {
  const int N = g.N, M = g.M;
  std::memset(gT, 0, sizeof(T) * (size_t)B * M * 3);
  std::memset(gR, 0, sizeof(T) * (size_t)B * M * 3);
  std::memset(gsT, 0, sizeof(T) * (size_t)B * M * 3);
  std::memset(gsR, 0, sizeof(T) * (size_t)B * M * 3);
  for (int b = 0; b < B; ++b) {
    std::vector<T> Rs((size_t)M * 9);
    for (int n = 0; n < M; ++n) euler(sA + ((size_t)b * M + n) * 3, &Rs[9 * n]);
    for (int v = 0; v < N; ++v) {
      T W[9] = {0};
      for (int e = g.v2n_off[v]; e < g.v2n_off[v + 1]; ++e)
        for (int k = 0; k < 9; ++k) W[k] += Rs[9 * g.v2n_node[e] + k] * wgt<T>(g, e);
      const T* gv = gV + ((size_t)b * N + v) * 3;
      T u[3];
      mtv(W, gv, u);  // (sum_s w_s R_s)^T g_v
      for (int i = 0; i < 3; ++i) gdisp[((size_t)b * N + v) * 3 + i] = u[i];
      const T* p = base + 3 * v;
      for (int e = g.v2n_off[v]; e < g.v2n_off[v + 1]; ++e) {
        const int n = g.v2n_node[e];
        const T w = wgt<T>(g, e);
        const T* gn = base + 3 * g.node_vertex[n];
        const T d[3] = {p[0] - gn[0], p[1] - gn[1], p[2] - gn[2]};
        for (int i = 0; i < 3; ++i) gT[((size_t)b * M + n) * 3 + i] += w * u[i];
        for (int k = 0; k < 3; ++k) {
          T Rk[9], r[3];
          euler_d(dA + ((size_t)b * M + n) * 3, k, Rk);
          mv(Rk, d, r);
          gR[((size_t)b * M + n) * 3 + k] += w * (r[0] * u[0] + r[1] * u[1] + r[2] * u[2]);
        }
      }
    }
    if (gM)
      for (int n = 0; n < M; ++n) {
        const int mk = g.node_marker ? g.node_marker[n] : -1;
        if (mk < 0) continue;
        const T* gm = gM + ((size_t)b * g.Mk + mk) * 3;
        T u[3];
        mtv(&Rs[9 * n], gm, u);
        const T* m = base_markers + 3 * mk;
        const T* gn = base + 3 * g.node_vertex[n];
        const T d[3] = {m[0] - gn[0], m[1] - gn[1], m[2] - gn[2]};
        for (int i = 0; i < 3; ++i) gT[((size_t)b * M + n) * 3 + i] += u[i];
        for (int k = 0; k < 3; ++k) {
          T Rk[9], r[3];
          euler_d(dA + ((size_t)b * M + n) * 3, k, Rk);
          mv(Rk, d, r);
          gR[((size_t)b * M + n) * 3 + k] += r[0] * u[0] + r[1] * u[1] + r[2] * u[2];
        }
      }
  }
}

// As-rigid-as-possible residuals of the node graph and their gradient
// (EmbeddedGraphArapGPUOp.cu:10-105, EmbeddedGraphArapGPUOpGrad.cu:11-100):
// r_jc = R_j (g_k - g_j) + g_j + t_j - (g_k + t_k) for the c-th neighbour k of
// node j (c < max_conn), weight (rig_j + rig_k) / 2; the backward takes dL/dr
// and returns dL/dt (each residual's +t_j and -t_k terms) and dL/dA (the
// rotation derivative applied to g_k - g_j).
template <class T>
void arap_fwd(const T* base, const int32_t* node_vertex, const int32_t* off, const int32_t* size, const int32_t* idx,
              const T* rig, int B, int M, int C, const T* dT, const T* dA, T* res, T* cw) {
  std::memset(res, 0, sizeof(T) * (size_t)B * M * C * 3);
  std::memset(cw, 0, sizeof(T) * (size_t)B * M * C);
  for (int b = 0; b < B; ++b)
    for (int j = 0; j < M; ++j) {
      T R[9];
      euler(dA + ((size_t)b * M + j) * 3, R);
      const T* gj = base + 3 * (size_t)node_vertex[j];
      const T* tj = dT + ((size_t)b * M + j) * 3;
      for (int c = 0; c < size[j] && c < C; ++c) {
        const int k = idx[off[j] + c];
        const T* gk = base + 3 * (size_t)node_vertex[k];
        const T* tk = dT + ((size_t)b * M + k) * 3;
        const T d[3] = {gk[0] - gj[0], gk[1] - gj[1], gk[2] - gj[2]};
        T* r = res + (((size_t)b * M + j) * C + c) * 3;
        for (int a = 0; a < 3; ++a)
          r[a] = R[3 * a] * d[0] + R[3 * a + 1] * d[1] + R[3 * a + 2] * d[2] + gj[a] + tj[a] - (gk[a] + tk[a]);
        cw[((size_t)b * M + j) * C + c] = (rig[j] + rig[k]) / T(2);
      }
    }
}

template <class T>
void arap_bwd(const T* base, const int32_t* node_vertex, const int32_t* off, const int32_t* size, const int32_t* idx,
              int B, int M, int C, const T* gres, const T* dA, T* gT, T* gA) {
  std::memset(gT, 0, sizeof(T) * (size_t)B * M * 3);
  std::memset(gA, 0, sizeof(T) * (size_t)B * M * 3);
  for (int b = 0; b < B; ++b)
    for (int j = 0; j < M; ++j) {
      T Rk[3][9];
      for (int k = 0; k < 3; ++k) euler_d(dA + ((size_t)b * M + j) * 3, k, Rk[k]);
      const T* gj = base + 3 * (size_t)node_vertex[j];
      for (int c = 0; c < size[j] && c < C; ++c) {
        const int k = idx[off[j] + c];
        const T* gk = base + 3 * (size_t)node_vertex[k];
        const T d[3] = {gk[0] - gj[0], gk[1] - gj[1], gk[2] - gj[2]};
        const T* G = gres + (((size_t)b * M + j) * C + c) * 3;
        for (int q = 0; q < 3; ++q)
          for (int a = 0; a < 3; ++a)
            gA[((size_t)b * M + j) * 3 + q] += (Rk[q][3 * a] * d[0] + Rk[q][3 * a + 1] * d[1] + Rk[q][3 * a + 2] * d[2]) * G[a];
        for (int a = 0; a < 3; ++a) {
          gT[((size_t)b * M + j) * 3 + a] += G[a];   // +t_j
          gT[((size_t)b * M + k) * 3 + a] -= G[a];   // -t_k
        }
      }
    }
}

}  // namespace orc

extern "C" void orc_arap_fwd(const double* base, const int32_t* node_vertex, const int32_t* off, const int32_t* size,
                             const int32_t* idx, const double* rig, int B, int M, int C, const double* dT,
                             const double* dA, double* res, double* cw) {
  orc::arap_fwd<double>(base, node_vertex, off, size, idx, rig, B, M, C, dT, dA, res, cw);
}
extern "C" void orc_arap_bwd(const double* base, const int32_t* node_vertex, const int32_t* off, const int32_t* size,
                             const int32_t* idx, int B, int M, int C, const double* gres, const double* dA, double* gT,
                             double* gA) {
  orc::arap_bwd<double>(base, node_vertex, off, size, idx, B, M, C, gres, dA, gT, gA);
}

#define ORC_ED_EXPORT(T, SUF)                                                                                 \
  extern "C" void orc_ed_fwd_##SUF(const int* dims, const int32_t* node_vertex, const int32_t* off,           \
                                   const int32_t* nodes, const T* w, const int32_t* marker_node,              \
                                   const T* base, const T* base_markers, const int32_t* faces, int F,        \
                                   const T* dT, const T* dA, const T* sT, const T* sA, const T* disp,        \
                                   T* verts, T* nrms, T* markers, T* graph) {                                 \
    orc::EdGraph g{dims[0], dims[1], dims[2], node_vertex, off, nodes, nullptr, nullptr, marker_node, nullptr}; \
    if (sizeof(T) == 8) g.v2n_w64 = (const double*)w; else g.v2n_w32 = (const float*)w;                      \
    orc::ed_fwd<T>(g, base, base_markers, faces, F, dT, dA, sT, sA, disp, dims[3], verts, nrms, markers,     \
                   graph);                                                                                    \
  }                                                                                                            \
  extern "C" void orc_ed_bwd_##SUF(const int* dims, const int32_t* node_vertex, const int32_t* off,           \
                                   const int32_t* nodes, const T* w, const int32_t* node_marker,              \
                                   const T* base, const T* base_markers, const T* gV, const T* gM,           \
                                   const T* dA, const T* sA, T* gT, T* gR, T* gsT, T* gsR, T* gdisp) {       \
    orc::EdGraph g{dims[0], dims[1], dims[2], node_vertex, off, nodes, nullptr, nullptr, nullptr, node_marker}; \
    if (sizeof(T) == 8) g.v2n_w64 = (const double*)w; else g.v2n_w32 = (const float*)w;                      \
    orc::ed_bwd<T>(g, base, base_markers, gV, gM, dA, sA, dims[3], gT, gR, gsT, gsR, gdisp);                 \
  }

ORC_ED_EXPORT(double, f64)
ORC_ED_EXPORT(float, f32)
