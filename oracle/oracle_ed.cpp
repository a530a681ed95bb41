// ORACLE — test infrastructure only.  CPU restatement of the embedded-graph
// deformation op used as the checker for the HIP kernels.
//
//   ed_fwd  EmbeddedGraphGPUOp.cu:11-225 (nodes, vertices, normals, markers)
//           with the Euler rotation R = Rz(g) Ry(b) Rx(a) of
//           EmbeddedGraphRotationHelper.h:53-77
//   ed_bwd  EmbeddedGraphGPUOpGrad.cu:9-201: node translation / rotation
//           gradients (node-centric sums), the marker term, displacement
//           gradient; skinned-node gradients are zero as in the reference
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

}  // namespace orc

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
