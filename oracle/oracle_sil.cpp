// ORACLE — test infrastructure only.  CPU restatement of the multi-view
// silhouette loss op and its feeders, used as the checker for the HIP kernels.
//
//   mvsil_fwd   MultiViewSilhouetteLossGPUOp.cu:11-240: model->data residual
//               per (batch, camera, point) from the uint8 DT image (remap of
//               values >= 127, 3x3 Sobel / 8, foreground sign flip, masks) and
//               data->model residual per DT pixel (boundary test: value 127
//               with >= 4 neighbours of 127 in the 5x5 window; brute-force
//               nearest boundary point within 30*scale, strict <, so the
//               lowest index wins ties).  Points whose 3x3 window leaves the
//               DT image are treated as out of view (the reference reads
//               outside the image there).
//   mvsil_bwd   MultiViewSilhouetteLossGPUOpGrad.cu:11-72
//   camproj     CameraProjectionGPUOp.cu:12-86 (point and vector modes),
//               CameraProjectionGPUOpGrad.cu:12-79 (point gradient)
//   boundary    ProjectedMeshBoundaryGpu semantics on a given coverage
//               bitmap: a vertex is on the boundary when its pixel's 3x3
//               window lies in the image and holds an uncovered pixel
//               (Skeletool/Mesh/CUDABasedRasterization.cu:168-200)
//   coverage    that op's own coverage bitmap (same file, :64-140): vertices
//               projected as getCamSpacePoint / projectPointFloat3 do (z
//               clamped to 1e-5 when not in front), per face the pixel box
//               [min - 0.5, max + 0.5] clamped to the image and truncated to
//               int, a pixel centre covered when its 2D barycentrics
//               (uv2barycentric, RendererUtil-style float formulas) all lie in
//               [-0.001, 1.001].  No depth test: the bitmap is the union of
//               the faces' pixels.
//
// Pinning (tests/test_oracle_sil.py): hand-built DT images (constant, ramps,
// sign flips, boundary bands), crop mapping, brute-force nearest search vs a
// numpy argmin with the tie rule, and FD of the projection gradient.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "oracle_common.hpp"

namespace orc {

static inline float remap(int v) { return v >= 127 ? (float)((v - 127) * 2) : (float)v; }

void mvsil_fwd(const float* pts, const float* nrm, const uint8_t* isb, const uint8_t* dt, const float* crops,
               int B, int C, int N, int U, int V, float* res, float* res1, float* dtg, float* cid) {
  for (int b = 0; b < B; ++b)
    for (int c = 0; c < C; ++c) {
      const float* cr = crops + ((size_t)b * C + c) * 7;
      const float bx = cr[0], by = cr[1], sc = cr[2], mnx = cr[3], mny = cr[4], mxx = cr[5], mxy = cr[6];
      const uint8_t* img = dt + ((size_t)b * C + c) * U * V;
      for (int v = 0; v < N; ++v) {
        const size_t pi = ((size_t)b * C + c) * N + v;
        const float pu = pts[2 * pi], pv = pts[2 * pi + 1];
        const bool in_view = pu - 2 > mnx && pu + 2 < mxx && pv - 2 > mny && pv + 2 < mxy;
        const float fx = (pu - mnx) / sc + bx, fy = (pv - mny) / sc + by;
        const int x = (int)fx, y = (int)fy;
        const bool inside_img = x - 1 >= 0 && x + 1 < U && y - 1 >= 0 && y + 1 < V;
        if (!in_view || !inside_img) {
          dtg[2 * pi] = dtg[2 * pi + 1] = 0.f;
          res[pi] = 0.f;
          continue;
        }
        auto at = [&](int xx, int yy) { return (int)img[(size_t)yy * U + xx]; };
        float dv = remap(at(x, y));
        const float f00 = remap(at(x - 1, y - 1)), f01 = remap(at(x, y - 1)), f02 = remap(at(x + 1, y - 1));
        const float f10 = remap(at(x - 1, y)), f12 = remap(at(x + 1, y));
        const float f20 = remap(at(x - 1, y + 1)), f21 = remap(at(x, y + 1)), f22 = remap(at(x + 1, y + 1));
        const float du = (1.f / 8.f) * ((-1.f) * f00 + (+1.f) * f02 + (-2.f) * f10 + (+2.f) * f12 + (-1.f) * f20 + (+1.f) * f22);
        const float dvv = (1.f / 8.f) * ((-1.f) * f00 + (-2.f) * f01 + (-1.f) * f02 + (+1.f) * f20 + (+2.f) * f21 + (+1.f) * f22);
        float nx = du, ny = dvv, ux = du, uy = dvv;
        const bool inside_fg = at(x, y) <= 127;
        const float nonvan = (du != 0.f || dvv != 0.f) ? 1.f : 0.f;
        const float bnd = isb[pi] ? 1.f : 0.f;
        if (inside_fg) { nx = -nx; ny = -ny; }
        const float dotn = nx * nrm[2 * pi] + ny * nrm[2 * pi + 1];
        if (!(dotn >= 0.f) && inside_fg) { ux = -ux; uy = -uy; }
        dv *= bnd;
        dv *= nonvan;
        ux *= bnd; uy *= bnd;
        ux *= nonvan; uy *= nonvan;
        dtg[2 * pi] = ux * (1.f / (255.f));
        dtg[2 * pi + 1] = uy * (1.f / (255.f));
        res[pi] = dv * (1.f / (255.f));
      }
      // data -> model
      for (int vi = 0; vi < V; ++vi)
        for (int ui = 0; ui < U; ++ui) {
          const size_t o = (((size_t)b * C + c) * V + vi) * U + ui;
          int counter = 0, closest = -1;
          if (img[(size_t)vi * U + ui] == 127) {
            for (int j = -2; j <= 2; ++j)
              for (int i = -2; i <= 2; ++i) {
                const int yy = vi + j, xx = ui + i;
                if (yy >= 0 && yy < V && xx >= 0 && xx < U && img[(size_t)yy * U + xx] == 127) counter++;
              }
            if (counter >= 4) {
              const float ix = ((float)ui - bx) * sc + mnx, iy = ((float)vi - by) * sc + mny;
              float best = 30.f * sc;
              for (int v = 0; v < N; ++v) {
                const size_t pi = ((size_t)b * C + c) * N + v;
                if (!isb[pi]) continue;
                const float dx = pts[2 * pi] - ix, dy = pts[2 * pi + 1] - iy;
                const float d = std::sqrt(dx * dx + dy * dy);
                if (d < best) { best = d; closest = v; }
              }
              if (closest >= 0) {
                const size_t pi = ((size_t)b * C + c) * N + closest;
                res1[2 * o] = (pts[2 * pi] - ix) / (30.f * sc);
                res1[2 * o + 1] = (pts[2 * pi + 1] - iy) / (30.f * sc);
                cid[o] = (float)closest;
              }
            }
          }
          if (counter < 4 || closest < 0) {
            res1[2 * o] = res1[2 * o + 1] = 0.f;
            cid[o] = -1.f;
          }
        }
    }
}

void mvsil_bwd(const float* g, const float* g1, const float* dtg, const float* cid, const float* crops, int B, int C,
               int N, int U, int V, double* out) {
  for (size_t i = 0; i < (size_t)B * C * N; ++i) {
    out[2 * i] = (double)dtg[2 * i] * g[i];
    out[2 * i + 1] = (double)dtg[2 * i + 1] * g[i];
  }
  for (int b = 0; b < B; ++b)
    for (int c = 0; c < C; ++c) {
      const double sc = crops[((size_t)b * C + c) * 7 + 2];
      for (size_t p = 0; p < (size_t)U * V; ++p) {
        const size_t o = ((size_t)b * C + c) * U * V + p;
        const int k = (int)cid[o];
        if (k < 0) continue;
        const size_t pi = ((size_t)b * C + c) * N + k;
        out[2 * pi] += g1[2 * o] / (30.0 * sc);
        out[2 * pi + 1] += g1[2 * o + 1] / (30.0 * sc);
      }
    }
}

template <class T>
void camproj(const T* P, const T* D, const T* ext, const T* intr, int B, int C, int N, int is_point, T* out) {
  for (int b = 0; b < B; ++b)
    for (int c = 0; c < C; ++c) {
      const T* E = ext + ((size_t)b * C + c) * 12;
      const T* K = intr + ((size_t)b * C + c) * 9;
      auto proj = [&](const T* x, T& u, T& v) {
        T h[3];
        for (int r = 0; r < 3; ++r) h[r] = K[3 * r] * x[0] + K[3 * r + 1] * x[1] + K[3 * r + 2] * x[2];
        T z = h[2];
        if (!(z > T(0.0000001))) z = T(0.00001);
        u = h[0] / z;
        v = h[1] / z;
      };
      for (int v = 0; v < N; ++v) {
        const T* p = P + ((size_t)b * N + v) * 3;
        T xc[3];
        for (int r = 0; r < 3; ++r) xc[r] = E[4 * r] * p[0] + E[4 * r + 1] * p[1] + E[4 * r + 2] * p[2] + E[4 * r + 3];
        T u, w;
        proj(xc, u, w);
        T* o = out + (((size_t)b * C + c) * N + v) * 2;
        if (is_point) {
          o[0] = u; o[1] = w;
        } else {
          const T* d = D + ((size_t)b * N + v) * 3;
          T x2[3];
          for (int r = 0; r < 3; ++r) x2[r] = xc[r] + (E[4 * r] * d[0] + E[4 * r + 1] * d[1] + E[4 * r + 2] * d[2]);
          T u2, w2;
          proj(x2, u2, w2);
          T du = u2 - u, dw = w2 - w;
          T l = std::sqrt(du * du + dw * dw);
          o[0] = du / l; o[1] = dw / l;
        }
      }
    }
}

// Reference point gradient: sum over cameras of (dI/dxc * R)^T g, with the
// intrinsics Jacobian of CameraProjectionGPUOpGrad.cu:49-59 (fx, skew, fy).
template <class T>
void camproj_bwd(const T* g, const T* P, const T* ext, const T* intr, int B, int C, int N, T* out) {
  for (int b = 0; b < B; ++b)
    for (int v = 0; v < N; ++v) {
      T acc[3] = {0, 0, 0};
      const T* p = P + ((size_t)b * N + v) * 3;
      for (int c = 0; c < C; ++c) {
        const T* E = ext + ((size_t)b * C + c) * 12;
        const T* K = intr + ((size_t)b * C + c) * 9;
        T xc[3];
        for (int r = 0; r < 3; ++r) xc[r] = E[4 * r] * p[0] + E[4 * r + 1] * p[1] + E[4 * r + 2] * p[2] + E[4 * r + 3];
        const T ax = K[0], ay = K[4], gm = K[1];
        const T J[2][3] = {{ax / xc[2], gm / xc[2], (-ax * xc[0] - gm * xc[1]) / (xc[2] * xc[2])},
                           {T(0), ay / xc[2], (-ay * xc[1]) / (xc[2] * xc[2])}};
        const T* gg = g + (((size_t)b * C + c) * N + v) * 2;
        for (int k = 0; k < 3; ++k) {
          T dp0 = J[0][0] * E[k] + J[0][1] * E[4 + k] + J[0][2] * E[8 + k];
          T dp1 = J[1][0] * E[k] + J[1][1] * E[4 + k] + J[1][2] * E[8 + k];
          acc[k] += dp0 * gg[0] + dp1 * gg[1];
        }
      }
      for (int k = 0; k < 3; ++k) out[((size_t)b * N + v) * 3 + k] = acc[k];
    }
}

// Boundary flags from projected points and a coverage bitmap [B,C,H,W].
void boundary_flags(const float* proj_uv, const uint8_t* covered, int B, int C, int N, int H, int W, uint8_t* out) {
  for (int b = 0; b < B; ++b)
    for (int c = 0; c < C; ++c)
      for (int v = 0; v < N; ++v) {
        const size_t pi = ((size_t)b * C + c) * N + v;
        const float pu = proj_uv[2 * pi], pw = proj_uv[2 * pi + 1];
        uint8_t f = 0;
        // (int)p - 1 >= 0 && (int)p + 1 < W, tested in float so that huge or
        // NaN positions never reach the integer conversion
        if (pu >= 1.f && pu < (float)(W - 1) && pw >= 1.f && pw < (float)(H - 1)) {
          const int u = (int)pu, w = (int)pw;
          for (int cc = u - 1; cc <= u + 1; ++cc)
            for (int r = w - 1; r <= w + 1; ++r)
              if (!covered[(((size_t)b * C + c) * H + r) * W + cc]) f = 1;
        }
        out[pi] = f;
      }
}

// 2D barycentrics of (u, v) in the projected triangle (uv2barycentric).
static void uv2bary(float u, float v, const float* a, const float* b, const float* c, float* e) {
  const float e2 = ((a[1] - b[1]) * u + (b[0] - a[0]) * v + a[0] * b[1] - b[0] * a[1]) /
                   ((a[1] - b[1]) * c[0] + (b[0] - a[0]) * c[1] + a[0] * b[1] - b[0] * a[1]);
  const float e1 = ((a[1] - c[1]) * u + (c[0] - a[0]) * v + a[0] * c[1] - c[0] * a[1]) /
                   ((a[1] - c[1]) * b[0] + (c[0] - a[0]) * b[1] + a[0] * c[1] - c[0] * a[1]);
  e[0] = 1.f - e2 - e1;
  e[1] = e1;
  e[2] = e2;
}

void mesh_coverage(const int32_t* faces, int F, const float* vpos, const float* ext, const float* intr, int B, int C,
                   int N, int H, int W, float* proj, uint8_t* cov) {
  std::memset(cov, 0, (size_t)B * C * H * W);
  for (int b = 0; b < B; ++b)
    for (int c = 0; c < C; ++c) {
      const float* E = ext + ((size_t)b * C + c) * 12;
      const float* K = intr + ((size_t)b * C + c) * 9;
      float* P = proj + ((size_t)b * C + c) * N * 2;
      for (int v = 0; v < N; ++v) {
        const float* X = vpos + ((size_t)b * N + v) * 3;
        const float xc = E[0] * X[0] + E[1] * X[1] + E[2] * X[2] + E[3];
        const float yc = E[4] * X[0] + E[5] * X[1] + E[6] * X[2] + E[7];
        const float zc = E[8] * X[0] + E[9] * X[1] + E[10] * X[2] + E[11];
        const float x = K[0] * xc + K[1] * yc + K[2] * zc;
        const float y = K[3] * xc + K[4] * yc + K[5] * zc;
        float z = K[6] * xc + K[7] * yc + K[8] * zc;
        if (!(z > 0.0000001f)) z = 0.00001f;
        P[2 * v] = x / z;
        P[2 * v + 1] = y / z;
      }
      uint8_t* cv = cov + ((size_t)b * C + c) * H * W;
      for (int f = 0; f < F; ++f) {
        const float* p0 = P + 2 * faces[3 * f];
        const float* p1 = P + 2 * faces[3 * f + 1];
        const float* p2 = P + 2 * faces[3 * f + 2];
        const float bx0 = std::fmax(std::fmin(p0[0], std::fmin(p1[0], p2[0])) - 0.5f, 0.f);
        const float by0 = std::fmax(std::fmin(p0[1], std::fmin(p1[1], p2[1])) - 0.5f, 0.f);
        const float bx1 = std::fmin(std::fmax(p0[0], std::fmax(p1[0], p2[0])) + 0.5f, (float)(W - 1));
        const float by1 = std::fmin(std::fmax(p0[1], std::fmax(p1[1], p2[1])) + 0.5f, (float)(H - 1));
        for (int u = (int)bx0; (float)u <= bx1; ++u)
          for (int v = (int)by0; (float)v <= by1; ++v) {
            float e[3];
            uv2bary((float)u + 0.5f, (float)v + 0.5f, p0, p1, p2, e);
            if (e[0] >= -0.001f && e[1] >= -0.001f && e[2] >= -0.001f && e[0] <= 1.001f && e[1] <= 1.001f &&
                e[2] <= 1.001f)
              cv[(size_t)v * W + u] = 1;
          }
      }
    }
}

}  // namespace orc

extern "C" void orc_mvsil_fwd(const float* pts, const float* nrm, const uint8_t* isb, const uint8_t* dt,
                              const float* crops, int B, int C, int N, int U, int V, float* res, float* res1,
                              float* dtg, float* cid) {
  orc::mvsil_fwd(pts, nrm, isb, dt, crops, B, C, N, U, V, res, res1, dtg, cid);
}
extern "C" void orc_mvsil_bwd(const float* g, const float* g1, const float* dtg, const float* cid, const float* crops,
                              int B, int C, int N, int U, int V, double* out) {
  orc::mvsil_bwd(g, g1, dtg, cid, crops, B, C, N, U, V, out);
}
extern "C" void orc_camproj_f64(const double* P, const double* D, const double* ext, const double* intr, int B, int C,
                                int N, int is_point, double* out) {
  orc::camproj<double>(P, D, ext, intr, B, C, N, is_point, out);
}
extern "C" void orc_camproj_f32(const float* P, const float* D, const float* ext, const float* intr, int B, int C,
                                int N, int is_point, float* out) {
  orc::camproj<float>(P, D, ext, intr, B, C, N, is_point, out);
}
extern "C" void orc_camproj_bwd_f64(const double* g, const double* P, const double* ext, const double* intr, int B,
                                    int C, int N, double* out) {
  orc::camproj_bwd<double>(g, P, ext, intr, B, C, N, out);
}
extern "C" void orc_mesh_coverage(const int32_t* faces, int F, const float* vpos, const float* ext, const float* intr,
                                  int B, int C, int N, int H, int W, float* proj, uint8_t* cov) {
  orc::mesh_coverage(faces, F, vpos, ext, intr, B, C, N, H, W, proj, cov);
}
extern "C" void orc_boundary_flags(const float* proj_uv, const uint8_t* covered, int B, int C, int N, int H, int W,
                                   uint8_t* out) {
  orc::boundary_flags(proj_uv, covered, B, C, N, H, W, out);
}
