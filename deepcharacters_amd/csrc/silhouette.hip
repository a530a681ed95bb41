// Multi-view silhouette loss (MultiViewSilhouetteLossGpu / ...Grad), camera
// projection (CameraProjectionGpu / ...Grad) and projected-mesh boundary
// flags (ProjectedMeshBoundaryGpu) on gfx950.
//
// The data->model term searches, for each DT boundary pixel, the nearest
// boundary point within 30*scale.  DCX_SIL_BRUTE scans every point as the
// reference does (MultiViewSilhouetteLossGPUOp.cu:195-212).  DCX_SIL_BINNED
// first buckets the view's boundary points into a uniform grid whose cells
// are slightly larger than the search radius, then scans only the 3x3 cells
// around the pixel; it evaluates the same float distance and keeps the
// lexicographic minimum of (distance, index), so both paths return the same
// ids as the reference's strict-< scan (lowest index among equal distances).
#include "raster_internal.hpp"

namespace dcx {

dcx_status exclusive_scan(const unsigned int* in, unsigned int* out, unsigned int* aux, long long n,
                          hipStream_t s);

namespace {

constexpr int kB = 256;

__device__ __forceinline__ float remap(int v) { return v >= 127 ? (float)((v - 127) * 2) : (float)v; }

struct SilArgs {
  int B, C, N, U, V;
  const float* pts;      // [B,C,N,2]
  const float* nrm;      // [B,C,N,2]
  const unsigned char* isb;  // [B,C,N]
  const unsigned char* dt;   // [B,C,U*V]
  const float* crops;    // [B,C,7]
  float* res;            // [B,C,N]
  float* res1;           // [B,C,U,V,2]
  float* dtg;            // [B,C,N,2]
  float* cid;            // [B,C,U,V]
};

struct GridArgs {
  int gx, gy, cells_per_view;
  unsigned int* cnt;     // [views*cells]
  unsigned int* off;     // [views*cells + 1]
  unsigned int* fill;
  unsigned int* items;   // point ids, capacity = B*C*N
};

// Model->data residual per (batch, camera, point): MultiViewSilhouetteLossGPUOp.cu:11-138.
__global__ __launch_bounds__(kB) void sil_point_kernel(SilArgs a) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)a.B * a.C * a.N) return;
  const int view = (int)(i / a.N);
  const float* cr = a.crops + (size_t)view * 7;
  const float pu = a.pts[2 * i], pv = a.pts[2 * i + 1];
  const bool in_view = pu - 2 > cr[3] && pu + 2 < cr[5] && pv - 2 > cr[4] && pv + 2 < cr[6];
  const float fx = (pu - cr[3]) / cr[2] + cr[0], fy = (pv - cr[4]) / cr[2] + cr[1];
  const int x = (int)fx, y = (int)fy;
  const bool inside_img = x - 1 >= 0 && x + 1 < a.U && y - 1 >= 0 && y + 1 < a.V;
  if (!in_view || !inside_img) {
    a.dtg[2 * i] = 0.f; a.dtg[2 * i + 1] = 0.f; a.res[i] = 0.f;
    return;
  }
  const unsigned char* img = a.dt + (size_t)view * a.U * a.V;
  auto at = [&](int xx, int yy) { return (int)img[(size_t)yy * a.U + xx]; };
  float dv = remap(at(x, y));
  const float f00 = remap(at(x - 1, y - 1)), f01 = remap(at(x, y - 1)), f02 = remap(at(x + 1, y - 1));
  const float f10 = remap(at(x - 1, y)), f12 = remap(at(x + 1, y));
  const float f20 = remap(at(x - 1, y + 1)), f21 = remap(at(x, y + 1)), f22 = remap(at(x + 1, y + 1));
  const float du = (1.f / 8.f) * ((-1.f) * f00 + (+1.f) * f02 + (-2.f) * f10 + (+2.f) * f12 + (-1.f) * f20 + (+1.f) * f22);
  const float dvv = (1.f / 8.f) * ((-1.f) * f00 + (-2.f) * f01 + (-1.f) * f02 + (+1.f) * f20 + (+2.f) * f21 + (+1.f) * f22);
  float nx = du, ny = dvv, ux = du, uy = dvv;
  const bool inside_fg = at(x, y) <= 127;
  const float nonvan = (du != 0.f || dvv != 0.f) ? 1.f : 0.f;
  const float bnd = a.isb[i] ? 1.f : 0.f;
  if (inside_fg) { nx = -nx; ny = -ny; }
  const float dotn = nx * a.nrm[2 * i] + ny * a.nrm[2 * i + 1];
  if (!(dotn >= 0.f) && inside_fg) { ux = -ux; uy = -uy; }
  dv *= bnd;
  dv *= nonvan;
  ux *= bnd; uy *= bnd;
  ux *= nonvan; uy *= nonvan;
  a.dtg[2 * i] = ux * (1.f / (255.f));
  a.dtg[2 * i + 1] = uy * (1.f / (255.f));
  a.res[i] = dv * (1.f / (255.f));
}

__device__ __forceinline__ bool is_dt_boundary(const unsigned char* img, int U, int V, int ui, int vi) {
  if (img[(size_t)vi * U + ui] != 127) return false;
  int counter = 0;
  for (int j = -2; j <= 2; ++j)
    for (int i = -2; i <= 2; ++i) {
      const int yy = vi + j, xx = ui + i;
      if (yy >= 0 && yy < V && xx >= 0 && xx < U && img[(size_t)yy * U + xx] == 127) counter++;
    }
  return counter >= 4;
}

__device__ __forceinline__ void write_d2m(const SilArgs& a, size_t o, int view, int best, float ix, float iy, float sc) {
  if (best >= 0) {
    const size_t pi = (size_t)view * a.N + best;
    a.res1[2 * o] = (a.pts[2 * pi] - ix) / (30.f * sc);
    a.res1[2 * o + 1] = (a.pts[2 * pi + 1] - iy) / (30.f * sc);
    a.cid[o] = (float)best;
  } else {
    a.res1[2 * o] = 0.f; a.res1[2 * o + 1] = 0.f;
    a.cid[o] = -1.f;
  }
}

// Data->model per DT pixel, brute force over all points.
__global__ __launch_bounds__(kB) void sil_pixel_brute_kernel(SilArgs a) {
  const long long o = (long long)blockIdx.x * kB + threadIdx.x;
  const long long UV = (long long)a.U * a.V;
  if (o >= (long long)a.B * a.C * UV) return;
  const int view = (int)(o / UV);
  const int p = (int)(o - (long long)view * UV);
  const int ui = p % a.U, vi = p / a.U;
  const unsigned char* img = a.dt + (size_t)view * UV;
  int best = -1;
  float ix = 0.f, iy = 0.f, sc = 1.f;
  if (is_dt_boundary(img, a.U, a.V, ui, vi)) {
    const float* cr = a.crops + (size_t)view * 7;
    sc = cr[2];
    ix = ((float)ui - cr[0]) * sc + cr[3];
    iy = ((float)vi - cr[1]) * sc + cr[4];
    float bd = 30.f * sc;
    const float* pv = a.pts + (size_t)view * a.N * 2;
    const unsigned char* bv = a.isb + (size_t)view * a.N;
    for (int v = 0; v < a.N; ++v) {
      if (!bv[v]) continue;
      const float dx = pv[2 * v] - ix, dy = pv[2 * v + 1] - iy;
      const float d = sqrtf(dx * dx + dy * dy);
      if (d < bd) { bd = d; best = v; }
    }
  }
  write_d2m(a, o, view, best, ix, iy, sc);
}

struct GridGeom {
  float ox, oy, cell;
};

__device__ __forceinline__ GridGeom grid_geom(const float* cr) {
  const float sc = cr[2];
  const float R = 30.f * sc;
  GridGeom g;
  g.cell = R * (1.f + 1.f / 1024.f);
  g.ox = (0.f - cr[0]) * sc + cr[3] - 2.f * g.cell;
  g.oy = (0.f - cr[1]) * sc + cr[4] - 2.f * g.cell;
  return g;
}

__device__ __forceinline__ int cell_of(const GridGeom& g, const GridArgs& ga, float x, float y) {
  const float fx = floorf((x - g.ox) / g.cell), fy = floorf((y - g.oy) / g.cell);
  if (!(fx >= 0.f && fy >= 0.f && fx < (float)ga.gx && fy < (float)ga.gy)) return -1;
  return (int)fy * ga.gx + (int)fx;
}

__global__ __launch_bounds__(kB) void sil_grid_count_kernel(SilArgs a, GridArgs g) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)a.B * a.C * a.N || !a.isb[i]) return;
  const int view = (int)(i / a.N);
  const GridGeom gg = grid_geom(a.crops + (size_t)view * 7);
  const int c = cell_of(gg, g, a.pts[2 * i], a.pts[2 * i + 1]);
  if (c >= 0) atomicAdd(g.cnt + (size_t)view * g.cells_per_view + c, 1u);
}

__global__ __launch_bounds__(kB) void sil_grid_scatter_kernel(SilArgs a, GridArgs g) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)a.B * a.C * a.N || !a.isb[i]) return;
  const int view = (int)(i / a.N);
  const GridGeom gg = grid_geom(a.crops + (size_t)view * 7);
  const int c = cell_of(gg, g, a.pts[2 * i], a.pts[2 * i + 1]);
  if (c < 0) return;
  const size_t cell = (size_t)view * g.cells_per_view + c;
  g.items[g.off[cell] + atomicAdd(g.fill + cell, 1u)] = (unsigned int)(i - (long long)view * a.N);
}

__global__ __launch_bounds__(kB) void sil_pixel_binned_kernel(SilArgs a, GridArgs g) {
  const long long o = (long long)blockIdx.x * kB + threadIdx.x;
  const long long UV = (long long)a.U * a.V;
  if (o >= (long long)a.B * a.C * UV) return;
  const int view = (int)(o / UV);
  const int p = (int)(o - (long long)view * UV);
  const int ui = p % a.U, vi = p / a.U;
  const unsigned char* img = a.dt + (size_t)view * UV;
  int best = -1;
  float ix = 0.f, iy = 0.f, sc = 1.f;
  if (is_dt_boundary(img, a.U, a.V, ui, vi)) {
    const float* cr = a.crops + (size_t)view * 7;
    sc = cr[2];
    ix = ((float)ui - cr[0]) * sc + cr[3];
    iy = ((float)vi - cr[1]) * sc + cr[4];
    const float R = 30.f * sc;
    float bd = R;
    const GridGeom gg = grid_geom(cr);
    const int cx = (int)floorf((ix - gg.ox) / gg.cell), cy = (int)floorf((iy - gg.oy) / gg.cell);
    const float* pv = a.pts + (size_t)view * a.N * 2;
    for (int yy = max(cy - 1, 0); yy <= min(cy + 1, g.gy - 1); ++yy)
      for (int xx = max(cx - 1, 0); xx <= min(cx + 1, g.gx - 1); ++xx) {
        const size_t cell = (size_t)view * g.cells_per_view + yy * g.gx + xx;
        for (unsigned int k = g.off[cell]; k < g.off[cell + 1]; ++k) {
          const int v = (int)g.items[k];
          const float dx = pv[2 * v] - ix, dy = pv[2 * v + 1] - iy;
          const float d = sqrtf(dx * dx + dy * dy);
          if (d < bd || (d == bd && best >= 0 && v < best)) { bd = d; best = v; }
        }
      }
  }
  write_d2m(a, o, view, best, ix, iy, sc);
}

__global__ __launch_bounds__(kB) void sil_bwd_point_kernel(const float* __restrict__ g, const float* __restrict__ dtg,
                                                            long long n, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = dtg[2 * i] * g[i];
  out[2 * i + 1] = dtg[2 * i + 1] * g[i];
}

__global__ __launch_bounds__(kB) void sil_bwd_pixel_kernel(const float* __restrict__ g1, const float* __restrict__ cid,
                                                            const float* __restrict__ crops, int N, long long UV,
                                                            long long n, float* __restrict__ out) {
  const long long o = (long long)blockIdx.x * kB + threadIdx.x;
  if (o >= n) return;
  const int k = (int)cid[o];
  if (k < 0) return;
  const int view = (int)(o / UV);
  const float sc = crops[(size_t)view * 7 + 2];
  const size_t pi = (size_t)view * N + k;
  atomicAdd(out + 2 * pi, g1[2 * o] / (30.f * sc));
  atomicAdd(out + 2 * pi + 1, g1[2 * o + 1] / (30.f * sc));
}

// CameraProjectionGPUOp.cu:12-86.
__global__ __launch_bounds__(kB) void camproj_kernel(const float* __restrict__ P, const float* __restrict__ D,
                                                      const float* __restrict__ ext, const float* __restrict__ intr,
                                                      int B, int C, int N, int is_point, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * C * N) return;
  const int view = (int)(i / N), v = (int)(i - (long long)view * N);
  const int b = view / C;
  const float* E = ext + (size_t)view * 12;
  const float* K = intr + (size_t)view * 9;
  const f3 p = ld3(P + ((size_t)b * N + v) * 3);
  const f3 xc = mk3(E[0] * p.x + E[1] * p.y + E[2] * p.z + E[3], E[4] * p.x + E[5] * p.y + E[6] * p.z + E[7],
                    E[8] * p.x + E[9] * p.y + E[10] * p.z + E[11]);
  auto proj = [&](f3 x) {
    const float h0 = K[0] * x.x + K[1] * x.y + K[2] * x.z;
    const float h1 = K[3] * x.x + K[4] * x.y + K[5] * x.z;
    float h2 = K[6] * x.x + K[7] * x.y + K[8] * x.z;
    if (!(h2 > 0.0000001f)) h2 = 0.00001f;
    return mk2(h0 / h2, h1 / h2);
  };
  const f2 uv = proj(xc);
  if (is_point) {
    out[2 * i] = uv.x; out[2 * i + 1] = uv.y;
  } else {
    const f3 d = ld3(D + ((size_t)b * N + v) * 3);
    const f3 x2 = mk3(xc.x + (E[0] * d.x + E[1] * d.y + E[2] * d.z), xc.y + (E[4] * d.x + E[5] * d.y + E[6] * d.z),
                      xc.z + (E[8] * d.x + E[9] * d.y + E[10] * d.z));
    const f2 uv2 = proj(x2);
    const float du = uv2.x - uv.x, dv = uv2.y - uv.y;
    const float l = sqrtf(du * du + dv * dv);
    out[2 * i] = du / l; out[2 * i + 1] = dv / l;
  }
}

// CameraProjectionGPUOpGrad.cu:12-79: sum over cameras of (dI/dxc R)^T g.
__global__ __launch_bounds__(kB) void camproj_bwd_kernel(const float* __restrict__ g, const float* __restrict__ P,
                                                          const float* __restrict__ ext, const float* __restrict__ intr,
                                                          int B, int C, int N, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * N) return;
  const int b = (int)(i / N), v = (int)(i - (long long)b * N);
  const f3 p = ld3(P + 3 * i);
  f3 acc = mk3(0.f, 0.f, 0.f);
  for (int c = 0; c < C; ++c) {
    const float* E = ext + ((size_t)b * C + c) * 12;
    const float* K = intr + ((size_t)b * C + c) * 9;
    const f3 xc = mk3(E[0] * p.x + E[1] * p.y + E[2] * p.z + E[3], E[4] * p.x + E[5] * p.y + E[6] * p.z + E[7],
                      E[8] * p.x + E[9] * p.y + E[10] * p.z + E[11]);
    const float ax = K[0], ay = K[4], gm = K[1];
    const float j00 = ax / xc.z, j01 = gm / xc.z, j02 = (-ax * xc.x - gm * xc.y) / (xc.z * xc.z);
    const float j11 = ay / xc.z, j12 = (-ay * xc.y) / (xc.z * xc.z);
    const float* gg = g + (((size_t)b * C + c) * N + v) * 2;
    for (int k = 0; k < 3; ++k) {
      const float dp0 = j00 * E[k] + j01 * E[4 + k] + j02 * E[8 + k];
      const float dp1 = j11 * E[4 + k] + j12 * E[8 + k];
      const float add = dp0 * gg[0] + dp1 * gg[1];
      if (k == 0) acc.x += add; else if (k == 1) acc.y += add; else acc.z += add;
    }
  }
  out[3 * i] = acc.x; out[3 * i + 1] = acc.y; out[3 * i + 2] = acc.z;
}

// ProjectedMeshBoundaryGpu semantics on a face buffer (face >= 0 = covered).
__global__ __launch_bounds__(kB) void boundary_kernel(const int* __restrict__ face, const float* __restrict__ uv, int B,
                                                       int C, int N, int H, int W, unsigned char* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * C * N) return;
  const int view = (int)(i / N);
  const float pu = uv[2 * i], pw = uv[2 * i + 1];
  unsigned char f = 0;
  // (int)p - 1 >= 0 && (int)p + 1 < W, tested in float so that huge or NaN
  // positions never reach the integer conversion
  if (pu >= 1.f && pu < (float)(W - 1) && pw >= 1.f && pw < (float)(H - 1)) {
    const int u = (int)pu, w = (int)pw;
    const int* fb = face + (size_t)view * H * W;
    for (int cc = u - 1; cc <= u + 1; ++cc)
      for (int r = w - 1; r <= w + 1; ++r)
        if (fb[(size_t)r * W + cc] < 0) f = 1;
  }
  out[i] = f;
}

// ProjectedMeshBoundaryGpu's own coverage raster (Skeletool/Mesh/
// CUDABasedRasterization.cu:64-140).  The vertices are projected as
// getCamSpacePoint / projectPointFloat3 do (project_vertex: z clamped to 1e-5
// when not in front); one lane per (view, face) walks the face's pixel box
// [min - 0.5, max + 0.5], clamped to the image and truncated, and marks every
// pixel centre whose 2D barycentrics (uv2barycentric) lie in [-0.001, 1.001].
// No depth test: the bitmap is the union of the faces' pixels, every writer
// stores the same byte.
__global__ __launch_bounds__(kB) void cov_project_kernel(const float* __restrict__ vpos, const float* __restrict__ ext,
                                                          const float* __restrict__ intr, int B, int C, int N,
                                                          float* __restrict__ uv) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * C * N) return;
  const int view = (int)(i / N), v = (int)(i - (long long)view * N);
  const int b = view / C;
  const f4 p = project_vertex(ext + (size_t)view * 12, intr + (size_t)view * 9, vpos + ((size_t)b * N + v) * 3);
  uv[2 * i] = p.x;
  uv[2 * i + 1] = p.y;
}

__device__ __forceinline__ f3 uv2bary(float u, float v, f2 a, f2 b, f2 c) {
  const float e2 = ((a.y - b.y) * u + (b.x - a.x) * v + a.x * b.y - b.x * a.y) /
                   ((a.y - b.y) * c.x + (b.x - a.x) * c.y + a.x * b.y - b.x * a.y);
  const float e1 = ((a.y - c.y) * u + (c.x - a.x) * v + a.x * c.y - c.x * a.y) /
                   ((a.y - c.y) * b.x + (c.x - a.x) * b.y + a.x * c.y - c.x * a.y);
  return mk3(1.f - e2 - e1, e1, e2);
}

__global__ __launch_bounds__(kB) void cov_raster_kernel(const float* __restrict__ uv, const int* __restrict__ faces,
                                                         int B, int C, int N, int F, int H, int W,
                                                         unsigned char* __restrict__ cov) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * C * F) return;
  const int view = (int)(i / F), f = (int)(i - (long long)view * F);
  const f2* P = (const f2*)uv + (size_t)view * N;
  const f2 p0 = P[faces[3 * f]], p1 = P[faces[3 * f + 1]], p2 = P[faces[3 * f + 2]];
  const float bx0 = fmaxf(fminf(p0.x, fminf(p1.x, p2.x)) - 0.5f, 0.f);
  const float by0 = fmaxf(fminf(p0.y, fminf(p1.y, p2.y)) - 0.5f, 0.f);
  const float bx1 = fminf(fmaxf(p0.x, fmaxf(p1.x, p2.x)) + 0.5f, (float)(W - 1));
  const float by1 = fminf(fmaxf(p0.y, fmaxf(p1.y, p2.y)) + 0.5f, (float)(H - 1));
  unsigned char* cv = cov + (size_t)view * H * W;
  for (int u = (int)bx0; (float)u <= bx1; ++u)
    for (int v = (int)by0; (float)v <= by1; ++v) {
      const f3 e = uv2bary((float)u + 0.5f, (float)v + 0.5f, p0, p1, p2);
      if (e.x >= -0.001f && e.y >= -0.001f && e.z >= -0.001f && e.x <= 1.001f && e.y <= 1.001f && e.z <= 1.001f)
        cv[(size_t)v * W + u] = 1;
    }
}

// The boundary test of boundary_kernel on the coverage bitmap (1 = covered).
__global__ __launch_bounds__(kB) void cov_boundary_kernel(const unsigned char* __restrict__ cov,
                                                           const float* __restrict__ uv, int B, int C, int N, int H,
                                                           int W, unsigned char* __restrict__ out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * C * N) return;
  const int view = (int)(i / N);
  const float pu = uv[2 * i], pw = uv[2 * i + 1];
  unsigned char f = 0;
  if (pu >= 1.f && pu < (float)(W - 1) && pw >= 1.f && pw < (float)(H - 1)) {
    const int u = (int)pu, w = (int)pw;
    const unsigned char* cv = cov + (size_t)view * H * W;
    for (int cc = u - 1; cc <= u + 1; ++cc)
      for (int r = w - 1; r <= w + 1; ++r)
        if (!cv[(size_t)r * W + cc]) f = 1;
  }
  out[i] = f;
}

__global__ __launch_bounds__(kB) void fill_u8_kernel(unsigned char* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i < n) p[i] = 0;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct SilWs {
  int gx, gy;
  size_t cnt, off, fill, aux, items, total;
};

SilWs sil_ws(int B, int C, int N, int U, int V) {
  SilWs w{};
  w.gx = U / 30 + 6;
  w.gy = V / 30 + 6;
  const size_t cells = (size_t)B * C * w.gx * w.gy;
  size_t o = 0;
  w.cnt = o; o = align256(o + 4 * cells);
  w.off = o; o = align256(o + 4 * (cells + 1));
  w.fill = o; o = align256(o + 4 * cells);
  w.aux = o; o = align256(o + 4 * ((cells + 1023) / 1024 + 1));
  w.items = o; o = align256(o + 4 * (size_t)B * C * N);
  w.total = o;
  return w;
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" size_t dcx_mvsil_workspace_size(int B, int C, int N, int U, int V) {
  if (B <= 0 || C <= 0 || N <= 0 || U <= 0 || V <= 0) return 0;
  return sil_ws(B, C, N, U, V).total;
}

extern "C" dcx_status dcx_mvsil_fwd(const float* points, const float* normals, const uint8_t* is_boundary,
                                    const uint8_t* dt_image, const float* crops, int B, int C, int N, int U, int V,
                                    int search, float* residual, float* residual1, float* dt_image_gradient,
                                    float* closest_vertex_id, void* workspace, size_t ws_bytes, void* stream) {
  DCX_REQUIRE(B > 0 && C > 0 && N > 0 && U > 0 && V > 0, "dcx_mvsil_fwd: bad sizes");
  DCX_REQUIRE(points && normals && is_boundary && dt_image && crops && residual && residual1 && dt_image_gradient &&
                  closest_vertex_id,
              "dcx_mvsil_fwd: null input/output");
  DCX_REQUIRE(search == DCX_SIL_BRUTE || search == DCX_SIL_BINNED, "dcx_mvsil_fwd: search mode %d", search);
  hipStream_t s = as_stream(stream);
  SilArgs a{B, C, N, U, V, points, normals, is_boundary, dt_image, crops, residual, residual1, dt_image_gradient,
            closest_vertex_id};
  const long long P = (long long)B * C * N, PIX = (long long)B * C * U * V;
  hipLaunchKernelGGL(sil_point_kernel, dim3(ceil_div(P, kB)), dim3(kB), 0, s, a);
  if (search == DCX_SIL_BRUTE) {
    hipLaunchKernelGGL(sil_pixel_brute_kernel, dim3(ceil_div(PIX, kB)), dim3(kB), 0, s, a);
    return check_launch("dcx_mvsil_fwd(brute)");
  }
  const SilWs L = sil_ws(B, C, N, U, V);
  DCX_REQUIRE(workspace && ws_bytes >= L.total, "dcx_mvsil_fwd: workspace %zu < %zu", ws_bytes, L.total);
  char* w = (char*)workspace;
  GridArgs g{L.gx, L.gy, L.gx * L.gy, (unsigned int*)(w + L.cnt), (unsigned int*)(w + L.off),
             (unsigned int*)(w + L.fill), (unsigned int*)(w + L.items)};
  const long long cells = (long long)B * C * g.cells_per_view;
  DCX_HIP(hipMemsetAsync(g.cnt, 0, 4 * cells, s));
  DCX_HIP(hipMemsetAsync(g.fill, 0, 4 * cells, s));
  hipLaunchKernelGGL(sil_grid_count_kernel, dim3(ceil_div(P, kB)), dim3(kB), 0, s, a, g);
  dcx_status st = exclusive_scan(g.cnt, g.off, (unsigned int*)(w + L.aux), cells, s);
  if (st) return st;
  hipLaunchKernelGGL(sil_grid_scatter_kernel, dim3(ceil_div(P, kB)), dim3(kB), 0, s, a, g);
  hipLaunchKernelGGL(sil_pixel_binned_kernel, dim3(ceil_div(PIX, kB)), dim3(kB), 0, s, a, g);
  return check_launch("dcx_mvsil_fwd(binned)");
}

extern "C" dcx_status dcx_mvsil_bwd(const float* mv_dt_grad, const float* mv_dt_grad_1, const float* dt_image_grad,
                                    const float* closest_vertex_id, const float* crops, int B, int C, int N, int U,
                                    int V, float* points_grad, void* stream) {
  DCX_REQUIRE(B > 0 && C > 0 && N > 0 && U > 0 && V > 0, "dcx_mvsil_bwd: bad sizes");
  DCX_REQUIRE(mv_dt_grad && mv_dt_grad_1 && dt_image_grad && closest_vertex_id && crops && points_grad,
              "dcx_mvsil_bwd: null input/output");
  hipStream_t s = as_stream(stream);
  const long long P = (long long)B * C * N, PIX = (long long)B * C * U * V;
  hipLaunchKernelGGL(sil_bwd_point_kernel, dim3(ceil_div(P, kB)), dim3(kB), 0, s, mv_dt_grad, dt_image_grad, P,
                     points_grad);
  hipLaunchKernelGGL(sil_bwd_pixel_kernel, dim3(ceil_div(PIX, kB)), dim3(kB), 0, s, mv_dt_grad_1, closest_vertex_id,
                     crops, N, (long long)U * V, PIX, points_grad);
  return check_launch("dcx_mvsil_bwd");
}

extern "C" dcx_status dcx_camproj_fwd(const float* points, const float* vectors, const float* extrinsics,
                                      const float* intrinsics, int B, int C, int N, int is_point, float* out,
                                      void* stream) {
  DCX_REQUIRE(B > 0 && C > 0 && N > 0, "dcx_camproj_fwd: bad sizes");
  DCX_REQUIRE(points && extrinsics && intrinsics && out && (is_point || vectors), "dcx_camproj_fwd: null input");
  hipLaunchKernelGGL(camproj_kernel, dim3(ceil_div((long long)B * C * N, kB)), dim3(kB), 0, as_stream(stream), points,
                     vectors, extrinsics, intrinsics, B, C, N, is_point, out);
  return check_launch("dcx_camproj_fwd");
}

extern "C" dcx_status dcx_camproj_bwd(const float* points_image_grad, const float* points, const float* extrinsics,
                                      const float* intrinsics, int B, int C, int N, float* points_grad, void* stream) {
  DCX_REQUIRE(B > 0 && C > 0 && N > 0, "dcx_camproj_bwd: bad sizes");
  DCX_REQUIRE(points_image_grad && points && extrinsics && intrinsics && points_grad, "dcx_camproj_bwd: null input");
  hipLaunchKernelGGL(camproj_bwd_kernel, dim3(ceil_div((long long)B * N, kB)), dim3(kB), 0, as_stream(stream),
                     points_image_grad, points, extrinsics, intrinsics, B, C, N, points_grad);
  return check_launch("dcx_camproj_bwd");
}

extern "C" size_t dcx_projected_mesh_boundary_workspace_size(int B, int C, int N, int H, int W) {
  if (B <= 0 || C <= 0 || N <= 0 || H <= 0 || W <= 0) return 0;
  return align256(sizeof(float) * 2 * (size_t)B * C * N) + (size_t)B * C * H * W;
}

extern "C" dcx_status dcx_projected_mesh_boundary(const dcx_topo* topo, int B, int C, int H, int W,
                                                  const float* points, const float* extrinsics,
                                                  const float* intrinsics, uint8_t* is_boundary, uint8_t* coverage,
                                                  void* workspace, size_t workspace_bytes, void* stream) {
  DCX_REQUIRE(topo && B > 0 && C > 0 && H > 0 && W > 0, "dcx_projected_mesh_boundary: bad sizes");
  DCX_REQUIRE(points && extrinsics && intrinsics && is_boundary && workspace,
              "dcx_projected_mesh_boundary: null argument");
  const int N = topo->N, F = topo->F;
  DCX_REQUIRE(workspace_bytes >= dcx_projected_mesh_boundary_workspace_size(B, C, N, H, W),
              "dcx_projected_mesh_boundary: workspace too small");
  hipStream_t s = as_stream(stream);
  float* uv = (float*)workspace;
  unsigned char* cov = coverage ? coverage : (unsigned char*)workspace + align256(sizeof(float) * 2 * (size_t)B * C * N);
  const long long npix = (long long)B * C * H * W;
  hipLaunchKernelGGL(fill_u8_kernel, dim3(ceil_div(npix, kB)), dim3(kB), 0, s, cov, npix);
  hipLaunchKernelGGL(cov_project_kernel, dim3(ceil_div((long long)B * C * N, kB)), dim3(kB), 0, s, points, extrinsics,
                     intrinsics, B, C, N, uv);
  hipLaunchKernelGGL(cov_raster_kernel, dim3(ceil_div((long long)B * C * F, kB)), dim3(kB), 0, s, uv, topo->faces, B,
                     C, N, F, H, W, cov);
  hipLaunchKernelGGL(cov_boundary_kernel, dim3(ceil_div((long long)B * C * N, kB)), dim3(kB), 0, s, cov, uv, B, C, N,
                     H, W, is_boundary);
  return check_launch("dcx_projected_mesh_boundary");
}

extern "C" dcx_status dcx_mesh_boundary(const int32_t* face, const float* points_image, int B, int C, int N, int H,
                                        int W, uint8_t* is_boundary, void* stream) {
  DCX_REQUIRE(B > 0 && C > 0 && N > 0 && H > 0 && W > 0, "dcx_mesh_boundary: bad sizes");
  DCX_REQUIRE(face && points_image && is_boundary, "dcx_mesh_boundary: null input");
  hipLaunchKernelGGL(boundary_kernel, dim3(ceil_div((long long)B * C * N, kB)), dim3(kB), 0, as_stream(stream), face,
                     points_image, B, C, N, H, W, is_boundary);
  return check_launch("dcx_mesh_boundary");
}
