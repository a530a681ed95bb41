// Distance transforms of silhouette masks and the native DT silhouette loss
// on gfx950.  The reference ships DT images computed offline
// (CreateTFRecordDataset.py:65-67 only loads them); here they are built on
// the GPU from the target masks, exactly:
//
//   edt: squared distance from every pixel centre to the nearest foreground
//        pixel centre, as integers: a column pass (vertical distance to the
//        nearest foreground pixel of the column, two sweeps) and a row pass
//        (min over columns x' of (x - x')^2 + g(x')^2, scanned outward from
//        x and stopped once (x - x')^2 reaches the best value).  Output
//        sqrt of it as float (+inf when the view has no foreground).
//   dt_encode: the uint8 encoding the silhouette op reads (value 127 on the
//        contour band, the inside distance below 127, 127 + half the
//        outside distance above), from the inside and outside transforms.
//   dtloss: L_view = sum over boundary points of D(q)^2 with D the bilinear
//        interpolation of the distance image at the point's DT-grid position
//        q = (p - crop_min) / scale + border_offset; dL/dp = 2 D grad D / scale.
#include "dcx_common.hpp"

namespace dcx {
namespace {

constexpr int kB = 256;
constexpr int kInfDist = 0x3fffffff;

// Column pass: g[v][y][x] = |y - y*| to the nearest pixel of class `fg` in
// column x (kInfDist if none).
__global__ __launch_bounds__(kB) void edt_col_kernel(const unsigned char* mask, int V, int H, int W, int fg,
                                                     int* g) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)V * W) return;
  const int v = (int)(i / W), x = (int)(i - (long long)v * W);
  const unsigned char* m = mask + (size_t)v * H * W;
  int* gv = g + (size_t)v * H * W;
  int last = -1;
  for (int y = 0; y < H; ++y) {
    if ((m[(size_t)y * W + x] != 0) == (fg != 0)) last = y;
    gv[(size_t)y * W + x] = last >= 0 ? y - last : kInfDist;
  }
  last = -1;
  for (int y = H - 1; y >= 0; --y) {
    if ((m[(size_t)y * W + x] != 0) == (fg != 0)) last = y;
    if (last >= 0 && last - y < gv[(size_t)y * W + x]) gv[(size_t)y * W + x] = last - y;
  }
}

// Row pass: exact squared distance, as a 64-bit integer (kInfDist^2 marks none).
__device__ __forceinline__ long long edt_row(const int* gr, int W, int x) {
  long long best = 0x7fffffffffffffffll;
  for (int d = 0; d < W; ++d) {
    const long long dd = (long long)d * d;
    if (dd >= best) break;
    if (x - d >= 0 && gr[x - d] != kInfDist) best = min(best, dd + (long long)gr[x - d] * gr[x - d]);
    if (d > 0 && x + d < W && gr[x + d] != kInfDist) best = min(best, dd + (long long)gr[x + d] * gr[x + d]);
  }
  return best;
}

__global__ __launch_bounds__(kB) void edt_row_kernel(const int* g, int V, int H, int W, float* dist) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)V * H * W) return;
  const long long row = i / W;
  const int x = (int)(i - row * W);
  const long long d2 = edt_row(g + row * W, W, x);
  dist[i] = d2 == 0x7fffffffffffffffll ? __int_as_float(0x7f800000) : sqrtf((float)d2);
}

// The DT image encoding: contour band (foreground pixels with a background
// or out-of-image 4-neighbour and background pixels with a foreground
// 4-neighbour) = 127;
// other foreground: clamp(rint(d_in) - 1, 0, 126); other background:
// 127 + clamp(rint(d_out / 2), 1, 128).  d_in is the distance to the
// nearest background pixel, d_out to the nearest foreground pixel.
__global__ __launch_bounds__(kB) void dt_encode_kernel(const unsigned char* mask, const int* g_fg, const int* g_bg,
                                                       int V, int H, int W, unsigned char* out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)V * H * W) return;
  const long long row = i / W;
  const int x = (int)(i - row * W);
  const int y = (int)(row % H);
  const unsigned char* m = mask + (i - (long long)y * W - x);
  const bool f = m[(size_t)y * W + x] != 0;
  // outside the image counts as background (a foreground pixel on the image
  // border is on the contour; a background one is not made so by the border)
  bool edge = false;
  const int nx[4] = {x - 1, x + 1, x, x}, ny[4] = {y, y, y - 1, y + 1};
  for (int k = 0; k < 4; ++k) {
    const bool in = nx[k] >= 0 && nx[k] < W && ny[k] >= 0 && ny[k] < H;
    const bool nf = in && m[(size_t)ny[k] * W + nx[k]] != 0;
    if (nf != f) edge = true;
  }
  int val;
  if (edge) {
    val = 127;
  } else if (f) {
    const long long d2 = edt_row(g_bg + row * W, W, x);
    const float dd = d2 == 0x7fffffffffffffffll ? 1e30f : sqrtf((float)d2);
    val = (int)fminf(fmaxf(rintf(dd) - 1.f, 0.f), 126.f);
  } else {
    const long long d2 = edt_row(g_fg + row * W, W, x);
    const float dd = d2 == 0x7fffffffffffffffll ? 1e30f : sqrtf((float)d2);
    val = 127 + (int)fminf(fmaxf(rintf(dd / 2.f), 1.f), 128.f);
  }
  out[i] = (unsigned char)val;
}

struct DtSample { float d, gx, gy; bool ok; };

// Bilinear sample of the distance image at DT-grid position (qx, qy) (pixel
// i's centre at i + 0.5) and its gradient; ok = false when the footprint
// leaves the image or touches an infinite value.
__device__ __forceinline__ DtSample dt_sample(const float* D, int H, int W, float qx, float qy) {
  DtSample s{0.f, 0.f, 0.f, false};
  const float fx = qx - 0.5f, fy = qy - 0.5f;
  if (!(fx >= 0.f && fy >= 0.f && fx < (float)(W - 1) && fy < (float)(H - 1))) return s;
  const int x0 = (int)fx, y0 = (int)fy;
  const float tx = fx - (float)x0, ty = fy - (float)y0;
  const float d00 = D[(size_t)y0 * W + x0], d10 = D[(size_t)y0 * W + x0 + 1];
  const float d01 = D[(size_t)(y0 + 1) * W + x0], d11 = D[(size_t)(y0 + 1) * W + x0 + 1];
  if (!(isfinite(d00) && isfinite(d10) && isfinite(d01) && isfinite(d11))) return s;
  s.d = (1.f - ty) * ((1.f - tx) * d00 + tx * d10) + ty * ((1.f - tx) * d01 + tx * d11);
  s.gx = (1.f - ty) * (d10 - d00) + ty * (d11 - d01);
  s.gy = (1.f - tx) * (d01 - d00) + tx * (d11 - d10);
  s.ok = true;
  return s;
}

// One workgroup per (view, block of points): residuals D(q), the block's sum
// of D^2 into a partial, and (bwd) dL/dp.
__global__ __launch_bounds__(kB) void dtloss_kernel(const float* points, const unsigned char* isb, const float* dist,
                                                    const float* crops, int N, int H, int W, float* residual,
                                                    float* partial, const float* g_loss, float* g_points) {
  __shared__ float s_red[kB / 64];
  const int view = blockIdx.y;
  const int v = blockIdx.x * kB + threadIdx.x;
  float l = 0.f;
  if (v < N) {
    const size_t i = (size_t)view * N + v;
    const float* cr = crops + (size_t)view * 7;
    const float sc = cr[2];
    const float qx = (points[2 * i] - cr[3]) / sc + cr[0], qy = (points[2 * i + 1] - cr[4]) / sc + cr[1];
    DtSample s = dt_sample(dist + (size_t)view * H * W, H, W, qx, qy);
    if (isb && !isb[i]) s.ok = false;
    const float r = s.ok ? s.d : 0.f;
    if (residual) residual[i] = r;
    l = r * r;
    if (g_points) {
      const float gl = g_loss[view];
      g_points[2 * i] = s.ok ? gl * 2.f * r * s.gx / sc : 0.f;
      g_points[2 * i + 1] = s.ok ? gl * 2.f * r * s.gy / sc : 0.f;
    }
  }
  if (!partial) return;
  l = wave_sum(l);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int k = 0; k < kB / 64; ++k) t += s_red[k];
    partial[(size_t)view * gridDim.x + blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kB) void dtloss_sum_kernel(const float* partial, int nblk, int V, float* loss) {
  const int view = blockIdx.x * kB + threadIdx.x;
  if (view >= V) return;
  float t = 0.f;
  for (int k = 0; k < nblk; ++k) t += partial[(size_t)view * nblk + k];
  loss[view] = t;
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" size_t dcx_edt_workspace_size(int V, int H, int W) {
  if (V <= 0 || H <= 0 || W <= 0) return 0;
  return 2 * sizeof(int) * (size_t)V * H * W;
}

extern "C" dcx_status dcx_edt(const uint8_t* mask, int V, int H, int W, float* dist, void* workspace,
                              size_t workspace_bytes, void* stream) {
  DCX_REQUIRE(mask && dist && V > 0 && H > 0 && W > 0, "dcx_edt: bad arguments");
  DCX_REQUIRE(workspace && workspace_bytes >= sizeof(int) * (size_t)V * H * W, "dcx_edt: workspace too small");
  hipStream_t s = as_stream(stream);
  int* g = (int*)workspace;
  hipLaunchKernelGGL(edt_col_kernel, dim3(ceil_div((long long)V * W, kB)), dim3(kB), 0, s, mask, V, H, W, 1, g);
  hipLaunchKernelGGL(edt_row_kernel, dim3(ceil_div((long long)V * H * W, kB)), dim3(kB), 0, s, g, V, H, W, dist);
  return check_launch("dcx_edt");
}

extern "C" dcx_status dcx_dt_encode(const uint8_t* mask, int V, int H, int W, uint8_t* dt_out, void* workspace,
                                    size_t workspace_bytes, void* stream) {
  DCX_REQUIRE(mask && dt_out && V > 0 && H > 0 && W > 0, "dcx_dt_encode: bad arguments");
  DCX_REQUIRE(workspace && workspace_bytes >= dcx_edt_workspace_size(V, H, W), "dcx_dt_encode: workspace too small");
  hipStream_t s = as_stream(stream);
  int* g_fg = (int*)workspace;
  int* g_bg = g_fg + (size_t)V * H * W;
  hipLaunchKernelGGL(edt_col_kernel, dim3(ceil_div((long long)V * W, kB)), dim3(kB), 0, s, mask, V, H, W, 1, g_fg);
  hipLaunchKernelGGL(edt_col_kernel, dim3(ceil_div((long long)V * W, kB)), dim3(kB), 0, s, mask, V, H, W, 0, g_bg);
  hipLaunchKernelGGL(dt_encode_kernel, dim3(ceil_div((long long)V * H * W, kB)), dim3(kB), 0, s, mask, g_fg, g_bg, V,
                     H, W, dt_out);
  return check_launch("dcx_dt_encode");
}

extern "C" size_t dcx_dtloss_workspace_size(int V, int N) {
  if (V <= 0 || N <= 0) return 0;
  return sizeof(float) * (size_t)V * ceil_div(N, kB);
}

extern "C" dcx_status dcx_dtloss_fwd(const float* points, const uint8_t* is_boundary, const float* dist,
                                     const float* crops, int V, int N, int H, int W, float* residual, float* loss,
                                     void* workspace, size_t workspace_bytes, void* stream) {
  DCX_REQUIRE(points && dist && crops && loss && V > 0 && N > 0 && H > 1 && W > 1, "dcx_dtloss_fwd: bad arguments");
  DCX_REQUIRE(workspace && workspace_bytes >= dcx_dtloss_workspace_size(V, N), "dcx_dtloss_fwd: workspace too small");
  hipStream_t s = as_stream(stream);
  const unsigned nblk = ceil_div(N, kB);
  float* partial = (float*)workspace;
  hipLaunchKernelGGL(dtloss_kernel, dim3(nblk, V), dim3(kB), 0, s, points, is_boundary, dist, crops, N, H, W, residual,
                     partial, nullptr, nullptr);
  hipLaunchKernelGGL(dtloss_sum_kernel, dim3(ceil_div(V, kB)), dim3(kB), 0, s, partial, (int)nblk, V, loss);
  return check_launch("dcx_dtloss_fwd");
}

extern "C" dcx_status dcx_dtloss_bwd(const float* points, const uint8_t* is_boundary, const float* dist,
                                     const float* crops, int V, int N, int H, int W, const float* g_loss,
                                     float* points_grad, void* stream) {
  DCX_REQUIRE(points && dist && crops && g_loss && points_grad && V > 0 && N > 0 && H > 1 && W > 1,
              "dcx_dtloss_bwd: bad arguments");
  hipLaunchKernelGGL(dtloss_kernel, dim3(ceil_div(N, kB), V), dim3(kB), 0, as_stream(stream), points, is_boundary,
                     dist, crops, N, H, W, nullptr, nullptr, g_loss, points_grad);
  return check_launch("dcx_dtloss_bwd");
}
