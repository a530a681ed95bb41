// Batched differentiable-character step on gfx950: many characters with
// uneven meshes, several views each, forward + backward in one launch
// sequence that can be captured once into a hipGraph and replayed.
//
// Per step (all characters at once, indexed through per-character and
// per-view descriptor tables):
//   clear            (an empty stage: the projection kernel zeroes the tile /
//                    grid counters and the fallback accumulators with plain
//                    stores, no memset, so the sequence is graph-safe)
//   node_fwd         per ED node: Euler delta rotation -> 15-float node row
//   vertex_fwd       per vertex: ED inner sum (4 nodes) then DQ skinning
//   normals          per vertex: normalised face-normal sum of the skinned mesh
//   view_normals     per (view, vertex): image-space normal direction
//                    (CameraProjection vector mode); with normals on a second
//                    stream beside project and the binning stages
//   project          per (view, vertex): image position / depth
//                    (CameraProjection point mode)
//   tri_count        per (view, face): triangle record, tiles touched
//   tile_scan, tri_scan, bin_scatter   per-tile triangle lists and, per
//                    triangle, the slots of its (triangle, tile) records
//   fine             per 32x8 tile: LDS depth (64-bit ds_min on packed
//                    (depth, face) keys); per pixel: face / barycentric /
//                    colour outputs, L1 colour loss vs the target image and
//                    its gradient (barycentric chain rule + the target-image
//                    model-to-data term), kept in LDS; then the thread that
//                    rasterised a triangle sums the pixels it won and writes
//                    one (triangle, tile) gradient record -- no atomics
//   sil_vertex       per (view, vertex): silhouette-boundary flag, the
//                    model->data DT residual, squared loss, its gradient, and
//                    the boundary vertex's count in the nearest-point grid
//   sil grid + sil_pixel   data->model residual via the binned nearest
//                    boundary-point search, squared loss, gradient
//   face_sum         per face: its (triangle, tile) records summed over tiles
//                    and views, each corner's sums stored at its slot of the
//                    vertex -> face-corner list
//   vertex_bwd       per vertex: sum of its corner slots, camera-projection
//                    adjoint, DQ-skinning adjoint, colour gradient
//   joint_partial    one wavefront per chunk of <= 512 (vertex, slot) entries
//                    of one joint (static CSR cut into chunks): DQ gradient
//                    partial
//   joint_reduce     per (joint, component): its chunk partials in chunk
//                    order (fixed order throughout: deterministic)
//   node_bwd         one wavefront per ED node: translation / Euler gradients
//   finalize         per character: losses from the per-block partial sums
#include <algorithm>
#include <vector>

#include "raster_internal.hpp"
#include "skin_common.hpp"

namespace dcx {
namespace {

constexpr int kB = 256;
constexpr int KS = 4;         // skinning joints per vertex
constexpr int kJChunk = 512;  // (vertex, slot) entries per joint-reduction chunk
constexpr int KE = 4;         // ED nodes per vertex
constexpr int kNodeRow = 15;  // Rd(9) g(3) td(3)
constexpr int kImgFs = 2;     // image_filter_size of the model-to-data term

// 1 / (L1 norm of the image-gradient tap weights xx / (xx^2 + yy^2)) over the
// (2 kImgFs + 1)^2 window, accumulated in the order imageGradient uses.
constexpr float img_grad_inv_norm() {
  float n = 0.f;
  for (int yy = -kImgFs; yy <= kImgFs; ++yy)
    for (int xx = -kImgFs; xx <= kImgFs; ++xx) {
      const int den = xx * xx + yy * yy;
      if (den != 0) n += (float)(xx < 0 ? -xx : xx) / (float)den;
    }
  return 1.f / n;
}
constexpr float kImgGradInvNorm = img_grad_inv_norm();
constexpr int kGrad = 18;     // per (triangle, tile) record: 3 vertices x (pos 3 + colour 3)
constexpr int kXcd = 8;       // L2 domains the workgroups are dealt round-robin to
constexpr int kFast = 256;    // tiles with at most this many triangles stage their list in LDS
// Tiles with longer lists (or past the record store) take the fine kernel's
// slow path, several times longer than a typical tile: the tile scan lists up
// to kHeavyCap of them and the first kHeavyCap workgroups of the launch take
// them, so they start first instead of ending the launch late.
constexpr int kHeavyCap = 4096;

struct CharD {
  int N, F, M, J;
  int v0, f0, m0, j0;  // global offsets
  int view0, C;
  int has_tc;          // texture coordinates given (the UV buffer is written)
};

struct ViewD {
  int ch, W, H, S;
  int tiles_x, tiles;
  long long vn0, vf0, pix0, sil0, tile0, cell0;
  long long mask0;  // first word of the view's background mask
  int mpitch;       // words per mask row: 2 * ceil(W / 64), so a row holds every tile shape's columns
  int gx, gy;       // silhouette grid
};

struct Dev {
  int B, V;
  int maxN, maxF, maxM, maxS2, maxJ, maxTiles;
  int tws, ths;           // tile shape: (1 << tws) x (1 << ths) pixels, 256 in all
  int gxN, gxS;           // blocks per (view|char) of the vertex kernels / DT-pixel tiles per view
  long long totVN, totVF, totPix, totTiles, totCells;
  const CharD* ch;
  const ViewD* vw;
  // static mesh data (global arrays, local indices)
  const float* rest;      // [sumN,3]
  const int* faces;       // [sumF,3] local vertex ids
  const int* jidx;        // [sumN,4] local joint ids
  const float* jw;        // [sumN,4]
  const int* enode;       // [sumN,4] local node ids
  const float* ew;        // [sumN,4]
  const int* node_vtx;    // [sumM] local vertex id
  const int* n2v_off;     // [sumM+1] global offsets into n2v arrays
  const int* n2v_vtx;     // local vertex ids
  const float* n2v_w;
  const int* vf_off;      // [sumN+1] global offsets
  const int* vf_list;     // local face ids, increasing per vertex (one entry per face corner)
  const int* corner_slot; // [sumF*3] global vf_list slot of each face corner
  const int* jv_off;      // [sumJ+1] global offsets: joint -> (vertex, slot) entries
  const int* jv_ent;      // local vertex * 4 + slot, increasing
  const int4* jchunk;     // [nChunk] (first entry, end entry, character vertex base, joint)
  const int* jc_off;      // [sumJ+1] joint -> its chunks
  const float* tc;        // [sumF,6] texture coordinates (0 for characters without)
  // scratch
  float* node_tab;        // [sumM,15]
  float* vdef;            // [sumN,3] skinned vertices
  float* vnrm;            // [sumN,3] unit vertex normals
  float* vsum;            // [sumN,3] the unnormalised sums of the face normals (as CudaRendererGpu's)
  f4* proj;               // [totVN]
  float* ke;              // [V,12] per view: K * R (3x3 row-major), then K * t
  f2* nrm2;               // [totVN]
  TriRec* tri;            // [totVF]
  uint2* trange;          // [totVF] tile range of the current shape: (tx0 | ty0 << 16, ntx | nty << 16),
                          // ntx = nty = 0 for a culled triangle (the binning reads these 8 bytes, not
                          // the 48-byte record)
  unsigned int* tri_blk;  // [V * ceil(maxF / kB)] tiles touched per binning workgroup, then their
                          // exclusive scan (the workgroup's first record slot)
  unsigned int* tri_base; // [totVF+1] first record slot per (view, face)
  unsigned int* tile_cnt; // [totTiles]
  unsigned int* tile_off; // [totTiles+1]
  unsigned int* tile_fill;
  unsigned int* bins;     // [capacity] tile-major triangle lists
  float* rec;             // [capacity,18] (triangle, tile) gradient records
  unsigned char* rec_vis; // [capacity] 1 if the triangle won a pixel of the tile
  unsigned int capacity;
  unsigned char* isb;     // [totVN]
  unsigned int* fgmask;   // [totMask] per view H rows of mpitch words: bit x of row y set <=> pixel (x, y)
                          // has a face (the boundary test's 3x3 reads stay in a mask 1/32 of the face
                          // buffer's size)
  long long totMask;
  // nearest-point grid, one of two layouts with the same search results:
  // cell_cap == 0, counting sort: cell_pts [totVN] grouped by cell at
  // cell_off; cell_cap > 0, fixed slots: a cell's first cell_cap points in
  // cell_slots [totCells * cell_cap] in arrival order, later ones in its
  // view's overflow list (ov_pts from vn0, ov_cnt).  Points are (x, y, point
  // id bits, 0).
  int cell_cap;
  unsigned int* cell_cnt; // [totCells] boundary points that fell in each cell
  unsigned int* cell_off; // [totCells+1]
  unsigned int* cell_fill;
  f4* cell_pts;
  f4* cell_slots;
  f4* ov_pts;
  unsigned int* ov_cnt;   // [V]
  unsigned int* scan_aux;
  unsigned int* heavy;     // [kHeavyCap] global indices of the slow-path tiles (any order)
  unsigned int* heavy_n;   // number of slow-path tiles (may exceed kHeavyCap: then none is listed)
  float* corner_sum;      // [sumF*3,6] raster gradient per face corner (position 3, colour 3), summed over
                          // tiles and views, stored at the corner's vf_list slot
  float* g_v;             // [sumN,3] position gradient from records beyond capacity
  float* g_pts;           // [totVN,2] model->data gradient w.r.t. the projected points
  long long* g_fx;        // [totVN,2] data->model residuals summed per point, fixed point (units of 2^-40)
  float* u;               // [sumN,3] gradient w.r.t. the ED-deformed position
  float* loss_tile;       // [totTiles] colour loss per tile
  float* loss_pt;         // [V*gxN] silhouette model->data loss per block
  float* loss_px;         // [V*gxS] silhouette data->model loss per block
  float* vq;              // [sumN,8] gradient w.r.t. the vertex's blended, normalised DQ
  float* jpart;           // [nChunk,8] per-chunk joint DQ gradient partials
  float* vws;             // [sumN,4] skinning weight x blend sign per slot
  float* node_skin;       // [sumM,12] skinned node rotation (row-major) and translation (graph skinning)
};

struct IO {
  const float* dq;        // [sumJ,8]
  const float* dT;        // [sumM,3]
  const float* dA;        // [sumM,3]
  const float* col;       // [sumN,3]
  const float* ext;       // [V,12]
  const float* intr;      // [V,9]
  const float* target;    // [totPix,3]
  const unsigned char* dt;  // [sum S^2]
  const float* crops;     // [V,7]
  float w_col, w_sil;
  int* face;
  float* bary;
  float* render;
  float* vout;            // [sumN,3] (may be null)
  float* loss;            // [B,2] colour, silhouette
  float* g_dq;
  float* g_dT;
  float* g_dA;
  float* g_col;
  float* uv;              // [totPix,2] (may be null)
  float* nrm_out;         // [totPix,3] (may be null)
  const float* skT;       // [sumM,3] skinned node translations, or null: dual-quaternion skinning
  const float* skA;       // [sumM,3] skinned node Euler angles (null with skT)
};

struct ClearList {
  unsigned int* p[8];
  long long n[8];
  int count;
};

// ---------------------------------------------------------------------------
// Exclusive scan of any length: 1024-element block scans, a recursive scan of
// the block totals, add-back; out[n] receives the total.
// ---------------------------------------------------------------------------

// Inclusive scan over a wavefront whose 64 lanes are all active: DPP row
// shifts by 1, 2, 4 and 8 scan each row of 16 lanes (a lane with no source
// in its row adds 0), then the end of row 0 is added into row 1 and the end
// of row 2 into row 3 (row_bcast:15), and the end of row 1 into rows 2 and 3
// (row_bcast:31).  Integer sums, so the result is exact.
__device__ __forceinline__ unsigned int wave_incl_scan(unsigned int v) {
  v += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);   // row_shr:1
  v += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);   // row_shr:2
  v += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);   // row_shr:4
  v += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);   // row_shr:8
  v += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
  v += (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
  return v;
}

// Exclusive scan over a 256-lane workgroup, all lanes active (every caller
// runs kB lanes and calls it from all of them); total receives the sum.
__device__ __forceinline__ unsigned int block_scan256(unsigned int v, unsigned int* s_wave, unsigned int& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const unsigned int inc = wave_incl_scan(v);
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  unsigned int base = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < kB / 64; ++w) {
    if (w < wid) base += s_wave[w];
    total += s_wave[w];
  }
  return base + inc - v;
}

// Exclusive scan of n counts into out[0..n] (out[n] = total) in two launches
// over chunks of kScanChunk entries (16 per lane): pass 1 writes each chunk's
// sum; pass 2 adds up the sums of the chunks before its own (at most a few
// thousand), scans its chunk and writes it.  No workgroup waits on another.
constexpr int kScanPer = 16;
constexpr int kScanChunk = kB * kScanPer;

__device__ __forceinline__ void scan_load16(const unsigned int* in, long long base, long long n, unsigned int* v) {
  if (base + kScanPer <= n) {
    const uint4* q = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int k = 0; k < kScanPer / 4; ++k) {
      const uint4 t = q[k];
      v[4 * k] = t.x; v[4 * k + 1] = t.y; v[4 * k + 2] = t.z; v[4 * k + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) v[k] = base + k < n ? in[base + k] : 0u;
  }
}

__device__ __forceinline__ void scan_sum_body(const unsigned int* in, unsigned int* sums, long long n, int block) {
  __shared__ unsigned int s_wave[kB / 64];
  unsigned int v[kScanPer], t = 0;
  scan_load16(in, (long long)block * kScanChunk + (long long)threadIdx.x * kScanPer, n, v);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) t += v[k];
  unsigned int total;
  (void)block_scan256(t, s_wave, total);
  if (threadIdx.x == 0) sums[block] = total;
}

__global__ __launch_bounds__(kB) void scan_sum_kernel(const unsigned int* in, unsigned int* sums, long long n) {
  scan_sum_body(in, sums, n, blockIdx.x);
}

// heavy != nullptr (the tile scan): also lists the tiles whose count exceeds
// kFast or whose list ends past `capacity` record slots.
__device__ __forceinline__ void scan_apply_body(const unsigned int* in, unsigned int* out, const unsigned int* sums,
                                                long long n, unsigned int* heavy, unsigned int* heavy_n,
                                                unsigned int capacity) {
  __shared__ unsigned int s_wave[kB / 64];
  unsigned int pre = 0;
  for (unsigned int i = threadIdx.x; i < blockIdx.x; i += kB) pre += sums[i];
  unsigned int offset;
  (void)block_scan256(pre, s_wave, offset);
  __syncthreads();  // s_wave is reused below
  const long long base = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanPer;
  unsigned int v[kScanPer], t = 0;
  scan_load16(in, base, n, v);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) t += v[k];
  unsigned int chunk;
  unsigned int ex = offset + block_scan256(t, s_wave, chunk);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) {
      out[base + k] = ex;
      if (heavy && (v[k] > (unsigned int)kFast || ex + v[k] > capacity)) {
        const unsigned int h = atomicAdd(heavy_n, 1u);
        if (h < (unsigned int)kHeavyCap) heavy[h] = (unsigned int)(base + k);
      }
    }
    ex += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = offset + chunk;
}

__global__ __launch_bounds__(kB) void scan_apply_kernel(const unsigned int* in, unsigned int* out,
                                                        const unsigned int* sums, long long n) {
  scan_apply_body(in, out, sums, n, nullptr, nullptr, 0u);
}

__global__ __launch_bounds__(kB) void tile_scan_apply_kernel(const unsigned int* in, unsigned int* out,
                                                             const unsigned int* sums, long long n,
                                                             unsigned int* heavy, unsigned int* heavy_n,
                                                             unsigned int capacity) {
  scan_apply_body(in, out, sums, n, heavy, heavy_n, capacity);
}

// In-place exclusive scan of a[0..n) by one workgroup, kOnePer entries per
// lane per pass (one pass up to 32768 entries: one launch, its loads all in
// flight together); *total = the sum.
constexpr int kOnePer = 128;
__device__ __forceinline__ void scan_one_body(unsigned int* a, long long n, unsigned int* total) {
  __shared__ unsigned int s_wave[kB / 64];
  unsigned int carry = 0;
  for (long long b = 0; b < n; b += (long long)kB * kOnePer) {
    const long long base = b + (long long)threadIdx.x * kOnePer;
    unsigned int v[kOnePer], t = 0;
#pragma unroll
    for (int q = 0; q < kOnePer / kScanPer; ++q) scan_load16(a, base + q * kScanPer, n, v + q * kScanPer);
#pragma unroll
    for (int k = 0; k < kOnePer; ++k) t += v[k];
    unsigned int chunk;
    unsigned int ex = carry + block_scan256(t, s_wave, chunk);  // each lane writes only the entries it read
    if (base + kOnePer <= n) {
      uint4* q = reinterpret_cast<uint4*>(a + base);
#pragma unroll
      for (int k = 0; k < kOnePer / 4; ++k) {
        uint4 o;
        o.x = ex; ex += v[4 * k];
        o.y = ex; ex += v[4 * k + 1];
        o.z = ex; ex += v[4 * k + 2];
        o.w = ex; ex += v[4 * k + 3];
        q[k] = o;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kOnePer; ++k) {
        if (base + k < n) a[base + k] = ex;
        ex += v[k];
      }
    }
    carry += chunk;
    __syncthreads();  // s_wave is reused by the next pass
  }
  if (threadIdx.x == 0) *total = carry;
}

// The chunk sums of one scan plus, in one more workgroup, the single-group
// scan of another array made by the same producer: one launch for both.
__global__ __launch_bounds__(kB) void scan_sum_with_one_kernel(const unsigned int* in, unsigned int* sums, long long n,
                                                               unsigned int* a1, long long n1, unsigned int* total1) {
  if (blockIdx.x == gridDim.x - 1) {  // uniform over the block
    scan_one_body(a1, n1, total1);
    return;
  }
  scan_sum_body(in, sums, n, blockIdx.x);
}

// Scratch elements scan_big needs for n inputs (the chunk sums).
long long scan_aux_elems(long long n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

dcx_status scan_big(const unsigned int* in, unsigned int* out, unsigned int* aux, long long n, hipStream_t s) {
  const long long nb = (n + kScanChunk - 1) / kScanChunk;
  if (nb < 1 || nb >= (1ll << 31)) return set_error(DCX_ERR_INVALID, "scan: bad length %lld", n);
  hipLaunchKernelGGL(scan_sum_kernel, dim3((unsigned)nb), dim3(kB), 0, s, in, aux, n);
  hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)nb), dim3(kB), 0, s, in, out, aux, n);
  return check_launch("scan");
}

// ---------------------------------------------------------------------------
// Forward: ED, skinning, normals, projection
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kB) void node_fwd_kernel(Dev d, IO io) {
  const CharD c = d.ch[blockIdx.y];
  const int n = blockIdx.x * kB + threadIdx.x;
  if (n >= c.M) return;
  const int gi = c.m0 + n;
  const m3 R = euler_mat(euler_sc(ld3(io.dA + 3 * (size_t)gi)));
  const f3 g = ld3(d.rest + 3 * ((size_t)c.v0 + d.node_vtx[gi]));
  const f3 t = ld3(io.dT + 3 * (size_t)gi);
  float* row = d.node_tab + (size_t)gi * kNodeRow;
  for (int k = 0; k < 9; ++k) row[k] = R.m[k];
  row[9] = g.x; row[10] = g.y; row[11] = g.z;
  row[12] = t.x; row[13] = t.y; row[14] = t.z;
  if (io.skA) {
    // the node's skinned transform (EmbeddedGraphGPUOp.cu:100-114 evaluates
    // it per vertex and node; here once per node)
    const m3 Rs = euler_mat(euler_sc(ld3(io.skA + 3 * (size_t)gi)));
    const f3 ts = ld3(io.skT + 3 * (size_t)gi);
    float* sk = d.node_skin + (size_t)gi * 12;
    for (int k = 0; k < 9; ++k) sk[k] = Rs.m[k];
    sk[9] = ts.x; sk[10] = ts.y; sk[11] = ts.z;
  }
}

__device__ __forceinline__ f3 ed_inner(const Dev& d, const CharD& c, int gv) {
  const f3 p = ld3(d.rest + 3 * (size_t)gv);
  f3 v = mk3(0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < KE; ++k) {
    const float* r = d.node_tab + ((size_t)c.m0 + d.enode[(size_t)gv * KE + k]) * kNodeRow;
    const f3 g = ld3(r + 9);
    m3 R;
    for (int q = 0; q < 9; ++q) R.m[q] = r[q];
    v += d.ew[(size_t)gv * KE + k] * (mul(R, p - g) + g + ld3(r + 12));
  }
  return v;
}

// The embedded graph's outer sum with the skinned node transforms
// (EmbeddedGraphGPUOp.cu:100-114): sum over the vertex's nodes s of
// w_s (R_s v_nr + t_s), in node order.
__device__ __forceinline__ f3 ed_outer(const Dev& d, const CharD& c, int gv, f3 vnr) {
  f3 o = mk3(0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < KE; ++k) {
    const float* r = d.node_skin + ((size_t)c.m0 + d.enode[(size_t)gv * KE + k]) * 12;
    m3 R;
    for (int q = 0; q < 9; ++q) R.m[q] = r[q];
    o += d.ew[(size_t)gv * KE + k] * (mul(R, vnr) + ld3(r + 9));
  }
  return o;
}

// Also clears the step's accumulators and counters (grid-stride over the
// whole grid, before any early return): the first of them is read by the
// binning that follows, so the clear needs no launch of its own.
__global__ __launch_bounds__(kB) void vertex_fwd_kernel(Dev d, IO io, ClearList L) {
  __shared__ float s_tab[256 * 8];
  {
    const long long stride = (long long)gridDim.x * gridDim.y * kB;
    const long long i0 = ((long long)blockIdx.y * gridDim.x + blockIdx.x) * kB + threadIdx.x;
    for (int k = 0; k < L.count; ++k)
      for (long long i = i0; i < L.n[k]; i += stride) L.p[k][i] = 0u;
  }
  const CharD c = d.ch[blockIdx.y];
  if ((long long)blockIdx.x * kB >= c.N) return;  // no vertices here (uniform over the block)
  if (!io.skA)
    for (int i = threadIdx.x; i < c.J * 8; i += kB) s_tab[i] = io.dq[(size_t)c.j0 * 8 + i];
  __syncthreads();
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= c.N) return;
  const int gv = c.v0 + v;
  const f3 vnr = ed_inner(d, c, gv);
  if (io.skA) {
    const f3 o = ed_outer(d, c, gv, vnr);
    d.vdef[3 * (size_t)gv] = o.x; d.vdef[3 * (size_t)gv + 1] = o.y; d.vdef[3 * (size_t)gv + 2] = o.z;
    if (io.vout) { io.vout[3 * (size_t)gv] = o.x; io.vout[3 * (size_t)gv + 1] = o.y; io.vout[3 * (size_t)gv + 2] = o.z; }
    return;
  }
  int idx[KS];
  float w[KS];
  for (int k = 0; k < KS; ++k) { idx[k] = d.jidx[(size_t)gv * KS + k]; w[k] = d.jw[(size_t)gv * KS + k]; }
  const DqBlend bl = dq_blend(s_tab, idx, w, KS);
  const f3 o = mul(quat_to_mat(bl.r), vnr) + dq_translation(bl.r, bl.d);
  d.vdef[3 * (size_t)gv] = o.x; d.vdef[3 * (size_t)gv + 1] = o.y; d.vdef[3 * (size_t)gv + 2] = o.z;
  if (io.vout) { io.vout[3 * (size_t)gv] = o.x; io.vout[3 * (size_t)gv + 1] = o.y; io.vout[3 * (size_t)gv + 2] = o.z; }
}

__global__ __launch_bounds__(kB) void normals_kernel(Dev d) {
  const CharD c = d.ch[blockIdx.y];
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= c.N) return;
  const int gv = c.v0 + v;
  const float* vb = d.vdef + 3 * (size_t)c.v0;
  f3 s = mk3(0.f, 0.f, 0.f);
  for (int k = d.vf_off[gv]; k < d.vf_off[gv + 1]; ++k) {
    const int* fv = d.faces + 3 * ((size_t)c.f0 + d.vf_list[k]);
    const f3 a = ld3(vb + 3 * fv[0]), b1 = ld3(vb + 3 * fv[1]), b2 = ld3(vb + 3 * fv[2]);
    s += cross(b1 - a, b2 - a);
  }
  const float l = sqrtf(dot(s, s));
  const f3 n = l > 0.f ? mk3(s.x / l, s.y / l, s.z / l) : mk3(0.f, 0.f, 0.f);
  d.vnrm[3 * (size_t)gv] = n.x; d.vnrm[3 * (size_t)gv + 1] = n.y; d.vnrm[3 * (size_t)gv + 2] = n.z;
  d.vsum[3 * (size_t)gv] = s.x; d.vsum[3 * (size_t)gv + 1] = s.y; d.vsum[3 * (size_t)gv + 2] = s.z;
}

__device__ __forceinline__ f2 proj2(const float* K, f3 x) {
  const float h0 = K[0] * x.x + K[1] * x.y + K[2] * x.z;
  const float h1 = K[3] * x.x + K[4] * x.y + K[5] * x.z;
  float h2 = K[6] * x.x + K[7] * x.y + K[8] * x.z;
  if (!(h2 > 0.0000001f)) h2 = 0.00001f;
  return mk2(h0 / h2, h1 / h2);
}

// Projected vertices and K [R | t] per view, for the fine kernel and the
// silhouette stages (the binning projects its corners itself).
__global__ __launch_bounds__(kB) void project_kernel(Dev d, IO io) {
  const int w = blockIdx.y;
  const ViewD vw = d.vw[w];
  const CharD c = d.ch[vw.ch];
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= c.N) return;
  const float* E = io.ext + (size_t)w * 12;
  const float* K = io.intr + (size_t)w * 9;
  const int gv = c.v0 + v;
  if (v == 0) {
    // K [R | t] for the gradient kernels (one lane per view)
    float* o = d.ke + (size_t)w * 12;
    for (int r = 0; r < 3; ++r) {
      for (int cc = 0; cc < 3; ++cc) o[3 * r + cc] = K[3 * r] * E[cc] + K[3 * r + 1] * E[4 + cc] + K[3 * r + 2] * E[8 + cc];
      o[9 + r] = K[3 * r] * E[3] + K[3 * r + 1] * E[7] + K[3 * r + 2] * E[11];
    }
  }
  d.proj[vw.vn0 + v] = project_vertex(E, K, d.vdef + 3 * (size_t)gv);
}

// Image-space normal direction per (view, vertex) for the silhouette
// (CameraProjectionGPUOp.cu:45-86); only the model->data residual reads it,
// so it runs beside the binning stages.
__global__ __launch_bounds__(kB) void view_normal_kernel(Dev d, IO io) {
  const int w = blockIdx.y;
  const ViewD vw = d.vw[w];
  const CharD c = d.ch[vw.ch];
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= c.N) return;
  const float* E = io.ext + (size_t)w * 12;
  const float* K = io.intr + (size_t)w * 9;
  const int gv = c.v0 + v;
  const f3 X = ld3(d.vdef + 3 * (size_t)gv);
  const f3 n = ld3(d.vnrm + 3 * (size_t)gv);
  const f3 xc = mk3(E[0] * X.x + E[1] * X.y + E[2] * X.z + E[3], E[4] * X.x + E[5] * X.y + E[6] * X.z + E[7],
                    E[8] * X.x + E[9] * X.y + E[10] * X.z + E[11]);
  const f3 x2 = mk3(xc.x + (E[0] * n.x + E[1] * n.y + E[2] * n.z), xc.y + (E[4] * n.x + E[5] * n.y + E[6] * n.z),
                    xc.z + (E[8] * n.x + E[9] * n.y + E[10] * n.z));
  const f2 a = proj2(K, xc), b = proj2(K, x2);
  const float du = b.x - a.x, dv = b.y - a.y;
  const float l = sqrtf(du * du + dv * dv);
  d.nrm2[vw.vn0 + v] = l > 0.f ? mk2(du / l, dv / l) : mk2(0.f, 0.f);
}

// ---------------------------------------------------------------------------
// Binning
// ---------------------------------------------------------------------------

struct TileRange { int tx0, ty0, ntx, nty; };
__device__ __forceinline__ TileRange tile_range(const TriRec& t, int tws, int ths) {
  TileRange r;
  const int x0 = t.bx0_by0 & 0xffff, y0 = t.bx0_by0 >> 16, x1 = t.bx1_by1 & 0xffff, y1 = t.bx1_by1 >> 16;
  r.tx0 = x0 >> tws; r.ty0 = y0 >> ths;
  r.ntx = (x1 >> tws) - r.tx0 + 1; r.nty = (y1 >> ths) - r.ty0 + 1;
  return r;
}

// Per-workgroup aggregation of (triangle, tile) pairs: an LDS open-addressing
// table keyed by the tile, so the workgroup issues one device-scope atomic per
// distinct tile instead of one per pair (device-scope atomics are executed in
// memory, not in an XCD's L2).  Pairs of large triangles, and pairs whose
// tile finds no free slot within kProbe probes, go straight to memory; every
// thread probing a given tile follows the same sequence, so all of them agree
// on whether and where that tile sits in the table.
constexpr int kHash = 1024;
constexpr int kProbe = 8;
constexpr int kBigTri = 32;  // tiles; larger triangles bypass the table
constexpr unsigned int kHashEmpty = 0xffffffffu;

__device__ __forceinline__ int hash_find(unsigned int* keys, unsigned int tile, bool insert) {
  unsigned int h = (tile * 2654435761u) >> 22;  // 10 bits
#pragma unroll 1
  for (int pr = 0; pr < kProbe; ++pr, h = (h + 1) & (kHash - 1)) {
    const unsigned int cur = insert ? atomicCAS(keys + h, kHashEmpty, tile) : keys[h];
    if (cur == tile || (insert && cur == kHashEmpty)) return (int)h;
    if (!insert && cur == kHashEmpty) return -1;
  }
  return -1;
}

__global__ __launch_bounds__(kB) void tri_count_kernel(Dev d, IO io) {
  __shared__ unsigned int s_keys[kHash];
  __shared__ unsigned int s_cnt[kHash];
  __shared__ unsigned int s_wave[kB / 64];
  const int w = blockIdx.y;
  const ViewD vw = d.vw[w];
  const CharD c = d.ch[vw.ch];
  // the grid is sized for the largest character: workgroups past this one's
  // faces (uniform over the block) only record that they need no slots
  if ((long long)blockIdx.x * kB >= c.F) {
    if (threadIdx.x == 0) d.tri_blk[(size_t)w * gridDim.x + blockIdx.x] = 0u;
    return;
  }
  unsigned int my_cnt = 0;
  for (int i = threadIdx.x; i < kHash; i += kB) { s_keys[i] = kHashEmpty; s_cnt[i] = 0u; }
  __syncthreads();
  const int f = blockIdx.x * kB + threadIdx.x;
  unsigned int* cnt = d.tile_cnt + vw.tile0;
  if (f < c.F) {
    const int* fv = d.faces + 3 * ((size_t)c.f0 + f);
    // the corners' projections, the same operations as project_kernel's (which
    // runs beside this kernel for the later stages)
    const float* E = io.ext + (size_t)w * 12;
    const float* K = io.intr + (size_t)w * 9;
    const float* vb = d.vdef + 3 * (size_t)c.v0;
    const f4 p0 = project_vertex(E, K, vb + 3 * fv[0]), p1 = project_vertex(E, K, vb + 3 * fv[1]);
    const f4 p2 = project_vertex(E, K, vb + 3 * fv[2]);
    TriRec t;
    const long long ri = vw.vf0 + f;
    if (!setup_tri(p0, p1, p2, f, vw.W, vw.H, t)) {
      t.bx0_by0 = 0x7fff7fff;
      t.bx1_by1 = 0;
      d.tri[ri] = t;
      d.trange[ri] = make_uint2(0u, 0u);
    } else {
      d.tri[ri] = t;
      const TileRange r = tile_range(t, d.tws, d.ths);
      d.trange[ri] = make_uint2((unsigned int)r.tx0 | ((unsigned int)r.ty0 << 16),
                                (unsigned int)r.ntx | ((unsigned int)r.nty << 16));
      my_cnt = (unsigned int)(r.ntx * r.nty);
      const bool big = r.ntx * r.nty > kBigTri;
      for (int ty = r.ty0; ty < r.ty0 + r.nty; ++ty)
        for (int tx = r.tx0; tx < r.tx0 + r.ntx; ++tx) {
          const unsigned int lt = (unsigned int)(ty * vw.tiles_x + tx);
          const int h = big ? -1 : hash_find(s_keys, lt, true);
          if (h >= 0) atomicAdd(s_cnt + h, 1u);
          else atomicAdd(cnt + lt, 1u);
        }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHash; i += kB)
    if (s_keys[i] != kHashEmpty) atomicAdd(cnt + s_keys[i], s_cnt[i]);
  // record slots this workgroup's faces need (scanned over the workgroups,
  // which run in (view, face) order, by the single-group record-slot scan)
  unsigned int tot;
  (void)block_scan256(my_cnt, s_wave, tot);
  if (threadIdx.x == 0) d.tri_blk[(size_t)w * gridDim.x + blockIdx.x] = tot;
}

__global__ __launch_bounds__(kB) void bin_scatter_kernel(Dev d) {
  __shared__ unsigned int s_keys[kHash];
  __shared__ unsigned int s_cnt[kHash];
  __shared__ unsigned int s_base[kHash];
  __shared__ unsigned int s_wave[kB / 64];
  const int w = blockIdx.y;
  const ViewD vw = d.vw[w];
  const CharD c = d.ch[vw.ch];
  if ((long long)blockIdx.x * kB >= c.F) return;  // no faces here (uniform over the block)
  for (int i = threadIdx.x; i < kHash; i += kB) { s_keys[i] = kHashEmpty; s_cnt[i] = 0u; }
  __syncthreads();
  const int f = blockIdx.x * kB + threadIdx.x;
  const long long ri = vw.vf0 + f;
  TileRange r{0, 0, 0, 0};
  bool big = false;
  if (f < c.F) {
    const uint2 tr = d.trange[ri];
    r = TileRange{(int)(tr.x & 0xffffu), (int)(tr.x >> 16), (int)(tr.y & 0xffffu), (int)(tr.y >> 16)};
    big = r.ntx * r.nty > kBigTri;
  }
  // first record slot of each face: the workgroup's offset plus the scan of
  // its faces' tile counts (the same counts tri_count summed)
  {
    unsigned int tot;
    const unsigned int ex = block_scan256((unsigned int)(r.ntx * r.nty), s_wave, tot);
    if (f < c.F) d.tri_base[ri] = d.tri_blk[(size_t)w * gridDim.x + blockIdx.x] + ex;
  }
  // pass 1: per-tile counts of this workgroup's pairs
  if (!big)
    for (int ty = r.ty0; ty < r.ty0 + r.nty; ++ty)
      for (int tx = r.tx0; tx < r.tx0 + r.ntx; ++tx) {
        const int h = hash_find(s_keys, (unsigned int)(ty * vw.tiles_x + tx), true);
        if (h >= 0) atomicAdd(s_cnt + h, 1u);
      }
  __syncthreads();
  // one reservation per distinct tile; a lane's (at most kHash / kB) memory
  // atomics are all issued before it waits for any of them
  {
    static_assert(kHash % kB == 0, "each lane owns kHash / kB slots");
    constexpr int kPer = kHash / kB;
    unsigned int key[kPer], base[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) key[j] = s_keys[threadIdx.x + j * kB];
#pragma unroll
    for (int j = 0; j < kPer; ++j)
      if (key[j] != kHashEmpty) base[j] = atomicAdd(d.tile_fill + vw.tile0 + key[j], s_cnt[threadIdx.x + j * kB]);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (key[j] != kHashEmpty) s_base[threadIdx.x + j * kB] = base[j];
      s_cnt[threadIdx.x + j * kB] = 0u;
    }
  }
  __syncthreads();
  // pass 2: slots
  for (int ty = r.ty0; ty < r.ty0 + r.nty; ++ty)
    for (int tx = r.tx0; tx < r.tx0 + r.ntx; ++tx) {
      const unsigned int lt = (unsigned int)(ty * vw.tiles_x + tx);
      const long long tile = vw.tile0 + lt;
      const unsigned int toff = d.tile_off[tile];  // in flight during the hash probe and the rank atomic
      const int h = big ? -1 : hash_find(s_keys, lt, false);
      const unsigned int rank = h >= 0 ? s_base[h] + atomicAdd(s_cnt + h, 1u) : atomicAdd(d.tile_fill + tile, 1u);
      const unsigned int slot = toff + rank;
      if (slot < d.capacity) d.bins[slot] = (unsigned int)ri;
    }
}

// ---------------------------------------------------------------------------
// Fine tile kernel
// ---------------------------------------------------------------------------

// Pixels of one triangle inside one tile into the LDS depth/id keys.  The
// edge functions are exact integers either way; for triangles whose snapped
// extent is at most kSmallSpan sub-pixels on both axes they are walked in
// 32-bit arithmetic: at any pixel centre of the bounding box (plus the one
// step past its end) |e| <= 2 * 2^14 * (2^14 + 2^9) < 2^31, and converting
// the same integer to float from 32 or 64 bits rounds identically, so the
// keys are bit-identical to the 64-bit walk.
constexpr int kSmallSpan = 1 << 14;  // 64 px

__device__ __forceinline__ float interp_iz32(int e0, int e1, int e2, const float* iz) {
  const float w0 = (float)e0, w1 = (float)e1, w2 = (float)e2;
  return w0 * iz[0] + w1 * iz[1] + w2 * iz[2];
}

__device__ __forceinline__ void raster_in_tile(const TriRec& t, int tx0, int ty0, int tws, int ths,
                                               unsigned int tag, unsigned long long* s_key) {
  const int x0 = max(t.bx0_by0 & 0xffff, tx0), y0 = max(t.bx0_by0 >> 16, ty0);
  const int x1 = min(t.bx1_by1 & 0xffff, tx0 + (1 << tws) - 1), y1 = min(t.bx1_by1 >> 16, ty0 + (1 << ths) - 1);
  if (x0 > x1 || y0 > y1) return;
  const int spx = max(t.X[0], max(t.X[1], t.X[2])) - min(t.X[0], min(t.X[1], t.X[2]));
  const int spy = max(t.Y[0], max(t.Y[1], t.Y[2])) - min(t.Y[0], min(t.Y[1], t.Y[2]));
  EdgeWalk w;
  w.init(t, x0, y0);
  // the record's 1/z in registers: the loops below would otherwise load it
  // again after every key update
  const float izv[3] = {t.iz[0], t.iz[1], t.iz[2]};
  if (spx <= kSmallSpan && spy <= kSmallSpan) {
    int row[3], dx[3], dy[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      row[i] = (int)w.e[i] + w.bias[i];  // bias folded in: inside <=> e + bias >= 0
      dx[i] = (int)w.dx[i];
      dy[i] = (int)w.dy[i];
    }
    for (int y = y0; y <= y1; ++y) {
      int e[3] = {row[0], row[1], row[2]};
      for (int x = x0; x <= x1; ++x) {
        if (e[0] >= 0 && e[1] >= 0 && e[2] >= 0)
          atomicMin(s_key + (((y - ty0) << tws) | (x - tx0)),
                    make_key(interp_iz32(e[0] - w.bias[0], e[1] - w.bias[1], e[2] - w.bias[2], izv), (int)tag));
        e[0] += dx[0]; e[1] += dx[1]; e[2] += dx[2];
      }
      row[0] += dy[0]; row[1] += dy[1]; row[2] += dy[2];
    }
    return;
  }
  long long row[3] = {w.e[0], w.e[1], w.e[2]};
  for (int y = y0; y <= y1; ++y) {
    long long e[3] = {row[0], row[1], row[2]};
    for (int x = x0; x <= x1; ++x) {
      if (w.inside(e))
        atomicMin(s_key + (((y - ty0) << tws) | (x - tx0)), make_key(interp_iz(e[0], e[1], e[2], izv), (int)tag));
      e[0] += w.dx[0]; e[1] += w.dx[1]; e[2] += w.dx[2];
    }
    row[0] += w.dy[0]; row[1] += w.dy[1]; row[2] += w.dy[2];
  }
}

// Barycentric VJP: gradient of (b0, b1) (third weight 1 - b0 - b1) w.r.t. the
// three world-space vertices through the projection (same derivation as the
// standalone rasterizer backward).
__device__ void bary_vjp_p(const float* E, const float* K, const f3 X[3], float px, float py, const float gb_in[3],
                           f3 out[3]) {
  float h2[3], u[3], v[3], iz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float xc = E[0] * X[i].x + E[1] * X[i].y + E[2] * X[i].z + E[3];
    const float yc = E[4] * X[i].x + E[5] * X[i].y + E[6] * X[i].z + E[7];
    const float zc = E[8] * X[i].x + E[9] * X[i].y + E[10] * X[i].z + E[11];
    const float h0 = K[0] * xc + K[1] * yc + K[2] * zc;
    const float h1 = K[3] * xc + K[4] * yc + K[5] * zc;
    h2[i] = K[6] * xc + K[7] * yc + K[8] * zc;
    u[i] = h0 / h2[i];
    v[i] = h1 / h2[i];
    iz[i] = 1.f / h2[i];
  }
  float e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
    e[i] = (u[k] - u[j]) * (py - v[j]) - (v[k] - v[j]) * (px - u[j]);
  }
  const float q0 = e[0] * iz[0], q1 = e[1] * iz[1], q2 = e[2] * iz[2];
  const float Q = q0 + q1 + q2;
  const float b0 = q0 / Q, b1 = q1 / Q;
  const float g0 = gb_in[0] - gb_in[2], g1 = gb_in[1] - gb_in[2];
  const float s = g0 * b0 + g1 * b1;
  const float gq[3] = {(g0 - s) / Q, (g1 - s) / Q, (0.f - s) / Q};
  float gu[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f}, giz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    giz[i] = gq[i] * e[i];
    const float ge = gq[i] * iz[i];
    const int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
    gu[j] += ge * (v[k] - py);
    gv[j] += ge * (px - u[k]);
    gu[k] += ge * (py - v[j]);
    gv[k] += ge * (u[j] - px);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float gh0 = gu[i] / h2[i], gh1 = gv[i] / h2[i];
    const float gh2 = -(gu[i] * u[i] + gv[i] * v[i]) / h2[i] - giz[i] * iz[i] * iz[i];
    const float gx = K[0] * gh0 + K[3] * gh1 + K[6] * gh2;
    const float gy = K[1] * gh0 + K[4] * gh1 + K[7] * gh2;
    const float gz = K[2] * gh0 + K[5] * gh1 + K[8] * gh2;
    out[i] = mk3(E[0] * gx + E[4] * gy + E[8] * gz, E[1] * gx + E[5] * gy + E[9] * gz,
                 E[2] * gx + E[6] * gy + E[10] * gz);
  }
}

// The same VJP from the stored projections P_i = (u_i, v_i, z_i, 1/z_i) of
// the three vertices and M = K R of the view, with hardware reciprocals in
// place of divisions (gradients carry a tolerance; the buffers the forward
// writes do not go through here).
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ void bary_vjp_fast(const f4 P[3], const float* M, float px, float py, const float gb_in[3],
                                              f3 out[3]) {
#pragma clang fp contract(fast)
  float e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
    e[i] = (P[k].x - P[j].x) * (py - P[j].y) - (P[k].y - P[j].y) * (px - P[j].x);
  }
  const float q0 = e[0] * P[0].w, q1 = e[1] * P[1].w, q2 = e[2] * P[2].w;
  const float rQ = rcp(q0 + q1 + q2);
  const float b0 = q0 * rQ, b1 = q1 * rQ;
  const float g0 = gb_in[0] - gb_in[2], g1 = gb_in[1] - gb_in[2];
  const float s = g0 * b0 + g1 * b1;
  const float gq[3] = {(g0 - s) * rQ, (g1 - s) * rQ, -s * rQ};
  float gu[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float ge = gq[i] * P[i].w;
    const int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
    gu[j] += ge * (P[k].y - py);
    gv[j] += ge * (px - P[k].x);
    gu[k] += ge * (py - P[j].y);
    gv[k] += ge * (P[j].x - px);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float iz = P[i].w;
    const float gh0 = gu[i] * iz, gh1 = gv[i] * iz;
    const float gh2 = -(gu[i] * P[i].x + gv[i] * P[i].y) * iz - gq[i] * e[i] * iz * iz;
    out[i] = mk3(M[0] * gh0 + M[3] * gh1 + M[6] * gh2, M[1] * gh0 + M[4] * gh1 + M[7] * gh2,
                 M[2] * gh0 + M[5] * gh1 + M[8] * gh2);
  }
}

// Sum over a wavefront whose 64 lanes are all active, the same in every
// lane: DPP adds within each row of 16 lanes (swap neighbours, swap pairs,
// half-row mirror, row mirror: after each step a lane and its partner hold
// the same sum), then the four rows' sums in order.  Fixed order, so the
// result repeats bit for bit; ~11 instructions against ~36 for a
// ds_bpermute butterfly.
__device__ __forceinline__ float wave_sum_full(float v) {
  auto dpp_add = [](float x, int ctrl) {
    int y = 0;
    switch (ctrl) {  // the control word must be a constant
      case 0: y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xf, 0xf, false); break;   // [1,0,3,2]
      case 1: y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xf, 0xf, false); break;   // [2,3,0,1]
      case 2: y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xf, 0xf, false); break;  // half mirror
      default: y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xf, 0xf, false); break; // mirror
    }
    return x + __int_as_float(y);
  };
  v = dpp_add(v, 0);
  v = dpp_add(v, 1);
  v = dpp_add(v, 2);
  v = dpp_add(v, 3);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 47));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
  return ((r0 + r1) + r2) + r3;
}

// Sum over a 256-lane workgroup (every kernel of the step that calls it runs
// 256 lanes and calls it with all of them).
static_assert(kTilePix == kB, "the fine kernel's workgroups are kB lanes too");
__device__ __forceinline__ float block_sum(float v, float* s_red) {
  v = wave_sum_full(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  float t = 0.f;
#pragma unroll
  for (int i = 0; i < kB / 64; ++i) t += s_red[i];
  return t;
}

// View holding global tile index t (the views' tiles are consecutive).
__device__ __forceinline__ int view_of_tile(const Dev& d, long long t) {
  int lo = 0, hi = d.V - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d.vw[mid].tile0 <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Workgroup (x, view) -> tile of the view.  Workgroups are dealt to the L2
// domains (XCDs) round-robin; (x, view) is workgroup kHeavyCap + view * n + x
// of the fine kernel's grid, with kHeavyCap and n multiples of kXcd, so it
// runs on domain x % kXcd.  Each view is cut into kXcd strips of
// n / kXcd consecutive tiles and domain k takes strip (k + view) % kXcd: a
// contiguous run, so triangle records and target pixels shared by
// neighbouring tiles are fetched into one L2, and a different strip in every
// view, so the heavy middle strips of centred characters spread over all
// domains (the same strip in every view left the domains holding them far
// behind the others).  Bijective on [0, n) for every view.
__device__ __forceinline__ int xcd_tile(unsigned int x, unsigned int n, unsigned int view) {
  return (int)((((x & (kXcd - 1)) + view) & (kXcd - 1)) * (n / kXcd) + x / kXcd);
}


// Fast-path walk record of one listed triangle, clipped to the tile: the
// three biased edge values (inside <=> all >= 0) at the centre of the clipped
// box's first pixel, the per-pixel steps (x in the low, y in the high 16 bits,
// in units of kSub), 1/z per vertex, the box (tile-relative x0 | x1 << 8 |
// y0 << 16 | y1 << 24), the depth-key tag and the three biases.  Only
// triangles whose snapped extent is at most kSmallSpan sub-pixels get one:
// their edge values at every pixel centre of the box fit 32 bits.
struct RowRec {
  int e[3];
  int dxy[3];
  float iz[3];
  int box;
  unsigned int tag;
  int bias;  // bit i < 3 set: edge i is not top-left (bias -1); bits 8..31: the record's first row in
             // the tile's row list
};
static_assert(sizeof(RowRec) == sizeof(TriRec), "the walk record takes the staged triangle's place");

__device__ __forceinline__ int step_x(int v) { return (v << 16) >> 8; }  // sign-extended low half * kSub
__device__ __forceinline__ int step_y(int v) { return (v >> 16) << 8; }

// One row of a walk record: pixels of row r of the clipped box into the LDS keys.
template <int TWS>
__device__ __forceinline__ void raster_row(const RowRec& q, int r, unsigned long long* s_key) {
  const int x0 = q.box & 0xff, x1 = (q.box >> 8) & 0xff, y = ((q.box >> 16) & 0xff) + r;
  const int b0 = -(q.bias & 1), b1 = -((q.bias >> 1) & 1), b2 = -((q.bias >> 2) & 1);
  int e0 = q.e[0] + step_y(q.dxy[0]) * r, e1 = q.e[1] + step_y(q.dxy[1]) * r, e2 = q.e[2] + step_y(q.dxy[2]) * r;
  const int d0 = step_x(q.dxy[0]), d1 = step_x(q.dxy[1]), d2 = step_x(q.dxy[2]);
  // the record lives in LDS beside the keys: its 1/z and tag are read once,
  // not again after every key update
  const float izv[3] = {q.iz[0], q.iz[1], q.iz[2]};
  const int tg = (int)q.tag;
  // wait for those LDS reads here (lgkmcnt 0, the other counters left
  // alone): waited for inside the loop, the wait would also hold every
  // iteration until the previous key update is done
  __builtin_amdgcn_s_waitcnt(0xC07F);
  unsigned long long* row = s_key + (y << TWS);
  for (int x = x0; x <= x1; ++x) {
    if ((e0 | e1 | e2) >= 0)
      atomicMin(row + x, make_key(interp_iz32(e0 - b0, e1 - b1, e2 - b2, izv), tg));
    e0 += d0; e1 += d1; e2 += d2;
  }
}
constexpr int kGPad = kTilePix + 1;  // row stride of the per-pixel gradient table (bank spread)
// Per-pixel gradient table rows: 0-8 position gradient (vertex k, axis a at
// row 3k + a), 9-11 the colour-loss gradient g, 12-13 the weights w0, w1.
constexpr int kGRows = 14;
// Fast-path depth keys carry (face << kTagBits) | list index: the order of
// equal depths stays the face order, and the winner's list slot comes free.
constexpr int kTagBits = 8;
static_assert((1 << kTagBits) >= kFast, "list index must fit the tag");

// Record component q of the (triangle, tile) sums at pixel p: q = 6 k + c,
// c < 3 the position gradient of vertex k, else g[c - 3] * w_k.
__device__ __forceinline__ float grad_at(const float* s_g, int q, int p) {
  const int k = q / 6, comp = q - 6 * k;
  if (comp < 3) return s_g[(3 * k + comp) * kGPad + p];
  const float w0 = s_g[12 * kGPad + p], w1 = s_g[13 * kGPad + p];
  const float wk = k == 0 ? w0 : (k == 1 ? w1 : 1.f - w0 - w1);
  return s_g[(9 + comp - 3) * kGPad + p] * wk;
}

// Sums of the per-pixel gradient table over the pixels each (triangle, tile)
// pair won, written as one record per pair (or, without a record slot,
// accumulated into the fallback buffers).
__device__ __forceinline__ void emit_record(const Dev& d, const IO& io, const CharD& c, unsigned int e, int face,
                                            int q, float a, bool won) {
  if (e < d.capacity) {
    if (q == 0) d.rec_vis[e] = won ? 1 : 0;
    if (won) d.rec[(size_t)e * kGrad + q] = a;
  } else if (won) {
    const int kk = q / 6, comp = q - 6 * kk;
    const int vtx = c.v0 + d.faces[3 * ((size_t)c.f0 + face) + kk];
    if (comp < 3) atomicAdd(d.g_v + 3 * (size_t)vtx + comp, a);
    else atomicAdd(io.g_col + 3 * (size_t)vtx + comp - 3, a);
  }
}

// The optional 'uv' and 'normal' buffers of pixel o (CudaRendererGpu's uv and
// normal albedo modes: the op renderer's shade_pixel, the same IEEE
// operations in the same order, so they match it bit for bit); f < 0: off
// the mesh, zeros.
__device__ __forceinline__ void write_extras(const Dev& d, const IO& io, const CharD& c, int view, long long o, int f,
                                             int i0, int i1, int i2, float a, float b, float px, float py) {
  if (io.uv) {
    float tu = 0.f, tv = 0.f;
    if (f >= 0 && c.has_tc) {
      const float cw = 1.f - a - b;
      const float* t6 = d.tc + ((size_t)c.f0 + f) * 6;
      tu = t6[0] * a + t6[2] * b + t6[4] * cw;
      tv = (1.f - t6[1]) * a + (1.f - t6[3]) * b + (1.f - t6[5]) * cw;
      tu = fminf(fmaxf(tu, 0.f), 1.f);
      tv = fminf(fmaxf(tv, 0.f), 1.f);
    }
    io.uv[2 * o] = tu;
    io.uv[2 * o + 1] = tv;
  }
  if (io.nrm_out) {
    f3 col = mk3(0.f, 0.f, 0.f);
    if (f >= 0) {
      const float cw = 1.f - a - b;
      const float* vs = d.vsum + 3 * (size_t)c.v0;
      const f3 n0 = ld3(vs + 3 * i0), n1 = ld3(vs + 3 * i1), n2 = ld3(vs + 3 * i2);
      f3 pn = mk3(n0.x * a + n1.x * b + n2.x * cw, n0.y * a + n1.y * b + n2.y * cw, n0.z * a + n1.z * b + n2.z * cw);
      const float pl = sqrtf(dot(pn, pn));
      pn = mk3(pn.x / pl, pn.y / pl, pn.z / pl);
      const f3 rd = world_ray_dir(io.ext + 12 * (size_t)view, io.intr + 9 * (size_t)view, px, py);
      if (dot(pn, rd) > 0.f) pn = -pn;
      col = mk3((1.f + pn.x) / 2.f, (1.f + pn.y) / 2.f, (1.f + pn.z) / 2.f);
    }
    io.nrm_out[3 * o] = col.x;
    io.nrm_out[3 * o + 1] = col.y;
    io.nrm_out[3 * o + 2] = col.z;
  }
}

// The pixel's face, barycentrics and colour, stored with the non-temporal
// hint: the step writes them once and reads them back never, so they stream
// to memory rather than taking L2 lines from the triangle records and target
// pixels that neighbouring tiles share.
__device__ __forceinline__ void store_pixel(const IO& io, long long o, int f, float a, float b, f3 col) {
  __builtin_nontemporal_store(f, io.face + o);
  __builtin_nontemporal_store(a, io.bary + 2 * o);
  __builtin_nontemporal_store(b, io.bary + 2 * o + 1);
  __builtin_nontemporal_store(col.x, io.render + 3 * o);
  __builtin_nontemporal_store(col.y, io.render + 3 * o + 1);
  __builtin_nontemporal_store(col.z, io.render + 3 * o + 2);
}

// Face bits of the tile's rows (the mask starts the step cleared, so tiles
// that no triangle touches store nothing): the first lane of each tile row in
// the wavefront stores that row's bits from the wavefront's ballot (lanes are
// pixels in row-major tile order), a 16-, 32- or 64-bit word at the tile's
// column of the row, never shared with another tile.  Rows below the image
// are not stored; bits right of it are 0 and never read.
template <int TWS>
__device__ __forceinline__ void store_fg(const Dev& d, const ViewD& vw, int tx, int y, int px_i, bool fg) {
  const unsigned long long m = __ballot(fg);
  if (px_i == 0 && y < vw.H) {
    unsigned int* row = d.fgmask + vw.mask0 + (long long)y * vw.mpitch;
    const int sh = (int)(threadIdx.x & 63u);
    if (TWS == 6) reinterpret_cast<unsigned long long*>(row)[tx] = m;
    else if (TWS == 5) row[tx] = (unsigned int)(m >> sh);
    else reinterpret_cast<unsigned short*>(row)[tx] = (unsigned short)(m >> sh);
  }
}

// TWS: log2 of the tile width (tiles are (1 << TWS) x (1 << (8 - TWS)) pixels).
// EXTRA: the step also writes the uv / normal buffers.
template <int TWS, bool EXTRA>
__global__ __launch_bounds__(kTilePix) __attribute__((amdgpu_waves_per_eu(7))) void fine_kernel(Dev d, IO io) {
  constexpr int tws = TWS, ths = 8 - TWS;
  constexpr int aw = (1 << tws) + 2 * kImgFs, an = aw * ((1 << ths) + 2 * kImgFs);  // apron row length, size
  // u1: triangle records and depth keys of the tile (raster) / per-pixel
  // gradient table (backward)
  __shared__ union U1 {
    struct { RowRec rec[kFast]; unsigned long long key[kTilePix]; unsigned char item[kTilePix]; } r;
    float g[kGRows * kGPad];
  } u1;
  // u2: target colours with apron (per-pixel phase) / counting-sort scratch
  // and the list's record slots and faces (fast path) / pixel faces (slow path)
  __shared__ union U2 {
    float tgt[3 * an];
    struct {
      unsigned short wcnt[kTilePix / 64][kFast];  // pixels per (wavefront, list entry), then their prefix
      unsigned short order[kTilePix];             // covered pixels sorted by triangle
      unsigned char won[kFast];                   // list entries that won pixels, compacted
      unsigned int e[kFast];                      // record slot of each list entry
      unsigned int start[kFast + 1];              // first sorted pixel of each list entry
    } srt;
    int face[kTilePix];
  } u2;
  __shared__ unsigned int s_wave[kTilePix / 64];
  __shared__ float s_red[kTilePix / 64];
  // workgroup -> (view, tile): the first kHeavyCap take the listed slow-path
  // tiles (when all of them are listed), the rest walk every view's tiles in
  // the domain-strip order of xcd_tile and skip the listed ones
  const unsigned int n_heavy = *d.heavy_n;
  const bool listed = n_heavy <= (unsigned int)kHeavyCap;
  int w, lt;
  if (blockIdx.x < (unsigned int)kHeavyCap) {
    if (!listed || blockIdx.x >= n_heavy) return;
    const long long t = d.heavy[blockIdx.x];
    w = view_of_tile(d, t);
    lt = (int)(t - d.vw[w].tile0);
  } else {
    const unsigned int nx = (unsigned int)((d.maxTiles + kXcd - 1) / kXcd * kXcd);
    const unsigned int j = blockIdx.x - (unsigned int)kHeavyCap;
    w = (int)(j / nx);
    lt = xcd_tile(j - (unsigned int)w * nx, nx, (unsigned int)w);
    if (lt >= d.vw[w].tiles) return;
    if (listed) {
      const long long t = d.vw[w].tile0 + lt;
      const unsigned int o0 = d.tile_off[t], o1 = d.tile_off[t + 1];
      if (o1 - o0 > (unsigned int)kFast || o1 > d.capacity) return;  // taken by a listed workgroup
    }
  }
  const ViewD vw = d.vw[w];
  const CharD c = d.ch[vw.ch];
  const long long tile = vw.tile0 + lt;
  const int tx = lt % vw.tiles_x, ty = lt / vw.tiles_x;
  const int tx0 = tx << tws, ty0 = ty << ths;
  const unsigned int off = d.tile_off[tile], end = d.tile_off[tile + 1];
  const bool overflow = end > d.capacity;
  const unsigned int n_items = overflow ? (unsigned int)c.F : end - off;
  const bool fast = !overflow && n_items <= (unsigned int)kFast;  // uniform over the block
  const TriRec* tv = d.tri + vw.vf0;
  const int px_i = threadIdx.x & ((1 << tws) - 1), py_i = threadIdx.x >> tws;
  const int x = tx0 + px_i, y = ty0 + py_i;
  const bool inimg = x < vw.W && y < vw.H;

  if (n_items == 0) {
    // no triangle touches the tile: background outputs and the colour loss
    // against the target, nothing staged
    float lsum = 0.f;
    if (inimg) {
      const long long o = vw.pix0 + (long long)y * vw.W + x;
      const f3 tgt = ld3(io.target + 3 * o);
      store_pixel(io, o, -1, 0.f, 0.f, mk3(0.f, 0.f, 0.f));
      lsum = io.w_col * (fabsf(tgt.x) + fabsf(tgt.y) + fabsf(tgt.z));
      if (EXTRA) write_extras(d, io, c, w, o, -1, 0, 0, 0, 0.f, 0.f, 0.f, 0.f);
    }
    const float lt_sum = block_sum(lsum, s_red);
    if (threadIdx.x == 0) d.loss_tile[tile] = lt_sum;
    return;
  }

  // target colours of the tile plus a kImgFs apron (clamped reads; the apron is
  // only used by pixels whose whole filter window lies inside the image): a
  // thread's kStage pixels are all loaded before any is written to LDS, and
  // the fast path issues those loads beside its tile-list read, so that
  // neither costs a round trip of its own.
  constexpr int kStage = (an + kTilePix - 1) / kTilePix;
  f3 stg[kStage];
  auto load_target = [&]() {
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
      const int i = threadIdx.x + k * kTilePix;
      if (i < an) {
        const int ay = i / aw, ax = i - ay * aw;
        const int gx = min(max(tx0 + ax - kImgFs, 0), vw.W - 1), gy = min(max(ty0 + ay - kImgFs, 0), vw.H - 1);
        stg[k] = ld3(io.target + 3 * (vw.pix0 + (long long)gy * vw.W + gx));
      }
    }
  };
  auto store_target = [&]() {
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
      const int i = threadIdx.x + k * kTilePix;
      if (i < an) { u2.tgt[i] = stg[k].x; u2.tgt[an + i] = stg[k].y; u2.tgt[2 * an + i] = stg[k].z; }
    }
  };

  int f = -1;   // winning face of this pixel
  int bj = -1;  // fast path: index of the winning triangle in the tile list
  unsigned int my_e = 0;  // fast path, thread j < n_items: record slot and face of list entry j
  int my_f = 0;
  unsigned long long* s_key = u1.r.key;
  if (fast) {
    // lane k < n_items builds the walk record of list entry k; the rows of
    // all records are then dealt out to the lanes, 256 rows per pass, so a
    // wavefront's time follows its widest row rather than its largest
    // triangle.  Triangles wider or taller than kSmallSpan sub-pixels are
    // walked whole by their own lane afterwards (64-bit edge values).
    s_key[threadIdx.x] = kEmptyKey;
    unsigned int nrows = 0, ri = 0;
    bool wide = false;
    TriRec t;
    unsigned int tb = 0;  // the triangle's first record slot, read beside the triangle
    if (threadIdx.x < n_items) ri = d.bins[off + threadIdx.x];
    load_target();
    if (threadIdx.x < n_items) {
      t = d.tri[ri];
      tb = d.tri_base[ri];
    }
    store_target();
    if (threadIdx.x < n_items) {
      my_f = t.face_flags & kFaceMask;
      const TileRange r = tile_range(t, tws, ths);
      my_e = tb + (unsigned int)((ty - r.ty0) * r.ntx + (tx - r.tx0));
      const int x0 = max(t.bx0_by0 & 0xffff, tx0), y0 = max(t.bx0_by0 >> 16, ty0);
      const int x1 = min(t.bx1_by1 & 0xffff, tx0 + (1 << tws) - 1), y1 = min(t.bx1_by1 >> 16, ty0 + (1 << ths) - 1);
      const int spx = max(t.X[0], max(t.X[1], t.X[2])) - min(t.X[0], min(t.X[1], t.X[2]));
      const int spy = max(t.Y[0], max(t.Y[1], t.Y[2])) - min(t.Y[0], min(t.Y[1], t.Y[2]));
      if (x0 <= x1 && y0 <= y1) {
        if (spx <= kSmallSpan && spy <= kSmallSpan) {
          EdgeWalk w;
          w.init(t, x0, y0);
          RowRec q;
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            q.e[i] = (int)w.e[i] + w.bias[i];
            q.dxy[i] = (((int)(w.dx[i] / kSub)) & 0xffff) | ((int)(w.dy[i] / kSub) << 16);
            q.iz[i] = t.iz[i];
          }
          q.box = (x0 - tx0) | ((x1 - tx0) << 8) | ((y0 - ty0) << 16) | ((y1 - ty0) << 24);
          q.tag = ((unsigned int)my_f << kTagBits) | threadIdx.x;
          q.bias = (w.bias[0] & 1) | ((w.bias[1] & 1) << 1) | ((w.bias[2] & 1) << 2);
          u1.r.rec[threadIdx.x] = q;
          nrows = (unsigned int)(y1 - y0 + 1);
        } else {
          wide = true;
        }
      }
    }
    unsigned int n_rows;
    const unsigned int row0 = block_scan256(nrows, s_wave, n_rows);
    if (nrows) u1.r.rec[threadIdx.x].bias |= (int)(row0 << 8);
    for (unsigned int base = 0; base < n_rows; base += kTilePix) {
      // entry k's rows [row0, row0 + nrows) that fall in this pass of 256
      const unsigned int lo = max(row0, base), hi = min(row0 + nrows, base + (unsigned int)kTilePix);
      for (unsigned int j = lo; j < hi; ++j) u1.r.item[j - base] = (unsigned char)threadIdx.x;
      __syncthreads();
      if (base + threadIdx.x < n_rows) {
        const int k = u1.r.item[threadIdx.x];
        raster_row<tws>(u1.r.rec[k], (int)(base + threadIdx.x - ((unsigned int)u1.r.rec[k].bias >> 8)), s_key);
      }
      __syncthreads();
    }
    if (__syncthreads_or(wide))
      if (wide) raster_in_tile(d.tri[ri], tx0, ty0, tws, ths, ((unsigned int)my_f << kTagBits) | threadIdx.x, s_key);
    __syncthreads();
    const unsigned long long key = s_key[threadIdx.x];
    if (inimg && key != kEmptyKey) {
      const unsigned int tag = (unsigned int)(key & 0xffffffffull);
      f = (int)(tag >> kTagBits);
      bj = (int)(tag & ((1u << kTagBits) - 1u));
    }
  } else {
    load_target();
    store_target();
    s_key[threadIdx.x] = kEmptyKey;
    __syncthreads();
    for (unsigned int k = threadIdx.x; k < n_items; k += kTilePix) {
      const TriRec& t = overflow ? tv[k] : d.tri[d.bins[off + k]];
      raster_in_tile(t, tx0, ty0, tws, ths, (unsigned int)(t.face_flags & kFaceMask), s_key);
    }
    __syncthreads();
    const unsigned long long key = s_key[threadIdx.x];
    f = (inimg && key != kEmptyKey) ? (int)(key & 0xffffffffull) : -1;
  }
  __syncthreads();  // the keys and staged triangles are overwritten by the gradient table below
  store_fg<tws>(d, vw, tx, y, px_i, f >= 0);

  // ---- per pixel: outputs, colour loss, gradient contributions -> LDS
  float* s_g = u1.g;
  const int pix = threadIdx.x;
  float lsum = 0.f;
  if (inimg) {
    const long long o = vw.pix0 + (long long)y * vw.W + x;
    const int ta = (py_i + kImgFs) * aw + (px_i + kImgFs);
    const f3 tgt = mk3(u2.tgt[ta], u2.tgt[an + ta], u2.tgt[2 * an + ta]);
    if (f < 0) {
      store_pixel(io, o, -1, 0.f, 0.f, mk3(0.f, 0.f, 0.f));
      lsum = io.w_col * (fabsf(tgt.x) + fabsf(tgt.y) + fabsf(tgt.z));
      if (EXTRA) write_extras(d, io, c, w, o, -1, 0, 0, 0, 0.f, 0.f, 0.f, 0.f);
    } else {
      const int* fv = d.faces + 3 * ((size_t)c.f0 + f);
      const int i0 = fv[0], i1 = fv[1], i2 = fv[2];
      const f4* pv = d.proj + vw.vn0;
      // the three projections, read once before the output stores (read after
      // them they would be loaded again)
      const f4 P[3] = {pv[i0], pv[i1], pv[i2]};
      const float px = (float)x + 0.5f, py = (float)y + 0.5f;
      const f2 ab = native_bary(P[0], P[1], P[2], px, py);
      const float wgt[3] = {ab.x, ab.y, 1.f - ab.x - ab.y};
      const float* cb = io.col + 3 * (size_t)c.v0;
      const f3 c0 = ld3(cb + 3 * i0), c1 = ld3(cb + 3 * i1), c2 = ld3(cb + 3 * i2);
      const f3 col = mk3(c0.x * wgt[0] + c1.x * wgt[1] + c2.x * wgt[2], c0.y * wgt[0] + c1.y * wgt[1] + c2.y * wgt[2],
                         c0.z * wgt[0] + c1.z * wgt[1] + c2.z * wgt[2]);
      store_pixel(io, o, f, ab.x, ab.y, col);
      if (EXTRA) write_extras(d, io, c, w, o, f, i0, i1, i2, ab.x, ab.y, px, py);
      const f3 df = col - tgt;
      lsum = io.w_col * (fabsf(df.x) + fabsf(df.y) + fabsf(df.z));
      auto sgn = [](float a) { return a > 0.f ? 1.f : (a < 0.f ? -1.f : 0.f); };
      const f3 g = mk3(io.w_col * sgn(df.x), io.w_col * sgn(df.y), io.w_col * sgn(df.z));
      // colour gradient g * w_k is formed when the records are summed
      s_g[9 * kGPad + pix] = g.x; s_g[10 * kGPad + pix] = g.y; s_g[11 * kGPad + pix] = g.z;
      s_g[12 * kGPad + pix] = ab.x; s_g[13 * kGPad + pix] = ab.y;
      {
      // gradients only from here on (they carry a tolerance; the buffers
      // above are written with uncontracted IEEE arithmetic): multiply-adds
      // may fuse
#pragma clang fp contract(fast)
      const float gw[3] = {dot(g, c0), dot(g, c1), dot(g, c2)};
      const float* M = d.ke + (size_t)w * 12;
      f3 gx[3];
      bary_vjp_fast(P, M, px, py, gw, gx);
      // model-to-data term (CUDABasedRasterizationGrad.cu:413-437): the
      // target's image gradient (imageGradient, RendererUtil.h:566-620, with
      // its tap weights xx / (xx^2 + yy^2) and their L1 norm folded into
      // constants) projected back through the fragment's position
      const int fs = kImgFs;
      if (x >= fs + 1 && y >= fs + 1 && x < vw.W - (fs + 1) && y < vw.H - (fs + 1)) {
        // g . I at each tap of the window, into both sums with the tap weights
        // xx / (xx^2 + yy^2) and yy / (xx^2 + yy^2); zero weights are skipped
        float tu = 0.f, tv = 0.f;  // -g . (dI/du), -g . (dI/dv) before the norm
#pragma unroll
        for (int yy = -fs; yy <= fs; ++yy)
#pragma unroll
          for (int xx = -fs; xx <= fs; ++xx) {
            if (xx == 0 && yy == 0) continue;
            const int a = (py_i + kImgFs + yy) * aw + (px_i + kImgFs + xx);
            const float gi = g.x * u2.tgt[a] + g.y * u2.tgt[an + a] + g.z * u2.tgt[2 * an + a];
            const float den = (float)(xx * xx + yy * yy);
            if (xx != 0) tu += gi * ((float)xx / den);
            if (yy != 0) tv += gi * ((float)yy / den);
          }
        const float gu = -tu * kImgGradInvNorm, gvv = -tv * kImgGradInvNorm;
        // the fragment's homogeneous image point M (sum_k w_k X_k) + m: the
        // projection is linear in homogeneous coordinates, so it is the same
        // blend of the vertices' (u z, v z, z) (a winning triangle has every
        // vertex in front of the camera, so these are its true projections)
        const float hx = P[0].x * P[0].z * wgt[0] + P[1].x * P[1].z * wgt[1] + P[2].x * P[2].z * wgt[2];
        const float hy = P[0].y * P[0].z * wgt[0] + P[1].y * P[1].z * wgt[1] + P[2].y * P[2].z * wgt[2];
        const float hz = P[0].z * wgt[0] + P[1].z * wgt[1] + P[2].z * wgt[2];
        if (fabsf(hz) > 0.0001f) {
          const float rz = rcp(hz);
          const float gh0 = gu * rz, gh1 = gvv * rz, gh2 = -(gu * hx + gvv * hy) * rz * rz;
          const f3 gX = mk3(M[0] * gh0 + M[3] * gh1 + M[6] * gh2, M[1] * gh0 + M[4] * gh1 + M[7] * gh2,
                            M[2] * gh0 + M[5] * gh1 + M[8] * gh2);
#pragma unroll
          for (int k = 0; k < 3; ++k) gx[k] = gx[k] + gX * wgt[k];
        }
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        s_g[(3 * k) * kGPad + pix] = gx[k].x;
        s_g[(3 * k + 1) * kGPad + pix] = gx[k].y;
        s_g[(3 * k + 2) * kGPad + pix] = gx[k].z;
      }
      }
    }
  }
  const float lt_sum = block_sum(lsum, s_red);  // its barriers also publish s_g and retire u2.tgt
  if (threadIdx.x == 0) d.loss_tile[tile] = lt_sum;

  if (fast) {
    // ---- counting sort of the covered pixels by winning triangle, pixel order
    // kept within each triangle: ranks within the wavefront from ballots,
    // then per-wavefront counts and one scan that gives both the triangles'
    // offsets and a compact list of the triangles that won pixels (those
    // whose first pixel -- rank 0 in the first wavefront holding any of it --
    // is a pixel of this tile; most listed triangles win none)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    static_assert(sizeof(u2.srt.wcnt) == kTilePix * 8, "one 8-byte clear per lane");
    reinterpret_cast<uint2*>(&u2.srt.wcnt[0][0])[threadIdx.x] = make_uint2(0u, 0u);
    if (threadIdx.x < n_items) u2.srt.e[threadIdx.x] = my_e;
    __syncthreads();
    // one pass per distinct triangle in the wavefront; the leader's triangle
    // is read into a scalar register (no LDS round trip per pass)
    int rank = 0;
    unsigned long long act = __ballot(bj >= 0);
    while (act) {
      const int leader = __ffsll((long long)act) - 1;
      const int jv = __builtin_amdgcn_readlane(bj, leader);
      const unsigned long long m = __ballot(bj == jv);
      if (bj == jv) rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
      if (lane == leader) u2.srt.wcnt[wv][jv] = (unsigned short)__popcll(m);
      act &= ~m;
    }
    __syncthreads();
    // counts of this pixel's triangle in the earlier wavefronts: every
    // wavefront's count is read, all together, and the later ones masked off
    // (a short-circuit loop waited for each read in turn)
    unsigned int before = 0;
#pragma unroll
    for (int q = 0; q < kTilePix / 64; ++q) {
      const unsigned int n = u2.srt.wcnt[q][max(bj, 0)];
      before |= q < wv ? n : 0u;
    }
    const bool first = bj >= 0 && rank == 0 && before == 0;
    unsigned int cnt = 0, pre[kTilePix / 64];
#pragma unroll
    for (int q = 0; q < kTilePix / 64; ++q) {
      pre[q] = cnt;  // list entry threadIdx.x: its pixels in earlier wavefronts
      cnt += u2.srt.wcnt[q][threadIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kTilePix / 64; ++q) u2.srt.wcnt[q][threadIdx.x] = (unsigned short)pre[q];
    // counts (<= 256) in the high half, first-pixel flags in the low half
    unsigned int tot2;
    const unsigned int ex2 = block_scan256((cnt << 16) | (first ? 1u : 0u), s_wave, tot2);
    u2.srt.start[threadIdx.x] = ex2 >> 16;
    if (threadIdx.x == 0) u2.srt.start[kFast] = tot2 >> 16;
    if (first) u2.srt.won[ex2 & 0xffffu] = (unsigned char)bj;
    const int n_won = (int)(tot2 & 0xffffu);
    __syncthreads();
    if (bj >= 0) u2.srt.order[u2.srt.start[bj] + u2.srt.wcnt[wv][bj] + rank] = (unsigned short)pix;
    if (threadIdx.x < n_items) {
      const unsigned int e = u2.srt.e[threadIdx.x];
      if (e < d.capacity) d.rec_vis[e] = cnt > 0 ? 1 : 0;
    }
    __syncthreads();
    // ---- one (won triangle, vertex) pair per work item: the vertex's 3
    // position and 3 colour components, each summed over the triangle's
    // pixels in pixel order (so the sums repeat bit for bit, whatever order
    // the tile list came in)
    for (int i = threadIdx.x; i < 3 * n_won; i += kTilePix) {
#pragma clang fp contract(fast)
      const int wi = i / 3, k = i - 3 * wi;
      const int j = u2.srt.won[wi];
      const unsigned int st = u2.srt.start[j], en = u2.srt.start[j + 1];
      // w_k = s0 w0 + s1 w1 + s2 (1 - w0 - w1) with one s = 1: exact, branch-free
      const float s0 = k == 0 ? 1.f : 0.f, s1 = k == 1 ? 1.f : 0.f, s2 = k == 2 ? 1.f : 0.f;
      const float* gk = s_g + (3 * k) * kGPad;
      float a[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      // kSumUnroll pixels' table reads are issued together (the LDS latency
      // of one pixel's reads is paid once per group), and the next group's
      // pixel indices beside them (one LDS round trip per group, not two);
      // the additions still run in pixel order
      constexpr int kSumUnroll = 4;
      int pn[kSumUnroll];
#pragma unroll
      for (int u = 0; u < kSumUnroll; ++u) pn[u] = u2.srt.order[min(st + (unsigned int)u, en - 1u)];
      for (unsigned int r0 = st; r0 < en; r0 += kSumUnroll) {
        float v[kSumUnroll][6];
        int pc[kSumUnroll];
#pragma unroll
        for (int u = 0; u < kSumUnroll; ++u) {
          pc[u] = pn[u];
          pn[u] = u2.srt.order[min(r0 + (unsigned int)(kSumUnroll + u), en - 1u)];
        }
#pragma unroll
        for (int u = 0; u < kSumUnroll; ++u) {
          const int p = pc[u];
          const float w0 = s_g[12 * kGPad + p], w1 = s_g[13 * kGPad + p];
          const float wk = (w0 * s0 + w1 * s1) + (1.f - w0 - w1) * s2;
          v[u][0] = gk[p];
          v[u][1] = gk[kGPad + p];
          v[u][2] = gk[2 * kGPad + p];
          v[u][3] = s_g[9 * kGPad + p] * wk;
          v[u][4] = s_g[10 * kGPad + p] * wk;
          v[u][5] = s_g[11 * kGPad + p] * wk;
        }
#pragma unroll
        for (int u = 0; u < kSumUnroll; ++u)
          if (r0 + (unsigned int)u < en) {
#pragma unroll
            for (int q = 0; q < 6; ++q) a[q] += v[u][q];
          }
      }
      const unsigned int e = u2.srt.e[j];
      if (e < d.capacity) {
        float2* rp = reinterpret_cast<float2*>(d.rec + (size_t)e * kGrad + 6 * k);
        rp[0] = make_float2(a[0], a[1]);
        rp[1] = make_float2(a[2], a[3]);
        rp[2] = make_float2(a[4], a[5]);
      } else {
        const int face = d.tri[d.bins[off + j]].face_flags & kFaceMask;
#pragma unroll
        for (int cmp = 0; cmp < 6; ++cmp) emit_record(d, io, c, e, face, 6 * k + cmp, a[cmp], true);
      }
    }
    return;
  }

  // ---- slow path (long or overflowed tile lists): the thread that rasterised
  // a triangle sums the pixels it won (row-major order)
  int* s_face = u2.face;
  s_face[threadIdx.x] = f;
  __syncthreads();
  for (unsigned int k = threadIdx.x; k < n_items; k += kTilePix) {
    const unsigned int ri_local = overflow ? k : d.bins[off + k] - (unsigned int)vw.vf0;
    const TriRec t = tv[ri_local];
    if (t.bx0_by0 == 0x7fff7fff) continue;
    const int bx0 = max(t.bx0_by0 & 0xffff, tx0), by0 = max(t.bx0_by0 >> 16, ty0);
    const int bx1 = min(t.bx1_by1 & 0xffff, tx0 + (1 << tws) - 1), by1 = min(t.bx1_by1 >> 16, ty0 + (1 << ths) - 1);
    if (bx0 > bx1 || by0 > by1) continue;
    const int face = t.face_flags & kFaceMask;
    float a[kGrad];
#pragma unroll
    for (int q = 0; q < kGrad; ++q) a[q] = 0.f;
    bool won = false;
    for (int yy = by0; yy <= by1; ++yy)
      for (int xx = bx0; xx <= bx1; ++xx) {
        const int p = ((yy - ty0) << tws) | (xx - tx0);
        if (s_face[p] == face) {
          won = true;
#pragma unroll
          for (int q = 0; q < kGrad; ++q) a[q] += grad_at(s_g, q, p);
        }
      }
    const TileRange r = tile_range(t, d.tws, d.ths);
    const unsigned int e = d.tri_base[vw.vf0 + ri_local] + (unsigned int)((ty - r.ty0) * r.ntx + (tx - r.tx0));
#pragma unroll
    for (int q = 0; q < kGrad; ++q) emit_record(d, io, c, e, face, q, a[q], won);
  }
}

// ---------------------------------------------------------------------------
// Silhouette
// ---------------------------------------------------------------------------

// Boundary flag of a projected vertex: inside the image with a 1-pixel
// margin and a background pixel (face < 0) in its 3x3 neighbourhood, read
// from the fine kernel's face mask: per row the two words holding columns
// u - 1 and u + 1 (the same word unless the three straddle a word boundary),
// shifted so that those three bits come lowest; a clear bit is background.
__device__ __forceinline__ bool boundary_p(const Dev& d, const ViewD& vw, f4 p) {
  // (int)p - 1 >= 0 && (int)p + 1 < W, tested in float (NaN / huge safe)
  if (!(p.x >= 1.f && p.x < (float)(vw.W - 1) && p.y >= 1.f && p.y < (float)(vw.H - 1))) return false;
  const int u = (int)p.x, r0 = (int)p.y;
  const int wa = (u - 1) >> 5, wb = (u + 1) >> 5, sh = (u - 1) & 31;
  const unsigned int* m = d.fgmask + vw.mask0 + (long long)(r0 - 1) * vw.mpitch;
  unsigned int all = 7u;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const unsigned int* row = m + (long long)r * vw.mpitch;
    all &= (unsigned int)((((unsigned long long)row[wb] << 32) | row[wa]) >> sh);
  }
  return (all & 7u) != 7u;
}

__device__ __forceinline__ float remap_p(int v) { return v >= 127 ? (float)((v - 127) * 2) : (float)v; }

// Uniform grid over a view's DT crop for the nearest-boundary-point search:
// kRing cells per search radius R = 30 * scale (cells a hair larger than
// R / kRing), kGridMargin cells of margin, so every point within R of a DT
// pixel lies inside the grid.
constexpr int kRing = 8;
constexpr int kGridMargin = kRing + 1;
// Largest slot count of the fixed-slot grid layout.
constexpr int kMaxCellCap = 1024;
struct GridGeomP { float ox, oy, cell; };
__device__ __forceinline__ GridGeomP grid_geom_p(const float* cr) {
  const float sc = cr[2];
  GridGeomP g;
  g.cell = 30.f * sc * (1.f + 1.f / 1024.f) / (float)kRing;
  g.ox = (0.f - cr[0]) * sc + cr[3] - (float)kGridMargin * g.cell;
  g.oy = (0.f - cr[1]) * sc + cr[4] - (float)kGridMargin * g.cell;
  return g;
}
__device__ __forceinline__ int cell_of_p(const GridGeomP& g, int gx, int gy, float x, float y) {
  const float fx = floorf((x - g.ox) / g.cell), fy = floorf((y - g.oy) / g.cell);
  if (!(fx >= 0.f && fy >= 0.f && fx < (float)gx && fy < (float)gy)) return -1;
  return (int)fy * gx + (int)fx;
}

// Per (view, vertex): the boundary flag, the model->data residual
// (MultiViewSilhouetteLossGPUOp.cu:11-138) with its loss and gradient, and
// the vertex's count in the nearest-point grid (boundary vertices only).
__global__ __launch_bounds__(kB) void sil_vertex_kernel_p(Dev d, IO io) {
  __shared__ float s_red[kB / 64];
  const int w = blockIdx.y;
  const ViewD vw = d.vw[w];
  const CharD c = d.ch[vw.ch];
  if ((long long)blockIdx.x * kB >= c.N) {  // no vertices here (uniform over the block)
    if (threadIdx.x == 0) d.loss_pt[(size_t)w * gridDim.x + blockIdx.x] = 0.f;
    return;
  }
  const int v = blockIdx.x * kB + threadIdx.x;
  float lsum = 0.f;
  // the view's crop record in registers, read before the atomics and stores
  // below (read after them, each value would wait for them)
  float crv[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) crv[k] = io.crops[(size_t)w * 7 + k];
  const float* cr = crv;
  if (v < c.N) {
    const long long pi = vw.vn0 + v;
    const f4 p = d.proj[pi];
    const bool isb = boundary_p(d, vw, p);
    d.isb[pi] = isb ? 1 : 0;
    if (isb) {
      // the boundary point's cell count; with fixed slots the point goes
      // straight into a slot of its cell (or the overflow list): the
      // nearest-point search takes the minimum of (distance, index), which
      // does not depend on the order of a cell's points
      const int cl = cell_of_p(grid_geom_p(cr), vw.gx, vw.gy, p.x, p.y);
      if (cl >= 0) {
        const long long cell = vw.cell0 + cl;
        const unsigned int k = atomicAdd(d.cell_cnt + cell, 1u);
        if (d.cell_cap > 0) {
          const f4 pt = mk4(p.x, p.y, __int_as_float(v), 0.f);
          if (k < (unsigned int)d.cell_cap) d.cell_slots[cell * d.cell_cap + k] = pt;
          else d.ov_pts[vw.vn0 + atomicAdd(d.ov_cnt + w, 1u)] = pt;  // at most one entry per vertex
        }
      }
    }
    const float pu = p.x, pv = p.y;
    const bool in_view = pu - 2 > cr[3] && pu + 2 < cr[5] && pv - 2 > cr[4] && pv + 2 < cr[6];
    const int xi = (int)((pu - cr[3]) / cr[2] + cr[0]), yi = (int)((pv - cr[4]) / cr[2] + cr[1]);
    const int U = vw.S, Vv = vw.S;
    float gx = 0.f, gy = 0.f;
    if (in_view && xi - 1 >= 0 && xi + 1 < U && yi - 1 >= 0 && yi + 1 < Vv) {
      const unsigned char* img = io.dt + vw.sil0;
      auto at = [&](int xx, int yy) { return (int)img[(size_t)yy * U + xx]; };
      float dv = remap_p(at(xi, yi));
      const float f00 = remap_p(at(xi - 1, yi - 1)), f01 = remap_p(at(xi, yi - 1)), f02 = remap_p(at(xi + 1, yi - 1));
      const float f10 = remap_p(at(xi - 1, yi)), f12 = remap_p(at(xi + 1, yi));
      const float f20 = remap_p(at(xi - 1, yi + 1)), f21 = remap_p(at(xi, yi + 1)), f22 = remap_p(at(xi + 1, yi + 1));
      const float du = (1.f / 8.f) * ((-1.f) * f00 + (+1.f) * f02 + (-2.f) * f10 + (+2.f) * f12 + (-1.f) * f20 + (+1.f) * f22);
      const float dvv = (1.f / 8.f) * ((-1.f) * f00 + (-2.f) * f01 + (-1.f) * f02 + (+1.f) * f20 + (+2.f) * f21 + (+1.f) * f22);
      float nx = du, ny = dvv, ux = du, uy = dvv;
      const bool inside_fg = at(xi, yi) <= 127;
      const float nonvan = (du != 0.f || dvv != 0.f) ? 1.f : 0.f;
      const float bnd = isb ? 1.f : 0.f;
      if (inside_fg) { nx = -nx; ny = -ny; }
      const f2 nn = d.nrm2[pi];
      if (!(nx * nn.x + ny * nn.y >= 0.f) && inside_fg) { ux = -ux; uy = -uy; }
      dv *= bnd;
      dv *= nonvan;
      ux *= bnd; uy *= bnd;
      ux *= nonvan; uy *= nonvan;
      const float res = dv * (1.f / (255.f));
      lsum = io.w_sil * res * res;
      const float gres = 2.f * io.w_sil * res;
      gx = ux * (1.f / (255.f)) * gres;
      gy = uy * (1.f / (255.f)) * gres;
    }
    d.g_pts[2 * pi] = gx;
    d.g_pts[2 * pi + 1] = gy;
    // the data->model residual sums of this point start at zero (the pixel
    // pass adds to them after this kernel)
    *reinterpret_cast<longlong2*>(d.g_fx + 2 * pi) = make_longlong2(0, 0);
  }
  const float t = block_sum(lsum, s_red);
  if (threadIdx.x == 0) d.loss_pt[(size_t)w * gridDim.x + blockIdx.x] = t;
}

// Counting-sort layout: the boundary points into their cells' ranges.
__global__ __launch_bounds__(kB) void sil_grid_scatter_p(Dev d, IO io) {
  const int w = blockIdx.y;
  const ViewD vw = d.vw[w];
  const CharD c = d.ch[vw.ch];
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= c.N || !d.isb[vw.vn0 + v]) return;
  const f4 p = d.proj[vw.vn0 + v];
  const int cl = cell_of_p(grid_geom_p(io.crops + (size_t)w * 7), vw.gx, vw.gy, p.x, p.y);
  if (cl < 0) return;
  const long long cell = vw.cell0 + cl;
  d.cell_pts[d.cell_off[cell] + atomicAdd(d.cell_fill + cell, 1u)] = mk4(p.x, p.y, __int_as_float(v), 0.f);
}

constexpr int kDtTile = 16;      // DT pixels per side of a silhouette tile (one lane each)
constexpr float kResidualFx = 1099511627776.f;  // 2^40: fixed-point unit of the data->model residual sums
constexpr int kDtApron = kDtTile + 4;  // the tile plus the 2-pixel border of the 5x5 boundary test
constexpr int kMaxRegionPts = 1280;  // regions with more points search the global grid
constexpr int kMaxRegionRows = 64;
constexpr int kMaxRegionCells = 1024;  // entries of the per-cell LDS offset table (rows x (cols + 1))

// Squared-distance bound of a best distance bd >= 0: a point whose squared
// distance d2 is at least this has a correctly rounded distance sqrtf(d2) of
// at least next_up(bd) (the bound is above next_up(bd)^2 exactly: next_up of
// a rounded value lies above the exact one), so it can neither beat nor tie
// bd.  Points below it are compared exactly, so the search skips the square
// root for most points and still returns the same point.
__device__ __forceinline__ float sq_bound(float bd) {
  const float nb = __int_as_float(__float_as_int(bd) + 1);
  const float t = nb * nb;
  if (!(t < 3.0e38f)) return INFINITY;
  return fmaxf(__int_as_float(__float_as_int(t) + 1), 1.17549435e-38f);  // at least the smallest normal
}

// Keep the point at squared distance d2 if it is nearer than (bd, best), or
// as near with a lower index (d2 below the bound tq of bd; tq follows bd).
__device__ __forceinline__ void nearest_take(float d2, int v, float x, float y, float& bd, int& best, float& bx,
                                             float& by, float& tq) {
  const float dd = sqrtf(d2);
  if (dd < bd || (dd == bd && best >= 0 && v < best)) {
    bd = dd; best = v; bx = x; by = y;
    tq = sq_bound(bd);
  }
}

// Data->model residual (MultiViewSilhouetteLossGPUOp.cu:142-240) + loss +
// gradient (atomic, as the reference).  One workgroup per 16x16 tile of DT
// pixels; the tile and a 2-pixel border of the DT image are staged in LDS for
// the boundary test (is-boundary: value 127 and at least four 127s in the
// 5x5 window, out-of-image pixels not counted).  The grid cells within R of
// the tile (each cell row of that region is one contiguous range of the
// grid) are copied to LDS with their per-row cumulative offsets.  Each
// boundary pixel then visits the cell rows nearest first (its own, then +-1,
// +-2, ...); in a row it scans the one contiguous range of cells whose
// x-distance lower bound does not exceed the best distance so far, keeping
// the minimum of (distance, index) with distance < R.  A row whose y-distance
// lower bound exceeds the best distance ends the search (all later rows are
// farther).  Points of skipped cells are strictly farther than the result,
// so it is the brute-force one; bounds carry a margin for float rounding.
// Regions larger than the LDS tables fall back to a ring search over the
// global grid.
__global__ __launch_bounds__(kB) void sil_pixel_kernel_p(Dev d, IO io) {
  __shared__ f2 s_pxy[kMaxRegionPts];
  __shared__ int s_pid[kMaxRegionPts];
  __shared__ unsigned int s_coff[kMaxRegionCells];
  // the DT tile with its apron (boundary test), then, for the counting-sort
  // layout, each region row's first point (staging): one LDS block, so the
  // kernel stays within 8 workgroups per CU
  __shared__ union {
    unsigned char dt[kDtApron * kDtApron];
    unsigned int ra[kMaxRegionRows];
  } s_u;
  static_assert(sizeof(s_u.ra) <= sizeof(s_u.dt), "the row table must not grow the DT tile's block");
  unsigned char* const s_dt = s_u.dt;
  unsigned int* const s_ra = s_u.ra;
  __shared__ unsigned char s_bp[kDtTile * kDtTile];  // the tile's boundary pixels, compacted
  __shared__ unsigned int s_wave[kB / 64];
  __shared__ float s_red[kB / 64];
  const int w = blockIdx.y;
  const ViewD vw = d.vw[w];
  const int S = vw.S;
  const int tpr = (S + kDtTile - 1) / kDtTile;
  const int tile = blockIdx.x;
  const bool tile_ok = tile < tpr * tpr;
  const int tu0 = (tile % tpr) * kDtTile, tv0 = (tile / tpr) * kDtTile;
  const int ui = tu0 + (int)(threadIdx.x % kDtTile);
  const int vi = tv0 + (int)(threadIdx.x / kDtTile);
  const unsigned char* img = io.dt + vw.sil0;
  // a boundary pixel has the contour value 127: a tile without one (most of
  // them) is done before anything is staged
  const bool contour = tile_ok && ui < S && vi < S && img[(size_t)vi * S + ui] == 127;
  if (!__syncthreads_or(contour)) {
    if (threadIdx.x == 0) d.loss_px[(size_t)w * gridDim.x + blockIdx.x] = 0.f;
    return;
  }
  if (tile_ok)
    for (int i = threadIdx.x; i < kDtApron * kDtApron; i += kB) {
      const int ay = i / kDtApron, ax = i - ay * kDtApron;
      const int gx = tu0 + ax - 2, gy = tv0 + ay - 2;
      s_dt[i] = (gx >= 0 && gx < S && gy >= 0 && gy < S) ? img[(size_t)gy * S + gx] : (unsigned char)0;
    }
  __syncthreads();
  bool bnd = false;
  if (tile_ok && ui < S && vi < S) {
    const int cx = (int)(threadIdx.x % kDtTile) + 2, cy = (int)(threadIdx.x / kDtTile) + 2;
    if (s_dt[cy * kDtApron + cx] == 127) {
      int counter = 0;
#pragma unroll
      for (int j = -2; j <= 2; ++j)
#pragma unroll
        for (int i = -2; i <= 2; ++i) counter += s_dt[(cy + j) * kDtApron + cx + i] == 127 ? 1 : 0;
      bnd = counter >= 4;
    }
  }
  float lsum = 0.f;
  if (__syncthreads_or(bnd)) {
    const float* cr = io.crops + (size_t)w * 7;
    const float sc = cr[2];
    const float R = 30.f * sc;
    const GridGeomP gg = grid_geom_p(cr);
    const float xlo = ((float)tu0 - cr[0]) * sc + cr[3] - R, xhi = ((float)(tu0 + kDtTile - 1) - cr[0]) * sc + cr[3] + R;
    const float ylo = ((float)tv0 - cr[1]) * sc + cr[4] - R, yhi = ((float)(tv0 + kDtTile - 1) - cr[1]) * sc + cr[4] + R;
    const int cx0 = max(0, (int)floorf((xlo - gg.ox) / gg.cell) - 1);
    const int cx1 = min(vw.gx - 1, (int)floorf((xhi - gg.ox) / gg.cell) + 1);
    const int cy0 = max(0, (int)floorf((ylo - gg.oy) / gg.cell) - 1);
    const int cy1 = min(vw.gy - 1, (int)floorf((yhi - gg.oy) / gg.cell) + 1);
    const int nrows = cy1 - cy0 + 1, ncols = cx1 - cx0 + 1;
    const int ncell = nrows * (ncols + 1);  // region cells row-major, plus an end entry per row
    const bool shape_ok = nrows > 0 && ncols > 0 && nrows <= kMaxRegionRows && ncell <= kMaxRegionCells;
    const int cap = d.cell_cap;
    // the region's cell counts (capped by the slots of the fixed-slot
    // layout), scanned in row-major order: entry (r, k) of s_coff is the LDS
    // offset of cell (cy0 + r, cx0 + k), entry (r, ncols) the end of row r;
    // thread t scans entries 4t .. 4t + 3
    unsigned int cnt4[4] = {0u, 0u, 0u, 0u}, part = 0u;
    if (shape_ok) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = 4 * (int)threadIdx.x + j;
        const int r = i / (ncols + 1), k = i - r * (ncols + 1);
        if (i < ncell && k < ncols) {
          const unsigned int n = d.cell_cnt[vw.cell0 + (long long)(cy0 + r) * vw.gx + cx0 + k];
          cnt4[j] = cap > 0 ? min(n, (unsigned int)cap) : n;
        }
        part += cnt4[j];
      }
      // counting-sort layout: a region row's cells are consecutive in the
      // grid, so their points are one contiguous range
      if (cap == 0 && (int)threadIdx.x < nrows)
        s_ra[threadIdx.x] = d.cell_off[vw.cell0 + (long long)(cy0 + (int)threadIdx.x) * vw.gx + cx0];
    }
    unsigned int total;
    unsigned int ex = block_scan256(part, s_wave, total);
    if (shape_ok)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = 4 * (int)threadIdx.x + j;
        if (i < ncell) s_coff[i] = ex;
        ex += cnt4[j];
      }
    __syncthreads();
    const bool in_lds = shape_ok && total <= (unsigned int)kMaxRegionPts;
    // fixed-slot layout: points of this view whose cell was full (searched by every pixel)
    const unsigned int n_ov = cap > 0 ? d.ov_cnt[w] : 0u;
    if (in_lds)
      for (unsigned int i = threadIdx.x; i < total; i += kB) {
        // the entry holding point i: the last one whose offset is <= i (an
        // entry holding points is followed by a larger offset; row ends and
        // empty cells share the offset of the next entry)
        int e = 0, eh = ncell - 1;
        while (e < eh) {
          const int mid = (e + eh + 1) >> 1;
          if (s_coff[mid] <= i) e = mid; else eh = mid - 1;
        }
        const int r = e / (ncols + 1), k = e - r * (ncols + 1);
        const long long cell = vw.cell0 + (long long)(cy0 + r) * vw.gx + cx0 + k;
        const f4 it = cap > 0 ? d.cell_slots[cell * cap + (i - s_coff[e])]
                              : d.cell_pts[s_ra[r] + (i - s_coff[r * (ncols + 1)])];
        s_pxy[i] = mk2(it.x, it.y);
        s_pid[i] = __float_as_int(it.z);
      }
    __syncthreads();
    // compact the boundary pixels, then give each a group of L lanes (a power
    // of two, L * count <= 256, at most a wavefront) that stride the points
    // of every cell row.  The group shares its best distance after each row
    // for pruning only; each lane keeps its own (distance, index) minimum and
    // a butterfly merge takes the minimum with the lower index on ties, so
    // the result is the single-lane one.
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long bm = __ballot(bnd);
    if (lane == 0) s_wave[wv] = (unsigned int)__popcll(bm);
    __syncthreads();
    int nb = 0, before = 0;
#pragma unroll
    for (int q = 0; q < kB / 64; ++q) {
      before += q < wv ? (int)s_wave[q] : 0;
      nb += (int)s_wave[q];
    }
    if (bnd) s_bp[before + __popcll(bm & ((1ull << lane) - 1ull))] = (unsigned char)threadIdx.x;
    __syncthreads();
    int lg = 0;  // log2 of the group size
    if (in_lds)
      while (lg < 6 && (nb << (lg + 1)) <= kB) ++lg;
    const int L = 1 << lg, item = (int)threadIdx.x >> lg, sub = (int)threadIdx.x & (L - 1);
    if (item < nb) {
      const int pp = s_bp[item];
      const int pu = tu0 + pp % kDtTile, pv = tv0 + pp / kDtTile;
      const float ix = ((float)pu - cr[0]) * sc + cr[3], iy = ((float)pv - cr[1]) * sc + cr[4];
      float bd = R, bx = 0.f, by = 0.f;
      float tq = sq_bound(R);
      int best = -1;
      const int cx = (int)floorf((ix - gg.ox) / gg.cell), cy = (int)floorf((iy - gg.oy) / gg.cell);
      if (in_lds) {
        const float step = gg.cell * (1.f - 1e-4f);  // lower bound of the distance per cell of separation
        float gbd = R;  // the group's best distance so far (pruning bound)
        for (int t = 0; t <= 2 * (kRing + 1); ++t) {
          const int dy = (t & 1) ? (t + 1) >> 1 : -(t >> 1);  // 0, +1, -1, +2, -2, ...
          const int ady = dy < 0 ? -dy : dy;
          if ((float)(ady > 0 ? ady - 1 : 0) * step > gbd) break;
          const int yy = cy + dy;
          if (yy < cy0 || yy > cy1) continue;
          const int kx = (int)(gbd / step) + 1;  // cells with (|xx - cx| - 1) * step <= gbd
          const int xa = max(cx - kx, cx0), xb = min(cx + kx, cx1);
          if (xa > xb) continue;
          const int rb = (yy - cy0) * (ncols + 1);
          const unsigned int q1 = s_coff[rb + xb - cx0 + 1];
          for (unsigned int q = s_coff[rb + xa - cx0] + (unsigned int)sub; q < q1; q += (unsigned int)L) {
            const f2 it = s_pxy[q];
            const float dx = it.x - ix, dyy = it.y - iy;
            const float d2 = dx * dx + dyy * dyy;
            if (d2 < tq) nearest_take(d2, s_pid[q], it.x, it.y, bd, best, bx, by, tq);
          }
          gbd = fminf(gbd, bd);
          for (int o = L >> 1; o > 0; o >>= 1) gbd = fminf(gbd, __shfl_xor(gbd, o, 64));
        }
        for (unsigned int q = (unsigned int)sub; q < n_ov; q += (unsigned int)L) {
          const f4 it = d.ov_pts[vw.vn0 + q];
          const float dx = it.x - ix, dyy = it.y - iy;
          const float d2 = dx * dx + dyy * dyy;
          if (d2 < tq) nearest_take(d2, __float_as_int(it.z), it.x, it.y, bd, best, bx, by, tq);
        }
        for (int o = L >> 1; o > 0; o >>= 1) {
          const float obd = __shfl_xor(bd, o, 64), obx = __shfl_xor(bx, o, 64), oby = __shfl_xor(by, o, 64);
          const int ob = __shfl_xor(best, o, 64);
          if (ob >= 0 && (best < 0 || obd < bd || (obd == bd && ob < best))) { bd = obd; best = ob; bx = obx; by = oby; }
        }
      } else {
        for (int k = 0; k <= kRing; ++k) {
          for (int oy = -k; oy <= k; ++oy) {
            const int yy = cy + oy;
            if (yy < 0 || yy >= vw.gy) continue;
            const int stp = (oy == -k || oy == k) ? 1 : 2 * k;
            for (int ox = -k; ox <= k; ox += stp) {
              const int xx = cx + ox;
              if (xx < 0 || xx >= vw.gx) continue;
              const long long cell = vw.cell0 + yy * vw.gx + xx;
              const f4* pts = cap > 0 ? d.cell_slots + cell * cap : d.cell_pts + d.cell_off[cell];
              const unsigned int q1 = cap > 0 ? min(d.cell_cnt[cell], (unsigned int)cap) : d.cell_cnt[cell];
              for (unsigned int q = 0; q < q1; ++q) {
                const f4 it = pts[q];
                const float dx = it.x - ix, dyy = it.y - iy;
                const float d2 = dx * dx + dyy * dyy;
                if (d2 < tq) nearest_take(d2, __float_as_int(it.z), it.x, it.y, bd, best, bx, by, tq);
              }
            }
          }
          const float reach = (float)k * gg.cell * (1.f - 1e-4f);
          if (bd < reach || reach >= R) break;
        }
        for (unsigned int q = 0; q < n_ov; ++q) {
          const f4 it = d.ov_pts[vw.vn0 + q];
          const float dx = it.x - ix, dyy = it.y - iy;
          const float d2 = dx * dx + dyy * dyy;
          if (d2 < tq) nearest_take(d2, __float_as_int(it.z), it.x, it.y, bd, best, bx, by, tq);
        }
      }
      if (sub == 0 && best >= 0) {
        const float r0 = (bx - ix) / (30.f * sc), r1 = (by - iy) / (30.f * sc);
        lsum = io.w_sil * (r0 * r0 + r1 * r1);
        // the residual (|r| <= 1) goes to the nearest point's sums in fixed
        // point: integer additions give the same bits in any order, so the
        // joint and node gradients repeat exactly; the gradient factor
        // 2 w_sil / (30 sc) is applied once, when the sums are read
        unsigned long long* q = reinterpret_cast<unsigned long long*>(d.g_fx + 2 * (vw.vn0 + best));
        atomicAdd(q, (unsigned long long)(long long)rintf(r0 * kResidualFx));
        atomicAdd(q + 1, (unsigned long long)(long long)rintf(r1 * kResidualFx));
      }
    }
  }
  const float t = block_sum(lsum, s_red);
  if (threadIdx.x == 0) d.loss_px[(size_t)w * gridDim.x + blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// Backward: vertices, ED nodes, per-character reductions
// ---------------------------------------------------------------------------

// Per face of every character: its (triangle, tile) records summed over the
// tiles of each view and over the views, in a fixed order.
__global__ __launch_bounds__(kB) void face_sum_kernel(Dev d) {
  const CharD c = d.ch[blockIdx.y];
  const int f = blockIdx.x * kB + threadIdx.x;
  if (f >= c.F) return;
  // the corners' slots are read first, so that their loads do not wait behind
  // the stores at the end
  const int* cs = d.corner_slot + 3 * ((size_t)c.f0 + f);
  const int slot[3] = {cs[0], cs[1], cs[2]};
  float a[kGrad];
#pragma unroll
  for (int q = 0; q < kGrad; ++q) a[q] = 0.f;
  // a character's views are consecutive, F (view, face) entries apart: four
  // views' record ranges are loaded before any is summed
  const long long ri0 = d.vw[c.view0].vf0 + f;
  for (int k0 = 0; k0 < c.C; k0 += 4) {
    unsigned int lo[4], hi[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      lo[j] = hi[j] = 0u;
      if (k0 + j < c.C) {
        const long long ri = ri0 + (long long)(k0 + j) * c.F;
        lo[j] = d.tri_base[ri];
        hi[j] = d.tri_base[ri + 1];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const unsigned int e1 = min(hi[j], d.capacity);
      for (unsigned int e = lo[j]; e < e1; ++e) {
        if (!d.rec_vis[e]) continue;
        const float* r = d.rec + (size_t)e * kGrad;
#pragma unroll
        for (int q = 0; q < kGrad; ++q) a[q] += r[q];
      }
    }
  }
  // each corner's six sums go to its slot of the vertex's incidence list, so
  // the vertex pass reads one contiguous range per vertex
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float2* o = reinterpret_cast<float2*>(d.corner_sum + 6 * (size_t)slot[k]);
    o[0] = make_float2(a[6 * k], a[6 * k + 1]);
    o[1] = make_float2(a[6 * k + 2], a[6 * k + 3]);
    o[2] = make_float2(a[6 * k + 4], a[6 * k + 5]);
  }
}

// dL/dv -> DQ-skinning adjoint, in two parts (the adjoint is linear in
// dL/dv, so the parts add up to the adjoint of the sum):
//   SIL = false: the raster part, the sums of the vertex's face corners
//     (ascending face, then corner) plus any fallback accumulations; writes
//     the colour gradient, the per-vertex joint-DQ terms, the blend weights
//     and R^T g.  It needs only the per-face sums, so it runs beside the
//     silhouette stages.
//   SIL = true: the silhouette part, the camera-projection adjoint over the
//     views (CameraProjectionGPUOpGrad.cu:12-79), added to those terms; a
//     vertex with no silhouette gradient in any view returns at once.
template <bool SIL>
__global__ __launch_bounds__(kB) void vertex_bwd_kernel(Dev d, IO io) {
#pragma clang fp contract(fast)
  __shared__ float s_tab[256 * 8];
  const CharD c = d.ch[blockIdx.y];
  if ((long long)blockIdx.x * kB >= c.N) return;  // no vertices here (uniform over the block)
  if (!io.skA)
    for (int i = threadIdx.x; i < c.J * 8; i += kB) s_tab[i] = io.dq[(size_t)c.j0 * 8 + i];
  __syncthreads();
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= c.N) return;
  const int gv = c.v0 + v;
  f3 g = mk3(0.f, 0.f, 0.f);
  if (!SIL) {
    g = ld3(d.g_v + 3 * (size_t)gv);
    f3 gc = ld3(io.g_col + 3 * (size_t)gv);
    for (int q = d.vf_off[gv]; q < d.vf_off[gv + 1]; ++q) {
      const float2* r = reinterpret_cast<const float2*>(d.corner_sum + 6 * (size_t)q);
      const float2 r0 = r[0], r1 = r[1], r2 = r[2];
      g += mk3(r0.x, r0.y, r1.x);
      gc += mk3(r1.y, r2.x, r2.y);
    }
    io.g_col[3 * (size_t)gv] = gc.x; io.g_col[3 * (size_t)gv + 1] = gc.y; io.g_col[3 * (size_t)gv + 2] = gc.z;
  } else {
    const f3 X = ld3(d.vdef + 3 * (size_t)gv);
    bool any = false;
    for (int k = 0; k < c.C; ++k) {
      const int w = c.view0 + k;
      const ViewD vw = d.vw[w];
      f2 gp = ((const f2*)d.g_pts)[vw.vn0 + v];
      const longlong2 q = *reinterpret_cast<const longlong2*>(d.g_fx + 2 * (vw.vn0 + v));
      if (q.x != 0 || q.y != 0) {
        // data->model part: the residual sums times 2 w_sil / (30 sc)
        // (MultiViewSilhouetteLossGPUOpGrad.cu:11-72)
        const float sc = io.crops[(size_t)w * 7 + 2];
        const float unit = 2.f * io.w_sil / (30.f * sc);
        gp.x += unit * (float)((double)q.x * (1.0 / (double)kResidualFx));
        gp.y += unit * (float)((double)q.y * (1.0 / (double)kResidualFx));
      }
      if (gp.x == 0.f && gp.y == 0.f) continue;
      any = true;
      const float* E = io.ext + (size_t)w * 12;
      const float* K = io.intr + (size_t)w * 9;
      const float xc = E[0] * X.x + E[1] * X.y + E[2] * X.z + E[3];
      const float yc = E[4] * X.x + E[5] * X.y + E[6] * X.z + E[7];
      const float zc = E[8] * X.x + E[9] * X.y + E[10] * X.z + E[11];
      const float ax = K[0], ay = K[4], gm = K[1];
      const float j00 = ax / zc, j01 = gm / zc, j02 = (-ax * xc - gm * yc) / (zc * zc);
      const float j11 = ay / zc, j12 = (-ay * yc) / (zc * zc);
      const float g0 = gp.x * j00, g1 = gp.x * j01 + gp.y * j11, g2 = gp.x * j02 + gp.y * j12;
      g = g + mk3(E[0] * g0 + E[4] * g1 + E[8] * g2, E[1] * g0 + E[5] * g1 + E[9] * g2, E[2] * g0 + E[6] * g1 + E[10] * g2);
    }
    if (!any) return;
  }
  if (io.skA) {
    // graph skinning: the outer sum's adjoint w.r.t. v_nr, sum_s w_s R_s^T g
    f3 u = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < KE; ++k) {
      const float* r = d.node_skin + ((size_t)c.m0 + d.enode[(size_t)gv * KE + k]) * 12;
      m3 R;
      for (int q = 0; q < 9; ++q) R.m[q] = r[q];
      u += d.ew[(size_t)gv * KE + k] * mulT(R, g);
    }
    float* up = d.u + 3 * (size_t)gv;
    if (!SIL) { up[0] = u.x; up[1] = u.y; up[2] = u.z; }
    else { up[0] += u.x; up[1] += u.y; up[2] += u.z; }
    return;
  }
  const f3 vnr = ed_inner(d, c, gv);
  int idx[KS];
  float w[KS];
  for (int k = 0; k < KS; ++k) { idx[k] = d.jidx[(size_t)gv * KS + k]; w[k] = d.jw[(size_t)gv * KS + k]; }
  const DqBlend bl = dq_blend(s_tab, idx, w, KS);
  const m3 R = quat_to_mat(bl.r);
  const float G[9] = {g.x * vnr.x, g.x * vnr.y, g.x * vnr.z, g.y * vnr.x, g.y * vnr.y, g.y * vnr.z,
                      g.z * vnr.x, g.z * vnr.y, g.z * vnr.z};
  const f4 r = bl.r, dd = bl.d;
  f4 gr = dR_dq_contract(r, G);
  gr.x += 2.f * (-dd.w * g.x - dd.z * g.y + dd.y * g.z);
  gr.y += 2.f * (dd.z * g.x - dd.w * g.y - dd.x * g.z);
  gr.z += 2.f * (-dd.y * g.x + dd.x * g.y - dd.w * g.z);
  gr.w += 2.f * (dd.x * g.x + dd.y * g.y + dd.z * g.z);
  const f4 gd = mk4(2.f * (r.w * g.x + r.z * g.y - r.y * g.z), 2.f * (-r.z * g.x + r.w * g.y + r.x * g.z),
                    2.f * (r.y * g.x - r.x * g.y + r.w * g.z), 2.f * (-r.x * g.x - r.y * g.y - r.z * g.z));
  f4 gbr, gbd;
  if (bl.small) {
    gbr = gr; gbd = gd;
  } else {
    const float proj = dot(gr, r) + dot(gd, dd);
    gbr = (gr - r * proj) * bl.scale;
    gbd = gd * bl.scale;
  }
  // the joint DQ gradients sum w_k * s_k * (gbr, gbd) over vertices: kept per
  // vertex and reduced per joint in a fixed order by joint_reduce_kernel
  float* q8 = d.vq + 8 * (size_t)gv;
  const f3 u = mulT(R, g);
  float* up = d.u + 3 * (size_t)gv;
  if (!SIL) {
    q8[0] = gbr.x; q8[1] = gbr.y; q8[2] = gbr.z; q8[3] = gbr.w;
    q8[4] = gbd.x; q8[5] = gbd.y; q8[6] = gbd.z; q8[7] = gbd.w;
    for (int k = 0; k < KS; ++k) d.vws[KS * (size_t)gv + k] = w[k] * bl.sgn[k];
    up[0] = u.x; up[1] = u.y; up[2] = u.z;
  } else {
    q8[0] += gbr.x; q8[1] += gbr.y; q8[2] += gbr.z; q8[3] += gbr.w;
    q8[4] += gbd.x; q8[5] += gbd.y; q8[6] += gbd.z; q8[7] += gbd.w;
    up[0] += u.x; up[1] += u.y; up[2] += u.z;
  }
}

// Joint DQ gradients in two deterministic passes.  The (vertex, slot) entries
// of every joint are cut into chunks of at most kJChunk (a static table), so
// the work spreads over thousands of wavefronts however skewed the joints'
// vertex counts are.  Pass 1: one wavefront per chunk, lane-strided then
// reduced across the wavefront.
__global__ __launch_bounds__(kB) void joint_partial_kernel(Dev d, int nChunk) {
#pragma clang fp contract(fast)
  const int ck = (blockIdx.x * kB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (ck >= nChunk) return;
  const int4 r = d.jchunk[ck];
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  // a chunk has at most kJChunk = 8 x 64 entries: every lane's (at most 8)
  // entry ids are loaded first, so the entries' data loads do not wait on
  // one another; the sum still runs over the lane's entries in order
  static_assert(kJChunk <= 8 * 64, "a lane takes at most 8 entries of a chunk");
  int ent[8];
#pragma unroll
  for (int it = 0; it < 8; ++it) {
    const int e = r.x + lane + 64 * it;
    ent[it] = e < r.y ? d.jv_ent[e] : -1;
  }
#pragma unroll
  for (int it = 0; it < 8; ++it) {
    if (ent[it] < 0) break;
    const size_t gv = (size_t)r.z + (ent[it] >> 2);
    const float ws = d.vws[KS * gv + (ent[it] & 3)];
    const f4* q8 = (const f4*)(d.vq + 8 * gv);
    const f4 qa = q8[0], qb = q8[1];
    acc[0] += ws * qa.x; acc[1] += ws * qa.y; acc[2] += ws * qa.z; acc[3] += ws * qa.w;
    acc[4] += ws * qb.x; acc[5] += ws * qb.y; acc[6] += ws * qb.z; acc[7] += ws * qb.w;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = wave_sum(acc[q]);
  if (lane < 8) {
    float v = acc[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) v = lane == q ? acc[q] : v;
    d.jpart[8 * (size_t)ck + lane] = v;
  }
}

// Pass 2: one wavefront per joint: lane l sums the joint's chunk partials
// l, l + 64, ... in order, then the lanes are reduced across the wavefront in
// a fixed order (a joint without entries gets 0).
__global__ __launch_bounds__(kB) void joint_reduce_kernel(Dev d, IO io, int totJ) {
#pragma clang fp contract(fast)
  const int j = (blockIdx.x * kB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (j >= totJ) return;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int k = d.jc_off[j] + lane; k < d.jc_off[j + 1]; k += 64) {
    const f4* pp = reinterpret_cast<const f4*>(d.jpart + 8 * (size_t)k);
    const f4 a = pp[0], b = pp[1];
    acc[0] += a.x; acc[1] += a.y; acc[2] += a.z; acc[3] += a.w;
    acc[4] += b.x; acc[5] += b.y; acc[6] += b.z; acc[7] += b.w;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = wave_sum(acc[q]);
  if (lane < 8) {
    float v = acc[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) v = lane == q ? acc[q] : v;
    io.g_dq[8 * (size_t)j + lane] = v;
  }
}

__global__ __launch_bounds__(kB) void node_bwd_kernel(Dev d, IO io, int totM) {
#pragma clang fp contract(fast)
  const int node = (blockIdx.x * kB + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (node >= totM) return;
  // character of this node: linear scan over the (few) characters
  int b = 0;
  while (b + 1 < d.B && d.ch[b + 1].m0 <= node) ++b;
  const CharD c = d.ch[b];
  const EulerSC sd = euler_sc(ld3(io.dA + 3 * (size_t)node));
  const m3 da = euler_dA(sd), db = euler_dB(sd), dg = euler_dG(sd);
  const f3 gn = ld3(d.rest + 3 * ((size_t)c.v0 + d.node_vtx[node]));
  float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int e = d.n2v_off[node] + lane; e < d.n2v_off[node + 1]; e += 64) {
    const int gv = c.v0 + d.n2v_vtx[e];
    const float wgt = d.n2v_w[e];
    const f3 uv = ld3(d.u + 3 * (size_t)gv);
    const f3 q = ld3(d.rest + 3 * (size_t)gv) - gn;
    acc[0] += wgt * uv.x; acc[1] += wgt * uv.y; acc[2] += wgt * uv.z;
    acc[3] += wgt * dot(mul(da, q), uv);
    acc[4] += wgt * dot(mul(db, q), uv);
    acc[5] += wgt * dot(mul(dg, q), uv);
  }
  for (int k = 0; k < 6; ++k) acc[k] = wave_sum(acc[k]);
  if (lane == 0) {
    io.g_dT[3 * (size_t)node] = acc[0]; io.g_dT[3 * (size_t)node + 1] = acc[1]; io.g_dT[3 * (size_t)node + 2] = acc[2];
    io.g_dA[3 * (size_t)node] = acc[3]; io.g_dA[3 * (size_t)node + 1] = acc[4]; io.g_dA[3 * (size_t)node + 2] = acc[5];
  }
}

// One block per character: the losses from the per-block partial sums, in a
// fixed order.
__global__ __launch_bounds__(kB) void finalize_kernel(Dev d, IO io) {
  __shared__ float s_red[kB / 64];
  const int b = blockIdx.x;
  const CharD c = d.ch[b];
  const ViewD v0 = d.vw[c.view0], v1 = d.vw[c.view0 + c.C - 1];
  // the character's tile partials: 8 independent running sums per lane (8
  // loads in flight), added in a fixed order
  float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const long long tend = v1.tile0 + v1.tiles;
  long long t = v0.tile0 + threadIdx.x;
  for (; t + 7 * kB < tend; t += 8 * kB)
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += d.loss_tile[t + u * kB];
  for (; t < tend; t += kB) a[0] += d.loss_tile[t];
  float s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  const float lc = block_sum(s, s_red);
  s = 0.f;
  for (long long i = (long long)c.view0 * d.gxN + threadIdx.x; i < (long long)(c.view0 + c.C) * d.gxN; i += kB)
    s += d.loss_pt[i];
  for (long long i = (long long)c.view0 * d.gxS + threadIdx.x; i < (long long)(c.view0 + c.C) * d.gxS; i += kB)
    s += d.loss_px[i];
  const float ls = block_sum(s, s_red);
  if (threadIdx.x == 0) { io.loss[2 * b] = lc; io.loss[2 * b + 1] = ls; }
}

// Tile shapes: (1 << tws) x (1 << (8 - tws)) pixels, tws in [4, 6] (16x16,
// 32x8, 64x4), one lane per pixel in the fine kernel.
constexpr int kMinTileShift = 4, kMaxTileShift = 6, kDefaultTileShift = 5;

// Per-view tile counts and offsets for one tile shape; returns the total.
long long layout_tiles(std::vector<ViewD>& vw, int tws, int& max_tiles) {
  const int tw = 1 << tws, th = 1 << (8 - tws);
  long long tot = 0;
  max_tiles = 0;
  for (ViewD& v : vw) {
    v.tiles_x = (v.W + tw - 1) / tw;
    v.tiles = v.tiles_x * ((v.H + th - 1) / th);
    v.tile0 = tot;
    tot += v.tiles;
    max_tiles = std::max(max_tiles, v.tiles);
  }
  return tot;
}

template <class T>
hipError_t up(T** dst, const std::vector<T>& src) {
  hipError_t e = hipMalloc(dst, sizeof(T) * std::max<size_t>(src.size(), 1));
  if (e != hipSuccess || src.empty()) return e;
  return hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
}

}  // namespace
}  // namespace dcx

using namespace dcx;

struct dcx_pipeline {
  Dev d{};
  int B = 0, V = 0, sumN = 0, sumF = 0, sumM = 0, sumJ = 0, nChunk = 0;
  int maxTilesAll = 0;           // over the tile shapes, for the tile arrays
  long long totTilesAll = 0;
  std::vector<ViewD> vw_host;    // view table of the current tile shape
  long long sumS2 = 0;
  std::vector<void*> allocs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // second stream (default priority), the fork event and the join events of
  // the work it takes beside the main stream (see pipe_enqueue)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_njoin = nullptr;
};

static void pipe_free(dcx_pipeline* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->ev_njoin) (void)hipEventDestroy(p->ev_njoin);
  if (p->side) (void)hipStreamDestroy(p->side);
  for (void* a : p->allocs) (void)hipFree(a);
  delete p;
}

extern "C" dcx_status dcx_pipeline_create(int n_chars, const dcx_char_desc* chars, long long bin_capacity,
                                          dcx_pipeline** out) {
  DCX_REQUIRE(out && chars && n_chars > 0, "dcx_pipeline_create: bad arguments");
  dcx_pipeline* p = new dcx_pipeline();
  p->B = n_chars;
  std::vector<CharD> ch(n_chars);
  std::vector<ViewD> vw;
  std::vector<float> rest, jw, ew, tcv;
  std::vector<int> faces, jidx, enode, node_vtx, n2v_off(1, 0), n2v_vtx, vf_off(1, 0), vf_list, jv_off(1, 0), jv_ent;
  std::vector<int> corner_slot;
  std::vector<int> jc_off(1, 0);
  std::vector<int4> jchunk;
  std::vector<float> n2v_w;
  long long vn = 0, vf = 0, pix = 0, sil = 0, tiles = 0, cells = 0, mask_words = 0;
  int maxN = 0, maxF = 0, maxM = 0, maxS2 = 0, maxJ = 0, maxTiles = 0;
  for (int b = 0; b < n_chars; ++b) {
    const dcx_char_desc& c = chars[b];
    if (!(c.num_vertices > 0 && c.num_faces > 0 && c.num_faces < (1 << (32 - kTagBits)) && c.num_nodes > 0 &&
          c.num_joints > 0 && c.num_joints <= 256 &&
          c.num_cameras > 0 && c.width > 0 && c.height > 0 && c.dt_size > 2 && c.dt_size <= 16384 &&
          c.width <= 32767 && c.height <= 32767)) {
      delete p;
      return set_error(DCX_ERR_INVALID, "dcx_pipeline_create: character %d has invalid sizes", b);
    }
    CharD& d = ch[b];
    d.N = c.num_vertices; d.F = c.num_faces; d.M = c.num_nodes; d.J = c.num_joints;
    d.v0 = p->sumN; d.f0 = p->sumF; d.m0 = p->sumM; d.j0 = p->sumJ;
    d.view0 = (int)vw.size(); d.C = c.num_cameras;
    d.has_tc = c.texcoords != nullptr;
    for (int i = 0; i < 6 * d.F; ++i) tcv.push_back(c.texcoords ? c.texcoords[i] : 0.f);
    for (int i = 0; i < 3 * d.N; ++i) rest.push_back(c.rest_pos[i]);
    for (int i = 0; i < 3 * d.F; ++i) {
      if (c.faces[i] < 0 || c.faces[i] >= d.N) { delete p; return set_error(DCX_ERR_INVALID, "character %d: bad face index", b); }
      faces.push_back(c.faces[i]);
    }
    for (int i = 0; i < KS * d.N; ++i) {
      if (c.joint_idx[i] < 0 || c.joint_idx[i] >= d.J) { delete p; return set_error(DCX_ERR_INVALID, "character %d: bad joint index", b); }
      jidx.push_back(c.joint_idx[i]);
      jw.push_back(c.joint_w[i]);
    }
    for (int i = 0; i < KE * d.N; ++i) {
      if (c.node_idx[i] < 0 || c.node_idx[i] >= d.M) { delete p; return set_error(DCX_ERR_INVALID, "character %d: bad node index", b); }
      enode.push_back(c.node_idx[i]);
      ew.push_back(c.node_w[i]);
    }
    for (int i = 0; i < d.M; ++i) {
      if (c.node_vertex[i] < 0 || c.node_vertex[i] >= d.N) { delete p; return set_error(DCX_ERR_INVALID, "character %d: bad node vertex", b); }
      node_vtx.push_back(c.node_vertex[i]);
    }
    // node -> vertex CSR (vertices ascending)
    std::vector<int> cnt(d.M + 1, 0);
    for (int i = 0; i < KE * d.N; ++i) cnt[c.node_idx[i] + 1]++;
    for (int m = 0; m < d.M; ++m) cnt[m + 1] += cnt[m];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    const int base = (int)n2v_vtx.size();
    n2v_vtx.resize(base + KE * d.N);
    n2v_w.resize(base + KE * d.N);
    for (int v = 0; v < d.N; ++v)
      for (int k = 0; k < KE; ++k) {
        const int m = c.node_idx[KE * v + k];
        const int slot = base + fill[m]++;
        n2v_vtx[slot] = v;
        n2v_w[slot] = c.node_w[KE * v + k];
      }
    for (int m = 0; m < d.M; ++m) n2v_off.push_back(base + cnt[m + 1]);
    // vertex -> face CSR (faces ascending per vertex)
    std::vector<int> vc(d.N + 1, 0);
    for (int i = 0; i < 3 * d.F; ++i) vc[c.faces[i] + 1]++;
    for (int v = 0; v < d.N; ++v) vc[v + 1] += vc[v];
    std::vector<int> vfill(vc.begin(), vc.end() - 1);
    const int vbase = (int)vf_list.size();
    vf_list.resize(vbase + 3 * d.F);
    corner_slot.resize((size_t)3 * (d.f0 + d.F));
    for (int f = 0; f < d.F; ++f)
      for (int k = 0; k < 3; ++k) {
        const int slot = vbase + vfill[c.faces[3 * f + k]]++;
        vf_list[slot] = f;
        corner_slot[(size_t)3 * (d.f0 + f) + k] = slot;
      }
    for (int v = 0; v < d.N; ++v) vf_off.push_back(vbase + vc[v + 1]);
    // joint -> (vertex, slot) CSR, entries in increasing (vertex, slot) order
    std::vector<int> jc(d.J + 1, 0);
    for (int i = 0; i < KS * d.N; ++i) jc[c.joint_idx[i] + 1]++;
    for (int j = 0; j < d.J; ++j) jc[j + 1] += jc[j];
    std::vector<int> jfill(jc.begin(), jc.end() - 1);
    const int jbase = (int)jv_ent.size();
    jv_ent.resize(jbase + KS * d.N);
    for (int i = 0; i < KS * d.N; ++i) jv_ent[jbase + jfill[c.joint_idx[i]]++] = i;  // i = v * 4 + slot
    for (int j = 0; j < d.J; ++j) {
      jv_off.push_back(jbase + jc[j + 1]);
      for (int e = jbase + jc[j]; e < jbase + jc[j + 1]; e += kJChunk)
        jchunk.push_back(make_int4(e, std::min(e + kJChunk, jbase + jc[j + 1]), p->sumN, p->sumJ + j));
      jc_off.push_back((int)jchunk.size());
    }
    for (int k = 0; k < d.C; ++k) {
      ViewD v{};
      v.ch = b; v.W = c.width; v.H = c.height; v.S = c.dt_size;
      v.vn0 = vn; v.vf0 = vf; v.pix0 = pix; v.sil0 = sil; v.cell0 = cells;
      v.mpitch = 2 * (int)ceil_div(c.width, 64);
      v.mask0 = mask_words;
      mask_words += (long long)v.mpitch * c.height;
      v.gx = c.dt_size * kRing / 30 + 2 * kGridMargin + 3; v.gy = v.gx;
      vn += d.N; vf += d.F; pix += (long long)c.width * c.height; sil += (long long)c.dt_size * c.dt_size;
      cells += (long long)v.gx * v.gy;
      vw.push_back(v);
    }
    p->sumN += d.N; p->sumF += d.F; p->sumM += d.M; p->sumJ += d.J;
    maxN = std::max(maxN, d.N); maxF = std::max(maxF, d.F); maxM = std::max(maxM, d.M); maxJ = std::max(maxJ, d.J);
    maxS2 = std::max(maxS2, c.dt_size * c.dt_size);
  }
  p->V = (int)vw.size();
  p->sumS2 = sil;
  // tile arrays are sized for the largest count over the tile shapes
  for (int tws = kMinTileShift; tws <= kMaxTileShift; ++tws) {
    int mt = 0;
    const long long t = layout_tiles(vw, tws, mt);
    tiles = std::max(tiles, t);
    maxTiles = std::max(maxTiles, mt);
  }
  p->maxTilesAll = maxTiles;
  p->totTilesAll = tiles;
  {
    int mt = 0;
    const long long t = layout_tiles(vw, kDefaultTileShift, mt);
    p->d.tws = kDefaultTileShift;
    p->d.ths = 8 - kDefaultTileShift;
    p->d.maxTiles = mt;
    p->d.totTiles = t;
  }
  p->vw_host = vw;
  Dev& d = p->d;
  d.B = n_chars; d.V = p->V; d.maxN = maxN; d.maxF = maxF; d.maxM = maxM; d.maxS2 = maxS2; d.maxJ = maxJ;
  d.gxN = (int)ceil_div(maxN, kB);
  {
    int maxS = 0;
    for (const ViewD& v : vw) maxS = std::max(maxS, v.S);
    const int tpr = (maxS + kDtTile - 1) / kDtTile;
    d.gxS = tpr * tpr;  // DT-pixel tiles per view
  }
  d.totVN = vn; d.totVF = vf; d.totPix = pix; d.totCells = cells;
  const long long cap = bin_capacity > 0 ? bin_capacity : std::max<long long>(4 * vf, 1 << 20);
  if (cap >= (1ll << 32) || vf >= (1ll << 31) || (long long)p->V * maxTiles >= (1ll << 31)) {
    delete p;
    return set_error(DCX_ERR_INVALID, "dcx_pipeline_create: batch too large (records %lld, tiles %lld)", cap, tiles);
  }
  if (p->V > 65535) {  // views index the grids' y extent
    delete p;
    return set_error(DCX_ERR_INVALID, "dcx_pipeline_create: %d views, at most 65535", p->V);
  }
  d.capacity = (unsigned int)cap;
  const long long aux = std::max(scan_aux_elems(tiles), scan_aux_elems(cells));
  hipError_t e = hipSuccess;
  auto track = [&](void* a) { p->allocs.push_back(a); };
  auto alloc = [&](auto** dst, size_t bytes) {
    if (e != hipSuccess) return;
    e = hipMalloc((void**)dst, std::max<size_t>(bytes, 4));
    if (e == hipSuccess) track((void*)*dst);
  };
  auto upload = [&](auto** dst, const auto& v) {
    if (e != hipSuccess) return;
    e = up(dst, v);
    if (e == hipSuccess) track((void*)*dst);
  };
  CharD* dch; ViewD* dvw;
  float* drest; int *dfaces, *djidx, *denode, *dnv, *dn2o, *dn2v, *dvfo, *dvfl, *djvo, *djve, *dcsl;
  float *djw, *dew, *dn2w;
  upload(&dch, ch); upload(&dvw, vw);
  upload(&drest, rest); upload(&dfaces, faces); upload(&djidx, jidx); upload(&djw, jw);
  upload(&denode, enode); upload(&dew, ew); upload(&dnv, node_vtx); upload(&dn2o, n2v_off);
  upload(&dn2v, n2v_vtx); upload(&dn2w, n2v_w); upload(&dvfo, vf_off); upload(&dvfl, vf_list);
  upload(&djvo, jv_off); upload(&djve, jv_ent); upload(&dcsl, corner_slot);
  int4* djck; int* djco;
  upload(&djck, jchunk); upload(&djco, jc_off);
  float* dtc;
  upload(&dtc, tcv);
  d.tc = dtc;
  d.jchunk = djck; d.jc_off = djco;
  p->nChunk = (int)jchunk.size();
  d.ch = dch; d.vw = dvw; d.rest = drest; d.faces = dfaces; d.jidx = djidx; d.jw = djw;
  d.enode = denode; d.ew = dew; d.node_vtx = dnv; d.n2v_off = dn2o; d.n2v_vtx = dn2v; d.n2v_w = dn2w;
  d.vf_off = dvfo; d.vf_list = dvfl; d.jv_off = djvo; d.jv_ent = djve; d.corner_slot = dcsl;
  alloc(&d.node_tab, sizeof(float) * kNodeRow * (size_t)p->sumM);
  alloc(&d.node_skin, sizeof(float) * 12 * (size_t)p->sumM);
  alloc(&d.vdef, sizeof(float) * 3 * (size_t)p->sumN);
  alloc(&d.vnrm, sizeof(float) * 3 * (size_t)p->sumN);
  alloc(&d.vsum, sizeof(float) * 3 * (size_t)p->sumN);
  alloc(&d.proj, sizeof(f4) * (size_t)vn);
  alloc(&d.ke, sizeof(float) * 12 * (size_t)p->V);
  alloc(&d.nrm2, sizeof(f2) * (size_t)vn);
  alloc(&d.tri, sizeof(TriRec) * (size_t)vf);
  alloc(&d.trange, sizeof(uint2) * (size_t)vf);
  alloc(&d.tri_blk, 4 * (size_t)p->V * (size_t)ceil_div(maxF, kB));
  alloc(&d.tri_base, 4 * (size_t)(vf + 1));
  alloc(&d.tile_cnt, 4 * (size_t)tiles);
  alloc(&d.tile_off, 4 * (size_t)(tiles + 1));
  alloc(&d.tile_fill, 4 * (size_t)tiles);
  alloc(&d.scan_aux, 4 * (size_t)aux);
  alloc(&d.heavy, 4 * (size_t)kHeavyCap);
  alloc(&d.heavy_n, 4);
  alloc(&d.bins, 4 * (size_t)cap);
  alloc(&d.rec, sizeof(float) * kGrad * (size_t)cap);
  alloc(&d.rec_vis, (size_t)cap);
  alloc(&d.isb, (size_t)vn);
  alloc(&d.fgmask, 4 * (size_t)mask_words);
  d.totMask = mask_words;
  alloc(&d.cell_cnt, 4 * (size_t)cells);
  alloc(&d.cell_off, 4 * (size_t)(cells + 1));
  alloc(&d.cell_fill, 4 * (size_t)cells);
  alloc(&d.cell_pts, sizeof(f4) * (size_t)vn);
  alloc(&d.ov_pts, sizeof(f4) * (size_t)vn);
  alloc(&d.ov_cnt, 4 * (size_t)p->V);
  d.cell_cap = 0;  // counting sort until the autotuner or dcx_pipeline_set_grid picks fixed slots
  d.cell_slots = nullptr;
  alloc(&d.corner_sum, sizeof(float) * kGrad * (size_t)p->sumF);
  alloc(&d.g_v, sizeof(float) * 3 * (size_t)p->sumN);
  alloc(&d.g_pts, sizeof(float) * 2 * (size_t)vn);
  alloc(&d.g_fx, sizeof(long long) * 2 * (size_t)vn);
  alloc(&d.u, sizeof(float) * 3 * (size_t)p->sumN);
  alloc(&d.loss_tile, sizeof(float) * (size_t)tiles);
  alloc(&d.loss_pt, sizeof(float) * (size_t)p->V * d.gxN);
  alloc(&d.loss_px, sizeof(float) * (size_t)p->V * d.gxS);
  alloc(&d.vq, sizeof(float) * 8 * (size_t)p->sumN);
  alloc(&d.vws, sizeof(float) * KS * (size_t)p->sumN);
  alloc(&d.jpart, sizeof(float) * 8 * (size_t)p->nChunk);
  if (e != hipSuccess) {
    pipe_free(p);
    return set_error(DCX_ERR_OOM, "dcx_pipeline_create: %s", hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_njoin, hipEventDisableTiming);
  if (e != hipSuccess) {
    pipe_free(p);
    return set_error(DCX_ERR_HIP, "dcx_pipeline_create: %s", hipGetErrorString(e));
  }
  *out = p;
  return DCX_OK;
}

extern "C" dcx_status dcx_pipeline_destroy(dcx_pipeline* p) {
  if (p) pipe_free(p);
  return DCX_OK;
}

extern "C" dcx_status dcx_pipeline_sizes(const dcx_pipeline* p, long long* sizes) {
  DCX_REQUIRE(p && sizes, "dcx_pipeline_sizes: null argument");
  sizes[0] = p->B; sizes[1] = p->V; sizes[2] = p->sumN; sizes[3] = p->sumF; sizes[4] = p->sumM; sizes[5] = p->sumJ;
  sizes[6] = p->d.totPix; sizes[7] = p->sumS2; sizes[8] = p->d.totTiles; sizes[9] = p->d.totVF;
  return DCX_OK;
}

// Launch order of one step; dcx_pipeline_profile times each entry.
static const char* const kStageNames[] = {
    "clear",          "node_fwd",      "vertex_fwd",       "normals",   "view_normals", "project",
    "tri_count",      "tile_scan",     "tri_scan",         "bin_scatter", "fine",       "sil_vertex",
    "sil_grid_scan",  "sil_grid_scatter", "sil_pixel",     "face_sum",  "vertex_bwd_raster",
    "vertex_bwd_sil", "joint_partial", "joint_reduce",     "node_bwd",  "finalize"};
constexpr int kNumStages = (int)(sizeof(kStageNames) / sizeof(kStageNames[0]));

// ev == nullptr: plain enqueue.  Otherwise ev[i] is recorded before stage i
// and ev[kNumStages] after the last one.
static dcx_status pipe_enqueue(dcx_pipeline* p, const dcx_pipeline_io* in, hipStream_t s,
                               hipEvent_t* ev = nullptr) {
  DCX_REQUIRE((in->skinned_t == nullptr) == (in->skinned_a == nullptr),
              "dcx_pipeline: skinned_t and skinned_a are both given or both NULL");
  const bool graph_skin = in->skinned_a != nullptr;
  DCX_REQUIRE((graph_skin || (in->dq && in->g_dq)) && in->delta_t && in->delta_a && in->colors && in->extrinsics &&
                  in->intrinsics && in->target && in->dt && in->crops && in->face && in->bary && in->render &&
                  in->loss && in->g_delta_t && in->g_delta_a && in->g_colors,
              "dcx_pipeline: null buffer in io");
  Dev d = p->d;
  IO io{in->dq, in->delta_t, in->delta_a, in->colors, in->extrinsics, in->intrinsics, in->target, in->dt, in->crops,
        in->w_color, in->w_sil, in->face, in->bary, in->render, in->vertices, in->loss, in->g_dq, in->g_delta_t,
        in->g_delta_a, in->g_colors, in->uv, in->normal, in->skinned_t, in->skinned_a};
  ClearList L{};
  auto add = [&](void* ptr, long long n) { L.p[L.count] = (unsigned int*)ptr; L.n[L.count] = n; L.count++; };
  add(d.g_v, 3ll * p->sumN);
  add(in->g_colors, 3ll * p->sumN);
  add(d.tile_cnt, d.totTiles);
  add(d.tile_fill, d.totTiles);
  add(d.cell_cnt, d.totCells);
  if (d.cell_cap > 0) add(d.ov_cnt, p->V);
  else add(d.cell_fill, d.totCells);
  add(d.heavy_n, 1);
  add(d.fgmask, d.totMask);
  int stage = 0;
  auto mark = [&]() { if (ev) (void)hipEventRecord(ev[stage], s); stage++; };
  const int B = p->B, V = p->V;
  // the launches of the stages, each on a given stream
  auto k_node_fwd = [&](hipStream_t q) {
    hipLaunchKernelGGL(node_fwd_kernel, dim3(ceil_div(d.maxM, kB), B), dim3(kB), 0, q, d, io);
  };
  auto k_vertex_fwd = [&](hipStream_t q) {
    hipLaunchKernelGGL(vertex_fwd_kernel, dim3(ceil_div(d.maxN, kB), B), dim3(kB), 0, q, d, io, L);
  };
  auto k_normals = [&](hipStream_t q) {
    hipLaunchKernelGGL(normals_kernel, dim3(ceil_div(d.maxN, kB), B), dim3(kB), 0, q, d);
  };
  auto k_view_normal = [&](hipStream_t q) {
    hipLaunchKernelGGL(view_normal_kernel, dim3(ceil_div(d.maxN, kB), V), dim3(kB), 0, q, d, io);
  };
  auto k_project = [&](hipStream_t q) {
    hipLaunchKernelGGL(project_kernel, dim3(ceil_div(d.maxN, kB), V), dim3(kB), 0, q, d, io);
  };
  auto k_tri_count = [&](hipStream_t q) {
    hipLaunchKernelGGL(tri_count_kernel, dim3(ceil_div(d.maxF, kB), V), dim3(kB), 0, q, d, io);
  };
  // tile counts -> tile offsets, and in the same first launch the
  // per-workgroup record-slot counts -> record-slot offsets
  auto k_tile_scan = [&](hipStream_t q) -> dcx_status {
    const long long nt = d.totTiles, nb = (nt + kScanChunk - 1) / kScanChunk;
    if (nb < 1 || nb >= (1ll << 31) - 1) return set_error(DCX_ERR_INVALID, "scan: bad length %lld", nt);
    hipLaunchKernelGGL(scan_sum_with_one_kernel, dim3((unsigned)nb + 1), dim3(kB), 0, q, d.tile_cnt, d.scan_aux, nt,
                       d.tri_blk, (long long)V * ceil_div(d.maxF, kB), d.tri_base + d.totVF);
    hipLaunchKernelGGL(tile_scan_apply_kernel, dim3((unsigned)nb), dim3(kB), 0, q, d.tile_cnt, d.tile_off, d.scan_aux,
                       nt, d.heavy, d.heavy_n, d.capacity);
    return check_launch("scan");
  };
  auto k_bin_scatter = [&](hipStream_t q) {
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(ceil_div(d.maxF, kB), V), dim3(kB), 0, q, d);
  };
  auto k_fine = [&](hipStream_t q) {
    // the slow-path tiles' workgroups first, then per view a multiple of kXcd
    const dim3 grid((unsigned)(kHeavyCap + (long long)ceil_div(d.maxTiles, kXcd) * kXcd * V));
    const dim3 blk(kTilePix);
    if (io.uv || io.nrm_out) {
      if (d.tws == 4) hipLaunchKernelGGL((fine_kernel<4, true>), grid, blk, 0, q, d, io);
      else if (d.tws == 6) hipLaunchKernelGGL((fine_kernel<6, true>), grid, blk, 0, q, d, io);
      else hipLaunchKernelGGL((fine_kernel<5, true>), grid, blk, 0, q, d, io);
    } else {
      if (d.tws == 4) hipLaunchKernelGGL((fine_kernel<4, false>), grid, blk, 0, q, d, io);
      else if (d.tws == 6) hipLaunchKernelGGL((fine_kernel<6, false>), grid, blk, 0, q, d, io);
      else hipLaunchKernelGGL((fine_kernel<5, false>), grid, blk, 0, q, d, io);
    }
  };
  auto k_face_sum = [&](hipStream_t q) {
    hipLaunchKernelGGL(face_sum_kernel, dim3(ceil_div(d.maxF, kB), B), dim3(kB), 0, q, d);
  };
  auto k_vertex_bwd_raster = [&](hipStream_t q) {
    hipLaunchKernelGGL(vertex_bwd_kernel<false>, dim3(d.gxN, B), dim3(kB), 0, q, d, io);
  };
  auto k_sil_vertex = [&](hipStream_t q) {
    hipLaunchKernelGGL(sil_vertex_kernel_p, dim3(d.gxN, V), dim3(kB), 0, q, d, io);
  };
  // the counting-sort layout's cell offsets and scatter (nothing to do with fixed slots)
  auto k_sil_grid_scan = [&](hipStream_t q) -> dcx_status {
    return d.cell_cap > 0 ? DCX_OK : scan_big(d.cell_cnt, d.cell_off, d.scan_aux, d.totCells, q);
  };
  auto k_sil_grid_scatter = [&](hipStream_t q) {
    if (d.cell_cap == 0)
      hipLaunchKernelGGL(sil_grid_scatter_p, dim3(ceil_div(d.maxN, kB), V), dim3(kB), 0, q, d, io);
  };
  auto k_sil_pixel = [&](hipStream_t q) {
    hipLaunchKernelGGL(sil_pixel_kernel_p, dim3(d.gxS, V), dim3(kB), 0, q, d, io);
  };
  auto k_vertex_bwd_sil = [&](hipStream_t q) {
    hipLaunchKernelGGL(vertex_bwd_kernel<true>, dim3(d.gxN, B), dim3(kB), 0, q, d, io);
  };
  // (graph skinning has no joint DQs: the two joint stages stay empty)
  auto k_joint_partial = [&](hipStream_t q) {
    if (!graph_skin)
      hipLaunchKernelGGL(joint_partial_kernel, dim3(ceil_div((long long)p->nChunk * 64, kB)), dim3(kB), 0, q, d,
                         p->nChunk);
  };
  auto k_joint_reduce = [&](hipStream_t q) {
    if (!graph_skin)
      hipLaunchKernelGGL(joint_reduce_kernel, dim3(ceil_div((long long)p->sumJ * 64, kB)), dim3(kB), 0, q, d, io,
                         p->sumJ);
  };
  auto k_node_bwd = [&](hipStream_t q) {
    hipLaunchKernelGGL(node_bwd_kernel, dim3(ceil_div((long long)p->sumM * 64, kB)), dim3(kB), 0, q, d, io, p->sumM);
  };
  auto k_finalize = [&](hipStream_t q) { hipLaunchKernelGGL(finalize_kernel, dim3(B), dim3(kB), 0, q, d, io); };

  if (ev) {
    // A profiled step: every stage on s, in stage order, bracketed by events.
    mark();
    mark();  // the clears run inside vertex_fwd_kernel
    k_node_fwd(s); mark();
    k_vertex_fwd(s); mark();
    k_normals(s); mark();
    k_view_normal(s); mark();
    k_project(s); mark();
    k_tri_count(s); mark();
    if (dcx_status e = k_tile_scan(s)) return e;
    mark();
    mark();  // the record-slot scan ran inside the tile scan's first launch
    k_bin_scatter(s); mark();
    k_fine(s); mark();
    k_sil_vertex(s); mark();
    if (dcx_status e = k_sil_grid_scan(s)) return e;
    mark();
    k_sil_grid_scatter(s); mark();
    k_sil_pixel(s); mark();
    k_face_sum(s); mark();
    k_vertex_bwd_raster(s); mark();
    k_vertex_bwd_sil(s); mark();
    k_joint_partial(s); mark();
    k_joint_reduce(s); mark();
    k_node_bwd(s); mark();
    k_finalize(s); mark();
    return check_launch("dcx_pipeline_step");
  }

  // Eager and captured steps run on two streams.  The side stream takes the
  // vertex normals, their image-space directions and the vertex projections
  // (read by the fine kernel and the silhouette stages; the binning projects
  // its corners itself) beside the binning, the per-face record sums and the
  // raster part of the vertex backward beside the silhouette stages, and the
  // loss totals and the node gradients beside the joint-DQ reduction.  Each
  // fork records an event on s and, before the side stream waits for it,
  // launches the next critical-path stage on s, so that in a captured graph
  // that stage is the fork's first dependent: the graph keeps a node's first
  // dependent on the node's own hardware queue, and every hand-over between
  // queues costs ~10-20 us, so the critical path stays on one queue.  Each
  // join is an event of its own.  The buffer clears run inside the vertex
  // kernel on s (a join would cost more than they do).
  hipStream_t sd = p->side;
  auto fork_rec = [&]() -> dcx_status {
    DCX_HIP(hipEventRecord(p->ev_fork, s));
    return DCX_OK;
  };
  auto fork_wait = [&]() -> dcx_status {
    DCX_HIP(hipStreamWaitEvent(sd, p->ev_fork, 0));
    return DCX_OK;
  };
  k_node_fwd(s);
  k_vertex_fwd(s);
  if (dcx_status e = fork_rec()) return e;
  k_tri_count(s);
  if (dcx_status e = fork_wait()) return e;
  k_normals(sd);
  k_view_normal(sd);
  k_project(sd);
  DCX_HIP(hipEventRecord(p->ev_njoin, sd));
  if (dcx_status e = k_tile_scan(s)) return e;
  k_bin_scatter(s);
  // the fine kernel reads the side stream's projections, K [R | t] and vertex normals
  DCX_HIP(hipStreamWaitEvent(s, p->ev_njoin, 0));
  k_fine(s);
  if (dcx_status e = fork_rec()) return e;
  k_sil_vertex(s);
  if (dcx_status e = fork_wait()) return e;
  // the raster part of the vertex backward follows the per-face sums there
  k_face_sum(sd);
  k_vertex_bwd_raster(sd);
  DCX_HIP(hipEventRecord(p->ev_join, sd));
  if (dcx_status e = k_sil_grid_scan(s)) return e;
  k_sil_grid_scatter(s);
  k_sil_pixel(s);
  // every loss partial is written: the totals go to the side stream
  if (dcx_status e = fork_rec()) return e;
  DCX_HIP(hipStreamWaitEvent(s, p->ev_join, 0));  // face_sum and the raster vertex part done
  k_vertex_bwd_sil(s);
  if (dcx_status e = fork_wait()) return e;
  k_finalize(sd);
  // the node gradients read the vertex pass's ED-position gradient
  if (dcx_status e = fork_rec()) return e;
  k_joint_partial(s);
  if (dcx_status e = fork_wait()) return e;
  k_node_bwd(sd);
  DCX_HIP(hipEventRecord(p->ev_join, sd));
  k_joint_reduce(s);
  DCX_HIP(hipStreamWaitEvent(s, p->ev_join, 0));  // the node gradients and the loss totals (all side work) done
  return check_launch("dcx_pipeline_step");
}

extern "C" dcx_status dcx_pipeline_step(dcx_pipeline* p, const dcx_pipeline_io* io, void* stream) {
  DCX_REQUIRE(p && io, "dcx_pipeline_step: null argument");
  return pipe_enqueue(p, io, as_stream(stream));
}

extern "C" dcx_status dcx_pipeline_capture(dcx_pipeline* p, const dcx_pipeline_io* io, void* stream) {
  DCX_REQUIRE(p && io, "dcx_pipeline_capture: null argument");
  DCX_REQUIRE(stream, "dcx_pipeline_capture: capture needs a non-default stream");
  hipStream_t s = as_stream(stream);
  if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
  if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
  DCX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  dcx_status st = pipe_enqueue(p, io, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (st) { if (g) (void)hipGraphDestroy(g); return st; }
  if (e != hipSuccess) return set_error(DCX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  p->graph = g;
  DCX_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
  return DCX_OK;
}

extern "C" dcx_status dcx_pipeline_replay(dcx_pipeline* p, void* stream) {
  DCX_REQUIRE(p && p->exec, "dcx_pipeline_replay: no captured graph");
  DCX_HIP(hipGraphLaunch(p->exec, as_stream(stream)));
  return DCX_OK;
}

extern "C" int dcx_pipeline_num_stages(void) { return kNumStages; }

extern "C" const char* dcx_pipeline_stage_name(int i) {
  return (i >= 0 && i < kNumStages) ? kStageNames[i] : nullptr;
}

extern "C" dcx_status dcx_pipeline_profile(dcx_pipeline* p, const dcx_pipeline_io* io, float* stage_ms, int n_stages,
                                           void* stream) {
  DCX_REQUIRE(p && io && stage_ms, "dcx_pipeline_profile: null argument");
  DCX_REQUIRE(n_stages == kNumStages, "dcx_pipeline_profile: n_stages must be %d", kNumStages);
  hipStream_t s = as_stream(stream);
  hipEvent_t ev[kNumStages + 1] = {};
  dcx_status st = DCX_OK;
  for (int i = 0; i <= kNumStages && st == DCX_OK; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) st = set_error(DCX_ERR_HIP, "dcx_pipeline_profile: hipEventCreate");
  if (st == DCX_OK) st = pipe_enqueue(p, io, s, ev);
  if (st == DCX_OK && hipEventSynchronize(ev[kNumStages]) != hipSuccess)
    st = set_error(DCX_ERR_HIP, "dcx_pipeline_profile: hipEventSynchronize");
  for (int i = 0; i < kNumStages && st == DCX_OK; ++i)
    if (hipEventElapsedTime(&stage_ms[i], ev[i], ev[i + 1]) != hipSuccess)
      st = set_error(DCX_ERR_HIP, "dcx_pipeline_profile: hipEventElapsedTime");
  for (int i = 0; i <= kNumStages; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  return st;
}

static void drop_graph(dcx_pipeline* p) {
  if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
  if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
}

extern "C" dcx_status dcx_pipeline_set_tile(dcx_pipeline* p, int tile_w) {
  DCX_REQUIRE(p, "dcx_pipeline_set_tile: null pipeline");
  int tws = -1;
  for (int k = kMinTileShift; k <= kMaxTileShift; ++k)
    if (tile_w == (1 << k)) tws = k;
  DCX_REQUIRE(tws >= 0, "dcx_pipeline_set_tile: tile width %d is not 16, 32 or 64", tile_w);
  if (tws == p->d.tws) return DCX_OK;
  // A step queued on any stream (the caller's, the side stream, a graph
  // replay) still reads the view table and runs kernels built for the old
  // shape: wait for all of it before the table or the graph changes.
  DCX_HIP(hipDeviceSynchronize());
  int mt = 0;
  const long long tot = layout_tiles(p->vw_host, tws, mt);
  DCX_HIP(hipMemcpy((void*)p->d.vw, p->vw_host.data(), sizeof(ViewD) * p->vw_host.size(), hipMemcpyHostToDevice));
  p->d.tws = tws;
  p->d.ths = 8 - tws;
  p->d.maxTiles = mt;
  p->d.totTiles = tot;
  drop_graph(p);  // kernel arguments changed: a captured step is stale
  return DCX_OK;
}

extern "C" int dcx_pipeline_tile_width(const dcx_pipeline* p) { return p ? 1 << p->d.tws : 0; }

extern "C" dcx_status dcx_pipeline_set_grid(dcx_pipeline* p, int cap) {
  DCX_REQUIRE(p, "dcx_pipeline_set_grid: null pipeline");
  DCX_REQUIRE(cap >= 0 && cap <= kMaxCellCap, "dcx_pipeline_set_grid: %d slots per cell is outside [0, %d]", cap,
              kMaxCellCap);
  Dev& d = p->d;
  if (cap == d.cell_cap) return DCX_OK;
  f4* slots = nullptr;
  if (cap > 0) {
    const hipError_t e = hipMalloc(&slots, sizeof(f4) * (size_t)cap * (size_t)d.totCells);
    if (e != hipSuccess) return set_error(DCX_ERR_OOM, "dcx_pipeline_set_grid: %s", hipGetErrorString(e));
  }
  // a queued step (any stream, replays included) may still use the old layout
  DCX_HIP(hipDeviceSynchronize());
  bool placed = false;
  for (void*& a : p->allocs)
    if (d.cell_slots && a == d.cell_slots) { (void)hipFree(a); a = slots; placed = true; }
  if (!placed && slots) p->allocs.push_back(slots);
  d.cell_slots = slots;
  d.cell_cap = cap;
  drop_graph(p);  // kernel arguments changed: a captured step is stale
  return DCX_OK;
}

extern "C" int dcx_pipeline_grid_cap(const dcx_pipeline* p) { return p ? p->d.cell_cap : -1; }

// Number of (triangle, tile) pairs of the last step (the bins' fill).
static dcx_status pairs_of_last_step(dcx_pipeline* p, hipStream_t s, unsigned long long* pairs) {
  unsigned int v = 0;
  DCX_HIP(hipMemcpyAsync(&v, p->d.tile_off + p->d.totTiles, sizeof(v), hipMemcpyDeviceToHost, s));
  DCX_HIP(hipStreamSynchronize(s));
  *pairs = v;
  return DCX_OK;
}

// Record storage for `cap` (triangle, tile) pairs.
static dcx_status grow_bins(dcx_pipeline* p, long long cap) {
  DCX_REQUIRE(cap < (1ll << 32), "dcx_pipeline_autotune: %lld records exceed the 32-bit record index", cap);
  Dev& d = p->d;
  unsigned int* bins = nullptr;
  float* rec = nullptr;
  unsigned char* vis = nullptr;
  hipError_t e = hipMalloc(&bins, 4 * (size_t)cap);
  if (e == hipSuccess) e = hipMalloc(&rec, sizeof(float) * kGrad * (size_t)cap);
  if (e == hipSuccess) e = hipMalloc(&vis, (size_t)cap);
  if (e != hipSuccess) {
    (void)hipFree(bins); (void)hipFree(rec); (void)hipFree(vis);
    return set_error(DCX_ERR_OOM, "dcx_pipeline_autotune: %s", hipGetErrorString(e));
  }
  DCX_HIP(hipDeviceSynchronize());  // the old buffers may still be read by queued work
  for (void*& a : p->allocs) {
    if (a == d.bins) { (void)hipFree(a); a = bins; }
    else if (a == d.rec) { (void)hipFree(a); a = rec; }
    else if (a == d.rec_vis) { (void)hipFree(a); a = vis; }
  }
  d.bins = bins; d.rec = rec; d.rec_vis = vis;
  d.capacity = (unsigned int)cap;
  drop_graph(p);
  return DCX_OK;
}

extern "C" dcx_status dcx_pipeline_autotune(dcx_pipeline* p, const dcx_pipeline_io* io, int reps, void* stream,
                                            int* tile_w_out, float* ms_out, long long* pairs_out) {
  DCX_REQUIRE(p && io && reps > 0, "dcx_pipeline_autotune: bad arguments");
  hipStream_t s = as_stream(stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DCX_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return set_error(DCX_ERR_HIP, "hipEventCreate"); }
  dcx_status st = DCX_OK;
  float best_ms = 0.f;
  int best = -1;
  unsigned long long most_pairs = 0;
  for (int tws = kMinTileShift; tws <= kMaxTileShift && st == DCX_OK; ++tws) {
    st = dcx_pipeline_set_tile(p, 1 << tws);
    // warm-up step; bins sized to this shape's pairs (plus a margin) so no
    // pair takes the overflow path
    if (st == DCX_OK) st = pipe_enqueue(p, io, s);
    unsigned long long pairs = 0;
    if (st == DCX_OK) st = pairs_of_last_step(p, s, &pairs);
    if (st == DCX_OK && pairs > p->d.capacity) {
      st = grow_bins(p, (long long)(pairs + pairs / 4 + 1024));
      if (st == DCX_OK) st = pipe_enqueue(p, io, s);
    }
    most_pairs = std::max(most_pairs, pairs);
    if (st == DCX_OK && hipEventRecord(e0, s) != hipSuccess) st = set_error(DCX_ERR_HIP, "hipEventRecord");
    for (int r = 0; r < reps && st == DCX_OK; ++r) st = pipe_enqueue(p, io, s);
    if (st == DCX_OK && hipEventRecord(e1, s) != hipSuccess) st = set_error(DCX_ERR_HIP, "hipEventRecord");
    float ms = 0.f;
    if (st == DCX_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      st = set_error(DCX_ERR_HIP, "dcx_pipeline_autotune: event timing failed");
    ms /= (float)reps;
    if (ms_out) ms_out[tws - kMinTileShift] = ms;
    if (st == DCX_OK && (best < 0 || ms < best_ms)) { best = tws; best_ms = ms; }
  }
  if (st == DCX_OK) st = dcx_pipeline_set_tile(p, 1 << best);
  // grid layout at that shape: the counting sort against fixed slots (32 or
  // 64 per cell, when the step's cells would send at most kGridOverflowMax
  // points of any view to the overflow list and the slots fit in
  // kGridSlotBytesMax); the fastest is kept
  constexpr long long kGridOverflowMax = 64;
  constexpr size_t kGridSlotBytesMax = (size_t)1 << 30;
  auto timed = [&](float* ms) -> dcx_status {
    dcx_status t = pipe_enqueue(p, io, s);
    if (t == DCX_OK && hipEventRecord(e0, s) != hipSuccess) t = set_error(DCX_ERR_HIP, "hipEventRecord");
    for (int r = 0; r < reps && t == DCX_OK; ++r) t = pipe_enqueue(p, io, s);
    if (t == DCX_OK && hipEventRecord(e1, s) != hipSuccess) t = set_error(DCX_ERR_HIP, "hipEventRecord");
    if (t == DCX_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(ms, e0, e1) != hipSuccess))
      t = set_error(DCX_ERR_HIP, "dcx_pipeline_autotune: event timing failed");
    *ms /= (float)reps;
    return t;
  };
  int best_cap = 0;
  float best_grid_ms = 0.f;
  if (st == DCX_OK) st = dcx_pipeline_set_grid(p, 0);
  if (st == DCX_OK) st = timed(&best_grid_ms);
  std::vector<unsigned int> cnt;
  if (st == DCX_OK) {
    cnt.resize((size_t)p->d.totCells);
    if (hipMemcpy(cnt.data(), p->d.cell_cnt, 4 * cnt.size(), hipMemcpyDeviceToHost) != hipSuccess)
      st = set_error(DCX_ERR_HIP, "dcx_pipeline_autotune: cell counts");
  }
  for (int cap = 32; cap <= 64 && st == DCX_OK; cap *= 2) {
    long long worst = 0;
    for (const ViewD& v : p->vw_host) {
      long long over = 0;
      for (long long i = v.cell0; i < v.cell0 + (long long)v.gx * v.gy; ++i)
        over += std::max(0ll, (long long)cnt[(size_t)i] - cap);
      worst = std::max(worst, over);
    }
    if (worst > kGridOverflowMax || sizeof(f4) * (size_t)cap * (size_t)p->d.totCells > kGridSlotBytesMax) continue;
    float ms = 0.f;
    st = dcx_pipeline_set_grid(p, cap);
    if (st == DCX_OK) st = timed(&ms);
    if (st == DCX_OK && ms < best_grid_ms) { best_grid_ms = ms; best_cap = cap; }
  }
  if (st == DCX_OK) st = dcx_pipeline_set_grid(p, best_cap);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) return st;
  if (tile_w_out) *tile_w_out = 1 << best;
  if (pairs_out) *pairs_out = (long long)most_pairs;
  return DCX_OK;
}
