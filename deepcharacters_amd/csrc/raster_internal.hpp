// Host/device structures shared by the rasterizer translation units.
#pragma once

#include "raster_common.hpp"

struct dcx_topo {
  int N = 0, F = 0;
  int* faces = nullptr;       // [F,3]
  float* texcoords = nullptr; // [F,6] or null
  int* vf_off = nullptr;      // [N+1] vertex -> face CSR offsets
  int* vf_list = nullptr;     // [3F] face ids, increasing per vertex
  float4* texel_tab = nullptr; // [tab_h, tab_w] (face, a, b, c) of the UV atlas, built on request
  int tab_h = 0, tab_w = 0;
};

namespace dcx {

constexpr int kTileW = 32;
constexpr int kTileH = 8;
constexpr int kTilePix = kTileW * kTileH;  // one lane per pixel in the resolve

struct RenderArgs {
  int N, F, C, H, W;
  int albedo, shaded, TH, TW;
  int view0, b0;              // first view / batch element held by proj / vnrm scratch
  const int* faces;
  const float* tc;
  const float* vpos;          // [B,N,3]
  const float* vcol;          // [B,N,3]
  const float* tex;           // [B,TH,TW,3]
  const float* sh;            // [B,C,27]
  const float* ext;           // [B,C*12]
  const float* intr;          // [B,C*9]
  f4* proj;                   // [views,N] projected vertices
  float* vnrm;                // [B,N,3] unnormalised vertex normals
  int* face_out;
  float* bary_out;
  float* render_out;
  int rule;                   // DCX_RULE_*
  const float* refcam;        // reference rule: [views held,32] camera inverses (invE, invP)
};

struct BinArgs {
  int views, tiles_x, tiles_per_view;
  unsigned int capacity;
  TriRec* tri;
  unsigned int* tile_cnt;
  unsigned int* tile_off;     // [tiles+1]
  unsigned int* tile_fill;
  unsigned int* bins;
};

struct WsLayout {
  size_t proj, fnrm, vnrm, keys, refcam, tri, tile_cnt, tile_off, tile_fill, aux, bins;
  long long tiles;
};

size_t render_ws_fixed(const dcx_topo* t, const dcx_raster_cfg* c, int B, WsLayout* L);
// Area-weighted, unnormalised vertex normals of B meshes (vnrm [B,N,3]; face
// normal scratch fnrm [B,F,3]); also broadcast into vn_out [B,C,N,3] if given.
dcx_status vertex_normals(const dcx_topo* t, int B, int C, const float* vpos, float* fnrm, float* vnrm,
                          float* vn_out, hipStream_t s);
dcx_status validate_cfg(const dcx_topo* t, const dcx_raster_cfg* c, int B);

}  // namespace dcx
