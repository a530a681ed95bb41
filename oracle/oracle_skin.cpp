// ORACLE — test infrastructure only.  CPU restatement of the reference's
// skinning algorithms, used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker; never part of the product path.
//
// Templated on the scalar: T=double is the FP64 oracle, T=float mirrors the
// reference's single-precision arithmetic.
//
//   dqs_fwd   DQSkinningGPUOp.cu:56-120, DQHelper.h:22-86
//   dqs_bwd   exact adjoint of dqs_fwd (no reference counterpart for the
//             joint-DQ gradient); the REF_COMPAT weight gradient restates
//             DQSkinningGPUOpGrad.cu:12-141 with DQHelper.h:93-235
//   dof_grad  DQSkinningGPUOpGrad.cu:145-206 (bounded at max_entries)
//   lbs_fwd   LBSkinningGPUOp.cu:49-96 (rest mesh indexed per vertex)
//   lbs_bwd   exact; REF_COMPAT restates LBSkinningGPUOpGrad.cu:11-88
//
// Pinning (tests/test_oracle_skin.py): closed forms (identity DQ, single
// joint rigid transform, equal joints), antipodal-sign invariance, unit
// normals, and FP64 central finite differences for every gradient.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_common.hpp"

namespace orc {

template <class T> struct Q { T x, y, z, w; };

template <class T> static Q<T> qld(const T* p) { return {p[0], p[1], p[2], p[3]}; }
template <class T> static T qdot(const Q<T>& a, const Q<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// Rotation matrix of quaternion (x,y,z,w), row-major.
template <class T> static void qmat(const Q<T>& q, T R[9]) {
  R[0] = T(1) - T(2) * q.y * q.y - T(2) * q.z * q.z;
  R[1] = T(2) * q.x * q.y - T(2) * q.z * q.w;
  R[2] = T(2) * q.x * q.z + T(2) * q.y * q.w;
  R[3] = T(2) * q.x * q.y + T(2) * q.z * q.w;
  R[4] = T(1) - T(2) * q.x * q.x - T(2) * q.z * q.z;
  R[5] = T(2) * q.y * q.z - T(2) * q.x * q.w;
  R[6] = T(2) * q.x * q.z - T(2) * q.y * q.w;
  R[7] = T(2) * q.y * q.z + T(2) * q.x * q.w;
  R[8] = T(1) - T(2) * q.x * q.x - T(2) * q.y * q.y;
}

// Derivatives dR/dq_c (c = x,y,z,w) of qmat, row-major.
template <class T> static void qmat_d(const Q<T>& q, T dR[4][9]) {
  const T x = q.x, y = q.y, z = q.z, w = q.w;
  T a[4][9] = {
      {0, 2 * y, 2 * z, 2 * y, -4 * x, -2 * w, 2 * z, 2 * w, -4 * x},
      {-4 * y, 2 * x, 2 * w, 2 * x, 0, 2 * z, -2 * w, 2 * z, -4 * y},
      {-4 * z, -2 * w, 2 * x, 2 * w, -4 * z, 2 * y, 2 * x, 2 * y, 0},
      {0, -2 * z, 2 * y, 2 * z, 0, -2 * x, -2 * y, 2 * x, 0}};
  std::memcpy(dR, a, sizeof(a));
}

// t = 2 vec(d * conj(r)).
template <class T> static void qtrans(const Q<T>& r, const Q<T>& d, T t[3]) {
  t[0] = T(2) * (-d.w * r.x + d.x * r.w - d.y * r.z + d.z * r.y);
  t[1] = T(2) * (-d.w * r.y + d.x * r.z + d.y * r.w - d.z * r.x);
  t[2] = T(2) * (-d.w * r.z - d.x * r.y + d.y * r.x + d.z * r.w);
}

template <class T> struct Blend {
  Q<T> br, bd, r, d;
  T sgn[16];
  T L;
  bool small;
};

template <class T>
static Blend<T> blend(const T* tab, const int32_t* idx, const T* w, int K) {
  Blend<T> b{};
  Q<T> q0 = qld(tab + 8 * idx[0]);
  b.br = {0, 0, 0, 0};
  b.bd = {0, 0, 0, 0};
  for (int k = 0; k < K; ++k) {
    Q<T> qr = qld(tab + 8 * idx[k]), qd = qld(tab + 8 * idx[k] + 4);
    T s = (k != 0 && qdot(q0, qr) < T(0)) ? T(-1) : T(1);
    b.sgn[k] = s;
    T ws = w[k] * s;
    b.br = {b.br.x + qr.x * ws, b.br.y + qr.y * ws, b.br.z + qr.z * ws, b.br.w + qr.w * ws};
    b.bd = {b.bd.x + qd.x * ws, b.bd.y + qd.y * ws, b.bd.z + qd.z * ws, b.bd.w + qd.w * ws};
  }
  b.L = std::sqrt(qdot(b.br, b.br));
  b.small = !(b.L > T(0.000001));
  T sc = b.small ? T(1) : T(1) / b.L;
  b.r = {b.br.x * sc, b.br.y * sc, b.br.z * sc, b.br.w * sc};
  b.d = {b.bd.x * sc, b.bd.y * sc, b.bd.z * sc, b.bd.w * sc};
  return b;
}

template <class T>
void dqs_fwd(const T* dq, const T* w, const int32_t* jidx, const T* rp, const T* rn, int B, int N,
             int K, int J, T* pos, T* nrm) {
  for (int b = 0; b < B; ++b)
    for (int v = 0; v < N; ++v) {
      Blend<T> bl = blend(dq + (size_t)b * J * 8, jidx + (size_t)v * K, w + ((size_t)b * N + v) * K, K);
      T R[9], t[3];
      qmat(bl.r, R);
      qtrans(bl.r, bl.d, t);
      const T* p = rp + 3 * v;
      T* o = pos + ((size_t)b * N + v) * 3;
      for (int i = 0; i < 3; ++i) o[i] = R[3 * i] * p[0] + R[3 * i + 1] * p[1] + R[3 * i + 2] * p[2] + t[i];
      if (nrm) {
        const T* n = rn + 3 * v;
        T m[3];
        for (int i = 0; i < 3; ++i) m[i] = R[3 * i] * n[0] + R[3 * i + 1] * n[1] + R[3 * i + 2] * n[2];
        T l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        T* no = nrm + ((size_t)b * N + v) * 3;
        for (int i = 0; i < 3; ++i) no[i] = m[i] / l;
      }
    }
}

// Reference skin-weight gradient restated from DQSkinningGPUOpGrad.cu:64-138.
template <class T>
static T dqs_ref_wgrad(const Blend<T>& bl, const Q<T>& qr0, const Q<T>& qd0, T sgn, const T* p, const T* g) {
  const T L = bl.small ? T(1) : bl.L;
  Q<T> qr = {qr0.x * sgn, qr0.y * sgn, qr0.z * sgn, qr0.w * sgn};
  Q<T> qd = {qd0.x * sgn, qd0.y * sgn, qd0.z * sgn, qd0.w * sgn};
  T dS = bl.small ? T(0) : qdot(bl.br, qr) / L;
  T inv = T(1) / (L * L);
  T dR[4] = {(qr.x * L - bl.br.x * dS) * inv, (qr.y * L - bl.br.y * dS) * inv,
             (qr.z * L - bl.br.z * dS) * inv, (qr.w * L - bl.br.w * dS) * inv};
  T dT[4] = {(qd.x * L - bl.bd.x * dS) * inv, (qd.y * L - bl.bd.y * dS) * inv,
             (qd.z * L - bl.bd.z * dS) * inv, (qd.w * L - bl.bd.w * dS) * inv};
  T dRq[4][9];
  qmat_d(bl.r, dRq);
  const Q<T>& d = bl.d;
  const Q<T>& r = bl.r;
  T TR[3][4] = {{-2 * d.w, 2 * d.z, -2 * d.y, 2 * d.x}, {-2 * d.z, -2 * d.w, 2 * d.x, 2 * d.y}, {2 * d.y, -2 * d.x, -2 * d.w, 2 * d.z}};
  T TT[3][4] = {{2 * r.w, -2 * r.z, 2 * r.y, -2 * r.x}, {2 * r.z, 2 * r.w, -2 * r.x, -2 * r.y}, {-2 * r.y, 2 * r.x, 2 * r.w, -2 * r.z}};
  T out = 0;
  for (int i = 0; i < 3; ++i) {
    T gi = 0;
    for (int c = 0; c < 4; ++c) {
      // The reference adds p.z where it means to multiply (DQHelper.h:121-174).
      T Jrot = dRq[c][3 * i] * p[0] + dRq[c][3 * i + 1] * p[1] + dRq[c][3 * i + 2] + p[2];
      gi += (Jrot + TR[i][c]) * dR[c] + TT[i][c] * dT[c];
    }
    out += gi * g[i];
  }
  return out;
}

template <class T>
void dqs_bwd(const T* gpos, const T* gnrm, const T* dq, const T* w, const int32_t* jidx, const T* rp,
             const T* rn, int B, int N, int K, int J, T* gdq, T* gw, T* grest, int flags) {
  if (gdq) std::memset(gdq, 0, sizeof(T) * (size_t)B * J * 8);
  for (int b = 0; b < B; ++b)
    for (int v = 0; v < N; ++v) {
      const T* tab = dq + (size_t)b * J * 8;
      const int32_t* idx = jidx + (size_t)v * K;
      const T* wv = w + ((size_t)b * N + v) * K;
      Blend<T> bl = blend(tab, idx, wv, K);
      T R[9], dR[4][9];
      qmat(bl.r, R);
      qmat_d(bl.r, dR);
      const T* p = rp + 3 * v;
      const T* g = gpos + ((size_t)b * N + v) * 3;
      T G[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) G[3 * i + j] = g[i] * p[j];
      if (gnrm) {
        const T* n = rn + 3 * v;
        const T* gn = gnrm + ((size_t)b * N + v) * 3;
        T m[3];
        for (int i = 0; i < 3; ++i) m[i] = R[3 * i] * n[0] + R[3 * i + 1] * n[1] + R[3 * i + 2] * n[2];
        T l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        T nn[3] = {m[0] / l, m[1] / l, m[2] / l};
        T pr = nn[0] * gn[0] + nn[1] * gn[1] + nn[2] * gn[2];
        for (int i = 0; i < 3; ++i) {
          T gm = (gn[i] - nn[i] * pr) / l;
          for (int j = 0; j < 3; ++j) G[3 * i + j] += gm * n[j];
        }
      }
      // dL/dr from R(r)p and from t(r,d); dL/dd from t.
      T gr[4], gd[4];
      for (int c = 0; c < 4; ++c) {
        T s = 0;
        for (int e = 0; e < 9; ++e) s += G[e] * dR[c][e];
        gr[c] = s;
      }
      const Q<T>& r = bl.r;
      const Q<T>& d = bl.d;
      // t_i = 2 * sum of bilinear terms; derivatives from qtrans.
      T dt_dr[3][4] = {{-2 * d.w, 2 * d.z, -2 * d.y, 2 * d.x}, {-2 * d.z, -2 * d.w, 2 * d.x, 2 * d.y}, {2 * d.y, -2 * d.x, -2 * d.w, 2 * d.z}};
      T dt_dd[3][4] = {{2 * r.w, -2 * r.z, 2 * r.y, -2 * r.x}, {2 * r.z, 2 * r.w, -2 * r.x, -2 * r.y}, {-2 * r.y, 2 * r.x, 2 * r.w, -2 * r.z}};
      for (int c = 0; c < 4; ++c) {
        gd[c] = 0;
        for (int i = 0; i < 3; ++i) {
          gr[c] += g[i] * dt_dr[i][c];
          gd[c] += g[i] * dt_dd[i][c];
        }
      }
      T gbr[4], gbd[4];
      const T rv[4] = {r.x, r.y, r.z, r.w}, dv[4] = {d.x, d.y, d.z, d.w};
      if (bl.small) {
        for (int c = 0; c < 4; ++c) { gbr[c] = gr[c]; gbd[c] = gd[c]; }
      } else {
        T proj = 0;
        for (int c = 0; c < 4; ++c) proj += gr[c] * rv[c] + gd[c] * dv[c];
        for (int c = 0; c < 4; ++c) {
          gbr[c] = (gr[c] - rv[c] * proj) / bl.L;
          gbd[c] = gd[c] / bl.L;
        }
      }
      if (grest) {
        T* o = grest + ((size_t)b * N + v) * 3;
        for (int j = 0; j < 3; ++j) o[j] = R[j] * g[0] + R[3 + j] * g[1] + R[6 + j] * g[2];
      }
      for (int k = 0; k < K; ++k) {
        const T* q = tab + 8 * idx[k];
        if (gw) {
          T* o = gw + ((size_t)b * N + v) * K + k;
          if (flags & 1) {
            *o = dqs_ref_wgrad(bl, qld(q), qld(q + 4), bl.sgn[k], p, g);
          } else {
            T s = 0;
            for (int c = 0; c < 4; ++c) s += gbr[c] * q[c] + gbd[c] * q[4 + c];
            *o = bl.sgn[k] * s;
          }
        }
        if (gdq) {
          T ws = wv[k] * bl.sgn[k];
          T* o = gdq + ((size_t)b * J + idx[k]) * 8;
          for (int c = 0; c < 4; ++c) { o[c] += ws * gbr[c]; o[4 + c] += ws * gbd[c]; }
        }
      }
    }
}

template <class T>
void dof_grad(const T* gpos, const T* sp, const T* jp, const T* ja, const T* infl, int B, int N, int Jall,
              int D, int maxE, T* out) {
  for (int b = 0; b < B; ++b)
    for (int dof = 0; dof < D; ++dof) {
      T acc = 0;
      for (int e = 0; e < maxE; ++e) {
        const T* row = infl + ((size_t)dof * maxE + e) * 4;
        if (row[0] == T(-1)) break;
        int vtx = (int)row[0], type = (int)row[1], j = (int)row[2];
        T wt = row[3];
        const T* a = ja + ((size_t)b * Jall + j) * 3;
        T rd[3];
        if (type == 0) {
          const T* c = jp + ((size_t)b * Jall + j) * 3;
          const T* vp = sp + ((size_t)b * N + vtx) * 3;
          T d[3] = {vp[0] - c[0], vp[1] - c[1], vp[2] - c[2]};
          rd[0] = a[1] * d[2] - a[2] * d[1];
          rd[1] = a[2] * d[0] - a[0] * d[2];
          rd[2] = a[0] * d[1] - a[1] * d[0];
        } else {
          rd[0] = a[0]; rd[1] = a[1]; rd[2] = a[2];
        }
        const T* g = gpos + ((size_t)b * N + vtx) * 3;
        acc += wt * (rd[0] * g[0] + rd[1] * g[1] + rd[2] * g[2]);
      }
      out[(size_t)b * D + dof] = acc;
    }
}

template <class T>
void lbs_fwd(const T* tr, const T* w, const int32_t* jidx, const T* rp, const T* rn, const T* disp, int B,
             int N, int K, int J, T* pos, T* nrm) {
  for (int b = 0; b < B; ++b)
    for (int v = 0; v < N; ++v) {
      T R[9] = {0}, t[3] = {0};
      for (int k = 0; k < K; ++k) {
        const T* M = tr + ((size_t)b * J + jidx[(size_t)v * K + k]) * 12;
        T wk = w[((size_t)b * N + v) * K + k];
        for (int i = 0; i < 9; ++i) R[i] += M[i] * wk;
        for (int i = 0; i < 3; ++i) t[i] += M[9 + i] * wk;
      }
      const T* dp = disp + ((size_t)b * N + v) * 3;
      T p[3] = {rp[3 * v] + dp[0], rp[3 * v + 1] + dp[1], rp[3 * v + 2] + dp[2]};
      T* o = pos + ((size_t)b * N + v) * 3;
      for (int i = 0; i < 3; ++i) o[i] = R[3 * i] * p[0] + R[3 * i + 1] * p[1] + R[3 * i + 2] * p[2] + t[i];
      if (nrm) {
        const T* n = rn + 3 * v;
        T* no = nrm + ((size_t)b * N + v) * 3;
        for (int i = 0; i < 3; ++i) no[i] = R[3 * i] * n[0] + R[3 * i + 1] * n[1] + R[3 * i + 2] * n[2];
      }
    }
}

template <class T>
void lbs_bwd(const T* gpos, const T* tr, const T* w, const int32_t* jidx, const T* rp, const T* disp, int B,
             int N, int K, int J, T* gw, T* gdisp, T* gtr, int flags) {
  const bool ref = flags & 1;
  if (gtr) std::memset(gtr, 0, sizeof(T) * (size_t)B * J * 12);
  for (int b = 0; b < B; ++b)
    for (int v = 0; v < N; ++v) {
      const T* g = gpos + ((size_t)b * N + v) * 3;
      const T* dp = disp + ((size_t)b * N + v) * 3;
      T base[3] = {rp[3 * v], rp[3 * v + 1], rp[3 * v + 2]};
      T p[3] = {base[0] + dp[0], base[1] + dp[1], base[2] + dp[2]};
      T Rb[9] = {0};
      for (int k = 0; k < K; ++k) {
        int j = jidx[(size_t)v * K + k];
        T wk = w[((size_t)b * N + v) * K + k];
        if (gw) {
          const T* M = tr + ((size_t)b * J + (ref ? k : j)) * 12;
          const T* pp = ref ? base : p;
          T s = 0;
          for (int i = 0; i < 3; ++i) s += (M[3 * i] * pp[0] + M[3 * i + 1] * pp[1] + M[3 * i + 2] * pp[2] + M[9 + i]) * g[i];
          gw[((size_t)b * N + v) * K + k] = s;
        }
        if (gdisp) {
          if (ref) {
            T wr = w[(size_t)v * K + k];
            if (!(wr < T(1e-4))) {
              const T* M = tr + ((size_t)b * J + k) * 12;
              for (int i = 0; i < 9; ++i) Rb[i] += M[i] * wr;
            }
          } else {
            const T* M = tr + ((size_t)b * J + j) * 12;
            for (int i = 0; i < 9; ++i) Rb[i] += M[i] * wk;
          }
        }
        if (gtr) {
          T* o = gtr + ((size_t)b * J + j) * 12;
          for (int i = 0; i < 3; ++i) {
            for (int c = 0; c < 3; ++c) o[3 * i + c] += wk * g[i] * p[c];
            o[9 + i] += wk * g[i];
          }
        }
      }
      if (gdisp) {
        T* o = gdisp + ((size_t)b * N + v) * 3;
        for (int c = 0; c < 3; ++c) o[c] = Rb[c] * g[0] + Rb[3 + c] * g[1] + Rb[6 + c] * g[2];
      }
    }
}

}  // namespace orc

#define ORC_SKIN_EXPORT(T, SUF)                                                                              \
  extern "C" void orc_dqs_fwd_##SUF(const T* dq, const T* w, const int32_t* ji, const T* rp, const T* rn,     \
                                    int B, int N, int K, int J, T* pos, T* nrm) {                            \
    orc::dqs_fwd<T>(dq, w, ji, rp, rn, B, N, K, J, pos, nrm);                                                 \
  }                                                                                                           \
  extern "C" void orc_dqs_bwd_##SUF(const T* gp, const T* gn, const T* dq, const T* w, const int32_t* ji,     \
                                    const T* rp, const T* rn, int B, int N, int K, int J, T* gdq, T* gw,      \
                                    T* gr, int flags) {                                                       \
    orc::dqs_bwd<T>(gp, gn, dq, w, ji, rp, rn, B, N, K, J, gdq, gw, gr, flags);                               \
  }                                                                                                           \
  extern "C" void orc_dof_grad_##SUF(const T* gp, const T* sp, const T* jp, const T* ja, const T* infl,       \
                                     int B, int N, int Jall, int D, int maxE, T* out) {                       \
    orc::dof_grad<T>(gp, sp, jp, ja, infl, B, N, Jall, D, maxE, out);                                         \
  }                                                                                                           \
  extern "C" void orc_lbs_fwd_##SUF(const T* tr, const T* w, const int32_t* ji, const T* rp, const T* rn,     \
                                    const T* disp, int B, int N, int K, int J, T* pos, T* nrm) {             \
    orc::lbs_fwd<T>(tr, w, ji, rp, rn, disp, B, N, K, J, pos, nrm);                                           \
  }                                                                                                           \
  extern "C" void orc_lbs_bwd_##SUF(const T* gp, const T* tr, const T* w, const int32_t* ji, const T* rp,     \
                                    const T* disp, int B, int N, int K, int J, T* gw, T* gd, T* gt,          \
                                    int flags) {                                                              \
    orc::lbs_bwd<T>(gp, tr, w, ji, rp, disp, B, N, K, J, gw, gd, gt, flags);                                  \
  }

ORC_SKIN_EXPORT(double, f64)
ORC_SKIN_EXPORT(float, f32)
