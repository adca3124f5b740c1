// stream_vec.h -- phase bodies of the streaming engine for instances of one
// or of four and more right-hand sides.  Each lane owns V adjacent columns of
// a row (up to 16 bytes of the inner precision), so a row's neighbour values
// arrive in K / V loads of up to 16 bytes instead of K scalar ones, and the
// column index and matrix entry of a nonzero are loaded by K / V lanes
// instead of K.
//
// Per element the arithmetic is the persistent kernel's (phase_* in
// step_kernel.h): the same expressions in the same precision and the same
// ascending summation over each row.  Only the order in which the per-lane
// partial dot products are combined differs, and BiCGSTAB's phases are cut
// differently (vphase_bi_p below).  The matrix and preconditioner
// are shared by all columns (one value column), which is the case for every
// instance with one or with more than two right-hand sides.
#pragma once

#include "step_kernel.h"

namespace femhip {

namespace {

template <typename T, int V>
struct VecOf {
  typedef T type __attribute__((ext_vector_type(V)));
};
template <typename T, int V>
using vec = typename VecOf<T, V>::type;

// Instances with one column or with four or more use this layout (their
// matrix is shared by all columns); two columns (Gray-Scott's u, v with their
// own diffusivities) keep the one-column-per-lane bodies.
template <int K>
constexpr bool vec_layout() {
  return K == 1 || K >= 4;
}

// Columns per lane: 16 bytes of the inner precision, at most K.
template <int K, typename TI>
constexpr int vec_cols() {
  return (int)(16 / sizeof(TI)) < K ? (int)(16 / sizeof(TI)) : K;
}

template <int V, typename T>
__device__ inline vec<T, V> vld(const T* p) {
  return *(const vec<T, V> GA*)p;
}
template <int V, typename T>
__device__ inline void vst(T* p, const vec<T, V>& v) {
  *(vec<T, V> GA*)p = v;
}
template <int V, typename T>
__device__ inline vec<T, V> splat(T x) {
  vec<T, V> r;
#pragma unroll
  for (int v = 0; v < V; v++) r[v] = x;
  return r;
}
template <int V, typename X>
__device__ inline vec<double, V> to_d(const X& x) {
  return __builtin_convertvector(x, vec<double, V>);
}
template <typename TI, int V, typename X>
__device__ inline vec<TI, V> to_i(const X& x) {
  return __builtin_convertvector(x, vec<TI, V>);
}
// Per-column scalars of the solve (alpha, beta, omega) for columns j0..j0+V-1.
template <typename TI, int V>
__device__ inline vec<TI, V> colv(const double* s, int j0) {
  vec<TI, V> r;
#pragma unroll
  for (int v = 0; v < V; v++) r[v] = to_ti<TI>(s[j0 + v]);
  return r;
}

template <typename T>
struct P2 {
  T a, b;
  __device__ P2& operator+=(const P2& o) {
    a += o.a;
    b += o.b;
    return *this;
  }
};
template <typename T>
struct P3 {
  T a, b, c;
  __device__ P3& operator+=(const P3& o) {
    a += o.a;
    b += o.b;
    c += o.c;
    return *this;
  }
};

// row_gather with U loads in flight: a vector lane already moves V columns
// per load, so fewer unrolled entries keep the same bytes in flight.
template <int U, typename A, typename F>
__device__ inline A vgather(int lo, int hi, A zero, F f) {
  A acc = zero;
  for (int s0 = lo; s0 < hi; s0 += U) {
    A t[U];
#pragma unroll
    for (int u = 0; u < U; u++) t[u] = f(min(s0 + u, hi - 1));
#pragma unroll
    for (int u = 0; u < U; u++)
      if (s0 + u < hi) acc += t[u];
  }
  return acc;
}

template <int V>
constexpr int vunroll() {
  return V >= 4 ? 2 : V == 2 ? 4 : 8;
}
// BiCGSTAB's gathers (one and two vectors per neighbour): twice the entries
// in flight at four columns per lane measured faster on 7-point rows.
template <int V>
constexpr int vunroll_bi() {
  return V >= 4 ? 4 : V == 2 ? 4 : 8;
}

// Lane layout: G = K / V lanes per row; lane columns j0 .. j0 + V - 1 with
// j0 = (tid % G) V; rows tid / G + m * (kBlock / G).
template <int K, int V>
struct VLay {
  static constexpr int G = K / V;
  static constexpr int RPP = kBlock / G;
  static_assert(K % V == 0 && (G & (G - 1)) == 0 && G <= kWave, "lane layout");
};

// A row block's pattern staged in LDS: row pointers relative to the block's
// first nonzero, column indices and the column-scaled inner matrix entries
// S_ij D^-1_jj (classic CG's phase A: S_ij).  The Krylov gathers then read
// only the vectors from global memory: one dependent memory level instead of
// three (row pointer -> column index -> vector).
#define LA __attribute__((address_space(3)))
template <typename TI>
struct Stage {
  const int LA* rp;  // [rows + 1]
  const int LA* ci;
  const TI LA* w;
  int row_lo;
};

// The right-hand side's steady cases after a launch's first step, where the
// previous update formed what the rows gather (vphase_rhs's nm and pre): 1
// Newmark (K and S rows, gathering hvec and the guess in x), 2 heat or
// advection-diffusion with a predictor (M and S rows, gathering u and the
// guess in x), 0 otherwise.
__device__ inline int rhs_case(const Inst& I, int step) {
  if (step <= 0) return 0;
  if (I.eq == EQ_NEWMARK) return I.hvec ? 1 : 0;
  return ((I.eq == EQ_HEAT || I.eq == EQ_ADVDIFF) && I.warm && I.extrap > 0) ? 2 : 0;
}

// A row block's pattern and its two FP64 value rows staged in LDS (the
// right-hand side's steady cases): local row pointers, columns, the K or M
// entries and the S entries.
struct RhsStage {
  const int LA* rp;
  const int LA* ci;
  const double LA* a;
  const double LA* s;
};

// vphase_rhs for rhs_case(I, step) = cs != 0 from the staged rows: the same
// per-row arithmetic, with the matrix rows read from LDS (one coalesced copy
// per block instead of one row per lane from memory).
template <int K, int V, typename TI>
__device__ void vphase_rhs_staged(const Inst& I, int row_lo, int row_hi, int cs, const RhsStage& g,
                                  vec<double, V> (&acc)[3]) {
  using DV = vec<double, V>;
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  constexpr int U = vunroll<V>();
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* r = (TI*)I.r;
  TI* xi = (TI*)I.xi;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const bool warm = I.warm != 0;
  const bool pre = cs == 2 || warm;  // x holds the guess already: not stored
  const double scale = cs == 1 ? -I.p1 * I.p1 : 1.0 / I.dt;
  const double* gv = cs == 1 ? I.hvec : I.u;
  const bool gs = cs == 2 || warm;  // S x is gathered
  const DV zd = splat<V>(0.0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    const P2<DV> q = vgather<U>(g.rp[row - row_lo], g.rp[row - row_lo + 1], P2<DV>{zd, zd}, [&](int e) {
      const size_t oc = (size_t)g.ci[e] * K + j0;
      return P2<DV>{g.a[e] * vld<V>(gv + oc), gs ? g.s[e] * vld<V>(I.x + oc) : zd};
    });
    const DV b = cs == 1 ? scale * q.a : q.a * scale;
    const DV r64 = b - q.b;
    vst<V>(I.b + o, b);
    if (!pre) vst<V>(I.x + o, zd);
    const TV rr = to_i<TI, V>(r64);
    vst<V>(r + o, rr);
    vst<V>(xi + o, splat<V>((TI)0));
    acc[0] += b * b;
    acc[1] += r64 * r64;
    const DV rd = to_d<V>(rr);
    acc[2] += (I.krylov == 1) ? rd * rd : rd * to_d<V>(dinv[row] * rr);
  }
}

template <int K, int V, typename TI>
__device__ void vphase_rhs(const Inst& I, int row_lo, int row_hi, int step, vec<double, V> (&acc)[3]) {
  using DV = vec<double, V>;
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  constexpr int U = vunroll<V>();
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* r = (TI*)I.r;
  TI* xi = (TI*)I.xi;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const double inv_dt = 1.0 / I.dt;
  const bool warm = I.warm != 0;
  // predictor order this step where the guess is gathered from the history
  // levels (the first step of a launch): quartic at most
  const int lv = warm ? min(min(I.extrap, kPersistentExtrap), I.hist + step) : 0;
  const bool ex = lv > 0;
  // after the first step of a launch the update wrote this step's guess into
  // x (stream_update_first, stream_update_newmark, stream.hip) and, for
  // Newmark, the displacement predictor u + dt v + (1/2 - beta) dt^2 a into
  // hvec
  const bool nm = I.eq == EQ_NEWMARK && step > 0 && I.hvec;
  const bool pre = (ex && step > 0 && (I.eq == EQ_HEAT || I.eq == EQ_ADVDIFF)) || (nm && warm);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  const double GA* M64 = gp(I.M64);
  const double GA* K64 = gp(I.K64);
  const double GA* S64 = gp(I.S64);
  const DV zd = splat<V>(0.0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const int lo = rp[row], hi = rp[row + 1];
    const size_t o = (size_t)row * K + j0;
    DV b = zd, sx = zd, x0 = zd;
    switch (I.eq) {
      case EQ_HEAT:
      case EQ_ADVDIFF:
      case EQ_GS: {  // b = (M/dt) u                                  CPUSolver.cpp:78,96
        if (pre) {  // the previous update formed the guess in x: gather it alone
          const P2<DV> g = vgather<U>(lo, hi, P2<DV>{zd, zd}, [&](int s) {
            const size_t oc = (size_t)ci[s] * K + j0;
            return P2<DV>{M64[s] * vld<V>(I.u + oc), S64[s] * vld<V>(I.x + oc)};
          });
          b = g.a * inv_dt;
          sx = g.b;
          break;
        }
        const P2<DV> g = vgather<U>(lo, hi, P2<DV>{zd, zd}, [&](int s) {
          const size_t oc = (size_t)ci[s] * K + j0;
          const DV uc = vld<V>(I.u + oc);
          const DV xc = lv == 4   ? 5.0 * (uc - vld<V>(I.h3 + oc)) + 10.0 * (vld<V>(I.v + oc) - vld<V>(I.a + oc)) +
                                        vld<V>(I.h4 + oc)
                        : lv == 3 ? 4.0 * (uc + vld<V>(I.v + oc)) - 6.0 * vld<V>(I.a + oc) - vld<V>(I.h3 + oc)
                        : lv == 2 ? 3.0 * (uc - vld<V>(I.a + oc)) + vld<V>(I.v + oc)
                        : ex      ? 2.0 * uc - vld<V>(I.a + oc)
                                  : uc;
          return P2<DV>{M64[s] * uc, warm ? S64[s] * xc : zd};
        });
        b = g.a * inv_dt;
        if (I.eq == EQ_GS) {  // CPUSolver.cpp:134,137 (two columns; kept for completeness)
          const double uu = gp(I.u)[(size_t)row * K], vv = gp(I.u)[(size_t)row * K + 1];
          const double uvv = uu * (vv * vv);
          const double w = I.weighting ? gp(I.lumped)[row] : 1.0;
#pragma unroll
          for (int v = 0; v < V; v++)
            b[v] += w * ((j0 + v == 0) ? (-uvv + I.p3 * (1.0 - uu)) : (uvv - (I.p3 + I.p4) * vv));
        }
        sx = g.b;
        const DV uo = vld<V>(I.u + o);
        x0 = !warm     ? zd
             : lv == 4 ? 5.0 * (uo - vld<V>(I.h3 + o)) + 10.0 * (vld<V>(I.v + o) - vld<V>(I.a + o)) + vld<V>(I.h4 + o)
             : lv == 3 ? 4.0 * (uo + vld<V>(I.v + o)) - 6.0 * vld<V>(I.a + o) - vld<V>(I.h3 + o)
             : lv == 2 ? 3.0 * (uo - vld<V>(I.a + o)) + vld<V>(I.v + o)
             : ex      ? 2.0 * uo - vld<V>(I.a + o)
                       : uo;
      } break;
      case EQ_WAVE: {  // b = (M/dt) v - c^2 K u                    CPUSolver.cpp:114
        const P3<DV> g = vgather<U>(lo, hi, P3<DV>{zd, zd, zd}, [&](int s) {
          const size_t oc = (size_t)ci[s] * K + j0;
          const DV vc = vld<V>(I.v + oc);
          const DV xc = ex ? 2.0 * vc - vld<V>(I.a + oc) : vc;
          return P3<DV>{M64[s] * vc, K64[s] * vld<V>(I.u + oc), warm ? S64[s] * xc : zd};
        });
        b = g.a * inv_dt - I.p1 * I.p1 * g.b;
        sx = g.c;
        const DV vo = vld<V>(I.v + o);
        x0 = warm ? (ex ? 2.0 * vo - vld<V>(I.a + o) : vo) : zd;
      } break;
      case EQ_NEWMARK: {  // b = -c^2 K (u + dt v + (1/2 - beta) dt^2 a)
        if (nm) {  // gather the predictor and the guess the previous update formed
          const P2<DV> g = vgather<U>(lo, hi, P2<DV>{zd, zd}, [&](int s) {
            const size_t oc = (size_t)ci[s] * K + j0;
            return P2<DV>{K64[s] * vld<V>(I.hvec + oc), warm ? S64[s] * vld<V>(I.x + oc) : zd};
          });
          b = -I.p1 * I.p1 * g.a;
          sx = g.b;
          x0 = warm ? vld<V>(I.x + o) : zd;
          break;
        }
        const double h = (0.5 - I.p2) * I.dt * I.dt;
        const P2<DV> g = vgather<U>(lo, hi, P2<DV>{zd, zd}, [&](int s) {
          const size_t oc = (size_t)ci[s] * K + j0;
          const DV ac = vld<V>(I.a + oc);
          const DV xc = lv == 2 ? 3.0 * (ac - vld<V>(I.h3 + oc)) + vld<V>(I.h4 + oc)
                        : ex    ? 2.0 * ac - vld<V>(I.h3 + oc)
                                : ac;
          return P2<DV>{K64[s] * (vld<V>(I.u + oc) + I.dt * vld<V>(I.v + oc) + h * ac), warm ? S64[s] * xc : zd};
        });
        b = -I.p1 * I.p1 * g.a;
        sx = g.b;
        const DV ao = vld<V>(I.a + o);
        x0 = !warm     ? zd
             : lv == 2 ? 3.0 * (ao - vld<V>(I.h3 + o)) + vld<V>(I.h4 + o)
             : ex      ? 2.0 * ao - vld<V>(I.h3 + o)
                       : ao;
      } break;
      default: {  // EQ_SOLVE: b supplied, x holds the guess
        b = vld<V>(I.b + o);
        if (warm)
          sx = vgather<U>(lo, hi, zd, [&](int s) { return S64[s] * vld<V>(I.x + (size_t)ci[s] * K + j0); });
        x0 = warm ? vld<V>(I.x + o) : zd;
      } break;
    }
    const DV r64 = b - sx;
    vst<V>(I.b + o, b);
    if (!pre) vst<V>(I.x + o, x0);  // pre: x holds the guess already (and is gathered)
    const TV rr = to_i<TI, V>(r64);
    vst<V>(r + o, rr);
    vst<V>(xi + o, splat<V>((TI)0));
    acc[0] += b * b;
    acc[1] += r64 * r64;
    // BiCGSTAB starts from rho = <r, r>; CG from <r, D^-1 r>
    const DV rd = to_d<V>(rr);
    acc[2] += (I.krylov == 1) ? rd * rd : rd * to_d<V>(dinv[row] * rr);
  }
}

// FP64 refinement residual r = b - S64 (x + xi); columns flagged in sh.reset
// restart from x alone (their correction is dropped).  ST: the block's
// pattern and S rows come from the LDS stage g (columns g.ci, entries g.a).
template <int K, int V, typename TI, bool ST = false>
__device__ void vphase_res(const Inst& I, int row_lo, int row_hi, const Shared& sh, vec<double, V> (&acc)[2],
                           const RhsStage& g = RhsStage{}) {
  using DV = vec<double, V>;
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  constexpr int U = vunroll<V>();
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* r = (TI*)I.r;
  TI* xi = (TI*)I.xi;
  bool dropv[V];
  bool drop = false;
#pragma unroll
  for (int v = 0; v < V; v++) drop |= (dropv[v] = sh.reset[j0 + v] != 0);
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const int GA* __restrict__ rp = gp(I.rp);
  const int GA* __restrict__ ci = gp(I.ci);
  const double GA* S64 = gp(I.S64);
  auto col = [&](int s) { if constexpr (ST) return g.ci[s]; else return ci[s]; };
  auto val = [&](int s) { if constexpr (ST) return g.a[s]; else return S64[s]; };
  const DV zd = splat<V>(0.0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    const DV bo = vld<V>(I.b + o);
    const int lo = ST ? g.rp[row - row_lo] : rp[row], hi = ST ? g.rp[row - row_lo + 1] : rp[row + 1];
    DV sx;
    if (!drop) {
      sx = vgather<U>(lo, hi, zd, [&](int s) {
        const size_t oc = (size_t)col(s) * K + j0;
        return val(s) * (vld<V>(I.x + oc) + to_d<V>(vld<V>(xi + oc)));
      });
    } else {
      sx = vgather<U>(lo, hi, zd, [&](int s) {
        const size_t oc = (size_t)col(s) * K + j0;
        DV c = to_d<V>(vld<V>(xi + oc));
#pragma unroll
        for (int v = 0; v < V; v++) c[v] = dropv[v] ? 0.0 : c[v];
        return val(s) * (vld<V>(I.x + oc) + c);
      });
    }
    const DV r64 = bo - sx;
    const TV rr = to_i<TI, V>(r64);
    vst<V>(r + o, rr);
    if (drop)
#pragma unroll
      for (int v = 0; v < V; v++)
        if (dropv[v]) gp(xi)[o + v] = (TI)0;
    acc[0] += r64 * r64;
    const DV rd = to_d<V>(rr);
    acc[1] += (I.krylov == 1) ? rd * rd : rd * to_d<V>(dinv[row] * rr);
  }
}

// Pipelined CG start: fold x += xi; w = S D^-1 r; <r, D^-1 r>, <D^-1 r, w>.
template <int K, int V, typename TI>
__device__ void vphase_pcg_init(const Inst& I, int row_lo, int row_hi, const Stage<TI>& g, const TI* rc, TI* wc,
                                vec<double, V> (&acc)[2]) {
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  constexpr int U = vunroll<V>();
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* xi = (TI*)I.xi;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TV zt = splat<V>((TI)0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    const TV ro = vld<V>(rc + o);
    const int lr = row - g.row_lo;
    const TV wa = vgather<U>(g.rp[lr], g.rp[lr + 1], zt, [&](int s) {
      return g.w[s] * vld<V>(rc + (size_t)g.ci[s] * K + j0);
    });
    vst<V>(I.x + o, vld<V>(I.x + o) + to_d<V>(vld<V>(xi + o)));
    vst<V>(xi + o, zt);
    vst<V>(wc + o, wa);
    const TV uo = dinv[row] * ro;
    acc[0] += to_d<V>(ro) * to_d<V>(uo);
    acc[1] += to_d<V>(uo) * to_d<V>(wa);
  }
}

// One Chronopoulos-Gear iteration.
template <int K, int V, typename TI>
__device__ void vphase_pcg_iter(const Inst& I, int row_lo, int row_hi, const Stage<TI>& g, const TI* rc, const TI* wc,
                                const TI* sc, TI* rn, TI* wn, TI* sn, bool first, const Shared& sh,
                                vec<double, V> (&acc)[3]) {
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  constexpr int U = vunroll<V>();
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* xi = (TI*)I.xi;
  TI* p = (TI*)I.p0;  // own rows only
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TV zt = splat<V>((TI)0);
  const TV al = colv<TI, V>(sh.alpha, j0);
  const TV be = first ? zt : colv<TI, V>(sh.beta, j0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    const TV ro = vld<V>(rc + o);
    const TV wo = vld<V>(wc + o);
    const TV so = first ? zt : vld<V>(sc + o);
    const TV po = first ? zt : vld<V>(p + o);
    const TV xio = vld<V>(xi + o);
    const int lr = row - g.row_lo;
    const TV wa = vgather<U>(g.rp[lr], g.rp[lr + 1], zt, [&](int s) {
      const size_t oc = (size_t)g.ci[s] * K + j0;
      const TV sv = cg_s<TV>(vld<V>(wc + oc), first ? zt : vld<V>(sc + oc), be, first);
      return g.w[s] * cg_r<TV>(vld<V>(rc + oc), sv, al);
    });
    const TI d = dinv[row];
    const TV uo = d * ro;
    const TV pn = first ? uo : (uo + be * po);
    const TV sv = cg_s<TV>(wo, so, be, first);
    const TV rr = cg_r<TV>(ro, sv, al);
    vst<V>(p + o, pn);
    vst<V>(xi + o, xio + al * pn);
    vst<V>(sn + o, sv);
    vst<V>(rn + o, rr);
    vst<V>(wn + o, wa);
    const TV un = d * rr;
    acc[0] += to_d<V>(rr) * to_d<V>(un);
    acc[1] += to_d<V>(un) * to_d<V>(wa);
    acc[2] += to_d<V>(rr) * to_d<V>(rr);
  }
}

// Classic CG phase A: p = z + beta p_old (z = D^-1 r), q = S p, <p, q>.
template <int K, int V, typename TI>
__device__ void vphase_cg_a(const Inst& I, int row_lo, int row_hi, const Stage<TI>& g, const TI* r, const TI* pold,
                            TI* pnew, bool first, const Shared& sh, vec<double, V> (&acc)[1]) {
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  constexpr int U = vunroll<V>();
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* q = (TI*)I.q;  // own rows only
  TI* xi = (TI*)I.xi;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TV zt = splat<V>((TI)0);
  const TV bj = first ? zt : colv<TI, V>(sh.beta, j0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    if (first) {
      vst<V>(I.x + o, vld<V>(I.x + o) + to_d<V>(vld<V>(xi + o)));
      vst<V>(xi + o, zt);
    }
    const int lr = row - g.row_lo;
    // p = D^-1 r + beta p_old is only partly column-scaled: g.w holds S itself here
    const TV qa = vgather<U>(g.rp[lr], g.rp[lr + 1], zt, [&](int s) {
      const int c = g.ci[s];
      const size_t oc = (size_t)c * K + j0;
      const TV zc = dinv[c] * vld<V>(r + oc);
      return g.w[s] * (first ? zc : (zc + bj * vld<V>(pold + oc)));
    });
    const TV zo = dinv[row] * vld<V>(r + o);
    const TV pn = first ? zo : (zo + bj * vld<V>(pold + o));
    vst<V>(pnew + o, pn);
    vst<V>(q + o, qa);
    acc[0] += to_d<V>(pn) * to_d<V>(qa);
  }
}

// Classic CG phase B: xi += alpha p, r -= alpha q, <r, D^-1 r>, <r, r>.
template <int K, int V, typename TI>
__device__ void vphase_cg_b(const Inst& I, int row_lo, int row_hi, TI* r, const TI* p, const Shared& sh,
                            vec<double, V> (&acc)[2]) {
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* xi = (TI*)I.xi;
  const TI* q = (const TI*)I.q;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TV al = colv<TI, V>(sh.alpha, j0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    vst<V>(xi + o, vld<V>(xi + o) + al * vld<V>(p + o));
    const TV rr = vld<V>(r + o) - al * vld<V>(q + o);
    vst<V>(r + o, rr);
    acc[0] += to_d<V>(rr) * to_d<V>(dinv[row] * rr);
    acc[1] += to_d<V>(rr) * to_d<V>(rr);
  }
}

// BiCGSTAB (right Jacobi preconditioning) in three phases, arranged so that
// the gathering phases read few vectors per neighbour.  Rho of the next
// iteration, <rh, s - omega t> = <rh, s> - omega <rh, t>, comes out of the s
// phase's dot products, so beta is known before the x phase, which then also
// forms the next search direction p = r + beta (p - omega v) on its own rows:
//   P: v = S D^-1 p; <rh, v>          (first iteration: p = rh = r, x += xi)
//   S: s = r - alpha v; t = S D^-1 s; <t,s>, <t,t>, <s,s>, <rh,s>, <rh,t>
//   X: s = r - alpha v again; xi += D^-1 (alpha p + omega s);
//      r = s - omega t; p = r + beta (p - omega v); <r,r>
// Phase P gathers one vector and S two (the persistent kernel's phase P
// rebuilds p from three), and s is never stored: phase X reads v anyway.
// Per element the expressions are the persistent kernel's (s rounded once
// more, below); rho is summed from s and t instead of from the rounded r.

// s = r - alpha v, product and difference rounded separately (no fused
// multiply-add), so that phases S and X form the same values.
template <typename TI, int V>
__device__ inline vec<TI, V> bi_s(const vec<TI, V>& r, const vec<TI, V>& al, const vec<TI, V>& v) {
#pragma clang fp contract(off)
  const vec<TI, V> av = al * v;
  return r - av;
}

template <int K, int V, typename TI>
__device__ void vphase_bi_p(const Inst& I, int row_lo, int row_hi, const Stage<TI>& g, const TI* r, TI* p, TI* v,
                            bool first, vec<double, V> (&acc)[1]) {
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  constexpr int U = vunroll_bi<V>();
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* rh = (TI*)I.rh;  // own rows only
  TI* xi = (TI*)I.xi;
  const TV zt = splat<V>((TI)0);
  const TI* src = first ? r : p;  // p is written on own rows only in the first iteration
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    TV ho;
    if (first) {
      ho = vld<V>(r + o);
      vst<V>(I.x + o, vld<V>(I.x + o) + to_d<V>(vld<V>(xi + o)));
      vst<V>(xi + o, zt);
      vst<V>(rh + o, ho);
      vst<V>(p + o, ho);
    } else {
      ho = vld<V>(rh + o);
    }
    const int lr = row - g.row_lo;
    const TV va = vgather<U>(g.rp[lr], g.rp[lr + 1], zt,
                             [&](int s) { return g.w[s] * vld<V>(src + (size_t)g.ci[s] * K + j0); });
    vst<V>(v + o, va);
    acc[0] += to_d<V>(ho) * to_d<V>(va);
  }
}

// BiCGSTAB phase S: s = r - alpha v; t = S D^-1 s; <t,s>, <t,t>, <s,s>,
// <rh,s>, <rh,t>.
template <int K, int V, typename TI>
__device__ void vphase_bi_s(const Inst& I, int row_lo, int row_hi, const Stage<TI>& g, const TI* r, const TI* v,
                            const Shared& sh, vec<double, V> (&acc)[5]) {
  using TV = vec<TI, V>;
  using DV = vec<double, V>;
  using L = VLay<K, V>;
  constexpr int U = vunroll_bi<V>();
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* tv = (TI*)I.t;  // own rows only
  const TI* rh = (const TI*)I.rh;
  const TV zt = splat<V>((TI)0);
  const TV al = colv<TI, V>(sh.alpha, j0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    const TV so = bi_s<TI, V>(vld<V>(r + o), al, vld<V>(v + o));
    const TV ho = vld<V>(rh + o);
    const int lr = row - g.row_lo;
    const TV ta = vgather<U>(g.rp[lr], g.rp[lr + 1], zt, [&](int s) {
      const size_t oc = (size_t)g.ci[s] * K + j0;
      return g.w[s] * bi_s<TI, V>(vld<V>(r + oc), al, vld<V>(v + oc));
    });
    vst<V>(tv + o, ta);
    const DV sd = to_d<V>(so), td = to_d<V>(ta), hd = to_d<V>(ho);
    acc[0] += td * sd;
    acc[1] += td * td;
    acc[2] += sd * sd;
    acc[3] += hd * sd;
    acc[4] += hd * td;
  }
}

// BiCGSTAB phase X: s = r - alpha v; xi += D^-1 (alpha p + omega s);
// r = s - omega t; p = r + beta (p - omega v); <r,r>.
template <int K, int V, typename TI>
__device__ void vphase_bi_x(const Inst& I, int row_lo, int row_hi, TI* r, TI* p, const TI* v, const Shared& sh,
                            vec<double, V> (&acc)[1]) {
  using TV = vec<TI, V>;
  using L = VLay<K, V>;
  const int j0 = ((int)threadIdx.x % L::G) * V;
  TI* xi = (TI*)I.xi;
  const TI* tv = (const TI*)I.t;
  const TI GA* dinv = gp((const TI*)I.dinvi);
  const TV al = colv<TI, V>(sh.alpha, j0);
  const TV om = colv<TI, V>(sh.omega, j0);
  const TV be = colv<TI, V>(sh.beta, j0);
  for (int row = row_lo + (int)threadIdx.x / L::G; row < row_hi; row += L::RPP) {
    const size_t o = (size_t)row * K + j0;
    const TI d = dinv[row];
    const TV vo = vld<V>(v + o);
    const TV so = bi_s<TI, V>(vld<V>(r + o), al, vo);  // phase S's s, bit for bit
    const TV to = vld<V>(tv + o);
    const TV po = vld<V>(p + o);
    vst<V>(xi + o, vld<V>(xi + o) + d * (al * po + om * so));
    const TV rr = so - om * to;
    vst<V>(r + o, rr);
    vst<V>(p + o, rr + be * (po - om * vo));
    acc[0] += to_d<V>(rr) * to_d<V>(rr);
  }
}

// Per-lane vector partial sums acc[0 .. NS) -> the block's sums at
// sh.red[q K + j], summed over the lanes of each column in a fixed order.
template <int K, int V, int NS>
__device__ inline void vblock_partials(const vec<double, V>* acc, Shared& sh) {
  static_assert(NS <= kMaxSlots, "block sums per round");
  constexpr int G = K / V;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NS; q++)
#pragma unroll
    for (int v = 0; v < V; v++) {
      double x = acc[q][v];
#pragma unroll
      for (int m = kWave / 2; m >= G; m >>= 1) x += __shfl_xor(x, m, kWave);
      if (lane < G) sh.wred[wave][q * K + lane * V + v] = x;
    }
  __syncthreads();
  if (threadIdx.x < NS * K) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; w++) s += sh.wred[w][threadIdx.x];
    sh.red[threadIdx.x] = s;
  }
}

}  // namespace

}  // namespace femhip
