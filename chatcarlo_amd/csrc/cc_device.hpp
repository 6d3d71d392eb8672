// Device-side building blocks of the gfx950 photon-transport path.
//
// Float32 everywhere except the FP64 tallies and collision accumulators.
// Cancellation-prone geometry terms (|o-c|^2 - r^2, o.d) are evaluated in
// float-float arithmetic (Dekker 1971: error-free TwoSum / FMA TwoProd), so a
// float32 kernel reproduces the FP64 roots to ~1e-7 relative.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/chatcarlo_hip.h"

namespace cc {

// v[a] for a run-time a, formed arithmetically: a select chain over the elements of
// one array or struct is folded by the compiler into a run-time index, which moves
// the whole aggregate from registers to scratch memory.
__device__ __forceinline__ float sel3(int a, float v0, float v1, float v2) {
  const uint32_t m0 = 0u - (uint32_t)(a == 0), m1 = 0u - (uint32_t)(a == 1), m2 = 0u - (uint32_t)(a == 2);
  return __uint_as_float((__float_as_uint(v0) & m0) | (__float_as_uint(v1) & m1) | (__float_as_uint(v2) & m2));
}
__device__ __forceinline__ int sel3i(int a, int v0, int v1, int v2) {
  return v0 * (a == 0) + v1 * (a == 1) + v2 * (a == 2);
}


constexpr float kHC = 12.3984193f;      // chatcarlo/physics.py:35
constexpr float kMEC2 = 511.0f;         // chatcarlo/physics.py:34
constexpr float kFluorCut = 5.0f;       // chatcarlo/physics.py:38
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr int kMaxEdges = 16;
constexpr int kRayGuide = 1024;   // guide-table buckets of each element's Rayleigh cumulative A(x)
constexpr int kFlStride = 32;   // edge, omega, 1/dlnE, ln edge, lines[8], cdf[8], K share[12]
constexpr int kKFracNodes = 12;
constexpr int kNParts = 5;
enum { P_PHOTO = 0, P_COMPT = 1, P_RAYL = 2, P_MUEN = 3, P_TOTAL = 4 };

struct Prim {
  int kind, axis, material, pad;
  float p[8];
};
struct World {
  float lo[3], hi[3];
  int background, n_prim;
};

struct Tables {
  int n_mat, n_e, n_el, n_q, n_h10;
  const float* density;
  const float* e_node;
  const float* dl_node;
  const float2* ss;          // [n_mat][5][n_e] (sig, slope)
  const int* edge_idx;       // [n_mat][n_e]
  const float* edge_e;       // [n_mat][16]
  const float* edge_dl;      // [n_mat][16][2]
  const float* edge_sig;     // [n_mat][16][3][2]
  const float* edge_slope;   // [n_mat][16][3][2]
  const int* mat_elem_off;
  const int* mat_elem;
  const int* sel_off;
  const float* sel;
  const float* sq_over_z;    // [n_el][n_q]
  const float* q_s;
  const float* ray_x;
  const float* ray_a;        // [n_el][n_q]
  const int* ray_guide;      // [n_el][kRayGuide]: the search start of ray_a for a target in bucket k
  const float* ray_gscale;   // [n_el]: kRayGuide / ray_a[el][n_q - 1] (bucket = floor(target scale))
  const float* fl;           // [n_el][20]
  const float* h10;          // [4][n_h10]
};

// ---------------------------------------------------------------------------
// float-float helpers
// ---------------------------------------------------------------------------
struct ff { float h, l; };
// FMA contraction is switched off inside every helper: fusing a rounded
// product into a later add would break the error-free transformations.
__device__ __forceinline__ ff two_sum(float a, float b) {
#pragma clang fp contract(off)
  float s = a + b;
  float bb = s - a;
  float e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
__device__ __forceinline__ ff two_prod(float a, float b) {
#pragma clang fp contract(off)
  float p = a * b;
  return {p, __builtin_fmaf(a, b, -p)};
}
__device__ __forceinline__ ff ff_add(ff a, ff b) {
#pragma clang fp contract(off)
  ff s = two_sum(a.h, b.h);
  float l = s.l + a.l + b.l;
  return two_sum(s.h, l);
}
__device__ __forceinline__ ff ff_neg(ff a) { return {-a.h, -a.l}; }
__device__ __forceinline__ ff ff_mul(ff a, ff b) {
#pragma clang fp contract(off)
  ff p = two_prod(a.h, b.h);
  float l = p.l + (a.h * b.l + a.l * b.h);
  return two_sum(p.h, l);
}
__device__ __forceinline__ ff ff_sqr(ff a) { return ff_mul(a, a); }
__device__ __forceinline__ float ff_f(ff a) { return a.h + a.l; }
__device__ __forceinline__ ff ff_sqrt(ff a) {
#pragma clang fp contract(off)
  if (a.h <= 0.f) return {0.f, 0.f};
  float s = sqrtf(a.h);
  // one Newton step: s' = s + (a - s^2) / (2 s)
  ff s2 = two_prod(s, s);
  float r = ((a.h - s2.h) - s2.l) + a.l;
  return two_sum(s, r / (2.f * s));
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}
__device__ __forceinline__ float u01(uint32_t x) {   // (0, 1], float32-exact
  return (float)((x >> 8) + 1u) * 5.9604644775390625e-8f;
}

// Per-history stream: counter (history lo, history hi, event counter, lane tag), key = seed.
struct PhiloxRegRng;
// Tag of a history's heel stream: the source sampler's heel-effect rejection draws from the
// history's Philox stream under tag ^ kHeelTag, counters from 0.
constexpr uint32_t kHeelTag = 0x48454C00u;

struct PhiloxRng {
  uint2 key;
  uint32_t h_lo, h_hi, tag, ctr;
  float buf[4];
  int pos;
  __device__ PhiloxRng(uint64_t seed, uint64_t hist, uint32_t tag_, uint32_t ctr_)
      : key(make_uint2((uint32_t)seed, (uint32_t)(seed >> 32))), h_lo((uint32_t)hist),
        h_hi((uint32_t)(hist >> 32)), tag(tag_), ctr(ctr_), pos(4) {}
  __device__ __forceinline__ float next() {
    if (pos == 4) {
      uint4 r = philox4x32_10(make_uint4(h_lo, h_hi, ctr, tag), key);
      ++ctr;
      buf[0] = u01(r.x); buf[1] = u01(r.y); buf[2] = u01(r.z); buf[3] = u01(r.w);
      pos = 0;
    }
    return buf[pos++];
  }
  __device__ __forceinline__ PhiloxRegRng heel_stream() const;
};

// The same stream as PhiloxRng, with the block's four uniforms in four registers picked by
// selects.  In the interaction kernels an indexed buffer lands in scratch memory, every draw
// then waiting on a scratch load; with the selects k_interact needs no scratch at five waves
// per SIMD.  (Where the buffer is promoted to LDS instead, as in the flight and splitting
// kernels, PhiloxRng is as fast or faster.)
struct PhiloxRegRng {
  uint2 key;
  uint32_t h_lo, h_hi, tag, ctr;
  float b0, b1, b2, b3;
  int pos;
  __device__ PhiloxRegRng(uint64_t seed, uint64_t hist, uint32_t tag_, uint32_t ctr_)
      : key(make_uint2((uint32_t)seed, (uint32_t)(seed >> 32))), h_lo((uint32_t)hist),
        h_hi((uint32_t)(hist >> 32)), tag(tag_), ctr(ctr_), b0(0.f), b1(0.f), b2(0.f), b3(0.f), pos(4) {}
  __device__ __forceinline__ float next() {
    if (pos == 4) {
      uint4 r = philox4x32_10(make_uint4(h_lo, h_hi, ctr, tag), key);
      ++ctr;
      b0 = u01(r.x); b1 = u01(r.y); b2 = u01(r.z); b3 = u01(r.w);
      pos = 0;
    }
    const float v = pos == 0 ? b0 : (pos == 1 ? b1 : (pos == 2 ? b2 : b3));
    ++pos;
    return v;
  }
  __device__ __forceinline__ PhiloxRegRng heel_stream() const {
    return PhiloxRegRng(((uint64_t)key.y << 32) | key.x, ((uint64_t)h_hi << 32) | h_lo, tag ^ kHeelTag, 0u);
  }
};

__device__ __forceinline__ PhiloxRegRng PhiloxRng::heel_stream() const {
  return PhiloxRegRng(((uint64_t)key.y << 32) | key.x, ((uint64_t)h_hi << 32) | h_lo, tag ^ kHeelTag, 0u);
}

// Caller-supplied uniforms (sub-kernel parity tests).
struct BufRng {
  const float* u;
  int n, pos;
  __device__ BufRng(const float* u_, int n_) : u(u_), n(n_), pos(0) {}
  __device__ __forceinline__ float next() { return pos < n ? u[pos++] : 1.0f; }
  __device__ __forceinline__ float at(int slot) const { return slot < n ? u[slot] : 1.0f; }
};

// ---------------------------------------------------------------------------
// geometry (chatcarlo/geometry.py:20-231)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void slab(const float* lo, const float* hi, const float o[3],
                                     const float d[3], const float inv_d[3], float& ten, float& tex, bool& hit) {
  ten = -INFINITY;
  tex = INFINITY;
  bool miss = false;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (fabsf(d[k]) < 1e-12f) {
      miss |= (o[k] < lo[k]) || (o[k] > hi[k]);
    } else {
      float ta = (lo[k] - o[k]) * inv_d[k];
      float tb = (hi[k] - o[k]) * inv_d[k];
      ten = fmaxf(ten, fminf(ta, tb));
      tex = fminf(tex, fmaxf(ta, tb));
    }
  }
  hit = !miss && (ten <= tex);
}

// Stable quadratic a t^2 + 2 b t + c = 0 with a, b, c, disc in float-float.
__device__ __forceinline__ bool quad_roots(ff aa, ff b, ff c, float& t1, float& t2) {
  ff disc = ff_add(ff_mul(b, b), ff_neg(ff_mul(aa, c)));
  const float a = ff_f(aa);
  if (disc.h < 0.f || (disc.h == 0.f && disc.l < 0.f)) return false;
  ff sq = ff_sqrt(disc);
  float sb = (b.h >= 0.f) ? 1.f : -1.f;
  ff q = ff_neg(ff_add(b, {sb * sq.h, sb * sq.l}));   // q = -(b + sign(b) sqrt(disc))
  float qf = ff_f(q);
  float r1, r2;
  if (qf == 0.f) {
    r1 = r2 = 0.f;
  } else {
    r1 = qf / a;
    r2 = ff_f(c) / qf;
  }
  t1 = fminf(r1, r2);
  t2 = fmaxf(r1, r2);
  return true;
}

__device__ __forceinline__ bool isect_sphere(const Prim& P, const float o[3], const float d[3],
                                             float& t1, float& t2) {
  ff b = {0.f, 0.f}, cc = {0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    ff oc = two_sum(o[k], -P.p[k]);
    b = ff_add(b, ff_mul(oc, {d[k], 0.f}));
    cc = ff_add(cc, ff_sqr(oc));
  }
  cc = ff_add(cc, ff_neg(two_prod(P.p[3], P.p[3])));
  // unit direction: a = 1 (chatcarlo/geometry.py:45-54)
  return quad_roots({1.0f, 0.0f}, b, cc, t1, t2);
}

__device__ __forceinline__ bool isect_cylinder(const Prim& P, const float o[3], const float d[3],
                                               float& t1, float& t2) {
  const int ax = P.axis;
  const int i0 = (ax == 0) ? 1 : 0;
  const int i1 = (ax == 2) ? 1 : 2;
  ff ox = two_sum(sel3(i0, o[0], o[1], o[2]), -sel3(i0, P.p[0], P.p[1], P.p[2]));
  ff oy = two_sum(sel3(i1, o[0], o[1], o[2]), -sel3(i1, P.p[0], P.p[1], P.p[2]));
  float dx = sel3(i0, d[0], d[1], d[2]), dy = sel3(i1, d[0], d[1], d[2]);
  ff aa = ff_add(two_prod(dx, dx), two_prod(dy, dy));
  float a = ff_f(aa);
  ff b = ff_add(ff_mul(ox, {dx, 0.f}), ff_mul(oy, {dy, 0.f}));
  ff cc = ff_add(ff_add(ff_sqr(ox), ff_sqr(oy)), ff_neg(two_prod(P.p[3], P.p[3])));
  float ten_r, tex_r;
  bool miss_r = false;
  if (a < 1e-12f) {
    if (cc.h < 0.f || (cc.h == 0.f && cc.l <= 0.f)) {
      ten_r = -INFINITY;
      tex_r = INFINITY;
    } else {
      miss_r = true;
      ten_r = tex_r = 0.f;
    }
  } else {
    if (!quad_roots(aa, b, cc, ten_r, tex_r)) miss_r = true;
  }
  // caps at c_ax +- half_h, plane offsets formed exactly in float-float
  float dz = sel3(ax, d[0], d[1], d[2]);
  const float oz = sel3(ax, o[0], o[1], o[2]);
  const float cz = sel3(ax, P.p[0], P.p[1], P.p[2]);
  ff lo_rel = ff_add(two_sum(cz, -P.p[4]), {-oz, 0.f});
  ff hi_rel = ff_add(two_sum(cz, P.p[4]), {-oz, 0.f});
  float ten_z, tex_z;
  bool miss_z = false;
  if (fabsf(dz) < 1e-12f) {
    // o outside [lo, hi]  <=>  lo - o > 0 or hi - o < 0
    miss_z = (lo_rel.h > 0.f || (lo_rel.h == 0.f && lo_rel.l > 0.f)) ||
             (hi_rel.h < 0.f || (hi_rel.h == 0.f && hi_rel.l < 0.f));
    ten_z = -INFINITY;
    tex_z = INFINITY;
  } else {
    float ta = ff_f(lo_rel) / dz, tb = ff_f(hi_rel) / dz;
    ten_z = fminf(ta, tb);
    tex_z = fmaxf(ta, tb);
  }
  if (miss_r || miss_z) return false;
  t1 = fmaxf(ten_r, ten_z);
  t2 = fminf(tex_r, tex_z);
  return !(t1 > t2);
}

__device__ __forceinline__ bool isect_prim(const Prim& P, const float o[3], const float d[3], const float inv_d[3],
                                           float& t1, float& t2) {
  if (P.kind == CC_PRIM_BOX) {
    bool hit;
    slab(&P.p[0], &P.p[3], o, d, inv_d, t1, t2, hit);
    return hit;
  }
  if (P.kind == CC_PRIM_SPHERE) return isect_sphere(P, o, d, t1, t2);
  return isect_cylinder(P, o, d, t1, t2);
}

// sign of (|p - c|^2 - r^2) <= 0 evaluated in float-float
__device__ __forceinline__ bool contains(const Prim& P, const float q[3]) {
  if (P.kind == CC_PRIM_BOX) {
    return q[0] >= P.p[0] && q[0] <= P.p[3] && q[1] >= P.p[1] && q[1] <= P.p[4] &&
           q[2] >= P.p[2] && q[2] <= P.p[5];
  }
  if (P.kind == CC_PRIM_SPHERE) {
    ff s = {0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 3; ++k) s = ff_add(s, ff_sqr(two_sum(q[k], -P.p[k])));
    s = ff_add(s, ff_neg(two_prod(P.p[3], P.p[3])));
    return s.h < 0.f || (s.h == 0.f && s.l <= 0.f);
  }
  const int ax = P.axis;
  const int i0 = (ax == 0) ? 1 : 0;
  const int i1 = (ax == 2) ? 1 : 2;
  ff s = ff_add(ff_sqr(two_sum(sel3(i0, q[0], q[1], q[2]), -sel3(i0, P.p[0], P.p[1], P.p[2]))),
                ff_sqr(two_sum(sel3(i1, q[0], q[1], q[2]), -sel3(i1, P.p[0], P.p[1], P.p[2]))));
  s = ff_add(s, ff_neg(two_prod(P.p[3], P.p[3])));
  if (!(s.h < 0.f || (s.h == 0.f && s.l <= 0.f))) return false;
  ff z = two_sum(sel3(ax, q[0], q[1], q[2]), -sel3(ax, P.p[0], P.p[1], P.p[2]));
  ff za = (z.h < 0.f || (z.h == 0.f && z.l < 0.f)) ? ff_neg(z) : z;
  ff e = ff_add(za, {-P.p[4], 0.f});
  return e.h < 0.f || (e.h == 0.f && e.l <= 0.f);
}

// A scene's primitives read through the constant address space.  In every loop over the
// primitives the index is uniform across the wave, so these loads are scalar (s_load, through
// the scalar cache) and the fields live in SGPRs: the kind of each primitive is a scalar branch,
// and no VGPRs or LDS round trips go to them.  Valid for memory that no kernel writes while
// it is read (a scene is uploaded before its launches).
struct ConstPrims {
  const __attribute__((address_space(4))) int* w;
};
__device__ __forceinline__ ConstPrims const_prims(const Prim* p) {
  return ConstPrims{(const __attribute__((address_space(4))) int*)p};
}
__device__ __forceinline__ Prim load_prim(ConstPrims s, int i) {
  static_assert(sizeof(Prim) == 12 * sizeof(int), "Prim is 12 words");
  const __attribute__((address_space(4))) int* q = s.w + 12 * i;
  Prim P;
  P.kind = q[0];
  P.axis = q[1];
  P.material = q[2];
  P.pad = q[3];
#pragma unroll
  for (int k = 0; k < 8; ++k) P.p[k] = __int_as_float(q[4 + k]);
  return P;
}
// per-lane primitives (batched scenes: each lane's scene has its own array)
__device__ __forceinline__ Prim load_prim(const Prim* s, int i) { return s[i]; }

// Last primitive containing the point wins (chatcarlo/geometry.py:205-212).
template <class PS>
__device__ __forceinline__ int locate(PS prims, int n_prim, int background, const float q[3]) {
  for (int i = n_prim - 1; i >= 0; --i) {
    const Prim P = load_prim(prims, i);
    if (contains(P, q)) return P.material;
  }
  return background;
}

// Min root beyond tmin over all primitives vs the world exit (chatcarlo/geometry.py:214-231,
// where tmin is the reference's 1e-6 cm threshold).
template <class PS>
__device__ __forceinline__ float next_boundary(PS prims, int n_prim, const World& W,
                                               const float o[3], const float d[3], float tmin,
                                               bool& escape) {
  // one reciprocal per axis and call, not two divisions per axis and box; every box root is
  // formed the same way on every call, so a crossed surface keeps the root it had
  const float inv_d[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
  float t_obj = INFINITY;
  for (int i = 0; i < n_prim; ++i) {
    float t1, t2;
    if (isect_prim(load_prim(prims, i), o, d, inv_d, t1, t2)) {
      if (t1 > tmin && t1 < t_obj) t_obj = t1;
      if (t2 > tmin && t2 < t_obj) t_obj = t2;
    }
  }
  float ten, tex;
  bool hit;
  slab(W.lo, W.hi, o, d, inv_d, ten, tex, hit);
  if (!hit) tex = INFINITY;
  escape = tex <= t_obj;
  return fminf(t_obj, tex);
}

// ---------------------------------------------------------------------------
// cross-section lookup (chatcarlo/materials.py:228-262 restated on the pack)
// ---------------------------------------------------------------------------
struct EPos {
  int i;       // interval
  float lr;    // ln(E / E_i)
  float f;     // lr / ln(E_{i+1}/E_i)
};

__device__ __forceinline__ EPos energy_pos(const Tables& T, float E) {
  const float inv_dln = (float)(T.n_e - 1) / 5.0106352940962555f;   // ln(150)
  int i = (int)(logf(E) * inv_dln);
  i = min(max(i, 0), T.n_e - 2);
  if (E < T.e_node[i] && i > 0) --i;
  else if (E >= T.e_node[i + 1] && i < T.n_e - 2) ++i;
  float ei = T.e_node[i];
  EPos p;
  p.i = i;
  p.lr = log1pf((E - ei) / ei);
  p.f = p.lr / T.dl_node[i];
  return p;
}

// part value; edge parts: P_PHOTO -> 0, P_MUEN -> 1, P_TOTAL -> 2
__device__ __forceinline__ float xs_part(const Tables& T, int m, int part, float E, const EPos& p) {
  float2 s = T.ss[((size_t)m * kNParts + part) * T.n_e + p.i];
  int ep = (part == P_PHOTO) ? 0 : (part == P_MUEN) ? 1 : (part == P_TOTAL) ? 2 : -1;
  if (ep >= 0) {
    int k = T.edge_idx[m * T.n_e + p.i];
    if (k >= 0) {
      const int r = m * kMaxEdges + k;
      float ee = T.edge_e[r];
      if (E < ee) {
        float fb = p.lr / T.edge_dl[r * 2 + 0];
        return s.x * expf(fb * T.edge_slope[(r * 3 + ep) * 2 + 0]);
      }
      float fa = log1pf((E - ee) / ee) / T.edge_dl[r * 2 + 1];
      return T.edge_sig[(r * 3 + ep) * 2 + 1] * expf(fa * T.edge_slope[(r * 3 + ep) * 2 + 1]);
    }
  }
  return s.x * expf(p.f * s.y);
}

__device__ __forceinline__ bool energy_ok(float E) { return E >= 1.0f && E <= 150.0f; }

// h*(10)/Phi [pSv cm^2], log-log with clamping (chatcarlo/dose_coefficients.py:32-42)
__device__ __forceinline__ float h10_coeff(const float* h10, int n, float E) {
  const float* he = h10;
  const float* hv = h10 + n;
  const float* hd = h10 + 2 * n;
  const float* hs = h10 + 3 * n;
  float ec = fminf(fmaxf(E, he[0]), he[n - 1]);
  // the last j <= n - 2 with he[j] <= ec (binary search; the table is sorted)
  int j = 0, hi = n - 2;
  while (j < hi) {
    const int mid = (j + hi + 1) >> 1;
    if (ec >= he[mid]) j = mid; else hi = mid - 1;
  }
  float f = logf(ec / he[j]) / hd[j];
  return hv[j] * expf(f * hs[j]);
}

// ---------------------------------------------------------------------------
// samplers (chatcarlo/physics.py), generic over the uniform source
// ---------------------------------------------------------------------------
__device__ __forceinline__ int select_elem(const Tables& T, int m, int kind, const EPos& p, float r) {
  const int off = T.mat_elem_off[m];
  const int ne = T.mat_elem_off[m + 1] - off;
  const float* c0 = T.sel + T.sel_off[m] + ((size_t)kind * T.n_e + p.i) * ne;
  const float* c1 = c0 + ne;
  int k = 0;
  for (int j = 0; j < ne; ++j) {
    float c = c0[j] + p.f * (c1[j] - c0[j]);
    k += (r > c) ? 1 : 0;
  }
  k = min(k, ne - 1);
  return T.mat_elem[off + k];
}

// np.interp(x, xp, fp) on a sorted table; caller gives a starting guess j0.
// The loads of the common case (the guess brackets x) are issued together, then checked.
__device__ __forceinline__ float interp_guess(const float* xp, const float* fp, int n, float x, int j) {
  j = min(max(j, 0), n - 2);
  const float xa = xp[0], xb = xp[n - 1], fa = fp[0], fb = fp[n - 1];
  float x0 = xp[j], x1 = xp[j + 1], f0 = fp[j], f1 = fp[j + 1];
  if (x <= xa) return fa;
  if (x >= xb) return fb;
  if (!(x >= x0 && x < x1)) {
    while (j > 0 && x < xp[j]) --j;
    while (j < n - 2 && x >= xp[j + 1]) ++j;
    x0 = xp[j]; x1 = xp[j + 1]; f0 = fp[j]; f1 = fp[j + 1];
  }
  return f0 + (f1 - f0) / (x1 - x0) * (x - x0);
}

template <class Rng>
__device__ __forceinline__ float compton_sample(const Tables& T, int m, float E, const EPos& p, Rng& rng,
                                                int kmax, float& cos_t, float& sin_t, int& el, int& trials) {
  el = select_elem(T, m, P_COMPT, p, rng.next());
  const float a = E / kMEC2;
  const float emin = 1.0f / (1.0f + 2.0f * a);
  const float env = 1.0f / emin + emin;
  const float* sq = T.sq_over_z + (size_t)el * T.n_q;
  const float dq = (20.0f - 1e-3f) / (float)(T.n_q - 2);
  for (int t = 0; t < kmax; ++t) {
    float x1 = rng.next(), x2 = rng.next(), x3 = rng.next();
    float eps = emin + x1 * (1.0f - emin);
    // 1 - cos = (1/eps - 1)/a written without the cancellation of 1/eps - 1:
    // 1 - eps = (1 - emin)(1 - x1) and 1 - emin = 2a/(1 + 2a)
    float omc = 2.0f * (1.0f - x1) / ((1.0f + 2.0f * a) * eps);
    float cs = 1.0f - omc;
    float sin2 = omc * (2.0f - omc);
    float g = 1.0f / eps + eps - sin2;
    bool acc = x2 * env <= g;
    float omcc = fminf(fmaxf(omc, 0.0f), 2.0f);
    float q = E * sqrtf(omcc * 0.5f) / kHC;
    int jg = (q < 1e-3f) ? 0 : 1 + (int)((q - 1e-3f) / dq);
    float soz = interp_guess(T.q_s, sq, T.n_q, q, jg);
    acc = acc && (x3 <= soz);
    if (acc) {
      cos_t = cs;
      // sin^2 = (1 - cos)(1 + cos) with 1 + cos = 2 x1 / eps
      sin_t = 2.0f * sqrtf(x1 * (1.0f - x1) / (1.0f + 2.0f * a)) / eps;
      trials = t + 1;
      return eps;
    }
  }
  trials = 0;
  cos_t = 1.0f;
  sin_t = 0.0f;
  return 1.0f;
}

// returns cos theta; trials = 0 when the cap is hit; qerr set if x_max beyond the table
template <class Rng>
__device__ __forceinline__ float rayleigh_sample(const Tables& T, int m, float E, const EPos& p, Rng& rng,
                                                 int kmax, float& sin_t, int& el, int& trials, bool& qerr) {
  el = select_elem(T, m, P_RAYL, p, rng.next());
  const int nq = T.n_q;
  const float* xg = T.ray_x;
  const float* ag = T.ray_a + (size_t)el * nq;
  float qm = E / kHC;
  float x_max = qm * qm;
  qerr = x_max > xg[nq - 1];
  const float dqr = 20.0f / (float)(nq - 1);
  float a_cut = interp_guess(xg, ag, nq, x_max, (int)(qm / dqr));
  const int* gd = T.ray_guide + (size_t)el * kRayGuide;
  const float gs = T.ray_gscale[el];
  for (int t = 0; t < kmax; ++t) {
    float x1 = rng.next(), x2 = rng.next();
    float target = x1 * a_cut;
    // largest j <= nq - 2 with A[j] <= target (np.interp bracket; A[0] = 0): from the guide entry of
    // the target's bucket, which no later bracket precedes, forward over A
    int lo = gd[min((int)(target * gs), kRayGuide - 1)];
    while (lo < nq - 2 && ag[lo + 1] <= target) ++lo;
    float x;
    if (target >= ag[nq - 1]) x = xg[nq - 1];
    else x = xg[lo] + (xg[lo + 1] - xg[lo]) / (ag[lo + 1] - ag[lo]) * (target - ag[lo]);
    x = fminf(x, x_max);
    float c = fminf(fmaxf(1.0f - 2.0f * x / x_max, -1.0f), 1.0f);
    if (x2 <= (1.0f + c * c) * 0.5f) {
      // sin^2 = (1 - c)(1 + c) = 4 x (x_max - x) / x_max^2
      sin_t = 2.0f * sqrtf(fmaxf(x * (x_max - x), 0.0f)) / x_max;
      trials = t + 1;
      return c;
    }
  }
  trials = 0;
  sin_t = 0.0f;
  return 1.0f;
}

// K fluorescence (chatcarlo/physics.py:155-212); slots u0..u3
template <class Rng>
__device__ __forceinline__ bool fluor_sample(const Tables& T, int m, float E, const EPos& p, Rng& rng,
                                             int& el, float& e_line) {
  float u0 = rng.next(), u1 = rng.next(), u2 = rng.next(), u3 = rng.next();
  el = select_elem(T, m, P_PHOTO, p, u0);
  const float* f = T.fl + (size_t)el * kFlStride;
  e_line = 0.0f;
  float edge = f[0];
  if (!(edge > 0.0f) || !(E > edge)) return false;
  // K-shell share, linear in ln E between nodes spaced 1/f[2] from ln(edge) = f[3]
  float t = (logf(E) - f[3]) * f[2];
  int j = min(max((int)t, 0), kKFracNodes - 2);
  float k0 = f[20 + j], k1 = f[21 + j];
  float kf = k0 + (t - (float)j) * (k1 - k0);
  if (!(u1 < kf)) return false;
  if (!(u2 < f[1])) return false;
  int li = 0;
  for (int l = 0; l < 8; ++l) li += (u3 > f[12 + l]) ? 1 : 0;
  li = min(li, 7);
  float ln = f[4 + li];
  if (ln >= kFluorCut) {
    e_line = ln;
    return true;
  }
  return false;
}

// New direction from (cos theta, sin theta) and phi = 2 pi u (chatcarlo/physics.py:354-367).
// Each sampler supplies sin theta in a form free of the cancellation of
// sqrt(1 - cos^2) near theta = 0 and pi.
__device__ __forceinline__ void rotate_dir(float d[3], float cos_t, float sin_t, float u_phi) {
  float st = sin_t;
  float sp, cp;
  sincosf(kTwoPi * u_phi, &sp, &cp);
  float up0, up1, up2;
  if (fabsf(d[2]) < 0.999f) { up0 = 0.f; up1 = 0.f; up2 = 1.f; }
  else { up0 = 1.f; up1 = 0.f; up2 = 0.f; }
  float u0 = up1 * d[2] - up2 * d[1];
  float u1 = up2 * d[0] - up0 * d[2];
  float u2 = up0 * d[1] - up1 * d[0];
  float inv = 1.0f / sqrtf(u0 * u0 + u1 * u1 + u2 * u2);
  u0 *= inv; u1 *= inv; u2 *= inv;
  float v0 = d[1] * u2 - d[2] * u1;
  float v1 = d[2] * u0 - d[0] * u2;
  float v2 = d[0] * u1 - d[1] * u0;
  float a = st * cp, b = st * sp;
  float n0 = a * u0 + b * v0 + cos_t * d[0];
  float n1 = a * u1 + b * v1 + cos_t * d[1];
  float n2 = a * u2 + b * v2 + cos_t * d[2];
  float inv2 = 1.0f / sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
  d[0] = n0 * inv2; d[1] = n1 * inv2; d[2] = n2 * inv2;
}

// chatcarlo/physics.py:215-220: cos = U(-1, 1); sin = 2 sqrt(u (1 - u)) = sqrt(1 - cos^2)
__device__ __forceinline__ void isotropic_dir(float d[3], float u1, float u2) {
  float c = -1.0f + 2.0f * u1;
  float s = 2.0f * sqrtf(fmaxf(u1 * (1.0f - u1), 0.0f));
  float sp, cp;
  sincosf(kTwoPi * u2, &sp, &cp);
  d[0] = s * cp; d[1] = s * sp; d[2] = c;
}

// ---------------------------------------------------------------------------
// track-length tally: exact 3-D DDA (Amanatides & Woo 1987) with FP64 atomics
// ---------------------------------------------------------------------------
// Batch tallies are held in cmask + 1 copies, cstride elements apart: workgroup b
// adds into copy b & cmask.  Workgroups are dealt to the 8 XCDs round-robin, so with
// 8 copies each XCD adds into its own copy and the hot voxels of a beam are not
// contended across XCDs; the batch end sums the copies.
struct GridDev {
  double* t;      // batch tallies, copy 0: kerma [keV] of voxel v at t[2v], H*(10) track integral [pSv cm^3] at t[2v + 1]
  int nx, ny, nz;
  float org[3];
  float h;
  float inv_h;
  uint32_t cmask; // copies - 1 (copies a power of two)
  uint32_t base;  // voxel 0's index in the tally-cache key space (batched scenes: its scene's offset)
  size_t cstride; // elements between copies
};

// FP64 add to device memory.  Every caller passes global memory, and the pointer is cast
// to say so: through a generic pointer the atomic is a flat one, which also counts
// against the LDS wait counter, so the wave's next LDS access would wait for its L2 round trip.
typedef __attribute__((address_space(1))) double global_f64;
__device__ __forceinline__ void atomic_add_f64(double* p, double v) {
  __hip_atomic_fetch_add((global_f64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-private cache of tally contributions in LDS: a direct-mapped table of
// (voxel, kerma, H*(10)) sums.  The photons of one block share the hot voxels of a beam
// across lanes, waves and steps; adds to a voxel that owns its slot cost LDS atomics
// instead of atomics beyond the L2.  A voxel whose slot another
// voxel owns goes straight to memory; the table is added into the grid once, when the
// block ends.  (Probing further slots was measured slower: when the table misses, the
// voxels are not shared anyway.)
constexpr int kCacheSlots = 2560;   // a multiple of 4; 50 KB, so that three blocks fit a CU's LDS (2688 fit two)
constexpr uint32_t kCacheEmpty = 0xffffffffu;
template <int kSlots>
struct TallyCacheT {
  static_assert(kSlots % 4 == 0, "slots come in groups of four");
  static constexpr int kN = kSlots;
  uint32_t key[kSlots];
  double val[kSlots][2];
};
using TallyCache = TallyCacheT<kCacheSlots>;

template <class C>
__device__ __forceinline__ void cache_init(C* c) {
  for (int k = threadIdx.x; k < C::kN; k += blockDim.x) {
    c->key[k] = kCacheEmpty;
    c->val[k][0] = 0.0;
    c->val[k][1] = 0.0;
  }
  __syncthreads();
}

template <class C>
__device__ __forceinline__ bool cache_add(C* c, uint32_t v, double a, double b) {
  // The four voxels whose tally pairs share a 64-byte segment (keys 4k .. 4k + 3) take four
  // consecutive slots, so the block-end flush adds them from neighbouring lanes of one
  // instruction, which the memory system merges into one request; groups are hashed.
  const uint32_t g = (uint32_t)(((uint64_t)((v >> 2) * 2654435761u) * (uint32_t)(C::kN / 4)) >> 32);
  const uint32_t s = (g << 2) | (v & 3u);
  uint32_t cur = c->key[s];
  if (cur == kCacheEmpty) cur = atomicCAS(&c->key[s], kCacheEmpty, v);
  if (cur != kCacheEmpty && cur != v) return false;
  atomicAdd(&c->val[s][0], a);
  atomicAdd(&c->val[s][1], b);
  return true;
}

// Sum of a double over all 64 lanes of the wave; every lane must be active.
__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// The value of the neighbouring lane (lane ^ 1): a DPP quad permutation (1, 0, 3, 2).
__device__ __forceinline__ uint32_t xor1_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ uint64_t xor1_u64(uint64_t x) {
  return (uint64_t)xor1_u32((uint32_t)x) | ((uint64_t)xor1_u32((uint32_t)(x >> 32)) << 32);
}
__device__ __forceinline__ double xor1_f64(double x) { return __longlong_as_double((long long)xor1_u64((uint64_t)__double_as_longlong(x))); }

// Adds a at p[0] and b at p[1] (one voxel's kerma and H*(10)) for every lane with pred, two
// lanes per voxel: in the first instruction each even lane adds its own a and its odd
// neighbour that voxel's b, in the second the odd lanes' voxels likewise.  Both adds of a
// voxel then fall in one 64-byte segment of one instruction, and the memory system merges
// such adds into a single atomic request; beyond the L2, where these atomics are
// performed, the request rate is the bound.  Every lane of the wave must call it.
__device__ __forceinline__ void pair_add(bool pred, double* p, double a, double b) {
  const bool odd = (__lane_id() & 1) != 0;
  const bool qp = xor1_u32(pred ? 1u : 0u) != 0u;
  double* const q = (double*)xor1_u64((uint64_t)p);
  const double qa = xor1_f64(a), qb = xor1_f64(b);
  if (odd ? qp : pred) atomic_add_f64(odd ? q + 1 : p, odd ? qb : a);   // the even lane's voxel
  if (odd ? pred : qp) atomic_add_f64(odd ? p + 1 : q, odd ? b : qa);   // the odd lane's voxel
}

// Adds (ak, ah) to voxel v for every lane with has = true: into the block's LDS tally cache
// when the voxel owns its slot, else to memory, both adds of a voxel in one request
// (pair_add).  Lanes that share a voxel meet in the cache; matching them across the wave
// first (ballots and lane shuffles at every DDA step) was measured 14 % slower once the
// cache was there.  Every lane of the wave must call this (converged control flow).
template <class C = TallyCache>
__device__ __forceinline__ void wave_tally(const GridDev& G, bool has, uint32_t v, double ak, double ah,
                                           C* tc = nullptr) {
  double* const t = G.t + (size_t)(blockIdx.x & G.cmask) * G.cstride;
  const bool direct = has && !(tc && cache_add(tc, G.base + v, ak, ah));
  pair_add(direct, t + 2 * (size_t)v, ak, ah);
}

// Clips the segment o + t d, t in [0, L], to the grid box: [t0, t1]; false if
// they do not overlap.
__device__ __forceinline__ bool dda_clip(const GridDev& G, const float o[3], const float d[3], float L,
                                         float& t0, float& t1) {
  const int n[3] = {G.nx, G.ny, G.nz};
  bool seg = true;
  t0 = 0.0f;
  t1 = L;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float lo = G.org[k], hi = G.org[k] + (float)n[k] * G.h;
    if (fabsf(d[k]) < 1e-30f) {
      if (o[k] < lo || o[k] > hi) seg = false;
    } else {
      float ta = (lo - o[k]) / d[k], tb = (hi - o[k]) / d[k];
      t0 = fmaxf(t0, fminf(ta, tb));
      t1 = fminf(t1, fmaxf(ta, tb));
    }
  }
  return seg && t1 > t0;
}

// Distance along o + t d to the plane of index k on axis a.  The walk and the
// chunk boundaries of dda_balanced use this one expression, so a chunk ends
// exactly where the walk meets that plane.
__device__ __forceinline__ float plane_t(float org_a, float h, float o_a, float inv_da, int k) {
  const float plane = org_a + (float)k * h;
  return (plane - o_a) * inv_da;
}



// Walks o + t d over [t0, t1] voxel by voxel (exact 3-D DDA, Amanatides & Woo
// 1987) for lanes with act = true, scoring wk*len and wh*len into every voxel
// overlapped.  fa >= 0 fixes the start index on axis fa to fidx (a walk that
// starts on a plane of that axis).  All lanes of the wave must call it.
// Returns the number of voxel visits of the calling lane.
template <class C = TallyCache>
__device__ __forceinline__ int dda_walk(const GridDev& G, const float o[3], const float d[3], float t0, float t1,
                                        double wk, double wh, bool act, int fa = -1, int fidx = 0,
                                        C* cache = nullptr) {
  const int n[3] = {G.nx, G.ny, G.nz};
  int idx[3] = {0, 0, 0}, step[3] = {0, 0, 0};
  float inv_d[3] = {0.f, 0.f, 0.f};
  if (act) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      // voxel holding the start point; a start that rounds onto the wrong side
      // of a plane only moves a zero-length piece
      int ik = (k == fa) ? fidx : (int)floorf((o[k] + d[k] * t0 - G.org[k]) * G.inv_h);
      idx[k] = min(max(ik, 0), n[k] - 1);
      step[k] = (d[k] > 0.f) ? 1 : ((d[k] < 0.f) ? -1 : 0);
      // one reciprocal per axis and segment, not a division per voxel
      inv_d[k] = step[k] != 0 ? 1.0f / d[k] : 0.f;
    }
  }
  float tc = t0;
  int visits = 0, it = 0;
  const int cap = G.nx + G.ny + G.nz + 4;
  while (__ballot(act)) {
    bool has = false;
    uint32_t v = 0xffffffffu;
    double ak = 0.0, ah = 0.0;
    if (act) {
      // parametric distance to the next plane on each axis, recomputed from the index
      float tm[3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
        tm[k] = step[k] != 0 ? plane_t(G.org[k], G.h, o[k], inv_d[k], idx[k] + (step[k] > 0 ? 1 : 0)) : INFINITY;
      int a = (tm[0] < tm[1]) ? ((tm[0] < tm[2]) ? 0 : 2) : ((tm[1] < tm[2]) ? 1 : 2);
      float tn = fminf(fmaxf(sel3(a, tm[0], tm[1], tm[2]), tc), t1);
      double len = (double)tn - (double)tc;
      if (len > 0.0) {
        v = (uint32_t)(((size_t)idx[0] * G.ny + idx[1]) * G.nz + idx[2]);
        ak = wk * len;
        ah = wh * len;
        has = true;
        ++visits;
      }
      tc = tn;
      if (tn >= t1 || ++it >= cap) {
        act = false;
      } else {
        // step on axis a (selects, not a run-time index into register arrays)
        bool out = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (k == a) {
            idx[k] += step[k];
            out = idx[k] < 0 || idx[k] >= n[k];
          }
        }
        if (out) act = false;
      }
    }
    wave_tally(G, has, v, ak, ah, cache);
  }
  return visits;
}

// Scores wk*len into the kerma and wh*len into the H*(10) tally of every voxel the segment
// o + t d, t in [0, L], overlaps, for lanes with seg = true; each lane walks
// its own segment.  All lanes of the wave must call it.
template <class C = TallyCache>
__device__ __forceinline__ int dda_score(const GridDev& G, const float o[3], const float d[3], float L,
                                         double wk, double wh, bool seg, C* cache = nullptr) {
  float t0 = 0.f, t1 = 0.f;
  if (seg) seg = dda_clip(G, o, d, L, t0, t1);
  return dda_walk(G, o, d, t0, t1, wk, wh, seg, -1, 0, cache);
}

constexpr int kLanes = 64;       // wavefront size of CDNA
constexpr int kChunkSlabs = 4;   // dominant-axis slabs per chunk

// The tally of dda_score with the work of the wave's segments spread over its
// lanes.  Each segment is cut at the planes of its dominant axis (largest |d|)
// into chunks of kChunkSlabs slabs; between two such planes the other axes
// cross at most one plane each, so a chunk holds at most 3 kChunkSlabs voxel
// pieces.  The wave's chunks are dealt to its 64 lanes in passes, so the lanes
// of one walk have similar trip counts; chunk ends lie on voxel faces, so the
// pieces (and the sums) are those of dda_score.  A lane reads its chunk's segment
// from the owning lane's registers (lane shuffles), so the walk needs no LDS of its
// own and the block's LDS goes to the tally cache.  grids: per-scene grids when
// kScene (key = scene index), else unused.  All lanes of the wave must call it.
template <bool kScene, class C = TallyCache>
__device__ __forceinline__ int dda_balanced(const GridDev& G0, const GridDev* grids, const float o[3], const float d[3],
                            float L, double wk, double wh, bool seg, uint32_t key, C* cache = nullptr,
                            size_t slot_stride = 0) {
  const int lane = __lane_id();
  float t0 = 0.f, t1 = 0.f;
  if (seg) seg = dda_clip(G0, o, d, L, t0, t1);
  const float ad[3] = {fabsf(d[0]), fabsf(d[1]), fabsf(d[2])};
  const int a = (ad[0] >= ad[1]) ? ((ad[0] >= ad[2]) ? 0 : 2) : ((ad[1] >= ad[2]) ? 1 : 2);
  int kf = 0, m = 0;
  uint32_t nch = 0;
  if (seg) {
    const int na = sel3i(a, G0.nx, G0.ny, G0.nz);
    const float oa = sel3(a, o[0], o[1], o[2]), da = sel3(a, d[0], d[1], d[2]);
    const float ga = sel3(a, G0.org[0], G0.org[1], G0.org[2]);
    const float c0 = (oa + da * t0 - ga) * G0.inv_h;
    const float c1 = (oa + da * t1 - ga) * G0.inv_h;
    if (da > 0.f) {
      kf = (int)floorf(c0) + 1;
      m = (int)ceilf(c1) - kf;
    } else {
      kf = (int)ceilf(c0) - 1;
      m = kf - (int)floorf(c1);
    }
    m = min(max(m, 0), na + 1);
    nch = (uint32_t)((m + kChunkSlabs) / kChunkSlabs);
  }
  uint32_t total = 0;
  uint32_t start = 0;
  {
    uint32_t x = nch;
    for (int off = 1; off < kLanes; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    total = __shfl(x, kLanes - 1);
    start = x - nch;
  }
  int visits = 0;
  for (uint32_t base = 0; base < total; base += kLanes) {
    const uint32_t q = base + (uint32_t)lane;
    // owner of chunk q: the last lane whose first chunk is at or before q, by a six-step search
    // that every lane runs whole, so that every shuffle reads an active lane
    int lo = 0, hi = kLanes - 1;
#pragma unroll
    for (int stp = 0; stp < 6; ++stp) {
      const int mid = (lo + hi + 1) >> 1;
      const uint32_t sm = (uint32_t)__shfl((int)start, mid);
      if (lo < hi) {
        if (sm <= q) lo = mid; else hi = mid - 1;
      }
    }
    const int s = lo;
    float so[3], sd[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      so[k] = __shfl(o[k], s);
      sd[k] = __shfl(d[k], s);
    }
    const float s0 = __shfl(t0, s), s1 = __shfl(t1, s);
    const double swk = __shfl(wk, s), swh = __shfl(wh, s);
    const uint32_t skey = (uint32_t)__shfl((int)key, s);
    int sa = __shfl(a, s);
    const int skf = __shfl(kf, s), sm = __shfl(m, s);
    const uint32_t sst = (uint32_t)__shfl((int)start, s);
    bool act = q < total;
    float ta = 0.f, tb = 0.f;
    int fidx = 0;
    if (act) {
      const int j = (int)(q - sst);
      const float da = sel3(sa, sd[0], sd[1], sd[2]), oa = sel3(sa, so[0], so[1], so[2]);
      const int sg = da > 0.f ? 1 : -1;
      const float inv_da = 1.0f / da;
      float ga, gh;
      if (kScene) {
        const GridDev& Gs = grids[skey];
        ga = sel3(sa, Gs.org[0], Gs.org[1], Gs.org[2]);
        gh = Gs.h;
      } else {
        ga = sel3(sa, G0.org[0], G0.org[1], G0.org[2]);
        gh = G0.h;
      }
      const int i0 = j * kChunkSlabs, i1 = i0 + kChunkSlabs;
      ta = i0 == 0 ? s0 : fmaxf(plane_t(ga, gh, oa, inv_da, skf + sg * (i0 - 1)), s0);
      tb = i1 >= sm + 1 ? s1 : fminf(plane_t(ga, gh, oa, inv_da, skf + sg * (i1 - 1)), s1);
      if (i0 == 0) {
        sa = -1;
      } else {
        const int kp = skf + sg * (i0 - 1);
        fidx = sg > 0 ? kp : kp - 1;
      }
      act = tb > ta;
    }
    if (kScene) {
      const GridDev Gs = grids[act ? skey : 0];
      visits += dda_walk(Gs, so, sd, ta, tb, swk, swh, act, sa, fidx);
    } else {
      if (slot_stride) {
        GridDev Gs = G0;
        const size_t off = (size_t)skey * slot_stride;
        Gs.t += 2 * off;
        Gs.base = (uint32_t)off;
        visits += dda_walk(Gs, so, sd, ta, tb, swk, swh, act, sa, fidx, cache);
      } else {
        visits += dda_walk(G0, so, sd, ta, tb, swk, swh, act, sa, fidx, cache);
      }
    }
  }
  return visits;
}

}  // namespace cc
