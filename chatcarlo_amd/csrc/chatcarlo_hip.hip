// Kernels and C-ABI of the gfx950 photon-transport path.
//
// Event-based analog transport (chatcarlo/transport.py:98-312 re-designed for
// one MI355X): a source kernel fills a structure-of-arrays particle pool; each
// round a flight kernel walks every active photon through analytic surfaces
// until it interacts or escapes (scoring the track-length tallies with an
// exact DDA voxel walk and FP64 atomics) and appends it to one of three
// interaction queues; photoelectric(+K fluorescence), bound-Compton and
// Rayleigh kernels then drain their own queue and append survivors to the
// next round's active list.  Queue appends and per-material energy deposits
// are wavefront-aggregated (one atomic per wave and key).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/chatcarlo_hip.h"
#include "cc_device.hpp"

using namespace cc;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(CC_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),      \
                  __FILE__, __LINE__);                                                    \
  } while (0)

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kMaxRounds = 100000;
constexpr int kMaxCrossings = 1 << 14;
constexpr int kTrialCap = 100000;

enum { F_ABSORBED = 1, F_ESCAPED = 2, F_DEAD = 4 };
// device counter slots (uint32)
enum { C_ACT_A = 0, C_ACT_B = 1, C_QP = 2, C_QC = 3, C_QR = 4, C_NCNT = 8 };
// device stats slots (u64)
enum { S_ABS = 0, S_ESC = 1, S_SCAT = 2, S_FLUOR = 3, S_EVT = 4, S_FLIGHT = 5, S_VOX = 6, S_NSTAT = 8 };

struct PState {
  float *x, *y, *z, *dx, *dy, *dz, *e, *tau, *wt;
  uint32_t* ctr;
  int *nsc, *mat, *flags;
};

struct Src {
  int field, n_spec, n_angles, rotation;
  float pos[3], dir[3], u[3], v[3];
  float sid, width, height, cos_half, omch;   // omch = 1 - cos_half
  float iso[3], axis[3];
  float scan;
  const float* spec_e;
  const float* spec_cdf;
};

struct Launch {
  Tables T;
  const Prim* prims;
  World W;
  GridDev G;
  int grid_on, fluor;
  PState S;
  uint32_t* cnt;
  unsigned long long* stats;
  double* edep;
  int* err;
  uint64_t seed, hist0;
};

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_append(bool pred, uint32_t* counter) {
  unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return -1;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  return pred ? (int)(base + __popcll(mask & ((1ull << lane) - 1ull))) : -1;
}

__device__ __forceinline__ void wave_count(bool pred, unsigned long long* counter) {
  unsigned long long mask = __ballot(pred);
  if (mask != 0ull && __lane_id() == __ffsll((long long)mask) - 1)
    atomicAdd(counter, (unsigned long long)__popcll(mask));
}

__device__ __forceinline__ void wave_sum_u64(unsigned long long v, unsigned long long* counter) {
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (__lane_id() == 0 && v) atomicAdd(counter, v);
}

// Sum `val` per distinct key over the wave and add each sum with one FP64 atomic.
__device__ __forceinline__ void wave_add_by_key(bool active, int key, double val, double* base) {
  const int lane = __lane_id();
  unsigned long long pending = __ballot(active);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int k = __shfl(key, leader);
    const bool mine = active && key == k && ((pending >> lane) & 1ull);
    const unsigned long long mm = __ballot(mine);
    double v = mine ? val : 0.0;
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == leader) atomic_add_f64(base + k, v);
    pending &= ~mm;
  }
}

__device__ __forceinline__ void raise_err(int* err, int bits) { atomicOr(err, bits); }

__device__ __forceinline__ void stage_prims(const Prim* g, Prim* s, int n) {
  const int* src = reinterpret_cast<const int*>(g);
  int* dst = reinterpret_cast<int*>(s);
  const int words = n * (int)(sizeof(Prim) / sizeof(int));
  for (int k = threadIdx.x; k < words; k += blockDim.x) dst[k] = src[k];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// source (chatcarlo/source.py:67-210, spectrum.py:82)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rodrigues(const float v[3], const float ax[3], float ca, float sa, float out[3]) {
  float cx = ax[1] * v[2] - ax[2] * v[1];
  float cy = ax[2] * v[0] - ax[0] * v[2];
  float cz = ax[0] * v[1] - ax[1] * v[0];
  float dt = ax[0] * v[0] + ax[1] * v[1] + ax[2] * v[2];
  out[0] = v[0] * ca + cx * sa + ax[0] * dt * (1.f - ca);
  out[1] = v[1] * ca + cy * sa + ax[1] * dt * (1.f - ca);
  out[2] = v[2] * ca + cz * sa + ax[2] * dt * (1.f - ca);
}

// Slots: [angle][shift][a][b] then [energy][tau] from the next block.
template <class Rng>
__device__ __forceinline__ void sample_source(const Src& s, Rng& rng, float p[3], float d[3], float& E) {
  float u0 = rng.next(), u1 = rng.next(), u2 = rng.next(), u3 = rng.next();
  float pos[3] = {s.pos[0], s.pos[1], s.pos[2]};
  float dd[3] = {s.dir[0], s.dir[1], s.dir[2]};
  float uu[3] = {s.u[0], s.u[1], s.u[2]};
  float vv[3] = {s.v[0], s.v[1], s.v[2]};
  if (s.rotation != 0) {
    float ang;
    if (s.rotation == 2) {
      int k = min((int)(u0 * (float)s.n_angles), s.n_angles - 1);
      ang = kTwoPi * (float)k / (float)s.n_angles;
    } else {
      ang = kTwoPi * u0;
    }
    float sa, ca;
    sincosf(ang, &sa, &ca);
    float rel[3] = {pos[0] - s.iso[0], pos[1] - s.iso[1], pos[2] - s.iso[2]};
    float r[3];
    rodrigues(rel, s.axis, ca, sa, r);
    pos[0] = s.iso[0] + r[0]; pos[1] = s.iso[1] + r[1]; pos[2] = s.iso[2] + r[2];
    float t[3];
    rodrigues(dd, s.axis, ca, sa, t); dd[0] = t[0]; dd[1] = t[1]; dd[2] = t[2];
    rodrigues(uu, s.axis, ca, sa, t); uu[0] = t[0]; uu[1] = t[1]; uu[2] = t[2];
    rodrigues(vv, s.axis, ca, sa, t); vv[0] = t[0]; vv[1] = t[1]; vv[2] = t[2];
    if (s.scan > 0.f) {
      float sh = -0.5f * s.scan + s.scan * u1;
      pos[0] += s.axis[0] * sh; pos[1] += s.axis[1] * sh; pos[2] += s.axis[2] * sh;
    }
  }
  if (s.field == CC_FIELD_CONE) {
    // c = cos_half + (1 - cos_half) u; 1 - c = (1 - cos_half)(1 - u) keeps sin accurate near the axis
    float omc = s.omch * (1.f - u2);
    float c = 1.f - omc;
    float sn = sqrtf(fmaxf(omc * (2.f - omc), 0.f));
    float sp, cp;
    sincosf(kTwoPi * u3, &sp, &cp);
    for (int k = 0; k < 3; ++k) d[k] = c * dd[k] + sn * cp * uu[k] + sn * sp * vv[k];
    for (int k = 0; k < 3; ++k) p[k] = pos[k];
  } else if (s.field == CC_FIELD_PARALLEL) {
    float a = -0.5f * s.width + s.width * u2;
    float b = -0.5f * s.height + s.height * u3;
    for (int k = 0; k < 3; ++k) { p[k] = pos[k] + a * uu[k] + b * vv[k]; d[k] = dd[k]; }
  } else {
    float a = -0.5f * s.width + s.width * u2;
    float b = -0.5f * s.height + s.height * u3;
    // target - origin = d sid + a u + b v  (chatcarlo/source.py:191-193)
    float t[3];
    for (int k = 0; k < 3; ++k) t[k] = dd[k] * s.sid + a * uu[k] + b * vv[k];
    float inv = 1.f / sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int k = 0; k < 3; ++k) { d[k] = t[k] * inv; p[k] = pos[k]; }
  }
  // rng.choice: first bin with cdf > U (searchsorted side="right")
  float ue = rng.next();
  int lo = 0, hi = s.n_spec - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (s.spec_cdf[mid] > ue) hi = mid; else lo = mid + 1;
  }
  E = s.spec_e[lo];
}

__global__ __launch_bounds__(kBlock) void k_source(Launch L, Src s, int n, int* act) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  PhiloxRng rng(L.seed, L.hist0 + (uint64_t)i, 0u, 0u);
  float p[3], d[3], E;
  sample_source(s, rng, p, d, E);
  float tau = -logf(rng.next());
  PState S = L.S;
  S.x[i] = p[0]; S.y[i] = p[1]; S.z[i] = p[2];
  S.dx[i] = d[0]; S.dy[i] = d[1]; S.dz[i] = d[2];
  S.e[i] = E; S.tau[i] = tau; S.wt[i] = 1.0f;
  S.ctr[i] = rng.ctr; S.nsc[i] = 0; S.mat[i] = -1; S.flags[i] = 0;
  act[i] = i;
}

// caller-supplied photons: only tau is drawn (chatcarlo/transport.py:124)
__global__ __launch_bounds__(kBlock) void k_init_given(Launch L, int n, int* act) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  PhiloxRng rng(L.seed, L.hist0 + (uint64_t)i, 0u, 0u);
  PState S = L.S;
  S.tau[i] = -logf(rng.next());
  S.wt[i] = 1.0f;
  S.ctr[i] = rng.ctr; S.nsc[i] = 0; S.mat[i] = -1; S.flags[i] = 0;
  act[i] = i;
}

__global__ void k_round_reset(uint32_t* cnt, int next_slot) {
  if (threadIdx.x == 0) {
    cnt[C_QP] = 0; cnt[C_QC] = 0; cnt[C_QR] = 0; cnt[next_slot] = 0;
  }
}

__global__ void k_batch_reset(uint32_t* cnt, unsigned long long* stats, double* edep, int cur_slot, uint32_t n) {
  const int t = threadIdx.x;
  if (t < C_NCNT) cnt[t] = 0;
  if (t < S_NSTAT) stats[t] = 0ull;
  if (t < CC_MAX_MAT) edep[t] = 0.0;
  __syncthreads();
  if (t == 0) cnt[cur_slot] = n;
}

// ---------------------------------------------------------------------------
// flight: boundary walk until interaction or escape
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_flight(Launch L, const int* act, const uint32_t* cnt_in, int* qp,
                                                   int* qc, int* qr) {
  __shared__ Prim sp[CC_MAX_PRIMS];
  stage_prims(L.prims, sp, L.W.n_prim);
  const uint32_t n = *cnt_in;
  const Tables& T = L.T;
  PState S = L.S;
  // Every lane of a wave stays in the segment loop until the whole wave is
  // done (flags instead of divergent exits), so the DDA tally can combine
  // lanes that share a voxel.
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + threadIdx.x;
    const bool valid = j < n;
    int type = -1, i = 0;
    bool escaped = false, alive = false;
    unsigned long long events = 0, visits = 0;
    float p[3] = {0.f, 0.f, 0.f}, d[3] = {0.f, 0.f, 1.f};
    float E = 1.0f, tau = 0.f, w = 0.f, mu = 0.f, muen = 0.f;
    double wh = 0.0;
    int mc = -1, it = 0;
    EPos ep{0, 0.f, 0.f};
    if (valid) {
      i = act[j];
      p[0] = S.x[i]; p[1] = S.y[i]; p[2] = S.z[i];
      d[0] = S.dx[i]; d[1] = S.dy[i]; d[2] = S.dz[i];
      E = S.e[i];
      tau = S.tau[i];
      w = S.wt[i];
      if (!energy_ok(E)) {
        raise_err(L.err, CC_DEV_ENERGY_RANGE);
        S.flags[i] |= F_DEAD;
      } else {
        alive = true;
        ep = energy_pos(T, E);
        if (L.grid_on) wh = (double)w * (double)h10_coeff(T.h10, T.n_h10, E);
      }
    }
    while (__ballot(alive)) {
      float ds = 0.f, taub = 0.f;
      bool inter = false, esc = false, seg = false;
      int m = 0;
      if (alive) {
        ++events;
        const float am = fmaxf(fabsf(p[0]), fmaxf(fabsf(p[1]), fabsf(p[2])));
        const float eps = 1e-6f + 1.2e-7f * am;
        float tb = next_boundary(sp, L.W.n_prim, L.W, p, d, eps, esc);
        if (!(tb < INFINITY) || tb < 0.f) {   // outside the world box
          escaped = true;
          alive = false;
        } else {
          const float mid[3] = {p[0] + d[0] * 0.5f * tb, p[1] + d[1] * 0.5f * tb, p[2] + d[2] * 0.5f * tb};
          m = locate(sp, L.W.n_prim, L.W.background, mid);
          if (m != mc) {
            mc = m;
            const float rho = T.density[m];
            mu = xs_part(T, m, P_TOTAL, E, ep) * rho;
            if (L.grid_on) muen = xs_part(T, m, P_MUEN, E, ep) * rho;
          }
          taub = mu * tb;
          inter = tau < taub;
          ds = inter ? tau / fmaxf(mu, 1e-30f) : tb;
          seg = L.grid_on != 0;
        }
      }
      if (L.grid_on) visits += dda_score(L.G, p, d, ds, (double)w * (double)E * (double)muen, wh, seg);
      if (alive) {
        p[0] += d[0] * ds; p[1] += d[1] * ds; p[2] += d[2] * ds;
        if (inter) {
          PhiloxRng rng(L.seed, L.hist0 + (uint64_t)i, 0u, S.ctr[i]);
          const float r = rng.next();
          S.ctr[i] = rng.ctr;
          const float ph = xs_part(T, m, P_PHOTO, E, ep);
          const float co = xs_part(T, m, P_COMPT, E, ep);
          const float ra = xs_part(T, m, P_RAYL, E, ep);
          const float tot = ph + co + ra;
          const float pp = tot > 0.f ? ph / tot : 0.f;
          const float pc = tot > 0.f ? co / tot : 0.f;
          type = (r < pp) ? 0 : ((r < pp + pc) ? 1 : 2);
          S.mat[i] = m;
          tau = 0.f;
          alive = false;
        } else {
          tau -= taub;
          if (esc) {
            escaped = true;
            alive = false;
          } else if (++it >= kMaxCrossings) {
            raise_err(L.err, CC_DEV_STEP_CAP);
            S.flags[i] |= F_DEAD;
            alive = false;
          }
        }
      }
    }
    if (valid && energy_ok(E)) {
      if (escaped) S.flags[i] |= F_ESCAPED;
      S.x[i] = p[0]; S.y[i] = p[1]; S.z[i] = p[2];
      S.tau[i] = tau;
    }
    wave_count(valid && escaped, L.stats + S_ESC);
    wave_count(valid, L.stats + S_FLIGHT);
    wave_sum_u64(events, L.stats + S_EVT);
    wave_sum_u64(visits, L.stats + S_VOX);
    int slot;
    slot = wave_append(type == 0, L.cnt + C_QP);
    if (slot >= 0) qp[slot] = i;
    slot = wave_append(type == 1, L.cnt + C_QC);
    if (slot >= 0) qc[slot] = i;
    slot = wave_append(type == 2, L.cnt + C_QR);
    if (slot >= 0) qr[slot] = i;
  }
}

// ---------------------------------------------------------------------------
// interaction kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_photo(Launch L, const int* q, const uint32_t* cnt_q, int* act_next,
                                                  uint32_t* cnt_next) {
  const uint32_t n = *cnt_q;
  const Tables& T = L.T;
  PState S = L.S;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + threadIdx.x;
    const bool valid = j < n;
    int i = 0, m = 0;
    bool emit = false;
    double dep = 0.0;
    if (valid) {
      i = q[j];
      m = S.mat[i];
      const float E = S.e[i];
      const float w = S.wt[i];
      PhiloxRng rng(L.seed, L.hist0 + (uint64_t)i, 0u, S.ctr[i]);
      float eline = 0.f;
      if (L.fluor) {
        const EPos ep = energy_pos(T, E);
        int el;
        emit = fluor_sample(T, m, E, ep, rng, el, eline);
      }
      dep = (double)w * (double)(emit ? E - eline : E);
      S.nsc[i] += 1;
      if (emit) {
        float d[3];
        isotropic_dir(d, rng.next(), rng.next());
        S.dx[i] = d[0]; S.dy[i] = d[1]; S.dz[i] = d[2];
        S.e[i] = eline;
        S.tau[i] = -logf(rng.next());
      } else {
        S.flags[i] |= F_ABSORBED;
      }
      S.ctr[i] = rng.ctr;
    }
    wave_add_by_key(valid, m, dep, L.edep);
    wave_count(valid && !emit, L.stats + S_ABS);
    wave_count(valid && emit, L.stats + S_FLUOR);
    wave_count(valid, L.stats + S_SCAT);
    const int slot = wave_append(valid && emit, cnt_next);
    if (slot >= 0) act_next[slot] = i;
  }
}

__global__ __launch_bounds__(kBlock) void k_compton(Launch L, const int* q, const uint32_t* cnt_q, int* act_next,
                                                    uint32_t* cnt_next) {
  const uint32_t n = *cnt_q;
  const Tables& T = L.T;
  PState S = L.S;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + threadIdx.x;
    const bool valid = j < n;
    int i = 0, m = 0;
    double dep = 0.0;
    if (valid) {
      i = q[j];
      m = S.mat[i];
      const float E = S.e[i];
      const float w = S.wt[i];
      PhiloxRng rng(L.seed, L.hist0 + (uint64_t)i, 0u, S.ctr[i]);
      const EPos ep = energy_pos(T, E);
      float cs, sn;
      int el, trials;
      const float eps = compton_sample(T, m, E, ep, rng, kTrialCap, cs, sn, el, trials);
      if (trials == 0) raise_err(L.err, CC_DEV_ITER_CAP);
      const float enew = E * eps;
      dep = (double)w * ((double)E - (double)enew);
      float d[3] = {S.dx[i], S.dy[i], S.dz[i]};
      rotate_dir(d, cs, sn, rng.next());
      S.dx[i] = d[0]; S.dy[i] = d[1]; S.dz[i] = d[2];
      S.e[i] = enew;
      S.tau[i] = -logf(rng.next());
      S.nsc[i] += 1;
      S.ctr[i] = rng.ctr;
    }
    wave_add_by_key(valid, m, dep, L.edep);
    wave_count(valid, L.stats + S_SCAT);
    const int slot = wave_append(valid, cnt_next);
    if (slot >= 0) act_next[slot] = i;
  }
}

__global__ __launch_bounds__(kBlock) void k_rayleigh(Launch L, const int* q, const uint32_t* cnt_q, int* act_next,
                                                     uint32_t* cnt_next) {
  const uint32_t n = *cnt_q;
  const Tables& T = L.T;
  PState S = L.S;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + threadIdx.x;
    const bool valid = j < n;
    int i = 0;
    if (valid) {
      i = q[j];
      const int m = S.mat[i];
      const float E = S.e[i];
      PhiloxRng rng(L.seed, L.hist0 + (uint64_t)i, 0u, S.ctr[i]);
      const EPos ep = energy_pos(T, E);
      int el, trials;
      bool qerr;
      float sn;
      const float cs = rayleigh_sample(T, m, E, ep, rng, kTrialCap, sn, el, trials, qerr);
      if (trials == 0) raise_err(L.err, CC_DEV_ITER_CAP);
      if (qerr) raise_err(L.err, CC_DEV_Q_RANGE);
      float d[3] = {S.dx[i], S.dy[i], S.dz[i]};
      rotate_dir(d, cs, sn, rng.next());
      S.dx[i] = d[0]; S.dy[i] = d[1]; S.dz[i] = d[2];
      S.tau[i] = -logf(rng.next());
      S.nsc[i] += 1;
      S.ctr[i] = rng.ctr;
    }
    wave_count(valid, L.stats + S_SCAT);
    const int slot = wave_append(valid, cnt_next);
    if (slot >= 0) act_next[slot] = i;
  }
}

// ---------------------------------------------------------------------------
// batch end: fold the batch buffers into totals and moments (chatcarlo/tally.py:59-94)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_end_batch(double* dk, double* dh, double* tk, double* th, double* qk,
                                                      double* qh, int* hits, size_t nvox, double inv_nb, int track) {
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x) {
    const double a = dk[v], b = dh[v];
    if (a == 0.0 && b == 0.0) continue;
    tk[v] += a;
    th[v] += b;
    if (track) {
      qk[v] += a * a * inv_nb;
      qh[v] += b * b * inv_nb;
      hits[v] += 1;
    }
    dk[v] = 0.0;
    dh[v] = 0.0;
  }
}

__global__ void k_fill_f64(double* p, size_t n, double v) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}
__global__ void k_fill_i32(int* p, size_t n, int v) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) p[k] = v;
}
__global__ void k_add_f64(double* dst, const double* src, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) dst[k] += src[k];
}
__global__ void k_add_i32(int* dst, const int* src, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x) dst[k] += src[k];
}

// ---------------------------------------------------------------------------
// test kernels: production device functions fed caller inputs
// ---------------------------------------------------------------------------
__global__ void k_test_boundary(const Prim* prims, World W, int n, const float* o, const float* d, float eps,
                                float* t, int* esc) {
  __shared__ Prim sp[CC_MAX_PRIMS];
  stage_prims(prims, sp, W.n_prim);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float oo[3] = {o[3 * i], o[3 * i + 1], o[3 * i + 2]};
  float dd[3] = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
  bool e;
  t[i] = next_boundary(sp, W.n_prim, W, oo, dd, eps, e);
  esc[i] = e ? 1 : 0;
}

__global__ void k_locate(const Prim* prims, World W, int n, const float* pts, int* mat) {
  __shared__ Prim sp[CC_MAX_PRIMS];
  stage_prims(prims, sp, W.n_prim);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float q[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  mat[i] = locate(sp, W.n_prim, W.background, q);
}

__global__ void k_test_xs(Tables T, int n, const int* mat, const float* e, float* out, int* err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float E = e[i];
  const int m = mat[i];
  if (!energy_ok(E)) { raise_err(err, CC_DEV_ENERGY_RANGE); return; }
  const EPos ep = energy_pos(T, E);
  const float rho = T.density[m];
  const float ph = xs_part(T, m, P_PHOTO, E, ep), co = xs_part(T, m, P_COMPT, E, ep),
              ra = xs_part(T, m, P_RAYL, E, ep);
  const float tot = ph + co + ra;
  float* o = out + 9 * (size_t)i;
  o[0] = tot * rho;
  o[1] = ph / tot;
  o[2] = co / tot;
  o[3] = ph; o[4] = co; o[5] = ra;
  o[6] = xs_part(T, m, P_MUEN, E, ep) * rho;
  o[7] = xs_part(T, m, P_TOTAL, E, ep) * rho;
  o[8] = h10_coeff(T.h10, T.n_h10, E);
}

__global__ void k_test_compton(Tables T, int n, const int* mat, const float* e, const float* u, int ku, int kmax,
                               float* eps, float* cs, int* el, int* tr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  BufRng rng(u + (size_t)i * ku, ku);
  const EPos ep = energy_pos(T, e[i]);
  float c, sn;
  int z, t;
  eps[i] = compton_sample(T, mat[i], e[i], ep, rng, kmax, c, sn, z, t);
  cs[i] = c; el[i] = z; tr[i] = t;
}

__global__ void k_test_rayleigh(Tables T, int n, const int* mat, const float* e, const float* u, int ku, int kmax,
                                float* cs, int* el, int* tr, int* err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  BufRng rng(u + (size_t)i * ku, ku);
  const EPos ep = energy_pos(T, e[i]);
  int z, t;
  bool qerr;
  float sn;
  cs[i] = rayleigh_sample(T, mat[i], e[i], ep, rng, kmax, sn, z, t, qerr);
  if (qerr) raise_err(err, CC_DEV_Q_RANGE);
  el[i] = z; tr[i] = t;
}

__global__ void k_test_fluor(Tables T, int n, const int* mat, const float* e, const float* u, int* el, int* emit,
                             float* eline) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  BufRng rng(u + (size_t)i * 4, 4);
  const EPos ep = energy_pos(T, e[i]);
  int z;
  float l;
  emit[i] = fluor_sample(T, mat[i], e[i], ep, rng, z, l) ? 1 : 0;
  el[i] = z; eline[i] = l;
}

__global__ void k_test_rotate(int n, const float* d, const float* cs, const float* up, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float dd[3] = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
  const float c = cs[i];
  rotate_dir(dd, c, sqrtf(fmaxf((1.0f - c) * (1.0f + c), 0.0f)), up[i]);
  out[3 * i] = dd[0]; out[3 * i + 1] = dd[1]; out[3 * i + 2] = dd[2];
}

__global__ void k_test_iso(int n, const float* u1, const float* u2, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float dd[3];
  isotropic_dir(dd, u1[i], u2[i]);
  out[3 * i] = dd[0]; out[3 * i + 1] = dd[1]; out[3 * i + 2] = dd[2];
}

__global__ void k_test_dda(GridDev G, int n, const float* o, const float* d, const float* len, const double* wk,
                           const double* wh) {
  // every lane takes part in the wave-aggregated tally; lanes past n score nothing
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  float oo[3] = {0.f, 0.f, 0.f}, dd[3] = {0.f, 0.f, 1.f};
  float L = 0.f;
  double a = 0.0, b = 0.0;
  if (valid) {
    oo[0] = o[3 * i]; oo[1] = o[3 * i + 1]; oo[2] = o[3 * i + 2];
    dd[0] = d[3 * i]; dd[1] = d[3 * i + 1]; dd[2] = d[3 * i + 2];
    L = len[i];
    a = wk[i];
    b = wh[i];
  }
  dda_score(G, oo, dd, L, a, b, valid);
}

__global__ void k_test_philox(int n, const uint32_t* c, const uint32_t* k, uint32_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint4 r = philox4x32_10(make_uint4(c[4 * i], c[4 * i + 1], c[4 * i + 2], c[4 * i + 3]),
                          make_uint2(k[2 * i], k[2 * i + 1]));
  out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
}

__global__ void k_test_source(Src s, uint64_t seed, uint64_t hist0, int n, float* pos, float* dir, float* en) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  PhiloxRng rng(seed, hist0 + (uint64_t)i, 0u, 0u);
  float p[3], d[3], E;
  sample_source(s, rng, p, d, E);
  for (int k = 0; k < 3; ++k) { pos[3 * i + k] = p[k]; dir[3 * i + k] = d[k]; }
  en[i] = E;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <class T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(CC_ERR_HIP, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  return CC_OK;
}

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // tables
  bool have_tables = false;
  Tables T{};
  std::vector<void*> tbufs;
  // scene
  bool have_scene = false;
  Prim* prims = nullptr;
  World W{};
  // grid
  bool have_grid = false;
  GridDev G{};
  double *tk = nullptr, *th = nullptr, *qk = nullptr, *qh = nullptr;
  int* hits = nullptr;
  size_t nvox = 0;
  bool track = false;
  int64_t g_batches = 0, g_hist = 0;
  // particle pool
  size_t cap = 0;
  PState S{};
  int *act0 = nullptr, *act1 = nullptr, *qp = nullptr, *qc = nullptr, *qr = nullptr;
  uint32_t* cnt = nullptr;
  unsigned long long* stats = nullptr;
  double* edep = nullptr;
  int* err = nullptr;
  uint32_t* h_cnt = nullptr;   // pinned
  // source scratch
  float* spec = nullptr;
  int spec_cap = 0;
  std::vector<hipEvent_t> ev;
};

void free_pool(Ctx* c) {
  void* ps[] = {c->S.x, c->S.y, c->S.z, c->S.dx, c->S.dy, c->S.dz, c->S.e, c->S.tau, c->S.wt,
                c->S.ctr, c->S.nsc, c->S.mat, c->S.flags, c->act0, c->act1, c->qp, c->qc, c->qr};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  c->S = PState{};
  c->act0 = c->act1 = c->qp = c->qc = c->qr = nullptr;
  c->cap = 0;
}

int ensure_pool(Ctx* c, size_t n) {
  if (n <= c->cap) return CC_OK;
  free_pool(c);
  int r;
  if ((r = dalloc(&c->S.x, n)) || (r = dalloc(&c->S.y, n)) || (r = dalloc(&c->S.z, n)) ||
      (r = dalloc(&c->S.dx, n)) || (r = dalloc(&c->S.dy, n)) || (r = dalloc(&c->S.dz, n)) ||
      (r = dalloc(&c->S.e, n)) || (r = dalloc(&c->S.tau, n)) || (r = dalloc(&c->S.wt, n)) ||
      (r = dalloc(&c->S.ctr, n)) || (r = dalloc(&c->S.nsc, n)) || (r = dalloc(&c->S.mat, n)) ||
      (r = dalloc(&c->S.flags, n)) || (r = dalloc(&c->act0, n)) || (r = dalloc(&c->act1, n)) ||
      (r = dalloc(&c->qp, n)) || (r = dalloc(&c->qc, n)) || (r = dalloc(&c->qr, n)))
    return r;
  c->cap = n;
  return CC_OK;
}

template <class T>
int upload(Ctx* c, const T* host, size_t n, T** dev) {
  int r = dalloc(dev, n);
  if (r) return r;
  c->tbufs.push_back((void*)*dev);
  if (n && host) HIP_TRY(hipMemcpy(*dev, host, n * sizeof(T), hipMemcpyHostToDevice));
  return CC_OK;
}

Launch make_launch(Ctx* c, uint64_t seed, uint64_t hist0, int grid_on, int fluor) {
  Launch L;
  L.T = c->T;
  L.prims = c->prims;
  L.W = c->W;
  L.G = c->G;
  L.grid_on = grid_on;
  L.fluor = fluor;
  L.S = c->S;
  L.cnt = c->cnt;
  L.stats = c->stats;
  L.edep = c->edep;
  L.err = c->err;
  L.seed = seed;
  L.hist0 = hist0;
  return L;
}

int grid_blocks(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (int)b;
}

struct Timer {
  Ctx* c;
  bool on;
  std::vector<std::pair<int, int>> marks;   // (kind, event index of start)
  int used = 0;
  hipEvent_t next() {
    if (used >= (int)c->ev.size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      c->ev.push_back(e);
    }
    return c->ev[used++];
  }
  int begin(int kind) {
    if (!on) return -1;
    int idx = used;
    hipEvent_t e = next();
    (void)hipEventRecord(e, c->stream);
    marks.push_back({kind, idx});
    return idx;
  }
  void end() {
    if (!on) return;
    hipEvent_t e = next();
    (void)hipEventRecord(e, c->stream);
  }
  // after a stream sync
  void collect(double* ms, uint64_t* launches) {
    if (!on) return;
    for (auto& m : marks) {
      float t = 0.f;
      (void)hipEventElapsedTime(&t, c->ev[m.second], c->ev[m.second + 1]);
      ms[m.first] += t;
      if (launches) launches[m.first] += 1;
    }
    marks.clear();
    used = 0;
  }
};

enum { K_SOURCE = 0, K_FLIGHT = 1, K_PHOTO = 2, K_COMPTON = 3, K_RAYLEIGH = 4, K_END = 5, K_OTHER = 6 };

// One batch of photons already in the pool [0, n): rounds until no photon is active.
// Rounds are issued in chunks without a host wait in between; the active count
// of every round is copied to pinned memory and inspected once per chunk.  The
// active population never grows, so the count at the start of a chunk bounds
// every grid of the chunk; rounds that find an empty list exit at once.
constexpr int kRoundChunk = 6;
int run_rounds(Ctx* c, const Launch& L, uint32_t n, Timer& tm, uint32_t* rounds_out) {
  int cur = C_ACT_A, nxt = C_ACT_B;
  int* act_cur = c->act0;
  int* act_nxt = c->act1;
  uint32_t active = n;
  uint32_t rounds = 0;
  while (active > 0) {
    if (rounds >= (uint32_t)kMaxRounds) return fail(CC_ERR_DEVICE, "round cap hit");
    const int nb = grid_blocks(active);
    for (int r = 0; r < kRoundChunk; ++r) {
      hipLaunchKernelGGL(k_round_reset, dim3(1), dim3(64), 0, c->stream, c->cnt, nxt);
      tm.begin(K_FLIGHT);
      hipLaunchKernelGGL(k_flight, dim3(nb), dim3(kBlock), 0, c->stream, L, act_cur, c->cnt + cur, c->qp, c->qc,
                         c->qr);
      tm.end();
      tm.begin(K_PHOTO);
      hipLaunchKernelGGL(k_photo, dim3(nb), dim3(kBlock), 0, c->stream, L, c->qp, c->cnt + C_QP, act_nxt,
                         c->cnt + nxt);
      tm.end();
      tm.begin(K_COMPTON);
      hipLaunchKernelGGL(k_compton, dim3(nb), dim3(kBlock), 0, c->stream, L, c->qc, c->cnt + C_QC, act_nxt,
                         c->cnt + nxt);
      tm.end();
      tm.begin(K_RAYLEIGH);
      hipLaunchKernelGGL(k_rayleigh, dim3(nb), dim3(kBlock), 0, c->stream, L, c->qr, c->cnt + C_QR, act_nxt,
                         c->cnt + nxt);
      tm.end();
      HIP_TRY(hipMemcpyAsync(c->h_cnt + r, c->cnt + nxt, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
      std::swap(cur, nxt);
      std::swap(act_cur, act_nxt);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int r = 0; r < kRoundChunk; ++r) {
      if (c->h_cnt[r] == 0) break;
      ++rounds;
    }
    active = c->h_cnt[kRoundChunk - 1];
  }
  if (rounds_out) *rounds_out = rounds + 1;
  return CC_OK;
}

int read_batch(Ctx* c, uint32_t nb, cc_run_out* out, int* err_bits) {
  unsigned long long st[S_NSTAT];
  double ed[CC_MAX_MAT];
  int e = 0;
  HIP_TRY(hipMemcpyAsync(st, c->stats, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(ed, c->edep, sizeof(ed), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&e, c->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  out->n_absorbed += st[S_ABS];
  out->n_escaped += st[S_ESC];
  out->scatter_sum += st[S_SCAT];
  out->n_fluorescence += st[S_FLUOR];
  out->n_events += st[S_EVT];
  out->n_flights += st[S_FLIGHT];
  out->n_voxel_visits += st[S_VOX];
  for (int m = 0; m < CC_MAX_MAT; ++m) {
    out->edep_keV[m] += ed[m];
    if (nb > 0) out->edep_sum2[m] += ed[m] * ed[m] / (double)nb;
  }
  *err_bits |= e;
  return CC_OK;
}

// ---- test and query entry points ------------------------------------------
template <class T>
int to_dev(const T* h, size_t n, T** d) {
  int r = dalloc(d, n);
  if (r) return r;
  if (n) HIP_TRY(hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return CC_OK;
}
template <class T>
int to_host(T* h, const T* d, size_t n) {
  if (n) HIP_TRY(hipMemcpy(h, d, n * sizeof(T), hipMemcpyDeviceToHost));
  return CC_OK;
}

struct DevBufs {
  std::vector<void*> v;
  ~DevBufs() {
    for (void* p : v) (void)hipFree(p);
  }
  template <class T>
  int in(const T* h, size_t n, T** d) {
    int r = to_dev(h, n, d);
    if (!r) v.push_back((void*)*d);
    return r;
  }
  template <class T>
  int out(size_t n, T** d) {
    int r = dalloc(d, n);
    if (!r) v.push_back((void*)*d);
    return r;
  }
};

int make_src_impl(Ctx* c, const cc_source* s, Src* d) {
  if (s->n_spec <= 0 || !s->spec_e || !s->spec_cdf) return fail(CC_ERR_ARG, "empty spectrum");
  for (int i = 0; i < s->n_spec; ++i)
    if (!(s->spec_e[i] >= 1.0 && s->spec_e[i] <= 150.0))
      return fail(CC_ERR_ARG, "spectrum energy %g keV outside [1, 150] keV", s->spec_e[i]);
  if (s->field < 0 || s->field > 2) return fail(CC_ERR_ARG, "field kind %d", s->field);
  if (s->rotation == 2 && s->n_angles < 2) return fail(CC_ERR_ARG, "n_angles must be >= 2");
  if (2 * s->n_spec > c->spec_cap) {
    if (c->spec) (void)hipFree(c->spec);
    c->spec = nullptr;
    int r = dalloc(&c->spec, 2 * (size_t)s->n_spec);
    if (r) return r;
    c->spec_cap = 2 * s->n_spec;
  }
  std::vector<float> h(2 * (size_t)s->n_spec);
  for (int i = 0; i < s->n_spec; ++i) {
    h[i] = (float)s->spec_e[i];
    h[s->n_spec + i] = (float)s->spec_cdf[i];
  }
  h[2 * s->n_spec - 1] = 1.0f;
  HIP_TRY(hipMemcpy(c->spec, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  d->field = s->field; d->n_spec = s->n_spec; d->n_angles = s->n_angles; d->rotation = s->rotation;
  for (int k = 0; k < 3; ++k) {
    d->pos[k] = (float)s->pos[k]; d->dir[k] = (float)s->dir[k];
    d->u[k] = (float)s->u[k]; d->v[k] = (float)s->v[k];
    d->iso[k] = (float)s->iso[k]; d->axis[k] = (float)s->axis[k];
  }
  d->sid = (float)s->sid; d->width = (float)s->width; d->height = (float)s->height;
  d->cos_half = (float)s->cos_half; d->omch = (float)(1.0 - s->cos_half);
  d->scan = (float)s->scan_length;
  d->spec_e = c->spec;
  d->spec_cdf = c->spec + s->n_spec;
  return CC_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* cc_last_error(void) { return g_err.c_str(); }

int cc_version(int32_t* major, int32_t* minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return CC_OK;
}

int cc_ctx_create(int32_t device, void** out) {
  if (!out) return fail(CC_ERR_ARG, "null ctx pointer");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(CC_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  Ctx* c = new Ctx();
  c->device = device;
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  int r;
  if ((r = dalloc(&c->cnt, C_NCNT)) || (r = dalloc(&c->stats, S_NSTAT)) || (r = dalloc(&c->edep, CC_MAX_MAT)) ||
      (r = dalloc(&c->err, 1))) {
    delete c;
    return r;
  }
  HIP_TRY(hipHostMalloc((void**)&c->h_cnt, sizeof(uint32_t) * 16, hipHostMallocDefault));
  HIP_TRY(hipMemset(c->err, 0, sizeof(int)));
  *out = c;
  return CC_OK;
}

int cc_ctx_destroy(void* ctx) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return CC_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_pool(c);
  for (void* p : c->tbufs) (void)hipFree(p);
  void* ps[] = {c->prims, c->tk, c->th, c->qk, c->qh, c->hits, c->G.dk, c->G.dh, c->cnt, c->stats, c->edep, c->err, c->spec};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  for (auto e : c->ev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CC_OK;
}

int cc_upload_tables(void* ctx, const cc_tables* t) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !t) return fail(CC_ERR_ARG, "null argument");
  if (t->n_mat <= 0 || t->n_mat > CC_MAX_MAT) return fail(CC_ERR_ARG, "n_mat %d out of range", t->n_mat);
  if (t->max_edges != kMaxEdges) return fail(CC_ERR_ARG, "max_edges must be %d", kMaxEdges);
  HIP_TRY(hipSetDevice(c->device));
  for (void* p : c->tbufs) (void)hipFree(p);
  c->tbufs.clear();
  const size_t nm = t->n_mat, ne = t->n_e, nel = t->n_el, nq = t->n_q;
  Tables T{};
  T.n_mat = t->n_mat; T.n_e = t->n_e; T.n_el = t->n_el; T.n_q = t->n_q; T.n_h10 = t->n_h10;
  // interleave (sig, slope) per node; slope of the last node is 0
  std::vector<float2> ss(nm * kNParts * ne);
  for (size_t m = 0; m < nm; ++m)
    for (size_t p = 0; p < (size_t)kNParts; ++p)
      for (size_t i = 0; i < ne; ++i) {
        const size_t r = (m * kNParts + p);
        ss[r * ne + i] = make_float2(t->sig[r * ne + i], i + 1 < ne ? t->slope[r * (ne - 1) + i] : 0.f);
      }
  float* f;
  int* k;
  float2* f2;
  int r;
  if ((r = upload(c, t->density, nm, &f))) return r; T.density = f;
  if ((r = upload(c, t->e_node, ne, &f))) return r; T.e_node = f;
  if ((r = upload(c, t->dl_node, ne - 1, &f))) return r; T.dl_node = f;
  if ((r = upload(c, ss.data(), ss.size(), &f2))) return r; T.ss = f2;
  if ((r = upload(c, t->edge_of_interval, nm * ne, &k))) return r; T.edge_idx = k;
  if ((r = upload(c, t->edge_e, nm * kMaxEdges, &f))) return r; T.edge_e = f;
  if ((r = upload(c, t->edge_dl, nm * kMaxEdges * 2, &f))) return r; T.edge_dl = f;
  if ((r = upload(c, t->edge_sig, nm * kMaxEdges * 6, &f))) return r; T.edge_sig = f;
  if ((r = upload(c, t->edge_slope, nm * kMaxEdges * 6, &f))) return r; T.edge_slope = f;
  if ((r = upload(c, t->mat_elem_off, nm + 1, &k))) return r; T.mat_elem_off = k;
  if ((r = upload(c, t->mat_elem, (size_t)t->elem_len, &k))) return r; T.mat_elem = k;
  if ((r = upload(c, t->sel_off, nm, &k))) return r; T.sel_off = k;
  if ((r = upload(c, t->sel, (size_t)t->sel_len, &f))) return r; T.sel = f;
  if ((r = upload(c, t->sq_over_z, nel * nq, &f))) return r; T.sq_over_z = f;
  if ((r = upload(c, t->q_s, nq, &f))) return r; T.q_s = f;
  if ((r = upload(c, t->ray_x, nq, &f))) return r; T.ray_x = f;
  if ((r = upload(c, t->ray_a, nel * nq, &f))) return r; T.ray_a = f;
  if ((r = upload(c, t->fl, nel * kFlStride, &f))) return r; T.fl = f;
  if ((r = upload(c, t->h10, (size_t)4 * t->n_h10, &f))) return r; T.h10 = f;
  c->T = T;
  c->have_tables = true;
  return CC_OK;
}

int cc_upload_scene(void* ctx, const cc_prim* prims, int32_t n_prim, const cc_world* world) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !world || (n_prim > 0 && !prims)) return fail(CC_ERR_ARG, "null argument");
  if (n_prim < 0 || n_prim > CC_MAX_PRIMS) return fail(CC_ERR_ARG, "n_prim %d out of range", n_prim);
  if (!c->have_tables) return fail(CC_ERR_STATE, "upload tables before the scene");
  for (int i = 0; i < n_prim; ++i) {
    if (prims[i].material < 0 || prims[i].material >= c->T.n_mat)
      return fail(CC_ERR_ARG, "primitive %d material id %d out of range", i, prims[i].material);
    if (prims[i].kind < 0 || prims[i].kind > 2) return fail(CC_ERR_ARG, "primitive %d kind %d", i, prims[i].kind);
    if (prims[i].axis < 0 || prims[i].axis > 2) return fail(CC_ERR_ARG, "primitive %d axis %d", i, prims[i].axis);
  }
  if (world->background < 0 || world->background >= c->T.n_mat) return fail(CC_ERR_ARG, "background id out of range");
  HIP_TRY(hipSetDevice(c->device));
  if (!c->prims) {
    int r = dalloc(&c->prims, CC_MAX_PRIMS);
    if (r) return r;
  }
  if (n_prim) HIP_TRY(hipMemcpy(c->prims, prims, sizeof(Prim) * n_prim, hipMemcpyHostToDevice));
  std::memcpy(&c->W, world, sizeof(World));
  c->W.n_prim = n_prim;
  c->have_scene = true;
  return CC_OK;
}

int cc_bind_grid(void* ctx, const double* origin, const int32_t* shape, double voxel_cm, int32_t flags) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !origin || !shape) return fail(CC_ERR_ARG, "null argument");
  if (shape[0] <= 0 || shape[1] <= 0 || shape[2] <= 0 || !(voxel_cm > 0))
    return fail(CC_ERR_ARG, "invalid grid shape or voxel size");
  HIP_TRY(hipSetDevice(c->device));
  cc_unbind_grid(ctx);
  const size_t nv = (size_t)shape[0] * shape[1] * shape[2];
  int r;
  if ((r = dalloc(&c->G.dk, nv)) || (r = dalloc(&c->G.dh, nv)) || (r = dalloc(&c->tk, nv)) || (r = dalloc(&c->th, nv)) ||
      (r = dalloc(&c->qk, nv)) || (r = dalloc(&c->qh, nv)) || (r = dalloc(&c->hits, nv)))
    return r;
  c->nvox = nv;
  c->G.nx = shape[0]; c->G.ny = shape[1]; c->G.nz = shape[2];
  for (int k = 0; k < 3; ++k) c->G.org[k] = (float)origin[k];
  c->G.h = (float)voxel_cm;
  c->G.inv_h = (float)(1.0 / voxel_cm);
  c->track = (flags & 1) != 0;
  c->have_grid = true;
  return cc_reset_grid(ctx);
}

int cc_unbind_grid(void* ctx) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return fail(CC_ERR_ARG, "null ctx");
  void* ps[] = {c->G.dk, c->G.dh, c->tk, c->th, c->qk, c->qh, c->hits};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  c->G = GridDev{};
  c->tk = c->th = c->qk = c->qh = nullptr;
  c->hits = nullptr;
  c->nvox = 0;
  c->have_grid = false;
  return CC_OK;
}

int cc_reset_grid(void* ctx) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !c->have_grid) return fail(CC_ERR_STATE, "no grid bound");
  const int nb = grid_blocks(c->nvox);
  double* ds[] = {c->G.dk, c->G.dh, c->tk, c->th, c->qk, c->qh};
  for (double* d : ds) hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(kBlock), 0, c->stream, d, c->nvox, 0.0);
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(kBlock), 0, c->stream, c->hits, c->nvox, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->g_batches = 0;
  c->g_hist = 0;
  return CC_OK;
}

int cc_read_grid(void* ctx, double* kerma, double* h10, double* ksum2, double* hsum2, int32_t* hits,
                 int64_t* n_batches, int64_t* n_hist) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !c->have_grid) return fail(CC_ERR_STATE, "no grid bound");
  HIP_TRY(hipSetDevice(c->device));
  const size_t b = c->nvox * sizeof(double);
  if (kerma) HIP_TRY(hipMemcpyAsync(kerma, c->tk, b, hipMemcpyDeviceToHost, c->stream));
  if (h10) HIP_TRY(hipMemcpyAsync(h10, c->th, b, hipMemcpyDeviceToHost, c->stream));
  if (ksum2) HIP_TRY(hipMemcpyAsync(ksum2, c->qk, b, hipMemcpyDeviceToHost, c->stream));
  if (hsum2) HIP_TRY(hipMemcpyAsync(hsum2, c->qh, b, hipMemcpyDeviceToHost, c->stream));
  if (hits) HIP_TRY(hipMemcpyAsync(hits, c->hits, c->nvox * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (n_batches) *n_batches = c->g_batches;
  if (n_hist) *n_hist = c->g_hist;
  return CC_OK;
}

int cc_load_grid(void* ctx, const double* kerma, const double* h10, const double* ksum2, const double* hsum2,
                 const int32_t* hits, int64_t n_batches, int64_t n_hist) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !c->have_grid) return fail(CC_ERR_STATE, "no grid bound");
  HIP_TRY(hipSetDevice(c->device));
  double* tmp;
  int r = dalloc(&tmp, c->nvox);
  if (r) return r;
  const int nb = grid_blocks(c->nvox);
  const double* hs[] = {kerma, h10, ksum2, hsum2};
  double* ds[] = {c->tk, c->th, c->qk, c->qh};
  for (int k = 0; k < 4; ++k) {
    if (!hs[k]) continue;
    HIP_TRY(hipMemcpyAsync(tmp, hs[k], c->nvox * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_add_f64, dim3(nb), dim3(kBlock), 0, c->stream, ds[k], tmp, c->nvox);
  }
  if (hits) {
    HIP_TRY(hipMemcpyAsync(tmp, hits, c->nvox * sizeof(int), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_add_i32, dim3(nb), dim3(kBlock), 0, c->stream, c->hits, (const int*)tmp, c->nvox);
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(tmp);
  c->g_batches += n_batches;
  c->g_hist += n_hist;
  return CC_OK;
}

int cc_run(void* ctx, const cc_source* src, uint64_t n_hist, uint64_t seed, uint64_t first, uint32_t batch,
           uint32_t flags, cc_run_out* out) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !src || !out) return fail(CC_ERR_ARG, "null argument");
  if (!c->have_tables || !c->have_scene) return fail(CC_ERR_STATE, "tables and scene must be uploaded");
  const bool grid_on = (flags & CC_RUN_GRID) != 0;
  if (grid_on && !c->have_grid) return fail(CC_ERR_STATE, "CC_RUN_GRID without a bound grid");
  if (batch == 0) return fail(CC_ERR_ARG, "batch must be > 0");
  HIP_TRY(hipSetDevice(c->device));
  std::memset(out, 0, sizeof(*out));
  Src s;
  int r = make_src_impl(c, src, &s);
  if (r) return r;
  r = ensure_pool(c, batch);
  if (r) return r;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->err, (size_t)1, 0);
  Timer tm{c, (flags & CC_RUN_TIMING) != 0};
  const bool track = c->track && (flags & CC_RUN_TRACK_UNCERTAINTY);
  int err_bits = 0;
  uint64_t remaining = n_hist, h0 = first;
  while (remaining > 0) {
    const uint32_t nb = (uint32_t)(remaining < batch ? remaining : batch);
    Launch L = make_launch(c, seed, h0, grid_on ? 1 : 0, (flags & CC_RUN_FLUORESCENCE) ? 1 : 0);
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, c->stream, c->cnt, c->stats, c->edep, (int)C_ACT_A, nb);
    tm.begin(K_SOURCE);
    hipLaunchKernelGGL(k_source, dim3((nb + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, L, s, (int)nb, c->act0);
    tm.end();
    HIP_TRY(hipGetLastError());
    uint32_t rounds = 0;
    r = run_rounds(c, L, nb, tm, &rounds);
    if (r) return r;
    if (rounds > out->rounds_max) out->rounds_max = rounds;
    if (grid_on) {
      tm.begin(K_END);
      hipLaunchKernelGGL(k_end_batch, dim3(grid_blocks(c->nvox)), dim3(kBlock), 0, c->stream, c->G.dk, c->G.dh, c->tk,
                         c->th, c->qk, c->qh, c->hits, c->nvox, 1.0 / (double)nb, track ? 1 : 0);
      tm.end();
      HIP_TRY(hipGetLastError());
      if (track) c->g_batches += 1;
      c->g_hist += nb;
    }
    r = read_batch(c, nb, out, &err_bits);
    if (r) return r;
    tm.collect(out->kernel_ms, out->kernel_launches);
    out->n_batches += 1;
    remaining -= nb;
    h0 += nb;
  }
  out->n_histories = n_hist;
  out->error_bits = (uint32_t)err_bits;
  if (err_bits) return fail(CC_ERR_DEVICE, "device error bits 0x%x", err_bits);
  return CC_OK;
}

int cc_transport(void* ctx, int64_t n, double* pos, double* dir, double* energy, uint64_t seed, uint64_t first,
                 uint32_t flags, int32_t* n_scatter, uint8_t* absorbed, uint8_t* escaped, cc_run_out* out) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !pos || !dir || !energy || !out) return fail(CC_ERR_ARG, "null argument");
  if (n < 0 || n > (1ll << 30)) return fail(CC_ERR_ARG, "n out of range");
  if (!c->have_tables || !c->have_scene) return fail(CC_ERR_STATE, "tables and scene must be uploaded");
  const bool grid_on = (flags & CC_RUN_GRID) != 0;
  if (grid_on && !c->have_grid) return fail(CC_ERR_STATE, "CC_RUN_GRID without a bound grid");
  for (int64_t i = 0; i < n; ++i)
    if (!(energy[i] >= 1.0 && energy[i] <= 150.0))
      return fail(CC_ERR_ARG, "energy %g keV outside the cross-section table range [1, 150] keV", energy[i]);
  HIP_TRY(hipSetDevice(c->device));
  std::memset(out, 0, sizeof(*out));
  if (n == 0) return CC_OK;
  int r = ensure_pool(c, (size_t)n);
  if (r) return r;
  std::vector<float> h((size_t)n);
  float* dsts[7] = {c->S.x, c->S.y, c->S.z, c->S.dx, c->S.dy, c->S.dz, c->S.e};
  for (int k = 0; k < 7; ++k) {
    for (int64_t i = 0; i < n; ++i) h[i] = (float)(k < 3 ? pos[3 * i + k] : (k < 6 ? dir[3 * i + k - 3] : energy[i]));
    HIP_TRY(hipMemcpy(dsts[k], h.data(), n * sizeof(float), hipMemcpyHostToDevice));
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->err, (size_t)1, 0);
  Timer tm{c, false};
  Launch L = make_launch(c, seed, first, grid_on ? 1 : 0, (flags & CC_RUN_FLUORESCENCE) ? 1 : 0);
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, c->stream, c->cnt, c->stats, c->edep, (int)C_ACT_A, (uint32_t)n);
  hipLaunchKernelGGL(k_init_given, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, L, (int)n, c->act0);
  HIP_TRY(hipGetLastError());
  uint32_t rounds = 0;
  r = run_rounds(c, L, (uint32_t)n, tm, &rounds);
  if (r) return r;
  out->rounds_max = rounds;
  if (grid_on) {
    hipLaunchKernelGGL(k_end_batch, dim3(grid_blocks(c->nvox)), dim3(kBlock), 0, c->stream, c->G.dk, c->G.dh, c->tk,
                       c->th, c->qk, c->qh, c->hits, c->nvox, 1.0 / (double)n, 0);
    HIP_TRY(hipGetLastError());
  }
  int err_bits = 0;
  r = read_batch(c, (uint32_t)n, out, &err_bits);
  if (r) return r;
  for (int k = 0; k < 7; ++k) {
    HIP_TRY(hipMemcpy(h.data(), dsts[k], n * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) {
      if (k < 3) pos[3 * i + k] = h[i];
      else if (k < 6) dir[3 * i + k - 3] = h[i];
      else energy[i] = h[i];
    }
  }
  std::vector<int> iv((size_t)n), fl((size_t)n);
  HIP_TRY(hipMemcpy(iv.data(), c->S.nsc, n * sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(fl.data(), c->S.flags, n * sizeof(int), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; ++i) {
    if (n_scatter) n_scatter[i] = iv[i];
    if (absorbed) absorbed[i] = (fl[i] & F_ABSORBED) ? 1 : 0;
    if (escaped) escaped[i] = (fl[i] & F_ESCAPED) ? 1 : 0;
  }
  out->n_histories = (uint64_t)n;
  out->n_batches = 1;
  out->error_bits = (uint32_t)err_bits;
  if (err_bits) return fail(CC_ERR_DEVICE, "device error bits 0x%x", err_bits);
  return CC_OK;
}

int cc_locate(void* ctx, int64_t n, const float* pts, int32_t* mat) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !c->have_scene) return fail(CC_ERR_STATE, "no scene");
  if (n <= 0) return CC_OK;
  HIP_TRY(hipSetDevice(c->device));
  DevBufs B;
  float* dp;
  int* dm;
  int r;
  if ((r = B.in(pts, 3 * (size_t)n, &dp)) || (r = B.out((size_t)n, &dm))) return r;
  hipLaunchKernelGGL(k_locate, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, c->prims, c->W, (int)n, dp, dm);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return to_host(mat, dm, (size_t)n);
}

int cc_test_boundary(void* ctx, int64_t n, const float* o, const float* d, float eps, float* t, int32_t* escape) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !c->have_scene) return fail(CC_ERR_STATE, "no scene");
  if (n <= 0) return CC_OK;
  HIP_TRY(hipSetDevice(c->device));
  DevBufs B;
  float *dO, *dD, *dT;
  int* dE;
  int r;
  if ((r = B.in(o, 3 * (size_t)n, &dO)) || (r = B.in(d, 3 * (size_t)n, &dD)) || (r = B.out((size_t)n, &dT)) ||
      (r = B.out((size_t)n, &dE)))
    return r;
  hipLaunchKernelGGL(k_test_boundary, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, c->prims, c->W, (int)n, dO,
                     dD, eps, dT, dE);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  if ((r = to_host(t, dT, (size_t)n))) return r;
  return to_host(escape, dE, (size_t)n);
}

static int check_mats(Ctx* c, int64_t n, const int32_t* mat, const float* e) {
  if (!c || !c->have_tables) return fail(CC_ERR_STATE, "no tables");
  for (int64_t i = 0; i < n; ++i) {
    if (mat[i] < 0 || mat[i] >= c->T.n_mat) return fail(CC_ERR_ARG, "material id %d out of range", mat[i]);
    if (!(e[i] >= 1.0f && e[i] <= 150.0f))
      return fail(CC_ERR_ARG, "energy %g keV outside the cross-section table range [1, 150] keV", (double)e[i]);
  }
  return CC_OK;
}

int cc_test_xs(void* ctx, int64_t n, const int32_t* mat, const float* e, float* out9) {
  Ctx* c = (Ctx*)ctx;
  int r = check_mats(c, n, mat, e);
  if (r || n <= 0) return r;
  HIP_TRY(hipSetDevice(c->device));
  DevBufs B;
  int *dm, *derr;
  float *de, *dout;
  if ((r = B.in(mat, (size_t)n, &dm)) || (r = B.in(e, (size_t)n, &de)) || (r = B.out(9 * (size_t)n, &dout)) ||
      (r = B.out(1, &derr)))
    return r;
  HIP_TRY(hipMemset(derr, 0, sizeof(int)));
  hipLaunchKernelGGL(k_test_xs, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, c->T, (int)n, dm, de, dout, derr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return to_host(out9, dout, 9 * (size_t)n);
}

int cc_test_compton(void* ctx, int64_t n, const int32_t* mat, const float* e, const float* u, int32_t k_trials,
                    float* eps, float* cos_t, int32_t* elem, int32_t* trials) {
  Ctx* c = (Ctx*)ctx;
  int r = check_mats(c, n, mat, e);
  if (r || n <= 0) return r;
  if (k_trials <= 0) return fail(CC_ERR_ARG, "k_trials must be > 0");
  HIP_TRY(hipSetDevice(c->device));
  DevBufs B;
  int *dm, *del, *dtr;
  float *de, *du, *deps, *dcs;
  const int ku = 1 + 3 * k_trials;
  if ((r = B.in(mat, (size_t)n, &dm)) || (r = B.in(e, (size_t)n, &de)) || (r = B.in(u, (size_t)n * ku, &du)) ||
      (r = B.out((size_t)n, &deps)) || (r = B.out((size_t)n, &dcs)) || (r = B.out((size_t)n, &del)) ||
      (r = B.out((size_t)n, &dtr)))
    return r;
  hipLaunchKernelGGL(k_test_compton, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, c->T, (int)n, dm, de, du, ku,
                     k_trials, deps, dcs, del, dtr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  if ((r = to_host(eps, deps, (size_t)n)) || (r = to_host(cos_t, dcs, (size_t)n)) || (r = to_host(elem, del, (size_t)n)))
    return r;
  return to_host(trials, dtr, (size_t)n);
}

int cc_test_rayleigh(void* ctx, int64_t n, const int32_t* mat, const float* e, const float* u, int32_t k_trials,
                     float* cos_t, int32_t* elem, int32_t* trials) {
  Ctx* c = (Ctx*)ctx;
  int r = check_mats(c, n, mat, e);
  if (r || n <= 0) return r;
  if (k_trials <= 0) return fail(CC_ERR_ARG, "k_trials must be > 0");
  HIP_TRY(hipSetDevice(c->device));
  DevBufs B;
  int *dm, *del, *dtr, *derr;
  float *de, *du, *dcs;
  const int ku = 1 + 2 * k_trials;
  if ((r = B.in(mat, (size_t)n, &dm)) || (r = B.in(e, (size_t)n, &de)) || (r = B.in(u, (size_t)n * ku, &du)) ||
      (r = B.out((size_t)n, &dcs)) || (r = B.out((size_t)n, &del)) || (r = B.out((size_t)n, &dtr)) ||
      (r = B.out(1, &derr)))
    return r;
  HIP_TRY(hipMemset(derr, 0, sizeof(int)));
  hipLaunchKernelGGL(k_test_rayleigh, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, c->T, (int)n, dm, de, du, ku,
                     k_trials, dcs, del, dtr, derr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  int e_bits = 0;
  if ((r = to_host(&e_bits, derr, 1))) return r;
  if (e_bits & CC_DEV_Q_RANGE) return fail(CC_ERR_ARG, "q(E) beyond the Rayleigh form-factor table");
  if ((r = to_host(cos_t, dcs, (size_t)n)) || (r = to_host(elem, del, (size_t)n))) return r;
  return to_host(trials, dtr, (size_t)n);
}

int cc_test_fluor(void* ctx, int64_t n, const int32_t* mat, const float* e, const float* u4, int32_t* elem,
                  int32_t* emit, float* e_line) {
  Ctx* c = (Ctx*)ctx;
  int r = check_mats(c, n, mat, e);
  if (r || n <= 0) return r;
  HIP_TRY(hipSetDevice(c->device));
  DevBufs B;
  int *dm, *del, *dem;
  float *de, *du, *dl;
  if ((r = B.in(mat, (size_t)n, &dm)) || (r = B.in(e, (size_t)n, &de)) || (r = B.in(u4, 4 * (size_t)n, &du)) ||
      (r = B.out((size_t)n, &del)) || (r = B.out((size_t)n, &dem)) || (r = B.out((size_t)n, &dl)))
    return r;
  hipLaunchKernelGGL(k_test_fluor, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, c->T, (int)n, dm, de, du, del,
                     dem, dl);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  if ((r = to_host(elem, del, (size_t)n)) || (r = to_host(emit, dem, (size_t)n))) return r;
  return to_host(e_line, dl, (size_t)n);
}

int cc_test_rotate(int64_t n, const float* d, const float* cos_t, const float* u_phi, float* out) {
  if (n <= 0) return CC_OK;
  DevBufs B;
  float *dd, *dc, *du, *dout;
  int r;
  if ((r = B.in(d, 3 * (size_t)n, &dd)) || (r = B.in(cos_t, (size_t)n, &dc)) || (r = B.in(u_phi, (size_t)n, &du)) ||
      (r = B.out(3 * (size_t)n, &dout)))
    return r;
  hipLaunchKernelGGL(k_test_rotate, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, (int)n, dd, dc, du, dout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return to_host(out, dout, 3 * (size_t)n);
}

int cc_test_isotropic(int64_t n, const float* u1, const float* u2, float* out) {
  if (n <= 0) return CC_OK;
  DevBufs B;
  float *d1, *d2, *dout;
  int r;
  if ((r = B.in(u1, (size_t)n, &d1)) || (r = B.in(u2, (size_t)n, &d2)) || (r = B.out(3 * (size_t)n, &dout))) return r;
  hipLaunchKernelGGL(k_test_iso, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, (int)n, d1, d2, dout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return to_host(out, dout, 3 * (size_t)n);
}

int cc_test_dda(void* ctx, int64_t n, const float* o, const float* d, const float* len, const double* wk,
                const double* wh) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !c->have_grid) return fail(CC_ERR_STATE, "no grid bound");
  if (n <= 0) return CC_OK;
  HIP_TRY(hipSetDevice(c->device));
  DevBufs B;
  float *dO, *dD, *dL;
  double *dk, *dh;
  int r;
  if ((r = B.in(o, 3 * (size_t)n, &dO)) || (r = B.in(d, 3 * (size_t)n, &dD)) || (r = B.in(len, (size_t)n, &dL)) ||
      (r = B.in(wk, (size_t)n, &dk)) || (r = B.in(wh, (size_t)n, &dh)))
    return r;
  hipLaunchKernelGGL(k_test_dda, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->G, (int)n, dO, dD, dL, dk,
                     dh);
  hipLaunchKernelGGL(k_end_batch, dim3(grid_blocks(c->nvox)), dim3(kBlock), 0, c->stream, c->G.dk, c->G.dh, c->tk, c->th,
                     c->qk, c->qh, c->hits, c->nvox, 1.0, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CC_OK;
}

int cc_test_philox(int64_t n, const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4) {
  if (n <= 0) return CC_OK;
  DevBufs B;
  uint32_t *dc, *dk, *dout;
  int r;
  if ((r = B.in(ctr4, 4 * (size_t)n, &dc)) || (r = B.in(key2, 2 * (size_t)n, &dk)) || (r = B.out(4 * (size_t)n, &dout)))
    return r;
  hipLaunchKernelGGL(k_test_philox, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, (int)n, dc, dk, dout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return to_host(out4, dout, 4 * (size_t)n);
}

int cc_test_source(void* ctx, const cc_source* src, int64_t n, uint64_t seed, uint64_t first, float* pos, float* dir,
                   float* energy) {
  Ctx* c = (Ctx*)ctx;
  if (!c || !src) return fail(CC_ERR_ARG, "null argument");
  if (n <= 0) return CC_OK;
  HIP_TRY(hipSetDevice(c->device));
  Src s;
  int r = make_src_impl(c, src, &s);
  if (r) return r;
  DevBufs B;
  float *dp, *dd, *de;
  if ((r = B.out(3 * (size_t)n, &dp)) || (r = B.out(3 * (size_t)n, &dd)) || (r = B.out((size_t)n, &de))) return r;
  hipLaunchKernelGGL(k_test_source, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, 0, s, seed, first, (int)n, dp, dd,
                     de);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  if ((r = to_host(pos, dp, 3 * (size_t)n)) || (r = to_host(dir, dd, 3 * (size_t)n))) return r;
  return to_host(energy, de, (size_t)n);
}

}  // extern "C"
