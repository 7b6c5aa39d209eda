// Learned refinement prediction: per-unit features from the particles, window labels from
// the plain criterion, the fused network + decision, and the activation of predicted cells.
// None of this has a counterpart in the reference, whose octree is rebuilt from the current
// particles every substep (adaptive_gimp.py:87-96, :323-337).
//
// A unit is a cell of level ulev = max(0, L-3): a block of 4^d finest cells.  Features are
// accumulated per unit with segmented wavefront reductions (particles are stored sorted by
// (level, leaf block, cell), so lanes of one unit are mostly contiguous) and one atomic per
// run; candidates are the units within one unit of an occupied one.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "launch.hpp"
#include "mlp.hpp"

namespace agimp {

template <int D>
struct P4 { static constexpr int v = (D == 2) ? 16 : 64; };

template <int D>
__device__ __forceinline__ int64_t unit_lidx(const LevelDesc& L, const int* I) {
  int J[3];
#pragma unroll
  for (int a = 0; a < D; ++a) J[a] = I[a] + PAD;
  if (!in_box<D>(L.E, J)) return -1;
  return (int64_t)bidx<D>(L.NB, J);
}

// padded coordinates of block-major cell index q
template <int D>
__device__ __forceinline__ void unit_coords(const LevelDesc& L, int q, int* J) {
  const int b = q / P4<D>::v, w = q % P4<D>::v;
  int B[3];
  block_coords<D>(L.NB, b, B);
  if (D == 2) {
    J[0] = B[0] * 4 + (w >> 2);
    J[1] = B[1] * 4 + (w & 3);
  } else {
    J[0] = B[0] * 4 + (w >> 4);
    J[1] = B[1] * 4 + ((w >> 2) & 3);
    J[2] = B[2] * 4 + (w & 3);
  }
}

// monotone map of floats to unsigned keys; zero-initialised memory is below every key
__device__ __forceinline__ uint32_t fkey(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funkey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ int udpp(int src, int old) {
  return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWMASK, 0xf, false);
}

// Segmented inclusive scan over runs of equal keys.  A key may occur in several runs of one
// wavefront (units interleave when particles are ordered by the cells of a coarser level),
// so a step combines with its source lane only if that lane lies in the lane's own run
// (source index >= the run's first lane) and the DPP move has a source at all (same key).
struct USeg {
  bool s1, s2, s4, s8, b15, b31, tail;
};
__device__ __forceinline__ USeg useg(int key) {
  USeg m;
  const int lane = __lane_id();
  const int prev = __shfl_up(key, 1);
  const unsigned long long heads = __ballot(lane == 0 || prev != key);
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const int head = 63 - __clzll(heads & upto);  // first lane of this lane's run
  m.s1 = udpp<0x111, 0xf>(key, -1) == key && lane - 1 >= head;
  m.s2 = udpp<0x112, 0xf>(key, -1) == key && lane - 2 >= head;
  m.s4 = udpp<0x114, 0xf>(key, -1) == key && lane - 4 >= head;
  m.s8 = udpp<0x118, 0xf>(key, -1) == key && lane - 8 >= head;
  m.b15 = udpp<0x142, 0xa>(key, -1) == key && (lane & ~15) - 1 >= head;  // source: last lane of the previous row
  m.b31 = udpp<0x143, 0xc>(key, -1) == key && 31 >= head;
  const int next = __shfl_down(key, 1);  // every lane takes part in the shuffle (no short-circuit)
  m.tail = (lane == 63) || (next != key);
  return m;
}
#define USEG_STEP(OP, v, m)                                                                               \
  do {                                                                                                   \
    float t_;                                                                                            \
    t_ = __int_as_float(udpp<0x111, 0xf>(__float_as_int(v), __float_as_int(v))); if (m.s1) v = OP(v, t_);  \
    t_ = __int_as_float(udpp<0x112, 0xf>(__float_as_int(v), __float_as_int(v))); if (m.s2) v = OP(v, t_);  \
    t_ = __int_as_float(udpp<0x114, 0xf>(__float_as_int(v), __float_as_int(v))); if (m.s4) v = OP(v, t_);  \
    t_ = __int_as_float(udpp<0x118, 0xf>(__float_as_int(v), __float_as_int(v))); if (m.s8) v = OP(v, t_);  \
    t_ = __int_as_float(udpp<0x142, 0xa>(__float_as_int(v), __float_as_int(v))); if (m.b15) v = OP(v, t_); \
    t_ = __int_as_float(udpp<0x143, 0xc>(__float_as_int(v), __float_as_int(v))); if (m.b31) v = OP(v, t_); \
  } while (0)
__device__ __forceinline__ float op_add(float a, float b) { return a + b; }
__device__ __forceinline__ float op_max(float a, float b) { return fmaxf(a, b); }

// a lane's request key: level g and the base cell of x at that level (20 bits per axis),
// or a value unique to the lane when the base cell does not fit the fields
template <int D>
__device__ __forceinline__ unsigned long long request_key(int g, const float* x, const LevelDesc& Lg, int pow2,
                                                          int lane) {
  int b[3] = {0, 0, 0};
  bool packs = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    b[a] = base_cell(x[a], Lg, pow2);
    packs = packs && b[a] >= -2 && b[a] < (1 << 20) - 2;
  }
  if (!packs) return ~(unsigned long long)lane;
  return ((unsigned long long)g << 60) | ((unsigned long long)(b[0] + 2) << 40) |
         ((unsigned long long)(b[1] + 2) << 20) | (unsigned long long)(D == 3 ? b[2] + 2 : 0);
}

// the particle's requests (adaptive_gimp.py:87-96): its 2^d cells at level g; record g+1 in
// the units they lie in, for g >= ulev (coarser requests are not predicted)
template <int D>
__device__ __forceinline__ void record_requests(const Topo& T, int ulev, const float* x, int g, int* dst,
                                                int stride, const int* near = nullptr) {
  if (g < ulev || g >= T.L) return;
  const LevelDesc Lg = T.lv[g];
  const LevelDesc Lu = T.lv[ulev];
  int b[3];
#pragma unroll
  for (int a = 0; a < D; ++a) b[a] = base_cell(x[a], Lg, T.pow2);
  int64_t seen[1 << D];
#pragma unroll
  for (int t = 0; t < (1 << D); ++t) {
    int Iu[3];
#pragma unroll
    for (int a = 0; a < D; ++a) Iu[a] = (b[a] + ((t >> (D - 1 - a)) & 1)) >> (g - ulev);
    bool far = false;  // only units within one unit of `near` (the candidates around it)
    if (near) {
#pragma unroll
      for (int a = 0; a < D; ++a) far = far || Iu[a] < near[a] - 1 || Iu[a] > near[a] + 1;
    }
    const int64_t q = far ? -1 : unit_lidx<D>(Lu, Iu);
    seen[t] = q;
    bool dup = q < 0;
#pragma unroll
    for (int s = 0; s < t; ++s) dup = dup || seen[s] == q;
    // most units already hold this level (many particles request the same units): a read of
    // the L2 value first (agent scope: a plain load may hit a stale per-CU cache line) skips
    // the contended atomic; a stale read only costs a redundant atomic
    if (!dup && __hip_atomic_load(&dst[q * stride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < g + 1)
      atomicMax(&dst[q * stride], g + 1);
  }
}

// ------------------------------------------------------------------------------ clears
// accumulators of the previous candidates (the only units ever written)
__global__ void k_unit_clear_acc(LearnDesc U) {
  const int nc = U.counts[1];
  const int64_t tot = (int64_t)nc * UNIT_ACC;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot; i += (int64_t)gridDim.x * blockDim.x)
    U.acc[(int64_t)U.cand[i / UNIT_ACC] * UNIT_ACC + i % UNIT_ACC] = 0u;
}

// flags, predicted levels (-1) and counters
__global__ void k_unit_clear_flags(LearnDesc U) {
  const int64_t nw = (U.nu + 3) / 4;
  uint32_t* uf = reinterpret_cast<uint32_t*>(U.uflag);
  uint32_t* cf = reinterpret_cast<uint32_t*>(U.cflag);
  uint32_t* up = reinterpret_cast<uint32_t*>(U.upred);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nw; i += (int64_t)gridDim.x * blockDim.x) {
    uf[i] = 0u;
    cf[i] = 0u;
    up[i] = 0xffffffffu;
  }
  if (blockIdx.x == 0 && threadIdx.x < 4) U.counts[threadIdx.x] = 0;
}

__global__ void k_unit_clear_labels(LearnDesc U) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < U.nu; i += (int64_t)gridDim.x * blockDim.x)
    U.lab[i] = 0;
}

// ------------------------------------------------------------------------------ accumulation
constexpr int ACC_TABLE_BITS = 8, ACC_TABLE = 1 << ACC_TABLE_BITS;  // >= threads per workgroup
constexpr int ACC_NSUM = 8;  // mass, count, |v|, |C|, det F - 1, v (3)

// add one unit's partial sums (order of `sums`: mass, count, |v|, |C|, det F - 1, v) and maxima
// (|v| bits, key of -det F, level + 1) to its accumulators
template <int D>
__device__ __forceinline__ void flush_unit(const LearnDesc& U, int u, const float* sums, const uint32_t* maxs) {
  float* af = reinterpret_cast<float*>(U.acc) + (int64_t)u * UNIT_ACC;
  uint32_t* au = U.acc + (int64_t)u * UNIT_ACC;
  atomicAdd(&af[UA_MASS], sums[0]);
  atomicAdd(&af[UA_COUNT], sums[1]);
  atomicAdd(&af[UA_VSUM], sums[2]);
  atomicAdd(&af[UA_CSUM], sums[3]);
  atomicAdd(&af[UA_JSUM], sums[4]);
  atomicAdd(&af[UA_V + 0], sums[5]);
  atomicAdd(&af[UA_V + 1], sums[6]);
  if (D == 3) atomicAdd(&af[UA_V + 2], sums[7]);
  atomicMax(&au[UA_VMAX], maxs[0]);
  atomicMax(&au[UA_JMIN], maxs[1]);
  atomicMax(reinterpret_cast<int*>(&au[UA_GMAX]), (int)maxs[2]);
  U.uflag[u] = 1;
}
template <int D>
__global__ void k_unit_accumulate(const SceneDesc* __restrict__ Dsc, int par, LearnDesc U) {
  const SceneDesc& sd = Dsc[0];
  const Topo& T = sd.T;
  const PBuf& P = sd.P[par];
  using PL = PLayout<D>;
  const int n = sd.n;
  if (blockIdx.x * blockDim.x >= (unsigned)n) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = k < n;
  const LevelDesc Lu = T.lv[U.ulev];
  int u = -1, g = -1;
  float x[3] = {0.f, 0.f, 0.f}, v[3] = {0.f, 0.f, 0.f};
  float m = 0.f, vn = 0.f, cn = 0.f, jm1 = 0.f, njac = 0.f;
  float xf[3] = {0.f, 0.f, 0.f};
  int Iown[3] = {0, 0, 0};
  unsigned long long k1 = ~(unsigned long long)__lane_id(), k2 = k1;  // unique unless set below
  if (live) {
    int I[3];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      x[a] = P.f[(size_t)(PL::X + a) * P.stride + k];
      v[a] = P.f[(size_t)(PL::V + a) * P.stride + k];
      I[a] = (int)floorf(x[a] * Lu.inv_dx);
      Iown[a] = I[a];
    }
    const int64_t q = unit_lidx<D>(Lu, I);
    u = (int)q;
    g = P.g[k];
    m = P.f[(size_t)PL::M * P.stride + k];
    float F[3][3], C2 = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        F[a][b] = P.f[(size_t)(PL::F + a * D + b) * P.stride + k];
        const float c = P.f[(size_t)(PL::C + a * D + b) * P.stride + k];
        C2 += c * c;
      }
    float J;
    if (D == 2) J = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    else
      J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
          F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
#pragma unroll
    for (int a = 0; a < D; ++a) vn += v[a] * v[a];
    vn = sqrtf(vn);
    cn = sqrtf(C2);
    jm1 = J - 1.f;
    njac = -J;
    // the same requests after a straight move over the window (one rounding: fma)
#pragma unroll
    for (int a = 0; a < D; ++a) xf[a] = __fmaf_rn(v[a], U.win, x[a]);
    if (u >= 0 && g >= U.ulev && g < T.L) {
      const LevelDesc Lg = T.lv[g];
      k1 = request_key<D>(g, x, Lg, T.pow2, __lane_id());
      k2 = request_key<D>(g, xf, Lg, T.pow2, __lane_id());
    }
  }
  // Particles arrive sorted by cell, so neighbours mostly record the same requests: a lane
  // whose left neighbour has the same key (level and base cell; for the swept requests also
  // the same unit, which bounds the recorded units) leaves the records to it.  Every lane
  // reaches the shuffles.
  const unsigned long long l1 = __shfl_up(k1, 1), l2 = __shfl_up(k2, 1);
  const int lu = __shfl_up(u, 1);
  const bool first = __lane_id() > 0;
  if (live && u >= 0) {
    // request cells lie within one unit of the particle's unit: all are candidates
    if (!(first && l1 == k1)) record_requests<D>(T, U.ulev, x, g, reinterpret_cast<int*>(U.acc) + UA_REQ, UNIT_ACC);
    if (!(first && l2 == k2 && lu == u))
      record_requests<D>(T, U.ulev, xf, g, reinterpret_cast<int*>(U.acc) + UA_REQF, UNIT_ACC, Iown);
  }
  const bool in = live && u >= 0;
  const int key = in ? u : -2 - (int)__lane_id();
  const USeg sm = useg(key);
  float s_m = m, s_n = in ? 1.f : 0.f, s_v = vn, s_c = cn, s_j = jm1;
  float s_v0 = v[0], s_v1 = v[1], s_v2 = v[2];
  float x_v = vn, x_j = njac, x_g = (float)(g + 1);
  USEG_STEP(op_add, s_m, sm);
  USEG_STEP(op_add, s_n, sm);
  USEG_STEP(op_add, s_v, sm);
  USEG_STEP(op_add, s_c, sm);
  USEG_STEP(op_add, s_j, sm);
  USEG_STEP(op_add, s_v0, sm);
  USEG_STEP(op_add, s_v1, sm);
  if (D == 3) USEG_STEP(op_add, s_v2, sm);
  USEG_STEP(op_max, x_v, sm);
  USEG_STEP(op_max, x_j, sm);
  USEG_STEP(op_max, x_g, sm);
  // run sums of the workgroup are merged per unit in an LDS table first (units of one
  // workgroup are few, but their runs are short: particles are ordered by base cell, units
  // by containing cell), then each unit is flushed once with global atomics
  __shared__ int tkey[ACC_TABLE];
  __shared__ float tsum[ACC_TABLE][ACC_NSUM];
  __shared__ uint32_t tmax[ACC_TABLE][3];
  for (int i = threadIdx.x; i < ACC_TABLE; i += blockDim.x) {
    tkey[i] = -1;
#pragma unroll
    for (int j = 0; j < ACC_NSUM; ++j) tsum[i][j] = 0.f;
    tmax[i][0] = tmax[i][1] = tmax[i][2] = 0u;
  }
  __syncthreads();
  if (in && sm.tail) {
    int slot = -1;
    uint32_t h = ((uint32_t)u * 2654435761u) >> (32 - ACC_TABLE_BITS);
    for (int probe = 0; probe < ACC_TABLE; ++probe) {
      const int e = (int)((h + probe) & (ACC_TABLE - 1));
      const int prev = atomicCAS(&tkey[e], -1, u);
      if (prev == -1 || prev == u) { slot = e; break; }
    }
    const float sums[ACC_NSUM] = {s_m, s_n, s_v, s_c, s_j, s_v0, s_v1, s_v2};
    const uint32_t maxs[3] = {__float_as_uint(x_v), fkey(x_j), (uint32_t)(int)x_g};
    if (slot >= 0) {
#pragma unroll
      for (int j = 0; j < ACC_NSUM; ++j)
        if (j < 7 || D == 3) atomicAdd(&tsum[slot][j], sums[j]);
#pragma unroll
      for (int j = 0; j < 3; ++j) atomicMax(&tmax[slot][j], maxs[j]);
    } else {  // table full (not expected for 256 particles): straight to the unit
      flush_unit<D>(U, u, sums, maxs);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ACC_TABLE; i += blockDim.x) {
    const int uu = tkey[i];
    if (uu < 0) continue;
    float sums[ACC_NSUM];
    uint32_t maxs[3];
#pragma unroll
    for (int j = 0; j < ACC_NSUM; ++j) sums[j] = tsum[i][j];
#pragma unroll
    for (int j = 0; j < 3; ++j) maxs[j] = tmax[i][j];
    flush_unit<D>(U, uu, sums, maxs);
  }
}

// the window labels: finest requested level + 1 per unit, max over the calls since the clear
template <int D>
__global__ void k_unit_requests(const SceneDesc* __restrict__ Dsc, int par, LearnDesc U) {
  const SceneDesc& sd = Dsc[0];
  const PBuf& P = sd.P[par];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  float x[3];
#pragma unroll
  for (int a = 0; a < D; ++a) x[a] = P.f[(size_t)a * P.stride + k];
  record_requests<D>(sd.T, U.ulev, x, P.g[k], U.lab, 1);
}

// ------------------------------------------------------------------------------ candidates
// append the indices of set byte flags to list, count in *count
__global__ void k_unit_compact(const uint8_t* __restrict__ flags, int n, int* list, int* count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = i < n && flags[i] != 0;
  const int idx = wave_agg_atomic_add(count, on);
  if (on) list[idx] = i;
}

template <int D>
__global__ void k_unit_dilate(const SceneDesc* __restrict__ Dsc, LearnDesc U) {
  const LevelDesc Lu = Dsc[0].T.lv[U.ulev];
  const int no = U.counts[0];
  constexpr int NN = (D == 2) ? 9 : 27;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)no * NN;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int q = U.occ[t / NN];
    int r = (int)(t % NN);
    int J[3];
    unit_coords<D>(Lu, q, J);
#pragma unroll
    for (int a = D - 1; a >= 0; --a) { J[a] += r % 3 - 1; r /= 3; }
    if (!in_box<D>(Lu.E, J)) continue;
    U.cflag[bidx<D>(Lu.NB, J)] = 1;
  }
}

// ------------------------------------------------------------------------------ features
__device__ __forceinline__ float lev_code(int lp1, int L) {
  // level + 1 (0: none) -> finest 1, one coarser 0.75, two coarser 0.5, coarser 0.25, none 0
  if (lp1 <= 0) return 0.f;
  return 0.25f * (float)(1 + max(0, 3 - (L - lp1)));
}

template <int D>
__global__ void k_unit_features(const SceneDesc* __restrict__ Dsc, LearnDesc U) {
  const Topo& T = Dsc[0].T;
  const LevelDesc Lu = T.lv[U.ulev];
  const int L = T.L;
  const int nc = U.counts[1];
  const float disp = U.win * Lu.inv_dx;  // velocity -> displacement in units over the window
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    const int q = U.cand[i];
    const uint32_t* au = U.acc + (int64_t)q * UNIT_ACC;
    const float* af = reinterpret_cast<const float*>(au);
    const float cnt = af[UA_COUNT];
    const bool occ = cnt > 0.f;
    const float ic = occ ? 1.f / cnt : 0.f;
    float f[UNIT_NFEAT];
    f[0] = af[UA_MASS] * U.inv_mref;
    f[1] = af[UA_VSUM] * ic * disp;
    f[2] = __uint_as_float(au[UA_VMAX]) * disp;
    f[3] = af[UA_CSUM] * ic * U.win;
    f[4] = af[UA_JSUM] * ic * 10.f;
    f[5] = occ ? (-funkey(au[UA_JMIN]) - 1.f) * 10.f : 0.f;
    f[6] = lev_code((int)au[UA_GMAX], L);
    f[7] = lev_code((int)au[UA_REQF], L);
    f[8] = lev_code((int)au[UA_REQ], L);
    float vb[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < D; ++a) vb[a] = af[UA_V + a] * ic;
    f[12] = vb[0] * disp;
    f[13] = vb[1] * disp;
    f[14] = (D == 3) ? vb[2] * disp : 0.f;
    int J[3];
    unit_coords<D>(Lu, q, J);
    float nm = 0.f, ng = 0.f, nr = 0.f, up = 0.f;
    constexpr int NN = (D == 2) ? 9 : 27;
    for (int r0 = 0; r0 < NN; ++r0) {
      if (r0 == NN / 2) continue;
      int K[3], dd[3], r = r0;
#pragma unroll
      for (int a = D - 1; a >= 0; --a) { dd[a] = r % 3 - 1; K[a] = J[a] + dd[a]; r /= 3; }
      if (!in_box<D>(Lu.E, K)) continue;
      const int64_t qn = bidx<D>(Lu.NB, K);
      const uint32_t* bu = U.acc + qn * UNIT_ACC;
      const float* bf = reinterpret_cast<const float*>(bu);
      nm += bf[UA_MASS];
      const float gc = lev_code((int)bu[UA_GMAX], L);
      ng = fmaxf(ng, gc);
      nr = fmaxf(nr, lev_code((int)bu[UA_REQ], L));
      const float cn = bf[UA_COUNT];
      if (cn > 0.f) {
        // the neighbour moves towards this unit: mean velocity . (this - neighbour) > 0
        float dot = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) dot -= bf[UA_V + a] * (float)dd[a];
        if (dot > 0.f) up = fmaxf(up, gc);
      }
    }
    f[9] = nm * U.inv_mref / (float)(NN - 1);
    f[10] = ng;
    f[11] = nr;
    f[15] = up;
    float4* out = reinterpret_cast<float4*>(U.feat + (int64_t)i * UNIT_NFEAT);
    out[0] = make_float4(f[0], f[1], f[2], f[3]);
    out[1] = make_float4(f[4], f[5], f[6], f[7]);
    out[2] = make_float4(f[8], f[9], f[10], f[11]);
    out[3] = make_float4(f[12], f[13], f[14], f[15]);
  }
}

// class of each candidate's window label: 0 finest, 1 one coarser, 2 two coarser, 3 none
__global__ void k_unit_label_classes(LearnDesc U, int L, int* out) {
  const int nc = U.counts[1];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    const int r = U.lab[U.cand[i]];
    out[i] = (r <= 0) ? 3 : min(3, L - r);
  }
}

// ------------------------------------------------------------------------------ prediction
constexpr int PRED_T = 256;

// network on the candidates' features, then the decision rule: class c = first whose
// cumulative probability reaches tau (3 if none); predicted level L-1-c if >= ulev
__global__ void __launch_bounds__(PRED_T) k_unit_predict(LearnDesc U, int L, int* cls_out) {
  __shared__ __align__(16) MlpLds W;
  mlp_load(W, U.params);
  __syncthreads();
  const int lane = __lane_id();
  const int64_t rows = U.counts[1];
  const int64_t ntiles = (rows + 15) / 16;
  const int64_t wave = (int64_t)blockIdx.x * (PRED_T / 64) + threadIdx.x / 64;
  const int64_t nwaves = (int64_t)gridDim.x * (PRED_T / 64);
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    MlpTile t;
    mlp_forward_tile(W, U.feat, rows, tile * 16, t);
    const int64_t row = tile * 16 + lane;
    int refined = 0;
    if (lane < 16 && row < rows) {
      const float mx = fmaxf(fmaxf(t.z[0], t.z[1]), fmaxf(t.z[2], t.z[3]));
      float e[4], se = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) { e[j] = __expf(t.z[j] - mx); se += e[j]; }
      const float thr = U.tau * se;
      int c = 3;
      float cum = 0.f;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        cum += e[j];
        if (c == 3 && cum >= thr) c = j;
      }
      const int lev = L - 1 - c;
      const int8_t pl = (c < 3 && lev >= U.ulev) ? (int8_t)lev : (int8_t)-1;
      U.upred[U.cand[row]] = pl;
      refined = pl >= 0;
      if (cls_out) cls_out[row] = c;
    }
    const unsigned long long b = __ballot(refined);
    if (lane == 0 && b) atomicAdd(&U.counts[2], __popcll(b));
  }
}

// The requests (adaptive_gimp.py:87-96) a particle makes while it moves in a straight line
// over the window: per axis the cells from min(b0, b1) to max(b0, b1) + 1, with b0 the base
// cell now and b1 the base cell of x + v * window (at most 3^d cells for moves shorter than a
// cell).  Part of every predicted octree (`mode` bit 1); the network's unit levels are the
// other part (`mode` bit 0).
template <int D>
__global__ void k_activate_ahead(const SceneDesc* __restrict__ Dsc, int par, LearnDesc U) {
  const SceneDesc& sd = Dsc[0];
  const Topo& T = sd.T;
  const PBuf& P = sd.P[par];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int g = P.g[k];
  if (g < 0 || g >= T.L) return;
  const LevelDesc Lg = T.lv[g];
  int lo[3], ext[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const float x = P.f[(size_t)a * P.stride + k];
    const float xf = __fmaf_rn(P.f[(size_t)(D + a) * P.stride + k], U.win, x);
    const int b0 = base_cell(x, Lg, T.pow2), b1 = base_cell(xf, Lg, T.pow2);
    lo[a] = min(b0, b1);
    ext[a] = min(max(b0, b1) - lo[a], 1) + 2;  // 2 or 3 cells (longer moves are left to the check)
  }
  constexpr int NMAX = (D == 2) ? 9 : 27;
  for (int t = 0; t < NMAX; ++t) {
    int J[3], r = t;
    bool in = true;
#pragma unroll
    for (int a = D - 1; a >= 0; --a) {
      const int o = r % 3;
      r /= 3;
      in = in && o < ext[a];
      J[a] = lo[a] + o + PAD;
    }
    if (!in || !in_box<D>(Lg.E, J)) continue;
    const uint32_t q = bidx<D>(Lg.NB, J);
    Lg.cell[q] = ACTIVATED;
    T.cflag[Lg.gbase + (int)(q / P4<D>::v)] = 1;
  }
}

// cells of the predicted level inside every refined unit without particles get ACTIVATED
// (between the particle activation and the culling of the rebuild)
template <int D>
__global__ void k_unit_activate(const SceneDesc* __restrict__ Dsc, LearnDesc U) {
  const Topo& T = Dsc[0].T;
  const LevelDesc Lu = T.lv[U.ulev];
  const int nc = U.counts[1];
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)nc * P4<D>::v;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int q = U.cand[t / P4<D>::v];
    const int w = (int)(t % P4<D>::v);
    const int lev = U.upred[q];
    if (lev < U.ulev) continue;
    // only units without particles now: requests inside occupied units come from their own
    // particles (current and swept requests), and finer cells over them would move those
    // particles' transfer level fl_p away from the plain method's
    if (U.uflag[q]) continue;
    const int s = lev - U.ulev;  // 0..2: 2^s cells per axis
    const int side = 1 << s;
    int r = w, off[3];
#pragma unroll
    for (int a = D - 1; a >= 0; --a) { off[a] = r % side; r /= side; }
    if (r != 0) continue;  // w >= side^d
    int J[3], I[3];
    unit_coords<D>(Lu, q, J);
#pragma unroll
    for (int a = 0; a < D; ++a) I[a] = (J[a] - PAD) * side + off[a];
    const LevelDesc Ll = T.lv[lev];
    int Jl[3];
#pragma unroll
    for (int a = 0; a < D; ++a) Jl[a] = I[a] + PAD;
    if (!in_box<D>(Ll.E, Jl)) continue;
    const uint32_t cq = bidx<D>(Ll.NB, Jl);
    Ll.cell[cq] = ACTIVATED;
    T.cflag[Ll.gbase + (int)(cq / P4<D>::v)] = 1;
  }
}

// ------------------------------------------------------------------------------ launchers
#define LEARN_DIM(D, ...) \
  do { if ((D) == 2) { constexpr int DD = 2; __VA_ARGS__; } else { constexpr int DD = 3; __VA_ARGS__; } } while (0)

static inline unsigned nblk(int64_t n, unsigned t, unsigned cap) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + t - 1) / t, cap));
}

void launch_unit_features(hipStream_t s, const SceneDesc* d, int dim, int n, int par, const LearnDesc& U) {
  LEARN_DIM(dim, {
    AGIMP_LAUNCH("k_unit_clear_acc", k_unit_clear_acc, dim3(nblk((int64_t)U.nu * UNIT_ACC, 256, 2048)), dim3(256), 0, s, U);
    AGIMP_LAUNCH("k_unit_clear_flags", k_unit_clear_flags, dim3(nblk(((int64_t)U.nu + 3) / 4, 256, 2048)), dim3(256), 0, s, U);
    AGIMP_LAUNCH("k_unit_accumulate", k_unit_accumulate<DD>, dim3(nblk(n, 256, 1u << 30)), dim3(256), 0, s, d, par, U);
    AGIMP_LAUNCH("k_unit_compact", k_unit_compact, dim3(nblk(U.nu, 256, 1u << 30)), dim3(256), 0, s, U.uflag, U.nu, U.occ,
                       U.counts + 0);
    AGIMP_LAUNCH("k_unit_dilate", k_unit_dilate<DD>, dim3(2048), dim3(256), 0, s, d, U);
    AGIMP_LAUNCH("k_unit_compact", k_unit_compact, dim3(nblk(U.nu, 256, 1u << 30)), dim3(256), 0, s, U.cflag, U.nu, U.cand,
                       U.counts + 1);
    AGIMP_LAUNCH("k_unit_features", k_unit_features<DD>, dim3(1024), dim3(256), 0, s, d, U);
  });
}

void launch_unit_predict(hipStream_t s, const LearnDesc& U, int L, int* cls_out) {
  AGIMP_LAUNCH("k_unit_predict", k_unit_predict, dim3(512), dim3(PRED_T), 0, s, U, L, cls_out);
}

void launch_unit_activate(hipStream_t s, const SceneDesc* d, int dim, int n, int par, const LearnDesc& U) {
  LEARN_DIM(dim, {
    if (U.mode & 1) AGIMP_LAUNCH("k_unit_activate", k_unit_activate<DD>, dim3(2048), dim3(256), 0, s, d, U);
    if (U.mode & 2) AGIMP_LAUNCH("k_activate_ahead", k_activate_ahead<DD>, dim3(nblk(n, 256, 1u << 30)), dim3(256), 0, s, d, par, U);
  });
}

void launch_unit_labels_clear(hipStream_t s, const LearnDesc& U) {
  AGIMP_LAUNCH("k_unit_clear_labels", k_unit_clear_labels, dim3(nblk(U.nu, 256, 2048)), dim3(256), 0, s, U);
}

void launch_unit_labels_accumulate(hipStream_t s, const SceneDesc* d, int dim, int n, int par, const LearnDesc& U) {
  LEARN_DIM(dim, AGIMP_LAUNCH("k_unit_requests", k_unit_requests<DD>, dim3(nblk(n, 256, 1u << 30)), dim3(256), 0, s, d, par, U));
}

void launch_unit_label_classes(hipStream_t s, const LearnDesc& U, int L, int* out) {
  AGIMP_LAUNCH("k_unit_label_classes", k_unit_label_classes, dim3(1024), dim3(256), 0, s, U, L, out);
}

}  // namespace agimp
