// HIP kernels of the adaptive-GIMP substep for gfx950 (wave64).
//
// Phase order per substep (adaptive_gimp.py:323-352):
//   refine:   clear -> activation (dynamic :87-96 | static :72-85) -> culling (:98-124)
//             -> node A marks (:142-153) -> T marks -> ghost marks (:160-187)
//   bucket:   level mapping (:189-216) as a counting sort by (level, leaf block)
//   physics:  P2G (:230-250) -> restriction (:289-301) -> grid update (:277-287)
//             -> prolongation (:303-321) -> G2P (:252-275) + plasticity + next stress
//
// Differences of execution (not of results) from the reference:
//  * culled cells keep their A|C bits instead of being zeroed (every later test is an exact
//    `== ACTIVATED` comparison, so A|C behaves as 0); readback strips them;
//  * all T marks of all levels are made before any ghost mark, so the ghost test
//    `node == UNACTIVATED` (:182) never races a T mark (:174): raw bits are canonical;
//  * restriction and prolongation are written as gathers over the receiving node (same
//    sums, no float atomics); prolongation overwrites instead of zero-then-add (:308-321);
//  * the grid update skips nodes that prolongation overwrites anyway (T, G, all-even
//    nodes on levels >= 1);
//  * the Kirchhoff stress (utils.py:11-16) of the next P2G is computed at the end of G2P
//    from the same FP64 SVD that the plasticity clamp needs, and stored (6 floats in 3D).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "launch.hpp"
#include "svd.hpp"

namespace agimp {

template <int D>
struct Pow4 { static constexpr int v = (D == 2) ? 16 : 64; };
template <int D>
struct Pow3 { static constexpr int v = (D == 2) ? 9 : 27; };

// the scene of this workgroup (blockIdx.y); kernels that touch particles also take the
// index `par` of the current particle buffer (the other one receives G2P's output)
#define AGIMP_SCENE                          \
  const SceneDesc& sd = Dsc[blockIdx.y];     \
  const Topo& T = sd.T;

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

__global__ void k_copy_u32(uint32_t* dst, const uint32_t* src, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = src[i];
}

// cell t of a compacted block list (global block ids): level, coordinates and index
template <int D>
__device__ __forceinline__ bool decode_list(const Topo& T, const int* list, int count, int t, int& l, int* I,
                                            uint32_t& q) {
  int e = t / Pow4<D>::v, w = t % Pow4<D>::v;
  if (e >= count) return false;
  int g = list[e];
  l = level_of_gbid(T, g);
  int b = g - T.lv[l].gbase;
  int B[3];
  block_coords<D>(T.lv[l].NB, b, B);
  if (D == 2) {
    I[0] = B[0] * 4 + (w >> 2) - PAD;
    I[1] = B[1] * 4 + (w & 3) - PAD;
  } else {
    I[0] = B[0] * 4 + (w >> 4) - PAD;
    I[1] = B[1] * 4 + ((w >> 2) & 3) - PAD;
    I[2] = B[2] * 4 + (w & 3) - PAD;
  }
  q = (uint32_t)b * Pow4<D>::v + w;
  return true;
}

// decode_list for a thread index t that runs over the list contiguously with lane 0 of every
// wavefront at a multiple of 64 (workgroups of a multiple of 64 threads): in 3D a leaf block
// has 64 cells, so all lanes of a wavefront share the list entry, and its load, level and
// block coordinates are computed once per wavefront in scalar registers
template <int D>
__device__ __forceinline__ bool decode_list_wave(const Topo& T, const int* list, int count, int t, int& l, int* I,
                                                 uint32_t& q) {
  if constexpr (D == 3) {
    const int e = __builtin_amdgcn_readfirstlane(t >> 6), w = t & 63;
    if (e >= count) return false;
    const int g = __builtin_amdgcn_readfirstlane(list[e]);
    l = level_of_gbid(T, g);
    const int NB = T.lv[l].NB, b = g - T.lv[l].gbase;
    int B[3];
    block_coords<D>(NB, b, B);
    I[0] = B[0] * 4 + (w >> 4) - PAD;
    I[1] = B[1] * 4 + ((w >> 2) & 3) - PAD;
    I[2] = B[2] * 4 + (w & 3) - PAD;
    q = (uint32_t)b * Pow4<D>::v + w;
    return true;
  } else {
    return decode_list<D>(T, list, count, t, l, I, q);
  }
}

// flag the leaf block holding cell q of level l as non-empty
template <int D>
__device__ __forceinline__ void mark_cell_block(const Topo& T, int l, int64_t q) {
  T.cflag[T.lv[l].gbase + (int)(q / Pow4<D>::v)] = 1;
}

// index of level-l coordinates I (unpadded), or -1 if outside the padded box
template <int D>
__device__ __forceinline__ int64_t lidx(const LevelDesc& L, const int* I) {
  int J[3];
#pragma unroll
  for (int a = 0; a < D; ++a) J[a] = I[a] + PAD;
  if (!in_box<D>(L.E, J)) return -1;
  return (int64_t)bidx<D>(L.NB, J);
}

// node_mask entry of level-l node I (adaptive_gimp.py:126-128 maps to the finest index)
template <int D>
__device__ __forceinline__ int64_t nidx(const Topo& T, int l, const int* I) {
  int f[3];
#pragma unroll
  for (int a = 0; a < D; ++a) f[a] = I[a] << T.lv[l].shift;
  return lidx<D>(T.lv[T.L - 1], f);
}

__device__ __forceinline__ void count(unsigned long long* cnt, int which, unsigned long long v = 1) {
  atomicAdd(&cnt[which], v);
}

// The 2^d corners (I + k, k in {0,1}^d) of level-l cell I as finest-level nodes: their
// node-mask addresses and in-box flags, from per-axis terms (corner u: bit D-1-a of u is k_a)
template <int D>
struct Corners {
  uint32_t t[3][2];
  bool in[3][2];
  __device__ __forceinline__ Corners(const Topo& T, int l, const int* I) {
    const LevelDesc Lf = T.lv[T.L - 1];
    const int sh = T.lv[l].shift;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int j = ((I[a] + k) << sh) + PAD;
        in[a][k] = j >= 0 && j < Lf.E;
        t[a][k] = bidx_axis<D>(Lf.NB, a, j);
      }
  }
  __device__ __forceinline__ bool addr(int u, uint32_t& q) const {
    bool ok = true;
    q = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const int k = (u >> (D - 1 - a)) & 1;
      ok = ok && in[a][k];
      q += t[a][k];
    }
    return ok;
  }
};

// Flag the grid-phase blocks of every level on which the corners of level-l cell I selected
// by the bits of `sel` are nodes: per level, a corner is a node there if its finest
// coordinates are multiples of that level's spacing, and its block flag index is a sum of
// per-axis terms.  Unrolled over the level bound, so the level descriptors are read once.
template <int D>
__device__ __forceinline__ void mark_valid_corners(const Topo& T, int l, const int* I, uint32_t sel) {
  if (!sel) return;
  const int sh = T.lv[l].shift;
#pragma unroll
  for (int i = 0; i < MAXL; ++i) {
    const int lv = T.L - 1 - i;
    if (lv >= 0) {
      const LevelDesc Lv = T.lv[lv];
      const int msk = (1 << Lv.shift) - 1;
      bool ok[3][2];
      int bt[3][2];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const int f = (I[a] + k) << sh;
          const int J = (f >> Lv.shift) + PAD;
          ok[a][k] = (f & msk) == 0 && J >= 0 && J < Lv.E;
          const int stride = (D == 2) ? (a == 0 ? Lv.NB : 1) : (a == 0 ? Lv.NB * Lv.NB : (a == 1 ? Lv.NB : 1));
          bt[a][k] = (J >> 2) * stride;
        }
#pragma unroll
      for (int u = 0; u < (1 << D); ++u) {
        if (!((sel >> u) & 1u)) continue;
        bool alive = true;
        int b = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
          const int k = (u >> (D - 1 - a)) & 1;
          alive = alive && ok[a][k];
          b += bt[a][k];
        }
        if (alive) Lv.sflag[b] = 1;
      }
    }
  }
}

// Clears done by a part of a launch: workgroup wg of the nwg workgroups that take part
// (grid-stride loops over those alone).

// zero n words
__device__ __forceinline__ void fill_zero(uint32_t* p, size_t n, int wg, int nwg) {
  const size_t st = (size_t)nwg * blockDim.x;
  for (size_t i = wg * (size_t)blockDim.x + threadIdx.x; i < n; i += st) p[i] = 0u;
}

// zero the grid of every block in the grid-phase lists (counts and lists as left by the
// previous substep); grid-stride over all listed nodes
template <int D>
__device__ __forceinline__ void clear_list_grid(const Topo& T, int wg, int nwg) {
  size_t total = 0;
  for (int l = 0; l < T.L; ++l) total += (size_t)T.scount[l] * Pow4<D>::v;
  const size_t st = (size_t)nwg * blockDim.x;
  for (size_t t = wg * (size_t)blockDim.x + threadIdx.x; t < total; t += st) {
    size_t acc = 0;
    for (int l = 0; l < T.L; ++l) {
      const size_t cnt = (size_t)T.scount[l] * Pow4<D>::v;
      if (t < acc + cnt) {
        const size_t q = t - acc;
        const int b = T.lv[l].slist[q / Pow4<D>::v];
        T.lv[l].grid[(size_t)b * Pow4<D>::v + q % Pow4<D>::v] = make_float4(0.f, 0.f, 0.f, 0.f);
        break;
      }
      acc += cnt;
    }
  }
}

// The octree's marks, before a rebuild: cell masks, corner and split maps, node mask, block
// flags, the compaction counters.  Only flagged blocks can hold marks: every cell, corner-map
// or split-map write flags its leaf block (cflag), and every node mark lies in a finest-level
// block whose
// grid-phase flag (sflag) is set (the node marks set it, it is saved after refine and
// restored, never cleared, by bucketing).  One thread per global block zeroes the block's
// cells, corners and splits if cflag is set and, for finest blocks with sflag set, its nodes;
// then both flags.  A thread reads and clears only its own block's flags.
template <int D>
__device__ __forceinline__ void clear_octree(const Topo& T, int wg, int nwg) {
  constexpr int WPB = Pow4<D>::v / 4;  // 32-bit words per leaf block of byte masks
  const int nb = T.total_blocks, fbase = T.lv[T.L - 1].gbase;
  uint8_t AGIMP_GLOBAL* sflag = T.lv[0].sflag;
  uint32_t AGIMP_GLOBAL* cells = (uint32_t AGIMP_GLOBAL*)T.lv[0].cell;  // block g at word g * WPB
  uint32_t AGIMP_GLOBAL* corners = (uint32_t AGIMP_GLOBAL*)T.lv[0].corner;  // the same layout
  uint32_t AGIMP_GLOBAL* splits = (uint32_t AGIMP_GLOBAL*)T.lv[0].split;    // the same layout
  uint32_t AGIMP_GLOBAL* nodes = (uint32_t AGIMP_GLOBAL*)T.node;
  const int st = nwg * blockDim.x;
  for (int g = wg * blockDim.x + threadIdx.x; g < nb; g += st) {
    if (T.cflag[g]) {
#pragma unroll
      for (int w = 0; w < WPB; ++w) cells[(size_t)g * WPB + w] = 0u;
#pragma unroll
      for (int w = 0; w < WPB; ++w) corners[(size_t)g * WPB + w] = 0u;
#pragma unroll
      for (int w = 0; w < WPB; ++w) splits[(size_t)g * WPB + w] = 0u;
      T.cflag[g] = 0;
    }
    if (sflag[g]) {
      if (g >= fbase) {
#pragma unroll
        for (int w = 0; w < WPB; ++w) nodes[(size_t)(g - fbase) * WPB + w] = 0u;
      }
      sflag[g] = 0;
    }
  }
  if (wg == 0 && threadIdx.x < 4) T.ccount[threadIdx.x] = 0;
}

// the bucketing pass's preparation that does not depend on the rebuild: zero the per-block
// particle counters and the frozen counter, count the substep
__device__ __forceinline__ void clear_bucket_counts(const SceneDesc& sd, int wg, int nwg) {
  fill_zero((uint32_t*)sd.S.bcnt, (size_t)sd.T.total_blocks, wg, nwg);
  fill_zero((uint32_t*)sd.S.frozen, 1, wg, nwg);
  if (wg == 0 && threadIdx.x == 0) atomicAdd(&sd.cnt[CNT_STEP], 1ull);
}

// Before refine: the octree's marks (clear_octree).
// chained (the bucketing pass follows in the same graph): also zero the grid of the previous
// grid-phase lists (k_clear_sgrid's work) and leave their counters, which the bucketing pass
// then neither re-reads for its clear nor needs; the level map resets them.  It also does the
// bucketing pass's preparation that does not depend on the rebuild (clear_bucket_counts); the
// level map saves the octree's grid-phase block flags.
template <int D>
__global__ void k_clear_refine(const SceneDesc* __restrict__ Dsc, int chained) {
  AGIMP_SCENE
  if (chained) {
    clear_list_grid<D>(T, blockIdx.x, gridDim.x);
    clear_bucket_counts(sd, blockIdx.x, gridDim.x);
  }
  clear_octree<D>(T, blockIdx.x, gridDim.x);
  if (!chained && blockIdx.x == 0 && threadIdx.x < MAXL) T.scount[threadIdx.x] = 0;
}

// the valid-node block flags of the octree: save after refine (to_saved) or restore
__global__ void k_copy_flags(const SceneDesc* __restrict__ Dsc, int to_saved) {
  AGIMP_SCENE
  uint32_t* a = (uint32_t*)T.lv[0].sflag;
  uint32_t* v = sd.vflag;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < sd.w_flags; i += st) {
    if (to_saved) v[i] = a[i];
    else a[i] = v[i];
  }
}

constexpr int COMPACT_T = 256;  // threads per workgroup of k_compact_cells (4 flags each)

// append the ids of set flags to list: each thread reads one word of 4 flags, a workgroup
// reserves its slots with one atomic, and ids are written in increasing order within it
__global__ void __launch_bounds__(COMPACT_T) k_compact_cells(const SceneDesc* __restrict__ Dsc, int which) {
  AGIMP_SCENE
  __shared__ int wsum[COMPACT_T / 64];
  __shared__ int wgbase;
  const int n = T.total_blocks;
  const int nw = (n + 3) >> 2;
  int* list = T.clist;
  const int w = blockIdx.x * COMPACT_T + threadIdx.x;
  const uint32_t word = w < nw ? ((const uint32_t AGIMP_GLOBAL*)T.cflag)[w] : 0u;
  const int g0 = w * 4;
  int c = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) c += (((word >> (8 * b)) & 0xffu) != 0u) && g0 + b < n;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  int incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(incl, o);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
#pragma unroll
    for (int i = 0; i < COMPACT_T / 64; ++i) {
      const int v = wsum[i];
      wsum[i] = tot;
      tot += v;
    }
    wgbase = tot ? atomicAdd(T.ccount + which, tot) : 0;
  }
  __syncthreads();
  int idx = wgbase + wsum[wave] + incl - c;
#pragma unroll
  for (int b = 0; b < 4; ++b)
    if ((((word >> (8 * b)) & 0xffu) != 0u) && g0 + b < n) list[idx++] = g0 + b;
}

// ------------------------------------------------------------------------------ clear
template <int D>
__global__ void k_clear_sgrid(const SceneDesc* __restrict__ Dsc) {
  AGIMP_SCENE
  clear_list_grid<D>(T, blockIdx.x, gridDim.x);
}

// before a bucketing pass that does not follow a rebuild in the same graph, in one launch:
// zero the grid of the previous grid-phase lists (read here, so the list counters are zeroed
// later, by the level map), the block counters and the frozen counter; restore the octree's
// valid-node block flags; count the substep.  (After a rebuild, k_clear_refine and k_levelmap
// do this work, chained.)
template <int D>
__global__ void k_bucket_prep(const SceneDesc* __restrict__ Dsc) {
  AGIMP_SCENE
  const size_t st = (size_t)gridDim.x * blockDim.x;
  const size_t t0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  clear_list_grid<D>(T, blockIdx.x, gridDim.x);
  clear_bucket_counts(sd, blockIdx.x, gridDim.x);
  uint32_t* a = (uint32_t*)T.lv[0].sflag;
  uint32_t* v = sd.vflag;
  for (size_t i = t0; i < sd.w_flags; i += st) a[i] = v[i];
}

// ------------------------------------------------------------------------------ activation
// adaptive_gimp.py:87-96: OR ACTIVATED into cell_mask[g_p][base + dI], dI in {0,1}^d, for
// the particle at x with request level g (live: g in [0, L)) and level descriptor Lg.  Every
// lane of the wavefront calls it (the duplicate test reads the left lane).
template <int D>
__device__ __forceinline__ void activate_requests(const Topo& T, unsigned long long* cnt, const LevelDesc& Lg,
                                                  const float* x, int g, bool live) {
  int b[3] = {0, 0, 0};
  if (live)
#pragma unroll
    for (int a = 0; a < D; ++a) b[a] = base_cell(x[a], Lg, T.pow2);
  // particles arrive sorted by cell: a lane whose left neighbour requests the same cells
  // (same level and base cell) leaves the stores to it; out-of-range cells are still counted
  // per particle.  Every lane reaches the shuffle.
  bool packs = live;  // base cells in [-2, 2^20 - 2): the 20-bit fields of the key are exact
#pragma unroll
  for (int a = 0; a < D; ++a) packs = packs && b[a] >= -2 && b[a] < (1 << 20) - 2;
  const unsigned long long key =
      packs ? (((unsigned long long)g << 60) | ((unsigned long long)(b[0] + 2) << 40) |
               ((unsigned long long)(b[1] + 2) << 20) | (unsigned long long)(D == 3 ? b[2] + 2 : 0))
            : ~(unsigned long long)__lane_id();
  const unsigned long long left = __shfl_up(key, 1);
  const bool dup = __lane_id() > 0 && left == key;
  if (!live) return;
  // corners whose block equals that of the corner with a lower bit on an axis where b and b+1
  // share a block flag nothing new
  int cross = 0;
#pragma unroll
  for (int a = 0; a < D; ++a)
    if (((b[a] + PAD) >> 2) != ((b[a] + 1 + PAD) >> 2)) cross |= 1 << (D - 1 - a);
  // the corners' addresses from per-axis terms (the block-major index is their sum)
  uint32_t t0[3], t1[3];
  bool in0[3], in1[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const int j = b[a] + PAD;
    in0[a] = j >= 0 && j < Lg.E;
    in1[a] = j + 1 >= 0 && j + 1 < Lg.E;
    t0[a] = bidx_axis<D>(Lg.NB, a, j);
    t1[a] = bidx_axis<D>(Lg.NB, a, j + 1);
  }
#pragma unroll
  for (int t = 0; t < (1 << D); ++t) {
    bool in = true;
    uint32_t q = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const bool hi = (t >> (D - 1 - a)) & 1;
      in = in && (hi ? in1[a] : in0[a]);
      q += hi ? t1[a] : t0[a];
    }
    if (!in) { count(cnt, CNT_OOB); continue; }
    if (dup) continue;
    Lg.cell[q] = ACTIVATED;  // cells only ever receive ACTIVATED in this phase
    if ((t & ~cross) == 0) T.cflag[Lg.gbase + (int)(q / Pow4<D>::v)] = 1;
  }
}

template <int D>
__global__ void k_activate_dynamic(const SceneDesc* __restrict__ Dsc, int par) {
  AGIMP_SCENE
  if (sd.static_mode) return;
  const PBuf& P = sd.P[par];
  const int n = sd.n;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  // the position is loaded beside the request level, not after the test on it
  float x[3] = {0.f, 0.f, 0.f};
  if (k < n)
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = pload(pfield(P, a), (uint32_t)k * 4u);
  const int g = k < n ? P.g[k] : -1;
  const bool live = g >= 0 && g < T.L;
  // particles are sorted by level, so a wavefront's live lanes nearly always request one
  // level: its descriptor then comes from scalar loads, not a per-lane load round
  const int g0 = __builtin_amdgcn_readfirstlane(live ? g : -1);
  if (__ballot(live && g != g0) == 0ull) {
    const LevelDesc Lg = T.lv[g0 >= 0 ? g0 : 0];
    activate_requests<D>(T, sd.cnt, Lg, x, g, live);
  } else {
    const LevelDesc Lg = T.lv[live ? g : 0];
    activate_requests<D>(T, sd.cnt, Lg, x, g, live);
  }
}

// adaptive_gimp.py:72-85 with the grid initializer given as finest-cell level maps
template <int D>
__device__ __forceinline__ int static_gi(const Topo& T, const int8_t* ret, const int8_t* act, const int* I,
                                         unsigned long long* cnt) {
  const LevelDesc F = T.lv[T.L - 1];
  int J[3];
#pragma unroll
  for (int a = 0; a < D; ++a) J[a] = I[a] + PAD;
  if (!in_box<D>(F.E, J)) { count(cnt, CNT_OOB); return T.L - 1; }
  size_t r = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) r = r * F.E + J[a];
  int al = act[r];
  if (al >= 0 && al < T.L) {
    int P[3];
#pragma unroll
    for (int a = 0; a < D; ++a) P[a] = I[a] >> (T.L - 1 - al);
    int64_t q = lidx<D>(T.lv[al], P);
    if (q < 0) {
      count(cnt, CNT_OOB);
    } else {
      T.lv[al].cell[q] = ACTIVATED;
      mark_cell_block<D>(T, al, q);
    }
  }
  return ret[r];
}

template <int D>
__global__ void k_activate_static(const SceneDesc* __restrict__ Dsc, int par) {
  AGIMP_SCENE
  if (!sd.static_mode) return;
  const PBuf& P = sd.P[par];
  const int n = sd.n;
  const int8_t* ret = sd.st_ret;
  const int8_t* act = sd.st_act;
  unsigned long long* cnt = sd.cnt;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const LevelDesc F = T.lv[T.L - 1];
  float x[3];
#pragma unroll
  for (int a = 0; a < D; ++a) x[a] = pload(pfield(P, a), (uint32_t)k * 4u);
  int fb[3];
#pragma unroll
  for (int a = 0; a < D; ++a) fb[a] = base_cell(x[a], F, T.pow2);
  constexpr int tot = (D == 2) ? 16 : 64;
  for (int t = 0; t < tot; ++t) {
    int I[3];
    int r = t;
#pragma unroll
    for (int a = D - 1; a >= 0; --a) { I[a] = fb[a] + (r & 3) - 1; r >>= 2; }
    static_gi<D>(T, ret, act, I, cnt);
  }
  int level = static_gi<D>(T, ret, act, fb, cnt);
  if (level < 0 || level >= T.L) { count(cnt, CNT_OOB); return; }
  int b[3];
#pragma unroll
  for (int a = 0; a < D; ++a) b[a] = base_cell(x[a], T.lv[level], T.pow2);
#pragma unroll
  for (int t = 0; t < (1 << D); ++t) {
    int I[3];
#pragma unroll
    for (int a = 0; a < D; ++a) I[a] = (b[a] + ((t >> (D - 1 - a)) & 1)) << (T.L - 1 - level);
    static_gi<D>(T, ret, act, I, cnt);
  }
}

// ------------------------------------------------------------------------------ culling
// adaptive_gimp.py:98-118 for all levels at once, in one pass.  Every cell of level l >= 1
// holding ACTIVATED marks CULLING on its ancestors from its finest activated ancestor (level
// l0) down to level l-1 (:98-109; ACTIVATED bits do not change here), and the 2^d children of
// each of those ancestors are split-activated (:111-118): every CULLING cell is such an
// ancestor, so this reaches the children of all of them.  The split goes to a map of its own
// (LevelDesc::split) and k_node_A ORs ACTIVATED into those cells before anything compares a
// cell with ACTIVATED again; so the ancestor searches here see only the activation's bits, as
// in the reference, where the split follows all culling marks.  The culled cell itself keeps
// A|C (it compares unequal to ACTIVATED from here on).  Children of one ancestor lie in one
// leaf block (2P and 2P+1 share a block), flagged once.  A cull's children always lie in the
// padded box (the finer boxes cover less than the coarser ones, and a culled cell has a child
// in the box), so the OOB count of a child outside (per cull here, per culled cell in the
// reference) is never reached.
template <int D>
__global__ void k_cull(const SceneDesc* __restrict__ Dsc) {
  AGIMP_SCENE
  if (sd.static_mode) return;
  const int* list = T.clist;
  const int* countp = T.ccount + 0;
  unsigned long long* cnt = sd.cnt;
  const int nlist = *countp;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nlist * Pow4<D>::v; t += gridDim.x * blockDim.x) {
    int l, I[3];
    uint32_t q;
    if (!decode_list_wave<D>(T, list, nlist, t, l, I, q) || l < 1) continue;
    if (!(T.lv[l].cell[q] & ACTIVATED)) continue;
    int l0 = -1;
    for (int lq = 0; lq < l; ++lq) {
      int P[3];
#pragma unroll
      for (int a = 0; a < D; ++a) P[a] = I[a] >> (l - lq);
      int64_t qq = lidx<D>(T.lv[lq], P);
      if (qq < 0) { count(cnt, CNT_OOB); continue; }
      if (T.lv[lq].cell[qq] & ACTIVATED) l0 = lq;
    }
    if (l0 < 0) continue;
    for (int lq = l0; lq < l; ++lq) {
      int P[3];
#pragma unroll
      for (int a = 0; a < D; ++a) P[a] = I[a] >> (l - lq);
      int64_t qq = lidx<D>(T.lv[lq], P);
      if (qq < 0) continue;
      T.lv[lq].cell[qq] |= CULLING;
      mark_cell_block<D>(T, lq, qq);
      const LevelDesc Lc = T.lv[lq + 1];
      int64_t qc = -1;
#pragma unroll
      for (int u = 0; u < (1 << D); ++u) {
        int Jc[3];
#pragma unroll
        for (int a = 0; a < D; ++a) Jc[a] = 2 * P[a] + ((u >> (D - 1 - a)) & 1);
        const int64_t qu = lidx<D>(Lc, Jc);
        if (qu < 0) { count(cnt, CNT_OOB); continue; }
        Lc.split[qu] = 1;
        qc = qu;
      }
      if (qc >= 0) mark_cell_block<D>(T, lq + 1, qc);
    }
  }
}

// ------------------------------------------------------------------------------ node marks
// The level-l corner map of exactly-ACTIVATED level-l cell I: its 2^d corners inside
// the padded box get 1, and the leaf blocks they reach beyond the cell's own get their cflag
// (so the next rebuild's clear zeroes them; the lists of this rebuild are already built).
template <int D>
__device__ __forceinline__ void mark_level_corners(const Topo& T, int l, const int* I) {
  const LevelDesc Ll = T.lv[l];
  uint32_t t[3][2];
  bool in1[3], cross[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const int j = I[a] + PAD;  // in [0, E): the cell lies in the box
    t[a][0] = bidx_axis<D>(Ll.NB, a, j);
    t[a][1] = bidx_axis<D>(Ll.NB, a, j + 1);
    in1[a] = j + 1 < Ll.E;
    cross[a] = (j & 3) == 3;
  }
#pragma unroll
  for (int u = 0; u < (1 << D); ++u) {
    bool in = true, other = false;
    uint32_t qq = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      const int k = (u >> (D - 1 - a)) & 1;
      in = in && (k == 0 || in1[a]);
      other = other || (k == 1 && cross[a]);
      qq += t[a][k];
    }
    if (!in) continue;
    Ll.corner[qq] = CORNER_HIT;
    if (other) T.cflag[Ll.gbase + (int)(qq / Pow4<D>::v)] = 1;
  }
}

// adaptive_gimp.py:142-153: corners of exactly-ACTIVATED cells get ACTIVATED; also the
// level-l corner maps of the level map.  The last `ngrid` workgroups instead zero the grid of
// the previous substep's grid-phase lists (when no earlier launch of this rebuild did).
template <int D>
__global__ void k_node_A(const SceneDesc* __restrict__ Dsc, int ngrid) {
  AGIMP_SCENE
  const int nwg = gridDim.x - ngrid;
  if ((int)blockIdx.x >= nwg) {
    clear_list_grid<D>(T, blockIdx.x - nwg, ngrid);
    return;
  }
  const int* list = T.clist;
  const int* countp = T.ccount + 2;
  unsigned long long* cnt = sd.cnt;
  const int nlist = *countp;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nlist * Pow4<D>::v; t += nwg * blockDim.x) {
  int l, I[3];
  uint32_t q;
  if (!decode_list_wave<D>(T, list, nlist, t, l, I, q)) continue;
  uint8_t c = T.lv[l].cell[q];
  if (T.lv[l].split[q]) {  // split-activated by the culling pass: the cell gets ACTIVATED now
    c |= ACTIVATED;
    T.lv[l].cell[q] = c;
  }
  if (c != ACTIVATED) continue;
  const Corners<D> C(T, l, I);
  uint32_t sel = 0;
#pragma unroll
  for (int u = 0; u < (1 << D); ++u) {
    uint32_t qq;
    if (!C.addr(u, qq)) { count(cnt, CNT_OOB); continue; }
    T.node[qq] = ACTIVATED;  // node_mask only receives ACTIVATED in this phase
    sel |= 1u << u;
  }
  mark_valid_corners<D>(T, l, I, sel);
  mark_level_corners<D>(T, l, I);
  }
}

// The face logic of adaptive_gimp.py:155-169 for the 2d faces of cell I at once: lp[j] is
// the level of the activated cell that holds face j's neighbour I +- e_{j/2} (the hierarchical
// check, the finest such level), or -1 if the face is skipped (neighbour outside the level,
// none activated, or the neighbour itself activated at level l).  Per level, the neighbours' cells differ from I's only on one axis, so their
// addresses are sums of the cell's own per-axis terms and one neighbour term.
template <int D>
__device__ __forceinline__ void face_parents(const Topo& T, int l, const int* I, int* lp) {
  const int n = T.lv[l].n;
  bool own[3], nb_in[3][2];  // I_a in [0, n); I_a + 1 / I_a - 1 in [0, n)
#pragma unroll
  for (int a = 0; a < D; ++a) {
    own[a] = I[a] >= 0 && I[a] < n;
    nb_in[a][0] = I[a] + 1 >= 0 && I[a] + 1 < n;
    nb_in[a][1] = I[a] - 1 >= 0 && I[a] - 1 < n;
  }
  bool inr[2 * D];
#pragma unroll
  for (int j = 0; j < 2 * D; ++j) {
    bool ok = nb_in[j >> 1][j & 1];
#pragma unroll
    for (int a = 0; a < D; ++a)
      if (a != (j >> 1)) ok = ok && own[a];
    inr[j] = ok;
    lp[j] = -1;
  }
  // unrolled over the level bound: the cell loads of all levels are independent and issue
  // together
#pragma unroll
  for (int lq = 0; lq < MAXL; ++lq) {
    if (lq <= l) {
      const int sft = l - lq;
      const LevelDesc Lq = T.lv[lq];
      uint32_t t0[3], tn[3][2];
      bool i0[3], in_n[3][2];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        const int c0 = (I[a] >> sft) + PAD, cp = ((I[a] + 1) >> sft) + PAD, cm = ((I[a] - 1) >> sft) + PAD;
        i0[a] = c0 >= 0 && c0 < Lq.E;
        in_n[a][0] = cp >= 0 && cp < Lq.E;
        in_n[a][1] = cm >= 0 && cm < Lq.E;
        t0[a] = bidx_axis<D>(Lq.NB, a, c0);
        tn[a][0] = bidx_axis<D>(Lq.NB, a, cp);
        tn[a][1] = bidx_axis<D>(Lq.NB, a, cm);
      }
#pragma unroll
      for (int j = 0; j < 2 * D; ++j) {
        const int ax = j >> 1, sd1 = j & 1;
        bool in = inr[j] && in_n[ax][sd1];
        uint32_t qq = tn[ax][sd1];
#pragma unroll
        for (int a = 0; a < D; ++a)
          if (a != ax) { in = in && i0[a]; qq += t0[a]; }
        if (in && Lq.cell[qq] == ACTIVATED) lp[j] = lq;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 2 * D; ++j)
    if (lp[j] == l) lp[j] = -1;  // cell_mask[l][I1] == ACTIVATED
}

// adaptive_gimp.py:170-174 (T-junction marks) and :176-183 (ghost marks) for all levels in
// one pass over the listed cells.  The ghost marks may run beside the T marks: they touch only
// cells that are not exactly ACTIVATED (every level in (lp, l] over the neighbour lies below
// its finest activated ancestor, at lp), which the T marks' face searches compare with
// ACTIVATED only, and only nodes still 0, while the T marks write nodes that the node A marks
// (the previous pass) already set.  Each cell's coarse-fine faces (3 bits each: activated
// parent level + 1, or 0) are found with its T marks, then each wavefront spreads the faces
// of its 64 cells over its lanes for the ghost marks, 64 faces per round.
// ghost marks of one coarse-fine face (neighbour I1 of level l, activated parent level lp)
template <int D>
__device__ __forceinline__ void mark_G_face(const Topo& T, unsigned long long* cnt, const int* I1, int l, int lp) {
  for (int l0 = lp + 1; l0 <= l; ++l0) {
    int P[3];
#pragma unroll
    for (int a = 0; a < D; ++a) P[a] = I1[a] >> (l - l0);
    int64_t qc = lidx<D>(T.lv[l0], P);
    if (qc < 0) { count(cnt, CNT_OOB); continue; }
    // |= GHOST as a plain store: the cell lies below the neighbour's finest exactly-ACTIVATED
    // ancestor (at lp), a leaf of the octree, so it holds no ACTIVATED bit (an activated
    // descendant would have culled that ancestor; static octrees have no finer activated cell
    // there by the choice of lp), no CULLING bit (set only on ancestors of activated cells) and
    // no COVERED bit yet (set after the marks): it holds 0 or GHOST.  No load to wait for.
    T.lv[l0].cell[qc] = GHOST;
    mark_cell_block<D>(T, l0, qc);
    const Corners<D> C(T, l0, P);
    uint32_t sel = 0;
#pragma unroll
    for (int u = 0; u < (1 << D); ++u) {
      uint32_t qq;
      if (!C.addr(u, qq)) { count(cnt, CNT_OOB); continue; }
      if (T.node[qq] == 0) {
        T.node[qq] = GHOST;
        sel |= 1u << u;
      }
    }
    mark_valid_corners<D>(T, l0, P, sel);
  }
}

// (workgroup wg of the nwg that take part)
template <int D>
__device__ __forceinline__ void mark_TG_wg(const SceneDesc& sd, int wg, int nwg) {
  const Topo& T = sd.T;
  const int* list = T.clist;
  unsigned long long* cnt = sd.cnt;
  const int nlist = T.ccount[2];
  const int N = nlist * Pow4<D>::v;
  const int lane = __lane_id();
  // the loop is uniform over each workgroup, so every lane takes part in the shuffles
  for (int base = wg * blockDim.x; base < N; base += nwg * blockDim.x) {
    const int t = base + threadIdx.x;
    int l = 0, I[3] = {0, 0, 0};
    uint32_t q;
    uint32_t word = 0;
    if (decode_list_wave<D>(T, list, nlist, t, l, I, q) && T.lv[l].cell[q] == ACTIVATED) {
      int lp[2 * D];
      face_parents<D>(T, l, I, lp);
      // the nodes of the face shared with neighbour I +- e_ax are corners of this cell (corner
      // bit 1 on axis ax for the + side, 0 for the - side); their node-mask addresses on the
      // finest level from per-axis terms
      const LevelDesc Lf = T.lv[T.L - 1];
      const int sh = T.lv[l].shift;
      uint32_t ct[3][2];
      bool cin[3][2];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const int j = ((I[a] + k) << sh) + PAD;
          cin[a][k] = j >= 0 && j < Lf.E;
          ct[a][k] = bidx_axis<D>(Lf.NB, a, j);
        }
#pragma unroll
      for (int j = 0; j < 2 * D; ++j) {
        if (lp[j] < 0) continue;
        word |= (uint32_t)(lp[j] + 1) << (3 * j);
        const int ax = j >> 1, cb = 1 - (j & 1);  // the face's corner bit on axis ax
        const int msk = (1 << (l - lp[j])) - 1;
#pragma unroll
        for (int u = 0; u < (1 << D); ++u) {
          int kk[3];
#pragma unroll
          for (int a = 0; a < D; ++a) kk[a] = (u >> (D - 1 - a)) & 1;
          if (kk[ax] != cb) continue;
          bool any = false, in = true;
          uint32_t qq = 0;
#pragma unroll
          for (int a = 0; a < D; ++a) {
            any = any || (((I[a] + kk[a]) & msk) != 0);
            in = in && cin[a][kk[a]];
            qq += ct[a][kk[a]];
          }
          if (!any) continue;
          if (!in) { count(cnt, CNT_OOB); continue; }
          // a T node lies on the face this cell shares with its neighbour, so it is a corner
          // of this exactly-activated cell: the node A marks have stored ACTIVATED there and
          // flagged its grid-phase blocks, and ghost marks never write a non-zero node, so the
          // byte holds A or A|T and |= T is the plain store of A|T (no read)
          T.node[qq] = ACTIVATED | T_JUNCTION;
        }
      }
    }
    // ghost marks of the wavefront's faces
    int nface = 0;
#pragma unroll
    for (int j = 0; j < 2 * D; ++j) nface += ((word >> (3 * j)) & 7u) != 0u;
    int incl = nface;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o);
      if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63);
    for (int f0 = 0; f0 < total; f0 += 64) {
      const int f = f0 + lane;
      // source lane: the first with incl > f
      int src = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const int probe = __shfl(incl, src + step - 1);
        if (probe <= f) src += step;
      }
      const uint32_t w = __shfl(word, src);
      const int k = f - (__shfl(incl, src) - __shfl(nface, src));  // rank of the face in its cell
      const int ls = __shfl(l, src);
      int I1[3];
#pragma unroll
      for (int a = 0; a < D; ++a) I1[a] = __shfl(I[a], src);
      if (f >= total) continue;
      int j = 0, seen = 0;
#pragma unroll
      for (int jj = 0; jj < 2 * D; ++jj) {
        const bool on = ((w >> (3 * jj)) & 7u) != 0u;
        if (on && seen == k) j = jj;
        seen += on;
      }
      const int lpj = (int)((w >> (3 * j)) & 7u) - 1;
      I1[j >> 1] += (j & 1) ? -1 : 1;
      mark_G_face<D>(T, cnt, I1, ls, lpj);
    }
  }
}

template <int D>
__global__ void k_mark_TG(const SceneDesc* __restrict__ Dsc) {
  mark_TG_wg<D>(Dsc[blockIdx.y], blockIdx.x, gridDim.x);
}

// Coverage of level-l cells by finer activated cells, level by level from the finest up (one
// launch per l = L-2 .. 0): a level-l cell without ACTIVATED whose 2^d children are all
// ACTIVATED or COVERED gets COVERED.  A parent's children lie in one leaf block of level l+1,
// so each thread owns one parent cell (no races); blocks that receive COVERED are flagged so
// the next coarser launch visits them.
template <int D>
__global__ void k_cover(const SceneDesc* __restrict__ Dsc, int l) {
  AGIMP_SCENE
  if (l + 1 >= T.L) return;
  const LevelDesc Lc = T.lv[l + 1];
  const LevelDesc Lp = T.lv[l];
  const int64_t tot = (int64_t)Lc.nblocks << D;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(t >> D);
    if (!T.cflag[Lc.gbase + b]) continue;
    const int pb = (int)(t & ((1 << D) - 1));
    int B[3], J[3];
    block_coords<D>(Lc.NB, b, B);
    bool all = true;
#pragma unroll
    for (int u = 0; u < (1 << D); ++u) {
      int Jc[3];
#pragma unroll
      for (int a = 0; a < D; ++a) Jc[a] = 4 * B[a] + 2 * ((pb >> (D - 1 - a)) & 1) + ((u >> (D - 1 - a)) & 1);
      all = all && (Lc.cell[bidx<D>(Lc.NB, Jc)] & (ACTIVATED | COVERED));
    }
    if (!all) continue;
    // padded parent coordinates: (padded child - PAD)/2 + PAD
#pragma unroll
    for (int a = 0; a < D; ++a) J[a] = ((4 * B[a] + 2 * ((pb >> (D - 1 - a)) & 1) - PAD) >> 1) + PAD;
    if (!in_box<D>(Lp.E, J)) continue;
    const uint32_t q = bidx<D>(Lp.NB, J);
    const uint8_t c = Lp.cell[q];
    if (c & ACTIVATED) continue;
    Lp.cell[q] = c | COVERED;
    T.cflag[Lp.gbase + (int)(q / Pow4<D>::v)] = 1;
  }
}

// After the coverage marks of a predicted octree: CORNER_FULL into the corner entries of every
// level whose in-box cells around the corner (cells N - d, d in {0,1}^d) all hold ACTIVATED or
// COVERED, the level map's fast test of the refinement criterion.  Cell N is one of them, so
// such corners lie in blocks with written cells (cflag set), and are cleared with those blocks.
// One thread per cell index of a flagged block; each corner entry has one writer.
template <int D>
__global__ void k_full_corners(const SceneDesc* __restrict__ Dsc) {
  AGIMP_SCENE
  constexpr int CB = Pow4<D>::v;  // cells per block
  const int64_t tot = (int64_t)T.total_blocks * CB;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int g = (int)(t / CB);
    if (!T.cflag[g]) continue;
    const int l = level_of_gbid(T, g);
    const LevelDesc Ll = T.lv[l];
    const int c = (int)(t % CB);
    int B[3], N[3];
    block_coords<D>(Ll.NB, g - Ll.gbase, B);
    bool in = true;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      N[a] = 4 * B[a] + ((c >> (2 * (D - 1 - a))) & 3);
      in = in && N[a] < Ll.E;
    }
    if (!in) continue;
    bool full = true;
#pragma unroll
    for (int u = 0; u < (1 << D); ++u) {
      int J[3];
      bool inb = true;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        J[a] = N[a] - ((u >> (D - 1 - a)) & 1);
        inb = inb && J[a] >= 0;
      }
      if (inb) full = full && (Ll.cell[bidx<D>(Ll.NB, J)] & (ACTIVATED | COVERED));
    }
    if (full) {
      const uint32_t q = bidx<D>(Ll.NB, N);
      Ll.corner[q] = Ll.corner[q] | CORNER_FULL;
    }
  }
}

// level-l cell I lies in a unit (cell of level ulev <= l) whose predicted level is >= l
template <int D>
__device__ __forceinline__ bool unit_covers(const Topo& T, int ulev, const int8_t* upred, int l, const int* I) {
  if (!upred || l < ulev) return false;
  int Iu[3];
#pragma unroll
  for (int a = 0; a < D; ++a) Iu[a] = I[a] >> (l - ulev);
  const int64_t q = lidx<D>(T.lv[ulev], Iu);
  return q >= 0 && (int)upred[q] >= l;
}

// ------------------------------------------------------------------------------ bucketing
constexpr int LM_T = 256, LM_P = 1;  // level map: threads per workgroup, particles per thread
constexpr int LM_TABLE_BITS = 8, LM_TABLE = 1 << LM_TABLE_BITS;  // >= particles per workgroup
static_assert(LM_TABLE >= LM_T * LM_P, "one table entry per particle at most");
constexpr int LM_EMPTY = -2;

// adaptive_gimp.py:189-216: fl_p = finest level with an exactly-ACTIVATED cell among the
// particle's 2^d cells.  Those cells are the ones of which the base cell + 1 is a corner, so
// per level one entry of the corner map answers.  lm_probe issues the loads of the entries of
// all levels for LM_P particles before any is used (inb: the corner lies in the padded box); a
// corner below the box has none of the cells in the box, and one beyond its high end (cells
// partly outside the box) is left to lm_finish, which tests the cells in the box.
template <int D, int NL>
__device__ __forceinline__ void lm_probe(const Topo& T, const float (*x)[3], uint32_t (*hv)[NL], uint32_t* inb,
                                         uint32_t* slow) {
#pragma unroll
  for (int p = 0; p < LM_P; ++p) inb[p] = slow[p] = 0;
  // NL: the most levels of the scenes of the launch; a scene with fewer probes its coarsest
  // level again and ignores the result (no branch that would split the loads)
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const bool has = T.L - 1 - i >= 0;
    const int l = max(T.L - 1 - i, 0);
    const LevelDesc Ll = T.lv[l];
#pragma unroll
    for (int p = 0; p < LM_P; ++p) {
      int N[3], Nc[3];
      bool lo = !has, hi = false;
#pragma unroll
      for (int a = 0; a < D; ++a) {
        N[a] = base_cell(x[p][a], Ll, T.pow2) + 1 + PAD;
        lo = lo || N[a] < 0;
        hi = hi || N[a] >= Ll.E;
        Nc[a] = min(max(N[a], 0), Ll.E - 1);
      }
      // every lane loads (a clamped address where the corner lies outside), and the byte stays
      // in a register of its own until every level's load is issued: a select or a branch on
      // it here would wait for it
      hv[p][i] = Ll.corner[bidx<D>(Ll.NB, Nc)];
      if (!lo && !hi) inb[p] |= 1u << i;
      if (!lo && hi) slow[p] |= 1u << i;
    }
  }
}

// The particle's (level, leaf block) at its level fl (descriptor Lf): its global block id, or
// -1 with an out-of-range count if its stencil leaves the padded box; stores its cell inside
// the block (the P2G sort key)
template <int D>
__device__ __forceinline__ int lm_block(const SceneDesc& sd, const LevelDesc& Lf, int k, const float* x) {
  int J[3];
  bool ok = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    J[a] = base_cell(x[a], Lf, sd.T.pow2) + PAD;
    ok = ok && J[a] >= 0 && J[a] + 2 < Lf.E;
  }
  if (!ok) {
    count(sd.cnt, CNT_OOB);
    return -1;
  }
  const int b = (D == 2) ? ((J[0] >> 2) * Lf.NB + (J[1] >> 2))
                         : (((J[0] >> 2) * Lf.NB + (J[1] >> 2)) * Lf.NB + (J[2] >> 2));
  int lc = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) lc = lc * 4 + (J[a] & 3);
  sd.S.lkey[k] = (uint8_t)lc;
  return Lf.gbase + b;
}

// The rest of the level map for the particle in slot k at x: the levels left to a cell test,
// fl, the refinement criterion check (when on), and the block.  Returns the global id of its
// (level, leaf block), or -1 (frozen: no such level, or its stencil leaves the padded box);
// stores its cell inside that block (the P2G sort key) and counts what the counters record.
template <int D, int NL>
__device__ __forceinline__ int lm_finish(const SceneDesc& sd, const PBuf& P, int k, const float* x, const uint32_t* hv,
                                         uint32_t inb, uint32_t slow) {
  const Topo& T = sd.T;
  unsigned long long* cnt = sd.cnt;
  int fl = -1, gbid = -1;
  bool hit[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) hit[i] = ((inb >> i) & 1u) && (hv[i] & CORNER_HIT);
  if (slow) {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      if (!((slow >> i) & 1u)) continue;
      const LevelDesc Ll = T.lv[T.L - 1 - i];
      int b[3];
#pragma unroll
      for (int a = 0; a < D; ++a) b[a] = base_cell(x[a], Ll, T.pow2) + PAD;
#pragma unroll
      for (int u = 0; u < (1 << D); ++u) {
        int J[3];
#pragma unroll
        for (int a = 0; a < D; ++a) J[a] = b[a] + ((u >> (D - 1 - a)) & 1);
        if (in_box<D>(Ll.E, J) && Ll.cell[bidx<D>(Ll.NB, J)] == ACTIVATED) hit[i] = true;
      }
    }
  }
  const int gp = sd.check ? P.g[k] : -1;  // level whose request cells are checked
  bool viol = false;
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    const int l = T.L - 1 - i;
    if (l < 0) continue;
    if (hit[i] && fl < 0) fl = l;
    // the criterion check: the in-box request cells of level gp must hold ACTIVATED or
    // COVERED, or lie in a unit predicted at gp or finer; a corner entry with CORNER_FULL
    // answers the first test for all of them
    if (l == gp && !(((inb >> i) & 1u) && (hv[i] & CORNER_FULL))) {
      const LevelDesc Ll = T.lv[l];
      int b[3];
#pragma unroll
      for (int a = 0; a < D; ++a) b[a] = base_cell(x[a], Ll, T.pow2);
#pragma unroll
      for (int u = 0; u < (1 << D); ++u) {
        int I[3], J[3];
#pragma unroll
        for (int a = 0; a < D; ++a) {
          I[a] = b[a] + ((u >> (D - 1 - a)) & 1);
          J[a] = I[a] + PAD;
        }
        if (in_box<D>(Ll.E, J) && !(Ll.cell[bidx<D>(Ll.NB, J)] & (ACTIVATED | COVERED)))
          viol = viol || !unit_covers<D>(T, sd.ulev, sd.upred, l, I);
      }
    }
  }
  if (viol) {
    count(cnt, CNT_VIOL);
    atomicMin(&cnt[CNT_FIRSTVIOL], cnt[CNT_STEP]);
  }
  // particles are sorted by level, so a wavefront's mapped lanes nearly always share fl: its
  // descriptor then comes from scalar loads, not a per-lane load round
  const int fl0 = __builtin_amdgcn_readfirstlane(fl);
  if (__ballot(fl >= 0 && fl != fl0) == 0ull) {
    if (fl >= 0) gbid = lm_block<D>(sd, T.lv[fl0], k, x);
  } else {
    if (fl >= 0) gbid = lm_block<D>(sd, T.lv[fl], k, x);
  }
  return gbid;
}

// Level mapping and the rank of every particle inside its (level, leaf block) bucket.  Each
// thread takes LM_P particles (slots a workgroup's LM_T apart, so each wavefront's 64 lanes
// hold consecutive slots), all loaded before any is used.  Ranks: runs of equal blocks inside
// each wavefront are merged per workgroup in an LDS table (local offsets from LDS atomics),
// then one global atomic per (workgroup, block) reserves the workgroup's range: consecutive
// wavefronts mostly fill the same block, and one returning global atomic per run made them
// queue on its counter.
// (workgroup wg of the launch's level-map part)
template <int D, int NL>
__device__ __forceinline__ void levelmap_wg(const SceneDesc& sd, int par, int wg) {
  const Topo& T = sd.T;
  const PBuf& P = sd.P[par];
  const int n = sd.n;
  int* gb = sd.S.gb;
  int* rk = sd.S.rk;
  int* bcnt = sd.S.bcnt;
  int* frozen = sd.S.frozen;
  const int k0 = wg * (LM_T * LM_P) + threadIdx.x;
  if (k0 - (int)threadIdx.x >= n) return;  // whole workgroup past this scene
  __shared__ int tkey[LM_TABLE], tcnt[LM_TABLE], tbase[LM_TABLE];
  for (int i = threadIdx.x; i < LM_TABLE; i += LM_T) {
    tkey[i] = LM_EMPTY;
    tcnt[i] = 0;
  }
  float x[LM_P][3];
#pragma unroll
  for (int p = 0; p < LM_P; ++p) {
    const int k = k0 + p * LM_T;
#pragma unroll
    for (int a = 0; a < 3; ++a) x[p][a] = 0.f;
    if (k < n)
#pragma unroll
      for (int a = 0; a < D; ++a) x[p][a] = pload(pfield(P, a), (uint32_t)k * 4u);
  }
  // the positions are consumed here, on the one path into the level loop: otherwise its
  // branches leave their loads pending at the joins
#pragma unroll
  for (int p = 0; p < LM_P; ++p) asm volatile("" ::"v"(x[p][0]), "v"(x[p][1]), "v"(x[p][2]));
  uint32_t hv[LM_P][NL], inb[LM_P], slow[LM_P];
  lm_probe<D, NL>(T, x, hv, inb, slow);
  int key[LM_P];  // block id, -1 frozen, LM_EMPTY past the end
#pragma unroll
  for (int p = 0; p < LM_P; ++p) {
    const int k = k0 + p * LM_T;
    key[p] = k < n ? lm_finish<D, NL>(sd, P, k, x[p], hv[p], inb[p], slow[p]) : LM_EMPTY;
    if (k < n) gb[k] = key[p];
  }
  __syncthreads();
  const int lane = __lane_id();
  const unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  int slot[LM_P], off[LM_P];
#pragma unroll
  for (int p = 0; p < LM_P; ++p) {
    const bool live = key[p] != LM_EMPTY;
    const int prev = __shfl_up(key[p], 1);
    const bool head = !live || lane == 0 || prev != key[p];
    const unsigned long long heads = __ballot(head);
    const int h = 63 - __clzll(heads & le);
    const unsigned long long above = heads & ~((h == 63) ? ~0ull : ((1ull << (h + 1)) - 1ull));
    const int nh = above ? (__ffsll((long long)above) - 1) : 64;
    int sl = 0, loc = 0;
    if (live && lane == h) {
      const uint32_t hs = ((uint32_t)(key[p] + 1) * 2654435761u) >> (32 - LM_TABLE_BITS);
      for (int probe = 0; probe < LM_TABLE; ++probe) {
        const int e = (int)((hs + probe) & (LM_TABLE - 1));
        const int old = atomicCAS(&tkey[e], LM_EMPTY, key[p]);
        if (old == LM_EMPTY || old == key[p]) { sl = e; break; }
      }
      loc = atomicAdd(&tcnt[sl], nh - h);
    }
    slot[p] = __shfl(sl, h);
    off[p] = __shfl(loc, h) + (lane - h);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < LM_TABLE; i += LM_T) {
    const int kk = tkey[i];
    if (kk != LM_EMPTY) tbase[i] = atomicAdd(kk >= 0 ? &bcnt[kk] : frozen, tcnt[i]);
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < LM_P; ++p) {
    const int k = k0 + p * LM_T;
    if (key[p] != LM_EMPTY) rk[k] = tbase[slot[p]] + off[p];
  }
}

// The last `ngrid` workgroups instead zero the grid of the previous substep's grid-phase lists
// (when the previous G2P launch did the rest of the bucketing preparation).
template <int D, int NL>
__global__ void __launch_bounds__(LM_T) k_levelmap(const SceneDesc* __restrict__ Dsc, int par, int ngrid) {
  const int nwg = gridDim.x - ngrid;
  if ((int)blockIdx.x >= nwg) clear_list_grid<D>(Dsc[blockIdx.y].T, blockIdx.x - nwg, ngrid);
  else levelmap_wg<D, NL>(Dsc[blockIdx.y], par, blockIdx.x);
}

// In a substep captured whole (the rebuild's last pass and the bucketing pass's first in one
// launch): the first ntg workgroups make the T and ghost marks, the others the level map.
// Independent: the level map reads the corner maps and cells that the node-marks pass left
// (ghost marks write GHOST only into cells that are not exactly ACTIVATED and into no corner
// map), and neither reads what the other writes; both are latency-bound and overlap.
template <int D, int NL>
__global__ void __launch_bounds__(LM_T) k_marks_levelmap(const SceneDesc* __restrict__ Dsc, int par, int ntg) {
  if ((int)blockIdx.x < ntg) mark_TG_wg<D>(Dsc[blockIdx.y], blockIdx.x, ntg);
  else levelmap_wg<D, NL>(Dsc[blockIdx.y], par, blockIdx.x - ntg);
}

// exclusive scan of (count, chunks) over all global blocks: 256 threads x 8 items per tile
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

__device__ __forceinline__ int2 add2(int2 a, int2 b) { return make_int2(a.x + b.x, a.y + b.y); }

// exclusive scan over the workgroup: shuffles inside each wavefront, then the SCAN_T / 64
// wavefront totals through LDS (sh holds at least SCAN_T / 64 entries)
__device__ __forceinline__ int2 block_excl_scan(int2 v, int2* sh, int2& total) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  int2 incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int ux = __shfl_up(incl.x, o), uy = __shfl_up(incl.y, o);
    if (lane >= o) { incl.x += ux; incl.y += uy; }
  }
  __syncthreads();  // sh may still be read by a previous call
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  int2 off = make_int2(0, 0);
  total = make_int2(0, 0);
#pragma unroll
  for (int i = 0; i < SCAN_T / 64; ++i) {
    const int2 t = sh[i];
    if (i < wave) off = add2(off, t);
    total = add2(total, t);
  }
  return make_int2(off.x + incl.x - v.x, off.y + incl.y - v.y);
}

// the (count, chunks) pairs of blocks base .. base + SCAN_I - 1 (0 past n), loaded without a
// branch per element so the loads issue together: 16-byte loads when all lie below n (base is
// a multiple of SCAN_I), else clamped element loads
__device__ __forceinline__ void load_cc(const int* bcnt, int base, int n, int CH, int2* v) {
  static_assert(SCAN_I % 4 == 0, "whole 16-byte loads");
  int c[SCAN_I];
  if (base + SCAN_I <= n) {
    const int4* p = (const int4*)(bcnt + base);
#pragma unroll
    for (int k = 0; k < SCAN_I / 4; ++k) {
      const int4 a = p[k];
      c[4 * k] = a.x; c[4 * k + 1] = a.y; c[4 * k + 2] = a.z; c[4 * k + 3] = a.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
      const int x = bcnt[min(base + i, n - 1)];
      c[i] = base + i < n ? x : 0;
    }
  }
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) v[i] = make_int2(c[i], (c[i] + CH - 1) / CH);
}

// save_flags (after a rebuild in the same graph): also save the octree's grid-phase block
// flags (the scatter adds to them next), the part of the bucketing preparation that must
// follow the rebuild's node, T and ghost marks.
__global__ void k_scan_reduce(const SceneDesc* __restrict__ Dsc, int save_flags) {
  AGIMP_SCENE
  if (save_flags) {
    const uint32_t* a = (const uint32_t*)T.lv[0].sflag;
    uint32_t* v = sd.vflag;
    const size_t st = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < sd.w_flags; i += st) v[i] = a[i];
  }
  const int* bcnt = sd.S.bcnt;
  const int n = T.total_blocks, CH = sd.CH;
  int2* tile_sum = sd.S.tiles;
  if ((int)blockIdx.x * SCAN_TILE >= n) return;
  __shared__ int2 sh[SCAN_T];
  int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_I;
  int2 v[SCAN_I];
  load_cc(bcnt, base, n, CH, v);
  int2 s = make_int2(0, 0);
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) s = add2(s, v[i]);
  int2 tot;
  block_excl_scan(s, sh, tot);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

// Final pass: each tile's offset is the sum of the totals of the tiles before it (a few
// hundred at most, read by this workgroup), so no separate pass scans the tile totals; the
// last tile also writes the grand totals (active particles, chunks).
__global__ void k_scan_final(const SceneDesc* __restrict__ Dsc) {
  AGIMP_SCENE
  const int* bcnt = sd.S.bcnt;
  const int n = T.total_blocks, CH = sd.CH;
  const int2* tile_sum = sd.S.tiles;
  int* boff = sd.S.boff;
  int* choff = sd.S.choff;
  const int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  // the grid-phase list counters: the block-list compaction in the P2G launch appends from
  // zero (every clear of the previous lists' grid has run by now)
  if (blockIdx.x == 0 && threadIdx.x < MAXL) T.scount[threadIdx.x] = 0;
  if ((int)blockIdx.x >= ntiles) return;
  __shared__ int2 sh[SCAN_T];
  // the preceding tiles' totals, 4 loads in flight per thread and round
  int2 o = make_int2(0, 0);
  const int nprev = (int)blockIdx.x;
  for (int i0 = threadIdx.x; i0 < nprev; i0 += 4 * SCAN_T) {
    int2 w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * SCAN_T;
      const int2 x = tile_sum[min(i, nprev - 1)];
      w[u] = i < nprev ? x : make_int2(0, 0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) o = add2(o, w[u]);
  }
  int2 tile_off;
  block_excl_scan(o, sh, tile_off);  // tile_off = the workgroup's total of o
  int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_I;
  int2 v[SCAN_I];
  load_cc(bcnt, base, n, CH, v);
  int2 s = make_int2(0, 0);
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) s = add2(s, v[i]);
  int2 tot;
  int2 run = add2(block_excl_scan(s, sh, tot), tile_off);
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) {
    if (base + i < n) { boff[base + i] = run.x; choff[base + i] = run.y; }
    run = add2(run, v[i]);
  }
  if ((int)blockIdx.x == ntiles - 1 && threadIdx.x == 0) {
    const int2 all = add2(tile_off, tot);
    sd.S.totals[0] = all.x;
    sd.S.totals[1] = all.y;
  }
}

// after the scan, in one launch: the sorted position of every particle (threads t < n) and,
// for every occupied leaf block (threads t < total_blocks), its P2G chunk descriptors and the
// blocks its 3^d stencils touch
template <int D>
__global__ void k_scatter_blocks(const SceneDesc* __restrict__ Dsc) {
  AGIMP_SCENE
  const int n = sd.n;
  const int* boff = sd.S.boff;
  const int* totals = sd.S.totals;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const int g = sd.S.gb[t];
    const int r = sd.S.rk[t];
    const int pos = (g >= 0) ? boff[g] + r : totals[0] + r;
    sd.S.perm[pos] = t;
    if (g >= 0) sd.S.pkey[pos] = sd.S.lkey[t];
    if (t == 0) {
      unsigned long long* cnt = sd.cnt;
      cnt[CNT_ACTIVE] += (unsigned long long)totals[0];
      cnt[CNT_FROZEN] += (unsigned long long)(n - totals[0]);
      cnt[CNT_CHUNKS] += (unsigned long long)totals[1];
    }
  }
  if (t >= T.total_blocks) return;
  const int c = sd.S.bcnt[t];
  if (c == 0) return;
  const int CH = sd.CH;
  const int l = level_of_gbid(T, t);
  const LevelDesc Ll = T.lv[l];
  const int b = t - Ll.gbase;
  const int nch = (c + CH - 1) / CH;
  const int o = sd.S.choff[t], s0 = boff[t];
  for (int j = 0; j < nch; ++j) sd.S.chunks[o + j] = make_int4(t, s0 + j * CH, min(CH, c - j * CH), 0);
  int B[3];
  block_coords<D>(Ll.NB, b, B);
#pragma unroll
  for (int u = 0; u < (1 << D); ++u) {
    int C2[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < D; ++a) { C2[a] = B[a] + ((u >> (D - 1 - a)) & 1); ok = ok && C2[a] < Ll.NB; }
    if (!ok) continue;
    int bb = (D == 2) ? (C2[0] * Ll.NB + C2[1]) : ((C2[0] * Ll.NB + C2[1]) * Ll.NB + C2[2]);
    Ll.sflag[bb] = 1;
  }
}

// compact the grid-phase block flags of every level into per-level lists (workgroup bx of
// COMPACT_T threads).  Only the grid phase reads the lists, and only the block flags that the
// scatter sets feed them, so the workgroups run in the P2G launch, after its own.
__device__ __forceinline__ void scompact_body(const Topo& T, int bx) {
  // each thread reads one word of 4 block flags (global block ids 4w .. 4w+3, levels laid out
  // one after the other); slots are counted per level in LDS and reserved with one atomic per
  // (workgroup, level)
  __shared__ int lcnt[MAXL], lbase[MAXL];
  const int n = T.total_blocks;
  const int nw = (n + 3) >> 2;
  if (bx * COMPACT_T >= nw) return;  // whole workgroup past this scene
  if (threadIdx.x < MAXL) lcnt[threadIdx.x] = 0;
  __syncthreads();
  const int w = bx * COMPACT_T + threadIdx.x;
  const uint32_t word = w < nw ? ((const uint32_t AGIMP_GLOBAL*)T.lv[0].sflag)[w] : 0u;
  int lv[4], off[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lv[k] = -1;
    const int g = 4 * w + k;
    if (((word >> (8 * k)) & 0xffu) != 0u && g < n) {
      lv[k] = level_of_gbid(T, g);
      off[k] = atomicAdd(&lcnt[lv[k]], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < MAXL) {
    const int c = lcnt[threadIdx.x];
    lbase[threadIdx.x] = c ? atomicAdd(&T.scount[threadIdx.x], c) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (lv[k] >= 0) T.lv[lv[k]].slist[lbase[lv[k]] + off[k]] = 4 * w + k - T.lv[lv[k]].gbase;
}

// ------------------------------------------------------------------------------ wave reductions
// DPP moves (GFX9 row_shr / row_bcast); lanes without a source keep `old`
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i(int src, int old) {
  return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWMASK, 0xf, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f(float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(src), CTRL, ROWMASK, 0xf, false));
}

// Which DPP steps of a segmented inclusive scan add their source: lanes sorted by key form
// contiguous segments; a step adds only when the source lane carries the same key.  With
// WAVE, the scan spans the wavefront (row_shr inside 16-lane rows, then the row_bcast steps
// across rows) and each segment has one tail, its last lane, holding the segment's sum.
// Without, it stays inside rows: the last lane of each row piece of a segment is a tail
// holding that piece's sum (a segment spanning rows has one tail per row).
struct SegMask {
  bool s1, s2, s4, s8, s16, s32, tail;
};

template <bool WAVE>
__device__ __forceinline__ SegMask seg_mask(int key) {
  SegMask m;
  m.s1 = dpp_i<0x111, 0xf>(key, -1) == key;
  m.s2 = dpp_i<0x112, 0xf>(key, -1) == key;
  m.s4 = dpp_i<0x114, 0xf>(key, -1) == key;
  m.s8 = dpp_i<0x118, 0xf>(key, -1) == key;
  m.s16 = m.s32 = false;
  const int nxt = __shfl_down(key, 1);
  if constexpr (WAVE) {
    // lane 15 of the row below (rows 1 and 3 take it) and lane 31 (rows 2 and 3)
    m.s16 = dpp_i<0x142, 0xa>(key, -1) == key;
    m.s32 = dpp_i<0x143, 0xc>(key, -1) == key;
    m.tail = (__lane_id() == 63) || (nxt != key);
  } else {
    m.tail = ((__lane_id() & 15) == 15) || (nxt != key);
  }
  return m;
}

// one step of the scan: lanes without a source read 0 (bound_ctrl), so the move folds into a
// DPP add (v_add_f32_dpp) and the step costs the add and a select
template <int CTRL>
__device__ __forceinline__ float seg_step(float v, bool on) {
  const float t = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
  const float s = v + t;
  return on ? s : v;
}

// Segmented inclusive scan: each tail ends up holding its segment's (or row piece's) sum.
template <bool WAVE>
__device__ __forceinline__ float seg_sum(float v, const SegMask& m) {
  v = seg_step<0x111>(v, m.s1);
  v = seg_step<0x112>(v, m.s2);
  v = seg_step<0x114>(v, m.s4);
  v = seg_step<0x118>(v, m.s8);
  if constexpr (WAVE) {
    // row_bcast over all rows, so the add folds into one v_add_f32_dpp; the rows that do not
    // take the broadcast lane (row 0 and 2 for lane 15, rows 0 and 1 for lane 31) have
    // s16 / s32 false and keep v
    v = seg_step<0x142>(v, m.s16);
    v = seg_step<0x143>(v, m.s32);
  }
  return v;
}

// The wavefront-wide scan of four quantities at once, each step one fused multiply-add with a
// DPP source: v += v[source lane] * on, with on = 1 where the step adds and 0 where it does
// not (bound_ctrl makes a missing source lane read 0).  The same sums, in the same order, as
// seg_sum<true> on each quantity (x + y * 1 is x + y in one rounding; x + y * 0 is x), in
// 24 instructions instead of 48 (the compiler does not fold a DPP move into the tied
// accumulator of v_fmac).  Inside the block each quantity's next step comes 4 instructions
// after its last write, more than the 2 wait states a DPP read of a VALU result needs; the
// leading s_nop 1 gives them for the writes before the block.  A non-finite value in one
// segment can reach its neighbour here (inf * 0), which the selecting scan avoids; such a
// particle is already counted as non-finite by G2P.
struct SegOn {
  float s1, s2, s4, s8, s16, s32;
};
__device__ __forceinline__ SegOn seg_on(const SegMask& m) {
  return SegOn{m.s1 ? 1.f : 0.f, m.s2 ? 1.f : 0.f, m.s4 ? 1.f : 0.f,
               m.s8 ? 1.f : 0.f, m.s16 ? 1.f : 0.f, m.s32 ? 1.f : 0.f};
}
#define AGIMP_SEG_STEP(ctrl, on)                                                      \
  "v_fmac_f32_dpp %0, %0, " on " " ctrl " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t" \
  "v_fmac_f32_dpp %1, %1, " on " " ctrl " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t" \
  "v_fmac_f32_dpp %2, %2, " on " " ctrl " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t" \
  "v_fmac_f32_dpp %3, %3, " on " " ctrl " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
__device__ __forceinline__ void seg_sum4_wave(float& a, float& b, float& c, float& d, const SegOn& o) {
  asm volatile("s_nop 1\n\t"
               AGIMP_SEG_STEP("row_shr:1", "%4")
               AGIMP_SEG_STEP("row_shr:2", "%5")
               AGIMP_SEG_STEP("row_shr:4", "%6")
               AGIMP_SEG_STEP("row_shr:8", "%7")
               AGIMP_SEG_STEP("row_bcast:15", "%8")
               AGIMP_SEG_STEP("row_bcast:31", "%9")
               : "+v"(a), "+v"(b), "+v"(c), "+v"(d)
               : "v"(o.s1), "v"(o.s2), "v"(o.s4), "v"(o.s8), "v"(o.s16), "v"(o.s32));
}
#undef AGIMP_SEG_STEP

// ------------------------------------------------------------------------------ transfers
// Quadratic B-spline weights and their t-derivatives: for r = 0.5 the GIMP overlap
// integrals of utils.py:18-42 summed by adaptive_gimp.py:218-228 reduce to these.
__device__ __forceinline__ void bspline(float t, float* w, float* dw) {
  float a = 0.5f - t, c = 0.5f + t;
  w[0] = 0.5f * a * a;
  w[1] = 0.75f - t * t;
  w[2] = 0.5f * c * c;
  dw[0] = -a;
  dw[1] = -2.0f * t;
  dw[2] = c;
}

constexpr int P2G_T = 256;       // threads per P2G workgroup
constexpr int P2G_MAX_CH = 2048; // largest work chunk (particles of one leaf block)

// adaptive_gimp.py:230-250.  One work chunk = up to CH particles of one (level, leaf block):
//  1. the chunk is counting-sorted by cell inside the block (LDS histogram); the order is
//     written back to perm, so G2P stores particles cell-sorted for the next substep;
//  2. each lane computes its particle's 3^d node contributions; lanes of one cell are summed
//     with a segmented DPP scan and only the scan's tails add into the (4+2)^d LDS node tile.
//     3D: the scan spans the wavefront, so one lane per cell and wavefront writes, and each
//     wavefront owns a tile that it updates by plain read-add-write.  2D: the scan stays in
//     16-lane rows and the row tails add with LDS float atomics into one shared tile (at most
//     4 lanes of one instruction hit the same LDS word);
//  3. the tiles are summed and flushed with one float atomic per non-zero node component.
// 5 waves per SIMD: 96 VGPRs in 3D (one 8-byte spill) instead of 112 at 4 waves; the 2D
// kernel needs 62 either way.
template <int D>
__global__ void __launch_bounds__(P2G_T) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_p2g(const SceneDesc* __restrict__ Dsc, float dt_arg, int par, int nwg) {
  AGIMP_SCENE
  // workgroups past the first nwg build the grid-phase block lists
  if ((int)blockIdx.x >= nwg) {
    scompact_body(T, (int)blockIdx.x - nwg);
    return;
  }
  const PBuf& P = sd.P[par];
  int* perm = sd.S.perm;
  const int4* chunks = sd.S.chunks;
  const int* totals = sd.S.totals;
  Physics ph = sd.ph;
  if (dt_arg > 0.f) ph.dt = dt_arg;
  using PL = PLayout<D>;
  constexpr int TS = 6;
  constexpr int TN = (D == 2) ? TS * TS : TS * TS * TS;
  constexpr int NC = Pow4<D>::v;  // cells per leaf block
  constexpr int PER = P2G_MAX_CH / P2G_T;
  // 3D: one node tile per wavefront and one tail per cell and wavefront, so a tail adds its
  // sums with a plain read-add-write of the node's float4 (LDS operations of a wavefront run
  // in order; tails of one instruction belong to different cells, hence different nodes).
  // The next node's read must stay after this node's write, as another lane's node may be
  // the same: the compiled code keeps the 27 read/write pairs in program order (the node
  // loop's DPP scans sit between them; tests/test_p2g_isa_cpu.py checks the assembly); a
  // wavefront fence to force it costs 6 us.
  // 2D: one tile, row-local tails, LDS float atomics (the extra scan steps cost more there
  // than the 3 atomics per node they would save).
  constexpr bool WAVE = (D == 3);
  constexpr int NWV = WAVE ? P2G_T / 64 : 1;
  __shared__ float4 tile[NWV][TN];
  __shared__ int hist[NC];
  __shared__ int sidx[P2G_MAX_CH];
  const int nchunks = totals[1];
  const int tid = threadIdx.x;
  for (int c = blockIdx.x; c < nchunks; c += nwg) {
    const int4 ck = chunks[c];
    const int l = level_of_gbid(T, ck.x);
    const LevelDesc Ll = T.lv[l];
    int B[3];
    block_coords<D>(Ll.NB, ck.x - Ll.gbase, B);
    for (int i = tid; i < NWV * TN; i += P2G_T) (&tile[0][0])[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (tid < NC) hist[tid] = 0;
    __syncthreads();
    // 1. counting sort of the chunk by cell
    int myk[PER], mylc[PER], myrk[PER];
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      int i = r * P2G_T + tid;
      bool valid = i < ck.z;
      int k = 0, lc = 0;
      if (valid) {
        // the particle's cell inside the block, as the level map found it (its base cell at
        // this level minus the block's origin)
        k = perm[ck.y + i];
        lc = sd.S.pkey[ck.y + i];
      }
      myk[r] = k;
      mylc[r] = lc;
      myrk[r] = 0;
      if (r * P2G_T < ck.z) myrk[r] = wave_run_atomic_add(&hist[lc], valid);
    }
    __syncthreads();
    if (tid < 64) {  // exclusive scan of the NC counters by one wavefront
      int v = (tid < NC) ? hist[tid] : 0, incl = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(incl, o);
        if (tid >= o) incl += u;
      }
      if (tid < NC) hist[tid] = incl - v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      int i = r * P2G_T + tid;
      if (i < ck.z) {
        int dst = hist[mylc[r]] + myrk[r];
        sidx[dst] = myk[r];
        perm[ck.y + dst] = myk[r];
      }
    }
    __syncthreads();
    // 2. contributions, reduced per cell
    const float scale = -ph.dt * Ll.inv_dx * ph.p_vol;  // -dt (0.5/r)^d / dx * V_p with r = 0.5
    for (int base = 0; base < ck.z; base += P2G_T) {
      // a wavefront whose lanes all lie past the chunk's end skips the pass (uniform per
      // wavefront; the pass has no barrier)
      if (base + (tid & ~63) >= ck.z) continue;
      const int i = base + tid;
      const bool valid = i < ck.z;
      const int k = valid ? sidx[i] : 0;
      const uint32_t ko = (uint32_t)k * 4u;  // byte offset of the slot in each field
      float x[3] = {0.f, 0.f, 0.f}, v[3] = {0.f, 0.f, 0.f}, Cm[3][3] = {}, tau[6] = {}, mp = 0.f;
      if (valid) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
          x[a] = pload(pfield(P, PL::X + a), ko);
          v[a] = pload(pfield(P, PL::V + a), ko);
        }
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int q = 0; q < D; ++q) Cm[a][q] = pload(pfield(P, PL::C + a * D + q), ko);
#pragma unroll
        for (int q = 0; q < PL::NTAU; ++q) tau[q] = pload(pfield(P, PL::TAU + q), ko);
        mp = pload(pfield(P, PL::M), ko);
      }
      int lb[3] = {0, 0, 0}, key = 0;
      float w[3][3], dw[3][3], tt[3];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        float qx = scaled(x[a], Ll, T.pow2);
        int b = (int)(qx + 0.5f) - 1;
        tt[a] = qx - (float)(b + 1);
        lb[a] = valid ? (b + PAD - 4 * B[a]) : 0;
        key = key * 4 + lb[a];
        bspline(tt[a], w[a], dw[a]);
      }
      if (!valid) key = 4096 + __lane_id();  // unique: never joins a neighbour's segment
      const SegMask sm = seg_mask<WAVE>(key);
      const bool writer = valid && sm.tail;
      float S[3][3];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int q = 0; q < D; ++q) S[a][q] = scale * tau[tau_idx<D>(a, q)];
      float vb[3];  // v + C (i - (t+1)) = vb + C i
#pragma unroll
      for (int a = 0; a < D; ++a) {
        float acc = v[a];
#pragma unroll
        for (int q = 0; q < D; ++q) acc -= Cm[a][q] * (tt[q] + 1.0f);
        vb[a] = acc;
      }
      if constexpr (D == 2) {
#pragma unroll
        for (int i0 = 0; i0 < 3; ++i0)
#pragma unroll
          for (int i1 = 0; i1 < 3; ++i1) {
            float wt = w[0][i0] * w[1][i1];
            float g0 = dw[0][i0] * w[1][i1], g1 = w[0][i0] * dw[1][i1];
            float mw = wt * mp;
            float p0 = mw * (vb[0] + Cm[0][0] * i0 + Cm[0][1] * i1) + S[0][0] * g0 + S[0][1] * g1;
            float p1 = mw * (vb[1] + Cm[1][0] * i0 + Cm[1][1] * i1) + S[1][0] * g0 + S[1][1] * g1;
            p0 = seg_sum<WAVE>(p0, sm);
            p1 = seg_sum<WAVE>(p1, sm);
            mw = seg_sum<WAVE>(mw, sm);
            if (writer) {
              float4& t = tile[0][(lb[0] + i0) * TS + (lb[1] + i1)];
              atomicAdd(&t.x, p0);
              atomicAdd(&t.y, p1);
              atomicAdd(&t.w, mw);
            }
          }
      } else {
        // p_a = mw (vb_a + C_a . i) + S_a . grad w with w = N0 N1 N2 (N_a, D_a: the weights along
        // axis a and their derivatives), regrouped by axis so that each node costs two fused
        // multiply-adds per component:
        //   p_a = N2 F_a(i0, i1) + N0 N1 G_a(i2),   F_a = N1 A_a(i0) + N0 B_a(i1),
        //   A_a(i0) = m N0 (vb_a + C_a0 i0) + S_a0 D0,  B_a(i1) = m N1 C_a1 i1 + S_a1 D1,
        //   G_a(i2) = m N2 C_a2 i2 + S_a2 D2
        // (i C written as sums: the compiler may not drop a product with the constant 0)
        const SegOn on = seg_on(sm);
        float A[3][3], Bq[3][3], G[3][3];  // [i][a]
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const float mN0 = mp * w[0][i], mN1 = mp * w[1][i], mN2 = mp * w[2][i];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const float c0 = i == 1 ? Cm[a][0] : Cm[a][0] + Cm[a][0];
            const float c1 = i == 1 ? Cm[a][1] : Cm[a][1] + Cm[a][1];
            const float c2 = i == 1 ? Cm[a][2] : Cm[a][2] + Cm[a][2];
            A[i][a] = mN0 * (i == 0 ? vb[a] : vb[a] + c0) + S[a][0] * dw[0][i];
            Bq[i][a] = i == 0 ? S[a][1] * dw[1][i] : mN1 * c1 + S[a][1] * dw[1][i];
            G[i][a] = i == 0 ? S[a][2] * dw[2][i] : mN2 * c2 + S[a][2] * dw[2][i];
          }
        }
        // the wavefront's tile at the lane's base node: every node below is a constant offset
        float4* tb = &tile[tid >> 6][(lb[0] * TS + lb[1]) * TS + lb[2]];
#pragma unroll
        for (int i0 = 0; i0 < 3; ++i0)
#pragma unroll
          for (int i1 = 0; i1 < 3; ++i1) {
            const float N01 = w[0][i0] * w[1][i1];
            const float m01 = mp * N01;
            float F[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) F[a] = w[1][i1] * A[i0][a] + w[0][i0] * Bq[i1][a];
#pragma unroll
            for (int i2 = 0; i2 < 3; ++i2) {
              // the tile read is issued first (after the previous node's write, which it must
              // follow), so the node's arithmetic and scan hide its latency.  Every lane reads
              // and adds (its node lies in the tile; only tails write), and the sum is used on
              // every path (the asm below): a read whose wait sits only in the tails' branch
              // makes the next read wait for it again before reusing its registers.
              float4* tn = tb + (i0 * TS + i1) * TS + i2;
              float4 u = *tn;
              const float w2 = w[2][i2];
              float mw = m01 * w2;
              float p[3];
#pragma unroll
              for (int a = 0; a < 3; ++a) p[a] = w2 * F[a] + N01 * G[i2][a];
              seg_sum4_wave(p[0], p[1], p[2], mw, on);
              u.x += p[0]; u.y += p[1]; u.z += p[2]; u.w += mw;
              if (writer) *tn = u;
              // compiler barrier ("memory"): the next node's tile read stays after this write
              asm volatile("" ::"v"(u.x), "v"(u.y), "v"(u.z), "v"(u.w) : "memory");
            }
          }
      }
    }
    __syncthreads();
    // 3. flush: four lanes per node, one per component, so that a wavefront's float atomics
    //    go to the consecutive words of 16 nodes (a few cache lines per instruction) rather
    //    than to one component of 64 nodes
    for (int i = tid; i < 4 * TN; i += P2G_T) {
      const int nd = i >> 2, comp = i & 3;
      float val = ((const float*)&tile[0][nd])[comp];
#pragma unroll
      for (int w = 1; w < NWV; ++w) val += ((const float*)&tile[w][nd])[comp];
      if (val == 0.f) continue;  // also the z slot in 2D, which stays 0
      int J[3];
      if (D == 2) { J[0] = 4 * B[0] + nd / TS; J[1] = 4 * B[1] + nd % TS; }
      else { J[0] = 4 * B[0] + nd / (TS * TS); J[1] = 4 * B[1] + (nd / TS) % TS; J[2] = 4 * B[2] + nd % TS; }
      if (!in_box<D>(Ll.E, J)) continue;
      unsafeAtomicAdd((float*)&Ll.grid[bidx<D>(Ll.NB, J)] + comp, val);
    }
    __syncthreads();
  }
}

// node flags of level-l node I through the finest node_mask (0 outside the box)
template <int D>
__device__ __forceinline__ uint8_t nflag(const Topo& T, int l, const int* I) {
  int64_t q = nidx<D>(T, l, I);
  return q < 0 ? (uint8_t)0 : T.node[q];
}
// the finest node mask and its extents, read once per kernel into scalar registers
struct NodeView {
  const uint8_t AGIMP_GLOBAL* node;
  int NB, E;
};
__device__ __forceinline__ NodeView node_view(const Topo& T) {
  const LevelDesc Ln = T.lv[T.L - 1];
  return NodeView{T.node, Ln.NB, Ln.E};
}
// node flags of the node I of the level whose index shift to the finest level is `shift`
template <int D>
__device__ __forceinline__ uint8_t nflag(const NodeView& N, int shift, const int* I) {
  int J[3];
#pragma unroll
  for (int a = 0; a < D; ++a) J[a] = (I[a] << shift) + PAD;
  if (!in_box<D>(N.E, J)) return 0;
  return N.node[bidx<D>(N.NB, J)];
}
// nflag without a branch: the load always issues (at index 0 when I lies outside the box)
template <int D>
__device__ __forceinline__ uint8_t nflag_nb(const NodeView& N, int shift, const int* I) {
  int J[3];
#pragma unroll
  for (int a = 0; a < D; ++a) J[a] = (I[a] << shift) + PAD;
  const bool ok = in_box<D>(N.E, J);
  const uint8_t m = N.node[ok ? bidx<D>(N.NB, J) : 0u];
  return ok ? m : (uint8_t)0;
}
__device__ __forceinline__ bool f_valid(uint8_t M) { return M > 0; }
__device__ __forceinline__ bool f_act(uint8_t M) { return (M & (ACTIVATED | T_JUNCTION)) == ACTIVATED; }
__device__ __forceinline__ bool f_T(uint8_t M) { return (M & T_JUNCTION) != 0; }
__device__ __forceinline__ bool f_ghost(uint8_t M) { return (M & (ACTIVATED | GHOST | T_JUNCTION)) == GHOST; }

// node I of list entry t of level l's grid-phase list
template <int D>
__device__ __forceinline__ uint32_t list_node(const LevelDesc& Ll, int e, int w, int* I) {
  int b = Ll.slist[e];
  int B[3];
  block_coords<D>(Ll.NB, b, B);
  if (D == 2) {
    I[0] = B[0] * 4 + (w >> 2) - PAD;
    I[1] = B[1] * 4 + (w & 3) - PAD;
  } else {
    I[0] = B[0] * 4 + (w >> 4) - PAD;
    I[1] = B[1] * 4 + ((w >> 2) & 3) - PAD;
    I[2] = B[2] * 4 + (w & 3) - PAD;
  }
  return (uint32_t)b * Pow4<D>::v + w;
}

template <int D>
__device__ __forceinline__ bool in_range_nodes(const int* I, int n) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < D; ++a) ok = ok && I[a] >= 0 && I[a] <= n;
  return ok;
}

// adaptive_gimp.py:277-287 for node I of level l, whose value is g (momentum, mass):
// momentum -> velocity, gravity, walls (+ sticky walls / slab collider).  Only active nodes
// with mass change; returns whether g changed.  Levels >= 1 exclude the nodes prolongation
// overwrites (T, ghost, all-even): their callers never pass those.
template <int D>
__device__ __forceinline__ bool grid_update(const Topo& T, const LevelDesc& Ll, const Physics& ph, const int* I,
                                            uint8_t M, float4& g) {
  if (!(g.w > 0.f && f_act(M))) return false;
  float v[3] = {g.x, g.y, g.z};
#pragma unroll
  for (int a = 0; a < D; ++a) { v[a] /= g.w; v[a] += ph.gravity[a] * ph.dt; }
  bool wall_any = false;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    int f = I[a] << Ll.shift;
    bool wl = f < ph.boundary_gap || f > T.finest_size - ph.boundary_gap;
    wall_any = wall_any || wl;
    if (wl && ph.wall_mode == 0) v[a] = 0.f;
  }
  if (ph.wall_mode == 1 && wall_any)
#pragma unroll
    for (int a = 0; a < D; ++a) v[a] = 0.f;
  if (ph.slab_enable) {
    double y = (double)I[ph.slab_axis] * (double)Ll.dx;
    if (y >= (double)ph.slab_lo && y <= (double)ph.slab_hi)
#pragma unroll
      for (int a = 0; a < D; ++a) v[a] = 0.f;
  }
  g.x = v[0]; g.y = v[1];
  if (D == 3) g.z = v[2];
  return true;
}

// adaptive_gimp.py:289-301 as a gather onto each valid coarse node I of level l-1: sources
// are the fine nodes 2I - d, d in {-1,0,1}^d, with weight 2^-|d|_1.  Whether a fine node
// contributes, and its weight (d_a != 0 exactly where the fine coordinate is odd), depend on
// the fine node alone, so one wavefront per coarse leaf block first stages the weighted
// contributions of the 9^d fine nodes around the block in LDS (each fine node loaded once
// instead of once per coarse neighbour), then every lane sums its coarse node's 3^d entries
// in the same order as a direct gather.
template <int D>
__global__ void __launch_bounds__(256) k_restrict(const SceneDesc* __restrict__ Dsc, int l, float dt_arg) {
  AGIMP_SCENE
  if (l >= T.L) return;
  Physics ph = sd.ph;
  if (dt_arg > 0.f) ph.dt = dt_arg;
  const LevelDesc Lc = T.lv[l - 1];
  const LevelDesc Lf = T.lv[l];
  const NodeView N = node_view(T);
  constexpr int NW = 4, R = 9, RN = (D == 2) ? R * R : R * R * R;
  __shared__ float4 tile[NW][RN];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nblk = T.scount[l - 1];
  for (int e = blockIdx.x * NW + wv; e < nblk; e += gridDim.x * NW) {  // uniform per wavefront
    const int b = Lc.slist[e];
    int B[3], f0[3];
    block_coords<D>(Lc.NB, b, B);
#pragma unroll
    for (int a = 0; a < D; ++a) f0[a] = 2 * (4 * B[a] - PAD) - 1;  // first fine node (unpadded)
    // the lane's coarse node of this block (lanes < 4^d): its flags load beside the fine
    // nodes' flags, and its grid value beside their values
    int I[3], w[3];
    if (D == 2) { w[0] = (lane >> 2) & 3; w[1] = lane & 3; }
    else { w[0] = lane >> 4; w[1] = (lane >> 2) & 3; w[2] = lane & 3; }
#pragma unroll
    for (int a = 0; a < D; ++a) I[a] = 4 * B[a] - PAD + w[a];
    const bool mine = lane < Pow4<D>::v;
    const uint8_t Mc = (mine && in_range_nodes<D>(I, Lc.n)) ? nflag<D>(N, Lc.shift, I) : (uint8_t)0;
    // the lane's fine nodes i = lane + 64 k, staged in rounds: every node flag load is issued
    // first, then every value load the flags ask for (a loop with a flag and a value load per
    // node waits for both, node after node)
    constexpr int NPL = (RN + 63) / 64;
    uint32_t qf[NPL];
    uint8_t M[NPL];
    int nodd[NPL];
    bool use[NPL];
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
      const int i = lane + 64 * k;
      int If[3], r = i;
      nodd[k] = 0;
#pragma unroll
      for (int a = D - 1; a >= 0; --a) { If[a] = f0[a] + r % R; r /= R; nodd[k] += If[a] & 1; }
      use[k] = false;
      qf[k] = 0;
      M[k] = 0;
      if (i < RN && in_range_nodes<D>(If, Lf.n)) {
        const int64_t q = lidx<D>(Lf, If);
        if (q >= 0) {
          qf[k] = (uint32_t)q;
          M[k] = nflag<D>(N, Lf.shift, If);
          use[k] = true;
        }
      }
    }
    float4 g[NPL];
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
      use[k] = use[k] && ((nodd[k] == 0 && f_act(M[k])) || f_T(M[k]) || f_ghost(M[k]));
      g[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (use[k]) g[k] = Lf.grid[qf[k]];
    }
    const uint32_t qc = (uint32_t)b * Pow4<D>::v + lane;
    float4 gc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (f_valid(Mc)) gc = Lc.grid[qc];
#pragma unroll
    for (int k = 0; k < NPL; ++k) {
      const int i = lane + 64 * k;
      const float wgt = (nodd[k] == 0) ? 1.f : (nodd[k] == 1 ? 0.5f : (nodd[k] == 2 ? 0.25f : 0.125f));
      if (i < RN) tile[wv][i] = make_float4(g[k].x * wgt, g[k].y * wgt, g[k].z * wgt, g[k].w * wgt);
    }
    wave_sync_lds();
    {
      if (f_valid(Mc)) {
        constexpr int NS = (D == 2) ? 9 : 27;
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) {
          int r = s2, idx = 0;
          int rr[3];
#pragma unroll
          for (int a = 0; a < D; ++a) { rr[a] = 2 * w[a] + 1 - ((r % 3) - 1); r /= 3; }  // If - f0
#pragma unroll
          for (int a = 0; a < D; ++a) idx = idx * R + rr[a];
          const float4 v = tile[wv][idx];
          acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
        }
        // no source contributed: the node keeps its value (adding zeros changes nothing);
        // level 0 is complete after this restriction, so its grid update is applied here
        const bool add = acc.x != 0.f || acc.y != 0.f || acc.z != 0.f || acc.w != 0.f;
        if (add || l == 1) {
          float4 g = gc;
          g.x += acc.x; g.y += acc.y; g.z += acc.z; g.w += acc.w;
          const bool upd = (l == 1) && grid_update<D>(T, Lc, ph, I, Mc, g);
          if (add || upd) Lc.grid[qc] = g;
        }
      }
    }
    wave_sync_lds();  // the tile is refilled for the next block
  }
}

// k_restrict with a workgroup per coarse leaf block (3D): the 9^3 fine nodes around the block
// are staged three per thread, each node's flags and value loaded in one round without a
// branch (the value whatever the flags say, a read of the dense grid), together with the
// coarse nodes' flags and values; after one barrier the first wavefront sums its node's 3^d
// entries in the same order as a direct gather.
template <int D>
__global__ void __launch_bounds__(256) k_restrict_wg(const SceneDesc* __restrict__ Dsc, int l, float dt_arg) {
  AGIMP_SCENE
  if (l >= T.L) return;
  Physics ph = sd.ph;
  if (dt_arg > 0.f) ph.dt = dt_arg;
  const LevelDesc Lc = T.lv[l - 1];
  const LevelDesc Lf = T.lv[l];
  const NodeView N = node_view(T);
  constexpr int R = 9, RN = (D == 2) ? R * R : R * R * R, NT = 256, NPT = (RN + NT - 1) / NT;
  __shared__ float4 tile[RN];
  const int tid = threadIdx.x, lane = tid & 63;
  const int nblk = T.scount[l - 1];
  for (int e = blockIdx.x; e < nblk; e += gridDim.x) {  // uniform per workgroup
    const int b = __builtin_amdgcn_readfirstlane(Lc.slist[e]);
    int B[3], f0[3];
    block_coords<D>(Lc.NB, b, B);
#pragma unroll
    for (int a = 0; a < D; ++a) f0[a] = 2 * (4 * B[a] - PAD) - 1;  // first fine node (unpadded)
    uint8_t M[NPT];
    int nodd[NPT];
    bool use[NPT];
    float4 g[NPT];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int i = tid + NT * k;
      int If[3], r = i;
      nodd[k] = 0;
#pragma unroll
      for (int a = D - 1; a >= 0; --a) { If[a] = f0[a] + r % R; r /= R; nodd[k] += If[a] & 1; }
      const int64_t q = lidx<D>(Lf, If);
      use[k] = i < RN && in_range_nodes<D>(If, Lf.n) && q >= 0;
      M[k] = nflag_nb<D>(N, Lf.shift, If);
      g[k] = Lf.grid[use[k] ? (uint32_t)q : 0u];
    }
    // the first wavefront's coarse node (lanes < 4^d)
    int I[3], w[3];
    if (D == 2) { w[0] = (lane >> 2) & 3; w[1] = lane & 3; }
    else { w[0] = lane >> 4; w[1] = (lane >> 2) & 3; w[2] = lane & 3; }
#pragma unroll
    for (int a = 0; a < D; ++a) I[a] = 4 * B[a] - PAD + w[a];
    const bool mine = tid < Pow4<D>::v;
    const uint8_t Mcl = nflag_nb<D>(N, Lc.shift, I);
    const uint8_t Mc = (mine && in_range_nodes<D>(I, Lc.n)) ? Mcl : (uint8_t)0;
    const uint32_t qc = (uint32_t)b * Pow4<D>::v + (lane & (Pow4<D>::v - 1));
    const float4 gc = Lc.grid[qc];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int i = tid + NT * k;
      const bool u = use[k] && ((nodd[k] == 0 && f_act(M[k])) || f_T(M[k]) || f_ghost(M[k]));
      const float wgt = (nodd[k] == 0) ? 1.f : (nodd[k] == 1 ? 0.5f : (nodd[k] == 2 ? 0.25f : 0.125f));
      if (i < RN)
        tile[i] = u ? make_float4(g[k].x * wgt, g[k].y * wgt, g[k].z * wgt, g[k].w * wgt)
                    : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    if (f_valid(Mc)) {
      constexpr int NS = (D == 2) ? 9 : 27;
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) {
        int r = s2, idx = 0;
        int rr[3];
#pragma unroll
        for (int a = 0; a < D; ++a) { rr[a] = 2 * w[a] + 1 - ((r % 3) - 1); r /= 3; }  // If - f0
#pragma unroll
        for (int a = 0; a < D; ++a) idx = idx * R + rr[a];
        const float4 v = tile[idx];
        acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
      }
      // no source contributed: the node keeps its value (adding zeros changes nothing);
      // level 0 is complete after this restriction, so its grid update is applied here
      const bool add = acc.x != 0.f || acc.y != 0.f || acc.z != 0.f || acc.w != 0.f;
      if (add || l == 1) {
        float4 gg = gc;
        gg.x += acc.x; gg.y += acc.y; gg.z += acc.z; gg.w += acc.w;
        const bool upd = (l == 1) && grid_update<D>(T, Lc, ph, I, Mc, gg);
        if (add || upd) Lc.grid[qc] = gg;
      }
    }
    __syncthreads();  // the tile is refilled for the next block
  }
}

// The grid update of single-level scenes.  With two or more levels it is fused into the
// last restriction (level 0) and into each prolongation (levels >= 1).
template <int D>
__global__ void k_gridop(const SceneDesc* __restrict__ Dsc, float dt_arg) {
  AGIMP_SCENE
  if (T.L >= 2) return;
  Physics ph = sd.ph;
  if (dt_arg > 0.f) ph.dt = dt_arg;
  const LevelDesc L0 = T.lv[0];
  const NodeView N = node_view(T);
  const int total = T.scount[0] * Pow4<D>::v;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    int I[3];
    uint32_t q = list_node<D>(L0, t / Pow4<D>::v, t % Pow4<D>::v, I);
    if (!in_range_nodes<D>(I, L0.n)) continue;
    float4 g = L0.grid[q];
    if (grid_update<D>(T, L0, ph, I, nflag<D>(N, L0.shift, I), g)) L0.grid[q] = g;
  }
}

// adaptive_gimp.py:303-321 as a gather onto each level-l node J that is T, ghost or
// all-even: J = 2I + d over valid coarse nodes I, weight 2^-|d|_1 (only d = 0 unless T/G).
// A wavefront takes 64 / 4^d leaf blocks of level l, one node per lane.  The sources of a
// block's nodes 4S - PAD + [0, 4) per axis are the coarse nodes 2S - PAD/2 + [0, 3): the
// 3^d of them are staged in LDS, their flags and values loaded together (the values
// whatever the flags say, a read of the dense grid), beside each lane's own node flags and
// value, so one round of loads precedes the sums.  Sums are taken in the same order as a
// direct gather.
template <int D>
__global__ void __launch_bounds__(256) k_prolong(const SceneDesc* __restrict__ Dsc, int l, float dt_arg) {
  AGIMP_SCENE
  if (l >= T.L) return;
  Physics ph = sd.ph;
  if (dt_arg > 0.f) ph.dt = dt_arg;
  const LevelDesc Lf = T.lv[l];
  const LevelDesc Lc = T.lv[l - 1];
  const NodeView N = node_view(T);
  constexpr int NPB = Pow4<D>::v, BPW = 64 / NPB, NCB = Pow3<D>::v, NW = 4;
  __shared__ float4 src[NW][BPW][NCB];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nblk = T.scount[l];
  const int sb = lane / NPB, w = lane % NPB;  // this lane's block in the wavefront, node in it
  const int cb = lane / NCB, cr = lane % NCB;  // the block and source this lane stages
  for (int e0 = (blockIdx.x * NW + wv) * BPW; e0 < nblk; e0 += gridDim.x * NW * BPW) {  // uniform per wavefront
    // every load below is issued without a branch (indices clamped, results selected), so
    // the list entries load in one round and the flags and values in the next
    const bool stager = lane < BPW * NCB && e0 + cb < nblk;  // coarse source cr of block e0 + cb
    const int e = e0 + sb;                                    // node w of block e0 + sb
    const int bs = Lf.slist[min(e0 + cb, nblk - 1)];
    const int bn = Lf.slist[min(e, nblk - 1)];
    int B[3], Ic[3], rr = cr;
    block_coords<D>(Lf.NB, bs, B);
#pragma unroll
    for (int a = D - 1; a >= 0; --a) { Ic[a] = 2 * B[a] - PAD / 2 + rr % 3; rr /= 3; }
    bool sok = stager && in_range_nodes<D>(Ic, Lc.n);
    const int64_t qs = lidx<D>(Lc, Ic);
    sok = sok && qs >= 0;
    const uint8_t Mcs = nflag_nb<D>(N, Lc.shift, Ic);
    const float4 sv = Lc.grid[sok ? (uint32_t)qs : 0u];
    int J[3], Bn[3];
    block_coords<D>(Lf.NB, bn, Bn);
    if (D == 2) { J[0] = Bn[0] * 4 + (w >> 2) - PAD; J[1] = Bn[1] * 4 + (w & 3) - PAD; }
    else { J[0] = Bn[0] * 4 + (w >> 4) - PAD; J[1] = Bn[1] * 4 + ((w >> 2) & 3) - PAD; J[2] = Bn[2] * 4 + (w & 3) - PAD; }
    const uint32_t q = (uint32_t)bn * NPB + w;
    const bool live = e < nblk && in_range_nodes<D>(J, Lf.n);
    const uint8_t M = nflag_nb<D>(N, Lf.shift, J);
    float4 g = Lf.grid[q];
    if (stager) src[wv][cb][cr] = (sok && f_valid(Mcs)) ? sv : make_float4(0.f, 0.f, 0.f, 0.f);
    wave_sync_lds();
    if (live) {
      const bool isT = f_T(M), isG = f_ghost(M), isA = f_act(M);
      bool even = true;
#pragma unroll
      for (int a = 0; a < D; ++a) even = even && ((J[a] & 1) == 0);
      if (!(isT || isG || even)) {  // not overwritten by prolongation: level l's grid update
        if (grid_update<D>(T, Lf, ph, J, M, g)) Lf.grid[q] = g;
      } else {
        float acc[3] = {0.f, 0.f, 0.f};
        if (isT || isG || isA) {
          int nodd = 0;
#pragma unroll
          for (int a = 0; a < D; ++a) nodd += (J[a] & 1);
          const int combos = (isT || isG) ? 1 << nodd : (nodd == 0 ? 1 : 0);
          const float wgt = (nodd == 0) ? 1.f : (nodd == 1 ? 0.5f : (nodd == 2 ? 0.25f : 0.125f));
          int c0[3];  // the block's first source, 2S - PAD/2, from the node's own coordinates
#pragma unroll
          for (int a = 0; a < D; ++a) c0[a] = ((J[a] + PAD) & ~3) / 2 - PAD / 2;
#pragma unroll
          for (int u = 0; u < (1 << D); ++u) {
            if (u < combos) {
              int idx = 0, bit = 0;
#pragma unroll
              for (int a = 0; a < D; ++a) {
                int d = 0;
                if (J[a] & 1) { d = ((u >> bit) & 1) ? 1 : -1; ++bit; }
                idx = idx * 3 + ((J[a] - d) >> 1) - c0[a];
              }
              const float4 vc = src[wv][sb][idx];
              acc[0] += vc.x * wgt; acc[1] += vc.y * wgt; acc[2] += vc.z * wgt;
            }
          }
        }
        g.x = acc[0]; g.y = acc[1];
        if (D == 3) g.z = acc[2];
        Lf.grid[q] = g;
      }
    }
    wave_sync_lds();  // the sources are restaged for the next blocks
  }
}

// G2P's gather: the quadratic B-spline weights of the particle at x on level Ll and the
// velocity and velocity gradient sums over its 3^d stencil nodes (nv = sum w v, G = sum v (x)
// grad w, in cell units)
template <int D>
__device__ __forceinline__ void g2p_gather(const LevelDesc& Ll, int pow2, const float* x, float* nv, float (*G)[3]) {
  float w[3][3], dw[3][3];
  int J[3];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float qx = scaled(x[a], Ll, pow2);
    int b = (int)(qx + 0.5f) - 1;
    bspline(qx - (float)(b + 1), w[a], dw[a]);
    J[a] = b + PAD;
  }
  uint32_t ca[3][3];  // per-axis address terms of the stencil nodes
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int i = 0; i < 3; ++i) ca[a][i] = bidx_axis<D>(Ll.NB, a, J[a] + i);
  if constexpr (D == 2) {
#pragma unroll
    for (int i0 = 0; i0 < 3; ++i0)
#pragma unroll
      for (int i1 = 0; i1 < 3; ++i1) {
        float4 gv = Ll.grid[ca[0][i0] + ca[1][i1]];
        float wt = w[0][i0] * w[1][i1];
        float g0 = dw[0][i0] * w[1][i1], g1 = w[0][i0] * dw[1][i1];
        nv[0] += gv.x * wt; nv[1] += gv.y * wt;
        G[0][0] += gv.x * g0; G[0][1] += gv.x * g1;
        G[1][0] += gv.y * g0; G[1][1] += gv.y * g1;
      }
  } else {
    // regrouped by axis: for each (i0, i1) the sums along axis 2, s_a = sum N2 v_a and
    // t_a = sum D2 v_a, then nv_a += N0 N1 s_a, G_a0 += D0 N1 s_a, G_a1 += N0 D1 s_a and
    // G_a2 += N0 N1 t_a (w, dw: the weights N and their derivatives D per axis)
#pragma unroll
    for (int i0 = 0; i0 < 3; ++i0)
#pragma unroll
      for (int i1 = 0; i1 < 3; ++i1) {
        float4 gv[3];
#pragma unroll
        for (int i2 = 0; i2 < 3; ++i2) gv[i2] = Ll.grid[ca[0][i0] + ca[1][i1] + ca[2][i2]];
        float sv[3], tv[3];
        sv[0] = w[2][0] * gv[0].x; sv[1] = w[2][0] * gv[0].y; sv[2] = w[2][0] * gv[0].z;
        tv[0] = dw[2][0] * gv[0].x; tv[1] = dw[2][0] * gv[0].y; tv[2] = dw[2][0] * gv[0].z;
#pragma unroll
        for (int i2 = 1; i2 < 3; ++i2) {
          sv[0] += w[2][i2] * gv[i2].x; sv[1] += w[2][i2] * gv[i2].y; sv[2] += w[2][i2] * gv[i2].z;
          tv[0] += dw[2][i2] * gv[i2].x; tv[1] += dw[2][i2] * gv[i2].y; tv[2] += dw[2][i2] * gv[i2].z;
        }
        const float w01 = w[0][i0] * w[1][i1], h0 = dw[0][i0] * w[1][i1], h1 = w[0][i0] * dw[1][i1];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          nv[a] += w01 * sv[a];
          G[a][0] += h0 * sv[a];
          G[a][1] += h1 * sv[a];
          G[a][2] += w01 * tv[a];
        }
      }
  }
}

// adaptive_gimp.py:252-275 (+ plasticity and the next substep's stress); reads slot
// perm[pos] of the current buffer, writes slot pos of the next buffer (re-sorting for free).
// The last `nclear` workgroups (G2P_NEXT_REBUILD: the next substep's rebuild follows in the same
// graph, so nothing reads this octree any more) clear the octree's marks and the bucketing
// counters for it; they start as G2P's own workgroups drain.
// G2P_NEXT_FIXED: the next substep, in the same graph, reuses this octree; the last `nclear`
// workgroups then do its bucketing preparation instead (counters, and the octree's grid-phase
// block flags restored from the saved copy).
template <int D>
__global__ void __launch_bounds__(256) k_g2p(const SceneDesc* __restrict__ Dsc, float dt_arg, int par, int nclear,
                                             int next) {
  AGIMP_SCENE
  {
    const int nwg = gridDim.x - nclear;
    if ((int)blockIdx.x >= nwg) {
      const int wg = blockIdx.x - nwg;
      if (next == G2P_NEXT_REBUILD) {
        clear_octree<D>(T, wg, nclear);
      } else {
        uint32_t* a = (uint32_t*)T.lv[0].sflag;
        const uint32_t* v = sd.vflag;
        const size_t st = (size_t)nclear * blockDim.x;
        for (size_t i = wg * (size_t)blockDim.x + threadIdx.x; i < sd.w_flags; i += st) a[i] = v[i];
      }
      clear_bucket_counts(sd, wg, nclear);
      return;
    }
  }
  const PBuf& Pin = sd.P[par];
  const PBuf& Pout = sd.P[par ^ 1];
  const int* perm = sd.S.perm;
  const int* gb = sd.S.gb;
  int* flv = sd.S.flv;
  const int n = sd.n;
  unsigned long long* cnt = sd.cnt;
  Physics ph = sd.ph;
  if (dt_arg > 0.f) ph.dt = dt_arg;
  using PL = PLayout<D>;
  int pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  int k = perm[pos];
  const uint32_t ko = (uint32_t)k * 4u, po = (uint32_t)pos * 4u;  // byte offsets of the slots
  // the position is loaded beside the block id, not after the frozen test; the frozen copy
  // uses it too, so it stays ahead of the branch
  float x[3];
#pragma unroll
  for (int a = 0; a < D; ++a) x[a] = pload(pfield(Pin, PL::X + a), ko);
  int g = gb[k];
  if (g < 0) {  // frozen (fl_p = -1, adaptive_gimp.py:211-216): carried over unchanged
#pragma unroll
    for (int a = 0; a < D; ++a) pstore(pfield(Pout, PL::X + a), po, x[a]);
    for (int c = PL::X + D; c < PL::NF; ++c) pstore(pfield(Pout, c), po, pload(pfield(Pin, c), ko));
    Pout.g[pos] = Pin.g[k];
    Pout.id[pos] = Pin.id[k];
    flv[pos] = -1;
    return;
  }
  const int l = level_of_gbid(T, g);
  // particles are sorted by level, so a wavefront's level is nearly always one: then its
  // descriptor comes from scalar loads (no per-lane load round before the gathers)
  const int l0 = __builtin_amdgcn_readfirstlane(l);
  float nv[3] = {0.f, 0.f, 0.f}, G[3][3] = {};
  float sc;
  if (__ballot(l != l0) == 0ull) {
    const LevelDesc Ll = T.lv[l0];
    g2p_gather<D>(Ll, T.pow2, x, nv, G);
    sc = T.pow2 ? ph.dt * Ll.inv_dx : ph.dt / Ll.dx;  // dt (0.5/r)^d / dx, r = 0.5
  } else {
    const LevelDesc Ll = T.lv[l];
    g2p_gather<D>(Ll, T.pow2, x, nv, G);
    sc = T.pow2 ? ph.dt * Ll.inv_dx : ph.dt / Ll.dx;
  }
  float Fo[3][3];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) Fo[i][j] = pload(pfield(Pin, PL::F + i * D + j), ko);
  double Fd[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      float acc = 0.f;
#pragma unroll
      for (int q = 0; q < D; ++q) acc += ((i == q ? 1.f : 0.f) + sc * G[i][q]) * Fo[q][j];
      Fd[i][j] = (double)acc;
    }
  double Jp = (double)pload(pfield(Pin, PL::JP), ko);
  // everything that does not depend on the SVD is stored first: shorter live ranges across
  // the FP64 SVD, and the stores drain while it runs
  bool finite = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float xn = x[a] + ph.dt * nv[a];
    finite = finite && isfinite(xn) && isfinite(nv[a]);
    pstore(pfield(Pout, PL::X + a), po, xn);
    pstore(pfield(Pout, PL::V + a), po, nv[a]);
  }
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) pstore(pfield(Pout, PL::C + i * D + j), po, G[i][j]);
  pstore(pfield(Pout, PL::M), po, pload(pfield(Pin, PL::M), ko));
  Pout.g[pos] = Pin.g[k];
  Pout.id[pos] = Pin.id[k];
  flv[pos] = l;
  if (!finite) count(cnt, CNT_NAN);
  double tau[D][D];
  constitutive<D>(Fd, Jp, ph, tau, true);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) pstore(pfield(Pout, PL::F + i * D + j), po, (float)Fd[i][j]);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = i; j < D; ++j) pstore(pfield(Pout, PL::TAU + tau_idx<D>(i, j)), po, (float)tau[i][j]);
  pstore(pfield(Pout, PL::JP), po, (float)Jp);
}

// Kirchhoff stress of the stored F (no plastic projection): after the caller sets state
template <int D>
__global__ void k_stress(const SceneDesc* __restrict__ Dsc, int par) {
  AGIMP_SCENE
  const PBuf& P = sd.P[par];
  const int n = sd.n;
  const Physics& ph = sd.ph;
  using PL = PLayout<D>;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double Fd[D][D], tau[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) Fd[i][j] = (double)P.f[(size_t)(PL::F + i * D + j) * P.stride + k];
  double Jp = (double)P.f[(size_t)PL::JP * P.stride + k];
  constitutive<D>(Fd, Jp, ph, tau, false);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = i; j < D; ++j) P.f[(size_t)(PL::TAU + tau_idx<D>(i, j)) * P.stride + k] = (float)tau[i][j];
}

// ------------------------------------------------------------------------------ learned windows
// (single scene) swap particle buffer p with the spare one in the descriptor; later launches
// read the descriptor anew
__global__ void k_swap_spare(SceneDesc* Dsc, int p) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  SceneDesc& sd = Dsc[0];
  const PBuf t = sd.P[p];
  sd.P[p] = sd.spare;
  sd.spare = t;
}

// window start: the counters saved to snap, then the refinement check's first failed substep
// (none) and the substep count reset
__global__ void k_window_start(const SceneDesc* Dsc, unsigned long long* snap) {
  unsigned long long* cnt = Dsc[0].cnt;
  const int i = threadIdx.x;
  if (blockIdx.x != 0 || i >= NCOUNTERS) return;
  const unsigned long long v = cnt[i];
  snap[i] = v;
  cnt[i] = (i == CNT_FIRSTVIOL) ? ~0ull : (i == CNT_STEP ? 0ull : v);
}

// window end: failed checks, first failed substep, substeps, and the failed checks at the
// window start, written to `out` (host memory mapped for the device)
__global__ void k_window_end(const SceneDesc* Dsc, const unsigned long long* snap, unsigned long long* out) {
  const unsigned long long* cnt = Dsc[0].cnt;
  const int i = threadIdx.x;
  if (blockIdx.x != 0 || i >= 4) return;
  out[i] = (i < 3) ? cnt[CNT_VIOL + i] : snap[CNT_VIOL];
}

void launch_swap_spare(hipStream_t s, SceneDesc* d, int p) {
  AGIMP_LAUNCH("k_swap_spare", k_swap_spare, dim3(1), dim3(64), 0, s, d, p);
}
void launch_window_start(hipStream_t s, const SceneDesc* d, unsigned long long* snap) {
  AGIMP_LAUNCH("k_window_start", k_window_start, dim3(1), dim3(64), 0, s, d, snap);
}
void launch_window_end(hipStream_t s, const SceneDesc* d, const unsigned long long* snap, unsigned long long* out) {
  AGIMP_LAUNCH("k_window_end", k_window_end, dim3(1), dim3(64), 0, s, d, snap, out);
}

// ------------------------------------------------------------------------------ host I/O
// caller order (AoS, like the reference's to_numpy layout) -> slot k = original index
template <int D>
__global__ void k_upload(const SceneDesc* __restrict__ Dsc, int par, const float* x, const float* v, const float* F, const float* C,
                         const float* m, const int* g, const float* Jp) {
  AGIMP_SCENE
  const PBuf& P = sd.P[par];
  const int n = sd.n;
  int* flv = sd.S.flv;
  using PL = PLayout<D>;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    P.f[(size_t)(PL::X + a) * P.stride + k] = x[(size_t)k * D + a];
    P.f[(size_t)(PL::V + a) * P.stride + k] = v[(size_t)k * D + a];
  }
#pragma unroll
  for (int q = 0; q < D * D; ++q) {
    P.f[(size_t)(PL::F + q) * P.stride + k] = F[(size_t)k * D * D + q];
    P.f[(size_t)(PL::C + q) * P.stride + k] = C[(size_t)k * D * D + q];
  }
  P.f[(size_t)PL::M * P.stride + k] = m[k];
  P.f[(size_t)PL::JP * P.stride + k] = Jp ? Jp[k] : 1.0f;
  P.g[k] = g[k];
  P.id[k] = k;
  flv[k] = 0;  // Taichi fields start zeroed (adaptive_gimp.py:62)
}

// slot order -> caller order (AoS)
template <int D>
__global__ void k_download(const SceneDesc* __restrict__ Dsc, int par, float* x, float* v, float* F, float* C, float* Jp, int* fl,
                           int* g) {
  AGIMP_SCENE
  const PBuf& P = sd.P[par];
  const int n = sd.n;
  const int* flv = sd.S.flv;
  using PL = PLayout<D>;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int o = P.id[k];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    x[(size_t)o * D + a] = P.f[(size_t)(PL::X + a) * P.stride + k];
    v[(size_t)o * D + a] = P.f[(size_t)(PL::V + a) * P.stride + k];
  }
#pragma unroll
  for (int q = 0; q < D * D; ++q) {
    F[(size_t)o * D * D + q] = P.f[(size_t)(PL::F + q) * P.stride + k];
    C[(size_t)o * D * D + q] = P.f[(size_t)(PL::C + q) * P.stride + k];
  }
  Jp[o] = P.f[(size_t)PL::JP * P.stride + k];
  fl[o] = flv[k];
  g[o] = P.g[k];
}

// dense readback of one level over [0, e)^d; culled cells (A|C kept for speed) read as the
// reference's 0 (or GHOST when a ghost mark landed on them later)
template <int D>
__global__ void k_dense_level(const SceneDesc* __restrict__ Dsc, int l, int* cell, float* m, float* v) {
  AGIMP_SCENE
  const LevelDesc Ll = T.lv[l];
  size_t tot = 1;
  for (int a = 0; a < D; ++a) tot *= (size_t)Ll.e;
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= tot) return;
  int I[3];
  size_t r = t;
  for (int a = D - 1; a >= 0; --a) { I[a] = (int)(r % Ll.e); r /= Ll.e; }
  int64_t q = lidx<D>(Ll, I);
  uint8_t c = Ll.cell[q] & (uint8_t)~COVERED;
  if (c & CULLING) c &= GHOST;
  cell[t] = c;
  float4 g = Ll.grid[q];
  m[t] = g.w;
  v[t * D + 0] = g.x;
  v[t * D + 1] = g.y;
  if (D == 3) v[t * D + 2] = g.z;
}

template <int D>
__global__ void k_dense_node(const SceneDesc* __restrict__ Dsc, int* out) {
  AGIMP_SCENE
  const LevelDesc Ll = T.lv[T.L - 1];
  size_t tot = 1;
  for (int a = 0; a < D; ++a) tot *= (size_t)Ll.e;
  size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (t >= tot) return;
  int I[3];
  size_t r = t;
  for (int a = D - 1; a >= 0; --a) { I[a] = (int)(r % Ll.e); r /= Ll.e; }
  out[t] = T.node[lidx<D>(Ll, I)];
}

// ------------------------------------------------------------------------------ launchers
static thread_local KernelTimer* g_kernel_timer = nullptr;
void set_kernel_timer(KernelTimer* t) { g_kernel_timer = t; }
KernelTimer* kernel_timer() { return g_kernel_timer; }

static inline unsigned blocks_for(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

#define AGIMP_DIM(D, ...) \
  do { if ((D) == 2) { constexpr int DD = 2; __VA_ARGS__; } else { constexpr int DD = 3; __VA_ARGS__; } } while (0)

size_t cells_total(const Topo& T, int lmin, int lmax) {
  size_t s = 0;
  int p4 = T.dim == 2 ? 16 : 64;
  for (int l = lmin; l <= lmax; ++l) s += (size_t)T.lv[l].nblocks * p4;
  return s;
}

static inline dim3 grid2(size_t x, int nscenes) {
  return dim3((unsigned)std::max<size_t>(1, std::min<size_t>(x, 1u << 20)), (unsigned)nscenes);
}

void launch_refine(hipStream_t s, const SceneDesc* d, const Dims& m, int par, const LearnDesc* learn, bool chained,
                   bool precleared) {
  const dim3 pg = grid2(blocks_for(m.max_n, 256), m.nscenes);
  const dim3 cg = grid2(blocks_for((m.max_total_blocks + 3) / 4, COMPACT_T), m.nscenes);
  const dim3 rg = grid2(std::min<size_t>(blocks_for(m.max_total_blocks, 256), (size_t)m.fill_grid), m.nscenes);
  const dim3 lg = grid2(m.list_grid, m.nscenes);
  const dim3 fg = grid2(std::min<size_t>(blocks_for(m.max_fill_words, 256), (size_t)m.fill_grid), m.nscenes);
  AGIMP_DIM(m.dim, {
    if (!chained) AGIMP_LAUNCH("k_clear_sgrid", k_clear_sgrid<DD>, lg, dim3(256), 0, s, d);
    // precleared (chained only): the previous substep's G2P launch cleared the octree and the
    // bucketing counters; the grid of its lists is zeroed beside the node marks
    if (!precleared) AGIMP_LAUNCH("k_clear_refine", k_clear_refine<DD>, rg, dim3(256), 0, s, d, chained ? 1 : 0);
    if (m.any_static) AGIMP_LAUNCH("k_activate_static", k_activate_static<DD>, pg, dim3(256), 0, s, d, par);
    if (m.any_dynamic) {
      AGIMP_LAUNCH("k_activate_dynamic", k_activate_dynamic<DD>, pg, dim3(256), 0, s, d, par);
      if (learn) launch_unit_activate(s, d, m.dim, (int)m.max_n, par, *learn);
      if (m.maxL > 1) {
        AGIMP_LAUNCH("k_compact_cells", k_compact_cells, cg, dim3(COMPACT_T), 0, s, d, 0);
        AGIMP_LAUNCH("k_cull", k_cull<DD>, lg, dim3(256), 0, s, d);
      }
    }
    AGIMP_LAUNCH("k_compact_cells", k_compact_cells, cg, dim3(COMPACT_T), 0, s, d, 2);
    {
      const int ngrid = precleared ? (int)rg.x : 0;
      AGIMP_LAUNCH("k_node_A", k_node_A<DD>, dim3(lg.x + ngrid, lg.y), dim3(256), 0, s, d, ngrid);
    }
    // chained: the T and ghost marks run in the bucketing pass's first launch, beside the level map
    if (!chained) AGIMP_LAUNCH("k_mark_TG", k_mark_TG<DD>, lg, dim3(256), 0, s, d);
    if (!chained) AGIMP_LAUNCH("k_copy_flags", k_copy_flags, fg, dim3(256), 0, s, d, 1);
    if (learn) {
      for (int l = m.maxL - 2; l >= 0; --l) AGIMP_LAUNCH("k_cover", k_cover<DD>, dim3(2048), dim3(256), 0, s, d, l);
      AGIMP_LAUNCH("k_full_corners", k_full_corners<DD>, dim3(2048), dim3(256), 0, s, d);
    }
  });
}

template <int D, int NL>
static void launch_levelmap(hipStream_t s, const SceneDesc* d, int par, bool chained, int ngrid, dim3 lg, dim3 lmg) {
  if (chained)
    AGIMP_LAUNCH("k_marks_levelmap", (k_marks_levelmap<D, NL>), dim3(lg.x + lmg.x, lmg.y), dim3(LM_T), 0, s, d, par,
                 (int)lg.x);
  else
    AGIMP_LAUNCH("k_levelmap", (k_levelmap<D, NL>), dim3(lmg.x + ngrid, lmg.y), dim3(LM_T), 0, s, d, par, ngrid);
}

void launch_bucket(hipStream_t s, const SceneDesc* d, const Dims& m, int par, bool chained, bool prepped) {
  const dim3 pg = grid2(blocks_for(m.max_n, 256), m.nscenes);
  const dim3 cg = grid2(blocks_for((m.max_total_blocks + 3) / 4, COMPACT_T), m.nscenes);
  const dim3 lg = grid2(m.list_grid, m.nscenes);
  const dim3 fg = grid2(std::min<size_t>(blocks_for(m.max_fill_words, 256), (size_t)m.fill_grid), m.nscenes);
  const dim3 tg = grid2(blocks_for(m.max_total_blocks, SCAN_TILE), m.nscenes);
  const dim3 sbg = grid2(std::max(blocks_for(m.max_n, 256), blocks_for(m.max_total_blocks, 256)), m.nscenes);
  AGIMP_DIM(m.dim, {
    // prepped: the previous G2P launch did the preparation but the grid clear, which runs
    // beside the level map
    if (!chained && !prepped) AGIMP_LAUNCH("k_bucket_prep", k_bucket_prep<DD>, fg, dim3(256), 0, s, d);
    const int ngrid = prepped ? (int)fg.x : 0;
    {
      const dim3 lmg = grid2(blocks_for(m.max_n, LM_T * LM_P), m.nscenes);
      // the level loop unrolled for the most levels of the launch's scenes; chained (after a
      // rebuild in the same graph): the rebuild's T and ghost marks in the same launch
      switch (m.maxL) {
        case 1: launch_levelmap<DD, 1>(s, d, par, chained, ngrid, lg, lmg); break;
        case 2: launch_levelmap<DD, 2>(s, d, par, chained, ngrid, lg, lmg); break;
        case 3: launch_levelmap<DD, 3>(s, d, par, chained, ngrid, lg, lmg); break;
        case 4: launch_levelmap<DD, 4>(s, d, par, chained, ngrid, lg, lmg); break;
        case 5: launch_levelmap<DD, 5>(s, d, par, chained, ngrid, lg, lmg); break;
        default: launch_levelmap<DD, MAXL>(s, d, par, chained, ngrid, lg, lmg); break;
      }
    }
    AGIMP_LAUNCH("k_scan_reduce", k_scan_reduce, tg, dim3(SCAN_T), 0, s, d, chained ? 1 : 0);
    AGIMP_LAUNCH("k_scan_final", k_scan_final, tg, dim3(SCAN_T), 0, s, d);
    AGIMP_LAUNCH("k_scatter_blocks", k_scatter_blocks<DD>, sbg, dim3(256), 0, s, d);
  });
}

void launch_p2g(hipStream_t s, const SceneDesc* d, const Dims& m, float dt, int par) {
  // P2G's workgroups, then those compacting the grid-phase block flags (COMPACT_T == P2G_T)
  static_assert(COMPACT_T == P2G_T, "one workgroup size for both parts of the launch");
  const int cw = (int)blocks_for((m.max_total_blocks + 3) / 4, COMPACT_T);
  AGIMP_DIM(m.dim, AGIMP_LAUNCH("k_p2g", k_p2g<DD>, grid2((size_t)m.p2g_grid + cw, m.nscenes), dim3(P2G_T), 0, s, d,
                                dt, par, m.p2g_grid));
}

void launch_grid(hipStream_t s, const SceneDesc* d, const Dims& m, float dt) {
  const dim3 gg = grid2(m.grid_grid, m.nscenes);
  AGIMP_DIM(m.dim, {
    for (int l = m.maxL - 1; l >= 1; --l) {
      if constexpr (DD == 3) AGIMP_LAUNCH("k_restrict", k_restrict_wg<DD>, gg, dim3(256), 0, s, d, l, dt);
      else AGIMP_LAUNCH("k_restrict", k_restrict<DD>, gg, dim3(256), 0, s, d, l, dt);
    }
    if (m.any_single) AGIMP_LAUNCH("k_gridop", k_gridop<DD>, gg, dim3(256), 0, s, d, dt);
    for (int l = 1; l < m.maxL; ++l) AGIMP_LAUNCH("k_prolong", k_prolong<DD>, gg, dim3(256), 0, s, d, l, dt);
  });
}

void launch_g2p(hipStream_t s, const SceneDesc* d, const Dims& m, float dt, int par, int next) {
  const int nclear = next != G2P_NEXT_NONE
                         ? (int)std::min<size_t>(blocks_for(next == G2P_NEXT_REBUILD ? m.max_total_blocks : m.max_fill_words, 256),
                                                 (size_t)m.fill_grid)
                         : 0;
  const dim3 g = grid2(blocks_for(m.max_n, 256) + nclear, m.nscenes);
  AGIMP_DIM(m.dim, AGIMP_LAUNCH("k_g2p", k_g2p<DD>, g, dim3(256), 0, s, d, dt, par, nclear, next));
}

void launch_stress(hipStream_t s, const SceneDesc* d, int dim, int n, int par) {
  AGIMP_DIM(dim, AGIMP_LAUNCH("k_stress", k_stress<DD>, grid2(blocks_for(n, 256), 1), dim3(256), 0, s, d, par));
}

void launch_upload(hipStream_t s, const SceneDesc* d, int par, int dim, int n, const float* x, const float* v,
                   const float* F, const float* C, const float* m, const int* g, const float* Jp) {
  AGIMP_DIM(dim, AGIMP_LAUNCH("k_upload", k_upload<DD>, grid2(blocks_for(n, 256), 1), dim3(256), 0, s, d, par, x, v, F, C, m,
                                    g, Jp));
}

void launch_download(hipStream_t s, const SceneDesc* d, int par, int dim, int n, float* x, float* v, float* F,
                     float* C, float* Jp, int* fl, int* g) {
  AGIMP_DIM(dim, AGIMP_LAUNCH("k_download", k_download<DD>, grid2(blocks_for(n, 256), 1), dim3(256), 0, s, d, par, x, v, F, C,
                                    Jp, fl, g));
}

void launch_dense_level(hipStream_t s, const SceneDesc* d, const Topo& T, int l, int* cell, float* m, float* v) {
  size_t tot = 1;
  for (int a = 0; a < T.dim; ++a) tot *= (size_t)T.lv[l].e;
  AGIMP_DIM(T.dim, AGIMP_LAUNCH("k_dense_level", k_dense_level<DD>, grid2(blocks_for(tot, 256), 1), dim3(256), 0, s, d, l, cell,
                                      m, v));
}

void launch_dense_node(hipStream_t s, const SceneDesc* d, const Topo& T, int* out) {
  size_t tot = 1;
  for (int a = 0; a < T.dim; ++a) tot *= (size_t)T.lv[T.L - 1].e;
  AGIMP_DIM(T.dim, AGIMP_LAUNCH("k_dense_node", k_dense_node<DD>, grid2(blocks_for(tot, 256), 1), dim3(256), 0, s, d, out));
}

}  // namespace agimp
