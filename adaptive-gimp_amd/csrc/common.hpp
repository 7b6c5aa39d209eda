// Shared definitions for the adaptive-GIMP substep on gfx950.
//
// Memory layout in HBM (one simulator handle):
//  * every level l keeps dense, block-major arrays over its padded index box
//    [-PAD, e_l + PAD)^d, e_l = 4*align_size(n_l/4 + 1, 4) (adaptive_gimp.py:36-37): a 4^d leaf
//    block (adaptive_gimp.py:28) is contiguous (64 nodes x 16 B = 1 KiB in 3D), blocks are
//    row-major.  Cell flags are one byte per cell, grid nodes one float4 (v.xyz, m).
//  * node_mask (adaptive_gimp.py:23-26) is one byte per finest node in the finest layout.
//  * particles are structure-of-arrays, one float array per component, kept in the order
//    of the last substep's (level, leaf block) sort; orig_id maps back to caller order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Device buffers live in global memory.  Typing the descriptors' pointers as such in the
// device pass makes every access through them a global_* instruction (SGPR base + VGPR
// offset, waited on by vmcnt alone) instead of a flat_* one, whose lgkmcnt share would
// make every scalar load of a descriptor field wait for all outstanding vector loads.
#ifdef __HIP_DEVICE_COMPILE__
#define AGIMP_GLOBAL __attribute__((address_space(1)))
#else
#define AGIMP_GLOBAL
#endif

namespace agimp {

// a device allocation as a descriptor pointer
template <class T>
__host__ __device__ inline T AGIMP_GLOBAL* as_global(T* p) {
  return (T AGIMP_GLOBAL*)p;
}

constexpr int MAXL = 6;  // octree levels supported
constexpr int PAD = 4;   // padding cells on each side of every level (one leaf block)

// cell / node flag bits, utils.py:5-9
constexpr uint8_t ACTIVATED = 1, GHOST = 2, T_JUNCTION = 4, CULLING = 8;
// not in the reference: a cell without ACTIVATED whose whole volume is covered by finer
// activated cells (set only by the rebuild of a predicted octree, read by the criterion check)
constexpr uint8_t COVERED = 16;
// corner map entries (a level's corner N is shared by its cells N - d, d in {0,1}^d):
// CORNER_HIT, one of those cells is exactly ACTIVATED; CORNER_FULL (predicted octrees only),
// every one of them that lies in the padded box holds ACTIVATED or COVERED
constexpr uint8_t CORNER_HIT = 1, CORNER_FULL = 2;

struct LevelDesc {
  int n;        // cells per axis in the domain, c * 2^l
  int e;        // reference extent per axis
  int E;        // padded extent per axis (e + 2*PAD)
  int NB;       // leaf blocks per axis (E / 4)
  int shift;    // L-1-l: level-l index -> finest index multiplier exponent
  int nblocks;  // NB^d
  int gbase;    // first global block id of this level
  float dx, inv_dx;
  uint8_t AGIMP_GLOBAL* cell;   // [nblocks * 4^d]
  // [nblocks * 4^d] in the cells' layout: entry N is 1 if level-l node N (the low corner of
  // cell N) is a corner of an exactly-ACTIVATED level-l cell.  A particle's 2^d cells at level
  // l are the cells of which its base cell + 1 is a corner, so one entry answers whether one
  // of them is exactly ACTIVATED (the level map's test).
  uint8_t AGIMP_GLOBAL* corner;
  // [nblocks * 4^d] in the cells' layout: 1 if the culling pass split-activated the cell (its
  // parent got CULLING); merged into the cell's ACTIVATED bit by the node-marks pass
  uint8_t AGIMP_GLOBAL* split;
  float4 AGIMP_GLOBAL* grid;    // [nblocks * 4^d]
  uint8_t AGIMP_GLOBAL* sflag;  // [nblocks]   block must be processed by the grid phase
  int AGIMP_GLOBAL* slist;      // [nblocks]   compacted sflag list
};

struct Topo {
  int dim, L, c;
  int finest_size;
  int total_blocks;  // sum of nblocks over levels
  int pow2;          // coarsest size is a power of two: x/dx == x*inv_dx exactly
  LevelDesc lv[MAXL];
  uint8_t AGIMP_GLOBAL* node;     // node_mask, finest layout
  int AGIMP_GLOBAL* scount;       // [MAXL] number of entries in each slist
  uint8_t AGIMP_GLOBAL* cflag;    // [total_blocks] leaf block holds a non-zero cell (by global block id)
  int AGIMP_GLOBAL* clist;        // [total_blocks] compacted cflag (global block ids)
  int AGIMP_GLOBAL* ccount;       // [4] list lengths of the successive compactions of cflag
};

struct Physics {
  float gravity[3];
  float dt;
  int boundary_gap;
  int wall_mode;  // 0: per-axis normal zeroing (adaptive_gimp.py:284-287), 1: sticky
  int plastic;    // 0 none, 1 singular-value clamp
  double theta_c, theta_s, hardening;
  double mu0, la0;
  float p_vol;
  int slab_enable, slab_axis;
  float slab_lo, slab_hi;
};

// Particle structure-of-arrays: component c of particle slot k is f[c*stride + k].
template <int D>
struct PLayout {
  static constexpr int X = 0, V = D, F = 2 * D, C = 2 * D + D * D, TAU = 2 * D + 2 * D * D;
  static constexpr int NTAU = D * (D + 1) / 2;
  static constexpr int M = TAU + NTAU, JP = M + 1, NF = JP + 1;
};
// tau packing (symmetric): 2D xx,yy,xy ; 3D xx,yy,zz,xy,xz,yz
template <int D>
__host__ __device__ inline int tau_idx(int i, int j) {
  if (i == j) return i;
  if (D == 2) return 2;
  int a = i < j ? i : j, b = i < j ? j : i;
  return (a == 0) ? (b == 1 ? 3 : 4) : 5;
}

struct PBuf {
  float AGIMP_GLOBAL* f;       // [NF * stride]
  int AGIMP_GLOBAL* g;         // [stride] prescribed level g_p
  int AGIMP_GLOBAL* id;        // [stride] original particle index
  int64_t stride;
};

// Field c of slot k: the field's base (the same for every lane, so it stays in scalar registers)
// plus the slot's byte offset k * 4 (one 32-bit vector register shared by all fields).  Each
// access is then a single global load or store with a scalar base, without 64-bit address
// arithmetic per field in vector registers.
__device__ __forceinline__ float AGIMP_GLOBAL* pfield(const PBuf& P, int c) {
  float AGIMP_GLOBAL* b = P.f + (size_t)c * P.stride;
#ifdef __HIP_DEVICE_COMPILE__
  asm volatile("" : "+s"(b));  // a scalar base the compiler cannot fold into the lane offsets
#endif
  return b;
}
__device__ __forceinline__ float pload(const float AGIMP_GLOBAL* base, uint32_t off) {
  return *(const float AGIMP_GLOBAL*)((const char AGIMP_GLOBAL*)base + off);
}
__device__ __forceinline__ void pstore(float AGIMP_GLOBAL* base, uint32_t off, float v) {
  *(float AGIMP_GLOBAL*)((char AGIMP_GLOBAL*)base + off) = v;
}

// scratch of the (level, leaf block) counting sort and the P2G chunk list
struct SortBufs {
  int AGIMP_GLOBAL* gb;       // [n]  global block id of each slot's particle (-1: frozen)
  int AGIMP_GLOBAL* rk;       // [n]  rank inside its block (or among frozen particles)
  uint8_t AGIMP_GLOBAL* lkey; // [n]  cell of each slot's particle inside its leaf block (P2G sort key)
  uint8_t AGIMP_GLOBAL* pkey; // [n]  the same key by sorted position
  int AGIMP_GLOBAL* perm;     // [n]  sorted position -> slot
  int AGIMP_GLOBAL* flv;      // [n]  fl_p of the particle now in each slot
  int AGIMP_GLOBAL* bcnt;     // [total_blocks] particles per block
  int AGIMP_GLOBAL* boff;     // [total_blocks] first sorted position of each block
  int AGIMP_GLOBAL* choff;    // [total_blocks] first chunk of each block
  int2 AGIMP_GLOBAL* tiles;   // scan tile partials
  int AGIMP_GLOBAL* totals;   // [2] active particles, chunks
  int AGIMP_GLOBAL* frozen;   // [1] frozen-particle counter
  int4 AGIMP_GLOBAL* chunks;  // (gbid, first position, count, -)
};

// Everything the kernels need about one scene.  Kernels take an array of these and pick
// theirs with blockIdx.y, so one launch advances a whole batch of scenes; a single
// simulator is a batch of one.  Which of the two particle buffers is current is a launch
// argument (the host tracks it; a captured substep exists for each of the two values).
struct SceneDesc {
  Topo T;
  PBuf P[2];
  // a third particle buffer (learned prediction only): a window of substeps on one predicted
  // octree swaps the window start's buffer with it after the first substep, so that state is
  // kept for a rollback without a copy
  PBuf spare;
  SortBufs S;
  Physics ph;
  int n;
  int static_mode;
  int CH;  // P2G particles per work chunk
  const int8_t AGIMP_GLOBAL* st_ret;
  const int8_t AGIMP_GLOBAL* st_act;
  unsigned long long AGIMP_GLOBAL* cnt;
  uint32_t AGIMP_GLOBAL* vflag;
  size_t w_cells, w_node, w_flags;  // 32-bit words of the cell masks, node mask, block flags
  // refinement criterion check (used while a predicted octree is reused): a particle's 2^d
  // request cells at level g_p must be covered by cells of level >= g_p, i.e. hold the
  // ACTIVATED or COVERED bit, or lie in a unit (cell of level ulev) whose predicted level
  // upred >= g_p
  int check;
  int ulev;
  const int8_t AGIMP_GLOBAL* upred;
};

// counters kept on the device (int64 each)
enum Counter : int {
  CNT_OOB = 0,        // accesses outside the padded level boxes (reference: undefined behaviour)
  CNT_FROZEN = 1,     // particles with fl_p = -1 this substep
  CNT_ACTIVE = 2,     // particles transferred this substep
  CNT_CHUNKS = 3,     // P2G work chunks
  CNT_NAN = 4,        // particles whose new state is not finite
  CNT_VIOL = 5,       // particles whose refinement criterion the octree failed (check on)
  CNT_FIRSTVIOL = 6,  // smallest substep number (CNT_STEP) with a failed check
  CNT_STEP = 7,       // substeps started (incremented by the bucketing phase)
  NCOUNTERS = 8
};

// Units of the learned refinement prediction: the cells of level ulev = max(0, L-3), i.e.
// blocks of 4^d finest cells, in the padded block-major layout of that level.
constexpr int UNIT_ACC = 16;  // accumulator words per unit
enum UnitAcc : int {
  UA_MASS = 0, UA_COUNT = 1, UA_VSUM = 2, UA_VMAX = 3, UA_CSUM = 4, UA_JSUM = 5, UA_JMIN = 6,  // JMIN: max key of -det F
  UA_GMAX = 7, UA_V = 9, UA_REQ = 12,  // GMAX, REQ: level + 1 (0: none); V: 3 words
  UA_REQF = 13  // requests of the particles moved by v * window (level + 1)
};
constexpr int UNIT_NFEAT = 16;

struct LearnDesc {
  int ulev;            // unit level
  int nu;              // units (cells of level ulev incl. padding)
  float win;           // window length in seconds (velocities -> displacements)
  float inv_mref;      // 1 / (reference density * unit volume)
  float tau;           // decision threshold on the cumulative class probability
  int mode;            // parts of a predicted octree: 1 network unit levels, 2 swept requests
  uint32_t AGIMP_GLOBAL* acc;       // [nu * UNIT_ACC]
  uint8_t AGIMP_GLOBAL* uflag;      // [nu] unit holds particles
  uint8_t AGIMP_GLOBAL* cflag;      // [nu] unit is a candidate (within one unit of an occupied one)
  int AGIMP_GLOBAL* occ;            // [nu] compacted occupied units
  int AGIMP_GLOBAL* cand;           // [nu] compacted candidates
  int AGIMP_GLOBAL* counts;         // [4] occupied, candidates, units refined by the prediction, -
  float AGIMP_GLOBAL* feat;         // [nu * UNIT_NFEAT] features of the candidates (candidate order)
  int AGIMP_GLOBAL* lab;            // [nu] finest requested level + 1 over a label window
  int8_t AGIMP_GLOBAL* upred;       // [nu] predicted level (-1: none)
  const float AGIMP_GLOBAL* params; // network parameters (5508 floats)
};

// block-major index of padded coordinates J (each in [0, E))
// bidx is a sum of per-axis terms: bidx(NB, J) = sum_a bidx_axis<D>(NB, a, J[a]), so a
// stencil's addresses need the 3 terms of each axis, not a full index per node
template <int D>
__device__ __forceinline__ uint32_t bidx_axis(int NB, int a, int j) {
  const uint32_t nb = (uint32_t)NB;
  const uint32_t S = (D == 2) ? (a == 0 ? nb * 16u : 16u) : (a == 0 ? nb * nb * 64u : (a == 1 ? nb * 64u : 64u));
  const uint32_t s = (D == 2) ? (a == 0 ? 4u : 1u) : (a == 0 ? 16u : (a == 1 ? 4u : 1u));
  return (uint32_t)(j >> 2) * S + (uint32_t)(j & 3) * s;
}

template <int D>
__device__ __forceinline__ uint32_t bidx(int NB, const int* J) {
  if (D == 2) {
    uint32_t b = (uint32_t)((J[0] >> 2) * NB + (J[1] >> 2));
    return (b << 4) | (uint32_t)(((J[0] & 3) << 2) | (J[1] & 3));
  } else {
    uint32_t b = (uint32_t)(((J[0] >> 2) * NB + (J[1] >> 2)) * NB + (J[2] >> 2));
    return (b << 6) | (uint32_t)(((J[0] & 3) << 4) | ((J[1] & 3) << 2) | (J[2] & 3));
  }
}

template <int D>
__device__ __forceinline__ bool in_box(int E, const int* J) {
  bool ok = true;
#pragma unroll
  for (int a = 0; a < D; ++a) ok = ok && (J[a] >= 0) && (J[a] < E);
  return ok;
}

// block id -> block coordinates
template <int D>
__device__ __forceinline__ void block_coords(int NB, int b, int* B) {
  if (D == 2) { B[0] = b / NB; B[1] = b - B[0] * NB; }
  else { B[2] = b % NB; int r = b / NB; B[1] = r % NB; B[0] = r / NB; }
}

// the floor((x/dx + 0.5)) - 1 base cell of a particle at level l, with the reference's
// f32 arithmetic: x/dx exact for power-of-two sizes, then +0.5 and truncation toward zero
__device__ __forceinline__ int base_cell(float x, const LevelDesc& L, int pow2) {
  float q = pow2 ? x * L.inv_dx : x / L.dx;
  return (int)(q + 0.5f) - 1;
}
__device__ __forceinline__ float scaled(float x, const LevelDesc& L, int pow2) {
  return pow2 ? x * L.inv_dx : x / L.dx;
}

__device__ __forceinline__ int level_of_gbid(const Topo& T, int gbid) {
  int l = 0;
#pragma unroll
  for (int q = 1; q < MAXL; ++q)
    if (q < T.L && gbid >= T.lv[q].gbase) l = q;
  return l;
}

// ---------------------------------------------------------------------------------------
// Run-aggregated atomic add: lanes whose (active, addr) equal their left neighbour's join
// its run; each run's first lane adds the run length once and every lane receives
// base + its offset in the run.  Must be called by all lanes of the wavefront.
__device__ __forceinline__ int wave_run_atomic_add(int* addr, bool active) {
  const int lane = __lane_id();
  unsigned long long a_bits = (unsigned long long)(uintptr_t)addr;
  unsigned long long prev = __shfl_up(a_bits, 1);
  int prev_act = __shfl_up((int)active, 1);
  bool head = !active || lane == 0 || !prev_act || prev != a_bits;
  unsigned long long heads = __ballot(head);
  unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
  int h = 63 - __clzll(heads & le);
  unsigned long long above = heads & ~((h == 63) ? ~0ull : ((1ull << (h + 1)) - 1ull));
  int nh = above ? (__ffsll((long long)above) - 1) : 64;
  int base = 0;
  if (active && lane == h) base = atomicAdd(addr, nh - h);
  base = __shfl(base, h);
  return base + (lane - h);
}

// LDS written by some lanes of a wavefront and read by others: order the accesses and keep
// the compiler from moving them across this point (the wavefront itself runs in lockstep)
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Slot reservation for scattered active lanes (compactions): lanes are grouped by address,
// one atomic per distinct address in the wavefront (one per counter, however the active
// lanes interleave), slots in lane order.  The loop is uniform over the wavefront.
__device__ __forceinline__ int wave_agg_atomic_add(int* addr, bool active) {
  const int lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1ull;
  const unsigned long long a_bits = (unsigned long long)(uintptr_t)addr;
  unsigned long long todo = __ballot(active);
  int slot = 0;
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const unsigned long long la = __shfl(a_bits, leader);
    const unsigned long long grp = __ballot(active && a_bits == la);
    int base = 0;
    if (lane == leader) base = atomicAdd(addr, __popcll(grp));
    base = __shfl(base, leader);
    if ((grp >> lane) & 1ull) slot = base + __popcll(grp & below);
    todo &= ~grp;
  }
  return slot;
}

}  // namespace agimp
