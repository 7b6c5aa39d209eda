// Host-side launchers of the substep kernels (kernels.hip) used by the C-ABI (capi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace agimp {

// scratch of the (level, leaf block) counting sort and the P2G chunk list
struct SortBufs {
  int* gb;       // [n]  global block id of each slot's particle (-1: frozen)
  int* rk;       // [n]  rank inside its block (or among frozen particles)
  int* perm;     // [n]  sorted position -> slot
  int* flv;      // [n]  fl_p of the particle now in each slot
  int* bcnt;     // [total_blocks] particles per block
  int* boff;     // [total_blocks] first sorted position of each block
  int* choff;    // [total_blocks] first chunk of each block
  int2* tiles;   // scan tile partials
  int* totals;   // [2] active particles, chunks
  int* frozen;   // [1] frozen-particle counter
  int4* chunks;  // (gbid, first position, count, -)
};

size_t cells_total(const Topo& T, int lmin, int lmax);
void launch_fill_u32(hipStream_t s, void* p, uint32_t v, size_t n);
// zero up to 8 buffers (ptrs[i], words[i] 32-bit words) in one launch
void launch_fill_ranges(hipStream_t s, void* const* ptrs, const size_t* words, int k);
void launch_copy_u32(hipStream_t s, void* dst, const void* src, size_t n);
void launch_clear_sgrid(hipStream_t s, const Topo& T, int grid_blocks);
void launch_refine(hipStream_t s, const Topo& T, const PBuf& P, int n, int static_mode, const int8_t* ret,
                   const int8_t* act, unsigned long long* cnt);
void launch_bucket(hipStream_t s, const Topo& T, const PBuf& P, int n, SortBufs& S, int CH, unsigned long long* cnt);
void launch_p2g(hipStream_t s, const Topo& T, const PBuf& Pin, SortBufs& S, const Physics& ph, int p2g_grid);
void launch_grid(hipStream_t s, const Topo& T, const Physics& ph, int grid_grid);
void launch_g2p(hipStream_t s, const Topo& T, const PBuf& Pin, const PBuf& Pout, int n, SortBufs& S,
                const Physics& ph, unsigned long long* cnt);
void launch_stress(hipStream_t s, int dim, const PBuf& P, int n, const Physics& ph);
void launch_upload(hipStream_t s, int dim, const PBuf& P, int n, const float* x, const float* v, const float* F,
                   const float* C, const float* m, const int* g, const float* Jp, int* flv);
void launch_download(hipStream_t s, int dim, const PBuf& P, int n, const int* flv, float* x, float* v, float* F,
                     float* C, float* Jp, int* fl, int* g);
void launch_dense_level(hipStream_t s, const Topo& T, int l, int* cell, float* m, float* v);
void launch_dense_node(hipStream_t s, const Topo& T, int* out);

}  // namespace agimp
