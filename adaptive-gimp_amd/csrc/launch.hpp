// Host-side launchers of the substep kernels (kernels.hip) used by the C-ABI (capi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace agimp {

// launch geometry of a batch of scenes (a single simulator is a batch of one)
struct Dims {
  int dim;
  int nscenes;
  int maxL;
  int any_single;           // some scene has one level (its grid update is not fused)
  int any_static, any_dynamic;
  size_t max_n;             // largest particle count
  size_t max_total_blocks;  // largest number of leaf blocks over all levels
  size_t max_fill_words;    // largest per-substep clear (32-bit words)
  int list_grid;            // workgroups per scene for block-list kernels
  int fill_grid;            // largest grid per scene of the clear / flag-copy kernels
  int p2g_grid;             // workgroups per scene for the P2G chunk loop
  int grid_grid;            // workgroups per scene for the grid passes
};

// Per-kernel timing, the counterpart of the Taichi kernel profiler (run_3d.py:12, :43): while a
// timer is installed on the calling thread, the launchers bracket every kernel launch with
// begin/end on the launch stream.  Never installed during graph capture.
struct KernelTimer {
  virtual void begin(const char* name, hipStream_t s) = 0;
  virtual void end(hipStream_t s) = 0;
  virtual ~KernelTimer() = default;
};
void set_kernel_timer(KernelTimer* t);  // per thread; nullptr switches it off
KernelTimer* kernel_timer();

#define AGIMP_LAUNCH(name, kern, grid, block, shmem, stream, ...)             \
  do {                                                                         \
    ::agimp::KernelTimer* kt_ = ::agimp::kernel_timer();                       \
    if (kt_) kt_->begin(name, stream);                                         \
    hipLaunchKernelGGL(kern, grid, block, shmem, stream, __VA_ARGS__);         \
    if (kt_) kt_->end(stream);                                                 \
  } while (0)

size_t cells_total(const Topo& T, int lmin, int lmax);
// par: index of the current particle buffer of every scene of the launch (G2P writes par ^ 1)
// learn (optional, single scene): cells of the predicted units are activated after the
// particles' own requests, before the culling
// chained: a rebuild whose bucketing pass follows directly in the same stream (both passed
// chained).  The rebuild's clear then also zeroes the previous grid-phase lists' grid, and the
// bucketing pass saves the octree's grid-phase block flags, which an unchained rebuild saves
// itself (k_copy_flags) and an unchained bucketing pass restores.
// precleared (chained only): the previous substep's G2P launch ran with clear_next, so the
// rebuild skips its clear launch and zeroes the previous lists' grid beside the node marks.
void launch_refine(hipStream_t s, const SceneDesc* d, const Dims& m, int par, const LearnDesc* learn = nullptr,
                   bool chained = false, bool precleared = false);
// prepped (unchained only): the previous substep's G2P launch ran with G2P_NEXT_FIXED, so the
// preparation launch is skipped and the previous lists' grid is zeroed beside the level map.
void launch_bucket(hipStream_t s, const SceneDesc* d, const Dims& m, int par, bool chained = false,
                   bool prepped = false);
void launch_p2g(hipStream_t s, const SceneDesc* d, const Dims& m, float dt, int par);
void launch_grid(hipStream_t s, const SceneDesc* d, const Dims& m, float dt);
// next: what follows in the same stream: G2P_NEXT_REBUILD, a chained rebuild with precleared;
// G2P_NEXT_FIXED, a bucketing pass with prepped on this octree.  G2P's launch then does that
// pass's preparation.
enum G2pNext : int { G2P_NEXT_NONE = 0, G2P_NEXT_REBUILD = 1, G2P_NEXT_FIXED = 2 };
void launch_g2p(hipStream_t s, const SceneDesc* d, const Dims& m, float dt, int par, int next = G2P_NEXT_NONE);
void launch_stress(hipStream_t s, const SceneDesc* d, int dim, int n, int par);
void launch_upload(hipStream_t s, const SceneDesc* d, int par, int dim, int n, const float* x, const float* v,
                   const float* F, const float* C, const float* m, const int* g, const float* Jp);
void launch_download(hipStream_t s, const SceneDesc* d, int par, int dim, int n, float* x, float* v, float* F,
                     float* C, float* Jp, int* fl, int* g);
void launch_dense_level(hipStream_t s, const SceneDesc* d, const Topo& T, int l, int* cell, float* m, float* v);
void launch_dense_node(hipStream_t s, const SceneDesc* d, const Topo& T, int* out);

// learned windows (kernels.hip), single scene: swap particle buffer p with the spare one; save
// the counters and reset the check's; report the check's counters to host-mapped `out` [4]
void launch_swap_spare(hipStream_t s, SceneDesc* d, int p);
void launch_window_start(hipStream_t s, const SceneDesc* d, unsigned long long* snap);
void launch_window_end(hipStream_t s, const SceneDesc* d, const unsigned long long* snap, unsigned long long* out);

// learned refinement prediction (learned.hip), single scene
void launch_unit_features(hipStream_t s, const SceneDesc* d, int dim, int n, int par, const LearnDesc& U);
void launch_unit_predict(hipStream_t s, const LearnDesc& U, int L, int* cls_out);
void launch_unit_activate(hipStream_t s, const SceneDesc* d, int dim, int n, int par, const LearnDesc& U);
void launch_unit_labels_clear(hipStream_t s, const LearnDesc& U);
void launch_unit_labels_accumulate(hipStream_t s, const SceneDesc* d, int dim, int n, int par, const LearnDesc& U);
void launch_unit_label_classes(hipStream_t s, const LearnDesc& U, int L, int* out);

}  // namespace agimp
