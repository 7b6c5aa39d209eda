// C-ABI entry points (include/agimp.h): device state of one simulator, batches of
// simulators, and the host-side orchestration of a substep (adaptive_gimp.py:323-352).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "agimp.h"
#include "common.hpp"
#include "launch.hpp"
#include "svd.hpp"

using namespace agimp;

// One captured substep (refine + bucketing + physics) of a set of scenes, replayed with
// hipGraphLaunch.  Kernels read dt from the device descriptors, so a graph stays valid when
// dt changes; one graph is captured per value of the current-buffer index (G2P swaps them).
struct Captured {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

struct agimp_handle {
  agimp_config cfg{};
  int D = 0, L = 0;
  int n = 0;
  hipStream_t stream = nullptr;
  SceneDesc hd{};             // host copy of the scene descriptor
  SceneDesc* dd = nullptr;    // device copy (one entry)
  Dims dims{};
  int host_parity = 0;        // index of the current particle buffer in hd.P
  int has_particles = 0;
  int topo_valid = 0;
  int max_chunks = 0;
  float graph_dt = -1.f;      // dt stored in the device descriptor for graph replays
  Captured graph[2];          // one substep starting from buffer 0 / buffer 1
  Captured chain[2];          // SUBSTEP_CHAIN substeps in one graph, from buffer 0 / buffer 1
  std::vector<void*> allocs;
  std::string err;
  // phase timing with HIP events on the handle's stream (agimp_set_timing)
  int timing = 0;
  struct Mark { int phase; hipEvent_t a, b; };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> event_pool;
  double phase_ms[AGIMP_NPHASES] = {};
  int64_t phase_calls[AGIMP_NPHASES] = {};
  // per-kernel timing: events around every launch while timing is on
  struct KMark { int kernel; hipEvent_t a, b; };
  struct KStat { double total_ms = 0.0, min_ms = 0.0, max_ms = 0.0; int64_t calls = 0; };
  std::vector<KMark> kmarks;
  std::vector<std::string> knames;  // kernel index -> launcher's kernel name
  std::vector<KStat> kstats;
  struct Timer : KernelTimer {
    agimp_handle* h = nullptr;
    int kernel = -1;
    hipEvent_t a = nullptr;
    void begin(const char* name, hipStream_t s) override;
    void end(hipStream_t s) override;
  } ktimer;
  double mean_mass = 1.0;     // mean particle mass (reference density of the unit features)
  // learned refinement prediction (agimp_learned_*)
  int learn_on = 0;
  int model_set = 0;
  int window = 20;
  LearnDesc U{};
  int* cls_dev = nullptr;     // [nu] per-candidate classes (labels or predictions)
  unsigned long long* snap_cnt = nullptr;  // the counters at a window start (the particles: hd.spare)
  unsigned long long* win_out = nullptr;   // [4] pinned host memory mapped for the device: a window's check
  unsigned long long* win_out_dev = nullptr;
  Captured fixed[2];          // substep without the rebuild (bucketing + physics), per buffer index
  Captured fixed_chain[2];    // a whole window of such substeps in one graph, per starting buffer
  int fixed_chain_len = 0;    // the window length they were captured for
  int64_t lstats[AGIMP_NLSTATS] = {};
};

struct agimp_batch {
  std::vector<agimp_handle*> scenes;
  std::vector<SceneDesc> host;  // batch copies of the descriptors, buffers ordered so that
  int par = 0;                  // host[i].P[par] is scene i's current buffer
  int device = 0;
  hipStream_t stream = nullptr;
  SceneDesc* dd = nullptr;
  Dims dims{};
  Captured graph[2];
  Captured chain[2];
  std::string err;
};

namespace {

int fail(agimp_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, call)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return fail((h), AGIMP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class Tp>
int dalloc(agimp_handle* h, Tp** p, size_t count) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(count * sizeof(Tp), 256);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(h, AGIMP_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  e = hipMemsetAsync(q, 0, bytes, h->stream);
  if (e != hipSuccess) return fail(h, AGIMP_EHIP, hipGetErrorString(e));
  h->allocs.push_back(q);
  *p = (Tp*)q;
  return 0;
}

int align_size(int x, int a) { return (x + (a - 1)) & ~(a - 1); }  // utils.py:44-45

Physics make_physics(const agimp_config& c) {
  Physics p{};
  for (int a = 0; a < 3; ++a) p.gravity[a] = c.gravity[a];
  p.boundary_gap = c.boundary_gap;
  p.wall_mode = c.wall_mode;
  p.plastic = c.plastic;
  p.theta_c = c.theta_c;
  p.theta_s = c.theta_s;
  p.hardening = c.hardening;
  double E = c.E, nu = c.nu;
  p.mu0 = E / (2.0 * (1.0 + nu));  // adaptive_gimp.py:53
  p.la0 = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  p.p_vol = c.p_vol;
  p.slab_enable = c.slab_enable;
  p.slab_axis = c.slab_axis;
  p.slab_lo = c.slab_lo;
  p.slab_hi = c.slab_hi;
  p.dt = 0.f;
  return p;
}

int check_ready(agimp_handle* h) {
  if (!h) return AGIMP_EINVAL;
  if (!h->has_particles) return fail(h, AGIMP_ESTATE, "particles not set (agimp_set_particles)");
  return 0;
}

int upload_desc(agimp_handle* h) {
  HIPCHK(h, hipMemcpyAsync(h->dd, &h->hd, sizeof(SceneDesc), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

hipEvent_t take_event(agimp_handle* h) {
  if (!h->event_pool.empty()) {
    hipEvent_t e = h->event_pool.back();
    h->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// begin/end a timed phase: events are recorded only while timing is on
struct PhaseScope {
  agimp_handle* h;
  int phase;
  hipEvent_t a = nullptr;
  PhaseScope(agimp_handle* hh, int p) : h(hh), phase(p) {
    if (h->timing) {
      a = take_event(h);
      if (a) (void)hipEventRecord(a, h->stream);
    }
  }
  ~PhaseScope() {
    if (!a) return;
    hipEvent_t b = take_event(h);
    if (!b) return;
    (void)hipEventRecord(b, h->stream);
    h->marks.push_back({phase, a, b});
  }
};

// fold recorded event pairs into the per-phase and per-kernel totals (synchronises the stream)
int collect_marks(agimp_handle* h) {
  if (h->marks.empty() && h->kmarks.empty()) return 0;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (auto& m : h->marks) {
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, m.a, m.b));
    h->phase_ms[m.phase] += ms;
    h->phase_calls[m.phase] += 1;
    h->event_pool.push_back(m.a);
    h->event_pool.push_back(m.b);
  }
  h->marks.clear();
  for (auto& m : h->kmarks) {
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, m.a, m.b));
    agimp_handle::KStat& st = h->kstats[m.kernel];
    st.min_ms = st.calls ? std::min(st.min_ms, (double)ms) : (double)ms;
    st.max_ms = st.calls ? std::max(st.max_ms, (double)ms) : (double)ms;
    st.total_ms += ms;
    st.calls += 1;
    h->event_pool.push_back(m.a);
    h->event_pool.push_back(m.b);
  }
  h->kmarks.clear();
  return 0;
}

// installs the handle's kernel timer on this thread for the scope (timing on, no capture)
struct KernelTimerScope {
  bool on;
  explicit KernelTimerScope(agimp_handle* h) : on(h->timing != 0) {
    if (on) {
      h->ktimer.h = h;
      set_kernel_timer(&h->ktimer);
    }
  }
  ~KernelTimerScope() {
    if (on) set_kernel_timer(nullptr);
  }
};

}  // namespace

void agimp_handle::Timer::begin(const char* name, hipStream_t s) {
  kernel = -1;
  for (size_t i = 0; i < h->knames.size(); ++i)
    if (h->knames[i] == name) { kernel = (int)i; break; }
  if (kernel < 0) {
    kernel = (int)h->knames.size();
    h->knames.push_back(name);
    h->kstats.emplace_back();
  }
  a = take_event(h);
  if (a) (void)hipEventRecord(a, s);
}

void agimp_handle::Timer::end(hipStream_t s) {
  if (!a) return;
  hipEvent_t b = take_event(h);
  if (!b) return;
  (void)hipEventRecord(b, s);
  h->kmarks.push_back({kernel, a, b});
  a = nullptr;
}

namespace {

Dims single_dims(const agimp_handle* h) {
  Dims m{};
  m.dim = h->D;
  m.nscenes = 1;
  m.maxL = h->L;
  m.any_single = h->L == 1;
  m.any_static = h->hd.static_mode;
  m.any_dynamic = !h->hd.static_mode;
  m.max_n = (size_t)h->n;
  m.max_total_blocks = (size_t)h->hd.T.total_blocks;
  m.max_fill_words = h->hd.w_cells + h->hd.w_node + 2 * h->hd.w_flags;
  m.list_grid = 2048;
  m.fill_grid = 4096;
  m.p2g_grid = std::max(1, std::min(h->max_chunks, 8192));
  m.grid_grid = 4096;
  return m;
}

int do_refine(agimp_handle* h) {
  KernelTimerScope kt(h);
  PhaseScope ps(h, AGIMP_PHASE_REFINE);
  launch_refine(h->stream, h->dd, h->dims, h->host_parity);
  h->topo_valid = 1;
  HIPCHK(h, hipGetLastError());
  return 0;
}

int do_physics(agimp_handle* h, float dt) {
  KernelTimerScope kt(h);
  {
    PhaseScope ps(h, AGIMP_PHASE_BUCKET);
    launch_bucket(h->stream, h->dd, h->dims, h->host_parity);
  }
  {
    PhaseScope ps(h, AGIMP_PHASE_P2G);
    launch_p2g(h->stream, h->dd, h->dims, dt, h->host_parity);
  }
  {
    PhaseScope ps(h, AGIMP_PHASE_GRID);
    launch_grid(h->stream, h->dd, h->dims, dt);
  }
  {
    PhaseScope ps(h, AGIMP_PHASE_G2P);
    launch_g2p(h->stream, h->dd, h->dims, dt, h->host_parity);
  }
  h->host_parity ^= 1;
  HIPCHK(h, hipGetLastError());
  if (h->marks.size() + h->kmarks.size() > 4096) {
    int rc = collect_marks(h);
    if (rc) return rc;
  }
  return 0;
}

// capture one full substep of the scenes behind `dd` into `cap` (dt taken from the descriptors)
hipError_t capture_substep(hipStream_t s, const SceneDesc* dd, const Dims& m, int par, Captured& cap) {
  cap.release();
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  launch_refine(s, dd, m, par, nullptr, true);  // chained: fewer clears and flag copies
  launch_bucket(s, dd, m, par, true);
  launch_p2g(s, dd, m, 0.f, par);
  launch_grid(s, dd, m, 0.f);
  launch_g2p(s, dd, m, 0.f, par);
  hipError_t le = hipGetLastError();
  e = hipStreamEndCapture(s, &cap.graph);
  if (le != hipSuccess) return le;
  if (e != hipSuccess) return e;
  return hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0);
}

// SUBSTEP_CHAIN full substeps in one graph, starting from buffer `par` (the buffer index flips
// after each substep, so an even count ends on `par` again): one host launch advances them all
constexpr int SUBSTEP_CHAIN = 8;
hipError_t capture_chain(hipStream_t s, const SceneDesc* dd, const Dims& m, int par, Captured& cap) {
  cap.release();
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  // inside the chain each G2P launch also clears its octree for the next substep's rebuild;
  // the last keeps it (the caller may read it), and the first rebuild clears for itself
  for (int i = 0; i < SUBSTEP_CHAIN; ++i) {
    const int p = par ^ (i & 1);
    launch_refine(s, dd, m, p, nullptr, true, i > 0);
    launch_bucket(s, dd, m, p, true);
    launch_p2g(s, dd, m, 0.f, p);
    launch_grid(s, dd, m, 0.f);
    launch_g2p(s, dd, m, 0.f, p, i + 1 < SUBSTEP_CHAIN ? G2P_NEXT_REBUILD : G2P_NEXT_NONE);
  }
  hipError_t le = hipGetLastError();
  e = hipStreamEndCapture(s, &cap.graph);
  if (le != hipSuccess) return le;
  if (e != hipSuccess) return e;
  return hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0);
}

// replay n substeps starting from buffer *par: whole chains while at least SUBSTEP_CHAIN remain,
// then single-substep graphs; captures what is missing on first use
hipError_t replay_substeps(hipStream_t s, const SceneDesc* dd, const Dims& m, Captured* one, Captured* chain,
                           int* par, int n) {
  static_assert(SUBSTEP_CHAIN % 2 == 0, "a chain must end on the buffer it starts from");
  for (int p = 0; p < 2; ++p) {
    if (!one[p].exec) {
      hipError_t e = capture_substep(s, dd, m, p, one[p]);
      if (e != hipSuccess) { one[p].release(); return e; }
    }
  }
  if (n >= SUBSTEP_CHAIN && !chain[*par].exec) {
    hipError_t e = capture_chain(s, dd, m, *par, chain[*par]);
    if (e != hipSuccess) { chain[*par].release(); return e; }
  }
  int i = 0;
  for (; i + SUBSTEP_CHAIN <= n; i += SUBSTEP_CHAIN) {
    hipError_t e = hipGraphLaunch(chain[*par].exec, s);
    if (e != hipSuccess) return e;
  }
  for (; i < n; ++i) {
    hipError_t e = hipGraphLaunch(one[*par].exec, s);
    if (e != hipSuccess) return e;
    *par ^= 1;
  }
  return hipSuccess;
}

// substep without the rebuild (bucketing + physics) of one scene, for reused octrees
hipError_t capture_fixed(hipStream_t s, const SceneDesc* dd, const Dims& m, int par, Captured& cap) {
  cap.release();
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  launch_bucket(s, dd, m, par);
  launch_p2g(s, dd, m, 0.f, par);
  launch_grid(s, dd, m, 0.f);
  launch_g2p(s, dd, m, 0.f, par);
  hipError_t le = hipGetLastError();
  e = hipStreamEndCapture(s, &cap.graph);
  if (le != hipSuccess) return le;
  if (e != hipSuccess) return e;
  return hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0);
}

// n substeps without the rebuild in one graph, starting from buffer `par` (a window of reuse of
// one predicted octree: the host checks the refinement criterion only at the window's end, so
// the window's substeps need no host launch each)
hipError_t capture_fixed_chain(hipStream_t s, SceneDesc* dd, const Dims& m, int par, int n, Captured& cap) {
  cap.release();
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  // each G2P launch but the last prepares the next substep's bucketing pass; after the first
  // substep the window start's buffer (par, read by it) is swapped with the spare one, which
  // the second substep then overwrites instead
  for (int i = 0; i < n; ++i) {
    const int p = par ^ (i & 1);
    launch_bucket(s, dd, m, p, false, i > 0);
    launch_p2g(s, dd, m, 0.f, p);
    launch_grid(s, dd, m, 0.f);
    launch_g2p(s, dd, m, 0.f, p, i + 1 < n ? G2P_NEXT_FIXED : G2P_NEXT_NONE);
    if (i == 0) launch_swap_spare(s, dd, par);
  }
  hipError_t le = hipGetLastError();
  e = hipStreamEndCapture(s, &cap.graph);
  if (le != hipSuccess) return le;
  if (e != hipSuccess) return e;
  return hipGraphInstantiate(&cap.exec, cap.graph, nullptr, nullptr, 0);
}

// dt of graph replays lives in the device descriptor
int set_graph_dt(agimp_handle* h, float dt) {
  if (dt == h->graph_dt) return 0;
  h->hd.ph.dt = dt;
  int rc = upload_desc(h);
  if (rc) return rc;
  h->graph_dt = dt;
  return 0;
}

int learn_ready(agimp_handle* h) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!h->learn_on) return fail(h, AGIMP_ESTATE, "learned prediction not enabled (agimp_learned_enable)");
  return 0;
}

// candidates' features of the current particles (device only)
int learn_features(agimp_handle* h, float dt) {
  KernelTimerScope kt(h);
  h->U.win = (float)h->window * dt;
  launch_unit_features(h->stream, h->dd, h->D, h->n, h->host_parity, h->U);
  HIPCHK(h, hipGetLastError());
  return 0;
}

int learn_count(agimp_handle* h, int which, int64_t* out) {
  int c = 0;
  HIPCHK(h, hipMemcpyAsync(&c, h->U.counts + which, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  *out = c;
  return 0;
}

// predicted octree: features -> network -> refine with the predicted units' cells
int learn_rebuild(agimp_handle* h, float dt) {
  PhaseScope ps(h, AGIMP_PHASE_REFINE);
  int rc = learn_features(h, dt);
  if (rc) return rc;
  KernelTimerScope kt(h);
  launch_unit_predict(h->stream, h->U, h->L, nullptr);
  launch_refine(h->stream, h->dd, h->dims, h->host_parity, &h->U);
  h->topo_valid = 1;
  HIPCHK(h, hipGetLastError());
  return 0;
}

size_t particle_floats(const agimp_handle* h) {
  const int NF = (h->D == 2) ? PLayout<2>::NF : PLayout<3>::NF;
  return (size_t)NF * (size_t)h->hd.P[0].stride;
}


// Staging of agimp_get_particles: x, v [n][d], F, C [n][d][d], Jp [n] float32, fl, g [n] int32
// in caller order, written by k_download, then copied to the host.
size_t particle_staging_bytes(const agimp_handle* h) {
  const size_t n = (size_t)h->n, nv = n * h->D, nm = n * h->D * h->D;
  return (2 * nv + 2 * nm + n) * sizeof(float) + 2 * n * sizeof(int);
}

int download_particles(agimp_handle* h, char* staging, float* x, float* v, float* F, float* C, float* Jp,
                       int32_t* fl, int32_t* g) {
  const int D = h->D, n = h->n;
  size_t nv = (size_t)n * D, nm = (size_t)n * D * D;
  float* dx = (float*)staging;
  float* dv = dx + nv;
  float* dF = dv + nv;
  float* dC = dF + nm;
  float* dJ = dC + nm;
  int* dfl = (int*)(dJ + n);
  int* dg = dfl + n;
  launch_download(h->stream, h->dd, h->host_parity, D, n, dx, dv, dF, dC, dJ, dfl, dg);
  hipError_t e = hipGetLastError();
  if (x && e == hipSuccess) e = hipMemcpyAsync(x, dx, nv * 4, hipMemcpyDeviceToHost, h->stream);
  if (v && e == hipSuccess) e = hipMemcpyAsync(v, dv, nv * 4, hipMemcpyDeviceToHost, h->stream);
  if (F && e == hipSuccess) e = hipMemcpyAsync(F, dF, nm * 4, hipMemcpyDeviceToHost, h->stream);
  if (C && e == hipSuccess) e = hipMemcpyAsync(C, dC, nm * 4, hipMemcpyDeviceToHost, h->stream);
  if (Jp && e == hipSuccess) e = hipMemcpyAsync(Jp, dJ, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (fl && e == hipSuccess) e = hipMemcpyAsync(fl, dfl, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (g && e == hipSuccess) e = hipMemcpyAsync(g, dg, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return fail(h, AGIMP_EHIP, std::string("get_particles: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace

// Scratch of agimp_get_particles_noalloc: the staging buffer that agimp_get_particles allocates
// on every call, held instead in one file-scope device array.  Each GPU the library runs on
// holds its own copy, at its own address, until the process ends.  In host code the name
// `scratch` is only a stand-in for the symbol: its device address is taken once per device
// (agimp_create, with the device current) and only that address is used.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::vector<char*> g_scratch_ptr;  // per device ordinal; nullptr until set up

// resolve the scratch array's address on the current device (`device`) once
static int setup_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0) return AGIMP_EINVAL;
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return AGIMP_EHIP;
  g_scratch_ptr[device] = (char*)p;
  return 0;
}

static char* scratch_of_device(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && (size_t)device < g_scratch_ptr.size()) ? g_scratch_ptr[device] : nullptr;
}

extern "C" {

int agimp_config_init(agimp_config* c, int32_t dim, int32_t level, int32_t coarsest_size, int64_t n_particles) {
  if (!c) return AGIMP_EINVAL;
  std::memset(c, 0, sizeof(*c));
  c->dim = dim;
  c->level = level;
  c->coarsest_size = coarsest_size;
  c->n_particles = n_particles;
  c->boundary_gap = 4;
  c->E = 5e3f;
  c->nu = 0.2f;
  c->gravity[1] = -100.0f;
  c->p_vol = 1.0f;
  c->radius = 0.5f;
  c->wall_mode = 0;
  c->plastic = 0;
  c->theta_c = 2.5e-2f;
  c->theta_s = 7.5e-3f;
  c->hardening = 0.0f;
  c->slab_axis = 1;
  // measured best on the 3D 1M and 4M scenes and on the 2D 200k band
  c->particles_per_chunk = 1024;
  c->device = 0;
  return 0;
}

int agimp_create(const agimp_config* cfg, agimp_handle** out) {
  if (!cfg || !out) return AGIMP_EINVAL;
  *out = nullptr;
  const agimp_config& c = *cfg;
  if (c.dim != 2 && c.dim != 3) return AGIMP_EINVAL;
  if (c.level < 1 || c.level > MAXL) return AGIMP_EINVAL;
  if (c.coarsest_size < 1 || c.n_particles < 1 || c.n_particles > (int64_t)1 << 30) return AGIMP_EINVAL;
  if (std::fabs(c.radius - 0.5f) > 0.0f) return AGIMP_EINVAL;  // the B-spline form needs r = 0.5
  if (c.particles_per_chunk < AGIMP_MIN_CHUNK || c.particles_per_chunk > AGIMP_MAX_CHUNK) return AGIMP_EINVAL;
  if (c.slab_enable && (c.slab_axis < 0 || c.slab_axis >= c.dim)) return AGIMP_EINVAL;
  agimp_handle* h = new agimp_handle();
  h->cfg = c;
  h->D = c.dim;
  h->L = c.level;
  h->n = (int)c.n_particles;
  int rc = 0;
  if (hipSetDevice(c.device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return AGIMP_EHIP;
  }
  if (setup_scratch(c.device) != 0) {  // the device address of agimp_get_particles_noalloc's scratch
    (void)hipStreamDestroy(h->stream);
    delete h;
    return AGIMP_EHIP;
  }
  SceneDesc& sd = h->hd;
  Topo& T = sd.T;
  T.dim = c.dim;
  T.L = c.level;
  T.c = c.coarsest_size;
  T.finest_size = c.coarsest_size << (c.level - 1);  // adaptive_gimp.py:17
  T.pow2 = (c.coarsest_size & (c.coarsest_size - 1)) == 0;
  const int p4 = c.dim == 2 ? 16 : 64;
  size_t cells = 0, flags = 0;
  int gbase = 0;
  for (int l = 0; l < T.L; ++l) {
    LevelDesc& Ll = T.lv[l];
    Ll.n = c.coarsest_size << l;
    Ll.e = 4 * align_size(Ll.n / 4 + 1, 4);  // adaptive_gimp.py:36-37
    Ll.E = Ll.e + 2 * PAD;
    Ll.NB = Ll.E / 4;
    Ll.shift = T.L - 1 - l;
    Ll.nblocks = (c.dim == 2) ? Ll.NB * Ll.NB : Ll.NB * Ll.NB * Ll.NB;
    Ll.gbase = gbase;
    gbase += Ll.nblocks;
    // coarsest_dx / 2**l in f32 (adaptive_gimp.py:16, :235)
    Ll.dx = (float)(1.0 / c.coarsest_size) / (float)(1 << l);
    Ll.inv_dx = (float)(c.coarsest_size << l);
    cells += (size_t)Ll.nblocks * p4;
    flags += (size_t)Ll.nblocks;
  }
  T.total_blocks = gbase;
  uint8_t* cellbuf = nullptr;
  uint8_t* cornerbuf = nullptr;
  uint8_t* splitbuf = nullptr;
  float4* gridbuf = nullptr;
  uint8_t* flagbuf = nullptr;
  int* slistbuf = nullptr;
  rc = rc ? rc : dalloc(h, &cellbuf, (cells + 3) & ~(size_t)3);
  rc = rc ? rc : dalloc(h, &cornerbuf, (cells + 3) & ~(size_t)3);
  rc = rc ? rc : dalloc(h, &splitbuf, (cells + 3) & ~(size_t)3);
  rc = rc ? rc : dalloc(h, &gridbuf, cells);
  rc = rc ? rc : dalloc(h, &flagbuf, (flags + 3) & ~(size_t)3);
  rc = rc ? rc : dalloc(h, &slistbuf, flags);
  rc = rc ? rc : dalloc(h, &sd.vflag, (flags + 3) / 4);
  rc = rc ? rc : dalloc(h, &T.node, (size_t)T.lv[T.L - 1].nblocks * p4);
  rc = rc ? rc : dalloc(h, &T.scount, MAXL);
  rc = rc ? rc : dalloc(h, &T.cflag, (flags + 3) & ~(size_t)3);
  rc = rc ? rc : dalloc(h, &T.clist, flags);
  rc = rc ? rc : dalloc(h, &T.ccount, 4);
  if (rc) { agimp_destroy(h); return rc; }
  size_t co = 0, fo = 0;
  for (int l = 0; l < T.L; ++l) {
    T.lv[l].cell = as_global(cellbuf + co);
    T.lv[l].corner = as_global(cornerbuf + co);
    T.lv[l].split = as_global(splitbuf + co);
    T.lv[l].grid = as_global(gridbuf + co);
    T.lv[l].sflag = as_global(flagbuf + fo);
    T.lv[l].slist = as_global(slistbuf + fo);
    co += (size_t)T.lv[l].nblocks * p4;
    fo += (size_t)T.lv[l].nblocks;
  }
  sd.w_cells = cells / 4;
  sd.w_node = (size_t)T.lv[T.L - 1].nblocks * p4 / 4;
  sd.w_flags = (flags + 3) / 4;
  // particles: two structure-of-arrays buffers
  const int NF = (c.dim == 2) ? PLayout<2>::NF : PLayout<3>::NF;
  int64_t stride = ((int64_t)h->n + 63) & ~(int64_t)63;
  for (int b = 0; b < 2; ++b) {
    PBuf& pb = sd.P[b];
    pb.stride = stride;
    rc = rc ? rc : dalloc(h, &pb.f, (size_t)NF * stride);
    rc = rc ? rc : dalloc(h, &pb.g, stride);
    rc = rc ? rc : dalloc(h, &pb.id, stride);
  }
  SortBufs& S = sd.S;
  rc = rc ? rc : dalloc(h, &S.gb, stride);
  rc = rc ? rc : dalloc(h, &S.rk, stride);
  rc = rc ? rc : dalloc(h, &S.lkey, stride);
  rc = rc ? rc : dalloc(h, &S.pkey, stride);
  rc = rc ? rc : dalloc(h, &S.perm, stride);
  rc = rc ? rc : dalloc(h, &S.flv, stride);
  rc = rc ? rc : dalloc(h, &S.bcnt, T.total_blocks);
  rc = rc ? rc : dalloc(h, &S.boff, T.total_blocks);
  rc = rc ? rc : dalloc(h, &S.choff, T.total_blocks);
  rc = rc ? rc : dalloc(h, &S.tiles, (T.total_blocks + 2047) / 2048 + 1);
  rc = rc ? rc : dalloc(h, &S.totals, 4);
  rc = rc ? rc : dalloc(h, &S.frozen, 4);
  // enough chunks for the smallest chunk size the handle can be switched to
  h->max_chunks = (int)std::min<int64_t>((int64_t)h->n / AGIMP_MIN_CHUNK + std::min<int64_t>(h->n, T.total_blocks) + 1,
                                         (int64_t)1 << 30);
  rc = rc ? rc : dalloc(h, &S.chunks, h->max_chunks);
  rc = rc ? rc : dalloc(h, &sd.cnt, NCOUNTERS);
  rc = rc ? rc : dalloc(h, &h->dd, 1);
  if (rc) { agimp_destroy(h); return rc; }
  sd.ph = make_physics(c);
  sd.n = h->n;
  sd.static_mode = 0;
  sd.CH = c.particles_per_chunk;
  h->dims = single_dims(h);
  rc = upload_desc(h);
  if (rc) { agimp_destroy(h); return rc; }
  *out = h;
  return 0;
}

int agimp_destroy(agimp_handle* h) {
  if (!h) return AGIMP_EINVAL;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->graph[0].release();
  h->graph[1].release();
  h->chain[0].release();
  h->chain[1].release();
  h->fixed[0].release();
  h->fixed[1].release();
  h->fixed_chain[0].release();
  h->fixed_chain[1].release();
  if (h->win_out) (void)hipHostFree(h->win_out);
  for (auto& m : h->marks) { (void)hipEventDestroy(m.a); (void)hipEventDestroy(m.b); }
  for (auto& m : h->kmarks) { (void)hipEventDestroy(m.a); (void)hipEventDestroy(m.b); }
  if (h->ktimer.a) (void)hipEventDestroy(h->ktimer.a);
  for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->hd.st_ret) (void)hipFree((void*)h->hd.st_ret);
  if (h->hd.st_act) (void)hipFree((void*)h->hd.st_act);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int agimp_set_particles(agimp_handle* h, const float* x, const float* v, const float* F, const float* C,
                        const float* m, const int32_t* g, const float* Jp) {
  if (!h || !x || !v || !F || !C || !m || !g) return AGIMP_EINVAL;
  const int D = h->D, n = h->n;
  size_t nv = (size_t)n * D, nm = (size_t)n * D * D;
  float *dx, *dv, *dF, *dC, *dm, *dJ = nullptr;
  int* dg;
  size_t bytes = (2 * nv + 2 * nm + 2 * (size_t)n) * sizeof(float) + (size_t)n * sizeof(int);
  char* tmp = nullptr;
  HIPCHK(h, hipMalloc(&tmp, bytes));
  dx = (float*)tmp;
  dv = dx + nv;
  dF = dv + nv;
  dC = dF + nm;
  dm = dC + nm;
  dJ = dm + n;
  dg = (int*)(dJ + n);
  hipError_t e = hipSuccess;
  e = e ? e : hipMemcpyAsync(dx, x, nv * 4, hipMemcpyHostToDevice, h->stream);
  e = e ? e : hipMemcpyAsync(dv, v, nv * 4, hipMemcpyHostToDevice, h->stream);
  e = e ? e : hipMemcpyAsync(dF, F, nm * 4, hipMemcpyHostToDevice, h->stream);
  e = e ? e : hipMemcpyAsync(dC, C, nm * 4, hipMemcpyHostToDevice, h->stream);
  e = e ? e : hipMemcpyAsync(dm, m, (size_t)n * 4, hipMemcpyHostToDevice, h->stream);
  e = e ? e : hipMemcpyAsync(dg, g, (size_t)n * 4, hipMemcpyHostToDevice, h->stream);
  if (Jp && e == hipSuccess) e = hipMemcpyAsync(dJ, Jp, (size_t)n * 4, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) {
    launch_upload(h->stream, h->dd, h->host_parity, D, n, dx, dv, dF, dC, dm, dg, Jp ? dJ : nullptr);
    launch_stress(h->stream, h->dd, D, n, h->host_parity);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(h, AGIMP_EHIP, std::string("set_particles: ") + hipGetErrorString(e));
  double msum = 0.0;
  for (int i = 0; i < n; ++i) msum += m[i];
  h->mean_mass = msum > 0.0 ? msum / n : 1.0;
  h->has_particles = 1;
  return 0;
}

int agimp_set_static_levels(agimp_handle* h, const int8_t* ret_level, const int8_t* act_level) {
  if (!h) return AGIMP_EINVAL;
  h->graph[0].release();
  h->graph[1].release();
  h->chain[0].release();
  h->chain[1].release();
  if (!ret_level || !act_level) {
    h->hd.static_mode = 0;
    h->dims = single_dims(h);
    return upload_desc(h);
  }
  size_t E = (size_t)h->hd.T.lv[h->L - 1].E;
  size_t cnt = (h->D == 2) ? E * E : E * E * E;
  for (size_t i = 0; i < cnt; ++i)
    if (ret_level[i] < 0 || ret_level[i] >= h->L || act_level[i] < -1 || act_level[i] >= h->L)
      return fail(h, AGIMP_EINVAL, "static level map entry outside [0, level)");
  if (!h->hd.st_ret) {
    int8_t *r = nullptr, *a = nullptr;
    HIPCHK(h, hipMalloc(&r, cnt));
    HIPCHK(h, hipMalloc(&a, cnt));
    h->hd.st_ret = as_global(r);
    h->hd.st_act = as_global(a);
  }
  HIPCHK(h, hipMemcpyAsync((void*)h->hd.st_ret, ret_level, cnt, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync((void*)h->hd.st_act, act_level, cnt, hipMemcpyHostToDevice, h->stream));
  h->hd.static_mode = 1;
  h->dims = single_dims(h);
  return upload_desc(h);
}

int agimp_set_particles_per_chunk(agimp_handle* h, int32_t ch) {
  if (!h) return AGIMP_EINVAL;
  if (ch < AGIMP_MIN_CHUNK || ch > AGIMP_MAX_CHUNK) return fail(h, AGIMP_EINVAL, "chunk size out of range");
  h->hd.CH = ch;
  return upload_desc(h);
}

int agimp_refine(agimp_handle* h) {
  int rc = check_ready(h);
  if (rc) return rc;
  return do_refine(h);
}

int agimp_substep(agimp_handle* h, float dt) {
  int rc = check_ready(h);
  if (rc) return rc;
  rc = do_refine(h);
  if (rc) return rc;
  return do_physics(h, dt);
}

int agimp_substep_fixed_topology(agimp_handle* h, float dt) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!h->topo_valid) return fail(h, AGIMP_ESTATE, "no octree yet: call agimp_refine or agimp_substep first");
  return do_physics(h, dt);
}

int agimp_advance(agimp_handle* h, float dt, int32_t n) {
  if (n < 0) return AGIMP_EINVAL;
  int rc = check_ready(h);
  if (rc) return rc;
  if (n == 0) return 0;
  if (h->timing) {  // phase timing needs the eager path
    for (int i = 0; i < n; ++i) {
      rc = agimp_substep(h, dt);
      if (rc) return rc;
    }
    return 0;
  }
  // replay captured substeps (chains of SUBSTEP_CHAIN, then single ones); dt lives in the
  // device descriptor
  rc = set_graph_dt(h, dt);
  if (rc) return rc;
  hipError_t e = replay_substeps(h->stream, h->dd, h->dims, h->graph, h->chain, &h->host_parity, n);
  if (e != hipSuccess) return fail(h, AGIMP_EHIP, std::string("graph replay: ") + hipGetErrorString(e));
  h->topo_valid = 1;
  return 0;
}

int agimp_get_particles(agimp_handle* h, float* x, float* v, float* F, float* C, float* Jp, int32_t* fl,
                        int32_t* g) {
  if (!h) return AGIMP_EINVAL;
  char* tmp = nullptr;
  HIPCHK(h, hipMalloc(&tmp, particle_staging_bytes(h)));
  const int rc = download_particles(h, tmp, x, v, F, C, Jp, fl, g);
  (void)hipFree(tmp);
  return rc;
}

int agimp_get_particles_noalloc(agimp_handle* h, float* x, float* v, float* F, float* C, float* Jp, int32_t* fl,
                                int32_t* g) {
  if (!h) return AGIMP_EINVAL;
  const size_t bytes = particle_staging_bytes(h);
  if (bytes > NOALLOC_SCRATCH_BYTES)
    return fail(h, AGIMP_EINVAL, "get_particles_noalloc: " + std::to_string(h->n) + " particles need " +
                                     std::to_string(bytes) + " bytes of staging, above the " +
                                     std::to_string(NOALLOC_SCRATCH_BYTES) + "-byte scratch array");
  char* p = scratch_of_device(h->cfg.device);
  if (!p) return fail(h, AGIMP_ESTATE, "get_particles_noalloc: no scratch address for this device");
  return download_particles(h, p, x, v, F, C, Jp, fl, g);
}

int agimp_get_level(agimp_handle* h, int32_t l, int32_t* cell_mask, float* m, float* v) {
  if (!h || l < 0 || l >= h->L) return AGIMP_EINVAL;
  size_t tot = 1;
  for (int a = 0; a < h->D; ++a) tot *= (size_t)h->hd.T.lv[l].e;
  char* tmp = nullptr;
  HIPCHK(h, hipMalloc(&tmp, tot * (4 + 4 + 4 * (size_t)h->D)));
  int* dc = (int*)tmp;
  float* dm = (float*)(dc + tot);
  float* dv = dm + tot;
  launch_dense_level(h->stream, h->dd, h->hd.T, l, dc, dm, dv);
  hipError_t e = hipGetLastError();
  if (cell_mask && e == hipSuccess) e = hipMemcpyAsync(cell_mask, dc, tot * 4, hipMemcpyDeviceToHost, h->stream);
  if (m && e == hipSuccess) e = hipMemcpyAsync(m, dm, tot * 4, hipMemcpyDeviceToHost, h->stream);
  if (v && e == hipSuccess) e = hipMemcpyAsync(v, dv, tot * 4 * h->D, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(h, AGIMP_EHIP, std::string("get_level: ") + hipGetErrorString(e));
  return 0;
}

int agimp_get_node_mask(agimp_handle* h, int32_t* out) {
  if (!h || !out) return AGIMP_EINVAL;
  size_t tot = 1;
  for (int a = 0; a < h->D; ++a) tot *= (size_t)h->hd.T.lv[h->L - 1].e;
  int* d = nullptr;
  HIPCHK(h, hipMalloc(&d, tot * 4));
  launch_dense_node(h->stream, h->dd, h->hd.T, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d, tot * 4, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(h, AGIMP_EHIP, std::string("get_node_mask: ") + hipGetErrorString(e));
  return 0;
}

int agimp_get_counters(agimp_handle* h, int64_t* out, int32_t n) {
  if (!h || !out || n < 0) return AGIMP_EINVAL;
  unsigned long long tmp[NCOUNTERS];
  HIPCHK(h, hipMemcpyAsync(tmp, h->hd.cnt, sizeof(tmp), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int i = 0; i < n && i < NCOUNTERS; ++i) out[i] = (int64_t)tmp[i];
  return 0;
}

int agimp_reset_counters(agimp_handle* h) {
  if (!h) return AGIMP_EINVAL;
  unsigned long long zero[NCOUNTERS] = {};
  HIPCHK(h, hipMemcpyAsync(h->hd.cnt, zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int agimp_set_timing(agimp_handle* h, int32_t on) {
  if (!h) return AGIMP_EINVAL;
  h->timing = on ? 1 : 0;
  return 0;
}

int agimp_get_timings(agimp_handle* h, double* ms, int64_t* calls, int32_t n) {
  if (!h || n < 0) return AGIMP_EINVAL;
  int rc = collect_marks(h);
  if (rc) return rc;
  for (int i = 0; i < n && i < AGIMP_NPHASES; ++i) {
    if (ms) ms[i] = h->phase_ms[i];
    if (calls) calls[i] = h->phase_calls[i];
    h->phase_ms[i] = 0.0;
    h->phase_calls[i] = 0;
  }
  return 0;
}

int agimp_get_kernel_timings(agimp_handle* h, char* names, int32_t name_len, double* total_ms, int64_t* calls,
                             double* min_ms, double* max_ms, int32_t cap, int32_t* count) {
  if (!h || cap < 0 || !count || (names && name_len < 1)) return AGIMP_EINVAL;
  int rc = collect_marks(h);
  if (rc) return rc;
  int k = 0;
  for (size_t i = 0; i < h->kstats.size(); ++i) {
    agimp_handle::KStat& st = h->kstats[i];
    if (!st.calls) continue;
    if (k < cap) {
      if (names) {
        std::strncpy(names + (size_t)k * name_len, h->knames[i].c_str(), (size_t)name_len - 1);
        names[(size_t)k * name_len + name_len - 1] = 0;
      }
      if (total_ms) total_ms[k] = st.total_ms;
      if (calls) calls[k] = st.calls;
      if (min_ms) min_ms[k] = st.min_ms;
      if (max_ms) max_ms[k] = st.max_ms;
    }
    ++k;
    st = agimp_handle::KStat{};
  }
  *count = k;
  return 0;
}

int agimp_level_extent(agimp_handle* h, int32_t l) {
  if (!h || l < 0 || l >= h->L) return AGIMP_EINVAL;
  return h->hd.T.lv[l].e;
}

int agimp_sync(agimp_handle* h) {
  if (!h) return AGIMP_EINVAL;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

const char* agimp_last_error(agimp_handle* h) { return h ? h->err.c_str() : "null handle"; }

// ------------------------------------------------------------------------------ learned prediction

int agimp_learned_enable(agimp_handle* h, int32_t window, float tau) {
  if (!h || window < 1 || !(tau > 0.f && tau <= 1.f)) return AGIMP_EINVAL;
  if (h->hd.static_mode) return fail(h, AGIMP_ESTATE, "learned prediction needs dynamic adaptivity");
  h->window = window;
  h->U.tau = tau;
  if (!h->learn_on) h->U.mode = AGIMP_PREDICT_NETWORK | AGIMP_PREDICT_SWEEP;
  if (!h->learn_on) {
    LearnDesc& U = h->U;
    U.ulev = std::max(0, h->L - 3);
    const LevelDesc& Lu = h->hd.T.lv[U.ulev];
    const int p4 = h->D == 2 ? 16 : 64;
    U.nu = Lu.nblocks * p4;
    const size_t nu = (size_t)U.nu, nu4 = (nu + 3) & ~(size_t)3;
    int rc = 0;
    rc = rc ? rc : dalloc(h, &U.acc, nu * UNIT_ACC);
    rc = rc ? rc : dalloc(h, &U.uflag, nu4);
    rc = rc ? rc : dalloc(h, &U.cflag, nu4);
    rc = rc ? rc : dalloc(h, &U.upred, nu4);
    rc = rc ? rc : dalloc(h, &U.occ, nu);
    rc = rc ? rc : dalloc(h, &U.cand, nu);
    rc = rc ? rc : dalloc(h, &U.counts, 4);
    rc = rc ? rc : dalloc(h, &U.feat, nu * UNIT_NFEAT);
    rc = rc ? rc : dalloc(h, &U.lab, nu);
    float* params = nullptr;
    rc = rc ? rc : dalloc(h, &params, (size_t)AGIMP_MLP_NPARAMS);
    rc = rc ? rc : dalloc(h, &h->cls_dev, nu);
    PBuf& sp = h->hd.spare;
    sp.stride = h->hd.P[0].stride;
    rc = rc ? rc : dalloc(h, &sp.f, particle_floats(h));
    rc = rc ? rc : dalloc(h, &sp.g, (size_t)sp.stride);
    rc = rc ? rc : dalloc(h, &sp.id, (size_t)sp.stride);
    rc = rc ? rc : dalloc(h, &h->snap_cnt, NCOUNTERS);
    if (rc) return rc;
    if (hipHostMalloc((void**)&h->win_out, 4 * sizeof(unsigned long long), hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void**)&h->win_out_dev, h->win_out, 0) != hipSuccess)
      return fail(h, AGIMP_ENOMEM, "hipHostMalloc failed for the window check's readback");
    HIPCHK(h, hipMemsetAsync(U.upred, 0xff, nu4, h->stream));
    U.params = as_global(params);
    double vol = 1.0;
    for (int a = 0; a < h->D; ++a) vol *= (double)Lu.dx;
    U.inv_mref = (float)((double)h->cfg.p_vol / (h->mean_mass * vol));
    h->hd.check = 1;
    h->hd.ulev = U.ulev;
    h->hd.upred = U.upred;
    int rc2 = upload_desc(h);
    if (rc2) return rc2;
    h->learn_on = 1;
  }
  return 0;
}

int agimp_learned_set_model(agimp_handle* h, const float* params, int32_t n) {
  if (!h || !params || n != AGIMP_MLP_NPARAMS) return AGIMP_EINVAL;
  if (!h->learn_on) return fail(h, AGIMP_ESTATE, "learned prediction not enabled (agimp_learned_enable)");
  HIPCHK(h, hipMemcpyAsync((void*)h->U.params, params, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->model_set = 1;
  return 0;
}

int agimp_learned_features(agimp_handle* h, float dt, float* feat, int32_t* units, int64_t cap, int64_t* count) {
  if (cap < 0 || !count) return AGIMP_EINVAL;
  int rc = learn_ready(h);
  if (rc) return rc;
  rc = learn_features(h, dt);
  if (rc) return rc;
  int64_t nc = 0;
  rc = learn_count(h, 1, &nc);
  if (rc) return rc;
  *count = nc;
  const int64_t k = std::min(nc, cap);
  if (k > 0 && feat)
    HIPCHK(h, hipMemcpyAsync(feat, h->U.feat, (size_t)k * UNIT_NFEAT * 4, hipMemcpyDeviceToHost, h->stream));
  if (k > 0 && units) HIPCHK(h, hipMemcpyAsync(units, h->U.cand, (size_t)k * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int agimp_learned_labels_reset(agimp_handle* h) {
  int rc = learn_ready(h);
  if (rc) return rc;
  launch_unit_labels_clear(h->stream, h->U);
  HIPCHK(h, hipGetLastError());
  return 0;
}

int agimp_learned_labels_accumulate(agimp_handle* h) {
  int rc = learn_ready(h);
  if (rc) return rc;
  launch_unit_labels_accumulate(h->stream, h->dd, h->D, h->n, h->host_parity, h->U);
  HIPCHK(h, hipGetLastError());
  return 0;
}

int agimp_learned_labels(agimp_handle* h, int32_t* cls, int64_t cap) {
  if (!cls || cap < 0) return AGIMP_EINVAL;
  int rc = learn_ready(h);
  if (rc) return rc;
  int64_t nc = 0;
  rc = learn_count(h, 1, &nc);
  if (rc) return rc;
  launch_unit_label_classes(h->stream, h->U, h->L, h->cls_dev);
  HIPCHK(h, hipGetLastError());
  const int64_t k = std::min(nc, cap);
  if (k > 0) HIPCHK(h, hipMemcpyAsync(cls, h->cls_dev, (size_t)k * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int agimp_learned_predict(agimp_handle* h, float dt, int32_t* cls, int8_t* level, int64_t cap, int64_t* count) {
  if (cap < 0 || !count) return AGIMP_EINVAL;
  int rc = learn_ready(h);
  if (rc) return rc;
  if (!h->model_set) return fail(h, AGIMP_ESTATE, "no model (agimp_learned_set_model)");
  rc = learn_features(h, dt);
  if (rc) return rc;
  launch_unit_predict(h->stream, h->U, h->L, h->cls_dev);
  HIPCHK(h, hipGetLastError());
  int64_t nc = 0;
  rc = learn_count(h, 1, &nc);
  if (rc) return rc;
  *count = nc;
  const int64_t k = std::min(nc, cap);
  if (k > 0 && cls) HIPCHK(h, hipMemcpyAsync(cls, h->cls_dev, (size_t)k * 4, hipMemcpyDeviceToHost, h->stream));
  if (level) HIPCHK(h, hipMemcpyAsync(level, h->U.upred, (size_t)h->U.nu, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int agimp_learned_advance(agimp_handle* h, float dt, int32_t n) {
  if (n < 0) return AGIMP_EINVAL;
  int rc = learn_ready(h);
  if (rc) return rc;
  if (!h->model_set) return fail(h, AGIMP_ESTATE, "no model (agimp_learned_set_model)");
  if (n == 0) return 0;
  rc = set_graph_dt(h, dt);
  if (rc) return rc;
  for (int p = 0; p < 2; ++p) {
    if (h->fixed[p].exec) continue;
    hipError_t e = capture_fixed(h->stream, h->dd, h->dims, p, h->fixed[p]);
    if (e != hipSuccess) {
      h->fixed[p].release();
      return fail(h, AGIMP_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
  }
  if (h->window > 1 && h->fixed_chain_len != h->window) {  // whole windows from one graph each
    for (int p = 0; p < 2; ++p) {
      hipError_t e = capture_fixed_chain(h->stream, h->dd, h->dims, p, h->window, h->fixed_chain[p]);
      if (e != hipSuccess) {
        h->fixed_chain[0].release();
        h->fixed_chain[1].release();
        h->fixed_chain_len = 0;
        return fail(h, AGIMP_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
      }
    }
    h->fixed_chain_len = h->window;
  }
  int t = 0;
  while (t < n) {
    rc = learn_rebuild(h, dt);
    if (rc) return rc;
    h->lstats[AGIMP_LSTAT_REBUILDS] += 1;
    const int w = std::min(h->window, n - t);
    const int par0 = h->host_parity;
    launch_window_start(h->stream, h->dd, h->snap_cnt);
    // the window start's particles stay in buffer par0 through the first substep, then in the
    // spare buffer, which takes par0's place (the window graph swaps them itself)
    if (w > 1 && w == h->fixed_chain_len) {
      HIPCHK(h, hipGraphLaunch(h->fixed_chain[h->host_parity].exec, h->stream));
      h->host_parity ^= w & 1;
    } else {
      for (int i = 0; i < w; ++i) {
        HIPCHK(h, hipGraphLaunch(h->fixed[h->host_parity].exec, h->stream));
        h->host_parity ^= 1;
        if (i == 0 && w > 1) launch_swap_spare(h->stream, h->dd, par0);
      }
    }
    if (w > 1) std::swap(h->hd.P[par0], h->hd.spare);  // the host copy follows the device's
    launch_window_end(h->stream, h->dd, h->snap_cnt, h->win_out_dev);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // failed checks, first failed substep, substeps; failed checks at the window start
    const unsigned long long st[3] = {h->win_out[0], h->win_out[1], h->win_out[2]}, v0 = h->win_out[3];
    if (st[1] <= (unsigned long long)w) {
      // the check failed before substep st[1] of the window: keep the st[1]-1 substeps that
      // ran on a valid octree, then rebuild from there.  Back to the window start: its buffer
      // returns to par0, and the counters are restored.
      h->lstats[AGIMP_LSTAT_VIOLATIONS] += (int64_t)(st[0] - v0);
      const int ok = (int)st[1] - 1;
      if (w > 1) {
        launch_swap_spare(h->stream, h->dd, par0);
        std::swap(h->hd.P[par0], h->hd.spare);
      }
      HIPCHK(h, hipMemcpyAsync(h->hd.cnt, h->snap_cnt, NCOUNTERS * 8, hipMemcpyDeviceToDevice, h->stream));
      h->host_parity = par0;
      for (int i = 0; i < ok; ++i) {
        HIPCHK(h, hipGraphLaunch(h->fixed[h->host_parity].exec, h->stream));
        h->host_parity ^= 1;
      }
      h->lstats[AGIMP_LSTAT_FORCED] += 1;
      h->lstats[AGIMP_LSTAT_DISCARDED] += w;
      h->lstats[AGIMP_LSTAT_REPLAYED] += ok;
      t += ok;
      if (ok == 0) {  // not reachable: the octree was just built from these particles
        rc = do_refine(h);
        if (rc) return rc;
        rc = do_physics(h, dt);
        if (rc) return rc;
        t += 1;
      }
    } else {
      t += w;
    }
  }
  h->lstats[AGIMP_LSTAT_SUBSTEPS] += n;
  return 0;
}

int agimp_learned_set_parts(agimp_handle* h, int32_t parts) {
  if (!h || parts < 0 || parts > (AGIMP_PREDICT_NETWORK | AGIMP_PREDICT_SWEEP)) return AGIMP_EINVAL;
  if (!h->learn_on) return fail(h, AGIMP_ESTATE, "learned prediction not enabled (agimp_learned_enable)");
  h->U.mode = parts;
  return 0;
}

int agimp_learned_stats(agimp_handle* h, int64_t* out, int32_t n) {
  if (!h || !out || n < 0) return AGIMP_EINVAL;
  if (h->learn_on) {
    int c[4] = {0, 0, 0, 0};
    HIPCHK(h, hipMemcpyAsync(c, h->U.counts, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->lstats[AGIMP_LSTAT_CANDIDATES] = c[1];
    h->lstats[AGIMP_LSTAT_REFINED] = c[2];
  }
  for (int i = 0; i < n && i < AGIMP_NLSTATS; ++i) out[i] = h->lstats[i];
  return 0;
}

int agimp_learned_reset_stats(agimp_handle* h) {
  if (!h) return AGIMP_EINVAL;
  for (int i = 0; i < AGIMP_NLSTATS; ++i) h->lstats[i] = 0;
  return 0;
}

// ------------------------------------------------------------------------------ batches

// launch geometry of a batch from its members' current settings (a member that switches
// between static and dynamic adaptivity after the batch was made changes it)
static Dims batch_dims(const std::vector<agimp_handle*>& scenes) {
  const int n = (int)scenes.size();
  Dims m{};
  m.dim = scenes[0]->D;
  m.nscenes = n;
  m.list_grid = std::max(16, 2048 / n);
  m.fill_grid = std::max(16, 4096 / n);  // the grid-stride loops absorb the rest
  m.grid_grid = std::max(16, 4096 / n);
  int p2g = 0;
  for (const agimp_handle* h : scenes) {
    m.maxL = std::max(m.maxL, h->L);
    m.any_single |= h->L == 1;
    m.any_static |= h->hd.static_mode;
    m.any_dynamic |= !h->hd.static_mode;
    m.max_n = std::max(m.max_n, (size_t)h->n);
    m.max_total_blocks = std::max(m.max_total_blocks, (size_t)h->hd.T.total_blocks);
    m.max_fill_words = std::max(m.max_fill_words, h->hd.w_cells + h->hd.w_node + 2 * h->hd.w_flags);
    p2g = std::max(p2g, h->max_chunks);
  }
  m.p2g_grid = std::max(16, std::min(p2g, 8192 / n));
  return m;
}

static bool same_dims(const Dims& a, const Dims& b) {
  return a.dim == b.dim && a.nscenes == b.nscenes && a.maxL == b.maxL && a.any_single == b.any_single &&
         a.any_static == b.any_static && a.any_dynamic == b.any_dynamic && a.max_n == b.max_n &&
         a.max_total_blocks == b.max_total_blocks && a.max_fill_words == b.max_fill_words &&
         a.list_grid == b.list_grid && a.fill_grid == b.fill_grid && a.p2g_grid == b.p2g_grid &&
         a.grid_grid == b.grid_grid;
}

int agimp_batch_create(agimp_handle* const* scenes, int32_t n, agimp_batch** out) {
  if (!scenes || n < 1 || !out) return AGIMP_EINVAL;
  *out = nullptr;
  agimp_batch* b = new agimp_batch();
  for (int i = 0; i < n; ++i) {
    agimp_handle* h = scenes[i];
    if (!h || !scenes[0] || h->D != scenes[0]->D || h->cfg.device != scenes[0]->cfg.device || !h->has_particles) {
      delete b;
      return AGIMP_EINVAL;
    }
    (void)hipStreamSynchronize(h->stream);
    b->scenes.push_back(h);
  }
  b->dims = batch_dims(b->scenes);
  b->device = scenes[0]->cfg.device;
  if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&b->dd, sizeof(SceneDesc) * (size_t)n) != hipSuccess) {
    agimp_batch_destroy(b);
    return AGIMP_EHIP;
  }
  b->host.resize(n);
  for (int i = 0; i < n; ++i) b->host[i] = b->scenes[i]->hd;
  if (hipMemcpy(b->dd, b->host.data(), sizeof(SceneDesc) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
    agimp_batch_destroy(b);
    return AGIMP_EHIP;
  }
  *out = b;
  return 0;
}

// make every scene's batch descriptor list its current buffer at index b->par (the scene may
// have been advanced on its own since the last batch call); uploads entries that changed, in
// order on the batch stream (b->host keeps the sources alive until the stream is synchronised)
static int batch_sync_buffers(agimp_batch* b) {
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    agimp_handle* h = b->scenes[i];
    SceneDesc want = h->hd;
    want.P[b->par] = h->hd.P[h->host_parity];
    want.P[b->par ^ 1] = h->hd.P[h->host_parity ^ 1];
    if (std::memcmp(&want, &b->host[i], sizeof(SceneDesc)) == 0) continue;
    b->host[i] = want;
    if (hipMemcpyAsync(b->dd + i, &b->host[i], sizeof(SceneDesc), hipMemcpyHostToDevice, b->stream) != hipSuccess)
      return AGIMP_EHIP;
  }
  return 0;
}

int agimp_batch_destroy(agimp_batch* b) {
  if (!b) return AGIMP_EINVAL;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->graph[0].release();
  b->graph[1].release();
  b->chain[0].release();
  b->chain[1].release();
  if (b->dd) (void)hipFree(b->dd);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int agimp_batch_set_particles_per_chunk(agimp_batch* b, const int32_t* ch) {
  if (!b || !ch) return AGIMP_EINVAL;
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    if (ch[i] < AGIMP_MIN_CHUNK || ch[i] > AGIMP_MAX_CHUNK) return AGIMP_EINVAL;
    b->scenes[i]->hd.CH = ch[i];
    int rc = upload_desc(b->scenes[i]);
    if (rc) return rc;
  }
  return batch_sync_buffers(b);
}

int agimp_batch_advance(agimp_batch* b, const float* dt, int32_t steps) {
  if (!b || !dt || steps < 0) return AGIMP_EINVAL;
  const int n = (int)b->scenes.size();
  // work a member queued on its own stream (agimp_advance and agimp_substep return before it
  // ends) must finish before the batch touches the same buffers
  for (agimp_handle* h : b->scenes)
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
      b->err = "batch_advance: a member's stream failed";
      return AGIMP_EHIP;
    }
  // a member switched between static and dynamic adaptivity since the graphs were captured:
  // their kernel set no longer fits, so recapture
  const Dims m = batch_dims(b->scenes);
  if (!same_dims(m, b->dims)) {
    b->dims = m;
    b->graph[0].release();
    b->graph[1].release();
    b->chain[0].release();
    b->chain[1].release();
  }
  // per-scene dt in the descriptors (read by the kernels at run time)
  for (int i = 0; i < n; ++i) {
    agimp_handle* h = b->scenes[i];
    if (h->hd.ph.dt != dt[i]) {
      h->hd.ph.dt = dt[i];
      h->graph_dt = -1.f;
      if (hipMemcpyAsync(h->dd, &h->hd, sizeof(SceneDesc), hipMemcpyHostToDevice, b->stream) != hipSuccess)
        return AGIMP_EHIP;
    }
  }
  if (batch_sync_buffers(b)) return AGIMP_EHIP;
  // all scenes advance together: chains of SUBSTEP_CHAIN substeps per host launch, then single
  // substeps
  hipError_t e = replay_substeps(b->stream, b->dd, b->dims, b->graph, b->chain, &b->par, steps);
  if (e != hipSuccess) {
    b->err = std::string("graph replay: ") + hipGetErrorString(e);
    return AGIMP_EHIP;
  }
  if (hipStreamSynchronize(b->stream) != hipSuccess) return AGIMP_EHIP;
  for (agimp_handle* h : b->scenes) {
    h->host_parity ^= (steps & 1);
    h->topo_valid = 1;
  }
  return 0;
}

const char* agimp_batch_last_error(agimp_batch* b) { return b ? b->err.c_str() : "null batch"; }

// ------------------------------------------------------------------------------ host maths

int agimp_host_svd(int32_t dim, const double* F, double* U, double* s, double* V) {
  if (!F || !U || !s || !V) return AGIMP_EINVAL;
  if (dim == 2) {
    double Fl[2][2], Ul[2][2], Vl[2][2];
    for (int i = 0; i < 4; ++i) Fl[i / 2][i % 2] = F[i];
    svd2(Fl, Ul, s, Vl);
    for (int i = 0; i < 4; ++i) { U[i] = Ul[i / 2][i % 2]; V[i] = Vl[i / 2][i % 2]; }
    return 0;
  }
  if (dim == 3) {
    double Fl[3][3], Ul[3][3], Vl[3][3];
    for (int i = 0; i < 9; ++i) Fl[i / 3][i % 3] = F[i];
    svd3(Fl, Ul, s, Vl);
    for (int i = 0; i < 9; ++i) { U[i] = Ul[i / 3][i % 3]; V[i] = Vl[i / 3][i % 3]; }
    return 0;
  }
  return AGIMP_EINVAL;
}

int agimp_host_constitutive(const agimp_config* cfg, double* F, double* Jp, double* tau, int32_t clamp) {
  if (!cfg || !F || !Jp || !tau) return AGIMP_EINVAL;
  Physics ph = make_physics(*cfg);
  if (cfg->dim == 2) {
    double Fl[2][2], T[2][2];
    for (int i = 0; i < 4; ++i) Fl[i / 2][i % 2] = F[i];
    constitutive<2>(Fl, *Jp, ph, T, clamp != 0);
    for (int i = 0; i < 4; ++i) { F[i] = Fl[i / 2][i % 2]; tau[i] = T[i / 2][i % 2]; }
    return 0;
  }
  if (cfg->dim == 3) {
    double Fl[3][3], T[3][3];
    for (int i = 0; i < 9; ++i) Fl[i / 3][i % 3] = F[i];
    constitutive<3>(Fl, *Jp, ph, T, clamp != 0);
    for (int i = 0; i < 9; ++i) { F[i] = Fl[i / 3][i % 3]; tau[i] = T[i / 3][i % 3]; }
    return 0;
  }
  return AGIMP_EINVAL;
}

}  // extern "C"
