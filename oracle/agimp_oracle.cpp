// TEST INFRASTRUCTURE ONLY -- the oracle.  Nothing in the product path may link or call
// this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the checker / the timed CPU restatement.
//
// CPU restatement of one adaptive-GIMP substep of g1n0st/adaptive-gimp, following the
// reference text line by line (it cannot be run here: Taichi is absent).  Every kernel
// below cites the reference function it restates.  Taichi semantics reproduced:
//   * an inactive sparse element reads as 0; any write activates its 4^d leaf block
//     (adaptive_gimp.py:37-42); struct-fors visit every element of active blocks (all
//     loop bodies test masks or mass, so visiting inactive zeros is equivalent);
//   * deactivate_all zeroes everything (adaptive_gimp.py:324-326);
//   * `.cast(int)` truncates toward zero; `//` on ints is floor division;
//   * `+=` / atomic_or inside parallel loops are atomic.
// Parallelisation keeps results deterministic: P2G scatters per leaf block in 2^d colour
// classes (blocks of one colour have disjoint 3^d stencils), grid scatters run serially.
//
// Template parameter R is the particle/grid float type (float mimics the reference's f32
// fields; double is the FP64 error oracle).  The SVD/polar/plasticity maths is always done
// in double (a one-sided Jacobi SVD written here), matching the product's FP64 choice.
//
// Extensions required by the workload configs and absent from the reference (each is off
// by default, which gives the reference behaviour): particle volume p_vol (reference: 1),
// sticky walls (reference: per-axis normal zeroing, adaptive_gimp.py:284-287), a slab
// collider, singular-value clamp plasticity with snow hardening (Stomakhin et al. 2013).
//
// Static-adaptivity grid initializers are given as two finest-cell maps evaluated by the
// caller: ret_level[I] (what grid_initializer(sim, I) returns) and act_level[I] (the level
// passed to activate_cell inside it, -1 for none), over the padded finest cell box.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int32_t UNACTIVATED = 0, ACTIVATED = 1, GHOST = 2, T_JUNCTION = 4, CULLING = 8;  // utils.py:5-9
constexpr int PAD = 4;  // cells of padding on each side of every level (one leaf block)

struct Params {
  double E, nu;
  double gravity[3];
  double p_vol, radius;
  int boundary_gap;
  int wall_mode;  // 0: reference per-axis normal zeroing, 1: sticky (all components)
  int plastic;    // 0: none (reference), 1: singular-value clamp (+ hardening)
  double theta_c, theta_s, hardening;
  int slab_enable, slab_axis;
  double slab_lo, slab_hi;
};

inline int floordiv_pow2(int a, int k) { return a >> k; }  // arithmetic shift == floor division

// ---------------------------------------------------------------- SVD (double precision)
// One-sided Jacobi (Hestenes 1958): rotate column pairs of A = F V until orthogonal.
// Output: rotation-variant SVD F = U diag(s) V^T with U, V in SO(d); the smallest singular
// value carries the sign of det F.
template <int D>
void svd_jacobi(const double F[D][D], double U[D][D], double s[D], double V[D][D]) {
  double A[D][D];
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) { A[i][j] = F[i][j]; V[i][j] = (i == j) ? 1.0 : 0.0; }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < D - 1; ++p)
      for (int q = p + 1; q < D; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < D; ++i) { alpha += A[i][p] * A[i][p]; beta += A[i][q] * A[i][q]; gamma += A[i][p] * A[i][q]; }
        if (gamma == 0.0) continue;
        double rel = std::fabs(gamma) / std::sqrt(alpha * beta);
        off = std::max(off, rel);
        if (rel < 1e-17) continue;
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (int i = 0; i < D; ++i) {
          double ap = A[i][p], aq = A[i][q];
          A[i][p] = c * ap - sn * aq; A[i][q] = sn * ap + c * aq;
          double vp = V[i][p], vq = V[i][q];
          V[i][p] = c * vp - sn * vq; V[i][q] = sn * vp + c * vq;
        }
      }
    if (off < 1e-16) break;
  }
  // singular values = column norms; sort descending
  int idx[D];
  double nrm[D];
  for (int j = 0; j < D; ++j) {
    idx[j] = j; nrm[j] = 0;
    for (int i = 0; i < D; ++i) nrm[j] += A[i][j] * A[i][j];
    nrm[j] = std::sqrt(nrm[j]);
  }
  std::sort(idx, idx + D, [&](int a, int b) { return nrm[a] > nrm[b]; });
  double As[D][D], Vs[D][D];
  for (int j = 0; j < D; ++j) {
    s[j] = nrm[idx[j]];
    for (int i = 0; i < D; ++i) { As[i][j] = A[i][idx[j]]; Vs[i][j] = V[i][idx[j]]; }
  }
  // V must be a rotation: flipping a column of V flips the same column of A = F V
  auto det = [](const double M[D][D]) {
    if (D == 2) return M[0][0] * M[1][1] - M[0][1] * M[1][0];
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  };
  if (det(Vs) < 0)
    for (int i = 0; i < D; ++i) { Vs[i][D - 1] = -Vs[i][D - 1]; As[i][D - 1] = -As[i][D - 1]; }
  // U columns = normalised columns of A; complete degenerate columns orthonormally
  for (int j = 0; j < D; ++j) {
    if (s[j] > 1e-300) {
      for (int i = 0; i < D; ++i) U[i][j] = As[i][j] / s[j];
    } else {
      for (int i = 0; i < D; ++i) U[i][j] = 0;
      if (D == 2) {
        if (j == 0) { U[0][0] = 1; U[1][0] = 0; }
        else { U[0][1] = -U[1][0]; U[1][1] = U[0][0]; }
      } else {
        if (j == 0) { U[0][0] = 1; }
        else if (j == 1) {  // any unit vector orthogonal to U[:,0]
          double a0 = U[0][0], a1 = U[1][0], a2 = U[2][0];
          double cx = 0, cy = -a2, cz = a1;  // e_x x u0
          if (std::fabs(a0) > 0.9) { cx = a2; cy = 0; cz = -a0; }  // e_y x u0
          double n = std::sqrt(cx * cx + cy * cy + cz * cz);
          U[0][1] = cx / n; U[1][1] = cy / n; U[2][1] = cz / n;
        } else {
          U[0][2] = U[1][0] * U[2][1] - U[2][0] * U[1][1];
          U[1][2] = U[2][0] * U[0][1] - U[0][0] * U[2][1];
          U[2][2] = U[0][0] * U[1][1] - U[1][0] * U[0][1];
        }
      }
    }
  }
  if (det(U) < 0) {
    for (int i = 0; i < D; ++i) U[i][D - 1] = -U[i][D - 1];
    s[D - 1] = -s[D - 1];
  }
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) V[i][j] = Vs[i][j];
}

// Fixed-corotated Kirchhoff stress, utils.py:11-16: tau = 2 mu (F - U V^T) F^T + la J (J-1) I
template <int D>
void kirchhoff(const double F[D][D], double mu, double la, double tau[D][D]) {
  double U[D][D], s[D], V[D][D];
  svd_jacobi<D>(F, U, s, V);
  double R[D][D];
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      R[i][j] = 0;
      for (int k = 0; k < D; ++k) R[i][j] += U[i][k] * V[j][k];
    }
  double J = (D == 2) ? F[0][0] * F[1][1] - F[0][1] * F[1][0]
                      : F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                            F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
  for (int i = 0; i < D; ++i)
    for (int j = 0; j < D; ++j) {
      double acc = 0;
      for (int k = 0; k < D; ++k) acc += (F[i][k] - R[i][k]) * F[j][k];
      tau[i][j] = 2.0 * mu * acc + ((i == j) ? la * J * (J - 1.0) : 0.0);
    }
}

// ---------------------------------------------------------------- simulator state
template <int D, class R>
struct Sim {
  using IV = std::array<int, D>;
  int L, c;
  int64_t N;
  Params prm;
  double coarsest_dx;
  int finest_size;
  double mu0, la0;
  std::vector<int> n, e, E, NB;                      // per level: size, ref extent, padded extent, blocks/axis
  std::vector<std::vector<int32_t>> cell;             // cell_mask[l]
  std::vector<std::vector<R>> gm;                     // ad_grid_m[l]
  std::vector<std::vector<R>> gv;                     // ad_grid_v[l] (D per node)
  std::vector<std::vector<uint8_t>> blk;              // leaf-block activation of grid[l]
  std::vector<int32_t> node;                          // node_mask (finest resolution)
  std::vector<int8_t> st_ret, st_act;                 // static level maps (finest cells, padded)
  bool static_mode = false;
  // particles (adaptive_gimp.py:55-62) + plastic J
  std::vector<R> x, v, F, C, m, Jp;
  std::vector<int32_t> g, fl;
  // pid lists (adaptive_gimp.py:44-45) as CSR per level: particles grouped by leaf block
  std::vector<std::vector<int64_t>> pid_off;          // per level: NB^d + 1 offsets
  std::vector<std::vector<int64_t>> pid;              // per level: particle ids
  int64_t oob = 0;

  static int align_size(int x, int a) { return (x + (a - 1)) & ~(a - 1); }  // utils.py:44-45

  Sim(int L_, int c_, int64_t N_, const Params& p) : L(L_), c(c_), N(N_), prm(p) {
    coarsest_dx = 1.0 / c;                 // adaptive_gimp.py:16
    finest_size = c * (1 << (L - 1));      // adaptive_gimp.py:17
    mu0 = p.E / (2 * (1 + p.nu));          // adaptive_gimp.py:53
    la0 = p.E * p.nu / ((1 + p.nu) * (1 - 2 * p.nu));
    for (int l = 0; l < L; ++l) {
      int nl = c << l;
      int el = 4 * align_size(nl / 4 + 1, 4);  // adaptive_gimp.py:36-37
      n.push_back(nl); e.push_back(el); E.push_back(el + 2 * PAD); NB.push_back((el + 2 * PAD) / 4);
      size_t cnt = 1;
      for (int a = 0; a < D; ++a) cnt *= (size_t)(el + 2 * PAD);
      cell.emplace_back(cnt, 0);
      gm.emplace_back(cnt, R(0));
      gv.emplace_back(cnt * D, R(0));
      size_t nb = 1;
      for (int a = 0; a < D; ++a) nb *= (size_t)NB[l];
      blk.emplace_back(nb, 0);
      pid_off.emplace_back(nb + 1, 0);
      pid.emplace_back();
    }
    node.assign(cell[L - 1].size(), 0);
    x.assign(N * D, R(0)); v.assign(N * D, R(0)); F.assign(N * D * D, R(0)); C.assign(N * D * D, R(0));
    m.assign(N, R(0)); Jp.assign(N, R(1)); g.assign(N, 0); fl.assign(N, -1);
  }

  bool inb(int l, const IV& I) const {
    for (int a = 0; a < D; ++a)
      if (I[a] < -PAD || I[a] >= e[l] + PAD) return false;
    return true;
  }
  size_t lin(int l, const IV& I) const {
    size_t k = 0;
    for (int a = 0; a < D; ++a) k = k * E[l] + (size_t)(I[a] + PAD);
    return k;
  }
  size_t blin(int l, const IV& I) const {
    size_t k = 0;
    for (int a = 0; a < D; ++a) k = k * NB[l] + (size_t)((I[a] + PAD) >> 2);
    return k;
  }
  void touch(int l, const IV& I) { blk[l][blin(l, I)] = 1; }
  int32_t cm(int l, const IV& I) {
    if (!inb(l, I)) { ++oob; return 0; }
    return cell[l][lin(l, I)];
  }
  void cm_or(int l, const IV& I, int32_t bits) {
    if (!inb(l, I)) { ++oob; return; }
    size_t k = lin(l, I);
#pragma omp atomic
    cell[l][k] |= bits;
    blk[l][blin(l, I)] = 1;
  }
  IV map(int l, const IV& I) const {  // adaptive_gimp.py:126-128
    IV r;
    for (int a = 0; a < D; ++a) r[a] = I[a] * (1 << (L - 1 - l));
    return r;
  }
  int32_t nm(int l, const IV& I) {
    IV f = map(l, I);
    if (!inb(L - 1, f)) { ++oob; return 0; }
    return node[lin(L - 1, f)];
  }
  void nm_or(int l, const IV& I, int32_t bits) {
    IV f = map(l, I);
    if (!inb(L - 1, f)) { ++oob; return; }
    size_t k = lin(L - 1, f);
#pragma omp atomic
    node[k] |= bits;
  }
  // predicates adaptive_gimp.py:130-140
  bool is_valid(int l, const IV& I) { return nm(l, I) > 0; }
  bool is_activated(int l, const IV& I) { return (nm(l, I) & (ACTIVATED | T_JUNCTION)) == ACTIVATED; }
  bool is_T(int l, const IV& I) { return (nm(l, I) & T_JUNCTION) == T_JUNCTION; }
  bool is_ghost(int l, const IV& I) { return (nm(l, I) & (ACTIVATED | GHOST | T_JUNCTION)) == GHOST; }

  // list of active-element coordinates at level l (struct-for over active leaf blocks)
  std::vector<IV> active_elems(int l) const {
    std::vector<IV> out;
    size_t nb = blk[l].size();
    for (size_t b = 0; b < nb; ++b) {
      if (!blk[l][b]) continue;
      IV B;
      size_t r = b;
      for (int a = D - 1; a >= 0; --a) { B[a] = (int)(r % NB[l]); r /= NB[l]; }
      int tot = 1 << (2 * D);
      for (int t = 0; t < tot; ++t) {
        IV I;
        int rr = t;
        for (int a = D - 1; a >= 0; --a) { I[a] = B[a] * 4 + (rr & 3) - PAD; rr >>= 2; }
        out.push_back(I);
      }
    }
    return out;
  }

  R dx_of(int l) const { return R(coarsest_dx) / R(1 << l); }  // coarsest_dx / (2**l) in the field type
  IV base_of(const R* xp, R dx) const {                          // (x/dx + 0.5).cast(int) - 1
    IV b;
    for (int a = 0; a < D; ++a) b[a] = (int)(xp[a] / dx + R(0.5)) - 1;
    return b;
  }

  // ----- adaptive_gimp.py:324-326 deactivate_all
  void clear() {
    for (int l = 0; l < L; ++l) {
      std::fill(cell[l].begin(), cell[l].end(), 0);
      std::fill(gm[l].begin(), gm[l].end(), R(0));
      std::fill(gv[l].begin(), gv[l].end(), R(0));
      std::fill(blk[l].begin(), blk[l].end(), 0);
      pid[l].clear();
      std::fill(pid_off[l].begin(), pid_off[l].end(), 0);
    }
    std::fill(node.begin(), node.end(), 0);
  }

  // ----- adaptive_gimp.py:68-70 activate_cell (static protocol, through the level maps)
  int static_gi(const IV& I) {
    if (!inb(L - 1, I)) { ++oob; return L - 1; }
    size_t k = lin(L - 1, I);
    int al = st_act[k];
    if (al >= 0 && al < L) {
      IV P;
      for (int a = 0; a < D; ++a) P[a] = floordiv_pow2(I[a], L - 1 - al);
      cm_or(al, P, ACTIVATED);
    }
    return st_ret[k];
  }

  // ----- adaptive_gimp.py:72-85
  void activate_cell_static() {
    // serial: grid initializer calls OR bits, order-independent
    for (int64_t p = 0; p < N; ++p) {
      R fdx = R(coarsest_dx) / R(1 << (L - 1));
      IV fb = base_of(&x[p * D], fdx);
      int tot = 1;
      for (int a = 0; a < D; ++a) tot *= 4;
      for (int t = 0; t < tot; ++t) {
        IV I;
        int r = t;
        for (int a = D - 1; a >= 0; --a) { I[a] = fb[a] + (r % 4) - 1; r /= 4; }
        static_gi(I);
      }
      int level = static_gi(fb);
      if (level < 0 || level >= L) { ++oob; continue; }
      R dx = R(coarsest_dx) / R(1 << level);
      IV b = base_of(&x[p * D], dx);
      for (int t = 0; t < (1 << D); ++t) {
        IV I;
        for (int a = 0; a < D; ++a) I[a] = (b[a] + ((t >> (D - 1 - a)) & 1)) * (1 << (L - 1 - level));
        static_gi(I);
      }
    }
  }

  // ----- adaptive_gimp.py:87-96
  void activate_cell_dynamic() {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < N; ++p) {
      int gp = g[p];
      if (gp < 0 || gp >= L) continue;
      R dx = R(coarsest_dx) / R(1 << gp);
      IV b = base_of(&x[p * D], dx);
      for (int t = 0; t < (1 << D); ++t) {
        IV I;
        for (int a = 0; a < D; ++a) I[a] = b[a] + ((t >> (D - 1 - a)) & 1);
        if (!inb(gp, I)) {
#pragma omp atomic
          ++oob;
          continue;
        }
        size_t k = lin(gp, I);
#pragma omp atomic
        cell[gp][k] |= ACTIVATED;
        blk[gp][blin(gp, I)] = 1;
      }
    }
  }

  // ----- adaptive_gimp.py:98-109
  void mark_hier(int l) {
    for (const IV& I : active_elems(l)) {
      if (!(cm(l, I) & ACTIVATED)) continue;
      int l0 = -1;
      for (int lq = 0; lq < l; ++lq) {
        IV P;
        for (int a = 0; a < D; ++a) P[a] = floordiv_pow2(I[a], l - lq);
        if (cm(lq, P) & ACTIVATED) l0 = std::max(l0, lq);
      }
      if (l0 != -1)
        for (int lq = l0; lq < l; ++lq) {
          IV P;
          for (int a = 0; a < D; ++a) P[a] = floordiv_pow2(I[a], l - lq);
          cm_or(lq, P, CULLING);
        }
    }
  }

  // ----- adaptive_gimp.py:111-118
  void split_hier(int l) {
    for (const IV& I : active_elems(l)) {
      if (!(cm(l, I) & CULLING)) continue;
      for (int t = 0; t < (1 << D); ++t) {
        IV J;
        for (int a = 0; a < D; ++a) J[a] = I[a] * 2 + ((t >> (D - 1 - a)) & 1);
        cm_or(l + 1, J, ACTIVATED);
      }
      cell[l][lin(l, I)] = 0;
    }
  }

  // ----- adaptive_gimp.py:142-153
  void activate_mask(int l) {
    for (const IV& I : active_elems(l)) {
      if (cm(l, I) != ACTIVATED) continue;
      for (int t = 0; t < (1 << D); ++t) {
        IV J;
        for (int a = 0; a < D; ++a) J[a] = I[a] + ((t >> (D - 1 - a)) & 1);
        nm_or(l, J, ACTIVATED);
      }
    }
  }

  // ----- adaptive_gimp.py:155-158
  int hier_check(int l, const IV& I) {
    int r = -1;
    for (int lq = 0; lq <= l; ++lq) {
      IV P;
      for (int a = 0; a < D; ++a) P[a] = floordiv_pow2(I[a], l - lq);
      if (cm(lq, P) == ACTIVATED) r = std::max(r, lq);
    }
    return r;
  }

  // ----- adaptive_gimp.py:160-183 (serial: one valid schedule of the parallel loop)
  void mark_ghost_T(int l) {
    for (const IV& I : active_elems(l)) {
      if (cm(l, I) != ACTIVATED) continue;
      for (int j = 0; j < 2 * D; ++j) {
        int ax = j / 2;
        IV I1 = I;
        I1[ax] += (j % 2) ? -1 : 1;
        bool inr = true;
        for (int a = 0; a < D; ++a) inr = inr && (0 <= I1[a] && I1[a] < n[l]);
        if (!inr) continue;  // l_p is only used when I1 is inside the level
        int lp = hier_check(l, I1);
        if (!(cm(l, I1) != ACTIVATED && lp > -1)) continue;
        for (int t = 0; t < (1 << D); ++t) {
          IV k;
          for (int a = 0; a < D; ++a) k[a] = (t >> (D - 1 - a)) & 1;
          if (k[ax] != j % 2) continue;
          bool any = false;
          for (int a = 0; a < D; ++a) any = any || (((I1[a] + k[a]) % (1 << (l - lp))) != 0);
          if (any) {
            IV J;
            for (int a = 0; a < D; ++a) J[a] = I1[a] + k[a];
            nm_or(l, J, T_JUNCTION);
          }
        }
        for (int l0 = lp + 1; l0 <= l; ++l0) {
          IV P;
          for (int a = 0; a < D; ++a) P[a] = floordiv_pow2(I1[a], l - l0);
          cm_or(l0, P, GHOST);
          for (int t = 0; t < (1 << D); ++t) {
            IV J;
            for (int a = 0; a < D; ++a) J[a] = P[a] + ((t >> (D - 1 - a)) & 1);
            if (nm(l0, J) == UNACTIVATED) nm_or(l0, J, GHOST);
          }
        }
      }
    }
  }

  // ----- adaptive_gimp.py:189-197, 206-216
  void level_mapping() {
    std::vector<int64_t> bidx(N, -1);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < N; ++p) {
      int f = -1;
      for (int l = 0; l < L; ++l) {
        IV b = base_of(&x[p * D], dx_of(l));
        for (int t = 0; t < (1 << D); ++t) {
          IV I;
          for (int a = 0; a < D; ++a) I[a] = b[a] + ((t >> (D - 1 - a)) & 1);
          if (inb(l, I) && cell[l][lin(l, I)] == ACTIVATED) f = std::max(f, l);
        }
      }
      // the particle's 3^d stencil must lie inside the padded level (else: counted, frozen)
      if (f >= 0) {
        IV b = base_of(&x[p * D], dx_of(f));
        IV b2 = b;
        for (int a = 0; a < D; ++a) b2[a] += 2;
        if (!inb(f, b) || !inb(f, b2)) {
#pragma omp atomic
          ++oob;
          f = -1;
        } else {
          bidx[p] = (int64_t)blin(f, b);
        }
      }
      fl[p] = f;
    }
    for (int l = 0; l < L; ++l) {
      std::vector<int64_t>& off = pid_off[l];
      std::fill(off.begin(), off.end(), 0);
      for (int64_t p = 0; p < N; ++p)
        if (fl[p] == l) off[bidx[p] + 1]++;
      for (size_t b = 1; b < off.size(); ++b) off[b] += off[b - 1];
      pid[l].assign(off.back(), 0);
      std::vector<int64_t> fill(off.begin(), off.end() - 1);
      for (int64_t p = 0; p < N; ++p)
        if (fl[p] == l) pid[l][fill[bidx[p]]++] = p;
    }
  }

  // ----- utils.py:18-42 get_linear_weight
  void linear_weight(const R t[D], const int dI[D], const int cl[D], R& weight, R gw_out[D]) const {
    R r = R(prm.radius);
    R w[D], gw[D];
    for (int a = 0; a < D; ++a) {
      R bmin = std::min(std::max(t[a] + R(1 - dI[a]) - r, R(0)), R(1));
      R bmax = std::min(std::max(t[a] + R(1 - dI[a]) + r, R(0)), R(1));
      R mx, mn;
      if (cl[a]) { mx = bmax; mn = bmin; }
      else { mx = R(1) - bmin; mn = R(1) - bmax; }
      w[a] = mx * mx - mn * mn;
      gw[a] = cl[a] ? (mx - mn) : (mn - mx);
    }
    R d2r = r * R(2);
    if (D == 2) {
      weight = w[0] * w[1] / (R(4) * d2r * d2r);
      gw_out[0] = gw[0] * w[1] / (R(2) * d2r * d2r);
      gw_out[1] = w[0] * gw[1] / (R(2) * d2r * d2r);
    } else {
      R den = d2r * d2r * d2r;
      weight = w[0] * w[1] * w[2] / (R(8) * den);
      gw_out[0] = gw[0] * w[1] * w[2] / (R(4) * den);
      gw_out[1] = w[0] * gw[1] * w[2] / (R(4) * den);
      gw_out[2] = w[0] * w[1] * gw[2] / (R(4) * den);
    }
  }

  // ----- adaptive_gimp.py:218-228 get_weight_stencil: W[3^d][d+1], index sum_a i_a 3^a
  void weight_stencil(const R t[D], R W[][D + 1]) const {
    constexpr int S = (D == 2) ? 9 : 27;
    for (int s = 0; s < S; ++s)
      for (int q = 0; q <= D; ++q) W[s][q] = R(0);
    for (int u = 0; u < (1 << D); ++u)
      for (int vv = 0; vv < (1 << D); ++vv) {
        int dI[D], cl[D];
        int idx = 0, conv = 1;
        for (int a = 0; a < D; ++a) {
          dI[a] = (u >> a) & 1; cl[a] = (vv >> a) & 1;
          idx += (dI[a] + cl[a]) * conv; conv *= 3;
        }
        R wt, gw[D];
        linear_weight(t, dI, cl, wt, gw);
        W[idx][0] += wt;
        for (int a = 0; a < D; ++a) W[idx][1 + a] += gw[a];
      }
  }

  double hardening_scale(int64_t p) const {
    if (prm.plastic && prm.hardening != 0.0) return std::exp(prm.hardening * (1.0 - (double)Jp[p]));
    return 1.0;
  }

  // ----- adaptive_gimp.py:230-250
  void p2g_particle(int l, int64_t p, R dt) {
    R dx = dx_of(l);
    IV b = base_of(&x[p * D], dx);
    R t[D];
    for (int a = 0; a < D; ++a) t[a] = x[p * D + a] / dx - R(b[a] + 1);
    double Fd[D][D], tau[D][D];
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) Fd[i][j] = (double)F[p * D * D + i * D + j];
    double h = hardening_scale(p);
    kirchhoff<D>(Fd, mu0 * h, la0 * h, tau);
    R r = R(prm.radius);
    R rd = (D == 2) ? r * r : r * r * r;
    R hd = (D == 2) ? R(0.25) : R(0.125);
    R scale = -dt * hd / (rd * dx) * R(prm.p_vol);
    R S[D][D];
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) S[i][j] = scale * R(tau[i][j]);
    constexpr int NS = (D == 2) ? 9 : 27;
    R W[NS][D + 1];
    weight_stencil(t, W);
    R mp = m[p];
    for (int s = 0; s < NS; ++s) {
      int off[D], rr = s;
      for (int a = 0; a < D; ++a) { off[a] = rr % 3; rr /= 3; }
      R w = W[s][0];
      R dpos[D];
      for (int a = 0; a < D; ++a) dpos[a] = R(off[a]) - (t[a] + R(1));
      IV I;
      for (int a = 0; a < D; ++a) I[a] = b[a] + off[a];
      size_t k = lin(l, I);
      gm[l][k] += w * mp;
      for (int a = 0; a < D; ++a) {
        R cd = R(0);
        for (int q = 0; q < D; ++q) cd += C[p * D * D + a * D + q] * dpos[q];
        R sg = R(0);
        for (int q = 0; q < D; ++q) sg += S[a][q] * W[s][1 + q];
        gv[l][k * D + a] += w * mp * (v[p * D + a] + cd) + sg;
      }
      blk[l][blin(l, I)] = 1;
    }
  }

  void p2g(int l, R dt) {
    // colour c: blocks whose coordinates have parity c; same-colour stencils are disjoint
    size_t nb = blk[l].size();
    for (int color = 0; color < (1 << D); ++color) {
#pragma omp parallel for schedule(dynamic, 4)
      for (int64_t bb = 0; bb < (int64_t)nb; ++bb) {
        int64_t s0 = pid_off[l][bb], s1 = pid_off[l][bb + 1];
        if (s0 == s1) continue;
        size_t r = (size_t)bb;
        int col = 0;
        for (int a = D - 1; a >= 0; --a) { col |= (int)((r % NB[l]) & 1) << a; r /= NB[l]; }
        if (col != color) continue;
        for (int64_t q = s0; q < s1; ++q) p2g_particle(l, pid[l][q], dt);
      }
    }
  }

  // ----- adaptive_gimp.py:252-275 (+ plasticity extension)
  void g2p(int l, R dt) {
    R dx = dx_of(l);
    int64_t cnt = (int64_t)pid[l].size();
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < cnt; ++q) {
      int64_t p = pid[l][q];
      IV b = base_of(&x[p * D], dx);
      R t[D];
      for (int a = 0; a < D; ++a) t[a] = x[p * D + a] / dx - R(b[a] + 1);
      constexpr int NS = (D == 2) ? 9 : 27;
      R W[NS][D + 1];
      weight_stencil(t, W);
      R nv[D] = {}, G[D][D] = {};
      for (int s = 0; s < NS; ++s) {
        int off[D], rr = s;
        for (int a = 0; a < D; ++a) { off[a] = rr % 3; rr /= 3; }
        IV I;
        for (int a = 0; a < D; ++a) I[a] = b[a] + off[a];
        size_t k = lin(l, I);
        for (int a = 0; a < D; ++a) {
          R gva = gv[l][k * D + a];
          nv[a] += gva * W[s][0];
          for (int q2 = 0; q2 < D; ++q2) G[a][q2] += gva * W[s][1 + q2];
        }
      }
      for (int a = 0; a < D; ++a) v[p * D + a] = nv[a];
      for (int a = 0; a < D; ++a)
        for (int q2 = 0; q2 < D; ++q2) C[p * D * D + a * D + q2] = G[a][q2];
      for (int a = 0; a < D; ++a) x[p * D + a] += dt * v[p * D + a];
      R r = R(prm.radius);
      R rd = (D == 2) ? r * r : r * r * r;
      R hd = (D == 2) ? R(0.25) : R(0.125);
      R sc = dt * hd / (rd * dx);
      R Fn[D][D];
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
          R acc = R(0);
          for (int k = 0; k < D; ++k) acc += ((i == k ? R(1) : R(0)) + sc * G[i][k]) * F[p * D * D + k * D + j];
          Fn[i][j] = acc;
        }
      if (prm.plastic) {
        double Fd[D][D], U[D][D], s[D], V[D][D];
        for (int i = 0; i < D; ++i)
          for (int j = 0; j < D; ++j) Fd[i][j] = (double)Fn[i][j];
        svd_jacobi<D>(Fd, U, s, V);
        double jp = (double)Jp[p];
        for (int a = 0; a < D; ++a) {
          double ns = std::min(std::max(s[a], 1.0 - prm.theta_c), 1.0 + prm.theta_s);
          jp *= s[a] / ns;
          s[a] = ns;
        }
        Jp[p] = R(jp);
        for (int i = 0; i < D; ++i)
          for (int j = 0; j < D; ++j) {
            double acc = 0;
            for (int k = 0; k < D; ++k) acc += U[i][k] * s[k] * V[j][k];
            Fn[i][j] = R(acc);
          }
      }
      for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) F[p * D * D + i * D + j] = Fn[i][j];
    }
  }

  // ----- adaptive_gimp.py:277-287 (+ sticky walls / slab collider extensions)
  void grid_op(int l, R dt) {
    std::vector<IV> el = active_elems(l);
    R dx = dx_of(l);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (int64_t)el.size(); ++q) {
      const IV& I = el[q];
      bool inr = true;
      for (int a = 0; a < D; ++a) inr = inr && (0 <= I[a] && I[a] <= n[l]);
      if (!inr) continue;
      size_t k = lin(l, I);
      if (!(gm[l][k] > R(0) && is_activated(l, I))) continue;
      for (int a = 0; a < D; ++a) gv[l][k * D + a] /= gm[l][k];
      for (int a = 0; a < D; ++a) gv[l][k * D + a] += R(prm.gravity[a]) * dt;
      IV f = map(l, I);
      bool wall_any = false;
      for (int a = 0; a < D; ++a) {
        bool w = f[a] < prm.boundary_gap || f[a] > finest_size - prm.boundary_gap;
        wall_any = wall_any || w;
        if (w && prm.wall_mode == 0) gv[l][k * D + a] = R(0);
      }
      if (prm.wall_mode == 1 && wall_any)
        for (int a = 0; a < D; ++a) gv[l][k * D + a] = R(0);
      if (prm.slab_enable) {
        double y = (double)I[prm.slab_axis] * (double)dx;
        if (y >= prm.slab_lo && y <= prm.slab_hi)
          for (int a = 0; a < D; ++a) gv[l][k * D + a] = R(0);
      }
    }
  }

  static bool all_even(const IV& I) {
    for (int a = 0; a < D; ++a)
      if (I[a] % 2 != 0) return false;
    return true;
  }

  // ----- adaptive_gimp.py:289-301
  void grid_restriction(int l) {
    for (const IV& I : active_elems(l)) {
      bool inr = true;
      for (int a = 0; a < D; ++a) inr = inr && (0 <= I[a] && I[a] <= n[l]);
      if (!inr) continue;
      if (!((is_activated(l, I) && all_even(I)) || is_T(l, I) || is_ghost(l, I))) continue;
      size_t k = lin(l, I);
      int tot = 1;
      for (int a = 0; a < D; ++a) tot *= 3;
      for (int t = 0; t < tot; ++t) {
        IV dI, J;
        int r = t, l1 = 0;
        bool ok = true;
        for (int a = D - 1; a >= 0; --a) { dI[a] = (r % 3) - 1; r /= 3; }
        for (int a = 0; a < D; ++a) {
          J[a] = I[a] + dI[a];
          ok = ok && (0 <= J[a] && J[a] <= n[l]) && (J[a] % 2 == 0);
          l1 += std::abs(dI[a]);
        }
        if (!ok) continue;
        IV I0;
        for (int a = 0; a < D; ++a) I0[a] = J[a] / 2;
        if (!is_valid(l - 1, I0)) continue;
        R w = R(std::pow(0.5, (double)l1));
        size_t k0 = lin(l - 1, I0);
        gm[l - 1][k0] += gm[l][k] * w;
        for (int a = 0; a < D; ++a) gv[l - 1][k0 * D + a] += gv[l][k * D + a] * w;
        touch(l - 1, I0);
      }
    }
  }

  // ----- adaptive_gimp.py:303-321
  void grid_prolongation(int l) {
    int l0 = l - 1;
    for (const IV& I : active_elems(l)) {
      bool inr = true;
      for (int a = 0; a < D; ++a) inr = inr && (0 <= I[a] && I[a] <= n[l]);
      if (!inr) continue;
      if (is_T(l, I) || is_ghost(l, I) || all_even(I)) {
        size_t k = lin(l, I);
        for (int a = 0; a < D; ++a) gv[l][k * D + a] = R(0);
      }
    }
    for (const IV& I : active_elems(l0)) {
      bool inr = true;
      for (int a = 0; a < D; ++a) inr = inr && (0 <= I[a] && I[a] <= n[l0]);
      if (!inr || !is_valid(l0, I)) continue;
      size_t k0 = lin(l0, I);
      int tot = 1;
      for (int a = 0; a < D; ++a) tot *= 3;
      for (int t = 0; t < tot; ++t) {
        IV dI, J;
        int r = t, l1 = 0;
        bool ok = true, zero = true;
        for (int a = D - 1; a >= 0; --a) { dI[a] = (r % 3) - 1; r /= 3; }
        for (int a = 0; a < D; ++a) {
          J[a] = 2 * I[a] + dI[a];
          ok = ok && (0 <= J[a] && J[a] <= n[l]);
          l1 += std::abs(dI[a]);
          zero = zero && dI[a] == 0;
        }
        if (!ok) continue;
        if (!(is_T(l, J) || is_ghost(l, J) || (is_activated(l, J) && zero))) continue;
        R w = R(std::pow(0.5, (double)l1));
        size_t k = lin(l, J);
        for (int a = 0; a < D; ++a) gv[l][k * D + a] += gv[l0][k0 * D + a] * w;
        touch(l, J);
      }
    }
  }

  // ----- adaptive_gimp.py:324-337 (the topology half of substep)
  void refine() {
    clear();
    if (static_mode) {
      activate_cell_static();
    } else {
      activate_cell_dynamic();
      for (int l = L - 1; l >= 1; --l) mark_hier(l);  // :120-124
      for (int l = 0; l < L - 1; ++l) split_hier(l);
    }
    for (int l = 0; l < L; ++l) activate_mask(l);     // :151-153
    for (int l = L - 1; l >= 0; --l) mark_ghost_T(l);  // :185-187
    level_mapping();
  }

  // substep on the current octree (no rebuild): zero the grid values, keep every mask,
  // redo level mapping (adaptive_gimp.py:206-216) and the physics
  void substep_fixed_topology(R dt) {
    for (int l = 0; l < L; ++l) {
      std::fill(gm[l].begin(), gm[l].end(), R(0));
      std::fill(gv[l].begin(), gv[l].end(), R(0));
    }
    level_mapping();
    physics(dt);
  }

  // ----- adaptive_gimp.py:339-352 (the physics half of substep)
  void physics(R dt) {
    for (int l = 0; l < L; ++l) p2g(l, dt);
    for (int l = L - 1; l >= 1; --l) grid_restriction(l);
    for (int l = 0; l < L; ++l) grid_op(l, dt);
    for (int l = 1; l < L; ++l) grid_prolongation(l);
    for (int l = 0; l < L; ++l) g2p(l, dt);
  }
};

struct Handle {
  int dim, fp64;
  void* sim;
};

template <int D, class R>
Sim<D, R>* S(Handle* h) { return reinterpret_cast<Sim<D, R>*>(h->sim); }

#define DISPATCH(h, CALL)                                                  \
  do {                                                                     \
    if ((h)->dim == 2 && !(h)->fp64) { auto* s = S<2, float>(h); CALL; }   \
    else if ((h)->dim == 2) { auto* s = S<2, double>(h); CALL; }           \
    else if (!(h)->fp64) { auto* s = S<3, float>(h); CALL; }               \
    else { auto* s = S<3, double>(h); CALL; }                              \
  } while (0)

template <int D, class R>
void copy_dense(Sim<D, R>* s, int l, const std::vector<int32_t>* src_i, int32_t* dst_i, const std::vector<R>* src_r,
                double* dst_r, int comps) {
  int el = s->e[l];
  int64_t tot = 1;
  for (int a = 0; a < D; ++a) tot *= el;
  for (int64_t q = 0; q < tot; ++q) {
    std::array<int, D> I;
    int64_t r = q;
    for (int a = D - 1; a >= 0; --a) { I[a] = (int)(r % el); r /= el; }
    size_t k = s->lin(l, I);
    if (src_i) dst_i[q] = (*src_i)[k];
    if (src_r)
      for (int cc = 0; cc < comps; ++cc) dst_r[q * comps + cc] = (double)(*src_r)[k * comps + cc];
  }
}

}  // namespace

extern "C" {

struct orc_params {
  double E, nu;
  double gravity[3];
  double p_vol, radius;
  int boundary_gap, wall_mode, plastic;
  double theta_c, theta_s, hardening;
  int slab_enable, slab_axis;
  double slab_lo, slab_hi;
};

void* orc_create(int dim, int level, int coarsest, int64_t n, int fp64, const orc_params* p) {
  Params q;
  static_assert(sizeof(Params) == sizeof(orc_params), "layout");
  std::memcpy(&q, p, sizeof(q));
  Handle* h = new Handle{dim, fp64, nullptr};
  if (dim == 2 && !fp64) h->sim = new Sim<2, float>(level, coarsest, n, q);
  else if (dim == 2) h->sim = new Sim<2, double>(level, coarsest, n, q);
  else if (!fp64) h->sim = new Sim<3, float>(level, coarsest, n, q);
  else h->sim = new Sim<3, double>(level, coarsest, n, q);
  return h;
}

void orc_destroy(void* hv) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, delete s);
  delete h;
}

void orc_set_threads(int nt) {
#ifdef _OPENMP
  omp_set_num_threads(nt);
#else
  (void)nt;
#endif
}

void orc_set_particles(void* hv, const double* x, const double* v, const double* F, const double* C, const double* m,
                       const double* Jp, const int32_t* g) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, {
    using R = typename std::remove_reference<decltype(s->x)>::type::value_type;
    int D = h->dim;
    for (int64_t i = 0; i < s->N * D; ++i) { s->x[i] = R(x[i]); s->v[i] = R(v[i]); }
    for (int64_t i = 0; i < s->N * D * D; ++i) { s->F[i] = R(F[i]); s->C[i] = R(C[i]); }
    for (int64_t i = 0; i < s->N; ++i) { s->m[i] = R(m[i]); s->Jp[i] = R(Jp[i]); s->g[i] = g[i]; }
  });
}

void orc_get_particles(void* hv, double* x, double* v, double* F, double* C, double* Jp, int32_t* fl) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, {
    int D = h->dim;
    for (int64_t i = 0; i < s->N * D; ++i) { x[i] = (double)s->x[i]; v[i] = (double)s->v[i]; }
    for (int64_t i = 0; i < s->N * D * D; ++i) { F[i] = (double)s->F[i]; C[i] = (double)s->C[i]; }
    for (int64_t i = 0; i < s->N; ++i) { Jp[i] = (double)s->Jp[i]; fl[i] = s->fl[i]; }
  });
}

void orc_set_static_map(void* hv, const int8_t* ret_level, const int8_t* act_level) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, {
    size_t cnt = s->cell[s->L - 1].size();
    s->st_ret.assign(ret_level, ret_level + cnt);
    s->st_act.assign(act_level, act_level + cnt);
    s->static_mode = true;
  });
}

void orc_refine(void* hv) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, s->refine());
}

void orc_physics(void* hv, double dt) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, {
    using R = typename std::remove_reference<decltype(s->x)>::type::value_type;
    s->physics(R((float)dt));  // dt0 is declared ti.f32 (adaptive_gimp.py:231,253,278)
  });
}

void orc_substep_fixed_topology(void* hv, double dt) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, {
    using R = typename std::remove_reference<decltype(s->x)>::type::value_type;
    s->substep_fixed_topology(R((float)dt));
  });
}

void orc_substep(void* hv, double dt) {
  orc_refine(hv);
  orc_physics(hv, dt);
}

// dense copies over the reference extent [0, e_l)^d
void orc_get_level(void* hv, int l, int32_t* cell, double* m, double* v) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, {
    copy_dense(s, l, &s->cell[l], cell, (decltype(&s->gm[l]))nullptr, nullptr, 0);
    copy_dense(s, l, (std::vector<int32_t>*)nullptr, nullptr, &s->gm[l], m, 1);
    copy_dense(s, l, (std::vector<int32_t>*)nullptr, nullptr, &s->gv[l], v, h->dim);
  });
}

void orc_get_node_mask(void* hv, int32_t* out) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, copy_dense(s, s->L - 1, &s->node, out, (decltype(&s->gm[0]))nullptr, nullptr, 0));
}

int64_t orc_get_oob(void* hv) {
  Handle* h = (Handle*)hv;
  int64_t r = 0;
  DISPATCH(h, r = s->oob);
  return r;
}

// padded particle-list view: counts per level
void orc_level_counts(void* hv, int64_t* out) {
  Handle* h = (Handle*)hv;
  DISPATCH(h, {
    for (int l = 0; l < s->L; ++l) out[l] = (int64_t)s->pid[l].size();
  });
}

// standalone helpers pinned by tests against closed forms / numpy
void orc_weight_stencil(int dim, double radius, const double* t, double* W) {
  Params p{};
  p.radius = radius;
  p.E = 1; p.nu = 0.2;
  if (dim == 2) {
    Sim<2, double> s(1, 4, 0, p);
    double tt[2] = {t[0], t[1]};
    double Wl[9][3];
    s.weight_stencil(tt, Wl);
    for (int i = 0; i < 9; ++i)
      for (int q = 0; q < 3; ++q) W[i * 3 + q] = Wl[i][q];
  } else {
    Sim<3, double> s(1, 4, 0, p);
    double tt[3] = {t[0], t[1], t[2]};
    double Wl[27][4];
    s.weight_stencil(tt, Wl);
    for (int i = 0; i < 27; ++i)
      for (int q = 0; q < 4; ++q) W[i * 4 + q] = Wl[i][q];
  }
}

void orc_svd(int dim, const double* F, double* U, double* s, double* V) {
  if (dim == 2) {
    double Fl[2][2], Ul[2][2], Vl[2][2];
    for (int i = 0; i < 4; ++i) Fl[i / 2][i % 2] = F[i];
    svd_jacobi<2>(Fl, Ul, s, Vl);
    for (int i = 0; i < 4; ++i) { U[i] = Ul[i / 2][i % 2]; V[i] = Vl[i / 2][i % 2]; }
  } else {
    double Fl[3][3], Ul[3][3], Vl[3][3];
    for (int i = 0; i < 9; ++i) Fl[i / 3][i % 3] = F[i];
    svd_jacobi<3>(Fl, Ul, s, Vl);
    for (int i = 0; i < 9; ++i) { U[i] = Ul[i / 3][i % 3]; V[i] = Vl[i / 3][i % 3]; }
  }
}

void orc_stress(int dim, const double* F, double mu, double la, double* tau) {
  if (dim == 2) {
    double Fl[2][2], T[2][2];
    for (int i = 0; i < 4; ++i) Fl[i / 2][i % 2] = F[i];
    kirchhoff<2>(Fl, mu, la, T);
    for (int i = 0; i < 4; ++i) tau[i] = T[i / 2][i % 2];
  } else {
    double Fl[3][3], T[3][3];
    for (int i = 0; i < 9; ++i) Fl[i / 3][i % 3] = F[i];
    kirchhoff<3>(Fl, mu, la, T);
    for (int i = 0; i < 9; ++i) tau[i] = T[i / 3][i % 3];
  }
}

}  // extern "C"
