// kmi_batch.hip — batched square lid-driven D2Q9 cavities of uneven size in one launch.
//
// Populations of every cavity live in one packed allocation (cavity c: 9 planes of
// n_c² cells at element offset foff_c, the reference layout per cavity).  A step is ONE
// launch of k_batch_step over a list of tile slots: each slot names (cavity, x0, y0) of a
// tw × th tile; a persistent grid of `groups` work-groups walks the slots with stride
// `groups`.  Slots are permuted so that slot s runs on XCD s mod 8 and each XCD owns one
// contiguous run of the cavity-major tile order (neighbouring tiles share an L2).
//
// The per-cell arithmetic is the single-lattice kernel's (kmi_cell.h), so a cavity of the
// batch is bit-identical to the same cavity run alone (tested).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "kmi_cell.h"
#include "kmi_internal.h"

namespace kmi {

struct CavRec {
  long long foff;  // element offset of the cavity's populations
  long long coff;  // cell offset (auxiliary per-cell arrays)
  int n;
  int pad;
  double om;  // BGK ω / MRT s_ν / TRT s⁺
  double sm;  // TRT s⁻
};

struct BatchArgs {
  const void* fi;
  void* fo;
  const CavRec* cav;
  const int* slots;  // 3 ints per slot: cavity (-1: empty), x0, y0
  int nslots;
  int tw;
  int lid;  // KMI_FACE_ZOUHE_VELOCITY or KMI_FACE_MOVING_WALL
  int pad;
  double u_lid;
  double s_e, s_eps, s_q;
  double* uprev;           // 2 × total cells (ux, uy)
  long long total_cells;
  double* partials;        // 4 per slot: (sum, error) of d2 and of u2
};

template <int ST, int COLL, bool NTS>
__global__ void __launch_bounds__(256) k_batch_step(const BatchArgs a) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  constexpr bool SH = S::shifted;
  const int lx = threadIdx.x % a.tw, ly = threadIdx.x / a.tw;
  for (int t = blockIdx.x; t < a.nslots; t += gridDim.x) {
    const int c = a.slots[3 * t];
    if (c < 0) continue;
    const CavRec cv = a.cav[c];
    const int n = cv.n;
    const int i = a.slots[3 * t + 1] + lx, j = a.slots[3 * t + 2] + ly;
    if (i >= n || j >= n) continue;
    const long long P = (long long)n * n;
    const E* __restrict__ src = static_cast<const E*>(a.fi) + cv.foff;
    E* __restrict__ dst = static_cast<E*>(a.fo) + cv.foff;
    const int self = i + n * j;
    R fp[9], fo[9];
    // pull with halfway bounce-back on all four faces (stream_2d_kernel!,
    // collide_stream_2d.jl:75-101); Ladd term on the north lid when it is a moving wall
    sfor<9>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = Lat<9>::cx[q], cy = Lat<9>::cy[q];
      constexpr int o = Lat<9>::opp[q];
      const bool ok = (cx > 0 ? i > 0 : (cx < 0 ? i < n - 1 : true)) &&
                      (cy > 0 ? j > 0 : (cy < 0 ? j < n - 1 : true));
      const long long addr = ok ? (long long)((i - cx) + n * (j - cy)) + q * P : (long long)self + o * P;
      R v = S::ld(src, addr);
      if constexpr (cy < 0) {
        if (!ok && a.lid == KMI_FACE_MOVING_WALL && j == n - 1)
          v = v + R(six_w(Lat<9>::w[q])) * R(1.0) * cdotu<cx, cy, 0>(R(a.u_lid), R(0.0), R(0.0));
      }
      fp[q] = v;
    });
    if (a.lid == KMI_FACE_ZOUHE_VELOCITY && j == n - 1) zh_north<SH>(fp, R(a.u_lid));
    if constexpr (COLL == KMI_BGK) collide_bgk<9, SH>(fp, fo, R(cv.om));
    else if constexpr (COLL == KMI_TRT) collide_trt<9, SH>(fp, fo, R(cv.om), R(cv.sm));
    else collide_mrt<SH>(fp, fo, R(a.s_e), R(a.s_eps), R(a.s_q), R(cv.om));
    sfor<9>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      S::template st<NTS>(dst, self + q * P, fo[q]);
    });
  }
}

// Pull of cell (i, j) of an n × n cavity: halfway bounce-back on the four walls (a link whose
// source lies outside takes the node's own opposite population), plus the Ladd term on the
// north lid when it is a moving wall — the rule of k_batch_step.  The source is element
// ld(off) of a store laid out as 9 planes of `plane` elements with rows of `stride`
// (HBM: the cavity, n and n²; LDS: the extended tile); `self` is the cell's own index.
template <class R, class LD>
KMI_DEV void cavity_pull(const BatchArgs& a, int n, int i, int j, unsigned self, unsigned stride,
                         unsigned plane, R (&fp)[9], LD&& ld) {
  sfor<9>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = Lat<9>::cx[q], cy = Lat<9>::cy[q];
    constexpr int o = Lat<9>::opp[q];
    const bool ok = (cx > 0 ? i > 0 : (cx < 0 ? i < n - 1 : true)) &&
                    (cy > 0 ? j > 0 : (cy < 0 ? j < n - 1 : true));
    const unsigned back = unsigned(cx) + unsigned(cy) * stride;  // source = self − back
    R v = ld(ok ? self - back + unsigned(q) * plane : self + unsigned(o) * plane);
    if constexpr (cy < 0) {
      if (!ok && a.lid == KMI_FACE_MOVING_WALL && j == n - 1)
        v = v + R(six_w(Lat<9>::w[q])) * R(1.0) * cdotu<cx, cy, 0>(R(a.u_lid), R(0.0), R(0.0));
    }
    fp[q] = v;
  });
}

// Lid closure and collision of a pulled cell (as in k_batch_step).
template <int ST, int COLL, class R>
KMI_DEV void cavity_collide(const BatchArgs& a, const CavRec& cv, int n, int j, R (&fp)[9], R (&fo)[9]) {
  constexpr bool SH = Store<ST>::shifted;
  if (a.lid == KMI_FACE_ZOUHE_VELOCITY && j == n - 1) zh_north<SH>(fp, R(a.u_lid));
  if constexpr (COLL == KMI_BGK) collide_bgk<9, SH>(fp, fo, R(cv.om));
  else if constexpr (COLL == KMI_TRT) collide_trt<9, SH>(fp, fo, R(cv.om), R(cv.sm));
  else collide_mrt<SH>(fp, fo, R(a.s_e), R(a.s_eps), R(a.s_q), R(cv.om));
}

// Two steps per launch over the same tile slots.  Phase 1 evaluates step t+1 on the tile and
// its one-cell ring (cells inside the cavity only), pulled from HBM, into LDS (9 values per
// cell, in the storage rounding); phase 2 evaluates step t+2 on the tile from LDS and writes
// HBM.  Bit-identical to two k_batch_step launches; HBM sees one read and one write of each
// population per two steps.  Dynamic LDS: 9 · (tw+2)(th+2) reals.
extern __shared__ double kmi_batch_lds[];

template <int ST, int COLL, bool NTS>
__global__ void __launch_bounds__(256) k_batch_step2(const BatchArgs a) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  R* const L = reinterpret_cast<R*>(kmi_batch_lds);
  const int tw = a.tw, th = int(blockDim.x) / a.tw;
  const int ew = tw + 2, ne = ew * (th + 2);
  const int lx = threadIdx.x % tw, ly = threadIdx.x / tw;
  for (int t = blockIdx.x; t < a.nslots; t += gridDim.x) {
    const int c = a.slots[3 * t];  // uniform over the work-group
    if (c < 0) continue;
    const CavRec cv = a.cav[c];
    const int n = cv.n;
    const int x0 = a.slots[3 * t + 1], y0 = a.slots[3 * t + 2];
    const long long P = (long long)n * n;
    const E* __restrict__ src = static_cast<const E*>(a.fi) + cv.foff;
    E* __restrict__ dst = static_cast<E*>(a.fo) + cv.foff;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      const int i = x0 - 1 + e % ew, j = y0 - 1 + e / ew;
      if (i < 0 || j < 0 || i >= n || j >= n) continue;
      R fp[9], fo[9];
      cavity_pull(a, n, i, j, unsigned(i + n * j), unsigned(n), unsigned(P), fp,
                  [&](unsigned off) { return S::ld(src, off); });
      cavity_collide<ST, COLL>(a, cv, n, j, fp, fo);
      sfor<9>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        L[q * ne + e] = storage_round<ST>(fo[q]);
      });
    }
    __syncthreads();
    const int i = x0 + lx, j = y0 + ly;
    if (i < n && j < n) {
      const int es = (ly + 1) * ew + lx + 1;
      R fp[9], fo[9];
      cavity_pull(a, n, i, j, unsigned(es), unsigned(ew), unsigned(ne), fp,
                  [&](unsigned off) { return L[off]; });
      cavity_collide<ST, COLL>(a, cv, n, j, fp, fo);
      const int self = i + n * j;
      sfor<9>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        S::template st<NTS>(dst, self + q * P, fo[q]);
      });
    }
    __syncthreads();  // the next slot's phase 1 overwrites LDS
  }
}

// K steps per launch (K = 3 or 4) over the same tile slots: the tile and a K-cell ring are
// loaded into LDS with one work-item per cell of that region ((tw+2K)(th+2K) ≤ 1024), then
// advanced K steps in place (step s updates the region less s cells on each side: read,
// barrier, write back in the storage rounding, barrier); the tile's t+K populations go to
// HBM.  Bit-identical to K single steps.  Dynamic LDS: 9 · (tw+2K)(th+2K) reals.
template <int ST, int COLL, int K>
__global__ void __launch_bounds__(1024) k_batch_stepK(const BatchArgs a) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  R* const L = reinterpret_cast<R*>(kmi_batch_lds);
  const int tw = a.tw, th = 256 / a.tw;
  const int ew = tw + 2 * K, eh = th + 2 * K, ne = ew * eh;
  const int e = threadIdx.x;  // one work-item per region cell (blockDim.x == ne)
  const int ex = e % ew, ey = e / ew;
  for (int t = blockIdx.x; t < a.nslots; t += gridDim.x) {
    const int c = a.slots[3 * t];  // uniform over the work-group
    if (c < 0) continue;
    const CavRec cv = a.cav[c];
    const int n = cv.n;
    const int i = a.slots[3 * t + 1] - K + ex, j = a.slots[3 * t + 2] - K + ey;
    const bool in = i >= 0 && j >= 0 && i < n && j < n;
    const unsigned P = unsigned(n) * unsigned(n);
    const E* __restrict__ src = static_cast<const E*>(a.fi) + cv.foff;
    E* __restrict__ dst = static_cast<E*>(a.fo) + cv.foff;
    if (in) {
      const unsigned self = unsigned(i + n * j);
      sfor<9>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        L[q * ne + e] = S::ld(src, self + unsigned(q) * P);
      });
    }
    __syncthreads();
    R fr[9];
    for (int s = 1; s <= K; ++s) {
      const bool valid = in && ex >= s && ex < ew - s && ey >= s && ey < eh - s;
      if (valid) {
        R fp[9];
        cavity_pull(a, n, i, j, unsigned(e), unsigned(ew), unsigned(ne), fp,
                    [&](unsigned off) { return L[off]; });
        cavity_collide<ST, COLL>(a, cv, n, j, fp, fr);
      }
      if (s == K) break;
      __syncthreads();  // every step-s read of L is done
      if (valid) {
        sfor<9>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          L[q * ne + e] = storage_round<ST>(fr[q]);
        });
      }
      __syncthreads();
    }
    if (in && ex >= K && ex < K + tw && ey >= K && ey < K + th) {
      const unsigned self = unsigned(i + n * j);
      sfor<9>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        S::template st<false>(dst, self + unsigned(q) * P, fr[q]);
      });
    }
    __syncthreads();  // the next slot's load overwrites LDS
  }
}

// Work-items of a K-step region for tile width tw (0: the region exceeds 1024).
KMI_HD int stepK_region(int tw, int K) {
  const int r = (tw + 2 * K) * (256 / tw + 2 * K);
  return r <= 1024 ? r : 0;
}

// Velocity of a cell from the stored populations, FP64 (compute_macroscopic_2d order).
template <int ST>
KMI_DEV void cell_u(const void* f, const CavRec& cv, int self, double& ux, double& uy) {
  using S = Store<ST>;
  using E = typename S::elem;
  const E* p = static_cast<const E*>(f) + cv.foff;
  const long long P = (long long)cv.n * cv.n;
  double fp[9];
  sfor<9>([&](auto qc) { fp[decltype(qc)::value] = double(S::ld(p, self + decltype(qc)::value * P)); });
  const double r = S::shifted ? 1.0 + zeroth_moment<9>(fp) : zeroth_moment<9>(fp);
  ux = first_moment<9, 0>(fp) / r;
  uy = first_moment<9, 1>(fp) / r;
}

// mode 0: u → uprev;  mode 1: per-slot partials of Σ|u − uprev|² and Σ|u|².
template <int ST>
__global__ void __launch_bounds__(256) k_batch_resid(const BatchArgs a, const void* f, int mode) {
  __shared__ double lds[4][4];
  const int lx = threadIdx.x % a.tw, ly = threadIdx.x / a.tw;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x; t < a.nslots; t += gridDim.x) {
    const int c = a.slots[3 * t];  // uniform over the work-group
    if (c < 0) continue;
    const CavRec cv = a.cav[c];
    const int i = a.slots[3 * t + 1] + lx, j = a.slots[3 * t + 2] + ly;
    double d2 = 0.0, u2 = 0.0;
    if (i < cv.n && j < cv.n) {
      const int self = i + cv.n * j;
      double ux, uy;
      cell_u<ST>(f, cv, self, ux, uy);
      const long long k = cv.coff + self;
      if (mode == 0) {
        a.uprev[k] = ux;
        a.uprev[a.total_cells + k] = uy;
      } else {
        const double dx = ux - a.uprev[k], dy = uy - a.uprev[a.total_cells + k];
        d2 = dx * dx + dy * dy;
        u2 = ux * ux + uy * uy;
      }
    }
    if (mode == 0) continue;
    // per-slot compensated pairs (sum, error) of d2 and u2: butterfly, then across waves
    double d2e = 0.0, u2e = 0.0;
    for (int off = 32; off > 0; off >>= 1) {
      const double y = __shfl_xor(d2, off, 64), ye = __shfl_xor(d2e, off, 64);
      const double z = __shfl_xor(u2, off, 64), ze = __shfl_xor(u2e, off, 64);
      nsum_merge(d2, d2e, y, ye);
      nsum_merge(u2, u2e, z, ze);
    }
    __syncthreads();  // previous slot's LDS reads are done
    if (lane == 0) {
      lds[wave][0] = d2;
      lds[wave][1] = d2e;
      lds[wave][2] = u2;
      lds[wave][3] = u2e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double s0 = lds[0][0], e0 = lds[0][1], s1 = lds[0][2], e1 = lds[0][3];
      for (int w = 1; w < (int)(blockDim.x + 63) / 64; ++w) {
        nsum_merge(s0, e0, lds[w][0], lds[w][1]);
        nsum_merge(s1, e1, lds[w][2], lds[w][3]);
      }
      a.partials[4 * t] = s0;
      a.partials[4 * t + 1] = e0;
      a.partials[4 * t + 2] = s1;
      a.partials[4 * t + 3] = e1;
    }
  }
}

// Consistent initialisation of a prescribed velocity field (Mei, Luo, Lallemand and
// d'Humières, "Consistent initial conditions for lattice Boltzmann simulations", Comput.
// Fluids 35 (2006) 855-862): between steps, every cell of a held cavity gets the equilibrium
// part of its populations moved back to the held velocity u_h (in uprev), keeping the density
// and the non-equilibrium part that the step produced:
//   f_q ← f_q + feq_q(ρ, u_h) − feq_q(ρ, u),   (ρ, u) the moments of f.
// Iterated, ρ and f − feq relax to the values consistent with u_h, walls and lid included;
// an feq-only start lacks them and seeds slowly decaying modes.  In place, FP64 arithmetic.
template <int ST>
__global__ void __launch_bounds__(256) k_batch_hold(const BatchArgs a, void* f, const int* __restrict__ hold) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  const int lx = threadIdx.x % a.tw, ly = threadIdx.x / a.tw;
  for (int t = blockIdx.x; t < a.nslots; t += gridDim.x) {
    const int c = a.slots[3 * t];
    if (c < 0 || !hold[c]) continue;
    const CavRec cv = a.cav[c];
    const int i = a.slots[3 * t + 1] + lx, j = a.slots[3 * t + 2] + ly;
    if (i >= cv.n || j >= cv.n) continue;
    const long long P = (long long)cv.n * cv.n;
    E* p = static_cast<E*>(f) + cv.foff;
    const int self = i + cv.n * j;
    double fp[9];
    sfor<9>([&](auto qc) { fp[decltype(qc)::value] = double(S::ld(p, self + decltype(qc)::value * P)); });
    const double r = S::shifted ? 1.0 + zeroth_moment<9>(fp) : zeroth_moment<9>(fp);
    const double ux = first_moment<9, 0>(fp) / r, uy = first_moment<9, 1>(fp) / r;
    const long long k = cv.coff + self;
    const double hx = a.uprev[k], hy = a.uprev[a.total_cells + k];
    const double du2 = (hx * hx + hy * hy) - (ux * ux + uy * uy);
    sfor<9>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr double cx = Lat<9>::cx[q], cy = Lat<9>::cy[q];
      const double cu = cx * ux + cy * uy, ch = cx * hx + cy * hy;
      const double d = Lat<9>::w[q] * r * (3.0 * (ch - cu) + 4.5 * (ch * ch - cu * cu) - 1.5 * du2);
      S::template st<false>(p, self + q * P, R(fp[q] + d));
    });
  }
}

// One wave per cavity: lane l folds the cavity's slot pairs l, l + 64, … in order, then a
// fixed butterfly merges the lane pairs; compensated (Neumaier) throughout → r.
__global__ void __launch_bounds__(64) k_batch_final(const double* __restrict__ partials,
                                                    const int* __restrict__ cav_slots,
                                                    const int* __restrict__ cav_range,
                                                    const int* __restrict__ active, int ncav, double window,
                                                    double* __restrict__ out) {
  const int c = blockIdx.x, lane = threadIdx.x;
  if (c >= ncav || !active[c]) return;  // uniform over the wave
  double s0 = 0.0, e0 = 0.0, s1 = 0.0, e1 = 0.0;
  for (int k = cav_range[c] + lane; k < cav_range[c + 1]; k += 64) {
    const int t = cav_slots[k];
    nsum_merge(s0, e0, partials[4 * t], partials[4 * t + 1]);
    nsum_merge(s1, e1, partials[4 * t + 2], partials[4 * t + 3]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double y = __shfl_xor(s0, off, 64), ye = __shfl_xor(e0, off, 64);
    const double z = __shfl_xor(s1, off, 64), ze = __shfl_xor(e1, off, 64);
    nsum_merge(s0, e0, y, ye);
    nsum_merge(s1, e1, z, ze);
  }
  if (lane == 0) {
    const double d2 = s0 + e0, u2 = s1 + e1;
    out[3 * c] = d2;
    out[3 * c + 1] = u2;
    out[3 * c + 2] = sqrt(d2 / u2) / window;  // window = 1: gnf_periodic_steady.jl:192
  }
}

#define KMI_BATCH_ST(st, CALL)                                               \
  switch (st) {                                                              \
    case KMI_STORE_F64: { constexpr int ST = KMI_STORE_F64; CALL; } break;   \
    case KMI_STORE_F32: { constexpr int ST = KMI_STORE_F32; CALL; } break;   \
    case KMI_STORE_F32S: { constexpr int ST = KMI_STORE_F32S; CALL; } break; \
    default: { constexpr int ST = KMI_STORE_F16S; CALL; } break;             \
  }

static hipError_t launch_batch_step(int st, int coll, bool nts, const BatchArgs& a, int groups,
                                    hipStream_t s) {
  const dim3 g(static_cast<unsigned>(groups)), b(256);
#define KMI_BATCH_COLL(ST)                                                                        \
  if (coll == KMI_BGK) {                                                                         \
    if (nts) hipLaunchKernelGGL((k_batch_step<ST, KMI_BGK, true>), g, b, 0, s, a);               \
    else hipLaunchKernelGGL((k_batch_step<ST, KMI_BGK, false>), g, b, 0, s, a);                  \
  } else if (coll == KMI_TRT) {                                                                  \
    if (nts) hipLaunchKernelGGL((k_batch_step<ST, KMI_TRT, true>), g, b, 0, s, a);               \
    else hipLaunchKernelGGL((k_batch_step<ST, KMI_TRT, false>), g, b, 0, s, a);                  \
  } else {                                                                                       \
    if (nts) hipLaunchKernelGGL((k_batch_step<ST, KMI_MRT, true>), g, b, 0, s, a);               \
    else hipLaunchKernelGGL((k_batch_step<ST, KMI_MRT, false>), g, b, 0, s, a);                  \
  }
  KMI_BATCH_ST(st, KMI_BATCH_COLL(ST));
#undef KMI_BATCH_COLL
  return hipGetLastError();
}

static hipError_t launch_batch_step2(int st, int coll, bool nts, const BatchArgs& a, int groups,
                                     hipStream_t s) {
  const dim3 g(static_cast<unsigned>(groups)), b(256);
  const int ne = (a.tw + 2) * (256 / a.tw + 2);
#define KMI_BATCH_COLL2(ST)                                                                          \
  {                                                                                                 \
    const size_t sh = size_t(9) * ne * sizeof(typename Store<ST>::real);                            \
    if (coll == KMI_BGK) {                                                                          \
      if (nts) hipLaunchKernelGGL((k_batch_step2<ST, KMI_BGK, true>), g, b, sh, s, a);              \
      else hipLaunchKernelGGL((k_batch_step2<ST, KMI_BGK, false>), g, b, sh, s, a);                 \
    } else if (coll == KMI_TRT) {                                                                   \
      if (nts) hipLaunchKernelGGL((k_batch_step2<ST, KMI_TRT, true>), g, b, sh, s, a);              \
      else hipLaunchKernelGGL((k_batch_step2<ST, KMI_TRT, false>), g, b, sh, s, a);                 \
    } else {                                                                                        \
      if (nts) hipLaunchKernelGGL((k_batch_step2<ST, KMI_MRT, true>), g, b, sh, s, a);              \
      else hipLaunchKernelGGL((k_batch_step2<ST, KMI_MRT, false>), g, b, sh, s, a);                 \
    }                                                                                               \
  }
  KMI_BATCH_ST(st, KMI_BATCH_COLL2(ST));
#undef KMI_BATCH_COLL2
  return hipGetLastError();
}

static hipError_t launch_batch_stepK(int st, int coll, int K, const BatchArgs& a, int groups, hipStream_t s) {
  const int ne = stepK_region(a.tw, K);
  if (ne == 0) return hipErrorInvalidValue;
  const dim3 g(static_cast<unsigned>(groups)), b(static_cast<unsigned>(ne));
#define KMI_BATCH_COLLK(ST)                                                                      \
  {                                                                                             \
    const size_t sh = size_t(9) * ne * sizeof(typename Store<ST>::real);                        \
    if (K == 3) {                                                                               \
      if (coll == KMI_BGK) hipLaunchKernelGGL((k_batch_stepK<ST, KMI_BGK, 3>), g, b, sh, s, a); \
      else if (coll == KMI_TRT) hipLaunchKernelGGL((k_batch_stepK<ST, KMI_TRT, 3>), g, b, sh, s, a); \
      else hipLaunchKernelGGL((k_batch_stepK<ST, KMI_MRT, 3>), g, b, sh, s, a);                 \
    } else {                                                                                    \
      if (coll == KMI_BGK) hipLaunchKernelGGL((k_batch_stepK<ST, KMI_BGK, 4>), g, b, sh, s, a); \
      else if (coll == KMI_TRT) hipLaunchKernelGGL((k_batch_stepK<ST, KMI_TRT, 4>), g, b, sh, s, a); \
      else hipLaunchKernelGGL((k_batch_stepK<ST, KMI_MRT, 4>), g, b, sh, s, a);                 \
    }                                                                                           \
  }
  KMI_BATCH_ST(st, KMI_BATCH_COLLK(ST));
#undef KMI_BATCH_COLLK
  return hipGetLastError();
}

static hipError_t launch_batch_resid(int st, const BatchArgs& a, const void* f, int mode,
                                     int groups, hipStream_t s) {
  KMI_BATCH_ST(st, hipLaunchKernelGGL((k_batch_resid<ST>), dim3(unsigned(groups)), dim3(256), 0, s, a, f, mode));
  return hipGetLastError();
}

static hipError_t launch_batch_hold(int st, const BatchArgs& a, void* f, const int* hold, int groups,
                                    hipStream_t s) {
  KMI_BATCH_ST(st, hipLaunchKernelGGL((k_batch_hold<ST>), dim3(unsigned(groups)), dim3(256), 0, s, a, f, hold));
  return hipGetLastError();
}

}  // namespace kmi

using namespace kmi;

struct kmi_batch {
  kmi_ctx* ctx = nullptr;
  kmi_batch_desc desc{};
  int ncav = 0;
  std::vector<int> n;
  std::vector<double> nu;
  std::vector<CavRec> cav;
  CavRec* cav_d = nullptr;
  size_t esize = 4;
  long long total_pop = 0, total_cells = 0;
  void* f[2] = {nullptr, nullptr};
  int cur = 0;
  // tiles
  int tw = 64, th = 4, groups = 0, nts = 0;
  int span = 1;  // steps per launch: 1, 2 (k_batch_step2), 3 or 4 (k_batch_stepK)
  bool tiles_dirty = true;
  int nslots = 0;
  int* slots_d = nullptr;
  size_t slots_cap = 0;
  int* cav_slots_d = nullptr;
  size_t cav_slots_cap = 0;
  int* cav_range_d = nullptr;
  int* active_d = nullptr;
  double* uprev = nullptr;
  double* partials = nullptr;
  size_t partials_cap = 0;
  double* resid_d = nullptr;  // 3 per cavity
  std::vector<int> status;
  std::vector<int64_t> steps;
  std::vector<double> resid;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_next[2] = {0, 1};  // buffer parity after a replay of graph[p]
  int64_t graph_check = 0, graph_window = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
};

static void batch_drop_graphs(kmi_batch* B) {
  for (auto& g : B->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

static int batch_groups(const kmi_batch* B) {
  int g = B->groups > 0 ? std::min(B->groups, B->nslots) : B->nslots;
  g = ((g + 7) / 8) * 8;  // multiple of 8: slot s keeps running on XCD s mod 8
  return std::max(g, 8);
}

// Cavity-major tile order, then slot permutation so XCD x owns run x of that order.
static int batch_build_tiles(kmi_batch* B) {
  kmi_ctx* c = B->ctx;
  std::vector<int> tiles;  // (cavity, x0, y0)
  std::vector<int> first(B->ncav + 1, 0);
  for (int cv = 0; cv < B->ncav; ++cv) {
    first[cv] = int(tiles.size() / 3);
    if (B->status[cv] != KMI_CAV_RUNNING) continue;
    const int n = B->n[cv];
    for (int y = 0; y < n; y += B->th)
      for (int x = 0; x < n; x += B->tw) {
        tiles.push_back(cv);
        tiles.push_back(x);
        tiles.push_back(y);
      }
  }
  const int T = int(tiles.size() / 3);
  first[B->ncav] = T;
  const int chunk = (T + 7) / 8;
  B->nslots = std::max(8 * chunk, 8);
  std::vector<int> slots(size_t(3) * B->nslots, -1);
  std::vector<int> slot_of(T, -1);
  for (int s = 0; s < B->nslots; ++s) {
    const int idx = (s % 8) * chunk + s / 8;
    if (idx < T) {
      for (int k = 0; k < 3; ++k) slots[3 * s + k] = tiles[3 * idx + k];
      slot_of[idx] = s;
    }
  }
  // per-cavity slot lists in tile order (for the deterministic residual sums)
  std::vector<int> cav_slots(std::max(T, 1));
  std::vector<int> range(B->ncav + 1, 0), active(B->ncav, 0);
  for (int cv = 0; cv < B->ncav; ++cv) {
    range[cv] = first[cv];
    active[cv] = B->status[cv] == KMI_CAV_RUNNING;
  }
  range[B->ncav] = T;
  for (int k = 0; k < T; ++k) cav_slots[k] = slot_of[k];
  auto grow = [&](void** p, size_t& cap, size_t need) -> hipError_t {
    if (need <= cap) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(p, need);
    if (e == hipSuccess) cap = need;
    return e;
  };
  KMI_HIP(c, grow((void**)&B->slots_d, B->slots_cap, slots.size() * sizeof(int)));
  KMI_HIP(c, grow((void**)&B->cav_slots_d, B->cav_slots_cap, cav_slots.size() * sizeof(int)));
  KMI_HIP(c, grow((void**)&B->partials, B->partials_cap, size_t(4) * B->nslots * sizeof(double)));
  KMI_HIP(c, hipMemcpy(B->slots_d, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice));
  KMI_HIP(c, hipMemcpy(B->cav_slots_d, cav_slots.data(), cav_slots.size() * sizeof(int), hipMemcpyHostToDevice));
  KMI_HIP(c, hipMemcpy(B->cav_range_d, range.data(), range.size() * sizeof(int), hipMemcpyHostToDevice));
  KMI_HIP(c, hipMemcpy(B->active_d, active.data(), active.size() * sizeof(int), hipMemcpyHostToDevice));
  B->tiles_dirty = false;
  batch_drop_graphs(B);
  return KMI_OK;
}

static BatchArgs batch_args(const kmi_batch* B) {
  BatchArgs a{};
  a.fi = B->f[B->cur];
  a.fo = B->f[B->cur ^ 1];
  a.cav = B->cav_d;
  a.slots = B->slots_d;
  a.nslots = B->nslots;
  a.tw = B->tw;
  a.lid = B->desc.lid;
  a.u_lid = B->desc.u_lid;
  a.s_e = B->desc.s_e;
  a.s_eps = B->desc.s_eps;
  a.s_q = B->desc.s_q;
  a.uprev = B->uprev;
  a.total_cells = B->total_cells;
  a.partials = B->partials;
  return a;
}

static int batch_launch_step(kmi_batch* B, hipStream_t s) {
  BatchArgs a = batch_args(B);
  hipError_t e = launch_batch_step(B->desc.storage, B->desc.collision, B->nts != 0, a,
                                   batch_groups(B), s);
  if (e != hipSuccess) return hip_fail(B->ctx, e, "batch step launch");
  B->cur ^= 1;
  return KMI_OK;
}

static int batch_launch_step2(kmi_batch* B, hipStream_t s) {
  BatchArgs a = batch_args(B);
  hipError_t e = launch_batch_step2(B->desc.storage, B->desc.collision, B->nts != 0, a,
                                    batch_groups(B), s);
  if (e != hipSuccess) return hip_fail(B->ctx, e, "batch two-step launch");
  B->cur ^= 1;
  return KMI_OK;
}

static int batch_launch_stepK(kmi_batch* B, hipStream_t s) {
  BatchArgs a = batch_args(B);
  hipError_t e = launch_batch_stepK(B->desc.storage, B->desc.collision, B->span, a, batch_groups(B), s);
  if (e != hipSuccess) return hip_fail(B->ctx, e, "batch multi-step launch");
  B->cur ^= 1;
  return KMI_OK;
}

// n steps: multi-step launches when enabled, then two-step and single steps for the rest.
static int batch_launch_n(kmi_batch* B, int64_t n, hipStream_t s) {
  if (B->span >= 3)
    for (; n >= B->span; n -= B->span)
      if (int rc = batch_launch_stepK(B, s)) return rc;
  if (B->span >= 2)
    for (; n >= 2; n -= 2)
      if (int rc = batch_launch_step2(B, s)) return rc;
  for (; n > 0; --n)
    if (int rc = batch_launch_step(B, s)) return rc;
  return KMI_OK;
}

static int batch_ensure_tiles(kmi_batch* B) {
  return B->tiles_dirty ? batch_build_tiles(B) : KMI_OK;
}

static int batch_active(const kmi_batch* B) {
  int k = 0;
  for (int s : B->status) k += s == KMI_CAV_RUNNING;
  return k;
}

// One check block: (check − m) steps, snapshot u, m steps, residual partials, per-cavity sums.
static int batch_record_block(kmi_batch* B, int64_t check, int64_t window, hipStream_t s) {
  if (int rc = batch_launch_n(B, check - window, s)) return rc;
  BatchArgs a = batch_args(B);
  hipError_t e = launch_batch_resid(B->desc.storage, a, B->f[B->cur], 0, batch_groups(B), s);
  if (e != hipSuccess) return hip_fail(B->ctx, e, "batch snapshot");
  if (int rc = batch_launch_n(B, window, s)) return rc;
  a = batch_args(B);
  e = launch_batch_resid(B->desc.storage, a, B->f[B->cur], 1, batch_groups(B), s);
  if (e != hipSuccess) return hip_fail(B->ctx, e, "batch residual");
  hipLaunchKernelGGL(k_batch_final, dim3(unsigned(B->ncav)), dim3(64), 0, s,
                     B->partials, B->cav_slots_d, B->cav_range_d, B->active_d, B->ncav,
                     double(window), B->resid_d);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(B->ctx, e, "batch residual sum");
  return KMI_OK;
}

extern "C" {

int kmi_batch_create(kmi_ctx* ctx, const kmi_batch_desc* desc, int32_t ncav, const int32_t* n,
                     const double* nu, kmi_batch** out) {
  if (!ctx) return fail(nullptr, KMI_ERR_ARG, "kmi_batch_create: ctx is NULL");
  if (!desc || !n || !nu || !out || ncav < 1)
    return fail(ctx, KMI_ERR_ARG, "kmi_batch_create: bad arguments");
  *out = nullptr;
  if (desc->storage < KMI_STORE_F64 || desc->storage > KMI_STORE_F16S)
    return fail(ctx, KMI_ERR_ARG, "kmi_batch_create: unknown storage");
  if (desc->collision < KMI_BGK || desc->collision > KMI_MRT)
    return fail(ctx, KMI_ERR_ARG, "kmi_batch_create: unknown collision");
  if (desc->lid != KMI_FACE_ZOUHE_VELOCITY && desc->lid != KMI_FACE_MOVING_WALL)
    return fail(ctx, KMI_ERR_ARG, "kmi_batch_create: lid must be Zou-He velocity or moving wall");
  long long cells = 0;
  for (int c = 0; c < ncav; ++c) {
    if (n[c] < 2 || n[c] > 16384) return fail(ctx, KMI_ERR_ARG, "kmi_batch_create: cavity side out of range");
    if (!(nu[c] > 0)) return fail(ctx, KMI_ERR_ARG, "kmi_batch_create: viscosity must be positive");
    cells += (long long)n[c] * n[c];
  }
  if (double(cells) * 9 >= 2147483647.0)
    return fail(ctx, KMI_ERR_UNSUPPORTED, "kmi_batch_create: more than 2^31 populations");
  KMI_HIP(ctx, hipSetDevice(ctx->device));
  kmi_batch* B = new (std::nothrow) kmi_batch();
  if (!B) return fail(ctx, KMI_ERR_NOMEM, "kmi_batch_create: out of host memory");
  B->ctx = ctx;
  B->desc = *desc;
  B->ncav = ncav;
  B->n.assign(n, n + ncav);
  B->nu.assign(nu, nu + ncav);
  B->esize = storage_bytes(desc->storage);
  long long off = 0, coff = 0;
  for (int c = 0; c < ncav; ++c) {
    CavRec r{};
    r.foff = off;
    r.coff = coff;
    r.n = n[c];
    r.om = 1.0 / (3.0 * nu[c] + 0.5);  // ω = s_ν = s⁺ = 1/(3ν + 1/2)
    r.sm = 1.0 / (desc->Lambda / (3.0 * nu[c]) + 0.5);
    B->cav.push_back(r);
    off += 9LL * n[c] * n[c];
    coff += (long long)n[c] * n[c];
  }
  B->total_pop = off;
  B->total_cells = coff;
  B->status.assign(ncav, KMI_CAV_RUNNING);
  B->steps.assign(ncav, 0);
  B->resid.assign(ncav, INFINITY);
  hipError_t e = hipMalloc(&B->f[0], size_t(off) * B->esize);
  if (e == hipSuccess) e = hipMalloc(&B->f[1], size_t(off) * B->esize);
  if (e == hipSuccess) e = hipMalloc(&B->cav_d, sizeof(CavRec) * ncav);
  if (e == hipSuccess) e = hipMalloc(&B->cav_range_d, sizeof(int) * (ncav + 1));
  if (e == hipSuccess) e = hipMalloc(&B->active_d, sizeof(int) * ncav);
  if (e == hipSuccess) e = hipMalloc(&B->resid_d, sizeof(double) * 3 * ncav);
  if (e == hipSuccess) e = hipMalloc(&B->uprev, sizeof(double) * 2 * coff);
  if (e == hipSuccess) e = hipMemcpy(B->cav_d, B->cav.data(), sizeof(CavRec) * ncav, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&B->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&B->ev[1]);
  if (e != hipSuccess) {
    kmi_batch_destroy(B);
    return hip_fail(ctx, e, "kmi_batch_create");
  }
  // rest state f = w_q for every cavity (initialize_2d, basic.jl:64-70)
  for (int c = 0; c < ncav; ++c) {
    if (int rc = kmi_batch_init_equilibrium(B, c, nullptr, nullptr, nullptr)) {
      kmi_batch_destroy(B);
      return rc;
    }
  }
  if (int rc = batch_build_tiles(B)) {
    kmi_batch_destroy(B);
    return rc;
  }
  *out = B;
  return KMI_OK;
}

int kmi_batch_destroy(kmi_batch* B) {
  if (!B) return KMI_OK;
  (void)hipSetDevice(B->ctx->device);
  (void)hipStreamSynchronize(B->ctx->stream);
  batch_drop_graphs(B);
  void* ptrs[] = {B->f[0], B->f[1], B->cav_d, B->slots_d, B->cav_slots_d, B->cav_range_d,
                  B->active_d, B->uprev, B->partials, B->resid_d};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto ev : B->ev)
    if (ev) (void)hipEventDestroy(ev);
  delete B;
  return KMI_OK;
}

// A cavity that is not running keeps its populations in both buffers (the parity moves on
// while the others step): after writing f[cur], copy the region to the other buffer.
static hipError_t batch_mirror_if_retired(kmi_batch* B, int32_t c, hipStream_t s) {
  if (B->status[c] == KMI_CAV_RUNNING) return hipSuccess;
  const size_t off = size_t(B->cav[c].foff) * B->esize;
  const size_t bytes = size_t(9) * B->n[c] * B->n[c] * B->esize;
  return hipMemcpyAsync(static_cast<char*>(B->f[B->cur ^ 1]) + off,
                        static_cast<const char*>(B->f[B->cur]) + off, bytes,
                        hipMemcpyDeviceToDevice, s);
}

static int batch_check_cav(kmi_batch* B, int32_t c) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  if (c < 0 || c >= B->ncav) return fail(B->ctx, KMI_ERR_ARG, "cavity index out of range");
  return KMI_OK;
}

int kmi_batch_upload(kmi_batch* B, int32_t c, const void* host, int32_t dt) {
  if (int rc = batch_check_cav(B, c)) return rc;
  kmi_ctx* x = B->ctx;
  if (!host || (dt != KMI_HOST_F64 && dt != KMI_HOST_F32)) return fail(x, KMI_ERR_ARG, "kmi_batch_upload: bad buffer");
  KMI_HIP(x, hipSetDevice(x->device));
  const long long nc = (long long)B->n[c] * B->n[c];
  std::vector<double> tmp;
  const double* src = static_cast<const double*>(host);
  if (dt == KMI_HOST_F32) {
    tmp.resize(size_t(9 * nc));
    for (long long i = 0; i < 9 * nc; ++i) tmp[i] = static_cast<const float*>(host)[i];
    src = tmp.data();
  }
  double* stage = nullptr;
  KMI_HIP(x, hipMalloc(&stage, size_t(9 * nc) * sizeof(double)));
  hipError_t e = hipMemcpyAsync(stage, src, size_t(9 * nc) * sizeof(double), hipMemcpyHostToDevice, x->stream);
  char* base = static_cast<char*>(B->f[B->cur]) + B->cav[c].foff * B->esize;
  if (e == hipSuccess) e = launch_to_storage(9, B->desc.storage, stage, base, nc, nc, 1.0, x->stream);
  if (e == hipSuccess) e = batch_mirror_if_retired(B, c, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  (void)hipFree(stage);
  if (e != hipSuccess) return hip_fail(x, e, "kmi_batch_upload");
  return KMI_OK;
}

int kmi_batch_download(kmi_batch* B, int32_t c, void* host, int32_t dt) {
  if (int rc = batch_check_cav(B, c)) return rc;
  kmi_ctx* x = B->ctx;
  if (!host || (dt != KMI_HOST_F64 && dt != KMI_HOST_F32)) return fail(x, KMI_ERR_ARG, "kmi_batch_download: bad buffer");
  KMI_HIP(x, hipSetDevice(x->device));
  const long long nc = (long long)B->n[c] * B->n[c];
  double* stage = nullptr;
  KMI_HIP(x, hipMalloc(&stage, size_t(9 * nc) * sizeof(double)));
  std::vector<double> tmp;
  double* dst = static_cast<double*>(host);
  if (dt == KMI_HOST_F32) {
    tmp.resize(size_t(9 * nc));
    dst = tmp.data();
  }
  const char* base = static_cast<const char*>(B->f[B->cur]) + B->cav[c].foff * B->esize;
  hipError_t e = launch_from_storage(9, B->desc.storage, base, stage, nc, nc, 1.0, x->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dst, stage, size_t(9 * nc) * sizeof(double), hipMemcpyDeviceToHost, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  (void)hipFree(stage);
  if (e != hipSuccess) return hip_fail(x, e, "kmi_batch_download");
  if (dt == KMI_HOST_F32)
    for (long long i = 0; i < 9 * nc; ++i) static_cast<float*>(host)[i] = float(tmp[i]);
  return KMI_OK;
}

int kmi_batch_init_equilibrium(kmi_batch* B, int32_t c, const double* rho, const double* ux,
                               const double* uy) {
  if (int rc = batch_check_cav(B, c)) return rc;
  kmi_ctx* x = B->ctx;
  KMI_HIP(x, hipSetDevice(x->device));
  const long long nc = (long long)B->n[c] * B->n[c];
  const double* host[3] = {rho, ux, uy};
  double* dev[3] = {nullptr, nullptr, nullptr};
  double* block = nullptr;
  int used = 0;
  for (auto h : host) used += h != nullptr;
  if (used) KMI_HIP(x, hipMalloc(&block, size_t(used) * nc * sizeof(double)));
  hipError_t e = hipSuccess;
  for (int k = 0, slot = 0; k < 3 && e == hipSuccess; ++k) {
    if (!host[k]) continue;
    dev[k] = block + size_t(slot++) * nc;
    e = hipMemcpyAsync(dev[k], host[k], size_t(nc) * sizeof(double), hipMemcpyHostToDevice, x->stream);
  }
  char* base = static_cast<char*>(B->f[B->cur]) + B->cav[c].foff * B->esize;
  if (e == hipSuccess) e = launch_init_eq(9, B->desc.storage, base, dev[0], dev[1], dev[2], nullptr, nc, nc, x->stream);
  if (e == hipSuccess) e = batch_mirror_if_retired(B, c, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  if (block) (void)hipFree(block);
  if (e != hipSuccess) return hip_fail(x, e, "kmi_batch_init_equilibrium");
  return KMI_OK;
}

int kmi_batch_moments(kmi_batch* B, int32_t c, double* rho, double* ux, double* uy) {
  if (int rc = batch_check_cav(B, c)) return rc;
  kmi_ctx* x = B->ctx;
  KMI_HIP(x, hipSetDevice(x->device));
  const long long nc = (long long)B->n[c] * B->n[c];
  double* block = nullptr;
  KMI_HIP(x, hipMalloc(&block, size_t(3) * nc * sizeof(double)));
  const char* base = static_cast<const char*>(B->f[B->cur]) + B->cav[c].foff * B->esize;
  hipError_t e = launch_moments(9, B->desc.storage, base, nullptr, block, block + nc, block + 2 * nc,
                                nullptr, nullptr, nc, nc, 0.0, MomForce{{0.0, 0.0, 0.0}, nullptr},
                                x->stream);
  double* host[3] = {rho, ux, uy};
  for (int k = 0; k < 3 && e == hipSuccess; ++k)
    if (host[k]) e = hipMemcpyAsync(host[k], block + k * nc, size_t(nc) * sizeof(double), hipMemcpyDeviceToHost, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  (void)hipFree(block);
  if (e != hipSuccess) return hip_fail(x, e, "kmi_batch_moments");
  return KMI_OK;
}

int kmi_batch_reset_status(kmi_batch* B) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  std::fill(B->status.begin(), B->status.end(), int(KMI_CAV_RUNNING));
  std::fill(B->steps.begin(), B->steps.end(), 0);
  std::fill(B->resid.begin(), B->resid.end(), INFINITY);
  B->tiles_dirty = true;
  return KMI_OK;
}

int kmi_batch_resume(kmi_batch* B) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  for (auto& s : B->status)
    if (s == KMI_CAV_MAX_STEPS) {
      s = KMI_CAV_RUNNING;
      B->tiles_dirty = true;
    }
  return KMI_OK;
}

int kmi_batch_step(kmi_batch* B, int64_t nsteps) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  if (nsteps < 0) return fail(B->ctx, KMI_ERR_ARG, "kmi_batch_step: nsteps must be >= 0");
  kmi_ctx* x = B->ctx;
  KMI_HIP(x, hipSetDevice(x->device));
  if (int rc = batch_ensure_tiles(B)) return rc;
  if (batch_active(B) == 0) return KMI_OK;
  if (int rc = batch_launch_n(B, nsteps, x->stream)) return rc;
  for (int c = 0; c < B->ncav; ++c)
    if (B->status[c] == KMI_CAV_RUNNING) B->steps[c] += nsteps;
  KMI_HIP(x, hipStreamSynchronize(x->stream));
  return KMI_OK;
}

int kmi_batch_consistent_init(kmi_batch* B, const int32_t* hold, int32_t iterations) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  kmi_ctx* x = B->ctx;
  if (iterations < 0) return fail(x, KMI_ERR_ARG, "kmi_batch_consistent_init: iterations must be >= 0");
  KMI_HIP(x, hipSetDevice(x->device));
  if (int rc = batch_ensure_tiles(B)) return rc;
  if (iterations == 0 || batch_active(B) == 0) return KMI_OK;
  std::vector<int> h(B->ncav, 1);
  if (hold)
    for (int c = 0; c < B->ncav; ++c) h[c] = hold[c] != 0;
  int* hold_d = nullptr;
  KMI_HIP(x, hipMalloc(&hold_d, sizeof(int) * B->ncav));
  hipStream_t s = x->stream;
  hipError_t e = hipMemcpyAsync(hold_d, h.data(), sizeof(int) * B->ncav, hipMemcpyHostToDevice, s);
  // the held velocities: the current moments, into uprev (scratch between residual checks)
  BatchArgs a = batch_args(B);
  if (e == hipSuccess) e = launch_batch_resid(B->desc.storage, a, B->f[B->cur], 0, batch_groups(B), s);
  int rc = KMI_OK;
  for (int32_t k = 0; k < iterations && e == hipSuccess && rc == KMI_OK; ++k) {
    rc = batch_launch_step(B, s);
    if (rc == KMI_OK) {
      a = batch_args(B);
      e = launch_batch_hold(B->desc.storage, a, B->f[B->cur], hold_d, batch_groups(B), s);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(hold_d);
  if (rc != KMI_OK) return rc;
  if (e != hipSuccess) return hip_fail(x, e, "kmi_batch_consistent_init");
  for (int c = 0; c < B->ncav; ++c)
    if (B->status[c] == KMI_CAV_RUNNING) B->steps[c] += iterations;
  return KMI_OK;
}

int kmi_batch_run(kmi_batch* B, double tol, int64_t check, int64_t window, int64_t max_steps,
                  double* ms) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  kmi_ctx* x = B->ctx;
  if (check < 2 || max_steps < 1 || window < 1 || window > check)
    return fail(x, KMI_ERR_ARG,
                "kmi_batch_run: need check_every >= 2, 1 <= window <= check_every, max_steps >= 1");
  KMI_HIP(x, hipSetDevice(x->device));
  hipStream_t s = x->stream;
  if (B->graph_check != check || B->graph_window != window) {
    batch_drop_graphs(B);
    B->graph_check = check;
    B->graph_window = window;
  }
  KMI_HIP(x, hipEventRecord(B->ev[0], s));
  std::vector<double> r(size_t(3) * B->ncav);
  while (batch_active(B) > 0) {
    if (int rc = batch_ensure_tiles(B)) return rc;
    const int p = B->cur;
    if (!B->graph[p]) {
      hipGraph_t gph = nullptr;
      KMI_HIP(x, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc = batch_record_block(B, check, window, s);
      hipError_t e = hipStreamEndCapture(s, &gph);
      B->graph_next[p] = B->cur;
      B->cur = p;  // capture recorded the launches only
      if (rc != KMI_OK) {
        if (gph) (void)hipGraphDestroy(gph);
        return rc;
      }
      if (e != hipSuccess) return hip_fail(x, e, "batch graph capture");
      e = hipGraphInstantiate(&B->graph[p], gph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gph);
      if (e != hipSuccess) return hip_fail(x, e, "batch graph instantiate");
    }
    KMI_HIP(x, hipGraphLaunch(B->graph[p], s));
    B->cur = B->graph_next[p];
    KMI_HIP(x, hipMemcpyAsync(r.data(), B->resid_d, r.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    KMI_HIP(x, hipStreamSynchronize(s));
    for (int c = 0; c < B->ncav; ++c) {
      if (B->status[c] != KMI_CAV_RUNNING) continue;
      B->steps[c] += check;
      const double ru = r[3 * c + 2];
      B->resid[c] = ru;
      int st = KMI_CAV_RUNNING;
      if (!std::isfinite(ru) && !(r[3 * c] == 0.0 && r[3 * c + 1] == 0.0)) st = KMI_CAV_DIVERGED;
      else if (ru < tol || (r[3 * c] == 0.0 && r[3 * c + 1] == 0.0)) st = KMI_CAV_CONVERGED;
      else if (B->steps[c] >= max_steps) st = KMI_CAV_MAX_STEPS;
      if (st != KMI_CAV_RUNNING) {
        B->status[c] = st;
        B->tiles_dirty = true;
        // no longer stepped while the buffer parity keeps changing: both buffers hold it
        KMI_HIP(x, batch_mirror_if_retired(B, c, s));
      }
    }
  }
  KMI_HIP(x, hipEventRecord(B->ev[1], s));
  KMI_HIP(x, hipEventSynchronize(B->ev[1]));
  float t = 0.f;
  KMI_HIP(x, hipEventElapsedTime(&t, B->ev[0], B->ev[1]));
  if (ms) *ms = t;
  return KMI_OK;
}

int kmi_batch_status(const kmi_batch* B, int32_t* status, int64_t* steps, double* residual) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  for (int c = 0; c < B->ncav; ++c) {
    if (status) status[c] = B->status[c];
    if (steps) steps[c] = B->steps[c];
    if (residual) residual[c] = B->resid[c];
  }
  return KMI_OK;
}

int kmi_batch_set_launch(kmi_batch* B, int32_t tw, int32_t th, int32_t groups, int32_t nts) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  if (!((tw == 16 || tw == 32 || tw == 64 || tw == 128 || tw == 256) && tw * th == 256) || groups < 0)
    return fail(B->ctx, KMI_ERR_ARG, "kmi_batch_set_launch: tiles of tw x th = 256 cells, tw in {16,32,64,128,256}");
  if (B->span >= 3 && stepK_region(tw, B->span) == 0)
    return fail(B->ctx, KMI_ERR_ARG, "kmi_batch_set_launch: tile too wide for the current steps per launch");
  KMI_HIP(B->ctx, hipSetDevice(B->ctx->device));
  KMI_HIP(B->ctx, hipStreamSynchronize(B->ctx->stream));
  B->tw = tw;
  B->th = th;
  B->groups = groups;
  B->nts = nts ? 1 : 0;
  B->tiles_dirty = true;
  batch_drop_graphs(B);
  return KMI_OK;
}

int kmi_batch_set_temporal(kmi_batch* B, int32_t steps_per_launch) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  if (steps_per_launch < 1 || steps_per_launch > 4)
    return fail(B->ctx, KMI_ERR_ARG, "kmi_batch_set_temporal: steps per launch must be 1, 2, 3 or 4");
  if (steps_per_launch >= 3 && stepK_region(B->tw, steps_per_launch) == 0)
    return fail(B->ctx, KMI_ERR_ARG,
                "kmi_batch_set_temporal: 3 or 4 steps per launch need tiles of width 16, 32 or 64");
  KMI_HIP(B->ctx, hipSetDevice(B->ctx->device));
  KMI_HIP(B->ctx, hipStreamSynchronize(B->ctx->stream));
  B->span = steps_per_launch;
  batch_drop_graphs(B);
  return KMI_OK;
}

int kmi_batch_get_temporal(const kmi_batch* B, int32_t* steps_per_launch) {
  if (!B || !steps_per_launch) return fail(nullptr, KMI_ERR_ARG, "kmi_batch_get_temporal: NULL argument");
  *steps_per_launch = B->span;
  return KMI_OK;
}

int kmi_batch_get_launch(const kmi_batch* B, int32_t* tw, int32_t* th, int32_t* groups, int32_t* nts) {
  if (!B || !tw || !th || !groups || !nts) return fail(nullptr, KMI_ERR_ARG, "kmi_batch_get_launch: NULL argument");
  *tw = B->tw;
  *th = B->th;
  *groups = B->groups;
  *nts = B->nts;
  return KMI_OK;
}

int kmi_batch_autotune(kmi_batch* B, int64_t steps_per_trial, double* best_ms) {
  if (!B) return fail(nullptr, KMI_ERR_ARG, "batch is NULL");
  kmi_ctx* x = B->ctx;
  KMI_HIP(x, hipSetDevice(x->device));
  KMI_HIP(x, hipStreamSynchronize(x->stream));
  if (batch_active(B) == 0) return fail(x, KMI_ERR_STATE, "kmi_batch_autotune: no running cavity");
  int64_t trial = std::max<int64_t>(2, steps_per_trial);
  trial = (trial + 11) / 12 * 12;  // whole launches of every span (1, 2, 3, 4)
  const size_t bytes = size_t(B->total_pop) * B->esize;
  void* save = nullptr;
  KMI_HIP(x, hipMalloc(&save, bytes));
  hipError_t e = hipMemcpyAsync(save, B->f[B->cur], bytes, hipMemcpyDeviceToDevice, x->stream);
  const int cur0 = B->cur;
  const int shapes[][2] = {{256, 1}, {128, 2}, {64, 4}, {32, 8}, {16, 16}};
  const int group_opts[] = {0, 8192, 4096, 2048, 1024};
  double best = 1e300;
  int bb[5] = {B->tw, B->th, B->groups, B->nts, B->span};
  for (const auto& sh : shapes) {
    for (int g : group_opts) {
      for (int nt = 0; nt < 8 && e == hipSuccess; ++nt) {
        B->tw = sh[0];
        B->th = sh[1];
        B->groups = g;
        B->nts = nt & 1;
        B->span = 1 + (nt >> 1);
        if (B->span >= 3 && (B->nts || stepK_region(B->tw, B->span) == 0)) continue;  // K-step stores are plain
        if (batch_build_tiles(B) != KMI_OK) { e = hipErrorUnknown; break; }
        if (g > 0 && g >= B->nslots) continue;  // same as one group per slot
        if (batch_launch_n(B, B->span, x->stream) != KMI_OK) { e = hipErrorUnknown; break; }
        double t = 1e300;
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
          e = hipEventRecord(B->ev[0], x->stream);
          if (e == hipSuccess && batch_launch_n(B, trial, x->stream) != KMI_OK) e = hipErrorUnknown;
          if (e == hipSuccess) e = hipEventRecord(B->ev[1], x->stream);
          if (e == hipSuccess) e = hipEventSynchronize(B->ev[1]);
          float ms = 0.f;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, B->ev[0], B->ev[1]);
          t = std::min(t, double(ms) / double(trial));
        }
        if (e == hipSuccess && t < best) {
          best = t;
          bb[0] = sh[0];
          bb[1] = sh[1];
          bb[2] = g;
          bb[3] = nt & 1;
          bb[4] = B->span;
        }
      }
    }
  }
  B->tw = bb[0];
  B->th = bb[1];
  B->groups = bb[2];
  B->nts = bb[3];
  B->span = bb[4];
  B->tiles_dirty = true;
  batch_drop_graphs(B);
  hipError_t e2 = hipMemcpyAsync(B->f[cur0], save, bytes, hipMemcpyDeviceToDevice, x->stream);
  if (e2 == hipSuccess) e2 = hipStreamSynchronize(x->stream);
  (void)hipFree(save);
  B->cur = cur0;
  if (e != hipSuccess) return hip_fail(x, e, "kmi_batch_autotune");
  if (e2 != hipSuccess) return hip_fail(x, e2, "kmi_batch_autotune (restore)");
  if (best_ms) *best_ms = best;
  return batch_ensure_tiles(B);
}

int kmi_batch_active_cells(const kmi_batch* B, int64_t* cells) {
  if (!B || !cells) return fail(nullptr, KMI_ERR_ARG, "kmi_batch_active_cells: NULL argument");
  int64_t s = 0;
  for (int c = 0; c < B->ncav; ++c)
    if (B->status[c] == KMI_CAV_RUNNING) s += int64_t(B->n[c]) * B->n[c];
  *cells = s;
  return KMI_OK;
}

}  // extern "C"
