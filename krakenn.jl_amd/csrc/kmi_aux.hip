// kmi_aux.hip — layout conversion, initial states, moments and FP64 reductions.
//
// Reductions are deterministic and compensated at every level: each thread accumulates its
// cells in FP64 as a Neumaier pair (sum, error), a wave64 butterfly (__shfl_xor) and an LDS
// pass across waves merge the pairs, each work-group writes its pair, and one 256-thread
// work-group folds the group pairs (strided per thread, then a fixed LDS tree).  No atomics,
// so repeated calls return identical bits.
#include "kmi_common.h"

namespace kmi {

constexpr int kRedBlock = 256;
constexpr int kRedMaxVals = 4;

// --- conversion host layout (double, [cell, q]) <-> storage ------------------------------
// shift_q: w_q for f, w_q T_ref for g (shifted storage only).
template <int Q, int ST>
__global__ void k_to_storage(const double* __restrict__ src, void* dst_, long long ncells,
                             long long plane, double shift_scale) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  E* dst = static_cast<E*>(dst_);
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    double v = src[c + q * ncells];
    if constexpr (S::shifted) v = v - Lat<Q>::w[q] * shift_scale;
    S::st(dst, c + q * plane, R(v));
  });
}

template <int Q, int ST>
__global__ void k_from_storage(const void* src_, double* __restrict__ dst, long long ncells,
                               long long plane, double shift_scale) {
  using S = Store<ST>;
  using E = typename S::elem;
  const E* src = static_cast<const E*>(src_);
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    double v = double(S::ld(src, c + q * plane));
    if constexpr (S::shifted) v = v + Lat<Q>::w[q] * shift_scale;
    dst[c + q * ncells] = v;
  });
}

// --- initial states -------------------------------------------------------------------
// f_q = feq_q(ρ, u), written as lattice/d2q9.jl:60-67 / d3q19.jl:67-75 (cu = Σ c_a u_a).
template <int Q, int ST>
__global__ void k_init_eq(void* dst_, const double* __restrict__ rho, const double* __restrict__ ux,
                          const double* __restrict__ uy, const double* __restrict__ uz,
                          long long ncells, long long plane) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  E* dst = static_cast<E*>(dst_);
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const double r = rho ? rho[c] : 1.0;
  const double u = ux ? ux[c] : 0.0, v = uy ? uy[c] : 0.0, w3 = uz ? uz[c] : 0.0;
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr double w = Lat<Q>::w[q];
    if constexpr (!S::shifted) {
      const R rr = R(r), a = R(u), b = R(v), d = R(w3);
      R cu = R(Lat<Q>::cx[q]) * a + R(Lat<Q>::cy[q]) * b;
      if constexpr (Q == 19) cu = cu + R(Lat<Q>::cz[q]) * d;
      R usq = a * a + b * b;
      if constexpr (Q == 19) usq = usq + d * d;
      S::st(dst, c + q * plane,
            R(w) * rr * (R(1) + R(3) * cu + R(4.5) * cu * cu - R(1.5) * usq));
    } else {
      double cu = Lat<Q>::cx[q] * u + Lat<Q>::cy[q] * v + Lat<Q>::cz[q] * w3;
      double usq = u * u + v * v + w3 * w3;
      double x = 3.0 * cu + 4.5 * cu * cu - 1.5 * usq;
      S::st(dst, c + q * plane, R(w * ((r - 1.0) + r * x)));
    }
  });
}

// g_q = w_q T (thermal.jl:94-96, 489-496).
template <int Q, int ST>
__global__ void k_init_temp(void* dst_, const double* __restrict__ T, long long ncells,
                            long long plane, double T0) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  E* dst = static_cast<E*>(dst_);
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const double t = T ? T[c] : T0;
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (!S::shifted) S::st(dst, c + q * plane, R(Lat<Q>::w[q]) * R(t));
    else S::st(dst, c + q * plane, R(Lat<Q>::w[q] * (t - T0)));
  });
}

// --- per-cell moments ------------------------------------------------------------------
// Plain storage: compute_macroscopic_2d/3d and compute_temperature_2d/3d in the storage type
// (macroscopic.jl:5-61, thermal_2d.jl:113-119, thermal_3d.jl:128-136).  Shifted: FP64 from
// the deviations.  MomForce: Guo body force, u = (Σcf + F/2)/ρ as
// compute_macroscopic_forced_2d!/_3d!/_field_3d! (macroscopic.jl:84-200); zero without
// forcing (adding zero leaves Σcf unchanged).
template <int Q, int ST>
KMI_DEV void cell_moments(const void* f_, const void* g_, long long c, long long plane, double T0,
                          const MomForce& F, double& rho, double& ux, double& uy, double& uz,
                          double& T) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  const E* f = static_cast<const E*>(f_);
  R fx = R(F.f[0]), fy = R(F.f[1]), fz = R(F.f[2]);
  if (F.field) {
    const R* fld = static_cast<const R*>(F.field);
    fx = fld[c];
    fy = fld[c + plane];
    fz = Q == 19 ? fld[c + 2 * plane] : R(0);
  }
  if constexpr (!S::shifted) {
    R fp[Q];
    sfor<Q>([&](auto qc) { fp[decltype(qc)::value] = S::ld(f, c + decltype(qc)::value * plane); });
    const R r = zeroth_moment<Q>(fp);
    const R inv = R(1) / r;
    rho = r;
    ux = (first_moment<Q, 0>(fp) + fx / R(2)) * inv;
    uy = (first_moment<Q, 1>(fp) + fy / R(2)) * inv;
    if constexpr (Q == 19) uz = (first_moment<Q, 2>(fp) + fz / R(2)) * inv;
    else uz = 0.0;
  } else {
    double fp[Q];
    sfor<Q>([&](auto qc) {
      fp[decltype(qc)::value] = double(S::ld(f, c + decltype(qc)::value * plane));
    });
    const double r = 1.0 + zeroth_moment<Q>(fp);
    rho = r;
    ux = (first_moment<Q, 0>(fp) + double(fx) / 2.0) / r;
    uy = (first_moment<Q, 1>(fp) + double(fy) / 2.0) / r;
    if constexpr (Q == 19) uz = (first_moment<Q, 2>(fp) + double(fz) / 2.0) / r;
    else uz = 0.0;
  }
  T = 0.0;
  if (g_) {
    using GS = Store<thermal_storage(ST)>;
    const auto* g = static_cast<const typename GS::elem*>(g_);
    if constexpr (!GS::shifted) {
      R gp[Q];
      sfor<Q>([&](auto qc) { gp[decltype(qc)::value] = GS::ld(g, c + decltype(qc)::value * plane); });
      T = zeroth_moment<Q>(gp);
    } else {
      double gp[Q];
      sfor<Q>([&](auto qc) {
        gp[decltype(qc)::value] = double(GS::ld(g, c + decltype(qc)::value * plane));
      });
      T = T0 + zeroth_moment<Q>(gp);
    }
  }
}

template <int Q, int ST>
__global__ void k_moments(const void* f, const void* g, double* rho, double* ux, double* uy,
                          double* uz, double* T, long long ncells, long long plane, double T0,
                          MomForce F) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double r, u, v, w, t;
  cell_moments<Q, ST>(f, g, c, plane, T0, F, r, u, v, w, t);
  if (rho) rho[c] = r;
  if (ux) ux[c] = u;
  if (uy) uy[c] = v;
  if (uz) uz[c] = w;
  if (T) T[c] = t;
}

// --- block reduction -------------------------------------------------------------------
// v[n] + c[n] is thread's compensated sum of value n; partials receive (sum, error) per value:
// partials[2 (group NV + n)], partials[2 (group NV + n) + 1].  is_max: plain maxima, error 0.
template <int NV>
KMI_DEV void block_reduce(double (&v)[NV], double (&c)[NV], bool is_max, double* partials) {
  __shared__ double lds[kRedBlock / 64][kRedMaxVals][2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    double x = v[n], e = c[n];
    for (int off = 32; off > 0; off >>= 1) {
      const double y = __shfl_xor(x, off, 64), ye = __shfl_xor(e, off, 64);
      if (is_max) x = fmax(x, y);
      else nsum_merge(x, e, y, ye);
    }
    if (lane == 0) {
      lds[wave][n][0] = x;
      lds[wave][n][1] = e;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) / 64;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      double x = lds[0][n][0], e = lds[0][n][1];
      for (int w = 1; w < nw; ++w) {
        if (is_max) x = fmax(x, lds[w][n][0]);
        else nsum_merge(x, e, lds[w][n][0], lds[w][n][1]);
      }
      partials[2 * (blockIdx.x * NV + n)] = x;
      partials[2 * (blockIdx.x * NV + n) + 1] = is_max ? 0.0 : e;
    }
  }
}

// kind: 0 mass, 1 momentum, 2 kinetic energy, 3 thermal, 5 max speed,
//       100 residual (Σ fluid |u − u_prev|², Σ fluid |u|²), 101 snapshot u → u_prev.
template <int Q, int ST>
__global__ void __launch_bounds__(kRedBlock)
    k_reduce_cells(const void* f, const void* g, const uint8_t* mask, double* uprev,
                   long long ncells, long long plane, double T0, MomForce F, int kind,
                   double* partials) {
  double acc[3] = {0.0, 0.0, 0.0}, err[3] = {0.0, 0.0, 0.0};
  const bool is_max = kind == KMI_RED_MAX_SPEED;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncells;
       c += (long long)gridDim.x * blockDim.x) {
    double r, u, v, w, t;
    cell_moments<Q, ST>(f, g, c, plane, T0, F, r, u, v, w, t);
    switch (kind) {
      case KMI_RED_MASS: nsum_add(acc[0], err[0], r); break;
      case KMI_RED_MOMENTUM:
        nsum_add(acc[0], err[0], r * u);
        nsum_add(acc[1], err[1], r * v);
        nsum_add(acc[2], err[2], r * w);
        break;
      case KMI_RED_KINETIC_ENERGY: nsum_add(acc[0], err[0], 0.5 * r * (u * u + v * v + w * w)); break;
      case KMI_RED_THERMAL: nsum_add(acc[0], err[0], t); break;
      case KMI_RED_MAX_SPEED: acc[0] = fmax(acc[0], sqrt(u * u + v * v + w * w)); break;
      case 100:
        if (!(mask && mask[c] == kCellSolid)) {
          const double du = u - uprev[c], dv = v - uprev[c + ncells], dw = w - uprev[c + 2 * ncells];
          nsum_add(acc[0], err[0], du * du + dv * dv + dw * dw);
          nsum_add(acc[1], err[1], u * u + v * v + w * w);
        }
        break;
      case 101:
        uprev[c] = u;
        uprev[c + ncells] = v;
        uprev[c + 2 * ncells] = w;
        break;
      default: break;
    }
  }
  if (kind == 101) return;
  block_reduce<3>(acc, err, is_max, partials);
}

// Hot-wall Nusselt number per wall node (thermal.jl:225-234, 525-533):
// Nu = −H (−8 T_w + 9 T_1 − T_2) / (3 dx) / ΔT, with T_1, T_2 the first two nodes along the
// normal and H the cell count along it.  The wall nodes are the rectangle
// [lo1, hi1) × [lo2, hi2) of the face's two tangent indices (axis order; 2D: one tangent).
template <int Q, int ST>
__global__ void __launch_bounds__(kRedBlock)
    k_nusselt(const void* g_, int nx, int ny, int nz, long long plane, double T0, int face,
              double Th, double dT, int lo1, int hi1, int lo2, int hi2, double* partials) {
  using S = Store<ST>;
  using E = typename S::elem;
  const E* g = static_cast<const E*>(g_);
  double acc[1] = {0.0}, err[1] = {0.0};
  const int n[3] = {nx, ny, nz};
  const int ax = face / 2, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
  const double H = n[ax];
  const int w1 = hi1 - lo1;
  const long long count = (long long)w1 * (hi2 - lo2);
  const long long stride[3] = {1, nx, (long long)nx * ny};
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < count;
       p += (long long)gridDim.x * blockDim.x) {
    const long long a = lo1 + p % w1, b = lo2 + p / w1;
    const long long base = a * stride[t1] + b * stride[t2];
    const bool hi = face % 2 == 1;
    const long long c1 = base + (hi ? (n[ax] - 1) : 0) * stride[ax];
    const long long c2 = c1 + (hi ? -stride[ax] : stride[ax]);
    double T1 = 0.0, T2 = 0.0;
    if constexpr (!S::shifted) {
      using R = typename S::real;
      R x = S::ld(g, c1), y = S::ld(g, c2);
      for (int q = 1; q < Q; ++q) { x = x + S::ld(g, c1 + q * plane); y = y + S::ld(g, c2 + q * plane); }
      T1 = x;
      T2 = y;
    } else {
      double x = 0.0, y = 0.0;
      for (int q = 0; q < Q; ++q) { x += double(S::ld(g, c1 + q * plane)); y += double(S::ld(g, c2 + q * plane)); }
      T1 = T0 + x;
      T2 = T0 + y;
    }
    nsum_add(acc[0], err[0], -H * (-8 * Th + 9 * T1 - T2) / 3.0 / dT);
  }
  block_reduce<1>(acc, err, false, partials);
}

// Force on the cut-link walls, accumulated in FP64 over the link list.
//   mea = 1: Mei–Luo–Shyy momentum exchange with the Bouzidi arriving population,
//            per link c_q (f_q + f~_qbar)   (_compute_drag_libb_mei_2d_host,
//            cylinder_libb.jl:61-104; drag_gpu.jl:346-376 for the link form);
//   mea = 0: halfway-BB exchange 2 c_q f_q   (cylinder_libb.jl:27-42, li_bb_3d_v2.jl:183-201).
// f is the stored (post-collision) field.  The neighbour behind the link is looked up inside
// the domain only, falling back to the cell's own opposite population, as the reference does.
template <int Q, int ST>
__global__ void __launch_bounds__(kRedBlock)
    k_drag(const void* f_, long long nlinks, const int* __restrict__ cell,
           const int* __restrict__ qv, const double* __restrict__ qwv,
           const double* __restrict__ uwv, int nx, int ny, int nz, long long plane, int mea,
           double* partials) {
  using S = Store<ST>;
  using E = typename S::elem;
  const E* f = static_cast<const E*>(f_);
  auto full = [&](long long idx, int q) {
    double v = double(S::ld(f, idx + q * plane));
    if constexpr (S::shifted) {
      double w = 0.0;
      sfor<Q>([&](auto qc) { if (decltype(qc)::value == q) w = Lat<Q>::w[decltype(qc)::value]; });
      v = w + v;
    }
    return v;
  };
  double acc[3] = {0.0, 0.0, 0.0}, err[3] = {0.0, 0.0, 0.0};
  for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nlinks;
       n += (long long)gridDim.x * blockDim.x) {
    const long long c = cell[n];
    const int q = qv[n];
    int cx = 0, cy = 0, cz = 0, o = 0;
    double w = 0.0;
    sfor<Q>([&](auto qc) {
      constexpr int k = decltype(qc)::value;
      if (k == q) {
        cx = Lat<Q>::cx[k], cy = Lat<Q>::cy[k], cz = Lat<Q>::cz[k], o = Lat<Q>::opp[k];
        w = Lat<Q>::w[k];
      }
    });
    const double fq = full(c, q);
    double F;
    if (!mea) {
      F = 2.0 * fq;
    } else {
      const int i = int(c % nx), j = int((c / nx) % ny), k = int(c / ((long long)nx * ny));
      const int im = i - cx, jm = j - cy, km = k - cz;
      const bool in = im >= 0 && im < nx && jm >= 0 && jm < ny && km >= 0 && km < nz;
      const double back = in ? full(im + (long long)nx * (jm + (long long)ny * km), q) : full(c, o);
      const double bar = full(c, o);
      const double qw = qwv[n];
      const double* u = uwv + 3 * n;
      const double cu = double(cx) * u[0] + double(cy) * u[1] + (Q == 19 ? double(cz) * u[2] : 0.0);
      const double d = -6.0 * w * cu;
      double arriving;
      if (qw <= 0.5) {
        arriving = 2 * qw * fq + (1 - 2 * qw) * back + d;
      } else {
        const double inv2q = 1.0 / (2 * qw);
        arriving = inv2q * fq + (1 - inv2q) * bar + inv2q * d;
      }
      F = fq + arriving;
    }
    nsum_add(acc[0], err[0], double(cx) * F);
    nsum_add(acc[1], err[1], double(cy) * F);
    nsum_add(acc[2], err[2], double(cz) * F);
  }
  block_reduce<3>(acc, err, false, partials);
}

// Combine the `nb` group pairs of `nv` values: thread t folds groups t, t + 256, … in order,
// then a fixed LDS tree merges the 256 thread pairs.  One work-group of 256 threads.
__global__ void __launch_bounds__(256) k_finalize(const double* __restrict__ partials, int nb, int nv,
                                                  int is_max, double* __restrict__ out) {
  __shared__ double ls[256], lc[256];
  const int t = threadIdx.x;
  for (int n = 0; n < nv; ++n) {
    double s = is_max ? -HUGE_VAL : 0.0, c = 0.0;
    for (int b = t; b < nb; b += 256) {
      const double x = partials[2 * (b * nv + n)], e = partials[2 * (b * nv + n) + 1];
      if (is_max) s = fmax(s, x);
      else nsum_merge(s, c, x, e);
    }
    ls[t] = s;
    lc[t] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) {
        if (is_max) {
          ls[t] = fmax(ls[t], ls[t + w]);
        } else {
          double a = ls[t], ac = lc[t];
          nsum_merge(a, ac, ls[t + w], lc[t + w]);
          ls[t] = a;
          lc[t] = ac;
        }
      }
      __syncthreads();
    }
    if (t == 0) out[n] = is_max ? ls[0] : ls[0] + lc[0];
    __syncthreads();
  }
}

// Plain device copy (float4 loads, non-temporal float4 stores): the practical 1:1
// read/write HBM ceiling the stream–collide kernel is compared against.
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_copy_probe(const f32x4* __restrict__ s,
                                                    f32x4* __restrict__ d, long long n4) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (long long)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(s[i], d + i);
}

hipError_t launch_copy_probe(const void* src, void* dst, long long bytes, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_probe, dim3(1024), dim3(256), 0, s, static_cast<const f32x4*>(src),
                     static_cast<f32x4*>(dst), bytes / 16);
  return hipGetLastError();
}

// ---- host-side launchers ----------------------------------------------------------------
#define KMI_ST_SWITCH(st, CALL)                                \
  switch (st) {                                                \
    case KMI_STORE_F64: { constexpr int ST = KMI_STORE_F64; CALL; } break;   \
    case KMI_STORE_F32: { constexpr int ST = KMI_STORE_F32; CALL; } break;   \
    case KMI_STORE_F32S: { constexpr int ST = KMI_STORE_F32S; CALL; } break; \
    default: { constexpr int ST = KMI_STORE_F16S; CALL; } break;             \
  }

static inline dim3 grid1(long long n) { return dim3(unsigned((n + 255) / 256)); }

hipError_t launch_to_storage(int Q, int st, const double* src, void* dst, long long ncells,
                             long long plane, double shift, hipStream_t s) {
  if (Q == 9) { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_to_storage<9, ST>), grid1(ncells), dim3(256), 0, s, src, dst, ncells, plane, shift)); }
  else { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_to_storage<19, ST>), grid1(ncells), dim3(256), 0, s, src, dst, ncells, plane, shift)); }
  return hipGetLastError();
}

hipError_t launch_from_storage(int Q, int st, const void* src, double* dst, long long ncells,
                               long long plane, double shift, hipStream_t s) {
  if (Q == 9) { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_from_storage<9, ST>), grid1(ncells), dim3(256), 0, s, src, dst, ncells, plane, shift)); }
  else { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_from_storage<19, ST>), grid1(ncells), dim3(256), 0, s, src, dst, ncells, plane, shift)); }
  return hipGetLastError();
}

hipError_t launch_init_eq(int Q, int st, void* dst, const double* rho, const double* ux,
                          const double* uy, const double* uz, long long ncells, long long plane,
                          hipStream_t s) {
  if (Q == 9) { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_init_eq<9, ST>), grid1(ncells), dim3(256), 0, s, dst, rho, ux, uy, uz, ncells, plane)); }
  else { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_init_eq<19, ST>), grid1(ncells), dim3(256), 0, s, dst, rho, ux, uy, uz, ncells, plane)); }
  return hipGetLastError();
}

hipError_t launch_init_temp(int Q, int st, void* dst, const double* T, long long ncells,
                            long long plane, double T0, hipStream_t s) {
  if (Q == 9) { KMI_ST_SWITCH(thermal_storage(st), hipLaunchKernelGGL((k_init_temp<9, ST>), grid1(ncells), dim3(256), 0, s, dst, T, ncells, plane, T0)); }
  else { KMI_ST_SWITCH(thermal_storage(st), hipLaunchKernelGGL((k_init_temp<19, ST>), grid1(ncells), dim3(256), 0, s, dst, T, ncells, plane, T0)); }
  return hipGetLastError();
}

hipError_t launch_moments(int Q, int st, const void* f, const void* g, double* rho, double* ux,
                          double* uy, double* uz, double* T, long long ncells, long long plane,
                          double T0, const MomForce& F, hipStream_t s) {
  if (Q == 9) { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_moments<9, ST>), grid1(ncells), dim3(256), 0, s, f, g, rho, ux, uy, uz, T, ncells, plane, T0, F)); }
  else { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_moments<19, ST>), grid1(ncells), dim3(256), 0, s, f, g, rho, ux, uy, uz, T, ncells, plane, T0, F)); }
  return hipGetLastError();
}

// Number of partial slots a reduction launch needs (3 values per group).
int reduce_groups(long long ncells) {
  long long g = (ncells + kRedBlock - 1) / kRedBlock;
  return int(g < 2048 ? (g < 1 ? 1 : g) : 2048);
}

hipError_t launch_reduce_cells(int Q, int st, const void* f, const void* g, const uint8_t* mask,
                               double* uprev, long long ncells, long long plane, double T0,
                               const MomForce& F, int kind, double* partials, double* out,
                               hipStream_t s) {
  const int nb = reduce_groups(ncells);
  if (Q == 9) { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_reduce_cells<9, ST>), dim3(nb), dim3(kRedBlock), 0, s, f, g, mask, uprev, ncells, plane, T0, F, kind, partials)); }
  else { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_reduce_cells<19, ST>), dim3(nb), dim3(kRedBlock), 0, s, f, g, mask, uprev, ncells, plane, T0, F, kind, partials)); }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || kind == 101) return e;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partials, nb, 3,
                     kind == KMI_RED_MAX_SPEED ? 1 : 0, out);
  return hipGetLastError();
}

hipError_t launch_drag(int Q, int st, const void* f, long long nlinks, const int* cell,
                       const int* q, const double* qw, const double* uw, int nx, int ny, int nz,
                       long long plane, int mea, double* partials, double* out, hipStream_t s) {
  const int nb = reduce_groups(nlinks);
  if (Q == 9) { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_drag<9, ST>), dim3(nb), dim3(kRedBlock), 0, s, f, nlinks, cell, q, qw, uw, nx, ny, nz, plane, mea, partials)); }
  else { KMI_ST_SWITCH(st, hipLaunchKernelGGL((k_drag<19, ST>), dim3(nb), dim3(kRedBlock), 0, s, f, nlinks, cell, q, qw, uw, nx, ny, nz, plane, mea, partials)); }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partials, nb, 3, 0, out);
  return hipGetLastError();
}

hipError_t launch_nusselt(int Q, int st, const void* g, int nx, int ny, int nz, long long plane,
                          double T0, int face, double Th, double dT, int lo1, int hi1, int lo2,
                          int hi2, double* partials, double* out, hipStream_t s) {
  const int nb = reduce_groups((long long)(hi1 - lo1) * (hi2 - lo2));
  if (Q == 9) { KMI_ST_SWITCH(thermal_storage(st), hipLaunchKernelGGL((k_nusselt<9, ST>), dim3(nb), dim3(kRedBlock), 0, s, g, nx, ny, nz, plane, T0, face, Th, dT, lo1, hi1, lo2, hi2, partials)); }
  else { KMI_ST_SWITCH(thermal_storage(st), hipLaunchKernelGGL((k_nusselt<19, ST>), dim3(nb), dim3(kRedBlock), 0, s, g, nx, ny, nz, plane, T0, face, Th, dT, lo1, hi1, lo2, hi2, partials)); }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, partials, nb, 1, 0, out);
  return hipGetLastError();
}

}  // namespace kmi
