// kraken_oracle.cpp — CPU restatement of Kraken.jl's stream–collide hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the MI355X library is checked
// against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it, and only as the checker or the timed CPU baseline.  The product library
// never links, loads or falls back to it.
//
// Every function restates one reference kernel (paths relative to the Kraken.jl
// tree) operation for operation, in the same floating-point evaluation order, so
// that a build with -ffp-contract=off reproduces Julia's IEEE results (Julia does not
// contract a*b+c without muladd/@fastmath).  Arrays use the reference's column-major
// layout f[i, j, (k,) q] with 0-based i, j, k, q here.
//
// Pinning (see tests/test_oracle_*.py): lattice identities, equilibrium moments,
// delta-streaming, bounce-back / Ladd / Zou–He closed forms, BGK/TRT/MRT relaxation
// of the (anti)symmetric and moment-space modes, TRT≡BGK at Λ=(1/s−½)², rest-state
// fixed point, exact Couette, Poiseuille and Taylor–Green decay, Ghia Re=100 and the
// reference's Ra=1e3 Nusselt number.  Equality of each generic descriptor-driven
// step with the restated reference kernel chain is checked bit for bit.
// Parity unpinned: equality with a run of the Julia reference itself (Julia is absent
// here); each restated function is pinned by the checks above instead.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/kraken_mi355x.h"

namespace ko {

// ----------------------------------------------------------------------------------
// Lattice constants — src/lattice/d2q9.jl:17-29, src/lattice/d3q19.jl:13-47
// (0-based here: q_here = q_reference - 1).
// ----------------------------------------------------------------------------------
static const double W9[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                             1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
static const int CX9[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY9[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP9[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

static const double W19[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                               1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                               1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                               1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
static const int CX19[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY19[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ19[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int OPP19[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

// 6 w_q as written in the reference's Ladd term (stream_periodic_3d.jl:126-127:
// T(1/6) for the w = 1/36 edge populations); axis populations use 6/18 and 6/9.
static double six_w(double w) {
  if (w == 1.0 / 36.0) return 1.0 / 6.0;
  if (w == 1.0 / 18.0) return 1.0 / 3.0;
  if (w == 1.0 / 9.0) return 2.0 / 3.0;
  return 2.0;  // rest population never crosses a face
}

// feq as written in src/kernels/equilibrium_helpers.jl:8-50 and
// equilibrium_helpers_3d.jl:26-100: w ρ (1 + 3cu + 4.5cu·cu − 1.5usq), left to right.
template <class T>
static inline T feq(double w, T rho, T cu, T usq) {
  return T(w) * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq);
}

// c·u in the reference's written form: first non-zero component, then ± the others
// (e.g. `-ux + uy`, `ux - uz`).
template <class T>
static inline T cdotu(int cx, int cy, int cz, T ux, T uy, T uz) {
  T acc = T(0);
  bool first = true;
  auto add = [&](int c, T u) {
    if (c == 0) return;
    if (first) {
      acc = (c > 0) ? u : -u;
      first = false;
    } else {
      acc = (c > 0) ? acc + u : acc - u;
    }
  };
  add(cx, ux);
  add(cy, uy);
  add(cz, uz);
  return acc;
}

static inline size_t I2(int64_t i, int64_t j, int64_t nx) { return size_t(i + nx * j); }
static inline size_t I3(int64_t i, int64_t j, int64_t k, int64_t nx, int64_t ny) {
  return size_t(i + nx * (j + ny * k));
}

// ----------------------------------------------------------------------------------
// D2Q9 moments — equilibrium_helpers.jl:54-61 / macroscopic.jl:5-18
// ----------------------------------------------------------------------------------
template <class T>
static inline void moments9(const T* f, T& rho, T& ux, T& uy) {
  rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  T inv = T(1) / rho;
  ux = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) * inv;
  uy = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) * inv;
}

// D3Q19 moments — equilibrium_helpers_3d.jl:102-112 / macroscopic.jl:42-61
template <class T>
static inline void moments19(const T* f, T& rho, T& ux, T& uy, T& uz) {
  rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] +
        f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
  T inv = T(1) / rho;
  ux = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14]) * inv;
  uy = (f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18]) * inv;
  uz = (f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18]) * inv;
}

// ----------------------------------------------------------------------------------
// Per-cell collisions (operate on a register copy fp[Q], write fo[Q])
// ----------------------------------------------------------------------------------
// BGK D2Q9 — collide_stream_2d.jl:105-131 (and fused_bgk_2d.jl:38-56)
template <class T>
static void bgk9(const T* fp, T* fo, T omega, T& rho, T& ux, T& uy) {
  moments9(fp, rho, ux, uy);
  T usq = ux * ux + uy * uy;
  for (int q = 0; q < 9; ++q) {
    T cu = cdotu<T>(CX9[q], CY9[q], 0, ux, uy, T(0));
    T fe = (q == 0) ? T(W9[0]) * rho * (T(1) - T(1.5) * usq) : feq<T>(W9[q], rho, cu, usq);
    fo[q] = fp[q] - omega * (fp[q] - fe);
  }
}

// TRT D2Q9 — fused_trt_2d.jl:137-170: f* = f − a(f−feq) − b(f̄−feq̄)
template <class T>
static void trt9(const T* fp, T* fo, T sp, T sm, T& rho, T& ux, T& uy) {
  moments9(fp, rho, ux, uy);
  T usq = ux * ux + uy * uy;
  T fe[9];
  fe[0] = T(W9[0]) * rho * (T(1) - T(1.5) * usq);
  for (int q = 1; q < 9; ++q)
    fe[q] = feq<T>(W9[q], rho, cdotu<T>(CX9[q], CY9[q], 0, ux, uy, T(0)), usq);
  T a = (sp + sm) * T(0.5);
  T b = (sp - sm) * T(0.5);
  fo[0] = fp[0] - sp * (fp[0] - fe[0]);
  for (int q = 1; q < 9; ++q) {
    int o = OPP9[q];
    fo[q] = fp[q] - a * (fp[q] - fe[q]) - b * (fp[o] - fe[o]);
  }
}

// MRT D2Q9 (Lallemand–Luo) — collide_mrt_2d.jl:56-107, hand-coded M and M^-1
template <class T>
static void mrt9(const T* fp, T* fo, T s_e, T s_eps, T s_q, T s_nu) {
  T f1 = fp[0], f2 = fp[1], f3 = fp[2], f4 = fp[3], f5 = fp[4], f6 = fp[5], f7 = fp[6],
    f8 = fp[7], f9 = fp[8];
  T rho = f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8 + f9;
  T e = -T(4) * f1 - f2 - f3 - f4 - f5 + T(2) * (f6 + f7 + f8 + f9);
  T eps = T(4) * f1 - T(2) * (f2 + f3 + f4 + f5) + f6 + f7 + f8 + f9;
  T jx = f2 - f4 + f6 - f7 - f8 + f9;
  T qx = -T(2) * f2 + T(2) * f4 + f6 - f7 - f8 + f9;
  T jy = f3 - f5 + f6 + f7 - f8 - f9;
  T qy = -T(2) * f3 + T(2) * f5 + f6 + f7 - f8 - f9;
  T pxx = f2 - f3 + f4 - f5;
  T pxy = f6 - f7 + f8 - f9;
  T inv = T(1) / rho;
  T ux = jx * inv;
  T uy = jy * inv;
  T usq = ux * ux + uy * uy;
  T e_eq = -T(2) * rho + T(3) * rho * usq;
  T eps_eq = rho - T(3) * rho * usq;
  T qx_eq = -rho * ux;
  T qy_eq = -rho * uy;
  T pxx_eq = rho * (ux * ux - uy * uy);
  T pxy_eq = rho * ux * uy;
  T es = e - s_e * (e - e_eq);
  T ep = eps - s_eps * (eps - eps_eq);
  T qxs = qx - s_q * (qx - qx_eq);
  T qys = qy - s_q * (qy - qy_eq);
  T ps = pxx - s_nu * (pxx - pxx_eq);
  T pxys = pxy - s_nu * (pxy - pxy_eq);
  T r = rho, jxs = jx, jys = jy;
  fo[0] = T(1.0 / 9) * r - T(1.0 / 9) * es + T(1.0 / 9) * ep;
  fo[1] = T(1.0 / 9) * r - T(1.0 / 36) * es - T(1.0 / 18) * ep + T(1.0 / 6) * jxs -
          T(1.0 / 6) * qxs + T(1.0 / 4) * ps;
  fo[2] = T(1.0 / 9) * r - T(1.0 / 36) * es - T(1.0 / 18) * ep + T(1.0 / 6) * jys -
          T(1.0 / 6) * qys - T(1.0 / 4) * ps;
  fo[3] = T(1.0 / 9) * r - T(1.0 / 36) * es - T(1.0 / 18) * ep - T(1.0 / 6) * jxs +
          T(1.0 / 6) * qxs + T(1.0 / 4) * ps;
  fo[4] = T(1.0 / 9) * r - T(1.0 / 36) * es - T(1.0 / 18) * ep - T(1.0 / 6) * jys +
          T(1.0 / 6) * qys - T(1.0 / 4) * ps;
  fo[5] = T(1.0 / 9) * r + T(1.0 / 18) * es + T(1.0 / 36) * ep + T(1.0 / 6) * jxs +
          T(1.0 / 12) * qxs + T(1.0 / 6) * jys + T(1.0 / 12) * qys + T(1.0 / 4) * pxys;
  fo[6] = T(1.0 / 9) * r + T(1.0 / 18) * es + T(1.0 / 36) * ep - T(1.0 / 6) * jxs -
          T(1.0 / 12) * qxs + T(1.0 / 6) * jys + T(1.0 / 12) * qys - T(1.0 / 4) * pxys;
  fo[7] = T(1.0 / 9) * r + T(1.0 / 18) * es + T(1.0 / 36) * ep - T(1.0 / 6) * jxs -
          T(1.0 / 12) * qxs - T(1.0 / 6) * jys - T(1.0 / 12) * qys + T(1.0 / 4) * pxys;
  fo[8] = T(1.0 / 9) * r + T(1.0 / 18) * es + T(1.0 / 36) * ep + T(1.0 / 6) * jxs +
          T(1.0 / 12) * qxs - T(1.0 / 6) * jys - T(1.0 / 12) * qys - T(1.0 / 4) * pxys;
}

// BGK D3Q19 — collide_stream_3d.jl:63-95 (same order as feq_3d)
template <class T>
static void bgk19(const T* fp, T* fo, T omega, T& rho, T& ux, T& uy, T& uz) {
  moments19(fp, rho, ux, uy, uz);
  T usq = ux * ux + uy * uy + uz * uz;
  for (int q = 0; q < 19; ++q) {
    T fe = (q == 0) ? T(W19[0]) * rho * (T(1) - T(1.5) * usq)
                    : feq<T>(W19[q], rho, cdotu<T>(CX19[q], CY19[q], CZ19[q], ux, uy, uz), usq);
    fo[q] = fp[q] - omega * (fp[q] - fe);
  }
}

// TRT D3Q19 — dsl/bricks_3d.jl:168-229 (Moments_3D + CollideTRTDirect_3D)
template <class T>
static void trt19(const T* fp, T* fo, T sp, T sm, T& rho, T& ux, T& uy, T& uz) {
  moments19(fp, rho, ux, uy, uz);
  T usq = ux * ux + uy * uy + uz * uz;
  T fe[19];
  fe[0] = T(W19[0]) * rho * (T(1) - T(1.5) * usq);
  for (int q = 1; q < 19; ++q)
    fe[q] = feq<T>(W19[q], rho, cdotu<T>(CX19[q], CY19[q], CZ19[q], ux, uy, uz), usq);
  T a = (sp + sm) * T(0.5);
  T b = (sp - sm) * T(0.5);
  fo[0] = fp[0] - sp * (fp[0] - fe[0]);
  for (int q = 1; q < 19; ++q) {
    int o = OPP19[q];
    fo[q] = fp[q] - a * (fp[q] - fe[q]) - b * (fp[o] - fe[o]);
  }
}

// Thermal BGK on g — thermal_2d.jl:16-60 / fused_thermal_2d.jl:77-112:
// geq = w T (1 + 3 c·u), no u² term.
template <class T>
static void thermal9(const T* gp, T* go, T ux, T uy, T omT) {
  T Temp = gp[0] + gp[1] + gp[2] + gp[3] + gp[4] + gp[5] + gp[6] + gp[7] + gp[8];
  T t3 = T(3);
  T geq;
  geq = T(4.0 / 9.0) * Temp;
  go[0] = gp[0] - omT * (gp[0] - geq);
  geq = T(1.0 / 9.0) * Temp * (T(1) + t3 * ux);
  go[1] = gp[1] - omT * (gp[1] - geq);
  geq = T(1.0 / 9.0) * Temp * (T(1) + t3 * uy);
  go[2] = gp[2] - omT * (gp[2] - geq);
  geq = T(1.0 / 9.0) * Temp * (T(1) - t3 * ux);
  go[3] = gp[3] - omT * (gp[3] - geq);
  geq = T(1.0 / 9.0) * Temp * (T(1) - t3 * uy);
  go[4] = gp[4] - omT * (gp[4] - geq);
  geq = T(1.0 / 36.0) * Temp * (T(1) + t3 * (ux + uy));
  go[5] = gp[5] - omT * (gp[5] - geq);
  geq = T(1.0 / 36.0) * Temp * (T(1) + t3 * (-ux + uy));
  go[6] = gp[6] - omT * (gp[6] - geq);
  geq = T(1.0 / 36.0) * Temp * (T(1) + t3 * (-ux - uy));
  go[7] = gp[7] - omT * (gp[7] - geq);
  geq = T(1.0 / 36.0) * Temp * (T(1) + t3 * (ux - uy));
  go[8] = gp[8] - omT * (gp[8] - geq);
}

// BGK + Guo Boussinesq forcing — thermal_2d.jl:421-480 / fused_thermal_2d.jl:116-157.
// ρ, ux, uy are the force-shifted moments computed by the caller.
template <class T>
static void guo9(const T* fp, T* fo, T rho, T ux, T uy, T fy, T om) {
  T fx = T(0);
  T usq = ux * ux + uy * uy;
  T guo = T(1) - om / T(2);
  T t3 = T(3);
  T Sq;
  Sq = T(4.0 / 9.0) * ((-ux) * fx + (-uy) * fy) * t3;
  fo[0] = fp[0] - om * (fp[0] - T(4.0 / 9.0) * rho * (T(1) - T(1.5) * usq)) + guo * Sq;
  Sq = T(1.0 / 9.0) * ((T(1) - ux) * fx + (-uy) * fy) * t3 + T(1.0 / 9.0) * ux * fx * T(9);
  fo[1] = fp[1] - om * (fp[1] - feq<T>(1.0 / 9.0, rho, ux, usq)) + guo * Sq;
  Sq = T(1.0 / 9.0) * ((-ux) * fx + (T(1) - uy) * fy) * t3 + T(1.0 / 9.0) * uy * fy * T(9);
  fo[2] = fp[2] - om * (fp[2] - feq<T>(1.0 / 9.0, rho, uy, usq)) + guo * Sq;
  Sq = T(1.0 / 9.0) * ((-T(1) - ux) * fx + (-uy) * fy) * t3 + T(1.0 / 9.0) * ux * fx * T(9);
  fo[3] = fp[3] - om * (fp[3] - feq<T>(1.0 / 9.0, rho, -ux, usq)) + guo * Sq;
  Sq = T(1.0 / 9.0) * ((-ux) * fx + (-T(1) - uy) * fy) * t3 + T(1.0 / 9.0) * uy * fy * T(9);
  fo[4] = fp[4] - om * (fp[4] - feq<T>(1.0 / 9.0, rho, -uy, usq)) + guo * Sq;
  Sq = T(1.0 / 36.0) * ((T(1) - ux) * fx + (T(1) - uy) * fy) * t3 +
       T(1.0 / 36.0) * (ux + uy) * (fx + fy) * T(9);
  fo[5] = fp[5] - om * (fp[5] - feq<T>(1.0 / 36.0, rho, ux + uy, usq)) + guo * Sq;
  Sq = T(1.0 / 36.0) * ((-T(1) - ux) * fx + (T(1) - uy) * fy) * t3 +
       T(1.0 / 36.0) * (-ux + uy) * (-fx + fy) * T(9);
  fo[6] = fp[6] - om * (fp[6] - feq<T>(1.0 / 36.0, rho, -ux + uy, usq)) + guo * Sq;
  Sq = T(1.0 / 36.0) * ((-T(1) - ux) * fx + (-T(1) - uy) * fy) * t3 +
       T(1.0 / 36.0) * (-ux - uy) * (-fx - fy) * T(9);
  fo[7] = fp[7] - om * (fp[7] - feq<T>(1.0 / 36.0, rho, -ux - uy, usq)) + guo * Sq;
  Sq = T(1.0 / 36.0) * ((T(1) - ux) * fx + (-T(1) - uy) * fy) * t3 +
       T(1.0 / 36.0) * (ux - uy) * (fx - fy) * T(9);
  fo[8] = fp[8] - om * (fp[8] - feq<T>(1.0 / 36.0, rho, ux - uy, usq)) + guo * Sq;
}

// collide_guo_2d_kernel! per-cell body (collide_guo_2d.jl:21-70), constant force (fx, fy);
// written with the reference's (−u)(−F) factors for the c = −e_x, −e_y populations.
template <class T>
static void guo9_force(const T* fp, T* fo, T rho, T ux, T uy, T fx, T fy, T om) {
  T usq = ux * ux + uy * uy;
  T pref = T(1) - om / T(2);
  T Sq;
  Sq = T(4.0 / 9.0) * ((-ux) * fx + (-uy) * fy) * T(3);
  fo[0] = fp[0] - om * (fp[0] - T(4.0 / 9.0) * rho * (T(1) - T(1.5) * usq)) + pref * Sq;
  Sq = T(1.0 / 9.0) * ((T(1) - ux) * fx + (-uy) * fy) * T(3) + T(1.0 / 9.0) * (ux) * (fx) * T(9);
  fo[1] = fp[1] - om * (fp[1] - feq<T>(1.0 / 9.0, rho, ux, usq)) + pref * Sq;
  Sq = T(1.0 / 9.0) * ((-ux) * fx + (T(1) - uy) * fy) * T(3) + T(1.0 / 9.0) * (uy) * (fy) * T(9);
  fo[2] = fp[2] - om * (fp[2] - feq<T>(1.0 / 9.0, rho, uy, usq)) + pref * Sq;
  Sq = T(1.0 / 9.0) * ((-T(1) - ux) * fx + (-uy) * fy) * T(3) + T(1.0 / 9.0) * (-ux) * (-fx) * T(9);
  fo[3] = fp[3] - om * (fp[3] - feq<T>(1.0 / 9.0, rho, -ux, usq)) + pref * Sq;
  Sq = T(1.0 / 9.0) * ((-ux) * fx + (-T(1) - uy) * fy) * T(3) + T(1.0 / 9.0) * (-uy) * (-fy) * T(9);
  fo[4] = fp[4] - om * (fp[4] - feq<T>(1.0 / 9.0, rho, -uy, usq)) + pref * Sq;
  Sq = T(1.0 / 36.0) * ((T(1) - ux) * fx + (T(1) - uy) * fy) * T(3) +
       T(1.0 / 36.0) * (ux + uy) * (fx + fy) * T(9);
  fo[5] = fp[5] - om * (fp[5] - feq<T>(1.0 / 36.0, rho, ux + uy, usq)) + pref * Sq;
  Sq = T(1.0 / 36.0) * ((-T(1) - ux) * fx + (T(1) - uy) * fy) * T(3) +
       T(1.0 / 36.0) * (-ux + uy) * (-fx + fy) * T(9);
  fo[6] = fp[6] - om * (fp[6] - feq<T>(1.0 / 36.0, rho, -ux + uy, usq)) + pref * Sq;
  Sq = T(1.0 / 36.0) * ((-T(1) - ux) * fx + (-T(1) - uy) * fy) * T(3) +
       T(1.0 / 36.0) * (-ux - uy) * (-fx - fy) * T(9);
  fo[7] = fp[7] - om * (fp[7] - feq<T>(1.0 / 36.0, rho, -ux - uy, usq)) + pref * Sq;
  Sq = T(1.0 / 36.0) * ((T(1) - ux) * fx + (-T(1) - uy) * fy) * T(3) +
       T(1.0 / 36.0) * (ux - uy) * (fx - fy) * T(9);
  fo[8] = fp[8] - om * (fp[8] - feq<T>(1.0 / 36.0, rho, ux - uy, usq)) + pref * Sq;
}

// Thermal BGK on g, D3Q19 — collide_thermal_3d_kernel! (thermal_3d.jl:12-107):
// T = g1 + … + g19 left to right, geq = w T (1 + 3 c·u) with c·u written as the reference
// does (`u_x`, `-u_x + u_y`, …; 1 + 3(−u) equals the written 1 − 3u exactly).
template <class T>
static void thermal19(const T* gp, T* go, T ux, T uy, T uz, T omT) {
  T Temp = gp[0];
  for (int q = 1; q < 19; ++q) Temp = Temp + gp[q];
  T t3 = T(3);
  T geq = T(1.0 / 3.0) * Temp;
  go[0] = gp[0] - omT * (gp[0] - geq);
  for (int q = 1; q < 19; ++q) {
    geq = T(W19[q]) * Temp * (T(1) + t3 * cdotu<T>(CX19[q], CY19[q], CZ19[q], ux, uy, uz));
    go[q] = gp[q] - omT * (gp[q] - geq);
  }
}

// (c_a − u_a) as the reference writes it: `(-ux)`, `(one(T)-ux)`, `(-one(T)-ux)`.
template <class T>
static inline T c_minus_u(int c, T u) {
  return c == 0 ? -u : (c > 0 ? T(1) - u : -T(1) - u);
}

// BGK + Guo source, D3Q19, one cell — the common body of collide_guo_3d_kernel!
// (collide_guo_3d.jl:21-185), collide_guo_field_3d_kernel! (:213-320) and
// collide_boussinesq_3d_kernel! (thermal_3d.jl:574-712, F = (0, fy, 0)):
//   S_q = w ((c−u)·F) 3 + w (c·u)(c·F) 9,  f* = f − ω(f − feq) + (1 − ω/2) S_q,
// with ρ, u the force-shifted moments of the caller.
template <class T>
static void guo19(const T* fp, T* fo, T rho, T ux, T uy, T uz, T fx, T fy, T fz, T om) {
  T usq = ux * ux + uy * uy + uz * uz;
  T pref = T(1) - om / T(2);
  T t3 = T(3), t9 = T(9);
  T fe = T(W19[0]) * rho * (T(1) - T(1.5) * usq);
  T Sq = T(W19[0]) * ((-ux) * fx + (-uy) * fy + (-uz) * fz) * t3;
  fo[0] = fp[0] - om * (fp[0] - fe) + pref * Sq;
  for (int q = 1; q < 19; ++q) {
    const int cx = CX19[q], cy = CY19[q], cz = CZ19[q];
    T cu = cdotu<T>(cx, cy, cz, ux, uy, uz);
    T cf = cdotu<T>(cx, cy, cz, fx, fy, fz);
    fe = feq<T>(W19[q], rho, cu, usq);
    Sq = T(W19[q]) * (c_minus_u(cx, ux) * fx + c_minus_u(cy, uy) * fy + c_minus_u(cz, uz) * fz) * t3 +
         T(W19[q]) * cu * cf * t9;
    fo[q] = fp[q] - om * (fp[q] - fe) + pref * Sq;
  }
}

// ----------------------------------------------------------------------------------
// Reference kernels restated one for one (whole-array operations)
// ----------------------------------------------------------------------------------

// stream_2d_kernel! — collide_stream_2d.jl:75-101: pull, halfway BB at all four edges.
template <class T>
static void stream_2d(T* fo, const T* fi, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      for (int q = 0; q < 9; ++q) {
        int64_t si = i - CX9[q], sj = j - CY9[q];
        bool ok = si >= 0 && si < nx && sj >= 0 && sj < ny;
        fo[c + q * P] = ok ? fi[I2(si, sj, nx) + q * P] : fi[c + OPP9[q] * P];
      }
    }
}

// stream_fully_periodic_2d_kernel! — stream_periodic_2d.jl:79-98
template <class T>
static void stream_fully_periodic_2d(T* fo, const T* fi, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i)
      for (int q = 0; q < 9; ++q) {
        int64_t si = (i - CX9[q] + nx) % nx, sj = (j - CY9[q] + ny) % ny;
        fo[I2(i, j, nx) + q * P] = fi[I2(si, sj, nx) + q * P];
      }
}

// stream_periodic_x_wall_y_2d_kernel! — stream_periodic_2d.jl:5-55
template <class T>
static void stream_periodic_x_wall_y_2d(T* fo, const T* fi, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      for (int q = 0; q < 9; ++q) {
        int64_t si = (i - CX9[q] + nx) % nx, sj = j - CY9[q];
        bool ok = sj >= 0 && sj < ny;
        fo[c + q * P] = ok ? fi[I2(si, sj, nx) + q * P] : fi[c + OPP9[q] * P];
      }
    }
}

// collide_2d_kernel! — collide_stream_2d.jl:105-131 (solid: swap pairs in place)
template <class T>
static void collide_2d(T* f, const uint8_t* solid, int64_t nx, int64_t ny, T omega) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      T fp[9], fo[9], r, ux, uy;
      for (int q = 0; q < 9; ++q) fp[q] = f[c + q * P];
      if (solid && solid[c]) {
        for (int q = 0; q < 9; ++q) fo[q] = fp[OPP9[q]];
      } else {
        bgk9(fp, fo, omega, r, ux, uy);
      }
      for (int q = 0; q < 9; ++q) f[c + q * P] = fo[q];
    }
}

// In-place TRT collide with the arithmetic of fused_trt_step_kernel! (fused_trt_2d.jl:125-171);
// the fused kernel equals stream_2d followed by this (checked bit for bit in the tests).
template <class T>
static void collide_trt_2d(T* f, const uint8_t* solid, int64_t nx, int64_t ny, T sp, T sm) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      T fp[9], fo[9], r, ux, uy;
      for (int q = 0; q < 9; ++q) fp[q] = f[c + q * P];
      if (solid && solid[c]) {
        for (int q = 0; q < 9; ++q) fo[q] = fp[OPP9[q]];
      } else {
        trt9(fp, fo, sp, sm, r, ux, uy);
      }
      for (int q = 0; q < 9; ++q) f[c + q * P] = fo[q];
    }
}

// collide_mrt_2d_kernel! — collide_mrt_2d.jl:43-109
template <class T>
static void collide_mrt_2d(T* f, const uint8_t* solid, int64_t nx, int64_t ny, T s_e, T s_eps,
                           T s_q, T s_nu) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      T fp[9], fo[9];
      for (int q = 0; q < 9; ++q) fp[q] = f[c + q * P];
      if (solid && solid[c]) {
        for (int q = 0; q < 9; ++q) fo[q] = fp[OPP9[q]];
      } else {
        mrt9(fp, fo, s_e, s_eps, s_q, s_nu);
      }
      for (int q = 0; q < 9; ++q) f[c + q * P] = fo[q];
    }
}

// fused_bgk_step_kernel! — fused_bgk_2d.jl:8-59 / fused_trt_step_kernel! — fused_trt_2d.jl:101-173
template <class T>
static void fused_2d(T* fo, const T* fi, T* rho, T* ux, T* uy, const uint8_t* solid, int64_t nx,
                     int64_t ny, bool trt, T a1, T a2) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      T fp[9], o[9], r, u, v;
      for (int q = 0; q < 9; ++q) {
        int64_t si = i - CX9[q], sj = j - CY9[q];
        bool ok = si >= 0 && si < nx && sj >= 0 && sj < ny;
        fp[q] = ok ? fi[I2(si, sj, nx) + q * P] : fi[c + OPP9[q] * P];
      }
      if (solid && solid[c]) {
        for (int q = 0; q < 9; ++q) o[q] = fp[OPP9[q]];
        r = T(1);
        u = T(0);
        v = T(0);
      } else if (trt) {
        trt9(fp, o, a1, a2, r, u, v);
      } else {
        bgk9(fp, o, a1, r, u, v);
      }
      for (int q = 0; q < 9; ++q) fo[c + q * P] = o[q];
      if (rho) rho[c] = r;
      if (ux) ux[c] = u;
      if (uy) uy[c] = v;
    }
}

// compute_macroscopic_2d_kernel! — macroscopic.jl:5-18
template <class T>
static void macroscopic_2d(T* rho, T* ux, T* uy, const T* f, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      T fp[9], r, u, v;
      for (int q = 0; q < 9; ++q) fp[q] = f[c + q * P];
      moments9(fp, r, u, v);
      if (rho) rho[c] = r;
      if (ux) ux[c] = u;
      if (uy) uy[c] = v;
    }
}

// zou_he_velocity_north_2d_kernel! — boundary_2d.jl:13-44 (uy_wall = 0)
template <class T>
static void zou_he_north(T* f, T u, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t i = 0; i < nx; ++i) {
    size_t c = I2(i, ny - 1, nx);
    T f1 = f[c], f2 = f[c + P], f3 = f[c + 2 * P], f4 = f[c + 3 * P], f6 = f[c + 5 * P],
      f7 = f[c + 6 * P];
    T rw = (f1 + f2 + f4 + T(2) * (f3 + f6 + f7)) / (T(1) + u * T(0));
    f[c + 4 * P] = f3;
    f[c + 8 * P] = f7 - T(0.5) * (f2 - f4) + T(0.5) * rw * u;
    f[c + 7 * P] = f6 + T(0.5) * (f2 - f4) - T(0.5) * rw * u;
  }
}

// zou_he_velocity_south_2d_kernel! — boundary_2d.jl:135-150
template <class T>
static void zou_he_south(T* f, T u, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t i = 0; i < nx; ++i) {
    size_t c = I2(i, 0, nx);
    T f1 = f[c], f2 = f[c + P], f4 = f[c + 3 * P], f5 = f[c + 4 * P], f8 = f[c + 7 * P],
      f9 = f[c + 8 * P];
    T rw = f1 + f2 + f4 + T(2) * (f5 + f8 + f9);
    f[c + 2 * P] = f5;
    f[c + 5 * P] = f8 - T(0.5) * (f2 - f4) + T(0.5) * rw * u;
    f[c + 6 * P] = f9 + T(0.5) * (f2 - f4) - T(0.5) * rw * u;
  }
}

// zou_he_velocity_west_2d_kernel! — boundary_2d.jl:171-185
template <class T>
static void zou_he_west(T* f, T u, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t j = 0; j < ny; ++j) {
    size_t c = I2(0, j, nx);
    T f1 = f[c], f3 = f[c + 2 * P], f4 = f[c + 3 * P], f5 = f[c + 4 * P], f7 = f[c + 6 * P],
      f8 = f[c + 7 * P];
    T rw = (f1 + f3 + f5 + T(2) * (f4 + f7 + f8)) / (T(1) - u);
    f[c + P] = f4 + T(2.0 / 3.0) * rw * u;
    f[c + 5 * P] = f8 - T(0.5) * (f3 - f5) + T(1.0 / 6.0) * rw * u;
    f[c + 8 * P] = f7 + T(0.5) * (f3 - f5) + T(1.0 / 6.0) * rw * u;
  }
}

// zou_he_pressure_east_2d_kernel! — boundary_2d.jl:206-223
template <class T>
static void zou_he_pressure_east(T* f, T rho_out, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t j = 0; j < ny; ++j) {
    size_t c = I2(nx - 1, j, nx);
    T f1 = f[c], f2 = f[c + P], f3 = f[c + 2 * P], f5 = f[c + 4 * P], f6 = f[c + 5 * P],
      f9 = f[c + 8 * P];
    T ux = -T(1) + (f1 + f3 + f5 + T(2) * (f2 + f6 + f9)) / rho_out;
    f[c + 3 * P] = f2 - T(2.0 / 3.0) * rho_out * ux;
    f[c + 6 * P] = f9 - T(0.5) * (f3 - f5) - T(1.0 / 6.0) * rho_out * ux;
    f[c + 7 * P] = f6 + T(0.5) * (f3 - f5) - T(1.0 / 6.0) * rho_out * ux;
  }
}

// zou_he_pressure_west_2d_kernel! — boundary_2d.jl:264-281
template <class T>
static void zou_he_pressure_west(T* f, T rho_in, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t j = 0; j < ny; ++j) {
    size_t c = I2(0, j, nx);
    T f1 = f[c], f3 = f[c + 2 * P], f4 = f[c + 3 * P], f5 = f[c + 4 * P], f7 = f[c + 6 * P],
      f8 = f[c + 7 * P];
    T ux = T(1) - (f1 + f3 + f5 + T(2) * (f4 + f7 + f8)) / rho_in;
    f[c + P] = f4 + T(2.0 / 3.0) * rho_in * ux;
    f[c + 5 * P] = f8 - T(0.5) * (f3 - f5) + T(1.0 / 6.0) * rho_in * ux;
    f[c + 8 * P] = f7 + T(0.5) * (f3 - f5) + T(1.0 / 6.0) * rho_in * ux;
  }
}

// Spatial Zou–He closures, per-node velocity / density arrays, corner nodes skipped
// (boundary_spatial_2d.jl).  f1..f9 of the reference are f[c + (q−1)P].
// zou_he_velocity_north_spatial_2d_kernel! — boundary_spatial_2d.jl:6-24
template <class T>
static void zou_he_north_spatial(T* f, const double* uxa, const double* uya, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t i = 1; i < nx - 1; ++i) {
    size_t c = I2(i, ny - 1, nx);
    T u_x = T(uxa[i]), u_y = T(uya[i]);
    T f1 = f[c], f2 = f[c + P], f3 = f[c + 2 * P], f4 = f[c + 3 * P], f6 = f[c + 5 * P],
      f7 = f[c + 6 * P];
    T rho_wall = (f1 + f2 + f4 + T(2) * (f3 + f6 + f7)) / (T(1) + u_y);
    f[c + 4 * P] = f3 - T(2.0 / 3.0) * rho_wall * u_y;
    f[c + 8 * P] = f7 - T(0.5) * (f2 - f4) + T(0.5) * rho_wall * u_x - T(1.0 / 6.0) * rho_wall * u_y;
    f[c + 7 * P] = f6 + T(0.5) * (f2 - f4) - T(0.5) * rho_wall * u_x - T(1.0 / 6.0) * rho_wall * u_y;
  }
}

// zou_he_velocity_south_spatial_2d_kernel! — boundary_spatial_2d.jl:40-57
template <class T>
static void zou_he_south_spatial(T* f, const double* uxa, const double* uya, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t i = 1; i < nx - 1; ++i) {
    size_t c = I2(i, 0, nx);
    T u_x = T(uxa[i]), u_y = T(uya[i]);
    T f1 = f[c], f2 = f[c + P], f4 = f[c + 3 * P], f5 = f[c + 4 * P], f8 = f[c + 7 * P],
      f9 = f[c + 8 * P];
    T rho_wall = (f1 + f2 + f4 + T(2) * (f5 + f8 + f9)) / (T(1) - u_y);
    f[c + 2 * P] = f5 + T(2.0 / 3.0) * rho_wall * u_y;
    f[c + 5 * P] = f8 - T(0.5) * (f2 - f4) + T(0.5) * rho_wall * u_x + T(1.0 / 6.0) * rho_wall * u_y;
    f[c + 6 * P] = f9 + T(0.5) * (f2 - f4) - T(0.5) * rho_wall * u_x + T(1.0 / 6.0) * rho_wall * u_y;
  }
  (void)ny;
}

// zou_he_velocity_west_spatial_2d_kernel! — boundary_spatial_2d.jl:72-89
template <class T>
static void zou_he_west_spatial(T* f, const double* uxa, const double* uya, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t j = 1; j < ny - 1; ++j) {
    size_t c = I2(0, j, nx);
    T u_x = T(uxa[j]), u_y = T(uya[j]);
    T f1 = f[c], f3 = f[c + 2 * P], f4 = f[c + 3 * P], f5 = f[c + 4 * P], f7 = f[c + 6 * P],
      f8 = f[c + 7 * P];
    T rho_wall = (f1 + f3 + f5 + T(2) * (f4 + f7 + f8)) / (T(1) - u_x);
    f[c + P] = f4 + T(2.0 / 3.0) * rho_wall * u_x;
    f[c + 5 * P] = f8 - T(0.5) * (f3 - f5) + T(1.0 / 6.0) * rho_wall * u_x + T(0.5) * rho_wall * u_y;
    f[c + 8 * P] = f7 + T(0.5) * (f3 - f5) + T(1.0 / 6.0) * rho_wall * u_x - T(0.5) * rho_wall * u_y;
  }
}

// zou_he_pressure_east_spatial_2d_kernel! — boundary_spatial_2d.jl:104-121
template <class T>
static void zou_he_pressure_east_spatial(T* f, const double* rhoa, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t j = 1; j < ny - 1; ++j) {
    size_t c = I2(nx - 1, j, nx);
    T rho_out = T(rhoa[j]);
    T f1 = f[c], f2 = f[c + P], f3 = f[c + 2 * P], f5 = f[c + 4 * P], f6 = f[c + 5 * P],
      f9 = f[c + 8 * P];
    T ux = -T(1) + (f1 + f3 + f5 + T(2) * (f2 + f6 + f9)) / rho_out;
    f[c + 3 * P] = f2 - T(2.0 / 3.0) * rho_out * ux;
    f[c + 6 * P] = f9 - T(0.5) * (f3 - f5) - T(1.0 / 6.0) * rho_out * ux;
    f[c + 7 * P] = f6 + T(0.5) * (f3 - f5) - T(1.0 / 6.0) * rho_out * ux;
  }
}

// ZouHeFace3D descriptors (boundary_3d.jl:56-137), kept 1-based as the reference writes
// them; rows in the face order W, E, S, N, B, T.
struct ZHFace3 {
  int axis, sign, qa, qao, qd[4], qdo[4], t1p, t1m, t2p, t2m, par[9], out[5];
};
static const ZHFace3 ZH3_TABLE[6] = {
    {1, -1, 2, 3, {8, 10, 12, 14}, {11, 9, 15, 13}, 4, 5, 6, 7,
     {1, 4, 5, 6, 7, 16, 17, 18, 19}, {3, 9, 11, 13, 15}},
    {1, +1, 3, 2, {9, 11, 13, 15}, {10, 8, 14, 12}, 4, 5, 6, 7,
     {1, 4, 5, 6, 7, 16, 17, 18, 19}, {2, 8, 10, 12, 14}},
    {2, -1, 4, 5, {8, 9, 16, 18}, {11, 10, 19, 17}, 2, 3, 6, 7,
     {1, 2, 3, 6, 7, 12, 13, 14, 15}, {5, 10, 11, 17, 19}},
    {2, +1, 5, 4, {10, 11, 17, 19}, {9, 8, 18, 16}, 2, 3, 6, 7,
     {1, 2, 3, 6, 7, 12, 13, 14, 15}, {4, 8, 9, 16, 18}},
    {3, -1, 6, 7, {12, 13, 16, 17}, {15, 14, 19, 18}, 2, 3, 4, 5,
     {1, 2, 3, 4, 5, 8, 9, 10, 11}, {7, 14, 15, 18, 19}},
    {3, +1, 7, 6, {14, 15, 18, 19}, {13, 12, 17, 16}, 2, 3, 4, 5,
     {1, 2, 3, 4, 5, 8, 9, 10, 11}, {6, 12, 13, 16, 17}}};

// Node list of a face: the fixed index on the normal axis, every node of the other two.
template <class Fn>
static void for_face_nodes(int face, int64_t nx, int64_t ny, int64_t nz, Fn fn) {
  const ZHFace3& Z = ZH3_TABLE[face];
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t x[3] = {i, j, k}, n[3] = {nx, ny, nz};
        const int64_t fixed = Z.sign < 0 ? 0 : n[Z.axis - 1] - 1;
        if (x[Z.axis - 1] == fixed) fn(I3(i, j, k, nx, ny));
      }
}

// zou_he_velocity_3d_kernel! — boundary_3d.jl:152-232.  un, ut1, ut2: the wall velocity
// along the normal axis and the face's two tangents (the _apply_zou_he_velocity_3d! args).
template <class T>
static void zou_he_velocity_3d(T* f, int face, T un, T ut1, T ut2, int64_t nx, int64_t ny,
                               int64_t nz) {
  const ZHFace3& Z = ZH3_TABLE[face];
  const size_t P = size_t(nx * ny * nz);
  for_face_nodes(face, nx, ny, nz, [&](size_t c) {
    auto F = [&](int q1) -> T& { return f[c + size_t(q1 - 1) * P]; };
    T sum_par = T(0), sum_out = T(0);
    for (int n = 0; n < 9; ++n) sum_par += F(Z.par[n]);
    for (int n = 0; n < 5; ++n) sum_out += F(Z.out[n]);
    const T sg = T(Z.sign);
    const T rho_w = (sum_par + T(2) * sum_out) / (T(1) + sg * un);
    F(Z.qa) = F(Z.qao) - sg * T(1.0 / 3.0) * rho_w * un;
    const T d1 = F(Z.par[5]), d2 = F(Z.par[6]), d3 = F(Z.par[7]), d4 = F(Z.par[8]);
    const T t1 = (F(Z.t1p) - F(Z.t1m)) + (d1 - d2 + d3 - d4);
    const T t2 = (F(Z.t2p) - F(Z.t2m)) + (d1 + d2 - d3 - d4);
    F(Z.qd[0]) = F(Z.qdo[0]) - T(0.5) * t1 + T(0.5) * rho_w * ut1 - sg * T(1.0 / 6.0) * rho_w * un;
    F(Z.qd[1]) = F(Z.qdo[1]) + T(0.5) * t1 - T(0.5) * rho_w * ut1 - sg * T(1.0 / 6.0) * rho_w * un;
    F(Z.qd[2]) = F(Z.qdo[2]) - T(0.5) * t2 + T(0.5) * rho_w * ut2 - sg * T(1.0 / 6.0) * rho_w * un;
    F(Z.qd[3]) = F(Z.qdo[3]) + T(0.5) * t2 - T(0.5) * rho_w * ut2 - sg * T(1.0 / 6.0) * rho_w * un;
  });
}

// zou_he_pressure_3d_kernel! — boundary_3d.jl:234-290
template <class T>
static void zou_he_pressure_3d(T* f, int face, T rho_out, int64_t nx, int64_t ny, int64_t nz) {
  const ZHFace3& Z = ZH3_TABLE[face];
  const size_t P = size_t(nx * ny * nz);
  for_face_nodes(face, nx, ny, nz, [&](size_t c) {
    auto F = [&](int q1) -> T& { return f[c + size_t(q1 - 1) * P]; };
    T sum_par = T(0), sum_out = T(0);
    for (int n = 0; n < 9; ++n) sum_par += F(Z.par[n]);
    for (int n = 0; n < 5; ++n) sum_out += F(Z.out[n]);
    const T sg = T(Z.sign);
    const T un = -sg * (T(1) - (sum_par + T(2) * sum_out) / rho_out);
    F(Z.qa) = F(Z.qao) - sg * T(1.0 / 3.0) * rho_out * un;
    const T t1 = F(Z.t1p) - F(Z.t1m);
    const T t2 = F(Z.t2p) - F(Z.t2m);
    F(Z.qd[0]) = F(Z.qdo[0]) - T(0.5) * t1 - sg * T(1.0 / 6.0) * rho_out * un;
    F(Z.qd[1]) = F(Z.qdo[1]) + T(0.5) * t1 - sg * T(1.0 / 6.0) * rho_out * un;
    F(Z.qd[2]) = F(Z.qdo[2]) - T(0.5) * t2 - sg * T(1.0 / 6.0) * rho_out * un;
    F(Z.qd[3]) = F(Z.qdo[3]) + T(0.5) * t2 - sg * T(1.0 / 6.0) * rho_out * un;
  });
}

// stream_3d_kernel! — collide_stream_3d.jl:5-43: pull, halfway BB at all six faces.
// stream_periodic_xz_wall_y_3d_kernel! — stream_periodic_3d.jl:24-75 (px = pz = true);
// PullHalfwayBB_3D{true} — dsl/bricks_3d.jl:106-136 (pz = true only).
template <class T>
static void stream_3d_flags(T* fo, const T* fi, int64_t nx, int64_t ny, int64_t nz, bool px,
                            bool py, bool pz) {
  const size_t P = size_t(nx * ny * nz);
#pragma omp parallel for
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        size_t c = I3(i, j, k, nx, ny);
        for (int q = 0; q < 19; ++q) {
          int64_t si = i - CX19[q], sj = j - CY19[q], sk = k - CZ19[q];
          bool ok = true;
          if (px) si = (si + nx) % nx; else ok = ok && si >= 0 && si < nx;
          if (py) sj = (sj + ny) % ny; else ok = ok && sj >= 0 && sj < ny;
          if (pz) sk = (sk + nz) % nz; else ok = ok && sk >= 0 && sk < nz;
          fo[c + q * P] = ok ? fi[I3(si, sj, sk, nx, ny) + q * P] : fi[c + OPP19[q] * P];
        }
      }
}

// stream_periodic_xz_movingwall_y_3d_kernel! — stream_periodic_3d.jl:106-166
template <class T>
static void stream_periodic_xz_movingwall_y_3d(T* fo, const T* fi, T Ub, T Ut, T rho_w,
                                               int64_t nx, int64_t ny, int64_t nz) {
  const size_t P = size_t(nx * ny * nz);
  T cbot = T(1.0 / 6.0) * rho_w * Ub;
  T ctop = T(1.0 / 6.0) * rho_w * Ut;
#pragma omp parallel for
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        size_t c = I3(i, j, k, nx, ny);
        for (int q = 0; q < 19; ++q) {
          int64_t si = (i - CX19[q] + nx) % nx, sj = j - CY19[q], sk = (k - CZ19[q] + nz) % nz;
          bool ok = sj >= 0 && sj < ny;
          T v = ok ? fi[I3(si, sj, sk, nx, ny) + q * P] : fi[c + OPP19[q] * P];
          if (!ok) {
            if (q == 7) v = v + cbot;       // q8  (+x,+y) at j = 1
            else if (q == 8) v = v - cbot;  // q9  (-x,+y) at j = 1
            else if (q == 9) v = v + ctop;  // q10 (+x,-y) at j = Ny
            else if (q == 10) v = v - ctop; // q11 (-x,-y) at j = Ny
          }
          fo[c + q * P] = v;
        }
      }
}

// collide_3d_kernel! — collide_stream_3d.jl:47-97
template <class T>
static void collide_3d(T* f, const uint8_t* solid, int64_t nx, int64_t ny, int64_t nz, T omega) {
  const size_t P = size_t(nx * ny * nz);
#pragma omp parallel for
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        size_t c = I3(i, j, k, nx, ny);
        T fp[19], fo[19], r, u, v, w;
        for (int q = 0; q < 19; ++q) fp[q] = f[c + q * P];
        if (solid && solid[c]) {
          for (int q = 0; q < 19; ++q) fo[q] = fp[OPP19[q]];
        } else {
          bgk19(fp, fo, omega, r, u, v, w);
        }
        for (int q = 0; q < 19; ++q) f[c + q * P] = fo[q];
      }
}

// In-place D3Q19 TRT collide with the arithmetic of CollideTRTDirect_3D (bricks_3d.jl:178-229);
// solid cells swap pairs like collide_3d_kernel!.
template <class T>
static void collide_trt_3d(T* f, const uint8_t* solid, int64_t nx, int64_t ny, int64_t nz, T sp,
                           T sm) {
  const size_t P = size_t(nx * ny * nz);
#pragma omp parallel for
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        size_t c = I3(i, j, k, nx, ny);
        T fp[19], fo[19], r, u, v, w;
        for (int q = 0; q < 19; ++q) fp[q] = f[c + q * P];
        if (solid && solid[c]) {
          for (int q = 0; q < 19; ++q) fo[q] = fp[OPP19[q]];
        } else {
          trt19(fp, fo, sp, sm, r, u, v, w);
        }
        for (int q = 0; q < 19; ++q) f[c + q * P] = fo[q];
      }
}

// Single pass PullHalfwayBB_3D (per-axis periodic wrap as stream_periodic_3d.jl:29-33)
// + Moments_3D + CollideTRTDirect_3D + WriteMoments_3D — the DSL composition of
// dsl/bricks_3d.jl:64-229, 476-483.  Used as the CPU baseline of the D3Q19 TRT case.
template <class T>
static void fused_trt_3d(T* fo, const T* fi, int64_t nx, int64_t ny, int64_t nz, bool px, bool py,
                         bool pz, T sp, T sm) {
  const size_t P = size_t(nx * ny * nz);
#pragma omp parallel for
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        size_t c = I3(i, j, k, nx, ny);
        T fp[19], o[19], r, u, v, w;
        for (int q = 0; q < 19; ++q) {
          int64_t si = i - CX19[q], sj = j - CY19[q], sk = k - CZ19[q];
          bool ok = true;
          if (px) si = (si + nx) % nx; else ok = ok && si >= 0 && si < nx;
          if (py) sj = (sj + ny) % ny; else ok = ok && sj >= 0 && sj < ny;
          if (pz) sk = (sk + nz) % nz; else ok = ok && sk >= 0 && sk < nz;
          fp[q] = ok ? fi[I3(si, sj, sk, nx, ny) + q * P] : fi[c + OPP19[q] * P];
        }
        trt19(fp, o, sp, sm, r, u, v, w);
        for (int q = 0; q < 19; ++q) fo[c + q * P] = o[q];
      }
}

// compute_macroscopic_3d_kernel! — macroscopic.jl:42-61
template <class T>
static void macroscopic_3d(T* rho, T* ux, T* uy, T* uz, const T* f, int64_t nx, int64_t ny,
                           int64_t nz) {
  const size_t P = size_t(nx * ny * nz);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    T fp[19], r, u, v, w;
    for (int q = 0; q < 19; ++q) fp[q] = f[c + q * P];
    moments19(fp, r, u, v, w);
    if (rho) rho[c] = r;
    if (ux) ux[c] = u;
    if (uy) uy[c] = v;
    if (uz) uz[c] = w;
  }
}

// --- thermal kernels ---------------------------------------------------------------

// collide_thermal_2d_kernel! — thermal_2d.jl:16-60 (u arrays from compute_macroscopic_2d!)
template <class T>
static void collide_thermal_2d(T* g, const T* ux, const T* uy, int64_t nx, int64_t ny, T omT) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    T gp[9], go[9];
    for (int q = 0; q < 9; ++q) gp[q] = g[c + q * P];
    thermal9(gp, go, ux[c], uy[c], omT);
    for (int q = 0; q < 9; ++q) g[c + q * P] = go[q];
  }
}

// compute_temperature_2d_kernel! — thermal_2d.jl:113-119
template <class T>
static void compute_temperature_2d(T* Temp, const T* g, int64_t nx, int64_t ny) {
  const size_t P = size_t(nx * ny);
  for (int64_t c = 0; c < int64_t(P); ++c)
    Temp[c] = g[c] + g[c + P] + g[c + 2 * P] + g[c + 3 * P] + g[c + 4 * P] + g[c + 5 * P] +
              g[c + 6 * P] + g[c + 7 * P] + g[c + 8 * P];
}

// _abb_slot — thermal_2d.jl:175-178
template <class T>
static inline T abb_slot(T gq, T w2, T Tface, bool on_corner, T Tcorner, bool skip) {
  T Tlink = on_corner ? Tcorner : Tface;
  return (on_corner && skip) ? gq : -gq + w2 * Tlink;
}

// apply_fixed_temp_{south,north,west,east}_2d_kernel! — thermal_2d.jl:182-290, with corner
// temperatures Tc[0..1] and skip flags sk[0..1] as passed by _apply_fixed_temp_faces_2d!.
template <class T>
static void fixed_temp_face(T* g, int face, T Tw, const T* Tc, const bool* sk, int64_t nx,
                            int64_t ny) {
  const size_t P = size_t(nx * ny);
  T wa = T(2) * T(1.0 / 9.0);
  T we = T(2) * T(1.0 / 36.0);
  if (face == KMI_SOUTH || face == KMI_NORTH) {
    int64_t j = (face == KMI_SOUTH) ? 0 : ny - 1;
    // south: 3 / 6 (corner i=1) / 7 (corner i=Nx); north: 5 / 9 (i=1) / 8 (i=Nx)
    int qa = (face == KMI_SOUTH) ? 2 : 4, q1 = (face == KMI_SOUTH) ? 5 : 8,
        q2 = (face == KMI_SOUTH) ? 6 : 7;
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      g[c + qa * P] = -g[c + qa * P] + wa * Tw;
      g[c + q1 * P] = abb_slot(g[c + q1 * P], we, Tw, i == 0, Tc[0], sk[0]);
      g[c + q2 * P] = abb_slot(g[c + q2 * P], we, Tw, i == nx - 1, Tc[1], sk[1]);
    }
  } else {
    int64_t i = (face == KMI_WEST) ? 0 : nx - 1;
    // west: 2 / 6 (corner j=1) / 9 (corner j=Ny); east: 4 / 7 (j=1) / 8 (j=Ny)
    int qa = (face == KMI_WEST) ? 1 : 3, q1 = (face == KMI_WEST) ? 5 : 6,
        q2 = (face == KMI_WEST) ? 8 : 7;
    for (int64_t j = 0; j < ny; ++j) {
      size_t c = I2(i, j, nx);
      g[c + qa * P] = -g[c + qa * P] + wa * Tw;
      g[c + q1 * P] = abb_slot(g[c + q1 * P], we, Tw, j == 0, Tc[0], sk[0]);
      g[c + q2 * P] = abb_slot(g[c + q2 * P], we, Tw, j == ny - 1, Tc[1], sk[1]);
    }
  }
}

// apply_fixed_temp_2d! / _apply_fixed_temp_faces_2d! — thermal_2d.jl:371-417:
// corner links take the mean temperature and belong to the x-face.
template <class T>
static void apply_fixed_temp_2d(T* g, const bool* dir, const double* Tf, int64_t nx, int64_t ny) {
  auto corner = [&](int a, int b) -> T {
    return dir[b] ? (T(Tf[a]) + T(Tf[b])) / T(2) : T(Tf[a]);
  };
  const int order[4] = {KMI_WEST, KMI_EAST, KMI_SOUTH, KMI_NORTH};
  for (int f : order) {
    if (!dir[f]) continue;
    T Tc[2];
    bool sk[2];
    if (f == KMI_WEST || f == KMI_EAST) {
      Tc[0] = corner(f, KMI_SOUTH);
      Tc[1] = corner(f, KMI_NORTH);
      sk[0] = sk[1] = false;
    } else {
      Tc[0] = Tc[1] = T(Tf[f]);
      sk[0] = dir[KMI_WEST];
      sk[1] = dir[KMI_EAST];
    }
    fixed_temp_face(g, f, T(Tf[f]), Tc, sk, nx, ny);
  }
}

// collide_boussinesq_2d_kernel! — thermal_2d.jl:421-480
template <class T>
static void collide_boussinesq_2d(T* f, const T* Temp, const uint8_t* solid, int64_t nx,
                                  int64_t ny, T om, T beta_g, T T_ref) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    T fp[9], fo[9];
    for (int q = 0; q < 9; ++q) fp[q] = f[c + q * P];
    if (solid && solid[c]) {
      for (int q = 0; q < 9; ++q) fo[q] = fp[OPP9[q]];
    } else {
      T fy = beta_g * (Temp[c] - T_ref);
      T fx = T(0);
      T rho = fp[0] + fp[1] + fp[2] + fp[3] + fp[4] + fp[5] + fp[6] + fp[7] + fp[8];
      T inv = T(1) / rho;
      T ux = ((fp[1] - fp[3] + fp[5] - fp[6] - fp[7] + fp[8]) + fx / T(2)) * inv;
      T uy = ((fp[2] - fp[4] + fp[5] + fp[6] - fp[7] - fp[8]) + fy / T(2)) * inv;
      guo9(fp, fo, rho, ux, uy, fy, om);
    }
    for (int q = 0; q < 9; ++q) f[c + q * P] = fo[q];
  }
}

// collide_guo_2d_kernel! — collide_guo_2d.jl:5-72 (constant body force)
template <class T>
static void collide_guo_2d(T* f, const uint8_t* solid, int64_t nx, int64_t ny, T om, T Fx, T Fy) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    T fp[9], fo[9];
    for (int q = 0; q < 9; ++q) fp[q] = f[c + q * P];
    if (solid && solid[c]) {
      for (int q = 0; q < 9; ++q) fo[q] = fp[OPP9[q]];
    } else {
      T rho = fp[0] + fp[1] + fp[2] + fp[3] + fp[4] + fp[5] + fp[6] + fp[7] + fp[8];
      T inv = T(1) / rho;
      T ux = ((fp[1] - fp[3] + fp[5] - fp[6] - fp[7] + fp[8]) + Fx / T(2)) * inv;
      T uy = ((fp[2] - fp[4] + fp[5] + fp[6] - fp[7] - fp[8]) + Fy / T(2)) * inv;
      guo9_force(fp, fo, rho, ux, uy, Fx, Fy, om);
    }
    for (int q = 0; q < 9; ++q) f[c + q * P] = fo[q];
  }
}

// compute_macroscopic_forced_2d_kernel! — macroscopic.jl:84-98
template <class T>
static void macroscopic_forced_2d(T* rho, T* ux, T* uy, const T* f, int64_t nx, int64_t ny, T Fx,
                                  T Fy) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    T fp[9];
    for (int q = 0; q < 9; ++q) fp[q] = f[c + q * P];
    T r = fp[0] + fp[1] + fp[2] + fp[3] + fp[4] + fp[5] + fp[6] + fp[7] + fp[8];
    T inv = T(1) / r;
    rho[c] = r;
    ux[c] = ((fp[1] - fp[3] + fp[5] - fp[6] - fp[7] + fp[8]) + Fx / T(2)) * inv;
    uy[c] = ((fp[2] - fp[4] + fp[5] + fp[6] - fp[7] - fp[8]) + Fy / T(2)) * inv;
  }
}

// fused_natconv_step_kernel! — fused_thermal_2d.jl:178-207 (Dirichlet west/east, adiabatic
// south/north, no solids).
template <class T>
static void fused_natconv_step(T* fo, const T* fi, T* go, const T* gi, T* Temp, int64_t nx,
                               int64_t ny, T om_f, T om_T, T beta_g, T T_ref, T T_hot,
                               T T_cold) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      size_t c = I2(i, j, nx);
      T fp[9], gp[9], f2[9], g2[9];
      for (int q = 0; q < 9; ++q) {
        int64_t si = i - CX9[q], sj = j - CY9[q];
        bool ok = si >= 0 && si < nx && sj >= 0 && sj < ny;
        fp[q] = ok ? fi[I2(si, sj, nx) + q * P] : fi[c + OPP9[q] * P];
        gp[q] = ok ? gi[I2(si, sj, nx) + q * P] : gi[c + OPP9[q] * P];
      }
      // apply_thermal_bc_cavity — fused_thermal_2d.jl:37-57
      if (i == 0) {
        gp[1] = -gp[1] + T(2) * T(1.0 / 9.0) * T_hot;
        gp[5] = -gp[5] + T(2) * T(1.0 / 36.0) * T_hot;
        gp[8] = -gp[8] + T(2) * T(1.0 / 36.0) * T_hot;
      }
      if (i == nx - 1) {
        gp[3] = -gp[3] + T(2) * T(1.0 / 9.0) * T_cold;
        gp[6] = -gp[6] + T(2) * T(1.0 / 36.0) * T_cold;
        gp[7] = -gp[7] + T(2) * T(1.0 / 36.0) * T_cold;
      }
      T Tl = gp[0] + gp[1] + gp[2] + gp[3] + gp[4] + gp[5] + gp[6] + gp[7] + gp[8];
      if (Temp) Temp[c] = Tl;
      // macroscopic_boussinesq — fused_thermal_2d.jl:61-72
      T fy = beta_g * (Tl - T_ref);
      T rho = fp[0] + fp[1] + fp[2] + fp[3] + fp[4] + fp[5] + fp[6] + fp[7] + fp[8];
      T inv = T(1) / rho;
      T ux = (fp[1] - fp[3] + fp[5] - fp[6] - fp[7] + fp[8]) * inv;
      T uy = ((fp[2] - fp[4] + fp[5] + fp[6] - fp[7] - fp[8]) + fy / T(2)) * inv;
      thermal9(gp, g2, ux, uy, om_T);
      guo9(fp, f2, rho, ux, uy, fy, om_f);
      for (int q = 0; q < 9; ++q) {
        go[c + q * P] = g2[q];
        fo[c + q * P] = f2[q];
      }
    }
}

// --- D3Q19 thermal and forced kernels ------------------------------------------------

// collide_thermal_3d_kernel! — thermal_3d.jl:12-107 (u arrays from compute_macroscopic_3d!)
template <class T>
static void collide_thermal_3d(T* g, const T* ux, const T* uy, const T* uz, int64_t nx,
                               int64_t ny, int64_t nz, T omT) {
  const size_t P = size_t(nx * ny * nz);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    T gp[19], go[19];
    for (int q = 0; q < 19; ++q) gp[q] = g[c + q * P];
    thermal19(gp, go, ux[c], uy[c], uz[c], omT);
    for (int q = 0; q < 19; ++q) g[c + q * P] = go[q];
  }
}

// compute_temperature_3d_kernel! — thermal_3d.jl:128-136
template <class T>
static void compute_temperature_3d(T* Temp, const T* g, int64_t nx, int64_t ny, int64_t nz) {
  const size_t P = size_t(nx * ny * nz);
  for (size_t c = 0; c < P; ++c) {
    T t = g[c];
    for (int q = 1; q < 19; ++q) t = t + g[c + q * P];
    Temp[c] = t;
  }
}

// apply_fixed_temp_{west,east,south,north,bottom,top}_3d_kernel! — thermal_3d.jl:193-480.
// Per face: the axis slot (pointing into the domain) and the four edge slots with the
// tangent index at which each is an edge slot, 0-based.  Edge n of a face uses T_e[n]
// and is left alone when skip_e[n] is set (_abb_slot_3d, thermal_3d.jl:185-188).
struct FixedTempFace3 {
  int axis_slot;
  int edge_slot[4];
  int edge_axis[4];  // 0: i, 1: j, 2: k
  int edge_hi[4];    // 0: index 0, 1: index n − 1
};
static const FixedTempFace3 FT3[6] = {
    {1, {7, 9, 11, 13}, {1, 1, 2, 2}, {0, 1, 0, 1}},     // west   (i = 0)
    {2, {8, 10, 12, 14}, {1, 1, 2, 2}, {0, 1, 0, 1}},    // east   (i = nx − 1)
    {3, {7, 8, 15, 17}, {0, 0, 2, 2}, {0, 1, 0, 1}},     // south  (j = 0)
    {4, {9, 10, 16, 18}, {0, 0, 2, 2}, {0, 1, 0, 1}},    // north  (j = ny − 1)
    {5, {11, 12, 15, 16}, {0, 0, 1, 1}, {0, 1, 0, 1}},   // bottom (k = 0)
    {6, {13, 14, 17, 18}, {0, 0, 1, 1}, {0, 1, 0, 1}},   // top    (k = nz − 1)
};

template <class T>
static void fixed_temp_face_3d(T* g, int face, T Tw, const T* Te, const bool* skip, int64_t nx,
                               int64_t ny, int64_t nz) {
  const int64_t n[3] = {nx, ny, nz};
  const size_t P = size_t(nx * ny * nz);
  const int ax = face / 2;
  const int64_t fixed = (face % 2) ? n[ax] - 1 : 0;
  const FixedTempFace3& F = FT3[face];
  const T wa = T(2) * T(1.0 / 18.0);
  const T we = T(2) * T(1.0 / 36.0);
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t x[3] = {i, j, k};
        if (x[ax] != fixed) continue;
        const size_t c = I3(i, j, k, nx, ny);
        T& ga = g[c + F.axis_slot * P];
        ga = -ga + wa * Tw;
        for (int e = 0; e < 4; ++e) {
          const int a = F.edge_axis[e];
          const bool on_edge = x[a] == (F.edge_hi[e] ? n[a] - 1 : 0);
          T& gq = g[c + F.edge_slot[e] * P];
          gq = abb_slot(gq, we, Tw, on_edge, Te[e], skip[e]);
        }
      }
}

// _apply_fixed_temp_faces_3d! — thermal_3d.jl:494-535: x-faces own their edges (mean
// temperature where the other face is Dirichlet), y-faces leave the edges with Dirichlet
// x-faces alone, z-faces leave every edge with a Dirichlet x- or y-face alone.
template <class T>
static void apply_fixed_temp_3d(T* g, const bool* dir, const double* Tf, int64_t nx, int64_t ny,
                                int64_t nz) {
  auto corner = [&](int a, int b) { return dir[b] ? (T(Tf[a]) + T(Tf[b])) / T(2) : T(Tf[a]); };
  for (int f = 0; f < 6; ++f) {
    if (!dir[f]) continue;
    T Te[4];
    bool sk[4] = {false, false, false, false};
    const T t = T(Tf[f]);
    if (f < 2) {
      Te[0] = corner(f, KMI_SOUTH), Te[1] = corner(f, KMI_NORTH);
      Te[2] = corner(f, KMI_BOTTOM), Te[3] = corner(f, KMI_TOP);
    } else if (f < 4) {
      Te[0] = t, Te[1] = t, Te[2] = corner(f, KMI_BOTTOM), Te[3] = corner(f, KMI_TOP);
      sk[0] = dir[KMI_WEST], sk[1] = dir[KMI_EAST];
    } else {
      Te[0] = Te[1] = Te[2] = Te[3] = t;
      sk[0] = dir[KMI_WEST], sk[1] = dir[KMI_EAST], sk[2] = dir[KMI_SOUTH], sk[3] = dir[KMI_NORTH];
    }
    fixed_temp_face_3d(g, f, t, Te, sk, nx, ny, nz);
  }
}

// Solid cell of the 3D collide kernels: swap opposite pairs (collide_guo_3d.jl:9-19).
template <class T>
static inline void swap_pairs_19(T* f, size_t c, size_t P) {
  for (int q = 1; q < 19; ++q)
    if (q < OPP19[q]) std::swap(f[c + q * P], f[c + OPP19[q] * P]);
}

// Force-shifted moments, D3Q19 — the common head of the 3D Guo kernels
// (collide_guo_3d.jl:34-39) and compute_macroscopic_forced(_field)_3d! (macroscopic.jl:121-200).
template <class T>
static inline void moments19_forced(const T* f, T fx, T fy, T fz, T& rho, T& ux, T& uy, T& uz) {
  rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] +
        f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
  T inv = T(1) / rho;
  ux = ((f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14]) + fx / T(2)) * inv;
  uy = ((f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18]) + fy / T(2)) * inv;
  uz = ((f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18]) + fz / T(2)) * inv;
}

// collide_guo_3d! (constant force) and collide_guo_field_3d! (per-node force):
// collide_guo_3d.jl:5-190, 199-332.  Fxa == nullptr selects the constant force.
template <class T>
static void collide_guo_3d(T* f, const uint8_t* solid, int64_t nx, int64_t ny, int64_t nz, T om,
                           T Fx, T Fy, T Fz, const T* Fxa = nullptr, const T* Fya = nullptr,
                           const T* Fza = nullptr) {
  const size_t P = size_t(nx * ny * nz);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    if (solid && solid[c]) {
      swap_pairs_19(f, c, P);
      continue;
    }
    T fx = Fxa ? Fxa[c] : Fx, fy = Fxa ? Fya[c] : Fy, fz = Fxa ? Fza[c] : Fz;
    T fp[19], fo[19], rho, ux, uy, uz;
    for (int q = 0; q < 19; ++q) fp[q] = f[c + q * P];
    moments19_forced(fp, fx, fy, fz, rho, ux, uy, uz);
    guo19(fp, fo, rho, ux, uy, uz, fx, fy, fz, om);
    for (int q = 0; q < 19; ++q) f[c + q * P] = fo[q];
  }
}

// collide_boussinesq_3d_kernel! — thermal_3d.jl:559-712: F = (0, β_g (T − T_ref), 0).
template <class T>
static void collide_boussinesq_3d(T* f, const T* Temp, const uint8_t* solid, int64_t nx,
                                  int64_t ny, int64_t nz, T om, T beta_g, T T_ref) {
  const size_t P = size_t(nx * ny * nz);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    if (solid && solid[c]) {
      swap_pairs_19(f, c, P);
      continue;
    }
    T fp[19], fo[19], rho, ux, uy, uz;
    for (int q = 0; q < 19; ++q) fp[q] = f[c + q * P];
    T fy = beta_g * (Temp[c] - T_ref), fx = T(0), fz = T(0);
    moments19_forced(fp, fx, fy, fz, rho, ux, uy, uz);
    guo19(fp, fo, rho, ux, uy, uz, fx, fy, fz, om);
    for (int q = 0; q < 19; ++q) f[c + q * P] = fo[q];
  }
}

// compute_macroscopic_forced_3d! / compute_macroscopic_forced_field_3d! — macroscopic.jl:121-200
template <class T>
static void macroscopic_forced_3d(T* rho, T* ux, T* uy, T* uz, const T* f, int64_t nx, int64_t ny,
                                  int64_t nz, T Fx, T Fy, T Fz, const T* Fxa = nullptr,
                                  const T* Fya = nullptr, const T* Fza = nullptr) {
  const size_t P = size_t(nx * ny * nz);
  for (size_t c = 0; c < P; ++c) {
    T fp[19];
    for (int q = 0; q < 19; ++q) fp[q] = f[c + q * P];
    T fx = Fxa ? Fxa[c] : Fx, fy = Fxa ? Fya[c] : Fy, fz = Fxa ? Fza[c] : Fz;
    moments19_forced(fp, fx, fy, fz, rho[c], ux[c], uy[c], uz[c]);
  }
}

// collide_guo_field_2d_kernel! — collide_guo_2d.jl:93-140 (the constant-force body with
// per-node (fx, fy)).
template <class T>
static void collide_guo_field_2d(T* f, const uint8_t* solid, int64_t nx, int64_t ny, T om,
                                 const T* Fxa, const T* Fya) {
  const size_t P = size_t(nx * ny);
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    T fp[9], fo[9];
    for (int q = 0; q < 9; ++q) fp[q] = f[c + q * P];
    if (solid && solid[c]) {
      for (int q = 0; q < 9; ++q) f[c + q * P] = fp[OPP9[q]];
      continue;
    }
    const T fx = Fxa[c], fy = Fya[c];
    T rho = fp[0] + fp[1] + fp[2] + fp[3] + fp[4] + fp[5] + fp[6] + fp[7] + fp[8];
    T inv = T(1) / rho;
    T ux = ((fp[1] - fp[3] + fp[5] - fp[6] - fp[7] + fp[8]) + fx / T(2)) * inv;
    T uy = ((fp[2] - fp[4] + fp[5] + fp[6] - fp[7] - fp[8]) + fy / T(2)) * inv;
    guo9_force(fp, fo, rho, ux, uy, fx, fy, om);
    for (int q = 0; q < 9; ++q) f[c + q * P] = fo[q];
  }
}

// ----------------------------------------------------------------------------------
// Generic descriptor-driven step: the semantics the MI355X kernels implement.
// Each piece is one of the restated reference kernels above, chained in the order of
// the reference drivers (basic.jl:121-136, thermal.jl:100-133, test_mrt.jl:19-25).
// ----------------------------------------------------------------------------------
struct Axis {
  int64_t n;
  bool periodic;
};

// Pull with per-face rules: periodic faces wrap; a source beyond any other face is
// replaced by the node's own opposite population (halfway bounce-back; the reference's
// `ifelse(inside, pull, f_in[opp])`), plus the Ladd term 6 w_q ρ_w (c_q·u_w) of every
// crossed MOVING_WALL face (stream_periodic_3d.jl:97-166).  Zou–He faces pull with
// bounce-back as in stream_2d! before the closure overwrites the unknowns.  The Ladd
// term is a momentum correction: thermal populations (ladd = false) bounce back plainly.
template <class T>
static void pull_generic(const kmi_desc& d, T* fo, const T* fi, bool ladd = true) {
  const int Q = d.stencil == KMI_D2Q9 ? 9 : 19;
  const int64_t nx = d.nx, ny = d.ny, nz = d.stencil == KMI_D2Q9 ? 1 : d.nz;
  const size_t P = size_t(nx * ny * nz);
  const int* cx = Q == 9 ? CX9 : CX19;
  const int* cy = Q == 9 ? CY9 : CY19;
  const int* cz = Q == 9 ? nullptr : CZ19;
  const int* opp = Q == 9 ? OPP9 : OPP19;
  const double* w = Q == 9 ? W9 : W19;
  const bool per[3] = {d.faces[KMI_WEST].kind == KMI_FACE_PERIODIC,
                       d.faces[KMI_SOUTH].kind == KMI_FACE_PERIODIC,
                       Q == 19 && d.faces[KMI_BOTTOM].kind == KMI_FACE_PERIODIC};
  const int64_t n[3] = {nx, ny, nz};
#pragma omp parallel for
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t x[3] = {i, j, k};
        size_t c = I3(i, j, k, nx, ny);
        for (int q = 0; q < Q; ++q) {
          const int cq[3] = {cx[q], cy[q], cz ? cz[q] : 0};
          int64_t s[3];
          bool ok = true;
          int crossed[3] = {-1, -1, -1};
          for (int a = 0; a < 3; ++a) {
            s[a] = x[a] - cq[a];
            if (s[a] < 0 || s[a] >= n[a]) {
              if (per[a]) {
                s[a] = (s[a] + n[a]) % n[a];
              } else {
                ok = false;
                crossed[a] = 2 * a + (s[a] < 0 ? 0 : 1);
              }
            }
          }
          T v;
          if (ok) {
            v = fi[I3(s[0], s[1], s[2], nx, ny) + q * P];
          } else {
            v = fi[c + opp[q] * P];
            for (int a = 0; a < 3; ++a) {
              if (crossed[a] < 0) continue;
              const kmi_face& F = d.faces[crossed[a]];
              if (!ladd || F.kind != KMI_FACE_MOVING_WALL) continue;
              T cu = cdotu<T>(cq[0], cq[1], cq[2], T(F.u[0]), T(F.u[1]), T(F.u[2]));
              v = v + T(six_w(w[q])) * T(F.rho) * cu;
            }
          }
          fo[c + q * P] = v;
        }
      }
}

// Per-node arrays of the profile faces, indexed [face][0: u_x or rho, 1: u_y].
struct Profiles {
  const double* p[4][2] = {};
};

template <class T>
static void zou_he_faces(const kmi_desc& d, T* f, const Profiles* pr = nullptr) {
  int order[6] = {0, 1, 2, 3, 4, 5};
  int n = 6;
  if (d.n_bc_order > 0) {
    n = d.n_bc_order;
    for (int i = 0; i < n; ++i) order[i] = d.bc_order[i];
  }
  for (int oi = 0; oi < n; ++oi) {
    int fid = order[oi];
    const kmi_face& F = d.faces[fid];
    if (d.stencil == KMI_D3Q19) {
      // _apply_zou_he_velocity_3d! argument mapping: normal component, then the two
      // tangents in axis order (boundary_3d.jl:317-390).
      const int ax = ZH3_TABLE[fid].axis - 1;
      const int a1 = ax == 0 ? 1 : 0, a2 = ax == 2 ? 1 : 2;
      if (F.kind == KMI_FACE_ZOUHE_VELOCITY)
        zou_he_velocity_3d(f, fid, T(F.u[ax]), T(F.u[a1]), T(F.u[a2]), d.nx, d.ny, d.nz);
      else if (F.kind == KMI_FACE_ZOUHE_PRESSURE)
        zou_he_pressure_3d(f, fid, T(F.rho), d.nx, d.ny, d.nz);
      continue;
    }
    if (F.kind == KMI_FACE_ZOUHE_VELOCITY_PROFILE && pr) {
      if (fid == KMI_NORTH) zou_he_north_spatial(f, pr->p[fid][0], pr->p[fid][1], d.nx, d.ny);
      else if (fid == KMI_SOUTH) zou_he_south_spatial(f, pr->p[fid][0], pr->p[fid][1], d.nx, d.ny);
      else if (fid == KMI_WEST) zou_he_west_spatial(f, pr->p[fid][0], pr->p[fid][1], d.nx, d.ny);
    } else if (F.kind == KMI_FACE_ZOUHE_PRESSURE_PROFILE && pr) {
      if (fid == KMI_EAST) zou_he_pressure_east_spatial(f, pr->p[fid][0], d.nx, d.ny);
    }
    if (F.kind == KMI_FACE_ZOUHE_VELOCITY) {
      if (fid == KMI_NORTH) zou_he_north(f, T(F.u[0]), d.nx, d.ny);
      else if (fid == KMI_SOUTH) zou_he_south(f, T(F.u[0]), d.nx, d.ny);
      else if (fid == KMI_WEST) zou_he_west(f, T(F.u[0]), d.nx, d.ny);
    } else if (F.kind == KMI_FACE_ZOUHE_PRESSURE) {
      if (fid == KMI_EAST) zou_he_pressure_east(f, T(F.rho), d.nx, d.ny);
      else if (fid == KMI_WEST) zou_he_pressure_west(f, T(F.rho), d.nx, d.ny);
    }
  }
}

// ---- face rebuild: apply_bc_rebuild_2d!/_3d! (src/bc/rebuild_2d.jl, rebuild_3d.jl) ----------
// _trt_collide_local — src/bc/specs.jl:7-50 (regularised TRT, 1-based names in comments)
template <class T>
static void trt_collide_local_2d(const T* f, T* o, T sp, T sm) {
  T rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  T ux = (f[1] - f[3] + f[5] - f[7] + f[8] - f[6]) / rho;  // f2 - f4 + f6 - f8 + f9 - f7
  T uy = (f[2] - f[4] + f[5] - f[7] + f[6] - f[8]) / rho;  // f3 - f5 + f6 - f8 + f7 - f9
  T usq = ux * ux + uy * uy;
  T fe[9];
  fe[0] = T(W9[0]) * rho * (T(1) - T(1.5) * usq);
  for (int q = 1; q < 9; ++q) fe[q] = feq<T>(W9[q], rho, cdotu<T>(CX9[q], CY9[q], 0, ux, uy, T(0)), usq);
  T Pxx = (f[1] - fe[1]) + (f[3] - fe[3]) + (f[5] - fe[5]) + (f[6] - fe[6]) + (f[7] - fe[7]) + (f[8] - fe[8]);
  T Pyy = (f[2] - fe[2]) + (f[4] - fe[4]) + (f[5] - fe[5]) + (f[6] - fe[6]) + (f[7] - fe[7]) + (f[8] - fe[8]);
  T Pxy = (f[5] - fe[5]) - (f[6] - fe[6]) + (f[7] - fe[7]) - (f[8] - fe[8]);
  T h = T(0.5);
  T fn[9];
  fn[0] = -h * T(2.0 / 9.0) * (Pxx + Pyy);
  fn[1] = h * T(1.0 / 9.0) * (T(2) * Pxx - Pyy);
  fn[2] = h * T(1.0 / 9.0) * (-Pxx + T(2) * Pyy);
  fn[3] = fn[1];
  fn[4] = fn[2];
  fn[5] = h * T(1.0 / 36.0) * (Pxx + Pyy) + T(1.0 / 4.0) * Pxy;
  fn[6] = h * T(1.0 / 36.0) * (Pxx + Pyy) - T(1.0 / 4.0) * Pxy;
  fn[7] = fn[5];
  fn[8] = fn[6];
  T a = (sp + sm) * h, b = (sp - sm) * h;
  o[0] = fe[0] + (T(1) - sp) * fn[0];
  for (int q = 1; q < 9; ++q) o[q] = fe[q] + (T(1) - a) * fn[q] - b * fn[OPP9[q]];
}

// _trt_collide_local_3d — src/bc/specs.jl:52-108
template <class T>
static void trt_collide_local_3d(const T* f, T* o, T sp, T sm) {
  T rho = f[0];
  for (int q = 1; q < 19; ++q) rho = rho + f[q];
  T ux = ((f[1] - f[2]) + (f[7] - f[10]) + (f[9] - f[8]) + (f[11] - f[14]) + (f[13] - f[12])) / rho;
  T uy = ((f[3] - f[4]) + (f[7] - f[10]) + (f[8] - f[9]) + (f[15] - f[18]) + (f[17] - f[16])) / rho;
  T uz = ((f[5] - f[6]) + (f[11] - f[14]) + (f[12] - f[13]) + (f[15] - f[18]) + (f[16] - f[17])) / rho;
  T a = (sp + sm) * T(0.5), b = (sp - sm) * T(0.5);
  T usq = ux * ux + uy * uy + uz * uz;
  T fe[19];
  fe[0] = T(W19[0]) * rho * (T(1) - T(1.5) * usq);
  for (int q = 1; q < 19; ++q)
    fe[q] = feq<T>(W19[q], rho, cdotu<T>(CX19[q], CY19[q], CZ19[q], ux, uy, uz), usq);
  o[0] = f[0] - sp * (f[0] - fe[0]);
  for (int q = 1; q < 19; ++q) o[q] = f[q] - a * (f[q] - fe[q]) - b * (f[OPP19[q]] - fe[OPP19[q]]);
}

// _bc_west_zh_velocity_2d! / _bc_east_zh_pressure_2d! — rebuild_2d.jl:1-50, rows j = 2..Ny-1;
// reads the pre-step populations fi at the reference's explicit neighbour indices.
template <class T>
static void rebuild_2d(T* fo, const T* fi, const kmi_desc& d, const double* prof, T sp, T sm) {
  const int64_t nx = d.nx, ny = d.ny;
  const size_t P = size_t(nx * ny);
  auto F = [&](int64_t i, int64_t j, int q) { return fi[I2(i, j, nx) + size_t(q) * P]; };
  for (int64_t j = 1; j < ny - 1; ++j) {
    T fp[9], o[9];
    if (d.faces[KMI_WEST].kind == KMI_FACE_ZOUHE_VELOCITY_REBUILD) {
      fp[0] = F(0, j, 0); fp[2] = F(0, j - 1, 2); fp[3] = F(1, j, 3);
      fp[4] = F(0, j + 1, 4); fp[6] = F(1, j - 1, 6); fp[7] = F(1, j + 1, 7);
      T u_in = T(prof[j]);
      T rho_w = (fp[0] + fp[2] + fp[4] + T(2) * (fp[3] + fp[6] + fp[7])) / (T(1) - u_in);
      fp[1] = fp[3] + T(2.0 / 3.0) * rho_w * u_in;
      fp[5] = fp[7] - T(0.5) * (fp[2] - fp[4]) + T(1.0 / 6.0) * rho_w * u_in;
      fp[8] = fp[6] + T(0.5) * (fp[2] - fp[4]) + T(1.0 / 6.0) * rho_w * u_in;
      trt_collide_local_2d(fp, o, sp, sm);
      for (int q = 0; q < 9; ++q) fo[I2(0, j, nx) + size_t(q) * P] = o[q];
    }
    if (d.faces[KMI_EAST].kind == KMI_FACE_ZOUHE_PRESSURE_REBUILD) {
      const int64_t e = nx - 1;
      fp[0] = F(e, j, 0); fp[1] = F(e - 1, j, 1); fp[2] = F(e, j - 1, 2);
      fp[4] = F(e, j + 1, 4); fp[5] = F(e - 1, j - 1, 5); fp[8] = F(e - 1, j + 1, 8);
      T rho = T(d.faces[KMI_EAST].rho);
      T u_x = -T(1) + (fp[0] + fp[2] + fp[4] + T(2) * (fp[1] + fp[5] + fp[8])) / rho;
      fp[3] = fp[1] - T(2.0 / 3.0) * rho * u_x;
      fp[6] = fp[8] - T(0.5) * (fp[2] - fp[4]) - T(1.0 / 6.0) * rho * u_x;
      fp[7] = fp[5] + T(0.5) * (fp[2] - fp[4]) - T(1.0 / 6.0) * rho * u_x;
      trt_collide_local_2d(fp, o, sp, sm);
      for (int q = 0; q < 9; ++q) fo[I2(e, j, nx) + size_t(q) * P] = o[q];
    }
  }
}

// _bc_west_zh_velocity_3d! / _bc_east_zh_pressure_3d! — rebuild_3d.jl:1-89, j = 2..Ny-1,
// k = 2..Nz-1; the west profile is the (Ny, Nz) matrix, index j + Ny k.
template <class T>
static void rebuild_3d(T* fo, const T* fi, const kmi_desc& d, const double* prof, T sp, T sm) {
  const int64_t nx = d.nx, ny = d.ny, nz = d.nz;
  const size_t P = size_t(nx * ny * nz);
  auto F = [&](int64_t i, int64_t j, int64_t k, int q) { return fi[I3(i, j, k, nx, ny) + size_t(q) * P]; };
  for (int64_t k = 1; k < nz - 1; ++k)
    for (int64_t j = 1; j < ny - 1; ++j) {
      for (int side = 0; side < 2; ++side) {
        const bool west = side == 0;
        if (west && d.faces[KMI_WEST].kind != KMI_FACE_ZOUHE_VELOCITY_REBUILD) continue;
        if (!west && d.faces[KMI_EAST].kind != KMI_FACE_ZOUHE_PRESSURE_REBUILD) continue;
        const int64_t b = west ? 0 : nx - 1, n = west ? 1 : nx - 2;  // boundary and next column
        T fp[19], o[19];
        fp[0] = F(b, j, k, 0);
        fp[3] = F(b, j - 1, k, 3);
        fp[4] = F(b, j + 1, k, 4);
        fp[5] = F(b, j, k - 1, 5);
        fp[6] = F(b, j, k + 1, 6);
        fp[15] = F(b, j - 1, k - 1, 15);
        fp[16] = F(b, j + 1, k - 1, 16);
        fp[17] = F(b, j - 1, k + 1, 17);
        fp[18] = F(b, j + 1, k + 1, 18);
        T sum_par = fp[0] + fp[3] + fp[4] + fp[5] + fp[6] + fp[15] + fp[16] + fp[17] + fp[18];
        T t1 = fp[3] - fp[4], t2 = fp[5] - fp[6];
        if (west) {
          fp[2] = F(n, j, k, 2);
          fp[8] = F(n, j - 1, k, 8);
          fp[10] = F(n, j + 1, k, 10);
          fp[12] = F(n, j, k - 1, 12);
          fp[14] = F(n, j, k + 1, 14);
          T sum_out = fp[2] + fp[8] + fp[10] + fp[12] + fp[14];
          T u_n = T(prof[j + ny * k]);
          T rho_w = (sum_par + T(2) * sum_out) / (T(1) - u_n);
          fp[1] = fp[2] + T(1.0 / 3.0) * rho_w * u_n;
          fp[7] = fp[10] - T(0.5) * t1 + T(1.0 / 6.0) * rho_w * u_n;
          fp[9] = fp[8] + T(0.5) * t1 + T(1.0 / 6.0) * rho_w * u_n;
          fp[11] = fp[14] - T(0.5) * t2 + T(1.0 / 6.0) * rho_w * u_n;
          fp[13] = fp[12] + T(0.5) * t2 + T(1.0 / 6.0) * rho_w * u_n;
        } else {
          fp[1] = F(n, j, k, 1);
          fp[7] = F(n, j - 1, k, 7);
          fp[9] = F(n, j + 1, k, 9);
          fp[11] = F(n, j, k - 1, 11);
          fp[13] = F(n, j, k + 1, 13);
          T sum_out = fp[1] + fp[7] + fp[9] + fp[11] + fp[13];
          T rho = T(d.faces[KMI_EAST].rho);
          T u_n = -(T(1) - (sum_par + T(2) * sum_out) / rho);
          fp[2] = fp[1] - T(1.0 / 3.0) * rho * u_n;
          fp[8] = fp[9] - T(0.5) * t1 - T(1.0 / 6.0) * rho * u_n;
          fp[10] = fp[7] + T(0.5) * t1 - T(1.0 / 6.0) * rho * u_n;
          fp[12] = fp[13] - T(0.5) * t2 - T(1.0 / 6.0) * rho * u_n;
          fp[14] = fp[11] + T(0.5) * t2 - T(1.0 / 6.0) * rho * u_n;
        }
        trt_collide_local_3d(fp, o, sp, sm);
        for (int q = 0; q < 19; ++q) fo[I3(b, j, k, nx, ny) + size_t(q) * P] = o[q];
      }
    }
}

// apply_bc_rebuild_2d!/_3d! after the fused step, at the descriptor's TRT rates.
template <class T>
static void apply_rebuilds(const kmi_desc& d, T* fo, const T* fi, const double* west_prof) {
  const bool any = d.faces[KMI_WEST].kind == KMI_FACE_ZOUHE_VELOCITY_REBUILD ||
                   d.faces[KMI_EAST].kind == KMI_FACE_ZOUHE_PRESSURE_REBUILD;
  if (!any) return;
  if (d.stencil == KMI_D2Q9) rebuild_2d(fo, fi, d, west_prof, T(d.s_plus), T(d.s_minus));
  else rebuild_3d(fo, fi, d, west_prof, T(d.s_plus), T(d.s_minus));
}

template <class T>
static void generic_step(const kmi_desc& d, const uint8_t* solid, T*& f, T*& f2, T*& g, T*& g2,
                         std::vector<T>& scratch, const Profiles* pr = nullptr,
                         const T* const* ffield = nullptr) {
  const bool d3 = d.stencil == KMI_D3Q19;
  const int64_t nx = d.nx, ny = d.ny, nz = d3 ? d.nz : 1;
  const size_t P = size_t(nx * ny * nz);
  pull_generic(d, f2, f);
  zou_he_faces(d, f2, pr);
  if (d.thermal) {
    pull_generic(d, g2, g, false);
    bool dir[6];
    double Tf[6];
    for (int a = 0; a < 6; ++a) {
      dir[a] = (d3 || a < 4) && d.faces[a].kind != KMI_FACE_PERIODIC &&
               d.faces[a].thermal == KMI_THERMAL_DIRICHLET;
      Tf[a] = d.faces[a].T;
    }
    scratch.resize(5 * P);
    T* Temp = scratch.data();
    T* rho = Temp + P;
    T* ux = rho + P;
    T* uy = ux + P;
    T* uz = uy + P;
    if (d3) {
      // run_natural_convection_3d step (thermal.jl:500-515): ABB faces, T, unforced u for
      // the thermal collision, Boussinesq collision with the force-shifted u.
      apply_fixed_temp_3d(g2, dir, Tf, nx, ny, nz);
      compute_temperature_3d(Temp, g2, nx, ny, nz);
      macroscopic_3d(rho, ux, uy, uz, f2, nx, ny, nz);
      collide_thermal_3d(g2, ux, uy, uz, nx, ny, nz, T(d.omega_T));
      collide_boussinesq_3d(f2, Temp, solid, nx, ny, nz, T(d.omega), T(d.beta_g), T(d.T_ref));
      std::swap(g, g2);
      std::swap(f, f2);
      return;
    }
    apply_fixed_temp_2d(g2, dir, Tf, nx, ny);
    compute_temperature_2d(Temp, g2, nx, ny);
    if (d.coupling == KMI_COUPLING_SPLIT) {
      // thermal.jl:123-125: u from compute_macroscopic_2d! (no force shift)
      macroscopic_2d(rho, ux, uy, f2, nx, ny);
    } else {
      // fused_thermal_2d.jl:61-72: u with the F/2 shift in y
      for (size_t c = 0; c < P; ++c) {
        T fp[9];
        for (int q = 0; q < 9; ++q) fp[q] = f2[c + q * P];
        T fy = T(d.beta_g) * (Temp[c] - T(d.T_ref));
        T r = fp[0] + fp[1] + fp[2] + fp[3] + fp[4] + fp[5] + fp[6] + fp[7] + fp[8];
        T inv = T(1) / r;
        ux[c] = (fp[1] - fp[3] + fp[5] - fp[6] - fp[7] + fp[8]) * inv;
        uy[c] = ((fp[2] - fp[4] + fp[5] + fp[6] - fp[7] - fp[8]) + fy / T(2)) * inv;
      }
    }
    collide_thermal_2d(g2, ux, uy, nx, ny, T(d.omega_T));
    collide_boussinesq_2d(f2, Temp, solid, nx, ny, T(d.omega), T(d.beta_g), T(d.T_ref));
    std::swap(g, g2);
  } else if (!d3) {
    if (d.collision == KMI_BGK && d.forcing == KMI_FORCE_GUO_FIELD)
      collide_guo_field_2d(f2, solid, nx, ny, T(d.omega), ffield[0], ffield[1]);
    else if (d.collision == KMI_BGK && d.forcing != KMI_FORCE_NONE)
      collide_guo_2d(f2, solid, nx, ny, T(d.omega), T(d.force[0]), T(d.force[1]));
    else if (d.collision == KMI_BGK) collide_2d(f2, solid, nx, ny, T(d.omega));
    else if (d.collision == KMI_TRT) collide_trt_2d(f2, solid, nx, ny, T(d.s_plus), T(d.s_minus));
    else collide_mrt_2d(f2, solid, nx, ny, T(d.s_e), T(d.s_eps), T(d.s_q), T(d.s_nu));
  } else {
    if (d.collision == KMI_BGK && d.forcing == KMI_FORCE_GUO_FIELD)
      collide_guo_3d(f2, solid, nx, ny, nz, T(d.omega), T(0), T(0), T(0), ffield[0], ffield[1],
                     ffield[2]);
    else if (d.collision == KMI_BGK && d.forcing != KMI_FORCE_NONE)
      collide_guo_3d(f2, solid, nx, ny, nz, T(d.omega), T(d.force[0]), T(d.force[1]),
                     T(d.force[2]));
    else if (d.collision == KMI_BGK) collide_3d(f2, solid, nx, ny, nz, T(d.omega));
    else collide_trt_3d(f2, solid, nx, ny, nz, T(d.s_plus), T(d.s_minus));
  }
  apply_rebuilds(d, f2, f, pr ? pr->p[KMI_WEST][0] : nullptr);
  std::swap(f, f2);
}

// ---- LI-BB V2 (li_bb_2d_v2.jl:35-110, li_bb_3d_v2.jl:13-43) --------------------------------
// _libb_branch — li_bb_2d.jl:70-83
template <class T>
static T libb_branch(T q_w, T f_post_here, T f_post_back, T f_bar_post_here, T delta) {
  const T half = T(0.5);
  if (q_w <= half) return T(2) * q_w * f_post_here + (T(1) - T(2) * q_w) * f_post_back + delta;
  const T inv_two_q = T(1) / (T(2) * q_w);
  return inv_two_q * f_post_here + (T(1) - inv_two_q) * f_bar_post_here + delta * inv_two_q;
}

// One flagged link of the pre-phase brick: 1-based q, its opposite, and the moving-wall
// term as the brick writes it: sign · coef · (s_a u_a  ±  s_b u_b) with u the link's wall
// velocity components (0 x, 1 y, 2 z; b = -1: single component).
struct LibbLink {
  int q, qbar, neg;       // neg: the leading constant is −coef
  double coef;
  int a, a_neg, b, b_sub;  // (±u_a) then (+ or −) u_b
};
// ApplyLiBBPrePhase — dsl/bricks.jl:518-581 (pairs (2,4), (3,5), (6,8), (7,9))
static const LibbLink LIBB9[8] = {
    {2, 4, 1, 2.0 / 3.0, 0, 0, -1, 0}, {4, 2, 0, 2.0 / 3.0, 0, 0, -1, 0},
    {3, 5, 1, 2.0 / 3.0, 1, 0, -1, 0}, {5, 3, 0, 2.0 / 3.0, 1, 0, -1, 0},
    {6, 8, 1, 1.0 / 6.0, 0, 0, 1, 0},  {8, 6, 0, 1.0 / 6.0, 0, 0, 1, 0},
    {7, 9, 1, 1.0 / 6.0, 0, 1, 1, 0},  {9, 7, 0, 1.0 / 6.0, 0, 1, 1, 0}};
// ApplyLiBBPrePhase_3D — dsl/bricks_3d.jl:368-473
static const LibbLink LIBB19[18] = {
    {2, 3, 1, 1.0 / 3.0, 0, 0, -1, 0},   {3, 2, 0, 1.0 / 3.0, 0, 0, -1, 0},
    {4, 5, 1, 1.0 / 3.0, 1, 0, -1, 0},   {5, 4, 0, 1.0 / 3.0, 1, 0, -1, 0},
    {6, 7, 1, 1.0 / 3.0, 2, 0, -1, 0},   {7, 6, 0, 1.0 / 3.0, 2, 0, -1, 0},
    {8, 11, 1, 1.0 / 6.0, 0, 0, 1, 0},   {11, 8, 0, 1.0 / 6.0, 0, 0, 1, 0},
    {9, 10, 1, 1.0 / 6.0, 0, 1, 1, 0},   {10, 9, 1, 1.0 / 6.0, 0, 0, 1, 1},
    {12, 15, 1, 1.0 / 6.0, 0, 0, 2, 0},  {15, 12, 0, 1.0 / 6.0, 0, 0, 2, 0},
    {13, 14, 1, 1.0 / 6.0, 0, 1, 2, 0},  {14, 13, 1, 1.0 / 6.0, 0, 0, 2, 1},
    {16, 19, 1, 1.0 / 6.0, 1, 0, 2, 0},  {19, 16, 0, 1.0 / 6.0, 1, 0, 2, 0},
    {17, 18, 1, 1.0 / 6.0, 1, 1, 2, 0},  {18, 17, 1, 1.0 / 6.0, 1, 0, 2, 1}};

// Pre-phase substitution on the pulled field f2 (fluid cells), reading the previous field fi;
// q_wall / uw arrays dense in the reference layout [cell + (q−1)P].
template <class T>
static void libb_prephase(const kmi_desc& d, T* f2, const T* fi, const uint8_t* solid,
                          const T* q_wall, const T* const* uw) {
  const bool d3 = d.stencil == KMI_D3Q19;
  const size_t P = size_t(d.nx * d.ny * (d3 ? d.nz : 1));
  const LibbLink* tab = d3 ? LIBB19 : LIBB9;
  const int nl = d3 ? 18 : 8;
#pragma omp parallel for
  for (int64_t c = 0; c < int64_t(P); ++c) {
    if (solid && solid[c]) continue;
    for (int n = 0; n < nl; ++n) {
      const LibbLink& L = tab[n];
      const size_t iq = c + size_t(L.q - 1) * P, ib = c + size_t(L.qbar - 1) * P;
      const T qw = q_wall[iq];
      if (!(qw > T(0))) continue;
      T ua = uw[L.a][iq];
      if (L.a_neg) ua = -ua;
      T comb = ua;
      if (L.b >= 0) comb = L.b_sub ? ua - uw[L.b][iq] : ua + uw[L.b][iq];
      const T k = L.neg ? -T(L.coef) : T(L.coef);
      const T delta = k * comb;
      f2[ib] = libb_branch(qw, fi[iq], f2[iq], fi[ib], delta);
    }
  }
}

// fused_trt_libb_v2_step! chain on the descriptor's faces: pull (halfway BB / periodic),
// pre-phase on fluid cells, collision, SolidInert rest equilibrium on solid cells.
template <class T>
static void libb_step(const kmi_desc& d, const uint8_t* solid, T*& f, T*& f2, const T* q_wall,
                      const T* const* uw, const double* west_prof = nullptr) {
  const bool d3 = d.stencil == KMI_D3Q19;
  const int64_t nx = d.nx, ny = d.ny, nz = d3 ? d.nz : 1;
  const size_t P = size_t(nx * ny * nz);
  pull_generic(d, f2, f);
  zou_he_faces(d, f2);
  libb_prephase(d, f2, f, solid, q_wall, uw);
  if (!d3) {
    if (d.collision == KMI_BGK) collide_2d(f2, (const uint8_t*)nullptr, nx, ny, T(d.omega));
    else collide_trt_2d(f2, (const uint8_t*)nullptr, nx, ny, T(d.s_plus), T(d.s_minus));
  } else {
    if (d.collision == KMI_BGK) collide_3d(f2, (const uint8_t*)nullptr, nx, ny, nz, T(d.omega));
    else collide_trt_3d(f2, (const uint8_t*)nullptr, nx, ny, nz, T(d.s_plus), T(d.s_minus));
  }
  const int Q = d3 ? 19 : 9;
  const double* w = d3 ? W19 : W9;
  for (size_t c = 0; c < P; ++c)  // SolidInert — dsl/bricks.jl:133-154: feq(ρ=1, u=0) = w_q
    if (solid && solid[c])
      for (int q = 0; q < Q; ++q) f2[c + q * P] = T(w[q]) * T(1) * (T(1) + T(0) + T(0) - T(0));
  apply_rebuilds(d, f2, f, west_prof);  // run_cylinder_libb_2d / run_sphere_libb_3d
  std::swap(f, f2);
}

// _compute_drag_libb_mei_2d_host — cylinder_libb.jl:61-104 (FP64 accumulation)
template <class T>
static void drag_mei_2d(const T* f, const T* qw, const T* uwx, const T* uwy, int64_t Nx,
                        int64_t Ny, double* out) {
  const size_t P = size_t(Nx * Ny);
  double Fx = 0.0, Fy = 0.0;
  for (int64_t j = 0; j < Ny; ++j)
    for (int64_t i = 0; i < Nx; ++i)
      for (int q = 1; q < 9; ++q) {
        const size_t c = I2(i, j, Nx);
        const double q_w = double(qw[c + q * P]);
        if (!(q_w > 0)) continue;
        const int qbar = OPP9[q];
        const int64_t im = i - CX9[q], jm = j - CY9[q];
        const double back = (im >= 0 && im < Nx && jm >= 0 && jm < Ny)
                                ? double(f[I2(im, jm, Nx) + q * P])
                                : double(f[c + qbar * P]);
        const double here = double(f[c + q * P]), bar = double(f[c + qbar * P]);
        const double w_q = (q <= 4) ? 1.0 / 9.0 : 1.0 / 36.0;
        const double cu = double(CX9[q]) * double(uwx[c + q * P]) + double(CY9[q]) * double(uwy[c + q * P]);
        const double delta = -6.0 * w_q * cu;
        double arriving;
        if (q_w <= 0.5) {
          arriving = 2 * q_w * here + (1 - 2 * q_w) * back + delta;
        } else {
          const double inv2q = 1.0 / (2 * q_w);
          arriving = inv2q * here + (1 - inv2q) * bar + inv2q * delta;
        }
        const double F = here + arriving;
        Fx += double(CX9[q]) * F;
        Fy += double(CY9[q]) * F;
      }
  out[0] = Fx;
  out[1] = Fy;
}

// _compute_drag_libb_3d_host — li_bb_3d_v2.jl:183-201 (2D: compute_drag_libb_2d,
// cylinder_libb.jl:27-42): Σ 2 c_q f_q over cut links.
template <class T>
static void drag_halfway(const kmi_desc& d, const T* f, const T* qw, double* out) {
  const bool d3 = d.stencil == KMI_D3Q19;
  const size_t P = size_t(d.nx * d.ny * (d3 ? d.nz : 1));
  const int Q = d3 ? 19 : 9;
  const int* cx = d3 ? CX19 : CX9;
  const int* cy = d3 ? CY19 : CY9;
  double F[3] = {0, 0, 0};
  for (size_t c = 0; c < P; ++c)
    for (int q = 1; q < Q; ++q)
      if (qw[c + q * P] > 0) {
        const double fv = double(f[c + q * P]);
        F[0] += 2.0 * double(cx[q]) * fv;
        F[1] += 2.0 * double(cy[q]) * fv;
        if (d3) F[2] += 2.0 * double(CZ19[q]) * fv;
      }
  out[0] = F[0], out[1] = F[1], out[2] = F[2];
}

// Neumaier-compensated sum.
struct KSum {
  double s = 0.0, c = 0.0;
  void add(double x) {
    double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  double val() const { return s + c; }
};

}  // namespace ko

// ==================================================================================
// Exported C entry points (double: _d, float: _f).  Array arguments follow the
// reference layout; `solid` may be NULL.
// ==================================================================================
extern "C" {

#define KO_BOTH(X) X(d, double) X(f, float)

#define KO_STREAM(S, T)                                                                  \
  void ko_stream_2d_##S(T* fo, const T* fi, int64_t nx, int64_t ny) {                   \
    ko::stream_2d(fo, fi, nx, ny);                                                       \
  }                                                                                      \
  void ko_stream_fully_periodic_2d_##S(T* fo, const T* fi, int64_t nx, int64_t ny) {    \
    ko::stream_fully_periodic_2d(fo, fi, nx, ny);                                        \
  }                                                                                      \
  void ko_stream_periodic_x_wall_y_2d_##S(T* fo, const T* fi, int64_t nx, int64_t ny) { \
    ko::stream_periodic_x_wall_y_2d(fo, fi, nx, ny);                                     \
  }                                                                                      \
  void ko_collide_2d_##S(T* f, const uint8_t* s, int64_t nx, int64_t ny, double om) {   \
    ko::collide_2d(f, s, nx, ny, T(om));                                                 \
  }                                                                                      \
  void ko_collide_trt_2d_##S(T* f, const uint8_t* s, int64_t nx, int64_t ny, double sp, \
                             double sm) {                                                \
    ko::collide_trt_2d(f, s, nx, ny, T(sp), T(sm));                                      \
  }                                                                                      \
  void ko_collide_mrt_2d_##S(T* f, const uint8_t* s, int64_t nx, int64_t ny, double se, \
                             double sp2, double sq, double sn) {                         \
    ko::collide_mrt_2d(f, s, nx, ny, T(se), T(sp2), T(sq), T(sn));                       \
  }                                                                                      \
  void ko_fused_bgk_step_2d_##S(T* fo, const T* fi, T* r, T* u, T* v, const uint8_t* s, \
                                int64_t nx, int64_t ny, double om) {                     \
    ko::fused_2d(fo, fi, r, u, v, s, nx, ny, false, T(om), T(0));                        \
  }                                                                                      \
  void ko_fused_trt_step_2d_##S(T* fo, const T* fi, T* r, T* u, T* v, const uint8_t* s, \
                                int64_t nx, int64_t ny, double sp, double sm) {          \
    ko::fused_2d(fo, fi, r, u, v, s, nx, ny, true, T(sp), T(sm));                        \
  }                                                                                      \
  void ko_macroscopic_2d_##S(T* r, T* u, T* v, const T* f, int64_t nx, int64_t ny) {    \
    ko::macroscopic_2d(r, u, v, f, nx, ny);                                              \
  }                                                                                      \
  void ko_zou_he_north_2d_##S(T* f, double u, int64_t nx, int64_t ny) {                 \
    ko::zou_he_north(f, T(u), nx, ny);                                                   \
  }                                                                                      \
  void ko_zou_he_south_2d_##S(T* f, double u, int64_t nx, int64_t ny) {                 \
    ko::zou_he_south(f, T(u), nx, ny);                                                   \
  }                                                                                      \
  void ko_zou_he_west_2d_##S(T* f, double u, int64_t nx, int64_t ny) {                  \
    ko::zou_he_west(f, T(u), nx, ny);                                                    \
  }                                                                                      \
  void ko_zou_he_pressure_east_2d_##S(T* f, double r, int64_t nx, int64_t ny) {         \
    ko::zou_he_pressure_east(f, T(r), nx, ny);                                           \
  }                                                                                      \
  void ko_zou_he_pressure_west_2d_##S(T* f, double r, int64_t nx, int64_t ny) {         \
    ko::zou_he_pressure_west(f, T(r), nx, ny);                                           \
  }                                                                                      \
  void ko_zou_he_velocity_3d_##S(T* f, int64_t face, double un, double ut1, double ut2,     \
                                 int64_t nx, int64_t ny, int64_t nz) {                   \
    ko::zou_he_velocity_3d(f, int(face), T(un), T(ut1), T(ut2), nx, ny, nz);                  \
  }                                                                                      \
  void ko_zou_he_pressure_3d_##S(T* f, int64_t face, double r, int64_t nx, int64_t ny,      \
                                 int64_t nz) {                                           \
    ko::zou_he_pressure_3d(f, int(face), T(r), nx, ny, nz);                                   \
  }                                                                                      \
  void ko_stream_3d_##S(T* fo, const T* fi, int64_t nx, int64_t ny, int64_t nz, int px,  \
                        int py, int pz) {                                                \
    ko::stream_3d_flags(fo, fi, nx, ny, nz, px != 0, py != 0, pz != 0);                  \
  }                                                                                      \
  void ko_stream_periodic_xz_movingwall_y_3d_##S(T* fo, const T* fi, double ub,         \
                                                 double ut, double rw, int64_t nx,       \
                                                 int64_t ny, int64_t nz) {               \
    ko::stream_periodic_xz_movingwall_y_3d(fo, fi, T(ub), T(ut), T(rw), nx, ny, nz);     \
  }                                                                                      \
  void ko_collide_3d_##S(T* f, const uint8_t* s, int64_t nx, int64_t ny, int64_t nz,    \
                         double om) {                                                    \
    ko::collide_3d(f, s, nx, ny, nz, T(om));                                             \
  }                                                                                      \
  void ko_collide_trt_3d_##S(T* f, const uint8_t* s, int64_t nx, int64_t ny, int64_t nz,\
                             double sp, double sm) {                                     \
    ko::collide_trt_3d(f, s, nx, ny, nz, T(sp), T(sm));                                  \
  }                                                                                      \
  void ko_fused_trt_3d_##S(T* fo, const T* fi, int64_t nx, int64_t ny, int64_t nz,      \
                           int px, int py, int pz, double sp, double sm) {               \
    ko::fused_trt_3d(fo, fi, nx, ny, nz, px != 0, py != 0, pz != 0, T(sp), T(sm));       \
  }                                                                                      \
  void ko_macroscopic_3d_##S(T* r, T* u, T* v, T* w, const T* f, int64_t nx, int64_t ny,\
                             int64_t nz) {                                               \
    ko::macroscopic_3d(r, u, v, w, f, nx, ny, nz);                                       \
  }                                                                                      \
  void ko_collide_thermal_2d_##S(T* g, const T* u, const T* v, int64_t nx, int64_t ny,  \
                                 double om) {                                            \
    ko::collide_thermal_2d(g, u, v, nx, ny, T(om));                                      \
  }                                                                                      \
  void ko_compute_temperature_2d_##S(T* t, const T* g, int64_t nx, int64_t ny) {        \
    ko::compute_temperature_2d(t, g, nx, ny);                                            \
  }                                                                                      \
  void ko_apply_fixed_temp_2d_##S(T* g, const int* dir, const double* Tf, int64_t nx,   \
                                  int64_t ny) {                                          \
    bool b[4] = {dir[0] != 0, dir[1] != 0, dir[2] != 0, dir[3] != 0};                    \
    ko::apply_fixed_temp_2d(g, b, Tf, nx, ny);                                           \
  }                                                                                      \
  void ko_collide_boussinesq_2d_##S(T* f, const T* t, const uint8_t* s, int64_t nx,     \
                                    int64_t ny, double om, double bg, double tr) {       \
    ko::collide_boussinesq_2d(f, t, s, nx, ny, T(om), T(bg), T(tr));                     \
  }                                                                                      \
  void ko_fused_natconv_step_2d_##S(T* fo, const T* fi, T* go, const T* gi, T* t,       \
                                    int64_t nx, int64_t ny, double of, double ot,        \
                                    double bg, double tr, double th, double tc) {        \
    ko::fused_natconv_step(fo, fi, go, gi, t, nx, ny, T(of), T(ot), T(bg), T(tr), T(th), \
                           T(tc));                                                       \
  }                                                                                      \
  void ko_pull_generic_##S(const kmi_desc* d, T* fo, const T* fi) {                     \
    ko::pull_generic(*d, fo, fi);                                                        \
  }                                                                                      \
  /* Advance `nsteps` steps of the generic chain; f (and g) are updated in place. */     \
  void ko_run_##S(const kmi_desc* d, const uint8_t* solid, T* f, T* g, int64_t nsteps) { \
    const int Q = d->stencil == KMI_D2Q9 ? 9 : 19;                                       \
    const size_t n = size_t(d->nx * d->ny * (d->stencil == KMI_D2Q9 ? 1 : d->nz)) * Q;   \
    std::vector<T> b1(n), b2(d->thermal ? n : 0), scratch;                               \
    T *fa = f, *fb = b1.data(), *ga = g, *gb = d->thermal ? b2.data() : nullptr;         \
    for (int64_t s = 0; s < nsteps; ++s) ko::generic_step(*d, solid, fa, fb, ga, gb, scratch); \
    if (fa != f) std::memcpy(f, fa, n * sizeof(T));                                      \
    if (d->thermal && ga != g) std::memcpy(g, ga, n * sizeof(T));                        \
  }                                                                                      \
  /* as ko_run, with the per-node arrays of profile faces: prof[2*face + k] */           \
  void ko_run_profiles_##S(const kmi_desc* d, const uint8_t* solid, T* f,                \
                           const double* const* prof, int64_t nsteps) {                  \
    const int Q = d->stencil == KMI_D2Q9 ? 9 : 19;                                       \
    const size_t n = size_t(d->nx * d->ny * (d->stencil == KMI_D2Q9 ? 1 : d->nz)) * Q;   \
    std::vector<T> b1(n), scratch;                                                       \
    ko::Profiles pr;                                                                     \
    for (int k = 0; k < 8; ++k) pr.p[k / 2][k % 2] = prof[k];                            \
    T *fa = f, *fb = b1.data(), *ga = nullptr, *gb = nullptr;                            \
    for (int64_t s = 0; s < nsteps; ++s)                                                 \
      ko::generic_step(*d, solid, fa, fb, ga, gb, scratch, &pr);                         \
    if (fa != f) std::memcpy(f, fa, n * sizeof(T));                                      \
  }                                                                                      \
  void ko_collide_guo_2d_##S(T* f, const uint8_t* s, int64_t nx, int64_t ny, double om, \
                             double fx, double fy) {                                     \
    ko::collide_guo_2d(f, s, nx, ny, T(om), T(fx), T(fy));                               \
  }                                                                                      \
  void ko_macroscopic_forced_2d_##S(T* r, T* u, T* v, const T* f, int64_t nx,           \
                                    int64_t ny, double fx, double fy) {                  \
    ko::macroscopic_forced_2d(r, u, v, f, nx, ny, T(fx), T(fy));                         \
  }                                                                                      \
  void ko_zou_he_velocity_spatial_2d_##S(T* f, int64_t face, const double* ux,          \
                                         const double* uy, int64_t nx, int64_t ny) {     \
    if (face == KMI_NORTH) ko::zou_he_north_spatial(f, ux, uy, nx, ny);                  \
    else if (face == KMI_SOUTH) ko::zou_he_south_spatial(f, ux, uy, nx, ny);             \
    else if (face == KMI_WEST) ko::zou_he_west_spatial(f, ux, uy, nx, ny);               \
  }                                                                                      \
  void ko_zou_he_pressure_east_spatial_2d_##S(T* f, const double* rho, int64_t nx,      \
                                              int64_t ny) {                              \
    ko::zou_he_pressure_east_spatial(f, rho, nx, ny);                                    \
  }

#define KO_3D(S, T)                                                                       \
  void ko_collide_thermal_3d_##S(T* g, const T* u, const T* v, const T* w, int64_t nx,   \
                                 int64_t ny, int64_t nz, double om) {                    \
    ko::collide_thermal_3d(g, u, v, w, nx, ny, nz, T(om));                               \
  }                                                                                      \
  void ko_compute_temperature_3d_##S(T* t, const T* g, int64_t nx, int64_t ny,           \
                                     int64_t nz) {                                       \
    ko::compute_temperature_3d(t, g, nx, ny, nz);                                        \
  }                                                                                      \
  /* dir[6], Tf[6] in face order W, E, S, N, B, T */                                     \
  void ko_apply_fixed_temp_3d_##S(T* g, const int* dir, const double* Tf, int64_t nx,   \
                                  int64_t ny, int64_t nz) {                              \
    bool b[6];                                                                           \
    for (int a = 0; a < 6; ++a) b[a] = dir[a] != 0;                                      \
    ko::apply_fixed_temp_3d(g, b, Tf, nx, ny, nz);                                       \
  }                                                                                      \
  /* the single-face kernel with its own temperature on every edge (thermal_3d.jl:225) */\
  void ko_apply_fixed_temp_face_3d_##S(T* g, int64_t face, double Tw, int64_t nx,       \
                                       int64_t ny, int64_t nz) {                         \
    T t = T(Tw), te[4] = {t, t, t, t};                                                   \
    bool sk[4] = {false, false, false, false};                                           \
    ko::fixed_temp_face_3d(g, int(face), t, te, sk, nx, ny, nz);                         \
  }                                                                                      \
  void ko_collide_boussinesq_3d_##S(T* f, const T* t, const uint8_t* s, int64_t nx,     \
                                    int64_t ny, int64_t nz, double om, double bg,        \
                                    double tr) {                                         \
    ko::collide_boussinesq_3d(f, t, s, nx, ny, nz, T(om), T(bg), T(tr));                 \
  }                                                                                      \
  void ko_collide_guo_3d_##S(T* f, const uint8_t* s, int64_t nx, int64_t ny, int64_t nz, \
                             double om, double fx, double fy, double fz) {               \
    ko::collide_guo_3d(f, s, nx, ny, nz, T(om), T(fx), T(fy), T(fz));                    \
  }                                                                                      \
  void ko_collide_guo_field_3d_##S(T* f, const uint8_t* s, const T* fx, const T* fy,    \
                                   const T* fz, int64_t nx, int64_t ny, int64_t nz,      \
                                   double om) {                                          \
    ko::collide_guo_3d(f, s, nx, ny, nz, T(om), T(0), T(0), T(0), fx, fy, fz);           \
  }                                                                                      \
  void ko_collide_guo_field_2d_##S(T* f, const uint8_t* s, const T* fx, const T* fy,    \
                                   int64_t nx, int64_t ny, double om) {                  \
    ko::collide_guo_field_2d(f, s, nx, ny, T(om), fx, fy);                               \
  }                                                                                      \
  /* fxa == NULL: constant force (fx, fy, fz) */                                         \
  void ko_macroscopic_forced_3d_##S(T* r, T* u, T* v, T* w, const T* f, int64_t nx,     \
                                    int64_t ny, int64_t nz, double fx, double fy,        \
                                    double fz, const T* fxa, const T* fya,               \
                                    const T* fza) {                                      \
    ko::macroscopic_forced_3d(r, u, v, w, f, nx, ny, nz, T(fx), T(fy), T(fz), fxa, fya,  \
                              fza);                                                      \
  }                                                                                      \
  /* _trt_collide_local(_3d) of one node (specs.jl:7-108): f, o of 9 or 19 values */     \
  void ko_trt_collide_local_##S(const T* f, T* o, int64_t Q, double sp, double sm) {    \
    if (Q == 9) ko::trt_collide_local_2d(f, o, T(sp), T(sm));                            \
    else ko::trt_collide_local_3d(f, o, T(sp), T(sm));                                   \
  }                                                                                      \
  /* ko_run with a per-node force field (KMI_FORCE_GUO_FIELD); fz unused in 2D */        \
  void ko_run_field_##S(const kmi_desc* d, const uint8_t* solid, T* f, const T* fx,     \
                        const T* fy, const T* fz, int64_t nsteps) {                      \
    const int Q = d->stencil == KMI_D2Q9 ? 9 : 19;                                       \
    const size_t n = size_t(d->nx * d->ny * (d->stencil == KMI_D2Q9 ? 1 : d->nz)) * Q;   \
    std::vector<T> b1(n), scratch;                                                       \
    const T* ff[3] = {fx, fy, fz};                                                       \
    T *fa = f, *fb = b1.data(), *ga = nullptr, *gb = nullptr;                            \
    for (int64_t s = 0; s < nsteps; ++s)                                                 \
      ko::generic_step(*d, solid, fa, fb, ga, gb, scratch, nullptr, ff);                 \
    if (fa != f) std::memcpy(f, fa, n * sizeof(T));                                      \
  }

#define KO_LIBB(S, T)                                                                     \
  void ko_run_libb_##S(const kmi_desc* d, const uint8_t* solid, T* f, const T* qw,       \
                       const T* uwx, const T* uwy, const T* uwz, int64_t nsteps,          \
                       const double* west_prof) {                                        \
    const int Q = d->stencil == KMI_D2Q9 ? 9 : 19;                                       \
    const size_t n = size_t(d->nx * d->ny * (d->stencil == KMI_D2Q9 ? 1 : d->nz)) * Q;   \
    std::vector<T> b1(n);                                                                \
    const T* uw[3] = {uwx, uwy, uwz};                                                    \
    T *fa = f, *fb = b1.data();                                                          \
    for (int64_t s = 0; s < nsteps; ++s) ko::libb_step(*d, solid, fa, fb, qw, uw, west_prof); \
    if (fa != f) std::memcpy(f, fa, n * sizeof(T));                                      \
  }                                                                                      \
  void ko_drag_mei_2d_##S(const T* f, const T* qw, const T* uwx, const T* uwy,          \
                          int64_t nx, int64_t ny, double* out) {                         \
    ko::drag_mei_2d(f, qw, uwx, uwy, nx, ny, out);                                       \
  }                                                                                      \
  void ko_drag_halfway_##S(const kmi_desc* d, const T* f, const T* qw, double* out) {   \
    ko::drag_halfway(*d, f, qw, out);                                                    \
  }

KO_BOTH(KO_STREAM)
KO_BOTH(KO_3D)
KO_BOTH(KO_LIBB)

/* Hot-wall Nusselt number — thermal.jl:225-234 (west wall, rows 2..Ny-1), generalised to
 * the hot face `face`; `periodic_parallel` != 0 averages over every wall-parallel index. */
double ko_nusselt_d(const double* Temp, int64_t nx, int64_t ny, int face, double T_hot,
                    double T_cold, int periodic_parallel) {
  double dT = T_hot - T_cold;
  bool xface = face == KMI_WEST || face == KMI_EAST;
  int64_t npar = xface ? ny : nx;
  double H = double(xface ? nx : ny);
  double dx = 1.0;
  int64_t lo = periodic_parallel ? 0 : 1, hi = periodic_parallel ? npar : npar - 1;
  double sum = 0.0;
  for (int64_t p = lo; p < hi; ++p) {
    double t1, t2;
    if (face == KMI_WEST) { t1 = Temp[ko::I2(0, p, nx)]; t2 = Temp[ko::I2(1, p, nx)]; }
    else if (face == KMI_EAST) { t1 = Temp[ko::I2(nx - 1, p, nx)]; t2 = Temp[ko::I2(nx - 2, p, nx)]; }
    else if (face == KMI_SOUTH) { t1 = Temp[ko::I2(p, 0, nx)]; t2 = Temp[ko::I2(p, 1, nx)]; }
    else { t1 = Temp[ko::I2(p, ny - 1, nx)]; t2 = Temp[ko::I2(p, ny - 2, nx)]; }
    double nu = -H * (-8 * T_hot + 9 * t1 - t2) / (3 * dx) / dT;
    sum += nu;
  }
  return sum / double(hi - lo);
}

/* 3D hot-wall Nusselt number — thermal.jl:525-533 (hot face `face`, T_1 and T_2 the first
 * two nodes along the normal, H the cell count along it), averaged over the face without
 * its edge rows (a periodic tangent axis keeps all of them). */
double ko_nusselt_3d_d(const double* Temp, int64_t nx, int64_t ny, int64_t nz, int face,
                       double T_hot, double T_cold, int per_t1, int per_t2) {
  const int64_t n[3] = {nx, ny, nz};
  const int ax = face / 2, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
  const double H = double(n[ax]), dT = T_hot - T_cold;
  const int64_t lo1 = per_t1 ? 0 : 1, hi1 = per_t1 ? n[t1] : n[t1] - 1;
  const int64_t lo2 = per_t2 ? 0 : 1, hi2 = per_t2 ? n[t2] : n[t2] - 1;
  double sum = 0.0;
  for (int64_t b = lo2; b < hi2; ++b)
    for (int64_t a = lo1; a < hi1; ++a) {
      int64_t x1[3], x2[3];
      x1[t1] = x2[t1] = a;
      x1[t2] = x2[t2] = b;
      x1[ax] = (face % 2) ? n[ax] - 1 : 0;
      x2[ax] = (face % 2) ? n[ax] - 2 : 1;
      const double T1 = Temp[ko::I3(x1[0], x1[1], x1[2], nx, ny)];
      const double T2 = Temp[ko::I3(x2[0], x2[1], x2[2], nx, ny)];
      sum += -H * ((-8 * T_hot + 9 * T1 - T2) / 3.0) / dT;
    }
  const int64_t cnt = (hi1 - lo1) * (hi2 - lo2);
  return sum / double(cnt > 1 ? cnt : 1);
}

/* Compensated sum of an array (for mass / energy checks). */
double ko_ksum_d(const double* x, int64_t n) {
  ko::KSum s;
  for (int64_t i = 0; i < n; ++i) s.add(x[i]);
  return s.val();
}

int ko_omp_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
