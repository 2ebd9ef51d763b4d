// kmi_common.h — lattice tables, storage traits and compile-time helpers shared by the
// MI355X stream–collide kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <utility>

#include "kraken_mi355x.h"

#define KMI_DEV __device__ __forceinline__
#define KMI_HD __host__ __device__ __forceinline__

namespace kmi {

// ---- compile-time loop --------------------------------------------------------------
template <class F, int... I>
KMI_HD void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
KMI_HD void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// ---- lattices (Kraken ordering, 0-based: src/lattice/d2q9.jl:17-29, d3q19.jl:13-47) --
template <int Q>
struct Lat;

template <>
struct Lat<9> {
  static constexpr int D = 2;
  static constexpr int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr double w[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

template <>
struct Lat<19> {
  static constexpr int D = 3;
  static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr double w[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                   1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

// ---- compensated FP64 sums (Neumaier): a running sum s with its error term c, value s + c.
// Built with -ffp-contract=off, so the error terms are exact.
KMI_HD void nsum_add(double& s, double& c, double x) {
  const double t = s + x;
  c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
  s = t;
}
// Fold the compensated pair (s2, c2) into (s, c).
KMI_HD void nsum_merge(double& s, double& c, double s2, double c2) {
  nsum_add(s, c, s2);
  c += c2;
}

// 6 w_q of the Ladd moving-wall term, as the reference writes it (T(1/6) for w = 1/36).
KMI_HD constexpr double six_w(double w) {
  return w == 1.0 / 36.0 ? 1.0 / 6.0 : (w == 1.0 / 18.0 ? 1.0 / 3.0 : (w == 1.0 / 9.0 ? 2.0 / 3.0 : 2.0));
}

// c·u in the reference's written form: the first non-zero component carries the sign
// of c, the others are added or subtracted in x, y, z order (`-ux + uy`, `ux - uz`).
template <int CX, int CY, int CZ, class R>
KMI_DEV R cdotu(R ux, R uy, R uz) {
  if constexpr (CX != 0) {
    R acc = CX > 0 ? ux : -ux;
    if constexpr (CY != 0) acc = CY > 0 ? acc + uy : acc - uy;
    if constexpr (CZ != 0) acc = CZ > 0 ? acc + uz : acc - uz;
    return acc;
  } else if constexpr (CY != 0) {
    R acc = CY > 0 ? uy : -uy;
    if constexpr (CZ != 0) acc = CZ > 0 ? acc + uz : acc - uz;
    return acc;
  } else if constexpr (CZ != 0) {
    return CZ > 0 ? uz : -uz;
  } else {
    return R(0);
  }
}

template <int Q, int AXIS>
KMI_HD constexpr int cval(int q) {
  return AXIS == 0 ? Lat<Q>::cx[q] : (AXIS == 1 ? Lat<Q>::cy[q] : Lat<Q>::cz[q]);
}
template <int Q, int AXIS>
KMI_HD constexpr int first_nonzero() {
  for (int q = 0; q < Q; ++q)
    if (cval<Q, AXIS>(q) != 0) return q;
  return Q;
}

// Σ_q c_{q,a} f_q in q order, first term carrying its sign — the written order of
// equilibrium_helpers.jl:54-61 and equilibrium_helpers_3d.jl:102-112.
template <int Q, int AXIS, class R>
KMI_DEV R first_moment(const R (&f)[Q]) {
  constexpr int q0 = first_nonzero<Q, AXIS>();
  static_assert(q0 < Q, "axis without velocities");
  R acc = cval<Q, AXIS>(q0) > 0 ? f[q0] : -f[q0];
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int c = cval<Q, AXIS>(q);
    if constexpr (q > q0 && c != 0) acc = c > 0 ? acc + f[q] : acc - f[q];
  });
  return acc;
}

template <int Q, class R>
KMI_DEV R zeroth_moment(const R (&f)[Q]) {
  R acc = f[0];
  sfor<Q - 1>([&](auto qc) { acc = acc + f[decltype(qc)::value + 1]; });
  return acc;
}

// ---- population storage -------------------------------------------------------------
// F64 / F32 store f itself; F32S / F16S store the deviation d = f - w_q (thermal:
// g - w_q T_ref).  F16S keeps d * 2^15 as IEEE binary16 (round to nearest even).
template <int S>
struct Store;

// st<true>: non-temporal store (streams past the caches; pays off when the populations
// do not fit in the last-level cache).
template <class E>
KMI_DEV void store_elem(E* p, size_t i, E v, bool nt) {
  if (nt) __builtin_nontemporal_store(v, p + i);
  else p[i] = v;
}

template <>
struct Store<KMI_STORE_F64> {
  using elem = double;
  using real = double;
  static constexpr bool shifted = false;
  KMI_DEV static real ld(const elem* p, size_t i) { return p[i]; }
  template <bool NT = false>
  KMI_DEV static void st(elem* p, size_t i, real v) { store_elem(p, i, v, NT); }
};

template <>
struct Store<KMI_STORE_F32> {
  using elem = float;
  using real = float;
  static constexpr bool shifted = false;
  KMI_DEV static real ld(const elem* p, size_t i) { return p[i]; }
  template <bool NT = false>
  KMI_DEV static void st(elem* p, size_t i, real v) { store_elem(p, i, v, NT); }
};

template <>
struct Store<KMI_STORE_F32S> {
  using elem = float;
  using real = float;
  static constexpr bool shifted = true;
  KMI_DEV static real ld(const elem* p, size_t i) { return p[i]; }
  template <bool NT = false>
  KMI_DEV static void st(elem* p, size_t i, real v) { store_elem(p, i, v, NT); }
};

constexpr float kF16Scale = 32768.0f;
constexpr float kF16InvScale = 1.0f / 32768.0f;

template <>
struct Store<KMI_STORE_F16S> {
  using elem = _Float16;
  using real = float;
  static constexpr bool shifted = true;
  KMI_DEV static real ld(const elem* p, size_t i) { return float(p[i]) * kF16InvScale; }
  template <bool NT = false>
  KMI_DEV static void st(elem* p, size_t i, real v) {
    store_elem(p, i, _Float16(v * kF16Scale), NT);
  }
};

inline size_t storage_bytes(int s) {
  return s == KMI_STORE_F64 ? 8 : (s == KMI_STORE_F16S ? 2 : 4);
}

// Storage of the thermal populations g.  Their deviations from w_q T_ref are O(w_q ΔT) while
// a step changes them by O(1e-6): binary16's 11-bit significand rounds such increments away
// and the temperature field freezes, so an F16S lattice keeps g in F32S (108 B per D2Q9
// cell update instead of 144 B for F32).
constexpr int thermal_storage(int s) { return s == KMI_STORE_F16S ? KMI_STORE_F32S : s; }

template <class T>
struct TypeTag {
  using type = T;
};

// ---- kernel arguments ---------------------------------------------------------------
struct FaceArgs {
  int kind;      // kmi_face_kind
  int dirichlet; // thermal Dirichlet
  double u[3];
  double rho;
  double T;
  // Corner (2D: entries 0, 1) or edge (3D: 0..3) link temperatures of the Dirichlet
  // face, in the order of the face's edge slots, and whether an adjacent face owns the
  // slot instead (apply_fixed_temp_2d corner rule, apply_fixed_temp_3d edge rule).
  double Tc[4];
  int skip[4];
  const double* prof[2];  // *_PROFILE faces: per-node (u_x, u_y) or (rho, -) on the device
};

struct StepArgs {
  const void* fi;
  void* fo;
  const void* gi;
  void* go;
  const uint8_t* mask;
  int nx, ny, nz;
  long long plane;  // elements between consecutive q planes
  int per[3];       // periodic axes
  int any_moving;   // some face is a MOVING_WALL
  int n_zh;         // Zou–He closures, applied in zh_face[] order
  int zh_face[6];
  double omega, s_plus, s_minus, s_e, s_eps, s_q, s_nu;
  double omega_T, beta_g, T_ref;
  int coupling;
  int forcing;      // 1: Guo body force (any kmi_forcing other than NONE)
  double fx, fy, fz;  // constant body force (KMI_FORCE_GUO)
  // Per-node body force (KMI_FORCE_GUO_FIELD): planes x, y(, z) of `plane` values in the
  // collision arithmetic (Store<ST>::real); nullptr: the constant force applies.
  const void* ffield;
  int solid_inert;  // masked cells hold the rest equilibrium (KMI_SOLID_INERT)
  // Cut links: mask value 2 marks a fluid cell with cut links; cut_slot[cell] indexes its
  // record in cut_rec: Q wall fractions (0: link not cut), then Q × 3 wall velocities.
  const int* cut_slot;
  const double* cut_rec;
  // Walled multi-step launches: work-group b marches tile tile_order[b] (-1: padding);
  // nullptr: the XCD-banded order of the periodic launches.
  const int* tile_order;
  FaceArgs face[6];
};

// Body force seen by the moment kernels: constant f[3], or per-node values (`field`, planes
// x, y(, z) in Store<ST>::real, as StepArgs::ffield) when field != nullptr.
struct MomForce {
  double f[3];
  const void* field;
};

// Cell classes held in the mask byte.
enum : unsigned char { kCellFluid = 0, kCellSolid = 1, kCellCut = 2 };

}  // namespace kmi
