// kmi_step_tb2d.hip — K time steps per launch for D2Q9 lattices (2, 4 or 8).
//
// Small 2D lattices (10^4–10^6 cells) do not fill the device: a single-step launch moves a
// few MB and its duration is the fixed cost of a dispatch (launch ramp, first-touch latency,
// the release of the written lines), 3–6 µs.  This kernel pays that cost once per K steps.
// A work-group loads its TW × TH tile plus a K-cell ring of the t populations into LDS,
// advances the whole region K steps in place (each step needs one cell less on every side),
// and writes the tile's t+K populations.  One work-item per cell of the region (1024), so
// each in-LDS step is one dependent chain per work-item with 16 waves per CU to hide it;
// the cost is the ring's re-evaluation.
//
// Scope: D2Q9 BGK / TRT / MRT with periodic, wall, moving-wall and constant Zou–He
// (velocity or pressure) faces, bounce-back or inert masks and Guo forcing (constant or
// per-node), and the thermal double distribution (THM: g beside f in LDS, Dirichlet /
// adiabatic walls, Boussinesq coupling split or fused; not in FP64) — the lattices of run_cavity_2d,
// run_poiseuille_2d, run_rayleigh_benard_2d and the .krk 2D loop without per-node faces.  Per-cell arithmetic and order are the single-step kernel's (kmi_cell.h),
// and every intermediate value passes through the storage rounding, so the result is
// bit-identical to K single-step launches in every storage.
#include "kmi_cell.h"
#include "kmi_internal.h"

namespace kmi {

template <int ST, int COLL, int TW, int TH, int K, bool THM>
__global__ void __launch_bounds__((TW + 2 * K) * (TH + 2 * K)) k_step_tb2d(const StepArgs a) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  using GS = Store<thermal_storage(ST)>;
  using GE = typename GS::elem;
  constexpr bool SH = S::shifted;
  constexpr int EW = TW + 2 * K, EH = TH + 2 * K, NE = EW * EH, NT = NE;
  constexpr int EPT = 1;
  __shared__ R L[(THM ? 18 : 9) * NE];
  R* const G = L + 9 * NE;  // thermal populations (THM)
  const GE* __restrict__ gi = static_cast<const GE*>(a.gi);
  GE* __restrict__ go = static_cast<GE*>(a.go);
  const int nx = a.nx, ny = a.ny;
  const long long P = a.plane;
  const E* __restrict__ fi = static_cast<const E*>(a.fi);
  E* __restrict__ fo = static_cast<E*>(a.fo);
  const int ox = int(blockIdx.x) * TW - K, oy = int(blockIdx.y) * TH - K;  // LDS origin
  const int tid = threadIdx.x;
  const bool perx = a.per[0] != 0, pery = a.per[1] != 0;
  // Unwrapped coordinates overshoot the lattice by at most K + TW (K < nx, ny is required).
  auto wx = [&](int u) { return u < 0 ? u + nx : (u >= nx ? u - nx : u); };
  auto wy = [&](int u) { return u < 0 ? u + ny : (u >= ny ? u - ny : u); };
  auto inside = [&](int ux, int uy) {
    return (perx || (ux >= 0 && ux < nx)) && (pery || (uy >= 0 && uy < ny));
  };

  for (int e = tid; e < NE; e += NT) {
    const int ux = ox + e % EW, uy = oy + e / EW;
    if (!inside(ux, uy)) continue;
    const long long self = wx(ux) + (long long)nx * wy(uy);
    sfor<9>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      L[q * NE + e] = S::ld(fi, self + q * P);
      if constexpr (THM) G[q * NE + e] = GS::ld(gi, self + q * P);
    });
  }
  __syncthreads();

  R nv[EPT][9], gv[EPT][9];
  for (int s = 1; s <= K; ++s) {
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      const int e = tid + r * NT;
      const int ex = e % EW, ey = e / EW;
      const int ux = ox + ex, uy = oy + ey;
      if (e >= NE || ex < s || ex >= EW - s || ey < s || ey >= EH - s || !inside(ux, uy)) continue;
      const int i = wx(ux), j = wy(uy);
      const long long self = i + (long long)nx * j;
      // a link whose source lies beyond a non-periodic face bounces back (k_step's rule)
      const bool vxm = perx || i > 0, vxp = perx || i < nx - 1;
      const bool vym = pery || j > 0, vyp = pery || j < ny - 1;
      R fp[9], fr[9], gp[9], gr[9];
      sfor<9>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = Lat<9>::cx[q], cy = Lat<9>::cy[q];
        constexpr int o = Lat<9>::opp[q];
        const bool ok = (cx > 0 ? vxm : (cx < 0 ? vxp : true)) && (cy > 0 ? vym : (cy < 0 ? vyp : true));
        if constexpr (THM) gp[q] = ok ? G[q * NE + e - cx - cy * EW] : G[o * NE + e];
        R v = ok ? L[q * NE + e - cx - cy * EW] : L[o * NE + e];
        if constexpr (q != 0) {
          if (a.any_moving && !ok) {
            auto add = [&](int fid) {
              const FaceArgs& F = a.face[fid];
              if (F.kind == KMI_FACE_MOVING_WALL)
                v = v + R(six_w(Lat<9>::w[q])) * R(F.rho) * cdotu<cx, cy, 0>(R(F.u[0]), R(F.u[1]), R(F.u[2]));
            };
            if (cx > 0 && !vxm) add(KMI_WEST);
            if (cx < 0 && !vxp) add(KMI_EAST);
            if (cy > 0 && !vym) add(KMI_SOUTH);
            if (cy < 0 && !vyp) add(KMI_NORTH);
          }
        }
        fp[q] = v;
      });
      for (int n = 0; n < a.n_zh; ++n) {
        const int fid = a.zh_face[n];
        const FaceArgs& F = a.face[fid];
        if (fid == KMI_NORTH && j == ny - 1) zh_north<SH>(fp, R(F.u[0]));
        else if (fid == KMI_SOUTH && j == 0) zh_south<SH>(fp, R(F.u[0]));
        else if (fid == KMI_WEST && i == 0) {
          if (F.kind == KMI_FACE_ZOUHE_VELOCITY) zh_west<SH>(fp, R(F.u[0]));
          else zh_pressure_west<SH>(fp, R(F.rho), R(F.rho - 1.0));
        } else if (fid == KMI_EAST && i == nx - 1) {
          zh_pressure_east<SH>(fp, R(F.rho), R(F.rho - 1.0));
        }
      }
      const bool solid = a.mask != nullptr && a.mask[self] == kCellSolid;
      if constexpr (THM) {
        // the thermal step of k_step: anti-bounce-back faces, T, Boussinesq force, thermal
        // BGK with the unforced (split) or forced (fused) u, Guo collision of f
        thermal_faces<SH>(gp, a, i, j);
        const R Tl = zeroth_moment<9>(gp);
        const R m0 = zeroth_moment<9>(fp);
        const R rho = SH ? R(1) + m0 : m0;
        const R inv = R(1) / rho;
        const R jx = first_moment<9, 0>(fp);
        const R jy = first_moment<9, 1>(fp);
        const R fy = SH ? R(a.beta_g) * Tl : R(a.beta_g) * (Tl - R(a.T_ref));
        const R fx = R(0);
        R uxf, uyf, uxt, uyt;
        if (a.coupling == KMI_COUPLING_SPLIT) {
          uxt = jx * inv;
          uyt = jy * inv;
          uxf = (jx + fx / R(2)) * inv;
          uyf = (jy + fy / R(2)) * inv;
        } else {
          uxf = jx * inv;
          uyf = (jy + fy / R(2)) * inv;
          uxt = uxf;
          uyt = uyf;
        }
        collide_thermal<SH>(gp, gr, uxt, uyt, R(a.omega_T), R(a.T_ref));
        if (solid) {
          sfor<9>([&](auto qc) { fr[decltype(qc)::value] = fp[Lat<9>::opp[decltype(qc)::value]]; });
        } else {
          collide_guo<SH>(fp, fr, m0, rho, uxf, uyf, fx, fy, R(a.omega));
        }
        sfor<9>([&](auto qc) { gv[r][decltype(qc)::value] = gr[decltype(qc)::value]; });
      } else if (solid && a.solid_inert) {
        sfor<9>([&](auto qc) { fr[decltype(qc)::value] = SH ? R(0) : R(Lat<9>::w[decltype(qc)::value]); });
      } else if (solid) {
        sfor<9>([&](auto qc) { fr[decltype(qc)::value] = fp[Lat<9>::opp[decltype(qc)::value]]; });
      } else if constexpr (COLL == KMI_BGK) {
        if (a.forcing) {
          R fx, fy, fz;
          node_force<9>(a, self, fx, fy, fz);
          const R m0 = zeroth_moment<9>(fp);
          const R rho = SH ? R(1) + m0 : m0;
          const R inv = R(1) / rho;
          const R ux2 = (first_moment<9, 0>(fp) + fx / R(2)) * inv;
          const R uy2 = (first_moment<9, 1>(fp) + fy / R(2)) * inv;
          collide_guo<SH>(fp, fr, m0, rho, ux2, uy2, fx, fy, R(a.omega));
        } else {
          collide_bgk<9, SH>(fp, fr, R(a.omega));
        }
      } else if constexpr (COLL == KMI_TRT) {
        collide_trt<9, SH>(fp, fr, R(a.s_plus), R(a.s_minus));
      } else {
        collide_mrt<SH>(fp, fr, R(a.s_e), R(a.s_eps), R(a.s_q), R(a.s_nu));
      }
      sfor<9>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        nv[r][q] = fr[q];
      });
    }
    if (s == K) break;
    __syncthreads();  // every step-s read of L is done
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      const int e = tid + r * NT;
      const int ex = e % EW, ey = e / EW;
      if (e >= NE || ex < s || ex >= EW - s || ey < s || ey >= EH - s || !inside(ox + ex, oy + ey)) continue;
      sfor<9>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        L[q * NE + e] = storage_round<ST>(nv[r][q]);
        if constexpr (THM) G[q * NE + e] = storage_round<thermal_storage(ST)>(gv[r][q]);
      });
    }
    __syncthreads();
  }

  // the tile's own cells: step t+K straight from registers
#pragma unroll
  for (int r = 0; r < EPT; ++r) {
    const int e = tid + r * NT;
    const int ex = e % EW, ey = e / EW;
    if (e >= NE || ex < K || ex >= K + TW || ey < K || ey >= K + TH) continue;
    const int ux = ox + ex, uy = oy + ey;
    if (ux >= nx || uy >= ny) continue;
    const long long self = ux + (long long)nx * uy;
    sfor<9>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      S::template st<false>(fo, self + q * P, nv[r][q]);
      if constexpr (THM) GS::template st<false>(go, self + q * P, gv[r][q]);
    });
  }
}

// ---- dispatch ---------------------------------------------------------------------------
// Variants: regions of 1024 cells, 1 = 28×28 tile, 2 steps; 2 = 24×24, 4 steps;
// 3 = 56×8, 4 steps; 4 = 16×16, 8 steps; 5 = 60×12, 2 steps; smaller regions, which spread
// a small lattice over more CUs: 6 = 8×8, 4 steps (256 cells); 7 = 16×16, 4 steps (576);
// 8 = 8×8, 8 steps (576).
bool tb2d_variant(int tile, int& TW, int& TH, int& K) {
  switch (tile) {
    case 1: TW = 28; TH = 28; K = 2; return true;
    case 2: TW = 24; TH = 24; K = 4; return true;
    case 3: TW = 56; TH = 8; K = 4; return true;
    case 4: TW = 16; TH = 16; K = 8; return true;
    case 5: TW = 60; TH = 12; K = 2; return true;
    case 6: TW = 8; TH = 8; K = 4; return true;
    case 7: TW = 16; TH = 16; K = 4; return true;
    case 8: TW = 8; TH = 8; K = 8; return true;
    default: return false;
  }
}

template <int ST, int COLL, int TW, int TH, int K>
static hipError_t launch_tb(const StepArgs& a, bool th, hipStream_t s) {
  const dim3 g(unsigned((a.nx + TW - 1) / TW), unsigned((a.ny + TH - 1) / TH)), b((TW + 2 * K) * (TH + 2 * K));
  if (th) {
    // the thermal chain collides f with Guo BGK; FP64 f and g exceed the registers of a
    // 1024-wide group, so FP64 thermal lattices keep single-step launches
    if constexpr (COLL == KMI_BGK && ST != KMI_STORE_F64)
      hipLaunchKernelGGL((k_step_tb2d<ST, COLL, TW, TH, K, true>), g, b, 0, s, a);
    else
      return hipErrorInvalidValue;
  } else {
    hipLaunchKernelGGL((k_step_tb2d<ST, COLL, TW, TH, K, false>), g, b, 0, s, a);
  }
  return hipGetLastError();
}

template <int ST, int COLL>
static hipError_t launch_tb_tile(int tile, const StepArgs& a, bool th, hipStream_t s) {
  switch (tile) {
    case 1: return launch_tb<ST, COLL, 28, 28, 2>(a, th, s);
    case 2: return launch_tb<ST, COLL, 24, 24, 4>(a, th, s);
    case 3: return launch_tb<ST, COLL, 56, 8, 4>(a, th, s);
    case 4: return launch_tb<ST, COLL, 16, 16, 8>(a, th, s);
    case 5: return launch_tb<ST, COLL, 60, 12, 2>(a, th, s);
    case 6: return launch_tb<ST, COLL, 8, 8, 4>(a, th, s);
    case 7: return launch_tb<ST, COLL, 16, 16, 4>(a, th, s);
    case 8: return launch_tb<ST, COLL, 8, 8, 8>(a, th, s);
    default: return hipErrorInvalidValue;
  }
}

template <int ST>
static hipError_t launch_tb_coll(int coll, int tile, const StepArgs& a, bool th, hipStream_t s) {
  if (th) return launch_tb_tile<ST, KMI_BGK>(tile, a, true, s);
  if (coll == KMI_MRT) return launch_tb_tile<ST, KMI_MRT>(tile, a, false, s);
  if (coll == KMI_TRT) return launch_tb_tile<ST, KMI_TRT>(tile, a, false, s);
  return launch_tb_tile<ST, KMI_BGK>(tile, a, false, s);
}

hipError_t launch_step_tb2d(int st, int coll, int tile, bool th, const StepArgs& a, hipStream_t s) {
  switch (st) {
    case KMI_STORE_F64: return launch_tb_coll<KMI_STORE_F64>(coll, tile, a, th, s);
    case KMI_STORE_F32: return launch_tb_coll<KMI_STORE_F32>(coll, tile, a, th, s);
    case KMI_STORE_F32S: return launch_tb_coll<KMI_STORE_F32S>(coll, tile, a, th, s);
    default: return launch_tb_coll<KMI_STORE_F16S>(coll, tile, a, th, s);
  }
}

}  // namespace kmi
