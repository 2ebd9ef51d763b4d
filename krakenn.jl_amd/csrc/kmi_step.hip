// kmi_step.hip — the single-pass stream–collide kernel for CDNA4 (gfx950).
//
// One work-item per cell, one launch per time step, A-B buffers in structure-of-arrays
// layout (one contiguous plane per velocity, x fastest — the reference's f[i,j,(k,)q]).
// Each work-item pulls its Q populations (halfway bounce-back / Ladd moving wall /
// periodic wrap folded into the address selection), applies the Zou–He closures of the
// faces it lies on, collides (BGK / TRT / MRT, or thermal DDF + Guo Boussinesq) and
// writes Q populations: 2·Q·sizeof(elem) bytes of HBM traffic per cell update.
//
// FP64 and FP32 storage evaluate every expression in the reference's written order
// (compiled with -ffp-contract=off) so they reproduce the reference kernels bit for bit:
//   pull + BB          src/kernels/collide_stream_2d.jl:75-101, collide_stream_3d.jl:5-43,
//                      stream_periodic_2d.jl:5-98, stream_periodic_3d.jl:24-166
//   BGK                collide_stream_2d.jl:105-131, collide_stream_3d.jl:47-97
//   TRT                fused_trt_2d.jl:101-173, dsl/bricks_3d.jl:168-229
//   MRT                collide_mrt_2d.jl:43-109
//   Zou–He             boundary_2d.jl:13-44, 135-185, 206-281; boundary_3d.jl:152-290
//   thermal DDF        thermal_2d.jl:16-60, 140-290, 421-480; fused_thermal_2d.jl:10-207;
//                      thermal_3d.jl:12-136, 164-535, 559-712
//   Guo forcing        collide_guo_2d.jl:5-140, collide_guo_3d.jl:5-325
//   face rebuild       bc/rebuild_2d.jl:1-50, bc/rebuild_3d.jl:1-89, bc/specs.jl:7-108
// F32S / F16S store deviations from the rest state and collide them in FP32 (the
// algebra is identical; only the rounding differs).
#include "kmi_cell.h"

namespace kmi {

// ---- the step kernel ----------------------------------------------------------------------
// BND = false: every face periodic, no mask, no closures (the Taylor–Green fast path).
// NTS: non-temporal population stores.  Work-group = bx × by cells of one z-plane.
template <int Q, int ST, int COLL, bool TH, bool BND, bool NTS>
__global__ void __launch_bounds__(256) k_step(const StepArgs a) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  constexpr bool SH = S::shifted;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= a.nx || j >= a.ny) return;
  const int k = (Q == 19) ? int(blockIdx.z) : 0;
  const int nx = a.nx, ny = a.ny;
  const long long P = a.plane;
  const E* __restrict__ fi = static_cast<const E*>(a.fi);
  E* __restrict__ fo = static_cast<E*>(a.fo);

  // Source coordinate of a +1 / −1 velocity component on each axis, and whether that
  // source exists (false: the link leaves through a non-periodic face).
  int xm = i - 1, xp = i + 1, ym = j - 1, yp = j + 1, zm = k - 1, zp = k + 1;
  bool vxm = true, vxp = true, vym = true, vyp = true, vzm = true, vzp = true;
  if constexpr (!BND) {
    xm = xm < 0 ? nx - 1 : xm;
    xp = xp >= nx ? 0 : xp;
    ym = ym < 0 ? ny - 1 : ym;
    yp = yp >= ny ? 0 : yp;
    if constexpr (Q == 19) {
      zm = zm < 0 ? a.nz - 1 : zm;
      zp = zp >= a.nz ? 0 : zp;
    }
  } else {
    if (xm < 0) { if (a.per[0]) xm = nx - 1; else { vxm = false; xm = i; } }
    if (xp >= nx) { if (a.per[0]) xp = 0; else { vxp = false; xp = i; } }
    if (ym < 0) { if (a.per[1]) ym = ny - 1; else { vym = false; ym = j; } }
    if (yp >= ny) { if (a.per[1]) yp = 0; else { vyp = false; yp = j; } }
    if constexpr (Q == 19) {
      if (zm < 0) { if (a.per[2]) zm = a.nz - 1; else { vzm = false; zm = k; } }
      if (zp >= a.nz) { if (a.per[2]) zp = 0; else { vzp = false; zp = k; } }
    }
  }
  const int self = i + nx * (j + ny * k);

  // tag: TypeTag<Store<…>> of the field pulled (the thermal populations may be stored
  // differently from the flow populations).
  auto pull = [&](auto tag, const auto* __restrict__ src, R (&out)[Q], bool ladd) {
    using SS = typename decltype(tag)::type;
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = Lat<Q>::cx[q], cy = Lat<Q>::cy[q], cz = Lat<Q>::cz[q];
      constexpr int o = Lat<Q>::opp[q];
      const int sx = cx > 0 ? xm : (cx < 0 ? xp : i);
      const int sy = cy > 0 ? ym : (cy < 0 ? yp : j);
      const int sz = cz > 0 ? zm : (cz < 0 ? zp : k);
      const bool ok = (cx > 0 ? vxm : (cx < 0 ? vxp : true)) &&
                      (cy > 0 ? vym : (cy < 0 ? vyp : true)) &&
                      (cz > 0 ? vzm : (cz < 0 ? vzp : true));
      const long long addr =
          ok ? (long long)(sx + nx * (sy + ny * sz)) + q * P : (long long)self + o * P;
      R v = SS::ld(src, addr);
      if constexpr (BND && q != 0) {
        if (ladd && !ok) {
          // Ladd term of every crossed moving face, x then y then z.
          auto add = [&](int fid) {
            const FaceArgs& F = a.face[fid];
            if (F.kind == KMI_FACE_MOVING_WALL)
              v = v + R(six_w(Lat<Q>::w[q])) * R(F.rho) *
                          cdotu<cx, cy, cz>(R(F.u[0]), R(F.u[1]), R(F.u[2]));
          };
          if (cx > 0 && !vxm) add(KMI_WEST);
          if (cx < 0 && !vxp) add(KMI_EAST);
          if (cy > 0 && !vym) add(KMI_SOUTH);
          if (cy < 0 && !vyp) add(KMI_NORTH);
          if (cz > 0 && !vzm) add(KMI_BOTTOM);
          if (cz < 0 && !vzp) add(KMI_TOP);
        }
      }
      out[q] = v;
    });
  };

  R fp[Q], fo_r[Q];
  pull(TypeTag<S>{}, fi, fp, BND && a.any_moving);

  if constexpr (BND && Q == 9) {
    // Face closures in the configured order.  Profile faces leave their two corner nodes
    // to the bounce-back pull (boundary_spatial_2d.jl: `if 1 < i < N`).
    for (int n = 0; n < a.n_zh; ++n) {
      const int fid = a.zh_face[n];
      const FaceArgs& F = a.face[fid];
      const bool prof = F.kind == KMI_FACE_ZOUHE_VELOCITY_PROFILE;
      if (fid == KMI_NORTH && j == ny - 1) {
        if (!prof) zh_north<SH>(fp, R(F.u[0]));
        else if (i > 0 && i < nx - 1) zh_north_profile<SH>(fp, R(F.prof[0][i]), R(F.prof[1][i]));
      } else if (fid == KMI_SOUTH && j == 0) {
        if (!prof) zh_south<SH>(fp, R(F.u[0]));
        else if (i > 0 && i < nx - 1) zh_south_profile<SH>(fp, R(F.prof[0][i]), R(F.prof[1][i]));
      } else if (fid == KMI_WEST && i == 0) {
        if (F.kind == KMI_FACE_ZOUHE_VELOCITY) zh_west<SH>(fp, R(F.u[0]));
        else if (F.kind == KMI_FACE_ZOUHE_PRESSURE) zh_pressure_west<SH>(fp, R(F.rho), R(F.rho - 1.0));
        else if (j > 0 && j < ny - 1) zh_west_profile<SH>(fp, R(F.prof[0][j]), R(F.prof[1][j]));
      } else if (fid == KMI_EAST && i == nx - 1) {
        if (F.kind == KMI_FACE_ZOUHE_PRESSURE) {
          zh_pressure_east<SH>(fp, R(F.rho), R(F.rho - 1.0));
        } else if (j > 0 && j < ny - 1) {
          const double r = F.prof[0][j];  // zou_he_pressure_east_spatial_2d_kernel!
          zh_pressure_east<SH>(fp, R(r), R(r - 1.0));
        }
      }
    }
  }
  if constexpr (BND && Q == 19) {
    // Face closures in the configured order; an edge node takes each of its faces in turn,
    // as the reference's one-launch-per-face sequence does.
    for (int n = 0; n < a.n_zh; ++n) {
      const int fid = a.zh_face[n];
      const FaceArgs& F = a.face[fid];
      switch (fid) {
        case KMI_WEST: if (i == 0) zh3_apply<SH, 0>(fp, F); break;
        case KMI_EAST: if (i == nx - 1) zh3_apply<SH, 1>(fp, F); break;
        case KMI_SOUTH: if (j == 0) zh3_apply<SH, 2>(fp, F); break;
        case KMI_NORTH: if (j == ny - 1) zh3_apply<SH, 3>(fp, F); break;
        case KMI_BOTTOM: if (k == 0) zh3_apply<SH, 4>(fp, F); break;
        default: if (k == a.nz - 1) zh3_apply<SH, 5>(fp, F); break;
      }
    }
  }

  // Rebuilt face node (apply_bc_rebuild_2d!/_3d!): the west / east face off its edges.
  int rb = 0;
  if constexpr (BND) {
    const bool inner = j > 0 && j < ny - 1 && (Q == 9 || (k > 0 && k < a.nz - 1));
    if (inner && i == 0 && a.face[KMI_WEST].kind == KMI_FACE_ZOUHE_VELOCITY_REBUILD) rb = 1;
    if (inner && i == nx - 1 && a.face[KMI_EAST].kind == KMI_FACE_ZOUHE_PRESSURE_REBUILD) rb = 2;
  }
  const unsigned char cls = (BND && a.mask != nullptr) ? a.mask[self] : kCellFluid;
  const bool solid = cls == kCellSolid;
  if constexpr (BND) {
    if (cls == kCellCut && !rb) libb_prephase<Q, S>(fp, fi, self, P, a.cut_rec + (long long)a.cut_slot[self] * (4 * Q));
  }

  if constexpr (TH) {
    using GS = Store<thermal_storage(ST)>;
    using GE = typename GS::elem;
    static_assert(std::is_same_v<typename GS::real, R> && GS::shifted == SH, "g and f share arithmetic");
    const GE* __restrict__ gi = static_cast<const GE*>(a.gi);
    GE* __restrict__ go = static_cast<GE*>(a.go);
    R gp[Q], go_r[Q];
    pull(TypeTag<GS>{}, gi, gp, false);
    if constexpr (BND) {
      if constexpr (Q == 9) thermal_faces<SH>(gp, a, i, j);
      else thermal_faces3<SH>(gp, a, i, j, k);
    }
    const R Tl = zeroth_moment<Q>(gp);  // T (shifted: T − T0)
    const R m0 = zeroth_moment<Q>(fp);
    const R rho = SH ? R(1) + m0 : m0;
    const R inv = R(1) / rho;
    const R jx = first_moment<Q, 0>(fp);
    const R jy = first_moment<Q, 1>(fp);
    const R fy = SH ? R(a.beta_g) * Tl : R(a.beta_g) * (Tl - R(a.T_ref));
    const R fx = R(0);
    if constexpr (Q == 9) {
      R uxf, uyf, uxt, uyt;
      if (a.coupling == KMI_COUPLING_SPLIT) {
        uxt = jx * inv;  // compute_macroscopic_2d! (thermal.jl:124)
        uyt = jy * inv;
        uxf = (jx + fx / R(2)) * inv;  // collide_boussinesq_2d_kernel! (thermal_2d.jl:440-441)
        uyf = (jy + fy / R(2)) * inv;
      } else {
        uxf = jx * inv;  // macroscopic_boussinesq (fused_thermal_2d.jl:67-69)
        uyf = (jy + fy / R(2)) * inv;
        uxt = uxf;
        uyt = uyf;
      }
      collide_thermal<SH>(gp, go_r, uxt, uyt, R(a.omega_T), R(a.T_ref));
      if (solid) {
        sfor<9>([&](auto qc) { fo_r[decltype(qc)::value] = fp[Lat<9>::opp[decltype(qc)::value]]; });
      } else {
        collide_guo<SH>(fp, fo_r, m0, rho, uxf, uyf, fx, fy, R(a.omega));
      }
    } else {
      // run_natural_convection_3d (thermal.jl:500-515): the thermal collision sees
      // compute_macroscopic_3d!'s u, the Boussinesq collision the F/2-shifted one
      // (thermal_3d.jl:593-596), F = (0, β_g (T − T_ref), 0).
      const R jz = first_moment<Q, 2>(fp);
      const R fz = R(0);
      collide_thermal19<SH>(gp, go_r, jx * inv, jy * inv, jz * inv, R(a.omega_T), R(a.T_ref));
      if (solid) {
        sfor<Q>([&](auto qc) { fo_r[decltype(qc)::value] = fp[Lat<Q>::opp[decltype(qc)::value]]; });
      } else {
        collide_guo19<SH>(fp, fo_r, m0, rho, (jx + fx / R(2)) * inv, (jy + fy / R(2)) * inv,
                          (jz + fz / R(2)) * inv, fx, fy, fz, R(a.omega));
      }
    }
    sfor<Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      GS::template st<NTS>(go, self + q * P, go_r[q]);
    });
  } else {
    if (rb) {
      const R v = rb == 1 ? R(a.face[KMI_WEST].prof[0][j + (Q == 19 ? ny * k : 0)]) : R(a.face[KMI_EAST].rho);
      rebuild_node<Q, SH>(fp, fo_r, rb, v, R(a.s_plus), R(a.s_minus));
    } else if (solid && a.solid_inert) {
      // SolidInert (dsl/bricks.jl:133-154): rest equilibrium, w_q (0 as a deviation)
      sfor<Q>([&](auto qc) { fo_r[decltype(qc)::value] = SH ? R(0) : R(Lat<Q>::w[decltype(qc)::value]); });
    } else if (solid) {
      sfor<Q>([&](auto qc) { fo_r[decltype(qc)::value] = fp[Lat<Q>::opp[decltype(qc)::value]]; });
    } else if constexpr (COLL == KMI_BGK) {
      bool forced = false;
      if constexpr (BND) {
        if (a.forcing) {
          // collide_guo_2d_kernel! / collide_guo_3d_kernel! and their per-node-field twins
          // (collide_guo_2d.jl:5-140, collide_guo_3d.jl:5-325): u = (Σcf + F/2)/ρ
          forced = true;
          R fx, fy, fz;
          node_force<Q>(a, self, fx, fy, fz);
          const R m0 = zeroth_moment<Q>(fp);
          const R rho = SH ? R(1) + m0 : m0;
          const R inv = R(1) / rho;
          const R ux = (first_moment<Q, 0>(fp) + fx / R(2)) * inv;
          const R uy = (first_moment<Q, 1>(fp) + fy / R(2)) * inv;
          if constexpr (Q == 9) {
            collide_guo<SH>(fp, fo_r, m0, rho, ux, uy, fx, fy, R(a.omega));
          } else {
            const R uz = (first_moment<Q, 2>(fp) + fz / R(2)) * inv;
            collide_guo19<SH>(fp, fo_r, m0, rho, ux, uy, uz, fx, fy, fz, R(a.omega));
          }
        }
      }
      if (!forced) collide_bgk<Q, SH>(fp, fo_r, R(a.omega));
    } else if constexpr (COLL == KMI_TRT) {
      collide_trt<Q, SH>(fp, fo_r, R(a.s_plus), R(a.s_minus));
    } else {
      static_assert(Q == 9 || COLL != KMI_MRT, "MRT is D2Q9");
      if constexpr (Q == 9)
        collide_mrt<SH>(fp, fo_r, R(a.s_e), R(a.s_eps), R(a.s_q), R(a.s_nu));
    }
  }
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    S::template st<NTS>(fo, self + q * P, fo_r[q]);
  });
}

// ---- dispatch ---------------------------------------------------------------------------
template <int Q, int ST, int COLL, bool TH, bool BND>
static hipError_t launch(bool nts, const StepArgs& a, dim3 g, dim3 b, hipStream_t s) {
  if (nts) hipLaunchKernelGGL((k_step<Q, ST, COLL, TH, BND, true>), g, b, 0, s, a);
  else hipLaunchKernelGGL((k_step<Q, ST, COLL, TH, BND, false>), g, b, 0, s, a);
  return hipGetLastError();
}

template <int Q, int ST>
static hipError_t launch_st(int coll, bool th, bool bnd, bool nts, const StepArgs& a, dim3 g,
                            dim3 b, hipStream_t s) {
  if (th) return bnd ? launch<Q, ST, KMI_BGK, true, true>(nts, a, g, b, s)
                     : launch<Q, ST, KMI_BGK, true, false>(nts, a, g, b, s);
  if constexpr (Q == 9) {
    if (coll == KMI_MRT) return bnd ? launch<9, ST, KMI_MRT, false, true>(nts, a, g, b, s)
                                    : launch<9, ST, KMI_MRT, false, false>(nts, a, g, b, s);
  }
  if (coll == KMI_TRT) return bnd ? launch<Q, ST, KMI_TRT, false, true>(nts, a, g, b, s)
                                  : launch<Q, ST, KMI_TRT, false, false>(nts, a, g, b, s);
  return bnd ? launch<Q, ST, KMI_BGK, false, true>(nts, a, g, b, s)
             : launch<Q, ST, KMI_BGK, false, false>(nts, a, g, b, s);
}

template <int Q>
static hipError_t launch_q(int st, int coll, bool th, bool bnd, bool nts, const StepArgs& a,
                           dim3 g, dim3 b, hipStream_t s) {
  switch (st) {
    case KMI_STORE_F64: return launch_st<Q, KMI_STORE_F64>(coll, th, bnd, nts, a, g, b, s);
    case KMI_STORE_F32: return launch_st<Q, KMI_STORE_F32>(coll, th, bnd, nts, a, g, b, s);
    case KMI_STORE_F32S: return launch_st<Q, KMI_STORE_F32S>(coll, th, bnd, nts, a, g, b, s);
    default: return launch_st<Q, KMI_STORE_F16S>(coll, th, bnd, nts, a, g, b, s);
  }
}

hipError_t launch_step(int q, int st, int coll, bool th, bool bnd, bool nts, const StepArgs& a,
                       dim3 g, dim3 b, hipStream_t s) {
  return q == 9 ? launch_q<9>(st, coll, th, bnd, nts, a, g, b, s)
                : launch_q<19>(st, coll, th, bnd, nts, a, g, b, s);
}

}  // namespace kmi
