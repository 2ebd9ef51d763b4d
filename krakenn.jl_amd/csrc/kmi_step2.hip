// kmi_step2.hip — two or three time steps per launch for D3Q19 BGK/TRT lattices: fully
// periodic ones (below), and walled ones (halfway and Ladd walls, Zou–He faces, solid masks;
// the BND variants further down), with the three-step march after the two-step one.
//
// A single-step kernel reads and writes every population once per step: 2·19·sizeof(elem)
// bytes of HBM traffic per cell update, and no kernel of that shape can beat the device's
// copy bandwidth.  This kernel advances TWO steps per pass: each work-group owns a column
// of TX × TY cells and a range of z-planes, marches up z, and keeps the intermediate
// (t+1) populations in LDS, so HBM sees one read and one write per population per two
// steps.  Cost: the t+1 step is also evaluated on a one-cell x–y halo ((TX+2)(TY+2)
// cells per plane) and on the two planes below and above the z-range.
//
// Per marched plane s:
//   phase 1  t+1 populations of plane s on the (TX+2)×(TY+2) extended tile: pull from the
//            t field in HBM (periodic wrap), collide; the c_z = −1 populations go to LDS at
//            once, the c_z = 0 and +1 populations stay in registers;
//   phase 2  t+2 populations of plane s−1 on the TX×TY tile: pull from LDS — c_z = +1
//            populations of plane s−2, c_z = 0 of s−1, c_z = −1 of s — collide, store to HBM;
//   phase 3  the registers kept in phase 1 replace the LDS values phase 2 has consumed
//            (c_z = 0 of s−1, c_z = +1 of s−2).
// LDS per extended cell: 9 (c_z = 0, one plane) + 10 (c_z = +1, two planes) + 5 (c_z = −1)
// = 24 values; two barriers per plane (after phases 1 and 2; phase 3's writes are read only
// after the next phase-1 barrier, and phase 1 writes only the c_z = −1 slot, whose readers
// finished before the phase-2 barrier).
//
// The arithmetic per cell and step is the single-step kernel's (same collide functions,
// -ffp-contract=off), and the t+1 values pass through the storage rounding before they
// enter LDS, so the result is bit-identical to two single-step launches in every storage.
#include <algorithm>
#include <vector>

#include "kmi_cell.h"

namespace kmi {

namespace tb {
// D3Q19 populations grouped by c_z (Kraken ordering, 0-based).
constexpr int kZ0[9] = {0, 1, 2, 3, 4, 7, 8, 9, 10};
constexpr int kZp[5] = {5, 11, 12, 15, 16};
constexpr int kZm[5] = {6, 13, 14, 17, 18};

// Index of q inside its c_z group.
constexpr int group_index(int q) {
  for (int n = 0; n < 9; ++n)
    if (kZ0[n] == q) return n;
  for (int n = 0; n < 5; ++n)
    if (kZp[n] == q || kZm[n] == q) return n;
  return -1;
}

KMI_DEV int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Walled lattices: the kernels' template flag BND is 0 for a fully periodic lattice, 1 for
// the tiles of a walled lattice whose cells at every level stay at least one cell away from
// every x and y wall (only z walls remain, and those are uniform across a plane), 2 for the
// other tiles (any wall, Zou–He face or solid cell), 3 for those tiles when the lattice has
// no solid mask and its x and y faces are periodic or resting walls (x–y walls bounce links
// back with no Ladd term and carry no Zou–He closure; the z faces are handled as in mode 1).
// Whether a cell's neighbour exists along each axis direction — a periodic axis always wraps;
// a wall face has none beyond it.
struct Nbr {
  bool xm, xp, ym, yp, zm, zp;
};
template <int BND>
KMI_DEV Nbr nbr(const StepArgs& a, int gx, int gy, int gz) {
  Nbr n;
  if constexpr (BND == 1) {
    n.xm = n.xp = n.ym = n.yp = true;
  } else {
    n.xm = a.per[0] || gx > 0;
    n.xp = a.per[0] || gx < a.nx - 1;
    n.ym = a.per[1] || gy > 0;
    n.yp = a.per[1] || gy < a.ny - 1;
  }
  n.zm = a.per[2] || gz > 0;
  n.zp = a.per[2] || gz < a.nz - 1;
  return n;
}

// The source x − c_q of population q exists (else the link bounces back from a wall).
template <int q>
KMI_DEV bool src_ok(const Nbr& n) {
  constexpr int cx = Lat<19>::cx[q], cy = Lat<19>::cy[q], cz = Lat<19>::cz[q];
  return (cx > 0 ? n.xm : (cx < 0 ? n.xp : true)) && (cy > 0 ? n.ym : (cy < 0 ? n.yp : true)) &&
         (cz > 0 ? n.zm : (cz < 0 ? n.zp : true));
}
// Ladd term of every moving face the bounced link q crosses, x then y then z: the
// single-step kernel's pull, operation for operation.
template <int q, int BND, class R>
KMI_DEV R ladd(R v, const StepArgs& a, const Nbr& n) {
  constexpr int cx = Lat<19>::cx[q], cy = Lat<19>::cy[q], cz = Lat<19>::cz[q];
  constexpr bool XY = BND == 2;  // mode 3: resting x–y walls, mode 1: none in reach
  auto add = [&](int fid) {
    const FaceArgs& F = a.face[fid];
    if (F.kind == KMI_FACE_MOVING_WALL)
      v = v + R(six_w(Lat<19>::w[q])) * R(F.rho) * cdotu<cx, cy, cz>(R(F.u[0]), R(F.u[1]), R(F.u[2]));
  };
  if (XY && cx > 0 && !n.xm) add(KMI_WEST);
  if (XY && cx < 0 && !n.xp) add(KMI_EAST);
  if (XY && cy > 0 && !n.ym) add(KMI_SOUTH);
  if (XY && cy < 0 && !n.yp) add(KMI_NORTH);
  if (cz > 0 && !n.zm) add(KMI_BOTTOM);
  if (cz < 0 && !n.zp) add(KMI_TOP);
  return v;
}

// Element at byte offset `off` of a population plane.  Plane bases are uniform (scalar
// registers) and the offsets 32-bit, so each access is one saddr-form global instruction
// with no 64-bit address arithmetic.  A field holds fewer than 2^31 / 19 cells, so every
// byte offset fits in 32 bits.
template <class E>
KMI_DEV const E& at(const E* plane, unsigned off) {
  return *reinterpret_cast<const E*>(reinterpret_cast<const char*>(plane) + off);
}
template <class E>
KMI_DEV E* at(E* plane, unsigned off) {
  return reinterpret_cast<E*>(reinterpret_cast<char*>(plane) + off);
}
}  // namespace tb

// Population access of the multi-step kernels.  F16S keeps the populations in the storage
// scale (d · 2^15) from load to store: the load and the on-chip rounding skip the scale
// multiplies and the collisions run scaled (collide_trt_shifted<.., true>,
// collide_bgk_f16u), which reproduces the single-step kernel's values bit for bit.
template <int ST>
struct Pops {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  static constexpr bool kScaled = ST == KMI_STORE_F16S;
  KMI_DEV static R ld(const E* p) {
    if constexpr (kScaled) return float(*p);
    else return S::ld(p, 0);
  }
  // The collision's last operation is often a fused multiply-add; converting its FP32
  // result straight to binary16 would let the compiler merge the two into one mixed-precision
  // FMA (v_fma_mix, a single rounding to binary16) and change the stored value.  The empty
  // register constraint keeps the FP32 rounding in between, as in the unscaled path.
  KMI_DEV static _Float16 to_half(R v) {
    asm volatile("" : "+v"(v));
    return _Float16(v);
  }
  KMI_DEV static R round(R v) {
    if constexpr (kScaled) return float(to_half(v));
    else return storage_round<ST>(v);
  }
  template <bool NT>
  KMI_DEV static void st(E* p, R v) {
    if constexpr (kScaled) store_elem(p, 0, to_half(v), NT);
    else S::template st<NT>(p, 0, v);
  }
};

template <int ST, int COLL, class R>
KMI_DEV void collide19(const R (&fp)[19], R (&fo)[19], const StepArgs& a) {
  constexpr bool SH = Store<ST>::shifted;
  if constexpr (ST == KMI_STORE_F16S) {
    if constexpr (COLL == KMI_TRT) collide_trt_shifted<19, true>(fp, fo, R(a.s_plus), R(a.s_minus));
    else collide_bgk_f16u<19>(fp, fo, R(a.omega));
  } else if constexpr (COLL == KMI_TRT) {
    collide_trt<19, SH>(fp, fo, R(a.s_plus), R(a.s_minus));
  } else {
    collide_bgk<19, SH>(fp, fo, R(a.omega));
  }
}

// After the pull, the single-step kernel's order for a D3Q19 cell without thermal populations,
// forcing, cut links or rebuilt faces: the Zou–He closures of the faces the cell lies on (in
// the configured order), then the solid-cell rule or the collision.
template <int ST, int COLL, int BND, class R>
KMI_DEV void update19(R (&fp)[19], R (&fo)[19], const StepArgs& a, int gx, int gy, int gz) {
  if constexpr (BND == 1 || BND == 3) {
    // interior tiles, or a lattice whose x–y faces are periodic or resting walls: only the
    // z faces can carry a Zou–He closure, and no solid mask exists
    constexpr bool SH = Store<ST>::shifted;
    for (int n = 0; n < a.n_zh; ++n) {
      const int fid = a.zh_face[n];
      if (fid == KMI_BOTTOM && gz == 0) zh3_apply<SH, 4>(fp, a.face[fid]);
      if (fid == KMI_TOP && gz == a.nz - 1) zh3_apply<SH, 5>(fp, a.face[fid]);
    }
  } else if constexpr (BND == 2) {
    constexpr bool SH = Store<ST>::shifted;
    for (int n = 0; n < a.n_zh; ++n) {
      const int fid = a.zh_face[n];
      const FaceArgs& F = a.face[fid];
      switch (fid) {
        case KMI_WEST: if (gx == 0) zh3_apply<SH, 0>(fp, F); break;
        case KMI_EAST: if (gx == a.nx - 1) zh3_apply<SH, 1>(fp, F); break;
        case KMI_SOUTH: if (gy == 0) zh3_apply<SH, 2>(fp, F); break;
        case KMI_NORTH: if (gy == a.ny - 1) zh3_apply<SH, 3>(fp, F); break;
        case KMI_BOTTOM: if (gz == 0) zh3_apply<SH, 4>(fp, F); break;
        default: if (gz == a.nz - 1) zh3_apply<SH, 5>(fp, F); break;
      }
    }
    const long long self = gx + (long long)a.nx * (gy + (long long)a.ny * gz);
    if (a.mask != nullptr && a.mask[self] == kCellSolid) {
      if (a.solid_inert) {
        sfor<19>([&](auto qc) { fo[decltype(qc)::value] = SH ? R(0) : R(Lat<19>::w[decltype(qc)::value]); });
      } else {
        sfor<19>([&](auto qc) { fo[decltype(qc)::value] = fp[Lat<19>::opp[decltype(qc)::value]]; });
      }
      return;
    }
  }
  collide19<ST, COLL>(fp, fo, a);
}

// A level's populations of one cell pulled from the previous level's values in LDS (grid
// width EXp, NEp cells; the cell sits at (ex, ey) of that grid): c_z = 0 from A, c_z = +1 from
// the plane below (Bs), c_z = −1 from the plane above (Cs).  Walled lattices: a link whose
// source lies beyond a wall takes the cell's own opposite population of the previous level
// (A, Bo, Co: the cell's own plane), plus the Ladd term of a moving wall.
template <int BND, class R>
KMI_DEV void pull_lds(R (&fp)[19], const R* A, const R* Bs, const R* Cs, const R* Bo, const R* Co, int NEp,
                      int EXp, int ex, int ey, const StepArgs& a, const tb::Nbr& nb) {
  const int own = ey * EXp + ex;
  sfor<19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = Lat<19>::cx[q], cy = Lat<19>::cy[q], cz = Lat<19>::cz[q];
    constexpr int n = tb::group_index(q), o = Lat<19>::opp[q], no = tb::group_index(o);
    bool ok = true;
    if constexpr (BND) ok = tb::src_ok<q>(nb);
    if (ok) {
      const int e = (ey - cy) * EXp + (ex - cx);
      if constexpr (cz == 0) fp[q] = A[n * NEp + e];
      else if constexpr (cz > 0) fp[q] = Bs[n * NEp + e];
      else fp[q] = Cs[n * NEp + e];
    } else if constexpr (BND) {
      R v;
      if constexpr (cz == 0) v = A[no * NEp + own];
      else if constexpr (cz > 0) v = Co[no * NEp + own];  // opp(q) has c_z = −1
      else v = Bo[no * NEp + own];
      if constexpr (q != 0)
        if (a.any_moving) v = tb::ladd<q, BND>(v, a, nb);
      fp[q] = v;
    }
  });
}

// The two-step march of one column tile (tx, ty) over z-range zc, LDS at `lds`.
// NT threads per work-group, at least TX·TY: one per interior cell (a thread then evaluates
// up to two extended cells in phase 1), or at least (TX+2)(TY+2), one per extended cell
// (threads beyond TX·TY idle in phase 2).
template <int ST, int COLL, int TX, int TY, int NT, bool NTS, int BND>
KMI_DEV void march2(const StepArgs& a, int zlen, int tx, int ty, int zc, typename Store<ST>::real* lds) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  constexpr int EX = TX + 2, EY = TY + 2, NE = EX * EY, NI = TX * TY;
  static_assert(NT >= NI && NT % 64 == 0, "one thread per interior cell at least");
  constexpr int EPT = (NE + NT - 1) / NT;  // extended cells per work-item in phase 1
  constexpr unsigned ES = sizeof(E);
  // walled lattices keep the c_z = −1 values of two planes (the cell's own plane feeds its
  // bounced links), 29 values per extended cell
  constexpr int NC = BND ? 2 : 1;
  R* const A = lds;            // c_z = 0:  [9][NE]
  R* const B = lds + 9 * NE;   // c_z = +1: [2 planes][5][NE]
  R* const C = lds + 19 * NE;  // c_z = −1: [NC planes][5][NE]

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long long P = a.plane;
  const int x0 = tx * TX, y0 = ty * TY;
  const int z0 = zc * zlen, z1 = min(z0 + zlen, nz);
  const E* __restrict__ fi = static_cast<const E*>(a.fi);
  E* __restrict__ fo = static_cast<E*>(a.fo);
  const int tid = threadIdx.x;

  // Pull of the t populations of plane s for the extended cells this work-item owns.  With
  // FP32 arithmetic the pull for plane s+1 is issued before plane s is processed, so its HBM
  // latency overlaps the collisions, the barriers and the phase-2 stores of plane s; FP64
  // and 1024-wide groups have no registers to spare for it.  Walled lattices: a link from
  // beyond a wall reads the cell's own opposite population (halfway bounce-back, with the
  // Ladd term of a moving wall); extended cells beyond a wall are never read and skipped.
  constexpr bool kPrefetch = sizeof(R) == 4 && (NT <= 512 || EPT == 1);
  R nxt[EPT][19];
  auto inside = [&](int s, int e) {
    if constexpr (BND == 0) return true;
    if constexpr (BND == 1) return a.per[2] || (s >= 0 && s < nz);
    const int ux = x0 - 1 + e % EX, uy = y0 - 1 + e / EX;
    return (a.per[0] || (ux >= 0 && ux < nx)) && (a.per[1] || (uy >= 0 && uy < ny)) &&
           (a.per[2] || (s >= 0 && s < nz));
  };
  auto pull_plane = [&](int s) {
    const int zs = tb::wrap(s, nz);
    const int zm = tb::wrap(zs - 1, nz), zp = tb::wrap(zs + 1, nz);
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      const int e = tid + r * NT;
      if (e < NE && inside(s, e)) {
        const int ex = e % EX, ey = e / EX;
        const int gx = tb::wrap(x0 - 1 + ex, nx), gy = tb::wrap(y0 - 1 + ey, ny);
        const int xm = tb::wrap(gx - 1, nx), xp = tb::wrap(gx + 1, nx);
        const int ym = tb::wrap(gy - 1, ny), yp = tb::wrap(gy + 1, ny);
        // byte offsets of the three source columns and of the nine source rows
        const unsigned bx[3] = {unsigned(xm) * ES, unsigned(gx) * ES, unsigned(xp) * ES};
        const int yy[3] = {ym, gy, yp}, zz[3] = {zm, zs, zp};
        tb::Nbr nb{};
        if constexpr (BND) nb = tb::nbr<BND>(a, gx, gy, zs);
        const unsigned own = (unsigned(gx) + unsigned(nx) * unsigned(gy + ny * zs)) * ES;
        sfor<19>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          constexpr int cx = Lat<19>::cx[q], cy = Lat<19>::cy[q], cz = Lat<19>::cz[q];
          bool ok = true;
          if constexpr (BND) ok = tb::src_ok<q>(nb);
          if (ok) {
            const unsigned row = unsigned(nx) * unsigned(yy[1 - cy] + ny * zz[1 - cz]) * ES;
            nxt[r][q] = Pops<ST>::ld(&tb::at(fi + q * P, row + bx[1 - cx]));
          } else if constexpr (BND) {
            R v = Pops<ST>::ld(&tb::at(fi + Lat<19>::opp[q] * P, own));
            if constexpr (q != 0)
              if (a.any_moving) v = tb::ladd<q, BND>(v, a, nb);
            nxt[r][q] = v;
          }
        });
      }
    }
  };
  if constexpr (kPrefetch) pull_plane(z0 - 1);

  for (int s = z0 - 1; s <= z1; ++s) {
    const int u = s + 3;  // non-negative plane counter for the LDS ring slot
    R* const Cw = C + (BND ? (u & 1) : 0) * 5 * NE;
    // ---- phase 1: t+1 on plane s, extended tile ----
    if constexpr (!kPrefetch) pull_plane(s);
    R cur[EPT][19];
#pragma unroll
    for (int r = 0; r < EPT; ++r)
#pragma unroll
      for (int q = 0; q < 19; ++q) cur[r][q] = nxt[r][q];
    if constexpr (kPrefetch)
      if (s < z1) pull_plane(s + 1);
    R keep0[EPT][9], keepp[EPT][5];
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      const int e = tid + r * NT;
      if (e < NE && inside(s, e)) {
        R fn[19];
        update19<ST, COLL, BND>(cur[r], fn, a, tb::wrap(x0 - 1 + e % EX, nx), tb::wrap(y0 - 1 + e / EX, ny),
                                tb::wrap(s, nz));
        sfor<19>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          constexpr int cz = Lat<19>::cz[q];
          constexpr int n = tb::group_index(q);
          const R v = Pops<ST>::round(fn[q]);
          if constexpr (cz == 0) keep0[r][n] = v;
          else if constexpr (cz > 0) keepp[r][n] = v;
          else Cw[n * NE + e] = v;
        });
      }
    }
    __syncthreads();
    // ---- phase 2: t+2 on plane s−1, interior tile ----
    if (s >= z0 + 1 && tid < NI) {
      const R* const Br = B + ((u - 2) & 1) * 5 * NE;  // c_z = +1 of plane s−2
      const R* const Bo = B + ((u - 1) & 1) * 5 * NE;  // c_z = +1 of plane s−1 (own)
      const R* const Co = C + (BND ? ((u - 1) & 1) : 0) * 5 * NE;  // c_z = −1 of plane s−1 (own)
      const int ix = tid % TX, iy = tid / TX;
      tb::Nbr nb{};
      if constexpr (BND) nb = tb::nbr<BND>(a, x0 + ix, y0 + iy, s - 1);
      R fp[19], fn[19];
      if constexpr (BND == 0) {
        // the periodic pull written out: through pull_lds the allocator keeps 1–4 more
        // VGPRs live, which takes the F16S 64×8 / 704-thread tile from 80 to 81 VGPRs and
        // from two groups per CU to one (0.361 against 0.479 ms per launch)
        sfor<19>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          constexpr int cx = Lat<19>::cx[q], cy = Lat<19>::cy[q], cz = Lat<19>::cz[q];
          constexpr int n = tb::group_index(q);
          const int e = (iy + 1 - cy) * EX + (ix + 1 - cx);
          if constexpr (cz == 0) fp[q] = A[n * NE + e];
          else if constexpr (cz > 0) fp[q] = Br[n * NE + e];
          else fp[q] = Cw[n * NE + e];
        });
      } else {
        pull_lds<BND>(fp, A, Br, Cw, Bo, Co, NE, EX, ix + 1, iy + 1, a, nb);
      }
      update19<ST, COLL, BND>(fp, fn, a, x0 + ix, y0 + iy, s - 1);
      const unsigned cell = (unsigned(x0 + ix) + unsigned(nx) * unsigned(y0 + iy + ny * (s - 1))) * ES;
      sfor<19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        Pops<ST>::template st<NTS>(tb::at(fo + q * P, cell), fn[q]);
      });
    }
    __syncthreads();
    // ---- phase 3: plane s's c_z = 0 and +1 populations into the slots phase 2 released ----
    {
      R* const Bw = B + (u & 1) * 5 * NE;
#pragma unroll
      for (int r = 0; r < EPT; ++r) {
        const int e = tid + r * NT;
        if (e < NE && inside(s, e)) {
#pragma unroll
          for (int n = 0; n < 9; ++n) A[n * NE + e] = keep0[r][n];
#pragma unroll
          for (int n = 0; n < 5; ++n) Bw[n * NE + e] = keepp[r][n];
        }
      }
    }
  }
}

// Grid: one work-group per (column tile, z-range), padded to a multiple of 8.  Work-group
// b runs on XCD b mod 8; it takes tile (b mod 8)·per_xcd + b/8, so each XCD works through a
// contiguous band of neighbouring tiles and their shared halo rows meet in its L2.  Walled
// lattices (BND = 2, or 3 when the x–y faces are periodic or resting walls and no mask
// exists): the tiles of the block `inner` ([x, y) × [z, w) in tile units) march on the
// interior variant (BND = 1), the others on the walled one.
template <int ST, int COLL, int TX, int TY, int NT, bool NTS, int BND>
__global__ void __launch_bounds__(NT)
    k_step2_d3q19(const StepArgs a, int zlen, int ntx, int nty, int ntiles, int per_xcd, const int4 inner) {
  using R = typename Store<ST>::real;
  __shared__ R lds[(19 + 5 * (BND ? 2 : 1)) * (TX + 2) * (TY + 2)];
  int t = int(blockIdx.x % 8u) * per_xcd + int(blockIdx.x / 8u);
  if constexpr (BND >= 2) {
    if (a.tile_order) t = a.tile_order[blockIdx.x];
    if (t < 0) return;  // padding groups of the table
  }
  // padding groups (uniform per group, before any barrier); the periodic kernels' prologue
  // is kept to this one test: a longer one moved the plane loop and cost the 64×8 /
  // 832-thread three-step launch 3 % (measured before loop heads were 64-byte aligned)
  if (t >= ntiles) return;
  const int tx = t % ntx, ty = (t / ntx) % nty, zc = t / (ntx * nty);
  if constexpr (BND >= 2) {
    if (tx >= inner.x && tx < inner.y && ty >= inner.z && ty < inner.w) {
      march2<ST, COLL, TX, TY, NT, NTS, 1>(a, zlen, tx, ty, zc, lds);
      return;
    }
  }
  march2<ST, COLL, TX, TY, NT, NTS, BND>(a, zlen, tx, ty, zc, lds);
}

// ---- three steps per launch ------------------------------------------------------------
// The same march with two LDS levels: the t+1 populations on a two-cell x–y halo
// ((TX+4)(TY+4) cells, level 1) and the t+2 populations on a one-cell halo ((TX+2)(TY+2),
// level 2), grouped by c_z as in the two-step kernel.  HBM sees one read and one write per
// population per three steps; the extra evaluations are the two halos and the two planes
// below and above the z-range at each level.  Per marched plane s (walled lattices):
//   P1  t+1 on plane s, level-1 tile, pulled from the t field in HBM: c_z = −1 → L1.C at
//       once, c_z = 0 and +1 kept in registers;                                    barrier
//   P2  t+2 on plane s−1, level-2 tile, pulled from L1 (c_z = +1 of s−2, 0 of s−1, −1 of s):
//       c_z = −1 → L2.C, c_z = 0 and +1 kept;                                      barrier
//   P3  t+3 on plane s−2, interior, pulled from L2 (c_z = +1 of s−3, 0 of s−2, −1 of s−1),
//       stored to HBM; the values kept in P1 replace the L1 slots P2 consumed;     barrier
//   P4  the values kept in P2 replace the L2 slots P3 consumed.
// P1 writes only L1.C, whose readers (P2) finished before the second barrier; P4's L2
// writes are read after the next iteration's second barrier.
// Periodic lattices drop the third barrier: level 2 keeps its c_z = 0 values of planes s−1
// and s−2 in two slots, and each level writes its c_z = +1 values one plane late (held in
// registers for a plane), so the single level-1 c_z = +1 slot holds plane s−2 when P2
// reads it and two level-2 slots hold planes s−3 (read by P3) and s−2 (written beside it).
// All L1 and L2 writes then follow P3 in the same barrier interval, and the threads with
// no P3 cell (and the waves that finish P3 first) run into the next plane's P1 while P3 is
// still being evaluated.  Every t+1 and t+2 value passes through the storage
// rounding, so the result is bit-identical to three single-step launches in every storage.
// NT threads per work-group (at least TX·TY): with NT = TX·TY a thread evaluates up to two
// level-1 and two level-2 cells one after the other; with NT ≥ (TX+4)(TY+4) every cell of a
// level has its own thread, which halves the registers a thread holds (more resident waves
// per CU) and runs the halo evaluations side by side; threads beyond TX·TY idle in P3.
// LDS values of one three-step work-group: periodic, 19 per level-1 cell and 33 per level-2
// cell (c_z = 0 and +1 of level 2 in two slots each); walled, 29 per cell of both levels.
template <int BND>
constexpr int march3_lds_values(int TX, int TY) {
  const int n1 = (TX + 4) * (TY + 4), n2 = (TX + 2) * (TY + 2);
  return BND ? 29 * (n1 + n2) : 19 * n1 + 33 * n2;
}

template <int ST, int COLL, int TX, int TY, int NT, bool NTS, int BND>
KMI_DEV void march3(const StepArgs& a, int zlen, int tx, int ty, int zc, typename Store<ST>::real* lds) {
  using S = Store<ST>;
  using E = typename S::elem;
  using R = typename S::real;
  constexpr int E1X = TX + 4, E1Y = TY + 4, N1 = E1X * E1Y;  // level 1 (t+1)
  constexpr int E2X = TX + 2, E2Y = TY + 2, N2 = E2X * E2Y;  // level 2 (t+2)
  constexpr int NI = TX * TY;
  static_assert(NT >= NI && NT % 64 == 0, "one thread per interior cell at least");
  constexpr int EPT1 = (N1 + NT - 1) / NT, EPT2 = (N2 + NT - 1) / NT;
  constexpr unsigned ES = sizeof(E);
  // Periodic lattices march with two barriers per plane (layout march3_lds_values<0>):
  // level 1 keeps one slot per set (its c_z = +1 values are written one plane late, see
  // below), level 2 two c_z = 0 slots and two c_z = +1 slots.  Walled lattices keep the
  // three-barrier schedule and 24 values per cell and level, plus a second c_z = −1 plane
  // (the cell's own plane feeds its bounced links).
  constexpr bool TWO = BND == 0;
  constexpr int NC = BND ? 2 : 1;  // c_z = −1 planes per level
  R* const A1 = lds;                         // c_z = 0:  [9][N1]
  R* const B1 = lds + 9 * N1;                // c_z = +1: [TWO ? 1 : 2 planes][5][N1]
  R* const C1 = B1 + (TWO ? 5 : 10) * N1;    // c_z = −1: [NC planes][5][N1]
  R* const L2 = C1 + 5 * NC * N1;
  R* const A2 = L2;                          // [TWO ? 2 : 1 planes][9][N2]
  R* const B2 = A2 + (TWO ? 18 : 9) * N2;    // [2][5][N2]
  R* const C2 = B2 + 10 * N2;                // [NC][5][N2]

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long long P = a.plane;
  const int x0 = tx * TX, y0 = ty * TY;
  const int z0 = zc * zlen, z1 = min(z0 + zlen, nz);
  const E* __restrict__ fi = static_cast<const E*>(a.fi);
  E* __restrict__ fo = static_cast<E*>(a.fo);
  const int tid = threadIdx.x;

  // cell (ux, uy) of plane s lies in the domain (walled lattices; periodic axes always)
  auto inside = [&](int ux, int uy, int s) {
    if constexpr (BND == 0) return true;
    if constexpr (BND == 1) return a.per[2] || (s >= 0 && s < nz);
    return (a.per[0] || (ux >= 0 && ux < nx)) && (a.per[1] || (uy >= 0 && uy < ny)) &&
           (a.per[2] || (s >= 0 && s < nz));
  };
  constexpr bool kPrefetch = sizeof(R) == 4 && (NT <= 512 || EPT1 == 1);
  R nxt[EPT1][19];
  auto pull_plane = [&](int s) {
    const int zs = tb::wrap(s, nz);
    const int zm = tb::wrap(zs - 1, nz), zp = tb::wrap(zs + 1, nz);
#pragma unroll
    for (int r = 0; r < EPT1; ++r) {
      const int e = tid + r * NT;
      if (e < N1 && inside(x0 - 2 + e % E1X, y0 - 2 + e / E1X, s)) {
        const int ex = e % E1X, ey = e / E1X;
        const int gx = tb::wrap(x0 - 2 + ex, nx), gy = tb::wrap(y0 - 2 + ey, ny);
        const int xm = tb::wrap(gx - 1, nx), xp = tb::wrap(gx + 1, nx);
        const int ym = tb::wrap(gy - 1, ny), yp = tb::wrap(gy + 1, ny);
        const unsigned bx[3] = {unsigned(xm) * ES, unsigned(gx) * ES, unsigned(xp) * ES};
        const int yy[3] = {ym, gy, yp}, zz[3] = {zm, zs, zp};
        tb::Nbr nb{};
        if constexpr (BND) nb = tb::nbr<BND>(a, gx, gy, zs);
        const unsigned own = (unsigned(gx) + unsigned(nx) * unsigned(gy + ny * zs)) * ES;
        sfor<19>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          constexpr int cx = Lat<19>::cx[q], cy = Lat<19>::cy[q], cz = Lat<19>::cz[q];
          bool ok = true;
          if constexpr (BND) ok = tb::src_ok<q>(nb);
          if (ok) {
            const unsigned row = unsigned(nx) * unsigned(yy[1 - cy] + ny * zz[1 - cz]) * ES;
            nxt[r][q] = Pops<ST>::ld(&tb::at(fi + q * P, row + bx[1 - cx]));
          } else if constexpr (BND) {
            R v = Pops<ST>::ld(&tb::at(fi + Lat<19>::opp[q] * P, own));
            if constexpr (q != 0)
              if (a.any_moving) v = tb::ladd<q, BND>(v, a, nb);
            nxt[r][q] = v;
          }
        });
      }
    }
  };
  if constexpr (kPrefetch) pull_plane(z0 - 2);
  // two-barrier schedule: the c_z = +1 values of the previous plane, held one plane longer
  R k1q[TWO ? EPT1 : 1][5] = {}, k2q[TWO ? EPT2 : 1][5] = {};

  for (int s = z0 - 2; s <= z1 + 1; ++s) {
    const int u = s + 4;  // non-negative plane counter for the ring slots
    R* const C1w = C1 + (BND ? (u & 1) : 0) * 5 * N1;        // c_z = −1, level 1, plane s
    R* const C2w = C2 + (BND ? ((u - 1) & 1) : 0) * 5 * N2;  // c_z = −1, level 2, plane s−1
    // ---- P1: t+1 on plane s, level-1 tile ----
    if constexpr (!kPrefetch) pull_plane(s);
    R cur[EPT1][19];
#pragma unroll
    for (int r = 0; r < EPT1; ++r)
#pragma unroll
      for (int q = 0; q < 19; ++q) cur[r][q] = nxt[r][q];
    if constexpr (kPrefetch)
      if (s <= z1) pull_plane(s + 1);
    R k1z[EPT1][9], k1p[EPT1][5];
#pragma unroll
    for (int r = 0; r < EPT1; ++r) {
      const int e = tid + r * NT;
      if (e < N1 && inside(x0 - 2 + e % E1X, y0 - 2 + e / E1X, s)) {
        R fn[19];
        update19<ST, COLL, BND>(cur[r], fn, a, tb::wrap(x0 - 2 + e % E1X, nx), tb::wrap(y0 - 2 + e / E1X, ny),
                                tb::wrap(s, nz));
        sfor<19>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          constexpr int cz = Lat<19>::cz[q];
          constexpr int n = tb::group_index(q);
          const R v = Pops<ST>::round(fn[q]);
          if constexpr (cz == 0) k1z[r][n] = v;
          else if constexpr (cz > 0) k1p[r][n] = v;
          else C1w[n * N1 + e] = v;
        });
      }
    }
    __syncthreads();
    // ---- P2: t+2 on plane s−1, level-2 tile ----
    R k2z[EPT2][9], k2p[EPT2][5];
    if (s >= z0) {
      const R* const B1r = B1 + (TWO ? 0 : ((u - 2) & 1) * 5 * N1);  // c_z = +1 of plane s−2
      const R* const B1o = B1 + ((u - 1) & 1) * 5 * N1;  // c_z = +1 of plane s−1 (own; walled only)
      const R* const C1o = C1 + (BND ? ((u - 1) & 1) : 0) * 5 * N1;  // c_z = −1 of plane s−1 (own)
#pragma unroll
      for (int r = 0; r < EPT2; ++r) {
        const int e = tid + r * NT;
        if (e < N2 && inside(x0 - 1 + e % E2X, y0 - 1 + e / E2X, s - 1)) {
          const int ex = e % E2X, ey = e / E2X;
          const int gx = tb::wrap(x0 - 1 + ex, nx), gy = tb::wrap(y0 - 1 + ey, ny), gz = tb::wrap(s - 1, nz);
          tb::Nbr nb{};
          if constexpr (BND) nb = tb::nbr<BND>(a, gx, gy, gz);
          R fp[19], fn[19];
          pull_lds<BND>(fp, A1, B1r, C1w, B1o, C1o, N1, E1X, ex + 1, ey + 1, a, nb);
          update19<ST, COLL, BND>(fp, fn, a, gx, gy, gz);
          sfor<19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cz = Lat<19>::cz[q];
            constexpr int n = tb::group_index(q);
            const R v = Pops<ST>::round(fn[q]);
            if constexpr (cz == 0) k2z[r][n] = v;
            else if constexpr (cz > 0) k2p[r][n] = v;
            else C2w[n * N2 + e] = v;
          });
        }
      }
    }
    __syncthreads();
    // ---- P3: t+3 on plane s−2, interior (walled: then the P1 values into level 1) ----
    if (s >= z0 + 2 && tid < NI) {
      // Two-barrier schedule: these waves also evaluate the next plane's P1 before the next
      // barrier, so their P3 issues ahead of the waves that already moved on to it (wave
      // priority only; 0.8 % shorter launches for 64×8 / 832 threads, measured same-box).
      if constexpr (TWO) __builtin_amdgcn_s_setprio(2);
      const R* const A2r = A2 + (TWO ? ((u - 2) & 1) * 9 * N2 : 0);  // c_z = 0 of plane s−2
      const R* const B2r = B2 + ((u - 3) & 1) * 5 * N2;  // c_z = +1 of plane s−3
      const R* const B2o = B2 + ((u - 2) & 1) * 5 * N2;  // c_z = +1 of plane s−2 (own)
      const R* const C2o = C2 + (BND ? ((u - 2) & 1) : 0) * 5 * N2;  // c_z = −1 of plane s−2 (own)
      const int ix = tid % TX, iy = tid / TX;
      tb::Nbr nb{};
      if constexpr (BND) nb = tb::nbr<BND>(a, x0 + ix, y0 + iy, s - 2);
      R fp[19], fn[19];
      pull_lds<BND>(fp, A2r, B2r, C2w, B2o, C2o, N2, E2X, ix + 1, iy + 1, a, nb);
      update19<ST, COLL, BND>(fp, fn, a, x0 + ix, y0 + iy, s - 2);
      const unsigned cell = (unsigned(x0 + ix) + unsigned(nx) * unsigned(y0 + iy + ny * (s - 2))) * ES;
      sfor<19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        Pops<ST>::template st<NTS>(tb::at(fo + q * P, cell), fn[q]);
      });
      if constexpr (TWO) __builtin_amdgcn_s_setprio(0);
    }
    if constexpr (TWO) {
      // The values kept in P1 and P2 go to LDS after P3 (so P3, which the next barrier
      // waits for, starts at once), into slots nobody reads until the next plane's P2
      // (level 1) and P3 (level 2): level 1 takes plane s's c_z = 0 values and plane s−1's
      // c_z = +1 values (P2 has read plane s−2's from that slot); level 2 takes plane s−1's
      // c_z = 0 values and plane s−2's c_z = +1 values, in the slots P3 does not read.  The
      // next plane's P1 writes only L1.C, whose reader (P2) finished before the last
      // barrier, so no third barrier is needed: threads without a P3 cell go straight on to
      // the next plane's P1.
#pragma unroll
      for (int r = 0; r < EPT1; ++r) {
        const int e = tid + r * NT;
        if (e < N1) {
#pragma unroll
          for (int n = 0; n < 9; ++n) A1[n * N1 + e] = k1z[r][n];
#pragma unroll
          for (int n = 0; n < 5; ++n) {
            B1[n * N1 + e] = k1q[r][n];
            k1q[r][n] = k1p[r][n];
          }
        }
      }
      if (s >= z0) {
        R* const A2w = A2 + ((u - 1) & 1) * 9 * N2;  // plane s−1
        R* const B2w = B2 + ((u - 2) & 1) * 5 * N2;  // plane s−2
#pragma unroll
        for (int r = 0; r < EPT2; ++r) {
          const int e = tid + r * NT;
          if (e < N2) {
#pragma unroll
            for (int n = 0; n < 9; ++n) A2w[n * N2 + e] = k2z[r][n];
#pragma unroll
            for (int n = 0; n < 5; ++n) {
              B2w[n * N2 + e] = k2q[r][n];
              k2q[r][n] = k2p[r][n];
            }
          }
        }
      }
    }
    if constexpr (!TWO) {
      {
        R* const B1w = B1 + (u & 1) * 5 * N1;
#pragma unroll
        for (int r = 0; r < EPT1; ++r) {
          const int e = tid + r * NT;
          if (e < N1 && inside(x0 - 2 + e % E1X, y0 - 2 + e / E1X, s)) {
#pragma unroll
            for (int n = 0; n < 9; ++n) A1[n * N1 + e] = k1z[r][n];
#pragma unroll
            for (int n = 0; n < 5; ++n) B1w[n * N1 + e] = k1p[r][n];
          }
        }
      }
      __syncthreads();
      // ---- P4: the P2 values (plane s−1) into level 2 ----
      if (s >= z0) {
        R* const B2w = B2 + ((u - 1) & 1) * 5 * N2;
#pragma unroll
        for (int r = 0; r < EPT2; ++r) {
          const int e = tid + r * NT;
          if (e < N2 && inside(x0 - 1 + e % E2X, y0 - 1 + e / E2X, s - 1)) {
#pragma unroll
            for (int n = 0; n < 9; ++n) A2[n * N2 + e] = k2z[r][n];
#pragma unroll
            for (int n = 0; n < 5; ++n) B2w[n * N2 + e] = k2p[r][n];
          }
        }
      }
    }
  }
}

// Grid and walled tiles as for k_step2_d3q19.
template <int ST, int COLL, int TX, int TY, int NT, bool NTS, int BND>
__global__ void __launch_bounds__(NT)
    k_step3_d3q19(const StepArgs a, int zlen, int ntx, int nty, int ntiles, int per_xcd, const int4 inner) {
  using R = typename Store<ST>::real;
  __shared__ R lds[march3_lds_values<BND>(TX, TY)];
  int t = int(blockIdx.x % 8u) * per_xcd + int(blockIdx.x / 8u);
  if constexpr (BND >= 2) {
    if (a.tile_order) t = a.tile_order[blockIdx.x];
    if (t < 0) return;  // padding groups of the table
  }
  // padding groups (uniform per group, before any barrier); the periodic kernels' prologue
  // is kept to this one test: a longer one moved the plane loop and cost the 64×8 /
  // 832-thread three-step launch 3 % (measured before loop heads were 64-byte aligned)
  if (t >= ntiles) return;
  const int tx = t % ntx, ty = (t / ntx) % nty, zc = t / (ntx * nty);
  if constexpr (BND >= 2) {
    if (tx >= inner.x && tx < inner.y && ty >= inner.z && ty < inner.w) {
      march3<ST, COLL, TX, TY, NT, NTS, 1>(a, zlen, tx, ty, zc, lds);
      return;
    }
  }
  march3<ST, COLL, TX, TY, NT, NTS, BND>(a, zlen, tx, ty, zc, lds);
}

// ---- dispatch ---------------------------------------------------------------------------
// Tile variants (TX × TY interior cells per work-group).  Two steps per launch: 1 = 32×8,
// 2 = 64×4, 3 = 64×8, 4 = 32×16, 5 = 64×16, not for FP64 (the larger the tile, the smaller
// the share of halo recomputation: (TX+2)(TY+2)/(TX·TY) = 1.55, 1.55, 1.29, 1.32, 1.16),
// with one thread per interior cell; 12 = 64×8 (704 threads), 13 = 32×16 (640), 14 = 32×8
// (384) with one thread per extended cell.
// Three steps per launch: 6 = 64×8, 7 = 32×8, 8 = 32×16 with one thread per interior cell;
// 9 = 32×8 (448 threads), 10 = 32×16 (768), 11 = 64×8 (832) with one thread per level-1
// cell; FP64 only 7 and 9 (LDS of the periodic march, march3_lds_values<0>: 149 / 78 / 136 KB
// for 64×8 / 32×8 / 32×16 in FP32, 155 KB for 32×8 in FP64; evaluations per update ((TX+4)(TY+4) + (TX+2)(TY+2) + TX·TY) / (3·TX·TY) = 1.29,
// 1.34, 1.20).
bool step2_tile(int tile, int st, int& TX, int& TY) {
  switch (tile) {
    case 1: TX = 32; TY = 8; return true;
    case 2: TX = 64; TY = 4; return true;
    case 3: TX = 64; TY = 8; return true;
    case 4: TX = 32; TY = 16; return true;
    case 5: TX = 64; TY = 16; return st != KMI_STORE_F64;  // FP64: > 160 KB of LDS
    case 6: TX = 64; TY = 8; return st != KMI_STORE_F64;
    case 7: TX = 32; TY = 8; return true;
    case 8: TX = 32; TY = 16; return st != KMI_STORE_F64;
    case 9: TX = 32; TY = 8; return true;
    case 10: TX = 32; TY = 16; return st != KMI_STORE_F64;
    case 11: TX = 64; TY = 8; return st != KMI_STORE_F64;
    case 12: TX = 64; TY = 8; return true;
    case 13: TX = 32; TY = 16; return true;
    case 14: TX = 32; TY = 8; return true;
    default: return false;
  }
}

int temporal_span(int tile) {
  if (tile >= 6 && tile <= 11) return 3;
  return tile >= 1 ? 2 : 1;
}

// Walled lattices (bnd): tiles 12 and 14 (two steps) and 9 (three steps, not FP64: 29 values
// per cell of both levels need 179 KB of LDS in FP64), non-temporal stores.
bool step2_tile_walled(int tile, int st) {
  if (st == KMI_STORE_F16S) return false;
  return tile == 12 || tile == 14 || (tile == 9 && st != KMI_STORE_F64);
}

// Along one axis, the tiles whose cells at every level (H halo cells either side of the tile)
// lie at least one cell from both walls: [lo, hi), the whole axis when it is periodic.
static void interior_tiles(bool per, int n, int T, int H, int ntiles, int& lo, int& hi) {
  lo = 0;
  hi = ntiles;
  if (per) return;
  while (lo < ntiles && lo * T - H < 1) ++lo;
  while (hi > lo && (hi - 1) * T + T - 1 + H > n - 2) --hi;
  if (hi <= lo) lo = hi = 0;
}

// The walled kernels' lean ring variant (BND = 3) applies: no solid mask, and every x and y
// face periodic or a resting wall.
static bool lean_walls(const StepArgs& a) {
  if (a.mask != nullptr) return false;
  for (int f : {KMI_WEST, KMI_EAST, KMI_SOUTH, KMI_NORTH})
    if (a.face[f].kind != KMI_FACE_PERIODIC && a.face[f].kind != KMI_FACE_WALL) return false;
  return true;
}

// One multi-step pass: every column tile × z-range in one launch, on the periodic kernel
// (BND = 0) or, for a walled lattice, the walled one (BND = 2) with the interior block of
// tiles (none when the lattice has a solid mask) marching on the interior variant.
template <class Go>
static hipError_t launch_tiles(bool bnd, int span, int TX, int TY, const StepArgs& a, int zchunks, Go&& go) {
  const int ntx = a.nx / TX, nty = a.ny / TY;
  const int zlen = (a.nz + zchunks - 1) / zchunks;
  const int nzc = (a.nz + zlen - 1) / zlen;
  const int ntiles = ntx * nty * nzc;
  const int per_xcd = (ntiles + 7) / 8;
  int4 inner = make_int4(0, 0, 0, 0);
  if (bnd && a.mask == nullptr) {
    interior_tiles(a.per[0] != 0, a.nx, TX, span - 1, ntx, inner.x, inner.y);
    interior_tiles(a.per[1] != 0, a.ny, TY, span - 1, nty, inner.z, inner.w);
  }
  return go(dim3(static_cast<unsigned>(8 * per_xcd)), zlen, ntx, nty, ntiles, per_xcd, inner);
}

// Work-group order of a walled multi-step launch: each XCD's band of tiles (the same bands
// as the periodic launch) with its ring tiles — the slower walled variant — first, so that
// when the groups outnumber the CUs the interior groups fill the rounds the ring groups
// leave (longest first).  order[b] = tile of work-group b, -1 for the padding groups.
void walled_tile_order(const StepArgs& a, int tile, int zchunks, std::vector<int>& order) {
  int TX = 0, TY = 0;
  order.clear();
  if (!step2_tile(tile, KMI_STORE_F32, TX, TY)) return;
  const int span = temporal_span(tile);
  const int ntx = a.nx / TX, nty = a.ny / TY;
  const int zlen = (a.nz + zchunks - 1) / zchunks;
  const int ntiles = ntx * nty * ((a.nz + zlen - 1) / zlen);
  const int per_xcd = (ntiles + 7) / 8;
  int ix0 = 0, ix1 = 0, iy0 = 0, iy1 = 0;
  if (a.mask == nullptr) {
    interior_tiles(a.per[0] != 0, a.nx, TX, span - 1, ntx, ix0, ix1);
    interior_tiles(a.per[1] != 0, a.ny, TY, span - 1, nty, iy0, iy1);
  }
  order.assign(size_t(8) * per_xcd, -1);
  std::vector<int> band;
  for (int x = 0; x < 8; ++x) {
    band.clear();
    const int t0 = x * per_xcd, t1 = std::min(t0 + per_xcd, ntiles);
    for (int pass = 0; pass < 2; ++pass)  // ring tiles, then interior tiles
      for (int t = t0; t < t1; ++t) {
        const int tx = t % ntx, ty = (t / ntx) % nty;
        const bool inner = tx >= ix0 && tx < ix1 && ty >= iy0 && ty < iy1;
        if (inner == (pass == 1)) band.push_back(t);
      }
    for (size_t i = 0; i < band.size(); ++i) order[size_t(x) + 8 * i] = band[i];
  }
}

template <int ST, int COLL, int TX, int TY, int NT = TX * TY>
static hipError_t launch2(bool nts, bool bnd, const StepArgs& a, int zchunks, hipStream_t s) {
  return launch_tiles(bnd, 2, TX, TY, a, zchunks,
                      [&](dim3 g, int zlen, int ntx, int nty, int ntiles, int per_xcd, int4 inner) -> hipError_t {
                        const dim3 b(static_cast<unsigned>(NT));
                        if (bnd) {
                          if constexpr (ST != KMI_STORE_F16S && (NT == 704 || NT == 384)) {
                            if (lean_walls(a))
                              hipLaunchKernelGGL((k_step2_d3q19<ST, COLL, TX, TY, NT, true, 3>), g, b, 0, s, a,
                                                 zlen, ntx, nty, ntiles, per_xcd, inner);
                            else
                              hipLaunchKernelGGL((k_step2_d3q19<ST, COLL, TX, TY, NT, true, 2>), g, b, 0, s, a,
                                                 zlen, ntx, nty, ntiles, per_xcd, inner);
                          } else {
                            return hipErrorInvalidValue;
                          }
                        } else if (nts) {
                          hipLaunchKernelGGL((k_step2_d3q19<ST, COLL, TX, TY, NT, true, 0>), g, b, 0, s, a, zlen, ntx,
                                             nty, ntiles, per_xcd, inner);
                        } else {
                          hipLaunchKernelGGL((k_step2_d3q19<ST, COLL, TX, TY, NT, false, 0>), g, b, 0, s, a, zlen,
                                             ntx, nty, ntiles, per_xcd, inner);
                        }
                        return hipGetLastError();
                      });
}

template <int ST, int COLL, int TX, int TY, int NT = TX * TY>
static hipError_t launch3(bool nts, bool bnd, const StepArgs& a, int zchunks, hipStream_t s) {
  return launch_tiles(bnd, 3, TX, TY, a, zchunks,
                      [&](dim3 g, int zlen, int ntx, int nty, int ntiles, int per_xcd, int4 inner) -> hipError_t {
                        const dim3 b(static_cast<unsigned>(NT));
                        if (bnd) {
                          if constexpr (ST != KMI_STORE_F16S && ST != KMI_STORE_F64 && NT == 448) {
                            if (lean_walls(a))
                              hipLaunchKernelGGL((k_step3_d3q19<ST, COLL, TX, TY, NT, true, 3>), g, b, 0, s, a,
                                                 zlen, ntx, nty, ntiles, per_xcd, inner);
                            else
                              hipLaunchKernelGGL((k_step3_d3q19<ST, COLL, TX, TY, NT, true, 2>), g, b, 0, s, a,
                                                 zlen, ntx, nty, ntiles, per_xcd, inner);
                          } else {
                            return hipErrorInvalidValue;
                          }
                        } else if (nts) {
                          hipLaunchKernelGGL((k_step3_d3q19<ST, COLL, TX, TY, NT, true, 0>), g, b, 0, s, a, zlen, ntx,
                                             nty, ntiles, per_xcd, inner);
                        } else {
                          hipLaunchKernelGGL((k_step3_d3q19<ST, COLL, TX, TY, NT, false, 0>), g, b, 0, s, a, zlen,
                                             ntx, nty, ntiles, per_xcd, inner);
                        }
                        return hipGetLastError();
                      });
}

template <int ST, int COLL>
static hipError_t launch2_tile(int tile, bool nts, bool bnd, const StepArgs& a, int zchunks, hipStream_t s) {
  switch (tile) {
    case 7: return launch3<ST, COLL, 32, 8>(nts, bnd, a, zchunks, s);
    case 9: return launch3<ST, COLL, 32, 8, 448>(nts, bnd, a, zchunks, s);
    case 6:
    case 8:
    case 10:
    case 11:
      if constexpr (ST != KMI_STORE_F64) {
        if (tile == 6) return launch3<ST, COLL, 64, 8>(nts, bnd, a, zchunks, s);
        if (tile == 8) return launch3<ST, COLL, 32, 16>(nts, bnd, a, zchunks, s);
        if (tile == 10) return launch3<ST, COLL, 32, 16, 768>(nts, bnd, a, zchunks, s);
        return launch3<ST, COLL, 64, 8, 832>(nts, bnd, a, zchunks, s);
      } else {
        return hipErrorInvalidValue;
      }
    case 1: return launch2<ST, COLL, 32, 8>(nts, bnd, a, zchunks, s);
    case 2: return launch2<ST, COLL, 64, 4>(nts, bnd, a, zchunks, s);
    case 3: return launch2<ST, COLL, 64, 8>(nts, bnd, a, zchunks, s);
    case 4: return launch2<ST, COLL, 32, 16>(nts, bnd, a, zchunks, s);
    case 12: return launch2<ST, COLL, 64, 8, 704>(nts, bnd, a, zchunks, s);
    case 13: return launch2<ST, COLL, 32, 16, 640>(nts, bnd, a, zchunks, s);
    case 14: return launch2<ST, COLL, 32, 8, 384>(nts, bnd, a, zchunks, s);
    case 5:
      if constexpr (ST != KMI_STORE_F64) return launch2<ST, COLL, 64, 16>(nts, bnd, a, zchunks, s);
      else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

template <int ST>
static hipError_t launch2_coll(int coll, int tile, bool nts, bool bnd, const StepArgs& a, int zchunks,
                               hipStream_t s) {
  return coll == KMI_TRT ? launch2_tile<ST, KMI_TRT>(tile, nts, bnd, a, zchunks, s)
                         : launch2_tile<ST, KMI_BGK>(tile, nts, bnd, a, zchunks, s);
}

hipError_t launch_step2(int st, int coll, int tile, int zchunks, bool nts, bool bnd, const StepArgs& a,
                        hipStream_t s) {
  switch (st) {
    case KMI_STORE_F64: return launch2_coll<KMI_STORE_F64>(coll, tile, nts, bnd, a, zchunks, s);
    case KMI_STORE_F32: return launch2_coll<KMI_STORE_F32>(coll, tile, nts, bnd, a, zchunks, s);
    case KMI_STORE_F32S: return launch2_coll<KMI_STORE_F32S>(coll, tile, nts, bnd, a, zchunks, s);
    default: return launch2_coll<KMI_STORE_F16S>(coll, tile, nts, bnd, a, zchunks, s);
  }
}

}  // namespace kmi
