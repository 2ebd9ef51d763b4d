// kmi_cell.h — per-cell arithmetic shared by the single-lattice and batched kernels:
// equilibria, moments, BGK / TRT / MRT, thermal BGK + Guo source, Zou–He closures and
// anti-bounce-back, each in the reference's evaluation order (see kmi_step.hip).
#pragma once

#include "kmi_common.h"

namespace kmi {

// Fused multiply-add a·b + c, rounded once.  The shifted storages (F32S, F16S), whose
// results are checked against FP64 with tolerances, use it explicitly in their collision
// arithmetic; the code is compiled without automatic contraction, so every kernel evaluates
// the same operations and the plain storages keep the reference's unfused order.
template <class R>
KMI_DEV R fmad(R a, R b, R c) {
  if constexpr (sizeof(R) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// Value a population takes after a store to and a load from storage ST (the multi-step
// kernels keep intermediate populations on chip in this rounding).
template <int ST, class R>
KMI_DEV R storage_round(R v) {
  if constexpr (ST == KMI_STORE_F16S) return float(_Float16(v * kF16Scale)) * kF16InvScale;
  else return v;
}

template <int Q, class R>
KMI_DEV R usq_of(R ux, R uy, R uz) {
  if constexpr (Q == 9) return ux * ux + uy * uy;
  else return ux * ux + uy * uy + uz * uz;
}

// ---- equilibria ----------------------------------------------------------------------
// Plain: feq_q = w ρ (1 + 3cu + 4.5cu² − 1.5u²).  Shifted: deq_q = w (δρ + ρ X_q).
// Each population q and its opposite share c·u up to sign: c_opp·u = −(c_q·u) exactly (the
// sums are negated term by term and round-to-nearest is symmetric), hence 3c_opp·u = −3c_q·u
// and (4.5 c_opp·u)(c_opp·u) = (4.5 c_q·u)(c_q·u) bit for bit, and (1 + (−t)) = (1 − t).  The
// pair is evaluated once; every value equals the reference's per-q expression.
template <int Q, bool SH, class R>
KMI_DEV void equilibria(R drho_or_rho, R rho, R ux, R uy, R uz, R (&fe)[Q]) {
  R usq;
  if constexpr (!SH) usq = usq_of<Q>(ux, uy, uz);
  else usq = Q == 9 ? fmad(uy, uy, ux * ux) : fmad(uz, uz, fmad(uy, uy, ux * ux));
  const R u15 = R(1.5) * usq;
  if constexpr (!SH) fe[0] = R(Lat<Q>::w[0]) * rho * (R(1) - u15);
  else fe[0] = R(Lat<Q>::w[0]) * fmad(rho, -u15, drho_or_rho);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = Lat<Q>::opp[q];
    if constexpr (q != 0 && q < o) {
      constexpr double w = Lat<Q>::w[q];
      const R cu = cdotu<Lat<Q>::cx[q], Lat<Q>::cy[q], Lat<Q>::cz[q]>(ux, uy, uz);
      const R t3 = R(3) * cu;
      if constexpr (!SH) {
        const R sq = R(4.5) * cu * cu;
        const R wr = R(w) * rho;
        fe[q] = wr * (((R(1) + t3) + sq) - u15);
        fe[o] = wr * (((R(1) - t3) + sq) - u15);
      } else {
        const R h = R(4.5) * cu;  // X_q = 3cu + 4.5cu² − 1.5u², deq = w (δρ + ρ X_q)
        fe[q] = R(w) * fmad(rho, fmad(h, cu, t3) - u15, drho_or_rho);
        fe[o] = R(w) * fmad(rho, fmad(h, cu, -t3) - u15, drho_or_rho);
      }
    }
  });
}

// Moments.  Plain: ρ = Σf, u = (Σcf)/ρ (macroscopic.jl).  Shifted: δρ = Σd, ρ = 1+δρ.
template <int Q, bool SH, class R>
KMI_DEV void moments(const R (&f)[Q], R& m0, R& rho, R& ux, R& uy, R& uz) {
  m0 = zeroth_moment<Q>(f);
  rho = SH ? R(1) + m0 : m0;
  R inv;
  if constexpr (SH && sizeof(R) == 4) inv = __builtin_amdgcn_rcpf(rho);  // 1 ulp, shifted only
  else inv = R(1) / rho;
  ux = first_moment<Q, 0>(f) * inv;
  uy = first_moment<Q, 1>(f) * inv;
  if constexpr (Q == 19) uz = first_moment<Q, 2>(f) * inv;
  else uz = R(0);
}

// Shifted BGK with deviations in the F16S storage scale (see collide_trt_shifted): bit for bit
// the scaled result of collide_bgk<Q, true>.
template <int Q, class R>
KMI_DEV void collide_bgk_f16u(const R (&fp)[Q], R (&fo)[Q], R om) {
  const R m0 = zeroth_moment<Q>(fp);
  const R rho = R(1) + m0 * R(kF16InvScale);
  const R inv = __builtin_amdgcn_rcpf(rho) * R(kF16InvScale);
  const R ux = first_moment<Q, 0>(fp) * inv, uy = first_moment<Q, 1>(fp) * inv;
  const R uz = Q == 19 ? first_moment<Q, 2>(fp) * inv : R(0);
  R fe[Q];
  equilibria<Q, true>(m0, rho * R(kF16Scale), ux, uy, uz, fe);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    fo[q] = fmad(-om, fp[q] - fe[q], fp[q]);
  });
}

template <int Q, bool SH, class R>
KMI_DEV void collide_bgk(const R (&fp)[Q], R (&fo)[Q], R om) {
  R m0, rho, ux, uy, uz, fe[Q];
  moments<Q, SH>(fp, m0, rho, ux, uy, uz);
  equilibria<Q, SH>(m0, rho, ux, uy, uz, fe);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (!SH) fo[q] = fp[q] - om * (fp[q] - fe[q]);
    else fo[q] = fmad(-om, fp[q] - fe[q], fp[q]);
  });
}

// Shifted-storage TRT in its even/odd form (the same relaxation as below, grouped per
// opposite pair): with S = d_q + d_q̄, D = d_q − d_q̄ of the deviations d = f − w,
//   d_q* = d_q − s⁺(S/2 − w E_q) − s⁻(D/2 − 3wρ c_q·u),   d_q̄* = the same with −s⁻ term,
// where w E_q = w(δρ + ρ(4.5(c_q·u)² − 1.5u²)) is the even part of the equilibrium deviation
// and 3wρ c_q·u its odd part; rest population d_0* = d_0 − s⁺(d_0 − w_0(δρ − 1.5ρu²)).
// About 16 % fewer operations per D3Q19 cell than evaluating both equilibria and relaxing
// each population; the plain storages keep the reference's order below.
//
// F16U = true: the deviations come in and go out in the F16S storage scale (d · 2^15), so
// the kernels skip a multiply per population on load, store and on-chip rounding.  Every
// operation is then the unscaled one times an exact power of two (ρ, u and c·u are formed
// as before), so the results are the unscaled collision's, scaled, bit for bit.
template <int Q, bool F16U = false, class R>
KMI_DEV void collide_trt_shifted(const R (&fp)[Q], R (&fo)[Q], R sp, R sm) {
  R m0, rho, ux, uy, uz;
  if constexpr (!F16U) {
    moments<Q, true>(fp, m0, rho, ux, uy, uz);
  } else {
    m0 = zeroth_moment<Q>(fp);  // 2^15 δρ
    rho = R(1) + m0 * R(kF16InvScale);
    const R inv = __builtin_amdgcn_rcpf(rho) * R(kF16InvScale);
    ux = first_moment<Q, 0>(fp) * inv;
    uy = first_moment<Q, 1>(fp) * inv;
    uz = Q == 19 ? first_moment<Q, 2>(fp) * inv : R(0);
  }
  const R rs = F16U ? rho * R(kF16Scale) : rho;  // ρ in the deviations' scale
  const R usq = Q == 9 ? fmad(uy, uy, ux * ux) : fmad(uz, uz, fmad(uy, uy, ux * ux));
  const R K = fmad(R(-1.5) * rs, usq, m0);  // δρ − 1.5ρu²
  const R r45 = R(4.5) * rs;
  const R hp = R(0.5) * sp, hm = R(0.5) * sm;
  fo[0] = fmad(-sp, fp[0] - R(Lat<Q>::w[0]) * K, fp[0]);
  sfor<Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = Lat<Q>::opp[q];
    if constexpr (q != 0 && q < o) {
      constexpr double w = Lat<Q>::w[q];
      const R cu = cdotu<Lat<Q>::cx[q], Lat<Q>::cy[q], Lat<Q>::cz[q]>(ux, uy, uz);
      const R E = fmad(r45, cu * cu, K);
      const R pe = fmad(hp, fp[q] + fp[o], -(R(w) * sp) * E);
      const R po = fmad(hm, fp[q] - fp[o], -((R(3.0 * w) * sm) * rs) * cu);
      fo[q] = fp[q] - (pe + po);
      fo[o] = fp[o] - (pe - po);
    }
  });
}

// f* = f − a(f − feq) − b(f̄ − feq̄), a = (s⁺+s⁻)/2, b = (s⁺−s⁻)/2 (fused_trt_2d.jl:150-166).
template <int Q, bool SH, class R>
KMI_DEV void collide_trt(const R (&fp)[Q], R (&fo)[Q], R sp, R sm) {
  if constexpr (SH) {
    collide_trt_shifted<Q>(fp, fo, sp, sm);
  } else {
    R m0, rho, ux, uy, uz, fe[Q];
    moments<Q, SH>(fp, m0, rho, ux, uy, uz);
    equilibria<Q, SH>(m0, rho, ux, uy, uz, fe);
    const R a = (sp + sm) * R(0.5);
    const R b = (sp - sm) * R(0.5);
    fo[0] = fp[0] - sp * (fp[0] - fe[0]);
    sfor<Q - 1>([&](auto qc) {
      constexpr int q = decltype(qc)::value + 1;
      constexpr int o = Lat<Q>::opp[q];
      fo[q] = fp[q] - a * (fp[q] - fe[q]) - b * (fp[o] - fe[o]);
    });
  }
}

// Shifted-storage D2Q9 MRT: the same relaxation as below on the moments of the deviations
// d = f − w, with the transforms written as butterflies over opposite pairs (axis pairs
// (1,3), (2,4), diagonal pairs (5,7), (6,8), 0-based) and fused multiply-adds, and
// q_eq = −ρu taken as −j; about half the floating-point operations of the reference's
// row-by-row products, which the plain storages keep.
template <class R>
KMI_DEV void collide_mrt_shifted(const R (&d)[9], R (&fo)[9], R s_e, R s_eps, R s_q, R s_nu) {
  const R a = d[1] + d[3], b = d[2] + d[4];  // axis pairs
  const R c = d[5] + d[7], g = d[6] + d[8];  // diagonal pairs
  const R dx = d[1] - d[3], dy = d[2] - d[4];
  const R dc = d[5] - d[7], dg = d[6] - d[8];
  const R sa = a + b, sd = c + g;
  const R m0 = d[0] + sa + sd;
  const R e = fmad(R(2), sd, fmad(R(-4), d[0], -sa));
  const R eps = fmad(R(4), d[0], fmad(R(-2), sa, sd));
  const R u1 = dc - dg, u2 = dc + dg;
  const R jx = dx + u1, jy = dy + u2;
  const R qx = fmad(R(-2), dx, u1), qy = fmad(R(-2), dy, u2);
  const R pxx = a - b, pxy = c - g;
  const R rho = R(1) + m0;
  R inv;
  if constexpr (sizeof(R) == 4) inv = __builtin_amdgcn_rcpf(rho);
  else inv = R(1) / rho;
  const R ux = jx * inv, uy = jy * inv;
  const R t = R(3) * rho * fmad(uy, uy, ux * ux);  // 3ρu²
  const R es = fmad(-s_e, e - fmad(R(-2), m0, t), e);  // e_eq + 2 = −2δρ + 3ρu²
  const R ep = fmad(-s_eps, eps - (m0 - t), eps);       // ε_eq − 1 = δρ − 3ρu²
  const R qxs = fmad(-s_q, qx + jx, qx);
  const R qys = fmad(-s_q, qy + jy, qy);
  const R ps = fmad(-s_nu, pxx - fmad(jx, ux, -(jy * uy)), pxx);  // p_xx,eq = ρ(ux² − uy²)
  const R pxys = fmad(-s_nu, pxy - jx * uy, pxy);                  // p_xy,eq = ρ ux uy
  const R r9 = R(1.0 / 9) * m0;
  fo[0] = fmad(R(1.0 / 9), ep - es, r9);
  const R A = fmad(R(-1.0 / 36), es, fmad(R(-1.0 / 18), ep, r9));
  const R B = fmad(R(1.0 / 18), es, fmad(R(1.0 / 36), ep, r9));
  const R tx = R(1.0 / 6) * (jx - qxs), ty = R(1.0 / 6) * (jy - qys), p4 = R(0.25) * ps;
  const R ap = A + p4, am = A - p4;
  fo[1] = ap + tx;
  fo[3] = ap - tx;
  fo[2] = am + ty;
  fo[4] = am - ty;
  const R X = fmad(R(1.0 / 6), jx, R(1.0 / 12) * qxs);
  const R Y = fmad(R(1.0 / 6), jy, R(1.0 / 12) * qys);
  const R P = R(0.25) * pxys;
  const R bp = B + X, bm = B - X, yp = Y + P, ym = Y - P;
  fo[5] = bp + yp;
  fo[6] = bm + ym;
  fo[7] = bm - ym;
  fo[8] = bp - yp;
}

// D2Q9 MRT (Lallemand–Luo), collide_mrt_2d.jl:56-107.  Shifted form: the moments of
// the deviation, with m_eq shifted by the rest-state moments (e = −2, ε = 1).
template <bool SH, class R>
KMI_DEV void collide_mrt(const R (&fp)[9], R (&fo)[9], R s_e, R s_eps, R s_q, R s_nu) {
  if constexpr (SH) {
    collide_mrt_shifted(fp, fo, s_e, s_eps, s_q, s_nu);
    return;
  }
  const R f1 = fp[0], f2 = fp[1], f3 = fp[2], f4 = fp[3], f5 = fp[4], f6 = fp[5], f7 = fp[6],
          f8 = fp[7], f9 = fp[8];
  const R m0 = f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8 + f9;
  const R e = -R(4) * f1 - f2 - f3 - f4 - f5 + R(2) * (f6 + f7 + f8 + f9);
  const R eps = R(4) * f1 - R(2) * (f2 + f3 + f4 + f5) + f6 + f7 + f8 + f9;
  const R jx = f2 - f4 + f6 - f7 - f8 + f9;
  const R qx = -R(2) * f2 + R(2) * f4 + f6 - f7 - f8 + f9;
  const R jy = f3 - f5 + f6 + f7 - f8 - f9;
  const R qy = -R(2) * f3 + R(2) * f5 + f6 + f7 - f8 - f9;
  const R pxx = f2 - f3 + f4 - f5;
  const R pxy = f6 - f7 + f8 - f9;
  const R rho = SH ? R(1) + m0 : m0;
  const R inv = R(1) / rho;
  const R ux = jx * inv;
  const R uy = jy * inv;
  const R usq = ux * ux + uy * uy;
  R e_eq, eps_eq;
  if constexpr (!SH) {
    e_eq = -R(2) * rho + R(3) * rho * usq;
    eps_eq = rho - R(3) * rho * usq;
  } else {
    e_eq = -R(2) * m0 + R(3) * rho * usq;  // e_eq(ρ) + 2
    eps_eq = m0 - R(3) * rho * usq;        // ε_eq(ρ) − 1
  }
  const R qx_eq = -rho * ux;
  const R qy_eq = -rho * uy;
  const R pxx_eq = rho * (ux * ux - uy * uy);
  const R pxy_eq = rho * ux * uy;
  const R es = e - s_e * (e - e_eq);
  const R ep = eps - s_eps * (eps - eps_eq);
  const R qxs = qx - s_q * (qx - qx_eq);
  const R qys = qy - s_q * (qy - qy_eq);
  const R ps = pxx - s_nu * (pxx - pxx_eq);
  const R pxys = pxy - s_nu * (pxy - pxy_eq);
  const R r = m0, jxs = jx, jys = jy;
  fo[0] = R(1.0 / 9) * r - R(1.0 / 9) * es + R(1.0 / 9) * ep;
  fo[1] = R(1.0 / 9) * r - R(1.0 / 36) * es - R(1.0 / 18) * ep + R(1.0 / 6) * jxs -
          R(1.0 / 6) * qxs + R(1.0 / 4) * ps;
  fo[2] = R(1.0 / 9) * r - R(1.0 / 36) * es - R(1.0 / 18) * ep + R(1.0 / 6) * jys -
          R(1.0 / 6) * qys - R(1.0 / 4) * ps;
  fo[3] = R(1.0 / 9) * r - R(1.0 / 36) * es - R(1.0 / 18) * ep - R(1.0 / 6) * jxs +
          R(1.0 / 6) * qxs + R(1.0 / 4) * ps;
  fo[4] = R(1.0 / 9) * r - R(1.0 / 36) * es - R(1.0 / 18) * ep - R(1.0 / 6) * jys +
          R(1.0 / 6) * qys - R(1.0 / 4) * ps;
  fo[5] = R(1.0 / 9) * r + R(1.0 / 18) * es + R(1.0 / 36) * ep + R(1.0 / 6) * jxs +
          R(1.0 / 12) * qxs + R(1.0 / 6) * jys + R(1.0 / 12) * qys + R(1.0 / 4) * pxys;
  fo[6] = R(1.0 / 9) * r + R(1.0 / 18) * es + R(1.0 / 36) * ep - R(1.0 / 6) * jxs -
          R(1.0 / 12) * qxs + R(1.0 / 6) * jys + R(1.0 / 12) * qys - R(1.0 / 4) * pxys;
  fo[7] = R(1.0 / 9) * r + R(1.0 / 18) * es + R(1.0 / 36) * ep - R(1.0 / 6) * jxs -
          R(1.0 / 12) * qxs - R(1.0 / 6) * jys - R(1.0 / 12) * qys + R(1.0 / 4) * pxys;
  fo[8] = R(1.0 / 9) * r + R(1.0 / 18) * es + R(1.0 / 36) * ep + R(1.0 / 6) * jxs +
          R(1.0 / 12) * qxs - R(1.0 / 6) * jys - R(1.0 / 12) * qys - R(1.0 / 4) * pxys;
}

// Thermal BGK, geq = w T (1 + 3 c·u) (thermal_2d.jl:16-60).  Shifted: g − w T0 stored,
// geq − w T0 = w (δT + T·3c·u).
template <bool SH, class R>
KMI_DEV void collide_thermal(const R (&gp)[9], R (&go)[9], R ux, R uy, R omT, R T0) {
  const R s = zeroth_moment<9>(gp);
  if constexpr (!SH) {
    const R Temp = s;
    const R t3 = R(3);
    R geq;
    geq = R(4.0 / 9.0) * Temp;
    go[0] = gp[0] - omT * (gp[0] - geq);
    geq = R(1.0 / 9.0) * Temp * (R(1) + t3 * ux);
    go[1] = gp[1] - omT * (gp[1] - geq);
    geq = R(1.0 / 9.0) * Temp * (R(1) + t3 * uy);
    go[2] = gp[2] - omT * (gp[2] - geq);
    geq = R(1.0 / 9.0) * Temp * (R(1) - t3 * ux);
    go[3] = gp[3] - omT * (gp[3] - geq);
    geq = R(1.0 / 9.0) * Temp * (R(1) - t3 * uy);
    go[4] = gp[4] - omT * (gp[4] - geq);
    geq = R(1.0 / 36.0) * Temp * (R(1) + t3 * (ux + uy));
    go[5] = gp[5] - omT * (gp[5] - geq);
    geq = R(1.0 / 36.0) * Temp * (R(1) + t3 * (-ux + uy));
    go[6] = gp[6] - omT * (gp[6] - geq);
    geq = R(1.0 / 36.0) * Temp * (R(1) + t3 * (-ux - uy));
    go[7] = gp[7] - omT * (gp[7] - geq);
    geq = R(1.0 / 36.0) * Temp * (R(1) + t3 * (ux - uy));
    go[8] = gp[8] - omT * (gp[8] - geq);
  } else {
    const R dT = s;
    const R Tt = T0 + dT;
    sfor<9>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr double w = Lat<9>::w[q];
      R geq;
      if constexpr (q == 0) geq = R(w) * dT;
      else geq = R(w) * (dT + Tt * (R(3) * cdotu<Lat<9>::cx[q], Lat<9>::cy[q], 0>(ux, uy, R(0))));
      go[q] = gp[q] - omT * (gp[q] - geq);
    });
  }
}

// Thermal BGK on D3Q19 g (collide_thermal_3d_kernel!, thermal_3d.jl:12-107):
// T = g1 + … + g19, geq = w T (1 + 3 c·u) with c·u in the reference's written form
// (1 + 3(−u) equals the written 1 − 3u bit for bit).  Shifted: as collide_thermal.
template <bool SH, class R>
KMI_DEV void collide_thermal19(const R (&gp)[19], R (&go)[19], R ux, R uy, R uz, R omT, R T0) {
  const R s = zeroth_moment<19>(gp);
  const R Tt = SH ? T0 + s : s;
  sfor<19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr double w = Lat<19>::w[q];
    R geq;
    if constexpr (q == 0) geq = R(w) * s;
    else if constexpr (!SH)
      geq = R(w) * s * (R(1) + R(3) * cdotu<Lat<19>::cx[q], Lat<19>::cy[q], Lat<19>::cz[q]>(ux, uy, uz));
    else
      geq = R(w) * (s + Tt * (R(3) * cdotu<Lat<19>::cx[q], Lat<19>::cy[q], Lat<19>::cz[q]>(ux, uy, uz)));
    go[q] = gp[q] - omT * (gp[q] - geq);
  });
}

// (c_a − u_a) as the reference writes it: `(-ux)`, `(one(T)-ux)`, `(-one(T)-ux)`.
template <int C, class R>
KMI_DEV R c_minus_u(R u) {
  if constexpr (C == 0) return -u;
  else if constexpr (C > 0) return R(1) - u;
  else return -R(1) - u;
}

// BGK + Guo source on D3Q19: collide_guo_3d_kernel! (collide_guo_3d.jl:21-185), its
// per-node-field twin (:213-320) and collide_boussinesq_3d_kernel! (thermal_3d.jl:574-712):
//   S_q = w ((c−u)·F) 3 + w (c·u)(c·F) 9,   f* = f − ω(f − feq) + (1 − ω/2) S_q,
// ux, uy, uz the force-shifted velocities; equilibria paired as in `equilibria`.
template <bool SH, class R>
KMI_DEV void collide_guo19(const R (&fp)[19], R (&fo)[19], R m0, R rho, R ux, R uy, R uz, R fx,
                           R fy, R fz, R om) {
  const R guo = R(1) - om / R(2);
  R fe[19];
  equilibria<19, SH>(m0, rho, ux, uy, uz, fe);
  sfor<19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = Lat<19>::cx[q], cy = Lat<19>::cy[q], cz = Lat<19>::cz[q];
    constexpr double w = Lat<19>::w[q];
    R Sq = R(w) * (c_minus_u<cx>(ux) * fx + c_minus_u<cy>(uy) * fy + c_minus_u<cz>(uz) * fz) * R(3);
    if constexpr (q != 0)
      Sq = Sq + R(w) * cdotu<cx, cy, cz>(ux, uy, uz) * cdotu<cx, cy, cz>(fx, fy, fz) * R(9);
    fo[q] = fp[q] - om * (fp[q] - fe[q]) + guo * Sq;
  });
}

// Body force at cell c: the per-node field when one is set, else the constant force.
template <int Q, class R>
KMI_DEV void node_force(const StepArgs& a, long long c, R& fx, R& fy, R& fz) {
  if (a.ffield) {
    const R* F = static_cast<const R*>(a.ffield);
    fx = F[c];
    fy = F[c + a.plane];
    fz = Q == 19 ? F[c + 2 * a.plane] : R(0);
  } else {
    fx = R(a.fx);
    fy = R(a.fy);
    fz = R(a.fz);
  }
}

// BGK + Guo source for a force (fx, fy): collide_guo_2d.jl:5-72 (constant force),
// thermal_2d.jl:443-478 and fused_thermal_2d.jl:116-157 (fx = 0, buoyancy fy).  ux, uy are
// the force-shifted velocities.  The reference writes the c_q = −e_x, −e_y terms as
// (−u)(−F) in collide_guo_2d and as u F in the Boussinesq kernels; negation is exact, so
// both are the same floating-point value.
template <bool SH, class R>
KMI_DEV void collide_guo(const R (&fp)[9], R (&fo)[9], R m0, R rho, R ux, R uy, R fx, R fy,
                         R om) {
  const R usq = ux * ux + uy * uy;
  const R guo = R(1) - om / R(2);
  const R t3 = R(3);
  R fe[9];
  equilibria<9, SH>(m0, rho, ux, uy, R(0), fe);
  R Sq;
  Sq = R(4.0 / 9.0) * ((-ux) * fx + (-uy) * fy) * t3;
  fo[0] = fp[0] - om * (fp[0] - fe[0]) + guo * Sq;
  Sq = R(1.0 / 9.0) * ((R(1) - ux) * fx + (-uy) * fy) * t3 + R(1.0 / 9.0) * ux * fx * R(9);
  fo[1] = fp[1] - om * (fp[1] - fe[1]) + guo * Sq;
  Sq = R(1.0 / 9.0) * ((-ux) * fx + (R(1) - uy) * fy) * t3 + R(1.0 / 9.0) * uy * fy * R(9);
  fo[2] = fp[2] - om * (fp[2] - fe[2]) + guo * Sq;
  Sq = R(1.0 / 9.0) * ((-R(1) - ux) * fx + (-uy) * fy) * t3 + R(1.0 / 9.0) * ux * fx * R(9);
  fo[3] = fp[3] - om * (fp[3] - fe[3]) + guo * Sq;
  Sq = R(1.0 / 9.0) * ((-ux) * fx + (-R(1) - uy) * fy) * t3 + R(1.0 / 9.0) * uy * fy * R(9);
  fo[4] = fp[4] - om * (fp[4] - fe[4]) + guo * Sq;
  Sq = R(1.0 / 36.0) * ((R(1) - ux) * fx + (R(1) - uy) * fy) * t3 +
       R(1.0 / 36.0) * (ux + uy) * (fx + fy) * R(9);
  fo[5] = fp[5] - om * (fp[5] - fe[5]) + guo * Sq;
  Sq = R(1.0 / 36.0) * ((-R(1) - ux) * fx + (R(1) - uy) * fy) * t3 +
       R(1.0 / 36.0) * (-ux + uy) * (-fx + fy) * R(9);
  fo[6] = fp[6] - om * (fp[6] - fe[6]) + guo * Sq;
  Sq = R(1.0 / 36.0) * ((-R(1) - ux) * fx + (-R(1) - uy) * fy) * t3 +
       R(1.0 / 36.0) * (-ux - uy) * (-fx - fy) * R(9);
  fo[7] = fp[7] - om * (fp[7] - fe[7]) + guo * Sq;
  Sq = R(1.0 / 36.0) * ((R(1) - ux) * fx + (-R(1) - uy) * fy) * t3 +
       R(1.0 / 36.0) * (ux - uy) * (fx - fy) * R(9);
  fo[8] = fp[8] - om * (fp[8] - fe[8]) + guo * Sq;
}

// ---- Zou–He closures on the node's post-stream populations (D2Q9) -----------------------
template <bool SH, class R>
KMI_DEV void zh_north(R (&f)[9], R u) {  // boundary_2d.jl:13-44
  const R f1 = f[0], f2 = f[1], f3 = f[2], f4 = f[3], f6 = f[5], f7 = f[6];
  R rw;
  if constexpr (!SH) rw = (f1 + f2 + f4 + R(2) * (f3 + f6 + f7)) / (R(1) + u * R(0));
  else rw = R(1) + (f1 + f2 + f4 + R(2) * (f3 + f6 + f7));
  f[4] = f3;
  f[8] = f7 - R(0.5) * (f2 - f4) + R(0.5) * rw * u;
  f[7] = f6 + R(0.5) * (f2 - f4) - R(0.5) * rw * u;
}
template <bool SH, class R>
KMI_DEV void zh_south(R (&f)[9], R u) {  // boundary_2d.jl:135-150
  const R f1 = f[0], f2 = f[1], f4 = f[3], f5 = f[4], f8 = f[7], f9 = f[8];
  R rw = f1 + f2 + f4 + R(2) * (f5 + f8 + f9);
  if constexpr (SH) rw = R(1) + rw;
  f[2] = f5;
  f[5] = f8 - R(0.5) * (f2 - f4) + R(0.5) * rw * u;
  f[6] = f9 + R(0.5) * (f2 - f4) - R(0.5) * rw * u;
}
template <bool SH, class R>
KMI_DEV void zh_west(R (&f)[9], R u) {  // boundary_2d.jl:171-185
  const R f1 = f[0], f3 = f[2], f4 = f[3], f5 = f[4], f7 = f[6], f8 = f[7];
  R s = f1 + f3 + f5 + R(2) * (f4 + f7 + f8);
  if constexpr (SH) s = R(1) + s;
  const R rw = s / (R(1) - u);
  f[1] = f4 + R(2.0 / 3.0) * rw * u;
  f[5] = f8 - R(0.5) * (f3 - f5) + R(1.0 / 6.0) * rw * u;
  f[8] = f7 + R(0.5) * (f3 - f5) + R(1.0 / 6.0) * rw * u;
}
// rho_m1 = ρ_out − 1, used by the shifted form only.
template <bool SH, class R>
KMI_DEV void zh_pressure_east(R (&f)[9], R rho, R rho_m1) {  // boundary_2d.jl:206-223
  const R f1 = f[0], f2 = f[1], f3 = f[2], f5 = f[4], f6 = f[5], f9 = f[8];
  R ux;
  if constexpr (!SH) ux = -R(1) + (f1 + f3 + f5 + R(2) * (f2 + f6 + f9)) / rho;
  else ux = ((f1 + f3 + f5 + R(2) * (f2 + f6 + f9)) - rho_m1) / rho;
  f[3] = f2 - R(2.0 / 3.0) * rho * ux;
  f[6] = f9 - R(0.5) * (f3 - f5) - R(1.0 / 6.0) * rho * ux;
  f[7] = f6 + R(0.5) * (f3 - f5) - R(1.0 / 6.0) * rho * ux;
}
template <bool SH, class R>
KMI_DEV void zh_pressure_west(R (&f)[9], R rho, R rho_m1) {  // boundary_2d.jl:264-281
  const R f1 = f[0], f3 = f[2], f4 = f[3], f5 = f[4], f7 = f[6], f8 = f[7];
  R ux;
  if constexpr (!SH) ux = R(1) - (f1 + f3 + f5 + R(2) * (f4 + f7 + f8)) / rho;
  else ux = (rho_m1 - (f1 + f3 + f5 + R(2) * (f4 + f7 + f8))) / rho;
  f[1] = f4 + R(2.0 / 3.0) * rho * ux;
  f[5] = f8 - R(0.5) * (f3 - f5) + R(1.0 / 6.0) * rho * ux;
  f[8] = f7 + R(0.5) * (f3 - f5) + R(1.0 / 6.0) * rho * ux;
}

// Per-node Zou–He velocity (u_x, u_y) on north / south / west (boundary_spatial_2d.jl:6-90).
template <bool SH, class R>
KMI_DEV void zh_north_profile(R (&f)[9], R ux, R uy) {
  R s = f[0] + f[1] + f[3] + R(2) * (f[2] + f[5] + f[6]);
  if constexpr (SH) s = R(1) + s;
  const R rw = s / (R(1) + uy);
  f[4] = f[2] - R(2.0 / 3.0) * rw * uy;
  f[8] = f[6] - R(0.5) * (f[1] - f[3]) + R(0.5) * rw * ux - R(1.0 / 6.0) * rw * uy;
  f[7] = f[5] + R(0.5) * (f[1] - f[3]) - R(0.5) * rw * ux - R(1.0 / 6.0) * rw * uy;
}

template <bool SH, class R>
KMI_DEV void zh_south_profile(R (&f)[9], R ux, R uy) {
  R s = f[0] + f[1] + f[3] + R(2) * (f[4] + f[7] + f[8]);
  if constexpr (SH) s = R(1) + s;
  const R rw = s / (R(1) - uy);
  f[2] = f[4] + R(2.0 / 3.0) * rw * uy;
  f[5] = f[7] - R(0.5) * (f[1] - f[3]) + R(0.5) * rw * ux + R(1.0 / 6.0) * rw * uy;
  f[6] = f[8] + R(0.5) * (f[1] - f[3]) - R(0.5) * rw * ux + R(1.0 / 6.0) * rw * uy;
}

template <bool SH, class R>
KMI_DEV void zh_west_profile(R (&f)[9], R ux, R uy) {
  R s = f[0] + f[2] + f[4] + R(2) * (f[3] + f[6] + f[7]);
  if constexpr (SH) s = R(1) + s;
  const R rw = s / (R(1) - ux);
  f[1] = f[3] + R(2.0 / 3.0) * rw * ux;
  f[5] = f[7] - R(0.5) * (f[2] - f[4]) + R(1.0 / 6.0) * rw * ux + R(0.5) * rw * uy;
  f[8] = f[6] + R(0.5) * (f[2] - f[4]) + R(1.0 / 6.0) * rw * ux - R(0.5) * rw * uy;
}

// ---- D3Q19 Zou–He closures (boundary_3d.jl:39-300, ZouHeFace3D restated 0-based) ------------
struct ZH3Face {
  int axis, sign;      // normal axis (0 x, 1 y, 2 z); −1 min face, +1 max face
  int qa, qao;         // unknown axis population and its known opposite
  int qd[4], qdo[4];   // unknown diagonals (tangent-1 pair, tangent-2 pair) and opposites
  int t1p, t1m, t2p, t2m;  // axis populations along the two tangents
  int par[9];          // populations with zero normal component
  int out[5];          // known populations leaving through the face
};
// Face order W, E, S, N, B, T; tangents: x-faces (y, z), y-faces (x, z), z-faces (x, y).
constexpr ZH3Face kZH3[6] = {
    {0, -1, 1, 2, {7, 9, 11, 13}, {10, 8, 14, 12}, 3, 4, 5, 6, {0, 3, 4, 5, 6, 15, 16, 17, 18}, {2, 8, 10, 12, 14}},
    {0, +1, 2, 1, {8, 10, 12, 14}, {9, 7, 13, 11}, 3, 4, 5, 6, {0, 3, 4, 5, 6, 15, 16, 17, 18}, {1, 7, 9, 11, 13}},
    {1, -1, 3, 4, {7, 8, 15, 17}, {10, 9, 18, 16}, 1, 2, 5, 6, {0, 1, 2, 5, 6, 11, 12, 13, 14}, {4, 9, 10, 16, 18}},
    {1, +1, 4, 3, {9, 10, 16, 18}, {8, 7, 17, 15}, 1, 2, 5, 6, {0, 1, 2, 5, 6, 11, 12, 13, 14}, {3, 7, 8, 15, 17}},
    {2, -1, 5, 6, {11, 12, 15, 16}, {14, 13, 18, 17}, 1, 2, 3, 4, {0, 1, 2, 3, 4, 7, 8, 9, 10}, {6, 13, 14, 17, 18}},
    {2, +1, 6, 5, {13, 14, 17, 18}, {12, 11, 16, 15}, 1, 2, 3, 4, {0, 1, 2, 3, 4, 7, 8, 9, 10}, {5, 11, 12, 15, 16}}};

// Σ parallel + 2 Σ outgoing, summed in table order from zero (the reference's loops).
template <int F, class R>
KMI_DEV R zh3_sum(const R (&f)[19]) {
  constexpr ZH3Face Z = kZH3[F];
  R sp = R(0), so = R(0);
  sfor<9>([&](auto n) { sp = sp + f[Z.par[decltype(n)::value]]; });
  sfor<5>([&](auto n) { so = so + f[Z.out[decltype(n)::value]]; });
  return sp + R(2) * so;
}

// zou_he_velocity_3d_kernel! (boundary_3d.jl:152-232): un, ut1, ut2 are the velocity
// components along the normal and the two tangents.
template <bool SH, int F, class R>
KMI_DEV void zh3_velocity(R (&f)[19], R un, R ut1, R ut2) {
  constexpr ZH3Face Z = kZH3[F];
  const R sg = R(Z.sign);
  const R denom = R(1) + sg * un;
  R s = zh3_sum<F>(f);
  if constexpr (SH) s = R(1) + s;
  const R rw = s / denom;
  f[Z.qa] = f[Z.qao] - sg * R(1.0 / 3.0) * rw * un;
  const R d1 = f[Z.par[5]], d2 = f[Z.par[6]], d3 = f[Z.par[7]], d4 = f[Z.par[8]];
  const R t1 = (f[Z.t1p] - f[Z.t1m]) + (d1 - d2 + d3 - d4);
  const R t2 = (f[Z.t2p] - f[Z.t2m]) + (d1 + d2 - d3 - d4);
  f[Z.qd[0]] = f[Z.qdo[0]] - R(0.5) * t1 + R(0.5) * rw * ut1 - sg * R(1.0 / 6.0) * rw * un;
  f[Z.qd[1]] = f[Z.qdo[1]] + R(0.5) * t1 - R(0.5) * rw * ut1 - sg * R(1.0 / 6.0) * rw * un;
  f[Z.qd[2]] = f[Z.qdo[2]] - R(0.5) * t2 + R(0.5) * rw * ut2 - sg * R(1.0 / 6.0) * rw * un;
  f[Z.qd[3]] = f[Z.qdo[3]] + R(0.5) * t2 - R(0.5) * rw * ut2 - sg * R(1.0 / 6.0) * rw * un;
}

// zou_he_pressure_3d_kernel! (boundary_3d.jl:234-290); rho_m1 = ρ − 1 (shifted form only).
template <bool SH, int F, class R>
KMI_DEV void zh3_pressure(R (&f)[19], R rho, R rho_m1) {
  constexpr ZH3Face Z = kZH3[F];
  const R sg = R(Z.sign);
  const R s = zh3_sum<F>(f);
  R un;
  if constexpr (!SH) un = -sg * (R(1) - s / rho);
  else un = -sg * ((rho_m1 - s) / rho);
  f[Z.qa] = f[Z.qao] - sg * R(1.0 / 3.0) * rho * un;
  const R t1 = f[Z.t1p] - f[Z.t1m];
  const R t2 = f[Z.t2p] - f[Z.t2m];
  f[Z.qd[0]] = f[Z.qdo[0]] - R(0.5) * t1 - sg * R(1.0 / 6.0) * rho * un;
  f[Z.qd[1]] = f[Z.qdo[1]] + R(0.5) * t1 - sg * R(1.0 / 6.0) * rho * un;
  f[Z.qd[2]] = f[Z.qdo[2]] - R(0.5) * t2 - sg * R(1.0 / 6.0) * rho * un;
  f[Z.qd[3]] = f[Z.qdo[3]] + R(0.5) * t2 - sg * R(1.0 / 6.0) * rho * un;
}

template <bool SH, int F, class R>
KMI_DEV void zh3_apply(R (&f)[19], const FaceArgs& Fa) {
  constexpr ZH3Face Z = kZH3[F];
  constexpr int a1 = Z.axis == 0 ? 1 : 0, a2 = Z.axis == 2 ? 1 : 2;
  if (Fa.kind == KMI_FACE_ZOUHE_VELOCITY)
    zh3_velocity<SH, F>(f, R(Fa.u[Z.axis]), R(Fa.u[a1]), R(Fa.u[a2]));
  else
    zh3_pressure<SH, F>(f, R(Fa.rho), R(Fa.rho - 1.0));
}

// ---- face rebuild (apply_bc_rebuild_2d!/_3d!, src/bc/rebuild_2d.jl, rebuild_3d.jl) ---------
// A rebuilt face node takes its pulled populations (the unknown ones are overwritten), the
// Zou–He closure of the rebuild kernels and a local TRT collision; plain populations.

// Regularised TRT of one D2Q9 node (_trt_collide_local, src/bc/specs.jl:7-50): the
// non-equilibrium part is rebuilt from the stress Π alone, then relaxed with TRT rates.
template <class R>
KMI_DEV void rebuild_collide9(const R (&f)[9], R (&o)[9], R sp, R sm) {
  const R rho = zeroth_moment<9>(f);
  const R ux = (f[1] - f[3] + f[5] - f[7] + f[8] - f[6]) / rho;
  const R uy = (f[2] - f[4] + f[5] - f[7] + f[6] - f[8]) / rho;
  R fe[9];
  equilibria<9, false>(rho, rho, ux, uy, R(0), fe);
  const R Pxx = (f[1] - fe[1]) + (f[3] - fe[3]) + (f[5] - fe[5]) + (f[6] - fe[6]) + (f[7] - fe[7]) + (f[8] - fe[8]);
  const R Pyy = (f[2] - fe[2]) + (f[4] - fe[4]) + (f[5] - fe[5]) + (f[6] - fe[6]) + (f[7] - fe[7]) + (f[8] - fe[8]);
  const R Pxy = (f[5] - fe[5]) - (f[6] - fe[6]) + (f[7] - fe[7]) - (f[8] - fe[8]);
  const R h = R(0.5);
  const R n0 = -h * R(2.0 / 9.0) * (Pxx + Pyy);
  const R n1 = h * R(1.0 / 9.0) * (R(2) * Pxx - Pyy);
  const R n2 = h * R(1.0 / 9.0) * (-Pxx + R(2) * Pyy);
  const R n5 = h * R(1.0 / 36.0) * (Pxx + Pyy) + R(1.0 / 4.0) * Pxy;
  const R n6 = h * R(1.0 / 36.0) * (Pxx + Pyy) - R(1.0 / 4.0) * Pxy;
  const R a = (sp + sm) * h, b = (sp - sm) * h;
  // n3 = n1, n4 = n2, n7 = n5, n8 = n6
  o[0] = fe[0] + (R(1) - sp) * n0;
  o[1] = fe[1] + (R(1) - a) * n1 - b * n1;
  o[2] = fe[2] + (R(1) - a) * n2 - b * n2;
  o[3] = fe[3] + (R(1) - a) * n1 - b * n1;
  o[4] = fe[4] + (R(1) - a) * n2 - b * n2;
  o[5] = fe[5] + (R(1) - a) * n5 - b * n5;
  o[6] = fe[6] + (R(1) - a) * n6 - b * n6;
  o[7] = fe[7] + (R(1) - a) * n5 - b * n5;
  o[8] = fe[8] + (R(1) - a) * n6 - b * n6;
}

// TRT of one D3Q19 node with the rebuild kernels' moments (_trt_collide_local_3d,
// src/bc/specs.jl:52-108): the velocity sums are grouped by opposite pairs and divided by ρ.
template <class R>
KMI_DEV void rebuild_collide19(const R (&f)[19], R (&o)[19], R sp, R sm) {
  const R rho = zeroth_moment<19>(f);
  const R ux = ((f[1] - f[2]) + (f[7] - f[10]) + (f[9] - f[8]) + (f[11] - f[14]) + (f[13] - f[12])) / rho;
  const R uy = ((f[3] - f[4]) + (f[7] - f[10]) + (f[8] - f[9]) + (f[15] - f[18]) + (f[17] - f[16])) / rho;
  const R uz = ((f[5] - f[6]) + (f[11] - f[14]) + (f[12] - f[13]) + (f[15] - f[18]) + (f[16] - f[17])) / rho;
  const R a = (sp + sm) * R(0.5), b = (sp - sm) * R(0.5);
  R fe[19];
  equilibria<19, false>(rho, rho, ux, uy, uz, fe);
  o[0] = f[0] - sp * (f[0] - fe[0]);
  sfor<18>([&](auto qc) {
    constexpr int q = decltype(qc)::value + 1;
    constexpr int p = Lat<19>::opp[q];
    o[q] = f[q] - a * (f[q] - fe[q]) - b * (f[p] - fe[p]);
  });
}

// Zou–He closures of the rebuild kernels (rebuild_2d.jl:1-50, rebuild_3d.jl:1-89):
// west normal velocity u / east density rho, zero tangential velocity.
template <class R>
KMI_DEV void rebuild_zh9(R (&f)[9], int kind, R v) {
  const R c23 = R(2.0 / 3.0), c16 = R(1.0 / 6.0), h = R(0.5);
  if (kind == 1) {
    const R rw = (f[0] + f[2] + f[4] + R(2) * (f[3] + f[6] + f[7])) / (R(1) - v);
    f[1] = f[3] + c23 * rw * v;
    f[5] = f[7] - h * (f[2] - f[4]) + c16 * rw * v;
    f[8] = f[6] + h * (f[2] - f[4]) + c16 * rw * v;
  } else {
    const R u = -R(1) + (f[0] + f[2] + f[4] + R(2) * (f[1] + f[5] + f[8])) / v;
    f[3] = f[1] - c23 * v * u;
    f[6] = f[8] - h * (f[2] - f[4]) - c16 * v * u;
    f[7] = f[5] + h * (f[2] - f[4]) - c16 * v * u;
  }
}

template <class R>
KMI_DEV void rebuild_zh19(R (&f)[19], int kind, R v) {
  const R c13 = R(1.0 / 3.0), c16 = R(1.0 / 6.0), h = R(0.5);
  const R par = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
  const R t1 = f[3] - f[4], t2 = f[5] - f[6];
  if (kind == 1) {
    const R out = f[2] + f[8] + f[10] + f[12] + f[14];
    const R rw = (par + R(2) * out) / (R(1) - v);
    f[1] = f[2] + c13 * rw * v;
    f[7] = f[10] - h * t1 + c16 * rw * v;
    f[9] = f[8] + h * t1 + c16 * rw * v;
    f[11] = f[14] - h * t2 + c16 * rw * v;
    f[13] = f[12] + h * t2 + c16 * rw * v;
  } else {
    const R out = f[1] + f[7] + f[9] + f[11] + f[13];
    const R u = -(R(1) - (par + R(2) * out) / v);
    f[2] = f[1] - c13 * v * u;
    f[8] = f[9] - h * t1 - c16 * v * u;
    f[10] = f[7] + h * t1 - c16 * v * u;
    f[12] = f[13] - h * t2 - c16 * v * u;
    f[14] = f[11] + h * t2 - c16 * v * u;
  }
}

// kind 1: west velocity v = u_n; kind 2: east pressure v = rho.  Shifted storage is
// converted to plain populations and back around the rebuild.
template <int Q, bool SH, class R>
KMI_DEV void rebuild_node(R (&fp)[Q], R (&fo)[Q], int kind, R v, R sp, R sm) {
  if constexpr (SH) sfor<Q>([&](auto qc) { fp[decltype(qc)::value] += R(Lat<Q>::w[decltype(qc)::value]); });
  if constexpr (Q == 9) {
    rebuild_zh9(fp, kind, v);
    rebuild_collide9(fp, fo, sp, sm);
  } else {
    rebuild_zh19(fp, kind, v);
    rebuild_collide19(fp, fo, sp, sm);
  }
  if constexpr (SH) sfor<Q>([&](auto qc) { fo[decltype(qc)::value] -= R(Lat<Q>::w[decltype(qc)::value]); });
}

// ---- interpolated bounce-back on cut links (LI-BB V2 pre-phase) ------------------------
// Order in which the reference's pre-phase visits the links: opposite pairs, q before
// opp(q) (dsl/bricks.jl:518-581, dsl/bricks_3d.jl:368-473).  A pair with both links cut
// substitutes the second population from the first substitution, as the reference does.
template <int Q>
struct LibbOrder;
template <>
struct LibbOrder<9> {
  static constexpr int q[8] = {1, 3, 2, 4, 5, 7, 6, 8};
};
template <>
struct LibbOrder<19> {
  static constexpr int q[18] = {1, 2, 3, 4, 5, 6, 7, 10, 8, 9, 11, 14, 12, 13, 15, 18, 16, 17};
};

// _libb_branch (li_bb_2d.jl:70-83).
template <class R>
KMI_DEV R libb_branch(R qw, R here, R back, R bar, R d) {
  if (qw <= R(0.5)) return R(2) * qw * here + (R(1) - R(2) * qw) * back + d;
  const R inv = R(1) / (R(2) * qw);
  return inv * here + (R(1) - inv) * bar + d * inv;
}

// fp: pulled populations of a cut cell; fi/self/P locate the cell's previous populations;
// rec: the cell's record (Q wall fractions, Q × 3 wall velocities).  The rule is linear with
// coefficients summing to one, so shifted storage applies it to the deviations unchanged.
template <int Q, class S, class R, class E>
KMI_DEV void libb_prephase(R (&fp)[Q], const E* __restrict__ fi, long long self, long long P,
                           const double* __restrict__ rec) {
  sfor<Q - 1>([&](auto nc) {
    constexpr int q = LibbOrder<Q>::q[decltype(nc)::value];
    constexpr int o = Lat<Q>::opp[q];
    constexpr int cx = Lat<Q>::cx[q], cy = Lat<Q>::cy[q], cz = Lat<Q>::cz[q];
    constexpr bool axis = cx * cx + cy * cy + cz * cz == 1;
    // 6 w_q: 2/3 (D2Q9 axis), 1/3 (D3Q19 axis), 1/6 (diagonals), written as the reference
    constexpr double six_wq = axis ? (Q == 9 ? 2.0 / 3.0 : 1.0 / 3.0) : 1.0 / 6.0;
    const R qw = R(rec[q]);
    if (qw > R(0)) {
      const double* u = rec + Q + 3 * q;
      const R d = R(-six_wq) * cdotu<cx, cy, cz>(R(u[0]), R(u[1]), R(u[2]));
      fp[o] = libb_branch(qw, S::ld(fi, self + q * P), fp[q], S::ld(fi, self + o * P), d);
    }
  });
}

// Anti-bounce-back of one slot (thermal_2d.jl:175-178); `tl` is T_link (minus T0 when shifted).
template <class R>
KMI_DEV R abb(R gq, R w2, R tface, bool on_corner, R tcorner, bool skip) {
  const R tl = on_corner ? tcorner : tface;
  return (on_corner && skip) ? gq : -gq + w2 * tl;
}

// Dirichlet faces W, E, S, N (thermal_2d.jl:182-290 slot tables, 371-417 corner rule).
template <bool SH, class R>
KMI_DEV void thermal_faces(R (&g)[9], const StepArgs& a, int i, int j) {
  const R wa = R(2) * R(1.0 / 9.0);
  const R we = R(2) * R(1.0 / 36.0);
  const double t0 = SH ? a.T_ref : 0.0;
  auto tv = [&](double t) { return SH ? R(t - t0) : R(t); };
  if (i == 0 && a.face[KMI_WEST].dirichlet) {
    const FaceArgs& F = a.face[KMI_WEST];
    g[1] = -g[1] + wa * tv(F.T);
    g[5] = abb(g[5], we, tv(F.T), j == 0, tv(F.Tc[0]), F.skip[0] != 0);
    g[8] = abb(g[8], we, tv(F.T), j == a.ny - 1, tv(F.Tc[1]), F.skip[1] != 0);
  }
  if (i == a.nx - 1 && a.face[KMI_EAST].dirichlet) {
    const FaceArgs& F = a.face[KMI_EAST];
    g[3] = -g[3] + wa * tv(F.T);
    g[6] = abb(g[6], we, tv(F.T), j == 0, tv(F.Tc[0]), F.skip[0] != 0);
    g[7] = abb(g[7], we, tv(F.T), j == a.ny - 1, tv(F.Tc[1]), F.skip[1] != 0);
  }
  if (j == 0 && a.face[KMI_SOUTH].dirichlet) {
    const FaceArgs& F = a.face[KMI_SOUTH];
    g[2] = -g[2] + wa * tv(F.T);
    g[5] = abb(g[5], we, tv(F.T), i == 0, tv(F.Tc[0]), F.skip[0] != 0);
    g[6] = abb(g[6], we, tv(F.T), i == a.nx - 1, tv(F.Tc[1]), F.skip[1] != 0);
  }
  if (j == a.ny - 1 && a.face[KMI_NORTH].dirichlet) {
    const FaceArgs& F = a.face[KMI_NORTH];
    g[4] = -g[4] + wa * tv(F.T);
    g[8] = abb(g[8], we, tv(F.T), i == 0, tv(F.Tc[0]), F.skip[0] != 0);
    g[7] = abb(g[7], we, tv(F.T), i == a.nx - 1, tv(F.Tc[1]), F.skip[1] != 0);
  }
}

// Dirichlet faces of D3Q19 g (apply_fixed_temp_*_3d_kernel!, thermal_3d.jl:193-480): the
// axis slot pointing into the domain and four edge slots per face, 0-based; Tc/skip carry
// the edge rule of _apply_fixed_temp_faces_3d! (thermal_3d.jl:494-535).
template <bool SH, class R>
KMI_DEV void thermal_faces3(R (&g)[19], const StepArgs& a, int i, int j, int k) {
  const R wa = R(2) * R(1.0 / 18.0);
  const R we = R(2) * R(1.0 / 36.0);
  const double t0 = SH ? a.T_ref : 0.0;
  auto tv = [&](double t) { return SH ? R(t - t0) : R(t); };
  const int xm = a.nx - 1, ym = a.ny - 1, zm = a.nz - 1;
  auto face = [&](int fid, int ax, int e0, int e1, int e2, int e3, bool c0, bool c1, bool c2,
                  bool c3) {
    const FaceArgs& F = a.face[fid];
    const R t = tv(F.T);
    g[ax] = -g[ax] + wa * t;
    g[e0] = abb(g[e0], we, t, c0, tv(F.Tc[0]), F.skip[0] != 0);
    g[e1] = abb(g[e1], we, t, c1, tv(F.Tc[1]), F.skip[1] != 0);
    g[e2] = abb(g[e2], we, t, c2, tv(F.Tc[2]), F.skip[2] != 0);
    g[e3] = abb(g[e3], we, t, c3, tv(F.Tc[3]), F.skip[3] != 0);
  };
  if (i == 0 && a.face[KMI_WEST].dirichlet)
    face(KMI_WEST, 1, 7, 9, 11, 13, j == 0, j == ym, k == 0, k == zm);
  if (i == xm && a.face[KMI_EAST].dirichlet)
    face(KMI_EAST, 2, 8, 10, 12, 14, j == 0, j == ym, k == 0, k == zm);
  if (j == 0 && a.face[KMI_SOUTH].dirichlet)
    face(KMI_SOUTH, 3, 7, 8, 15, 17, i == 0, i == xm, k == 0, k == zm);
  if (j == ym && a.face[KMI_NORTH].dirichlet)
    face(KMI_NORTH, 4, 9, 10, 16, 18, i == 0, i == xm, k == 0, k == zm);
  if (k == 0 && a.face[KMI_BOTTOM].dirichlet)
    face(KMI_BOTTOM, 5, 11, 12, 15, 16, i == 0, i == xm, j == 0, j == ym);
  if (k == zm && a.face[KMI_TOP].dirichlet)
    face(KMI_TOP, 6, 13, 14, 17, 18, i == 0, i == xm, j == 0, j == ym);
}

}  // namespace kmi
