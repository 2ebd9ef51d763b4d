// kmi_nn.hip — the learned warm start: a small conv decoder trained and run on the GPU.
//
// Network (no reference counterpart; the injection point is the initial populations, as
// f_init in src/drivers/gnf_periodic_steady.jl:114-150):
//   x (in_dim) → dense → SiLU → dense → SiLU → latent hw×hw×c0 (NHWC)
//   → [nearest ×2, 3×3 conv, SiLU] → [×2, conv, SiLU] → [×2, conv] → (8hw)²×out_c.
// Every matrix product is one kernel (k_mm): BF16 operands on gfx950's
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation, staged through LDS in K-steps of 32 by
// loaders that compute the im2col / transposed-conv addresses incrementally (no im2col
// buffers), with the bias / SiLU / activation-gradient epilogues fused where K is not split.
// Weights are kept in FP32 (Adam master copy) and mirrored to BF16 for the products;
// activations and gradients enter the products as BF16, everything else is FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "kmi_internal.h"

namespace kmi {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 → BF16, round to nearest even: on the device gfx950's v_cvt_pk_bf16_f32, on the host
// the same rounding in integer arithmetic (inf and NaN kept).
KMI_HD uint16_t f2bf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
#else
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
#endif
}
// The logistic function 1 / (1 + e^−z) with the hardware exponential and reciprocal (a few
// ulp; the network's products round to BF16 anyway): SiLU(z) = z·σ(z), SiLU'(z) =
// σ(z)(1 + z(1 − σ(z))).
KMI_DEV float sigmoid_fast(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
KMI_HD float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, uint32_t(h) << 16);
}

// ---- operand loaders ------------------------------------------------------------------------
// A product C[M×N] = Σ_k A(m, k) B(k, n) reads each operand through a loader in units of eight
// BF16 elements (16 bytes).  A K-major loader returns the eight consecutive k of one row
// (m for A, n for B); a row-major one returns eight consecutive rows at one k.  Each thread
// keeps a cursor per unit: the row decode happens once, and every K-step of 32 advances the
// cursor incrementally (no division per element).  Elements outside the matrix read as 0.
KMI_DEV bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
KMI_DEV uint4 ld16(const uint16_t* p) { return *reinterpret_cast<const uint4*>(p); }
KMI_DEV uint4 pack8(const uint16_t (&v)[8]) {
  return make_uint4(uint32_t(v[0]) | (uint32_t(v[1]) << 16), uint32_t(v[2]) | (uint32_t(v[3]) << 16),
                    uint32_t(v[4]) | (uint32_t(v[5]) << 16), uint32_t(v[6]) | (uint32_t(v[7]) << 16));
}
constexpr uint4 kZero4 = {0u, 0u, 0u, 0u};

// (row, k) = p[row·s0 + k], k contiguous.
struct DenseK {
  static constexpr bool kmajor = true;
  const uint16_t* p;
  long long s0;
  int nrow, nk;
  struct Cur {
    const uint16_t* q;
    int k;
    bool ok;
  };
  KMI_DEV void init(Cur& c, int row, int k) const {
    c.ok = row < nrow;
    c.k = k;
    c.q = p + (long long)(c.ok ? row : 0) * s0 + k;
  }
  KMI_DEV uint4 load(const Cur& c) const {
    if (!c.ok || c.k >= nk) return kZero4;
    if (c.k + 8 <= nk && aligned16(c.q)) return ld16(c.q);
    uint16_t v[8];
    for (int e = 0; e < 8; ++e) v[e] = c.k + e < nk ? c.q[e] : uint16_t(0);
    return pack8(v);
  }
  KMI_DEV void next(Cur& c) const {
    c.q += 32;
    c.k += 32;
  }
};

// (row, k) = p[row + k·s1], rows contiguous (eight consecutive rows per unit).
struct DenseR {
  static constexpr bool kmajor = false;
  const uint16_t* p;
  long long s1;
  int nrow, nk;
  struct Cur {
    const uint16_t* q;
    int row, k;
  };
  KMI_DEV void init(Cur& c, int row, int k) const {
    c.row = row;
    c.k = k;
    c.q = p + row + (long long)k * s1;
  }
  KMI_DEV uint4 load(const Cur& c) const {
    if (c.k >= nk || c.row >= nrow) return kZero4;
    if (c.row + 8 <= nrow && aligned16(c.q)) return ld16(c.q);
    uint16_t v[8];
    for (int e = 0; e < 8; ++e) v[e] = c.row + e < nrow ? c.q[e] : uint16_t(0);
    return pack8(v);
  }
  KMI_DEV void next(Cur& c) const {
    c.q += 32 * s1;
    c.k += 32;
  }
};

// (row, k) = p[row·s0 + k·s1], any strides: element loads along k.
struct DenseS {
  static constexpr bool kmajor = true;
  const uint16_t* p;
  long long s0, s1;
  int nrow, nk;
  struct Cur {
    const uint16_t* q;
    int k;
    bool ok;
  };
  KMI_DEV void init(Cur& c, int row, int k) const {
    c.ok = row < nrow;
    c.k = k;
    c.q = p + (long long)(c.ok ? row : 0) * s0 + (long long)k * s1;
  }
  KMI_DEV uint4 load(const Cur& c) const {
    uint16_t v[8];
    for (int e = 0; e < 8; ++e) v[e] = (c.ok && c.k + e < nk) ? c.q[(long long)e * s1] : uint16_t(0);
    return pack8(v);
  }
  KMI_DEV void next(Cur& c) const {
    c.q += 32 * s1;
    c.k += 32;
  }
};

// The row-major matrix of LA with row M0 replaced by ones (BF16 1.0) for k < nk (LA reads
// zeros there): a weight-gradient product Σ_k A(m, k) B(k, n) then also yields
// Σ_k B(k, n) — the bias gradient — in its row M0.
template <class LA>
struct WithOnes {
  static_assert(!LA::kmajor, "rows of eight");
  static constexpr bool kmajor = false;
  LA la;
  int M0, nk;
  struct Cur {
    typename LA::Cur c;
    int row, k;
  };
  KMI_DEV void init(Cur& u, int row, int k) const {
    u.row = row;
    u.k = k;
    la.init(u.c, row, k);
  }
  KMI_DEV uint4 load(const Cur& u) const {
    uint4 v = la.load(u.c);
    const int e = M0 - u.row;
    if (e >= 0 && e < 8 && u.k < nk) {
      uint32_t* w = reinterpret_cast<uint32_t*>(&v);
      w[e >> 1] |= 0x3F80u << (16 * (e & 1));
    }
    return v;
  }
  KMI_DEV void next(Cur& u) const {
    la.next(u.c);
    u.k += 32;
  }
};

// Convolution after nearest ×2 up-sampling (3×3, zero padding 1) on an NHWC input
// [B, Hi, Hi, C] (C a multiple of 8): row m = output pixel (b, y, x) of the Ho = 2 Hi grid,
// k = (ky, kx, c) = tap·C + c, element in[b, (y+ky−1)/2, (x+kx−1)/2, c].
struct ConvFwdA {
  static constexpr bool kmajor = true;
  const uint16_t* in;
  int B, Hi, C;
  struct Cur {
    const uint16_t* base;  // image b
    int y, x, ky, kx, c;
    bool ok;
  };
  KMI_DEV void init(Cur& u, int m, int k) const {
    const int Ho = 2 * Hi, HW = Ho * Ho;
    u.ok = m < B * HW;
    const int mm = u.ok ? m : 0, b = mm / HW, r = mm - b * HW;
    u.y = r / Ho;
    u.x = r - u.y * Ho;
    u.base = in + (long long)b * Hi * Hi * C;
    const int t = k / C;
    u.c = k - t * C;
    u.ky = t / 3;
    u.kx = t - 3 * u.ky;
  }
  KMI_DEV uint4 load(const Cur& u) const {
    const int Ho = 2 * Hi, yy = u.y + u.ky - 1, xx = u.x + u.kx - 1;
    if (!u.ok || u.ky > 2 || yy < 0 || yy >= Ho || xx < 0 || xx >= Ho) return kZero4;
    return ld16(u.base + ((long long)(yy >> 1) * Hi + (xx >> 1)) * C + u.c);
  }
  KMI_DEV void next(Cur& u) const {
    u.c += 32;
    while (u.c >= C) {
      u.c -= C;
      if (++u.kx == 3) {
        u.kx = 0;
        ++u.ky;
      }
    }
  }
};

// The same matrix read transposed for the weight gradient: row r = (tap, c) (eight
// consecutive channels of one tap per unit), k = output pixel (b, y, x).
struct ConvWgradA {
  static constexpr bool kmajor = false;
  const uint16_t* in;
  int B, Hi, C;
  struct Cur {
    int b, y, x, p, ky, kx, c;
    bool ok;
  };
  KMI_DEV void init(Cur& u, int r, int k) const {
    const int Ho = 2 * Hi, HW = Ho * Ho;
    u.ok = r < 9 * C;
    const int t = u.ok ? r / C : 0;
    u.c = u.ok ? r - t * C : 0;
    u.ky = t / 3;
    u.kx = t - 3 * u.ky;
    u.p = k;
    u.b = k / HW;
    const int rr = k - u.b * HW;
    u.y = rr / Ho;
    u.x = rr - u.y * Ho;
  }
  KMI_DEV uint4 load(const Cur& u) const {
    const int Ho = 2 * Hi, yy = u.y + u.ky - 1, xx = u.x + u.kx - 1;
    if (!u.ok || u.p >= B * Ho * Ho || yy < 0 || yy >= Ho || xx < 0 || xx >= Ho) return kZero4;
    return ld16(in + (((long long)u.b * Hi + (yy >> 1)) * Hi + (xx >> 1)) * C + u.c);
  }
  KMI_DEV void next(Cur& u) const {
    const int Ho = 2 * Hi;
    u.p += 32;
    u.x += 32;
    while (u.x >= Ho) {
      u.x -= Ho;
      if (++u.y == Ho) {
        u.y = 0;
        ++u.b;
      }
    }
  }
};

// Data gradient through the convolution and the up-sampling, on the input grid: row
// m = input pixel (b, yi, xi), k = (d, tap, o) = (4·9 index)·Cp + o with d = (dy, dx) the
// output pixel's position in its 2×2 block; element dZ[b, 2yi+dy−ky+1, 2xi+dx−kx+1, o] of the
// output gradient stored [B, Ho, Ho, Cp] (Cp a multiple of 8, channels ≥ Co zero).
struct ConvDgradA {
  static constexpr bool kmajor = true;
  const uint16_t* dz;
  int B, Hi, Cp;
  struct Cur {
    const uint16_t* base;
    int yi, xi, d, ky, kx, o;
    bool ok;
  };
  KMI_DEV void init(Cur& u, int m, int k) const {
    const int HW = Hi * Hi;
    u.ok = m < B * HW;
    const int mm = u.ok ? m : 0, b = mm / HW, r = mm - b * HW;
    u.yi = r / Hi;
    u.xi = r - u.yi * Hi;
    u.base = dz + (long long)b * 4 * HW * Cp;
    const int s = k / Cp;
    u.o = k - s * Cp;
    u.d = s / 9;
    const int t = s - 9 * u.d;
    u.ky = t / 3;
    u.kx = t - 3 * u.ky;
  }
  KMI_DEV uint4 load(const Cur& u) const {
    const int Ho = 2 * Hi;
    const int y = 2 * u.yi + (u.d >> 1) - u.ky + 1, x = 2 * u.xi + (u.d & 1) - u.kx + 1;
    if (!u.ok || u.d > 3 || y < 0 || y >= Ho || x < 0 || x >= Ho) return kZero4;
    return ld16(u.base + ((long long)y * Ho + x) * Cp + u.o);
  }
  KMI_DEV void next(Cur& u) const {
    u.o += 32;
    while (u.o >= Cp) {
      u.o -= Cp;
      if (++u.kx == 3) {
        u.kx = 0;
        if (++u.ky == 3) {
          u.ky = 0;
          ++u.d;
        }
      }
    }
  }
};

// Its second operand: row n = input channel c, k = (d, tap, o), element W[(tap·Ci + c)·Cp + o]
// of the weights stored [9 Ci, Cp] (the same for the four d).
struct ConvDgradB {
  static constexpr bool kmajor = true;
  const uint16_t* w;
  int Ci, Cp;
  struct Cur {
    int c, d, t, o;
    bool ok;
  };
  KMI_DEV void init(Cur& u, int n, int k) const {
    u.ok = n < Ci;
    u.c = u.ok ? n : 0;
    const int s = k / Cp;
    u.o = k - s * Cp;
    u.d = s / 9;
    u.t = s - 9 * u.d;
  }
  KMI_DEV uint4 load(const Cur& u) const {
    if (!u.ok || u.d > 3) return kZero4;
    return ld16(w + ((long long)u.t * Ci + u.c) * Cp + u.o);
  }
  KMI_DEV void next(Cur& u) const {
    u.o += 32;
    while (u.o >= Cp) {
      u.o -= Cp;
      if (++u.t == 9) {
        u.t = 0;
        ++u.d;
      }
    }
  }
};

// ---- epilogues: element (m, n) of the block's result ------------------------------------------
// In two parts: pre(N, m, n) loads what the element needs from memory (a bias, a
// pre-activation), and operator() combines it with the product value and stores.  The kernels
// call pre() for all of a thread's elements before any store: the stores may alias the inputs
// as far as the compiler can tell, so interleaving the two would serialise one global-memory
// round trip per element.
struct EpiStore {  // P[(z·M + m)·N + n] (split-K partials, or the result itself with one split)
  float* P;
  KMI_DEV float pre(int, int, int) const { return 0.f; }
  KMI_DEV void operator()(int M, int N, int z, int m, int n, float v, float) const {
    P[((long long)z * M + m) * N + n] = v;
  }
};
struct EpiBiasAct {  // z = v + bias; Z = z; h = SiLU(z) or z; H (FP32) and Hb (BF16) = h
  const float* bias;
  float* Z;
  float* H;
  uint16_t* Hb;
  int silu;
  KMI_DEV float pre(int, int, int n) const { return bias[n]; }
  KMI_DEV void operator()(int, int N, int, int m, int n, float v, float b) const {
    const long long i = (long long)m * N + n;
    const float z = v + b;
    if (Z) Z[i] = z;
    const float h = silu ? z * sigmoid_fast(z) : z;
    if (H) H[i] = h;
    if (Hb) Hb[i] = f2bf(h);
  }
};
struct EpiActGrad {  // g = v ⊙ SiLU'(zin) (or v); FP32 (when dz is set) and BF16 copies
  const float* zin;
  float* dz;
  uint16_t* dzb;
  int silu;
  KMI_DEV float pre(int N, int m, int n) const { return silu ? zin[(long long)m * N + n] : 0.f; }
  KMI_DEV void operator()(int, int N, int, int m, int n, float v, float z) const {
    const long long i = (long long)m * N + n;
    float g = v;
    if (silu) {
      const float s = sigmoid_fast(z);
      g = g * (s * (1.f + z * (1.f - s)));
    }
    if (dz) dz[i] = g;
    dzb[i] = f2bf(g);
  }
};

// ---- the product kernel ---------------------------------------------------------------------
// 256 threads = 4 waves arranged WM × (4/WM) over a BM × BN block tile; each wave owns a
// (BM/WM) × (BN·WM/4) tile of 16×16 MFMA blocks (v_mfma_f32_16x16x32_bf16: lane l holds A row
// l%16 and B column l%16 with k = 8(l/16)..+7, and D rows 4(l/16)..+3 of column l%16).  Each
// K-step of 32 stages the A and B slabs into LDS as [row][32 k] (80-byte rows: the 16-byte
// fragment reads of a quarter-wave hit 64 distinct banks); K-major units are copied as they
// come, row-major units are scattered into their eight rows.  Two LDS buffers: the loads of
// step s+1 are in flight while the MFMAs of step s run, with one barrier per step.  Split-K:
// blockIdx.z covers K range [z·kchunk, (z+1)·kchunk).
constexpr int kLdsRow = 40;

template <int BM, int BN, int WM, class LA, class LB, class EPI>
__global__ void __launch_bounds__(256) k_mm(const int M, const int N, const int K, const int kchunk, const LA la,
                                            const LB lb, const EPI epi) {
  constexpr int WN = 4 / WM, TM = BM / WM, TN = BN / WN, RM = TM / 16, RN = TN / 16;
  static_assert(WM * WN == 4 && TM % 16 == 0 && TN % 16 == 0, "tile shape");
  constexpr int UA = BM * 4, UB = BN * 4, PA = (UA + 255) / 256, PB = (UB + 255) / 256;
  __shared__ __align__(16) uint16_t As[2][BM][kLdsRow];
  __shared__ __align__(16) uint16_t Bs[2][BN][kLdsRow];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w / WN, wn = w % WN;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
  // unit u: K-major → row u/4, k 8(u%4); row-major → rows 8(u/32)..+7, k u%32
  auto urow = [](bool km, int u) { return km ? (u >> 2) : ((u >> 5) << 3); };
  auto uk = [](bool km, int u) { return km ? ((u & 3) << 3) : (u & 31); };
  typename LA::Cur ca[PA];
  typename LB::Cur cb[PB];
#pragma unroll
  for (int p = 0; p < PA; ++p) {
    const int u = t + 256 * p;
    if (u < UA) la.init(ca[p], m0 + urow(LA::kmajor, u), kbeg + uk(LA::kmajor, u));
  }
#pragma unroll
  for (int p = 0; p < PB; ++p) {
    const int u = t + 256 * p;
    if (u < UB) lb.init(cb[p], n0 + urow(LB::kmajor, u), kbeg + uk(LB::kmajor, u));
  }
  uint4 ra[PA], rb[PB];
  auto fetch = [&]() {
#pragma unroll
    for (int p = 0; p < PA; ++p)
      if (t + 256 * p < UA) ra[p] = la.load(ca[p]);
#pragma unroll
    for (int p = 0; p < PB; ++p)
      if (t + 256 * p < UB) rb[p] = lb.load(cb[p]);
  };
  auto advance = [&]() {
#pragma unroll
    for (int p = 0; p < PA; ++p)
      if (t + 256 * p < UA) la.next(ca[p]);
#pragma unroll
    for (int p = 0; p < PB; ++p)
      if (t + 256 * p < UB) lb.next(cb[p]);
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int p = 0; p < PA; ++p) {
      const int u = t + 256 * p;
      if (u >= UA) continue;
      const int r = urow(LA::kmajor, u), k = uk(LA::kmajor, u);
      if constexpr (LA::kmajor) {
        *reinterpret_cast<uint4*>(&As[buf][r][k]) = ra[p];
      } else {
        const uint32_t v[4] = {ra[p].x, ra[p].y, ra[p].z, ra[p].w};
#pragma unroll
        for (int e = 0; e < 8; ++e) As[buf][r + e][k] = uint16_t(v[e >> 1] >> (16 * (e & 1)));
      }
    }
#pragma unroll
    for (int p = 0; p < PB; ++p) {
      const int u = t + 256 * p;
      if (u >= UB) continue;
      const int r = urow(LB::kmajor, u), k = uk(LB::kmajor, u);
      if constexpr (LB::kmajor) {
        *reinterpret_cast<uint4*>(&Bs[buf][r][k]) = rb[p];
      } else {
        const uint32_t v[4] = {rb[p].x, rb[p].y, rb[p].z, rb[p].w};
#pragma unroll
        for (int e = 0; e < 8; ++e) Bs[buf][r + e][k] = uint16_t(v[e >> 1] >> (16 * (e & 1)));
      }
    }
  };
  v4f acc[RM][RN];
#pragma unroll
  for (int a = 0; a < RM; ++a)
#pragma unroll
    for (int b = 0; b < RN; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  if (kbeg < kend) {
    fetch();
    stage(0);
  }
  __syncthreads();
  int buf = 0;
  const int fr = lane & 15, fk = (lane >> 4) << 3;
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    const bool more = k0 + 32 < kend;
    if (more) {
      advance();
      fetch();
    }
    v8bf af[RM], bfr[RN];
#pragma unroll
    for (int a = 0; a < RM; ++a)
      af[a] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(&As[buf][wm * TM + a * 16 + fr][fk]));
#pragma unroll
    for (int b = 0; b < RN; ++b)
      bfr[b] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(&Bs[buf][wn * TN + b * 16 + fr][fk]));
#pragma unroll
    for (int a = 0; a < RM; ++a)
#pragma unroll
      for (int b = 0; b < RN; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    if (more) stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  float pre[RM][RN][4];
#pragma unroll
  for (int a = 0; a < RM; ++a)
#pragma unroll
    for (int b = 0; b < RN; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = m0 + wm * TM + a * 16 + ((lane >> 4) << 2) + i, n = n0 + wn * TN + b * 16 + fr;
        pre[a][b][i] = m < M && n < N ? epi.pre(N, m, n) : 0.f;
      }
#pragma unroll
  for (int a = 0; a < RM; ++a)
#pragma unroll
    for (int b = 0; b < RN; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = m0 + wm * TM + a * 16 + ((lane >> 4) << 2) + i, n = n0 + wn * TN + b * 16 + fr;
        if (m < M && n < N) epi(M, N, blockIdx.z, m, n, acc[a][b][i], pre[a][b][i]);
      }
}

// K split into `splits` chunks of whole K-steps; the number of chunks that results.
KMI_HD int split_chunk(int K, int splits, int* nsplit) {
  int kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 31) / 32 * 32;
  *nsplit = (K + kchunk - 1) / kchunk;
  return kchunk;
}

template <int BM, int BN, int WM, class LA, class LB, class EPI>
static hipError_t mm(int M, int N, int K, int splits, const LA& la, const LB& lb, const EPI& epi, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  int sp = 1;
  const int kchunk = split_chunk(K, splits, &sp);
  const dim3 g(unsigned((M + BM - 1) / BM), unsigned((N + BN - 1) / BN), unsigned(sp));
  hipLaunchKernelGGL((k_mm<BM, BN, WM, LA, LB, EPI>), g, dim3(256), 0, s, M, N, K, kchunk, la, lb, epi);
  return hipGetLastError();
}

// ---- direct convolutions ----------------------------------------------------------------------
// The three products of a decoder stage (forward, data gradient, weight gradient) as direct
// convolutions: a work-group stages the input patch its tile reads, zero halo included, and
// the weights in LDS once, so each input element crosses L2 about once per tile rather than
// once per tap (the im2col loaders of k_mm), and the K loop runs from LDS with no global loads
// and no barriers.  They cover 32 or 64 input channels and up to 64 output channels (the
// weight gradient also needs a 16-, 32- or 64-pixel input grid); other shapes take k_mm.

// Row stride (BF16 elements) for n elements per row such that 16 lanes reading 16-byte (or
// 4-byte) fragments from 16 consecutive rows hit 16 distinct groups of four LDS banks: a
// multiple of 8 elements whose count of 8-element units is odd.
constexpr int lds_stride(int n) {
  const int s = (n + 7) / 8 * 8;
  return (s / 8) % 2 == 1 ? s : s + 8;
}

// Stages n units of 16 bytes into LDS: unit u is loaded by src(u) and stored by dst(u, v).
// Each thread issues the loads of up to BATCH of its units before their stores, so a batch's
// global-memory latency is paid once rather than once per unit.
template <int NT, int BATCH, class Src, class Dst>
KMI_DEV void stage_units(int n, const Src& src, const Dst& dst) {
  for (int u0 = int(threadIdx.x); u0 < n; u0 += BATCH * NT) {
    uint4 v[BATCH];
#pragma unroll
    for (int j = 0; j < BATCH; ++j) {
      const int u = u0 + j * NT;
      v[j] = u < n ? src(u) : kZero4;
    }
#pragma unroll
    for (int j = 0; j < BATCH; ++j) {
      const int u = u0 + j * NT;
      if (u < n) dst(u, v[j]);
    }
  }
}

// Forward: one work-group per TY × TX tile of output pixels of one image (TX a multiple of 16,
// TY even), all COP ≥ Co output channels (a multiple of 16).  Up-sampled pixel (Y, X) of the
// tile reads source pixel ((Y + ky − 1) >> 1, (X + kx − 1) >> 1), all inside the staged patch of
// (TY/2 + 2) × (TX/2 + 2) source pixels (zero outside the image: the up-grid's zero padding).
// The 4 waves own TY·TX/64 blocks of 16 pixels along x each.  A fragment (16 pixels × 8
// channels of one tap) is one 16-byte LDS read per lane at its pixel's source position; a B
// fragment (8 consecutive k of one output channel) one read of the transposed weights
// [COP][9 CI] (k = tap·CI + c).
template <int CI, int COP, int TY, int TX, class EPI>
__global__ void __launch_bounds__(256) k_conv_fwd(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wt,
                                                  const int Hi, const int Co, const int groups, const EPI epi) {
  constexpr int PY = TY / 2 + 2, PX = TX / 2 + 2, PS = lds_stride(CI), KW = lds_stride(9 * CI);
  constexpr int XB = TX / 16, MB = TY * XB, MBW = MB / 4, NB = COP / 16, UC = CI / 8, UW = 9 * CI / 8;
  static_assert(CI % 32 == 0 && COP % 16 == 0 && TX % 16 == 0 && TY % 2 == 0 && MB % 4 == 0, "conv tile");
  __shared__ __align__(16) uint16_t patch[PY * PX * PS];
  __shared__ __align__(16) uint16_t wl[COP * KW];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // blockIdx.z = image · groups + channel group (COP output channels each)
  const int Ho = 2 * Hi, x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, b = blockIdx.z / groups;
  const int o0 = (blockIdx.z - b * groups) * COP;
  wt += (long long)o0 * 9 * CI;
  const uint16_t* img = in + (long long)b * Hi * Hi * CI;
  // the patch's units, then the weights'
  constexpr int NP = PY * PX * UC;
  stage_units<256, 12>(
      NP + COP * UW,
      [&](int u) {
        if (u >= NP) return ld16(wt + (long long)(u - NP) * 8);  // [COP][9 CI] is contiguous
        const int cu = u % UC, p = u / UC, px = p % PX, py = p / PX;
        const int sy = y0 / 2 - 1 + py, sx = x0 / 2 - 1 + px;
        return (sy >= 0 && sy < Hi && sx >= 0 && sx < Hi) ? ld16(img + ((long long)sy * Hi + sx) * CI + cu * 8) : kZero4;
      },
      [&](int u, uint4 v) {
        if (u >= NP) {
          const int o = (u - NP) / UW, ku = (u - NP) - o * UW;
          *reinterpret_cast<uint4*>(&wl[o * KW + ku * 8]) = v;
        } else {
          const int cu = u % UC, p = u / UC;
          *reinterpret_cast<uint4*>(&patch[p * PS + cu * 8]) = v;
        }
      });
  __syncthreads();
  const int fr = lane & 15, fk = (lane >> 4) << 3;
  v4f acc[MBW][NB];
#pragma unroll
  for (int j = 0; j < MBW; ++j)
#pragma unroll
    for (int n = 0; n < NB; ++n) acc[j][n] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const int ky = tap / 3, kx = tap % 3;
    int pa[MBW];
#pragma unroll
    for (int j = 0; j < MBW; ++j) {
      const int blk = w * MBW + j, r = blk / XB, cx = (blk % XB) * 16;
      pa[j] = ((((r + ky - 1) >> 1) + 1) * PX + ((cx + fr + kx - 1) >> 1) + 1) * PS + fk;
    }
#pragma unroll
    for (int cs = 0; cs < CI; cs += 32) {
      v8bf bfr[NB];
#pragma unroll
      for (int n = 0; n < NB; ++n)
        bfr[n] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(&wl[(n * 16 + fr) * KW + tap * CI + cs + fk]));
#pragma unroll
      for (int j = 0; j < MBW; ++j) {
        const v8bf af = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(&patch[pa[j] + cs]));
#pragma unroll
        for (int n = 0; n < NB; ++n) acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr[n], acc[j][n], 0, 0, 0);
      }
    }
  }
  float pre[MBW][NB][4];
#pragma unroll
  for (int j = 0; j < MBW; ++j) {
    const int blk = w * MBW + j, y = y0 + blk / XB, xb = x0 + (blk % XB) * 16 + ((lane >> 4) << 2);
#pragma unroll
    for (int n = 0; n < NB; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = xb + i, o = o0 + n * 16 + fr;
        pre[j][n][i] = y < Ho && x < Ho && o < Co ? epi.pre(Co, (b * Ho + y) * Ho + x, o) : 0.f;
      }
  }
#pragma unroll
  for (int j = 0; j < MBW; ++j) {
    const int blk = w * MBW + j, y = y0 + blk / XB, xb = x0 + (blk % XB) * 16 + ((lane >> 4) << 2);
#pragma unroll
    for (int n = 0; n < NB; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = xb + i, o = o0 + n * 16 + fr;
        if (y < Ho && x < Ho && o < Co) epi(0, Co, 0, (b * Ho + y) * Ho + x, o, acc[j][n][i], pre[j][n][i]);
      }
  }
}

// Data gradient on the input grid: one work-group per TY × TX tile of input pixels of one image
// (TX a multiple of 16), all CI input channels.  dIn[s, c] = Σ_{d, tap, o} dZ[2s + d − k + 1, o]
// · W[tap, c, o], d = (dy, dx) the output pixel's place in its 2 × 2 block and k = (ky, kx) the
// tap; K = 36·CP in the order (d, tap, o).  The dZ patch (up-grid rows 2y0 − 1 … 2y0 + 2TY,
// columns 2x0 − 2 … 2x0 + 2TX + 1, zero outside the image) is stored split by column parity, so
// the 16 pixels of a fragment read 16 consecutive LDS pixels; the weights as [tap·CI + c][CP].
template <int CI, int CG, int CP, int TY, int TX, class EPI>
__global__ void __launch_bounds__(256) k_conv_dgrad(const uint16_t* __restrict__ dz, const uint16_t* __restrict__ wsrc,
                                                    const int Hi, const EPI epi) {
  // CG input channels per work-group (CI / CG groups: blockIdx.z = image · groups + group)
  constexpr int QY = 2 * TY + 2, HX = TX + 2, PS = lds_stride(CP), UC = CP / 8, GROUPS = CI / CG;
  constexpr int XB = TX / 16, MB = TY * XB, MBW = MB / 4, NB = CG / 16, KSTEPS = 36 * CP / 32;
  static_assert(CG % 16 == 0 && CI % CG == 0 && CP % 8 == 0 && TX % 16 == 0 && MB % 4 == 0, "conv tile");
  __shared__ __align__(16) uint16_t dzs[QY * 2 * HX * PS];
  __shared__ __align__(16) uint16_t wl[9 * CG * PS];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int Ho = 2 * Hi, x0 = blockIdx.x * TX, y0 = blockIdx.y * TY, b = blockIdx.z / GROUPS;
  const int c0 = (blockIdx.z - b * GROUPS) * CG;
  const uint16_t* img = dz + (long long)b * Ho * Ho * CP;
  // slot (q, par, h) holds up-grid pixel (2y0 − 1 + q, 2x0 − 2 + 2h + par); then the weights
  constexpr int NP = QY * HX * 2 * UC;
  stage_units<256, 12>(
      NP + 9 * CG * UC,
      [&](int u) {
        if (u >= NP) {  // rows tap·CI + c0 … + CG − 1 of the weights [9 CI][CP]
          const int cu = (u - NP) % UC, row = (u - NP) / UC, tap = row / CG;
          return ld16(wsrc + ((long long)tap * CI + c0 + (row - tap * CG)) * CP + cu * 8);
        }
        const int cu = u % UC, p = u / UC, par = p & 1, h = (p >> 1) % HX, q = (p >> 1) / HX;
        const int Y = 2 * y0 - 1 + q, X = 2 * x0 - 2 + 2 * h + par;
        return (Y >= 0 && Y < Ho && X >= 0 && X < Ho) ? ld16(img + ((long long)Y * Ho + X) * CP + cu * 8) : kZero4;
      },
      [&](int u, uint4 v) {
        if (u >= NP) {
          const int cu = (u - NP) % UC, row = (u - NP) / UC;
          *reinterpret_cast<uint4*>(&wl[row * PS + cu * 8]) = v;
        } else {
          const int cu = u % UC, p = u / UC, par = p & 1, h = (p >> 1) % HX, q = (p >> 1) / HX;
          *reinterpret_cast<uint4*>(&dzs[((q * 2 + par) * HX + h) * PS + cu * 8]) = v;
        }
      });
  __syncthreads();
  const int fr = lane & 15, fk = (lane >> 4) << 3;
  v4f acc[MBW][NB];
#pragma unroll
  for (int j = 0; j < MBW; ++j)
#pragma unroll
    for (int n = 0; n < NB; ++n) acc[j][n] = v4f{0.f, 0.f, 0.f, 0.f};
  if constexpr (CP % 32 == 0) {
    // a K-step of 32 channels lies in one (d, tap): with both loops unrolled every offset
    // beyond the lane's base is a compile-time constant
    int abase[MBW];
#pragma unroll
    for (int j = 0; j < MBW; ++j) {
      const int blk = w * MBW + j, r = blk / XB, cx = (blk % XB) * 16;
      abase[j] = (4 * r * HX + cx + fr) * PS + fk;
    }
    const int bbase = fr * PS + fk;
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int ky = tap / 3, kx = tap % 3, e = (d & 1) - kx + 1;
        const int aoff = ((2 * ((d >> 1) - ky + 2) + (e & 1)) * HX + ((e + 2) >> 1)) * PS;
#pragma unroll
        for (int cs = 0; cs < CP; cs += 32) {
          v8bf bfr[NB];
#pragma unroll
          for (int n = 0; n < NB; ++n)
            bfr[n] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(&wl[bbase + (tap * CG + n * 16) * PS + cs]));
#pragma unroll
          for (int j = 0; j < MBW; ++j) {
            const v8bf af = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(&dzs[abase[j] + aoff + cs]));
#pragma unroll
            for (int n = 0; n < NB; ++n)
              acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr[n], acc[j][n], 0, 0, 0);
          }
        }
      }
  } else {
  for (int ks = 0; ks < KSTEPS; ++ks) {
    // this lane's 8 consecutive k: (d, tap) = s, channels o … o + 7
    const int k = ks * 32 + fk, s = k / CP, o = k - s * CP, d = s / 9, tap = s - 9 * d;
    const int ky = tap / 3, kx = tap - 3 * ky, e = (d & 1) - kx + 1;
    const int par = e & 1, hoff = (e + 2) >> 1, qoff = (d >> 1) - ky + 2;
    v8bf bfr[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n)
      bfr[n] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(&wl[(tap * CG + n * 16 + fr) * PS + o]));
#pragma unroll
    for (int j = 0; j < MBW; ++j) {
      const int blk = w * MBW + j, r = blk / XB, cx = (blk % XB) * 16;
      const v8bf af = __builtin_bit_cast(
          v8bf, *reinterpret_cast<const uint4*>(&dzs[(((2 * r + qoff) * 2 + par) * HX + cx + fr + hoff) * PS + o]));
#pragma unroll
      for (int n = 0; n < NB; ++n) acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr[n], acc[j][n], 0, 0, 0);
    }
  }
  }
  float pre[MBW][NB][4];
#pragma unroll
  for (int j = 0; j < MBW; ++j) {
    const int blk = w * MBW + j, ys = y0 + blk / XB, xb = x0 + (blk % XB) * 16 + ((lane >> 4) << 2);
#pragma unroll
    for (int n = 0; n < NB; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int xs = xb + i;
        pre[j][n][i] = ys < Hi && xs < Hi ? epi.pre(CI, (b * Hi + ys) * Hi + xs, c0 + n * 16 + fr) : 0.f;
      }
  }
#pragma unroll
  for (int j = 0; j < MBW; ++j) {
    const int blk = w * MBW + j, ys = y0 + blk / XB, xb = x0 + (blk % XB) * 16 + ((lane >> 4) << 2);
#pragma unroll
    for (int n = 0; n < NB; ++n)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int xs = xb + i;
        if (ys < Hi && xs < Hi) epi(0, CI, 0, (b * Hi + ys) * Hi + xs, c0 + n * 16 + fr, acc[j][n][i], pre[j][n][i]);
      }
  }
}

// Weight and bias gradients over a chunk of input rows, for one block of 16 output channels:
// dW[tap, c, o] = Σ in_up[Y + ky − 1, X + kx − 1, c] dZ[Y, X, o] and Σ dZ[Y, X, o] over the
// chunk's output pixels, into the split partials P[z] laid out [9 CI + 1][Co] (bias row last).
// On the input grid, output pixel (Y, X) = (2s + dy, 2x + dx) under tap (ky, kx) reads input
// pixel (s + oy, x + ox) with oy = (dy + ky − 1) >> 1 ∈ {−1, 0, 1} (likewise ox): the 36
// (tap, d) products are 16 distinct products Q[d, oy, ox] (oy − dy + 1 and ox − dx + 1 in
// {0, 1}), accumulated once each and summed per tap at the end.  MFMA rows are channels c,
// columns o, k input pixels along a row.  Each step stages TS = 128/HI input rows: the input
// transposed to [c][row][8 + x] with zero margins (rows s0 − 1 … s0 + TS), the ±1 column
// shifts assembled from the aligned fragment and the dwords either side of it; dZ transposed
// per parity d to [d][o][row·HI + x].  One wave per 16 input channels (CI/16 waves); the
// first also forms the bias row from a fragment of ones.
template <int CI, int HI>
__global__ void __launch_bounds__(64 * (CI / 16)) k_conv_wgrad(const uint16_t* __restrict__ in, const uint16_t* __restrict__ dz,
                                                            const int CP, const int Co, const int steps,
                                                            const int steps_per_chunk, float* __restrict__ P) {
  constexpr int NT = 64 * (CI / 16), TS = 128 / HI, RL = HI + 16, CS = lds_stride((TS + 2) * RL);
  constexpr int DS = lds_stride(TS * HI), UC = CI / 8, HO = 2 * HI;
  static_assert(CI % 16 == 0 && 128 % HI == 0 && HI % 8 == 0 && TS <= HI, "wgrad tile");
  __shared__ __align__(16) uint16_t xt[CI * CS];
  __shared__ __align__(16) uint16_t dt[4 * 16 * DS];
  const int t = threadIdx.x, lane = t & 63, cb = t >> 6, ob = blockIdx.y;
  const int fr = lane & 15, g = lane >> 4;
  for (int u = t; u < CI * CS / 2; u += NT) reinterpret_cast<uint32_t*>(xt)[u] = 0u;  // margins stay zero
  v4f q[16], qb = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 16; ++i) q[i] = v4f{0.f, 0.f, 0.f, 0.f};
  const v8bf ones = __builtin_bit_cast(v8bf, make_uint4(0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u));
  const int st0 = blockIdx.x * steps_per_chunk, st1 = min(steps, st0 + steps_per_chunk);
  for (int st = st0; st < st1; ++st) {
    const int row0 = st * TS, b = row0 / HI, s0 = row0 - b * HI;
    __syncthreads();  // the previous step's reads are done
    // input rows s0 − 1 … s0 + TS (channels scattered to their transposed rows), then dZ of
    // the 2TS output rows, 16 channels of block ob (zero beyond CP)
    const uint16_t* img = in + (long long)b * HI * HI * CI;
    const uint16_t* dimg = dz + (long long)b * HO * HO * CP;
    constexpr int NX = (TS + 2) * HI * UC;
    stage_units<NT, 8>(
        NX + 2 * TS * HO * 2,
        [&](int u) {
          if (u < NX) {
            const int cu = u % UC, p = u / UC, x = p % HI, s = s0 - 1 + p / HI;
            return (s >= 0 && s < HI) ? ld16(img + ((long long)s * HI + x) * CI + cu * 8) : kZero4;
          }
          const int w = u - NX, cu = w & 1, X = (w >> 1) % HO, Yr = (w >> 1) / HO, o0 = ob * 16 + cu * 8;
          return o0 < CP ? ld16(dimg + ((long long)(2 * s0 + Yr) * HO + X) * CP + o0) : kZero4;
        },
        [&](int u, uint4 v) {
          const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
          if (u < NX) {
            const int cu = u % UC, p = u / UC, x = p % HI, qr = p / HI;
#pragma unroll
            for (int e = 0; e < 8; ++e) xt[(cu * 8 + e) * CS + qr * RL + 8 + x] = uint16_t(vv[e >> 1] >> (16 * (e & 1)));
          } else {
            const int w = u - NX, cu = w & 1, X = (w >> 1) % HO, Yr = (w >> 1) / HO;
            const int d = ((Yr & 1) << 1) | (X & 1), r = Yr >> 1, x = X >> 1;
#pragma unroll
            for (int e = 0; e < 8; ++e)
              dt[(d * 16 + cu * 8 + e) * DS + r * HI + x] = uint16_t(vv[e >> 1] >> (16 * (e & 1)));
          }
        });
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TS * HI / 32; ++kk) {
      const int p = kk * 32 + g * 8, r = p / HI, x = p - r * HI;
      v8bf bd[4];
#pragma unroll
      for (int d = 0; d < 4; ++d)
        bd[d] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4*>(&dt[(d * 16 + fr) * DS + p]));
      if (cb == 0)
#pragma unroll
        for (int d = 0; d < 4; ++d) qb = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ones, bd[d], qb, 0, 0, 0);
#pragma unroll
      for (int oy = -1; oy <= 1; ++oy) {
        const uint16_t* base = &xt[(cb * 16 + fr) * CS + (r + oy + 1) * RL + 8 + x];
        const uint4 c0 = *reinterpret_cast<const uint4*>(base);
        const uint32_t lf = *reinterpret_cast<const uint32_t*>(base - 2), rt = *reinterpret_cast<const uint32_t*>(base + 8);
        v8bf f[3];
        f[0] = __builtin_bit_cast(v8bf, make_uint4(__builtin_amdgcn_alignbyte(c0.x, lf, 2), __builtin_amdgcn_alignbyte(c0.y, c0.x, 2),
                                                   __builtin_amdgcn_alignbyte(c0.z, c0.y, 2), __builtin_amdgcn_alignbyte(c0.w, c0.z, 2)));
        f[1] = __builtin_bit_cast(v8bf, c0);
        f[2] = __builtin_bit_cast(v8bf, make_uint4(__builtin_amdgcn_alignbyte(c0.y, c0.x, 2), __builtin_amdgcn_alignbyte(c0.z, c0.y, 2),
                                                   __builtin_amdgcn_alignbyte(c0.w, c0.z, 2), __builtin_amdgcn_alignbyte(rt, c0.w, 2)));
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
          const int oyi = oy - dy + 1;
          if (oyi < 0 || oyi > 1) continue;
#pragma unroll
          for (int ox = -1; ox <= 1; ++ox)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
              const int oxi = ox - dx + 1;
              if (oxi < 0 || oxi > 1) continue;
              const int qi = ((dy * 2 + oyi) * 2 + dx) * 2 + oxi;
              q[qi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(f[ox + 1], bd[dy * 2 + dx], q[qi], 0, 0, 0);
            }
        }
      }
    }
  }
  // dW[tap] = Σ_d Q[d, oy(dy, ky), ox(dx, kx)] in the order d = 0 … 3
  const int Mw = 9 * CI + 1, o = ob * 16 + fr;
  float* Pz = P + (long long)blockIdx.x * Mw * Co;
  if (o < Co) {
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int ky = tap / 3, kx = tap % 3;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = 0.f;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const int dy = d >> 1, dx = d & 1;
          const int oyi = ((dy + ky - 1) >> 1) - dy + 1, oxi = ((dx + kx - 1) >> 1) - dx + 1;
          v += q[((dy * 2 + oyi) * 2 + dx) * 2 + oxi][i];
        }
        Pz[(long long)(tap * CI + cb * 16 + g * 4 + i) * Co + o] = v;
      }
    }
    if (cb == 0 && g == 0) Pz[(long long)(9 * CI) * Co + o] = qb[0];
  }
}

// Σ over splits (fixed order) + bias; optional SiLU; writes Z (pre-activation, FP32), the
// activation H (FP32, optional) and its BF16 copy (optional).
__global__ void k_epilogue(const float* __restrict__ P, int splits, long long MN, int N,
                           const float* __restrict__ bias, int silu, float* __restrict__ Z,
                           float* __restrict__ H, uint16_t* __restrict__ Hb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MN) return;
  float z = 0.f;
  for (int s = 0; s < splits; ++s) z += P[s * MN + i];
  if (bias) z += bias[i % N];
  if (Z) Z[i] = z;
  float h = z;
  if (silu) h = z * sigmoid_fast(z);
  if (H) H[i] = h;
  if (Hb) Hb[i] = f2bf(h);
}

// Σ over split-K partials of products with a long reduction and a small output (weight
// gradients): 256/SL outputs × SL split lanes per work-group, lane g summing splits g, g + SL,
// …, the lane sums added in lane order (a fixed order: the result is deterministic).
template <int SL>
__global__ void __launch_bounds__(256) k_split_sum(const float* __restrict__ P, int splits, long long MN,
                                                    float* __restrict__ out) {
  constexpr int NO = 256 / SL;
  __shared__ float part[SL][NO + 1];
  const int o = threadIdx.x % NO, g = threadIdx.x / NO;
  const long long i = (long long)blockIdx.x * NO + o;
  float z = 0.f;
  if (i < MN)
    for (int sp = g; sp < splits; sp += SL) z += P[(long long)sp * MN + i];
  part[g][o] = z;
  __syncthreads();
  if (g == 0 && i < MN) {
    float t = 0.f;
    for (int k = 0; k < SL; ++k) t += part[k][o];
    out[i] = t;
  }
}
static hipError_t launch_split_sum(const float* P, int splits, long long mn, float* out, hipStream_t s) {
  if (splits > 256)
    hipLaunchKernelGGL(k_split_sum<64>, dim3(unsigned((mn + 3) / 4)), dim3(256), 0, s, P, splits, mn, out);
  else
    hipLaunchKernelGGL(k_split_sum<16>, dim3(unsigned((mn + 15) / 16)), dim3(256), 0, s, P, splits, mn, out);
  return hipGetLastError();
}

// Σ over `splits` partials, then dZ = Σ ⊙ SiLU'(Z) (or Σ); FP32 (when dZ is set) and BF16
// copies.  16 outputs × 16 split lanes per work-group, lane g summing splits g, g + 16, …,
// the lane sums added in lane order (fixed order).
__global__ void __launch_bounds__(256) k_split_act(const float* __restrict__ P, int splits, long long n,
                                                    const float* __restrict__ Z, int silu, float* __restrict__ dZ,
                                                    uint16_t* __restrict__ dZb) {
  __shared__ float part[16][17];
  const int o = threadIdx.x & 15, g = threadIdx.x >> 4;
  const long long i = (long long)blockIdx.x * 16 + o;
  float z = 0.f;
  if (i < n)
    for (int sp = g; sp < splits; sp += 16) z += P[(long long)sp * n + i];
  part[g][o] = z;
  __syncthreads();
  if (g != 0 || i >= n) return;
  float v = 0.f;
  for (int k = 0; k < 16; ++k) v += part[k][o];
  if (silu) {
    const float zz = Z[i], sg = sigmoid_fast(zz);
    v = v * (sg * (1.f + zz * (1.f - sg)));
  }
  if (dZ) dZ[i] = v;
  dZb[i] = f2bf(v);
}

// MSE loss partials and dY = 2 (Y − T) / count, stored in BF16 as [rows][Cp] (the padding
// columns C … Cp − 1 are zero from allocation on).
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             long long n, int C, int Cp, uint16_t* __restrict__ dYb,
                                             double* __restrict__ part) {
  __shared__ double lds[4];
  double acc = 0.0;
  const float scale = 2.f / float(n);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    acc += double(d) * double(d);
    const int ii = int(i), r = ii / C;  // element counts stay below 2^31 (kmi_nn_create)
    dYb[(long long)r * Cp + (ii - r * C)] = f2bf(scale * d);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (lds[0] + lds[1] + lds[2] + lds[3]) / double(n);
}
// One work-group: thread-strided sums in fixed order, then a fixed LDS tree.
__global__ void __launch_bounds__(256) k_sum_parts(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double lds[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int i = t; i < n; i += 256) s += part[i];
  lds[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) lds[t] += lds[t + w];
    __syncthreads();
  }
  if (t == 0) out[0] = lds[0];
}

__global__ void k_to_bf16(const float* __restrict__ x, long long n, uint16_t* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// The BF16 weight copies the products read, in one launch: job k writes n elements of dst
// from a BF16 [rows][cols] source — transposed to [outc][rows] (transpose = 1), or with its
// columns padded to outc (transpose = 0); elements beyond cols are zero.
struct WCopy {
  const uint16_t* src;
  uint16_t* dst;
  int rows, cols, outc, transpose;
};
struct WCopies {
  WCopy job[4];
  long long end[4];  // cumulative element counts
  int n;
};
__global__ void k_weight_copies(const WCopies c) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < c.n; ++k) {
    if (i >= c.end[k]) continue;
    const WCopy& J = c.job[k];
    const long long e = i - (k ? c.end[k - 1] : 0);
    if (J.transpose) {
      const int o = int(e / J.rows), r = int(e - (long long)o * J.rows);
      J.dst[e] = o < J.cols ? J.src[(long long)r * J.cols + o] : uint16_t(0);
    } else {
      const long long r = e / J.outc;
      const int col = int(e - r * J.outc);
      J.dst[e] = col < J.cols ? J.src[r * J.cols + col] : uint16_t(0);
    }
    return;
  }
}

// Adam (β1 = 0.9, β2 = 0.999, ε = 1e-8) on the FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wb, long long n, float lr,
                       float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
  w[i] = wi;
  wb[i] = f2bf(wi);
}

// Gather rows idx[0..B) of the device dataset into the batch buffers: grid (chunks, B), one
// float4 of a target row per thread (a row holds (8·latent_hw)²·out_c floats, a multiple of
// 4), and the first group of each row copies its input features.
__global__ void k_gather(const float* __restrict__ X, const float* __restrict__ Y, const int* __restrict__ idx,
                         int din, int dout4, float* __restrict__ xb, float* __restrict__ yb) {
  const int b = blockIdx.y, src = idx[b], t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < dout4)
    reinterpret_cast<float4*>(yb)[(long long)b * dout4 + t] = reinterpret_cast<const float4*>(Y)[(long long)src * dout4 + t];
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < din; i += blockDim.x) xb[(long long)b * din + i] = X[(long long)src * din + i];
}

static inline dim3 g1(long long n) { return dim3(unsigned((n + 255) / 256)); }

// ---- static workspace of the allocation-free entry point ----------------------------------
// kmi_gemm_bf16_noalloc stages its operands and split-K partials here; each GPU holds its own
// copy of the array until the process ends.  In host code `scratch` names a host-side stand-in,
// not a device address: its address on each device comes from hipGetSymbolAddress, taken once
// per device (with that device current) when a context is created, and only that address is
// ever handed to kernels or copies.
static __device__ __align__(256) unsigned char scratch[KMI_NOALLOC_WORKSPACE_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kMaxDevices] = {};

hipError_t resolve_noalloc_workspace(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_dev[device] = ptr;
  return hipSuccess;
}

void* noalloc_workspace(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_dev[device];
}

}  // namespace kmi

using namespace kmi;

// ---- the network object -------------------------------------------------------------------
struct kmi_nn {
  kmi_ctx* ctx = nullptr;
  kmi_nn_desc d{};
  int L = 0;  // hw·hw·c0
  int H1 = 0, H2 = 0, H3 = 0;  // spatial sizes after conv 1, 2, 3
  int Cp3 = 0;  // output channels padded to a multiple of 8 (BF16 gradient and weight copies)
  long long nparam = 0;
  // parameter offsets in the flat vector
  long long oW1, ob1, oW2, ob2, oC1, obc1, oC2, obc2, oC3, obc3;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* wb = nullptr;
  uint16_t* w3p = nullptr;  // BF16 C3 as [9 c2][Cp3], columns >= out_c zero (null: Cp3 == out_c)
  uint16_t* wt[3] = {nullptr, nullptr, nullptr};  // BF16 conv weights transposed to [COP][9 Ci] (direct forward)
  int64_t adam_t = 0;
  // activations (max_batch)
  float *x = nullptr, *tgt = nullptr;
  uint16_t* xb = nullptr;
  float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *z4 = nullptr, *y = nullptr;
  uint16_t *h1b = nullptr, *h2b = nullptr, *h3b = nullptr, *h4b = nullptr;
  uint16_t* dyb = nullptr; // BF16 loss gradient [B·H3², Cp3] (columns ≥ out_c zero)
  uint16_t* dzb[2] = {nullptr, nullptr};  // ping-pong gradients of the stage inputs (BF16)
  float* P = nullptr;  // split-K partials
  size_t P_elems = 0;
  double* lpart = nullptr;
  double* lout = nullptr;
  // device dataset, the epoch's permutation and the per-step losses of an epoch
  float *dsX = nullptr, *dsY = nullptr;
  int dsN = 0;
  int* idx = nullptr;
  int* perm_d = nullptr;
  double* lhist = nullptr;
};

// Output channels of the direct kernels' N dimension (0: the shape takes k_mm).
static int direct_cop(int Co) { return Co <= 16 ? 16 : Co <= 32 ? 32 : Co <= 64 ? 64 : 0; }
static bool direct_ci(int Ci) { return Ci == 32 || Ci == 64; }
static bool fwd_direct(int Ci, int Co) { return direct_ci(Ci) && direct_cop(Co) != 0; }
static bool dgrad_direct(int Ci, int Cp) { return direct_ci(Ci) && (Cp == 8 || Cp == 16 || Cp == 32 || Cp == 64); }
static bool wgrad_direct(int Ci, int Hi) { return direct_ci(Ci) && (Hi == 16 || Hi == 32 || Hi == 64); }

static long long out_elems(const kmi_nn* N, int B) { return (long long)B * N->H3 * N->H3 * N->d.out_c; }
static const uint16_t* w3_bf16(const kmi_nn* N) { return N->w3p ? N->w3p : N->wb + N->oC3; }

static int nn_alloc(kmi_nn* N) {
  const kmi_nn_desc& d = N->d;
  const int B = d.max_batch;
  N->L = d.latent_hw * d.latent_hw * d.latent_c;
  N->H1 = 2 * d.latent_hw;
  N->H2 = 4 * d.latent_hw;
  N->H3 = 8 * d.latent_hw;
  N->Cp3 = (d.out_c + 7) / 8 * 8;
  long long o = 0;
  auto take = [&](long long n) { long long r = o; o += n; return r; };
  N->oW1 = take((long long)d.in_dim * d.hidden);
  N->ob1 = take(d.hidden);
  N->oW2 = take((long long)d.hidden * N->L);
  N->ob2 = take(N->L);
  N->oC1 = take(9LL * d.latent_c * d.c1);
  N->obc1 = take(d.c1);
  N->oC2 = take(9LL * d.c1 * d.c2);
  N->obc2 = take(d.c2);
  N->oC3 = take(9LL * d.c2 * d.out_c);
  N->obc3 = take(d.out_c);
  N->nparam = o;
  const long long a1 = (long long)B * d.hidden, a2 = (long long)B * N->L;
  const long long a3 = (long long)B * N->H1 * N->H1 * d.c1, a4 = (long long)B * N->H2 * N->H2 * d.c2;
  const long long a5 = out_elems(N, B), a5p = (long long)B * N->H3 * N->H3 * N->Cp3;
  const long long big = std::max({a1, a2, a3, a4, a5});
  // P holds one split of every product it receives (pick_splits adds splits only while they
  // fit): the weight gradients with their bias row ((9 Ci + 1) × Co, (hidden + 1) × L,
  // in_dim × hidden), the data gradients (B·Hi²·Ci, at most `big`) and dH1 (B × hidden).
  const long long wmax = std::max({(long long)d.in_dim * d.hidden, (long long)(d.hidden + 1) * N->L,
                                   (9LL * d.latent_c + 1) * d.c1, (9LL * d.c1 + 1) * d.c2,
                                   (9LL * d.c2 + 1) * d.out_c, a1, big});
  N->P_elems = size_t(std::max({wmax, 16LL * 1024 * 1024}));
  hipError_t e = hipSuccess;
  auto al = [&](void** p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 16)); };
  al((void**)&N->w, o * 4);
  al((void**)&N->g, o * 4);
  al((void**)&N->m, o * 4);
  al((void**)&N->v, o * 4);
  al((void**)&N->wb, o * 2);
  if (N->Cp3 != d.out_c) al((void**)&N->w3p, 9LL * d.c2 * N->Cp3 * 2);
  {
    const int ci[3] = {d.latent_c, d.c1, d.c2}, co[3] = {d.c1, d.c2, d.out_c};
    for (int k = 0; k < 3; ++k)
      if (fwd_direct(ci[k], co[k])) al((void**)&N->wt[k], (size_t)direct_cop(co[k]) * 9 * ci[k] * 2);
  }
  al((void**)&N->x, (size_t)B * d.in_dim * 4);
  al((void**)&N->xb, (size_t)B * d.in_dim * 2);
  al((void**)&N->tgt, a5 * 4);
  al((void**)&N->z1, a1 * 4);
  al((void**)&N->h1b, a1 * 2);
  al((void**)&N->z2, a2 * 4);
  al((void**)&N->h2b, a2 * 2);
  al((void**)&N->z3, a3 * 4);
  al((void**)&N->h3b, a3 * 2);
  al((void**)&N->z4, a4 * 4);
  al((void**)&N->h4b, a4 * 2);
  al((void**)&N->y, a5 * 4);
  al((void**)&N->dyb, a5p * 2);
  for (int k = 0; k < 2; ++k) al((void**)&N->dzb[k], big * 2);
  al((void**)&N->P, N->P_elems * 4);
  al((void**)&N->lpart, 1024 * 8);
  al((void**)&N->lout, 64);
  al((void**)&N->idx, (size_t)B * 4);
  if (e != hipSuccess) return hip_fail(N->ctx, e, "kmi_nn allocation");
  if ((e = hipMemset(N->m, 0, o * 4)) != hipSuccess || (e = hipMemset(N->v, 0, o * 4)) != hipSuccess ||
      (e = hipMemset(N->dyb, 0, a5p * 2)) != hipSuccess)
    return hip_fail(N->ctx, e, "kmi_nn init");
  return KMI_OK;
}

// Splits of a product on BM × BN tiles: about `groups` work-groups (enough resident groups
// per CU to hide the operand latency of the K loop), at least `kmin` of K per split, within
// P's size.
static int pick_splits(const kmi_nn* N, int BM, int BN, long long M, long long Nn, long long K, long long groups,
                       long long kmin) {
  const long long tiles = ((M + BM - 1) / BM) * ((Nn + BN - 1) / BN);
  long long s = std::max<long long>(1, std::min<long long>((groups + tiles - 1) / tiles, K / kmin));
  s = std::min<long long>(s, (long long)N->P_elems / std::max<long long>(1, M * Nn));
  return int(std::max<long long>(1, s));
}

// Σ over the splits of P into out (fixed order).
static hipError_t split_sum(const float* P, int M, int Nn, int K, int splits, float* out, hipStream_t s) {
  int sp = 1;
  split_chunk(K, splits, &sp);
  return launch_split_sum(P, sp, (long long)M * Nn, out, s);
}

// The BF16 copies derived from the weights: C3 padded to Cp3 columns.
static hipError_t refresh_weight_copies(kmi_nn* N, hipStream_t s) {
  const kmi_nn_desc& d = N->d;
  WCopies c{};
  long long total = 0;
  auto add = [&](const uint16_t* src, uint16_t* dst, int rows, int cols, int outc, int transpose) {
    c.job[c.n] = WCopy{src, dst, rows, cols, outc, transpose};
    total += (long long)rows * outc;
    c.end[c.n++] = total;
  };
  if (N->w3p) add(N->wb + N->oC3, N->w3p, 9 * d.c2, d.out_c, N->Cp3, 0);
  const int ci[3] = {d.latent_c, d.c1, d.c2}, co[3] = {d.c1, d.c2, d.out_c};
  const long long off[3] = {N->oC1, N->oC2, N->oC3};
  for (int k = 0; k < 3; ++k)
    if (N->wt[k]) add(N->wb + off[k], N->wt[k], 9 * ci[k], co[k], direct_cop(co[k]), 1);
  if (!c.n) return hipSuccess;
  hipLaunchKernelGGL(k_weight_copies, g1(total), dim3(256), 0, s, c);
  return hipGetLastError();
}

#define NN_CHECK(expr)                                            \
  do {                                                            \
    hipError_t _e = (expr);                                       \
    if (_e != hipSuccess) return hip_fail(N->ctx, _e, #expr);     \
  } while (0)

// One convolution stage of the decoder: input [B, Hi, Hi, Ci] (BF16), output channels Co,
// weights [9 Ci, Cp] in BF16 (Cp = Co rounded up to 8, padded columns zero).
struct Stage {
  const uint16_t* in;
  int Hi, Ci, Co, Cp;
  const uint16_t* wbf;
  long long ow, ob;
};

// Products with N = Co output channels, tiles chosen by the channel count.
template <class LA, class LB, class EPI>
static hipError_t mm_by_n(int M, int Nn, int K, int splits, const LA& la, const LB& lb, const EPI& epi, hipStream_t s) {
  if (Nn > 32) return mm<128, 64, 4>(M, Nn, K, splits, la, lb, epi, s);
  if (Nn > 16) return mm<128, 32, 4>(M, Nn, K, splits, la, lb, epi, s);
  return mm<256, 16, 4>(M, Nn, K, splits, la, lb, epi, s);
}
// Weight-gradient products (M = 9 Ci rows, N = Co, long K): 64-row tiles.
template <class LA, class LB, class EPI>
static hipError_t mm_wgrad(int M, int Nn, int K, int splits, const LA& la, const LB& lb, const EPI& epi,
                           hipStream_t s) {
  if (Nn > 32) return mm<64, 64, 2>(M, Nn, K, splits, la, lb, epi, s);
  if (Nn > 16) return mm<64, 32, 2>(M, Nn, K, splits, la, lb, epi, s);
  return mm<64, 16, 4>(M, Nn, K, splits, la, lb, epi, s);
}

// Direct-convolution launches (shapes checked by fwd_direct / dgrad_direct / wgrad_direct).
// COP output channels per work-group, `groups` groups (the transposed weights hold
// groups · COP rows).
template <int CI, int COP, int TY, int TX, class EPI>
static hipError_t launch_fwd(const Stage& S, int B, const uint16_t* wt, const EPI& epi, hipStream_t s,
                             int groups = 1) {
  const int Ho = 2 * S.Hi;
  const dim3 g(unsigned((Ho + TX - 1) / TX), unsigned((Ho + TY - 1) / TY), unsigned(B * groups));
  hipLaunchKernelGGL((k_conv_fwd<CI, COP, TY, TX, EPI>), g, dim3(256), 0, s, S.in, wt, S.Hi, S.Co, groups, epi);
  return hipGetLastError();
}
template <int CI, class EPI>
static hipError_t conv_fwd_ci(const Stage& S, int B, const uint16_t* wt, const EPI& epi, hipStream_t s) {
  switch (direct_cop(S.Co)) {
    case 64: return launch_fwd<CI, 32, 8, 16>(S, B, wt, epi, s, 2);  // two groups of 32 channels
    case 32: return launch_fwd<CI, 32, 8, 16>(S, B, wt, epi, s);
    default: return launch_fwd<CI, 16, 8, 32>(S, B, wt, epi, s);
  }
}
template <class EPI>
static hipError_t conv_fwd(const Stage& S, int B, const uint16_t* wt, const EPI& epi, hipStream_t s) {
  return S.Ci == 64 ? conv_fwd_ci<64>(S, B, wt, epi, s) : conv_fwd_ci<32>(S, B, wt, epi, s);
}

// Tiles: 64 input channels in groups of 32 (16 with 64 output channels), so the weight slab
// a group stages stays small and two or three groups fit per CU.
template <int CI, int CP, class EPI>
static hipError_t launch_dgrad(const Stage& S, int B, const uint16_t* dzb, const EPI& epi, hipStream_t s) {
  constexpr int CG = CI == 64 ? (CP >= 64 ? 16 : 32) : CI, TY = CP >= 32 ? 4 : 8, TX = CI == 32 ? 32 : 16;
  const dim3 g(unsigned((S.Hi + TX - 1) / TX), unsigned((S.Hi + TY - 1) / TY), unsigned(B * (CI / CG)));
  hipLaunchKernelGGL((k_conv_dgrad<CI, CG, CP, TY, TX, EPI>), g, dim3(256), 0, s, dzb, S.wbf, S.Hi, epi);
  return hipGetLastError();
}
template <int CI, class EPI>
static hipError_t conv_dgrad_ci(const Stage& S, int B, const uint16_t* dzb, const EPI& epi, hipStream_t s) {
  switch (S.Cp) {
    case 8: return launch_dgrad<CI, 8>(S, B, dzb, epi, s);
    case 16: return launch_dgrad<CI, 16>(S, B, dzb, epi, s);
    case 32: return launch_dgrad<CI, 32>(S, B, dzb, epi, s);
    default: return launch_dgrad<CI, 64>(S, B, dzb, epi, s);
  }
}
template <class EPI>
static hipError_t conv_dgrad(const Stage& S, int B, const uint16_t* dzb, const EPI& epi, hipStream_t s) {
  return S.Ci == 64 ? conv_dgrad_ci<64>(S, B, dzb, epi, s) : conv_dgrad_ci<32>(S, B, dzb, epi, s);
}

// Weight and bias gradients of a stage into out = g + ow ([9 Ci][Co], then the Co biases):
// chunks of input rows per work-group (about 256 work-groups, within P), partials summed in
// chunk order.
template <int CI, int HI>
static hipError_t launch_wgrad(kmi_nn* N, const Stage& S, int B, const uint16_t* dzb, float* out, hipStream_t s) {
  const int nob = (S.Co + 15) / 16, steps = B * HI * HI / 128;
  const long long mn = (9LL * CI + 1) * S.Co;
  // about 1024 work-groups, with at most 8 MB of partials
  long long nz = std::min<long long>(steps, std::max(1, 1024 / nob));
  nz = std::min<long long>(nz, std::max<long long>(2 * 1024 * 1024 / mn, 256 / nob));
  nz = std::max<long long>(1, std::min<long long>(nz, (long long)N->P_elems / mn));
  const int spc = int((steps + nz - 1) / nz);
  nz = (steps + spc - 1) / spc;
  hipLaunchKernelGGL((k_conv_wgrad<CI, HI>), dim3(unsigned(nz), unsigned(nob)), dim3(64 * (CI / 16)), 0, s, S.in, dzb,
                     S.Cp, S.Co, steps, spc, N->P);
  if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
  return launch_split_sum(N->P, int(nz), mn, out, s);
}
template <int CI>
static hipError_t conv_wgrad_ci(kmi_nn* N, const Stage& S, int B, const uint16_t* dzb, float* out, hipStream_t s) {
  switch (S.Hi) {
    case 16: return launch_wgrad<CI, 16>(N, S, B, dzb, out, s);
    case 32: return launch_wgrad<CI, 32>(N, S, B, dzb, out, s);
    default: return launch_wgrad<CI, 64>(N, S, B, dzb, out, s);
  }
}
static hipError_t conv_wgrad(kmi_nn* N, const Stage& S, int B, const uint16_t* dzb, float* out, hipStream_t s) {
  return S.Ci == 64 ? conv_wgrad_ci<64>(N, S, B, dzb, out, s) : conv_wgrad_ci<32>(N, S, B, dzb, out, s);
}

static void stages(const kmi_nn* N, Stage (&st)[3]) {
  const kmi_nn_desc& d = N->d;
  st[0] = {N->h2b, d.latent_hw, d.latent_c, d.c1, d.c1, N->wb + N->oC1, N->oC1, N->obc1};
  st[1] = {N->h3b, N->H1, d.c1, d.c2, d.c2, N->wb + N->oC2, N->oC2, N->obc2};
  st[2] = {N->h4b, N->H2, d.c2, d.out_c, N->Cp3, w3_bf16(N), N->oC3, N->obc3};
}

// Forward pass for batch B from N->x (FP32) into N->y.
static int nn_forward(kmi_nn* N, int B) {
  const kmi_nn_desc& d = N->d;
  hipStream_t s = N->ctx->stream;
  const uint16_t* wb = N->wb;
  const float* w = N->w;
  hipLaunchKernelGGL(k_to_bf16, g1((long long)B * d.in_dim), dim3(256), 0, s, N->x, (long long)B * d.in_dim, N->xb);
  // dense 1: [B, in] · W1[in, hidden], bias, SiLU
  NN_CHECK((mm<64, 64, 2>(B, d.hidden, d.in_dim, 1, DenseK{N->xb, d.in_dim, B, d.in_dim},
                          DenseR{wb + N->oW1, d.hidden, d.hidden, d.in_dim},
                          EpiBiasAct{w + N->ob1, N->z1, nullptr, N->h1b, 1}, s)));
  // dense 2 → latent (NHWC)
  NN_CHECK((mm<64, 64, 2>(B, N->L, d.hidden, 1, DenseK{N->h1b, d.hidden, B, d.hidden},
                          DenseR{wb + N->oW2, N->L, N->L, d.hidden},
                          EpiBiasAct{w + N->ob2, N->z2, nullptr, N->h2b, 1}, s)));
  // conv 1..3 on the ×2 up-sampled inputs; bias and SiLU fused (none after the last)
  Stage st[3];
  stages(N, st);
  float* zout[3] = {N->z3, N->z4, N->y};
  uint16_t* hout[3] = {N->h3b, N->h4b, nullptr};
  for (int k = 0; k < 3; ++k) {
    const Stage& S = st[k];
    const int Ho = 2 * S.Hi, M = B * Ho * Ho;
    const EpiBiasAct epi{w + S.ob, zout[k], nullptr, hout[k], k < 2 ? 1 : 0};
    if (N->wt[k])
      NN_CHECK(conv_fwd(S, B, N->wt[k], epi, s));
    else
      NN_CHECK(mm_by_n(M, S.Co, 9 * S.Ci, 1, ConvFwdA{S.in, B, S.Hi, S.Ci}, DenseR{S.wbf, S.Cp, S.Cp, 9 * S.Ci}, epi,
                       s));
  }
  return KMI_OK;
}

// Backward pass from N->dyb (BF16 loss gradient, [B·H3²][Cp3]) into N->g (all parameter gradients).
static int nn_backward(kmi_nn* N, int B) {
  const kmi_nn_desc& d = N->d;
  hipStream_t s = N->ctx->stream;
  const uint16_t* wb = N->wb;
  float* g = N->g;
  Stage st[3];
  stages(N, st);
  const float* zin[3] = {N->z2, N->z3, N->z4};  // pre-activations of the stage inputs (SiLU)
  // gradient w.r.t. the stage output pre-activation in BF16 [M, Cp]
  const uint16_t* dzb = N->dyb;
  int pp = 0;
  for (int k = 2; k >= 0; --k) {
    const Stage& S = st[k];
    const int Ho = 2 * S.Hi, Mo = B * Ho * Ho;
    // weight gradient dW[(ky,kx,c), o] = Σ_pixels A(pixel, (ky,kx,c)) dZ[pixel, o], with a row
    // of ones below A: its result row is the bias gradient Σ_pixels dZ[pixel, o], which the
    // flat layout stores right after the weights (g + ob = g + ow + 9 Ci Co)
    if (wgrad_direct(S.Ci, S.Hi)) {
      NN_CHECK(conv_wgrad(N, S, B, dzb, g + S.ow, s));
    } else {
      const int Mw = 9 * S.Ci + 1;
      const int sp = pick_splits(N, 64, S.Co > 32 ? 64 : (S.Co > 16 ? 32 : 16), Mw, S.Co, Mo, 2048, 512);
      NN_CHECK(mm_wgrad(Mw, S.Co, Mo, sp, WithOnes<ConvWgradA>{ConvWgradA{S.in, B, S.Hi, S.Ci}, 9 * S.Ci, Mo},
                        DenseR{dzb, S.Cp, S.Cp, Mo}, EpiStore{N->P}, s));
      NN_CHECK(split_sum(N->P, Mw, S.Co, Mo, sp, g + S.ow, s));
    }
    // data gradient into the stage input (per input pixel, up-sampling folded in), times
    // the SiLU derivative of the input's pre-activation; split over K when the tiles alone
    // leave the device idle
    const int Mi = B * S.Hi * S.Hi, Kd = 36 * S.Cp;
    const int spd = pick_splits(N, S.Ci > 16 ? 128 : 256, S.Ci > 32 ? 64 : (S.Ci > 16 ? 32 : 16), Mi, S.Ci, Kd,
                                1024, 256);
    if (dgrad_direct(S.Ci, S.Cp)) {
      NN_CHECK(conv_dgrad(S, B, dzb, EpiActGrad{zin[k], nullptr, N->dzb[pp], 1}, s));
    } else if (spd == 1) {
      NN_CHECK(mm_by_n(Mi, S.Ci, Kd, 1, ConvDgradA{dzb, B, S.Hi, S.Cp}, ConvDgradB{S.wbf, S.Ci, S.Cp},
                       EpiActGrad{zin[k], nullptr, N->dzb[pp], 1}, s));
    } else {
      NN_CHECK(mm_by_n(Mi, S.Ci, Kd, spd, ConvDgradA{dzb, B, S.Hi, S.Cp}, ConvDgradB{S.wbf, S.Ci, S.Cp},
                       EpiStore{N->P}, s));
      int sp2 = 1;
      split_chunk(Kd, spd, &sp2);
      const long long ni = (long long)Mi * S.Ci;
      hipLaunchKernelGGL(k_split_act, dim3(unsigned((ni + 15) / 16)), dim3(256), 0, s, N->P, sp2, ni, zin[k], 1,
                         nullptr, N->dzb[pp]);
      NN_CHECK(hipGetLastError());
    }
    dzb = N->dzb[pp];
    pp ^= 1;
  }
  // dense 2: dZ2 is in dzb as [B, L]; weight and bias gradients (a row of ones below
  // h1ᵀ; b2 follows W2 in the flat layout)
  NN_CHECK((mm<64, 64, 2>(d.hidden + 1, N->L, B, 1,
                          WithOnes<DenseR>{DenseR{N->h1b, d.hidden, d.hidden, B}, d.hidden, B},
                          DenseR{dzb, N->L, N->L, B}, EpiStore{g + N->oW2}, s)));
  // dH1 = dZ2 W2ᵀ (long reduction over the latent: split-K)
  {
    const int sp = pick_splits(N, 64, 64, B, d.hidden, N->L, 512, 512);
    NN_CHECK((mm<64, 64, 2>(B, d.hidden, N->L, sp, DenseK{dzb, N->L, B, N->L}, DenseK{wb + N->oW2, N->L, d.hidden, N->L},
                            EpiStore{N->P}, s)));
    int sp2 = 1;
    split_chunk(N->L, sp, &sp2);
    const long long nh = (long long)B * d.hidden;
    hipLaunchKernelGGL(k_split_act, dim3(unsigned((nh + 15) / 16)), dim3(256), 0, s, N->P, sp2, nh, N->z1, 1,
                       nullptr, N->dzb[pp]);
    NN_CHECK(hipGetLastError());
  }
  // dense 1: weight and bias gradients (a row of ones below xᵀ; b1 follows W1 in the flat
  // layout)
  NN_CHECK((mm<64, 64, 2>(d.in_dim + 1, d.hidden, B, 1,
                          WithOnes<DenseR>{DenseR{N->xb, d.in_dim, d.in_dim, B}, d.in_dim, B},
                          DenseR{N->dzb[pp], d.hidden, d.hidden, B}, EpiStore{g + N->oW1}, s)));
  return KMI_OK;
}

// MSE loss and its gradient dY; the loss lands in `slot` (device; default lout) and, when
// `loss` is given, is read back.
static int nn_loss(kmi_nn* N, int B, double* loss, double* slot = nullptr) {
  hipStream_t s = N->ctx->stream;
  const long long n = out_elems(N, B);
  const int nb = int(std::min<long long>(1024, (n + 255) / 256));
  if (!slot) slot = N->lout;
  hipLaunchKernelGGL(k_mse, dim3(unsigned(nb)), dim3(256), 0, s, N->y, N->tgt, n, N->d.out_c, N->Cp3, N->dyb,
                     N->lpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, s, N->lpart, nb, slot);
  NN_CHECK(hipGetLastError());
  if (loss) {
    NN_CHECK(hipMemcpyAsync(loss, slot, 8, hipMemcpyDeviceToHost, s));
    NN_CHECK(hipStreamSynchronize(s));
  }
  return KMI_OK;
}

static int nn_adam(kmi_nn* N, float lr) {
  N->adam_t += 1;
  const float c1 = 1.f - powf(0.9f, float(N->adam_t)), c2 = 1.f - powf(0.999f, float(N->adam_t));
  hipLaunchKernelGGL(k_adam, g1(N->nparam), dim3(256), 0, N->ctx->stream, N->w, N->g, N->m, N->v, N->wb, N->nparam,
                     lr, c1, c2);
  NN_CHECK(hipGetLastError());
  NN_CHECK(refresh_weight_copies(N, N->ctx->stream));
  return KMI_OK;
}

// splitmix64 → N(0, 1) (Box–Muller), for reproducible weight initialisation.
static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static double gauss(uint64_t& s) {
  const double u1 = (double(splitmix(s) >> 11) + 0.5) / 9007199254740992.0;
  const double u2 = (double(splitmix(s) >> 11) + 0.5) / 9007199254740992.0;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

extern "C" {

void kmi_nn_desc_init(kmi_nn_desc* d) {
  if (!d) return;
  std::memset(d, 0, sizeof(*d));
  d->in_dim = 2;
  d->hidden = 256;
  d->latent_hw = 16;
  d->latent_c = 64;
  d->c1 = 64;
  d->c2 = 32;
  d->out_c = 3;
  d->max_batch = 32;
  d->seed = 1234;
}

int kmi_nn_create(kmi_ctx* ctx, const kmi_nn_desc* d, kmi_nn** out) {
  if (!ctx) return fail(nullptr, KMI_ERR_ARG, "kmi_nn_create: ctx is NULL");
  if (!d || !out) return fail(ctx, KMI_ERR_ARG, "kmi_nn_create: NULL argument");
  *out = nullptr;
  if (d->in_dim < 1 || d->hidden < 1 || d->latent_hw < 1 || d->latent_c < 1 || d->c1 < 1 || d->c2 < 1 ||
      d->out_c < 1 || d->max_batch < 1 || d->max_batch > 4096 || 8LL * d->latent_hw > 4096)
    return fail(ctx, KMI_ERR_ARG, "kmi_nn_create: sizes out of range");
  if (d->latent_c % 8 || d->c1 % 8 || d->c2 % 8)
    return fail(ctx, KMI_ERR_ARG, "kmi_nn_create: latent_c, c1 and c2 must be multiples of 8");
  {
    // the kernels index pixels and activation elements in 32 bits
    const long long h3 = 8LL * d->latent_hw, cmax = std::max({d->latent_c, d->c1, d->c2, d->out_c + 7, d->hidden});
    if ((long long)d->max_batch * h3 * h3 * cmax >= (1LL << 31) ||
        (long long)d->hidden * d->latent_hw * d->latent_hw * d->latent_c >= (1LL << 31))
      return fail(ctx, KMI_ERR_ARG, "kmi_nn_create: max_batch·(8·latent_hw)²·channels must stay below 2^31");
  }
  KMI_HIP(ctx, hipSetDevice(ctx->device));
  kmi_nn* N = new (std::nothrow) kmi_nn();
  if (!N) return fail(ctx, KMI_ERR_NOMEM, "kmi_nn_create: out of host memory");
  N->ctx = ctx;
  N->d = *d;
  if (int rc = nn_alloc(N)) {
    kmi_nn_destroy(N);
    return rc;
  }
  // He-normal weights (std √(2/fan_in)), zero biases; the output layer scaled by 0.1
  std::vector<float> w(N->nparam, 0.f);
  uint64_t s = d->seed;
  auto fill = [&](long long off, long long n, double fan_in, double gain) {
    const double sd = gain * std::sqrt(2.0 / fan_in);
    for (long long i = 0; i < n; ++i) w[off + i] = float(sd * gauss(s));
  };
  fill(N->oW1, (long long)d->in_dim * d->hidden, d->in_dim, 1.0);
  fill(N->oW2, (long long)d->hidden * N->L, d->hidden, 1.0);
  fill(N->oC1, 9LL * d->latent_c * d->c1, 9.0 * d->latent_c, 1.0);
  fill(N->oC2, 9LL * d->c1 * d->c2, 9.0 * d->c1, 1.0);
  fill(N->oC3, 9LL * d->c2 * d->out_c, 9.0 * d->c2, 0.1);
  if (int rc = kmi_nn_set_params(N, w.data())) {
    kmi_nn_destroy(N);
    return rc;
  }
  *out = N;
  return KMI_OK;
}

int kmi_nn_destroy(kmi_nn* N) {
  if (!N) return KMI_OK;
  (void)hipSetDevice(N->ctx->device);
  (void)hipStreamSynchronize(N->ctx->stream);
  void* ptrs[] = {N->w, N->g, N->m, N->v, N->wb, N->w3p, N->x, N->xb, N->tgt, N->z1, N->h1b, N->z2, N->h2b,
                  N->z3, N->h3b, N->z4, N->h4b, N->y, N->dyb, N->dzb[0],
                  N->dzb[1], N->P, N->lpart, N->lout, N->idx, N->dsX, N->dsY, N->perm_d,
                  N->lhist, N->wt[0], N->wt[1], N->wt[2]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete N;
  return KMI_OK;
}

int kmi_nn_param_count(const kmi_nn* N, int64_t* n) {
  if (!N || !n) return fail(nullptr, KMI_ERR_ARG, "kmi_nn_param_count: NULL argument");
  *n = N->nparam;
  return KMI_OK;
}

int kmi_nn_get_params(kmi_nn* N, float* host) {
  if (!N || !host) return fail(nullptr, KMI_ERR_ARG, "kmi_nn_get_params: NULL argument");
  KMI_HIP(N->ctx, hipSetDevice(N->ctx->device));
  KMI_HIP(N->ctx, hipMemcpy(host, N->w, N->nparam * 4, hipMemcpyDeviceToHost));
  return KMI_OK;
}

int kmi_nn_set_params(kmi_nn* N, const float* host) {
  if (!N || !host) return fail(nullptr, KMI_ERR_ARG, "kmi_nn_set_params: NULL argument");
  KMI_HIP(N->ctx, hipSetDevice(N->ctx->device));
  KMI_HIP(N->ctx, hipMemcpy(N->w, host, N->nparam * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf16, g1(N->nparam), dim3(256), 0, N->ctx->stream, N->w, N->nparam, N->wb);
  KMI_HIP(N->ctx, hipGetLastError());
  KMI_HIP(N->ctx, refresh_weight_copies(N, N->ctx->stream));
  KMI_HIP(N->ctx, hipStreamSynchronize(N->ctx->stream));
  N->adam_t = 0;
  KMI_HIP(N->ctx, hipMemset(N->m, 0, N->nparam * 4));
  KMI_HIP(N->ctx, hipMemset(N->v, 0, N->nparam * 4));
  return KMI_OK;
}

static int nn_batch_ok(kmi_nn* N, int B) {
  if (B < 1 || B > N->d.max_batch) return fail(N->ctx, KMI_ERR_ARG, "batch size must be in 1..max_batch");
  return KMI_OK;
}

int kmi_nn_forward(kmi_nn* N, int32_t B, const float* x, float* y) {
  if (!N || !x || !y) return fail(nullptr, KMI_ERR_ARG, "kmi_nn_forward: NULL argument");
  if (int rc = nn_batch_ok(N, B)) return rc;
  kmi_ctx* c = N->ctx;
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipMemcpyAsync(N->x, x, (size_t)B * N->d.in_dim * 4, hipMemcpyHostToDevice, c->stream));
  if (int rc = nn_forward(N, B)) return rc;
  KMI_HIP(c, hipMemcpyAsync(y, N->y, out_elems(N, B) * 4, hipMemcpyDeviceToHost, c->stream));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  return KMI_OK;
}

int kmi_nn_loss_grad(kmi_nn* N, int32_t B, const float* x, const float* target, double* loss, float* grad) {
  if (!N || !x || !target) return fail(nullptr, KMI_ERR_ARG, "kmi_nn_loss_grad: NULL argument");
  if (int rc = nn_batch_ok(N, B)) return rc;
  kmi_ctx* c = N->ctx;
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipMemcpyAsync(N->x, x, (size_t)B * N->d.in_dim * 4, hipMemcpyHostToDevice, c->stream));
  KMI_HIP(c, hipMemcpyAsync(N->tgt, target, out_elems(N, B) * 4, hipMemcpyHostToDevice, c->stream));
  if (int rc = nn_forward(N, B)) return rc;
  if (int rc = nn_loss(N, B, loss)) return rc;
  if (int rc = nn_backward(N, B)) return rc;
  if (grad) KMI_HIP(c, hipMemcpyAsync(grad, N->g, N->nparam * 4, hipMemcpyDeviceToHost, c->stream));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  return KMI_OK;
}

int kmi_nn_train_step(kmi_nn* N, int32_t B, const float* x, const float* target, float lr, double* loss) {
  if (int rc = kmi_nn_loss_grad(N, B, x, target, loss, nullptr)) return rc;
  if (int rc = nn_adam(N, lr)) return rc;
  KMI_HIP(N->ctx, hipStreamSynchronize(N->ctx->stream));
  return KMI_OK;
}

int kmi_nn_set_dataset(kmi_nn* N, int32_t n, const float* x, const float* y) {
  if (!N || !x || !y || n < 1) return fail(nullptr, KMI_ERR_ARG, "kmi_nn_set_dataset: bad arguments");
  kmi_ctx* c = N->ctx;
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  for (void** p : {(void**)&N->dsX, (void**)&N->dsY, (void**)&N->perm_d, (void**)&N->lhist}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  N->dsN = 0;
  const long long dout = out_elems(N, 1);
  KMI_HIP(c, hipMalloc(&N->dsX, (size_t)n * N->d.in_dim * 4));
  KMI_HIP(c, hipMalloc(&N->dsY, (size_t)n * dout * 4));
  KMI_HIP(c, hipMalloc(&N->perm_d, (size_t)n * 4));
  KMI_HIP(c, hipMalloc(&N->lhist, (size_t)n * 8));
  KMI_HIP(c, hipMemcpy(N->dsX, x, (size_t)n * N->d.in_dim * 4, hipMemcpyHostToDevice));
  KMI_HIP(c, hipMemcpy(N->dsY, y, (size_t)n * dout * 4, hipMemcpyHostToDevice));
  N->dsN = n;
  return KMI_OK;
}

// Epochs of Adam over the device dataset in mini-batches of B (a Fisher–Yates shuffle per
// epoch from `seed`; the last partial batch is dropped).  Returns the mean loss of the
// final epoch.  An epoch's permutation is uploaded once and the training steps queue on the
// stream without host round trips; each step's loss stays on the device until the epoch's
// end.
int kmi_nn_train_epochs(kmi_nn* N, int32_t epochs, int32_t B, float lr, uint64_t seed, double* mean_loss) {
  if (!N) return fail(nullptr, KMI_ERR_ARG, "kmi_nn_train_epochs: NULL argument");
  kmi_ctx* c = N->ctx;
  if (N->dsN < 1) return fail(c, KMI_ERR_STATE, "kmi_nn_train_epochs: no dataset");
  if (int rc = nn_batch_ok(N, B)) return rc;
  if (B > N->dsN) return fail(c, KMI_ERR_ARG, "kmi_nn_train_epochs: batch larger than the dataset");
  KMI_HIP(c, hipSetDevice(c->device));
  const long long dout = out_elems(N, 1);
  std::vector<int> perm(N->dsN);
  uint64_t s = seed;
  const int nb = N->dsN / B;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < N->dsN; ++i) perm[i] = i;
    for (int i = N->dsN - 1; i > 0; --i) std::swap(perm[i], perm[int(splitmix(s) % uint64_t(i + 1))]);
    // the previous epoch's gathers have read the permutation before it is replaced
    KMI_HIP(c, hipStreamSynchronize(c->stream));
    KMI_HIP(c, hipMemcpyAsync(N->perm_d, perm.data(), (size_t)N->dsN * 4, hipMemcpyHostToDevice, c->stream));
    for (int b = 0; b < nb; ++b) {
      hipLaunchKernelGGL(k_gather, dim3(unsigned((dout / 4 + 255) / 256), unsigned(B)), dim3(256), 0, c->stream,
                         N->dsX, N->dsY, N->perm_d + (size_t)b * B, N->d.in_dim, int(dout / 4), N->x, N->tgt);
      KMI_HIP(c, hipGetLastError());
      if (int rc = nn_forward(N, B)) return rc;
      if (int rc = nn_loss(N, B, nullptr, N->lhist + b)) return rc;
      if (int rc = nn_backward(N, B)) return rc;
      if (int rc = nn_adam(N, lr)) return rc;
    }
  }
  std::vector<double> lh(nb);
  KMI_HIP(c, hipMemcpyAsync(lh.data(), N->lhist, (size_t)nb * 8, hipMemcpyDeviceToHost, c->stream));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) acc += lh[b];
  if (mean_loss) *mean_loss = epochs > 0 ? acc / nb : 0.0;
  return KMI_OK;
}

}  // extern "C"

// C = A·B for host FP32 matrices, products in BF16 on MFMA with FP32 accumulation.
// A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn], C row-major M×N.
// Device buffers: dA, dB (FP32 copies), bA, bB (BF16 copies), P (split-K partials, `sp`
// splits of M·N; the sum lands in P's first M·N).
struct GemmBufs {
  float *dA, *dB, *P;
  uint16_t *bA, *bB;
};
struct GemmShape {
  long long na, nb;  // elements of A and B spanned by the strides
  int splits;        // splits asked for
  int sp;            // splits mm() runs after rounding the K chunks to whole K-steps
};
static bool gemm_shape(int32_t M, int32_t Nn, int32_t K, int64_t sam, int64_t sak, int64_t sbk, int64_t sbn,
                       int32_t splits, GemmShape& g) {
  if (M < 1 || Nn < 1 || K < 1 || splits < 1 || sam < 0 || sak < 0 || sbk < 0 || sbn < 0) return false;
  g.na = (long long)(M - 1) * sam + (long long)(K - 1) * sak + 1;
  g.nb = (long long)(K - 1) * sbk + (long long)(Nn - 1) * sbn + 1;
  g.splits = splits;
  split_chunk(K, splits, &g.sp);  // as mm() rounds the chunks
  return true;
}
// A(m, k) = a[m·sam + k·sak] and B(k, n) = b[k·sbk + n·sbn] read through the loader their
// strides allow: K-major (unit stride along k), row-major (unit stride along the rows), or
// element by element.
template <class LA>
static hipError_t gemm_strided_b(int M, int Nn, int K, const LA& la, const uint16_t* b, int64_t sbk, int64_t sbn,
                                 int splits, float* P, hipStream_t s) {
  if (sbk == 1) return mm<64, 64, 2>(M, Nn, K, splits, la, DenseK{b, sbn, Nn, K}, EpiStore{P}, s);
  if (sbn == 1) return mm<64, 64, 2>(M, Nn, K, splits, la, DenseR{b, sbk, Nn, K}, EpiStore{P}, s);
  return mm<64, 64, 2>(M, Nn, K, splits, la, DenseS{b, sbn, sbk, Nn, K}, EpiStore{P}, s);
}
static hipError_t gemm_strided(int M, int Nn, int K, const uint16_t* a, int64_t sam, int64_t sak, const uint16_t* b,
                               int64_t sbk, int64_t sbn, int splits, float* P, hipStream_t s) {
  if (sak == 1) return gemm_strided_b(M, Nn, K, DenseK{a, sam, M, K}, b, sbk, sbn, splits, P, s);
  if (sam == 1) return gemm_strided_b(M, Nn, K, DenseR{a, sak, M, K}, b, sbk, sbn, splits, P, s);
  return gemm_strided_b(M, Nn, K, DenseS{a, sam, sak, M, K}, b, sbk, sbn, splits, P, s);
}

static hipError_t gemm_bf16_run(hipStream_t s, int32_t M, int32_t Nn, int32_t K, const float* A, int64_t sam,
                                int64_t sak, const float* B, int64_t sbk, int64_t sbn, const GemmShape& g,
                                const GemmBufs& d, float* C) {
  hipError_t e = hipMemcpyAsync(d.dA, A, g.na * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d.dB, B, g.nb * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, g1(g.na), dim3(256), 0, s, d.dA, g.na, d.bA);
    hipLaunchKernelGGL(k_to_bf16, g1(g.nb), dim3(256), 0, s, d.dB, g.nb, d.bB);
    e = gemm_strided(M, Nn, K, d.bA, sam, sak, d.bB, sbk, sbn, g.splits, d.P, s);
  }
  if (e == hipSuccess) {
    const long long mn = (long long)M * Nn;
    hipLaunchKernelGGL(k_epilogue, g1(mn), dim3(256), 0, s, d.P, g.sp, mn, Nn, (const float*)nullptr, 0, d.P,
                       (float*)nullptr, (uint16_t*)nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(C, d.P, (size_t)M * Nn * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

extern "C" {

int kmi_gemm_bf16(kmi_ctx* ctx, int32_t M, int32_t Nn, int32_t K, const float* A, int64_t sam, int64_t sak,
                  const float* B, int64_t sbk, int64_t sbn, int32_t splits, float* C) {
  GemmShape g;
  if (!ctx || !A || !B || !C || !gemm_shape(M, Nn, K, sam, sak, sbk, sbn, splits, g))
    return fail(ctx, KMI_ERR_ARG, "kmi_gemm_bf16: bad arguments");
  KMI_HIP(ctx, hipSetDevice(ctx->device));
  GemmBufs d{nullptr, nullptr, nullptr, nullptr, nullptr};
  hipError_t e = hipMalloc(&d.dA, g.na * 4);
  if (e == hipSuccess) e = hipMalloc(&d.dB, g.nb * 4);
  if (e == hipSuccess) e = hipMalloc(&d.bA, g.na * 2);
  if (e == hipSuccess) e = hipMalloc(&d.bB, g.nb * 2);
  if (e == hipSuccess) e = hipMalloc(&d.P, (size_t)g.sp * M * Nn * 4);
  if (e == hipSuccess) e = gemm_bf16_run(ctx->stream, M, Nn, K, A, sam, sak, B, sbk, sbn, g, d, C);
  for (void* p : {(void*)d.dA, (void*)d.dB, (void*)d.bA, (void*)d.bB, (void*)d.P})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return hip_fail(ctx, e, "kmi_gemm_bf16");
  return KMI_OK;
}

// The same product with every device buffer carved from the static workspace of this GPU:
// no allocation.  Calls on one GPU share the workspace, so calls that overlap on one GPU race.
int kmi_gemm_bf16_noalloc(kmi_ctx* ctx, int32_t M, int32_t Nn, int32_t K, const float* A, int64_t sam,
                          int64_t sak, const float* B, int64_t sbk, int64_t sbn, int32_t splits, float* C) {
  GemmShape g;
  if (!ctx || !A || !B || !C || !gemm_shape(M, Nn, K, sam, sak, sbk, sbn, splits, g))
    return fail(ctx, KMI_ERR_ARG, "kmi_gemm_bf16_noalloc: bad arguments");
  auto up = [](unsigned long long b) { return (b + 255ull) / 256ull * 256ull; };
  const unsigned long long oA = 0, oB = oA + up(g.na * 4ull), obA = oB + up(g.nb * 4ull),
                           obB = obA + up(g.na * 2ull), oP = obB + up(g.nb * 2ull),
                           end = oP + up((unsigned long long)g.sp * M * Nn * 4ull);
  if (end > KMI_NOALLOC_WORKSPACE_BYTES)
    return fail(ctx, KMI_ERR_ARG, "kmi_gemm_bf16_noalloc: operands and partials need " + std::to_string(end) +
                                      " bytes, more than the " + std::to_string(KMI_NOALLOC_WORKSPACE_BYTES) +
                                      "-byte workspace; use kmi_gemm_bf16");
  unsigned char* base = static_cast<unsigned char*>(noalloc_workspace(ctx->device));
  if (!base) return fail(ctx, KMI_ERR_STATE, "kmi_gemm_bf16_noalloc: no workspace address for this device");
  KMI_HIP(ctx, hipSetDevice(ctx->device));
  const GemmBufs d{reinterpret_cast<float*>(base + oA), reinterpret_cast<float*>(base + oB),
                   reinterpret_cast<float*>(base + oP), reinterpret_cast<uint16_t*>(base + obA),
                   reinterpret_cast<uint16_t*>(base + obB)};
  hipError_t e = gemm_bf16_run(ctx->stream, M, Nn, K, A, sam, sak, B, sbk, sbn, g, d, C);
  if (e != hipSuccess) return hip_fail(ctx, e, "kmi_gemm_bf16_noalloc");
  return KMI_OK;
}

}  // extern "C"
