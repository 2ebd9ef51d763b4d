// kmi_nn.hip — the learned warm start: a small conv decoder trained and run on the GPU.
//
// Network (no reference counterpart; the injection point is the initial populations, as
// f_init in src/drivers/gnf_periodic_steady.jl:114-150):
//   x (in_dim) → dense → SiLU → dense → SiLU → latent hw×hw×c0 (NHWC)
//   → [nearest ×2, 3×3 conv, SiLU] → [×2, conv, SiLU] → [×2, conv] → (8hw)²×out_c.
// Every matrix product is one kernel: an implicit GEMM on BF16 MFMA tiles
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate) whose operand loaders compute the
// im2col / transposed-conv addresses on the fly (no im2col buffers).  Weights are kept
// in FP32 (Adam master copy) and mirrored to BF16 for the products; activations and
// gradients enter the products as BF16, everything else is FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "kmi_internal.h"

namespace kmi {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 → BF16, round to nearest even (finite inputs).
KMI_HD uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
KMI_HD float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, uint32_t(h) << 16);
}

// ---- operand loaders: element (i, j) of a logical matrix, 0 outside ------------------------
struct LoadDense {  // p[i*s0 + j*s1]
  const uint16_t* p;
  long long s0, s1;
  int n0, n1;
  KMI_DEV uint16_t operator()(int i, int j) const {
    return (i < n0 && j < n1) ? p[(long long)i * s0 + (long long)j * s1] : uint16_t(0);
  }
};

// Forward conv after nearest ×2 up-sampling, 3×3, zero padding 1, NHWC input [B, Hi, Hi, C]:
// A(m = (b, y, x) on the Ho = 2Hi grid, k = (ky, kx, c)).  TRANS swaps (i, j) → (k, m)
// (the weight-gradient product reads the same matrix transposed).
template <bool TRANS>
struct LoadConvUp {
  const uint16_t* in;
  int B, Hi, C;
  KMI_DEV uint16_t operator()(int i, int j) const {
    const int m = TRANS ? j : i, k = TRANS ? i : j;
    const int Ho = 2 * Hi;
    if (m >= B * Ho * Ho || k >= 9 * C) return 0;
    const int b = m / (Ho * Ho), r = m % (Ho * Ho), y = r / Ho, x = r % Ho;
    const int t = k / C, c = k % C, yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    if (yy < 0 || yy >= Ho || xx < 0 || xx >= Ho) return 0;
    return in[(((long long)b * Hi + (yy >> 1)) * Hi + (xx >> 1)) * C + c];
  }
};

// Data gradient through conv + up-sampling: A(m = (b, yi, xi) on the Hi grid,
// k = (dy, dx, ky, kx, o)) = dZ[b, 2yi+dy−ky+1, 2xi+dx−kx+1, o] on the Ho grid.
struct LoadConvUpGradIn {
  const uint16_t* dz;
  int B, Hi, Co;
  KMI_DEV uint16_t operator()(int m, int k) const {
    const int Ho = 2 * Hi;
    if (m >= B * Hi * Hi || k >= 36 * Co) return 0;
    const int b = m / (Hi * Hi), r = m % (Hi * Hi), yi = r / Hi, xi = r % Hi;
    const int o = k % Co, s = k / Co, t = s % 9, d = s / 9;
    const int y = 2 * yi + (d >> 1) - t / 3 + 1, x = 2 * xi + (d & 1) - t % 3 + 1;
    if (y < 0 || y >= Ho || x < 0 || x >= Ho) return 0;
    return dz[(((long long)b * Ho + y) * Ho + x) * Co + o];
  }
};
// B(k = (dy, dx, ky, kx, o), n = c) = W[(ky, kx, c), o] with W row-major [9 Ci, Co].
struct LoadConvWeightT {
  const uint16_t* w;
  int Ci, Co;
  KMI_DEV uint16_t operator()(int k, int n) const {
    if (k >= 36 * Co || n >= Ci) return 0;
    const int o = k % Co, t = (k / Co) % 9;
    return w[((long long)t * Ci + n) * Co + o];
  }
};

// ---- four consecutive elements (i, j..j+3), 0 at j+e >= jend, packed into a uint2 ---------
KMI_DEV uint2 pack4(const uint16_t (&v)[4]) {
  return make_uint2(uint32_t(v[0]) | (uint32_t(v[1]) << 16), uint32_t(v[2]) | (uint32_t(v[3]) << 16));
}
KMI_DEV bool aligned8(const uint16_t* p) { return (reinterpret_cast<uintptr_t>(p) & 7u) == 0; }
KMI_DEV uint2 ld8(const uint16_t* p) { return *reinterpret_cast<const uint2*>(p); }

// Generic: four calls of the loader.
template <class LD>
KMI_DEV uint2 load4(const LD& ld, int i, int j, int jend) {
  uint16_t v[4];
  for (int e = 0; e < 4; ++e) v[e] = j + e < jend ? ld(i, j + e) : uint16_t(0);
  return pack4(v);
}
// Dense row-major operand: one 8-byte load when the four elements are contiguous.
KMI_DEV uint2 load4(const LoadDense& ld, int i, int j, int jend) {
  if (i >= ld.n0) return make_uint2(0u, 0u);
  const uint16_t* q = ld.p + (long long)i * ld.s0 + (long long)j * ld.s1;
  if (ld.s1 == 1 && j + 3 < min(jend, ld.n1) && aligned8(q)) return ld8(q);
  uint16_t v[4];
  for (int e = 0; e < 4; ++e) v[e] = (j + e < jend && j + e < ld.n1) ? q[(long long)e * ld.s1] : uint16_t(0);
  return pack4(v);
}
// Forward conv operand: the pixel m is decoded once; four consecutive k = (tap, c..c+3) of
// one tap are four consecutive channels of one input pixel (C a multiple of 4).
KMI_DEV uint2 load4(const LoadConvUp<false>& ld, int m, int k, int jend) {
  const int Ho = 2 * ld.Hi, C = ld.C;
  if (m >= ld.B * Ho * Ho) return make_uint2(0u, 0u);
  const int b = m / (Ho * Ho), r = m % (Ho * Ho), y = r / Ho, x = r % Ho;
  const uint16_t* base = ld.in + (long long)b * ld.Hi * ld.Hi * C;
  if (C % 4 == 0 && k % 4 == 0 && k + 3 < min(jend, 9 * C)) {
    const int t = k / C, c = k % C, yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    if (yy < 0 || yy >= Ho || xx < 0 || xx >= Ho) return make_uint2(0u, 0u);
    const uint16_t* q = base + ((long long)(yy >> 1) * ld.Hi + (xx >> 1)) * C + c;
    if (aligned8(q)) return ld8(q);
  }
  uint16_t v[4];
  for (int e = 0; e < 4; ++e) {
    const int kk = k + e;
    v[e] = 0;
    if (kk >= jend || kk >= 9 * C) continue;
    const int t = kk / C, c = kk % C, yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    if (yy < 0 || yy >= Ho || xx < 0 || xx >= Ho) continue;
    v[e] = base[((long long)(yy >> 1) * ld.Hi + (xx >> 1)) * C + c];
  }
  return pack4(v);
}
// Data-gradient operand: four consecutive k = (d, tap, o..o+3) are four consecutive output
// channels of one output pixel (Co a multiple of 4).
KMI_DEV uint2 load4(const LoadConvUpGradIn& ld, int m, int k, int jend) {
  const int Hi = ld.Hi, Ho = 2 * Hi, Co = ld.Co;
  if (m >= ld.B * Hi * Hi) return make_uint2(0u, 0u);
  const int b = m / (Hi * Hi), r = m % (Hi * Hi), yi = r / Hi, xi = r % Hi;
  const uint16_t* base = ld.dz + (long long)b * Ho * Ho * Co;
  if (Co % 4 == 0 && k % 4 == 0 && k + 3 < min(jend, 36 * Co)) {
    const int o = k % Co, s = k / Co, t = s % 9, d = s / 9;
    const int y = 2 * yi + (d >> 1) - t / 3 + 1, x = 2 * xi + (d & 1) - t % 3 + 1;
    if (y < 0 || y >= Ho || x < 0 || x >= Ho) return make_uint2(0u, 0u);
    const uint16_t* q = base + ((long long)y * Ho + x) * Co + o;
    if (aligned8(q)) return ld8(q);
  }
  uint16_t v[4];
  for (int e = 0; e < 4; ++e) {
    const int kk = k + e;
    v[e] = 0;
    if (kk >= jend || kk >= 36 * Co) continue;
    const int o = kk % Co, s = kk / Co, t = s % 9, d = s / 9;
    const int y = 2 * yi + (d >> 1) - t / 3 + 1, x = 2 * xi + (d & 1) - t % 3 + 1;
    if (y < 0 || y >= Ho || x < 0 || x >= Ho) continue;
    v[e] = base[((long long)y * Ho + x) * Co + o];
  }
  return pack4(v);
}

// Weight-gradient operand A(k, m..m+3): the tap and channel are decoded once; four
// consecutive pixels (m a multiple of 4, Ho a multiple of 4) lie on one output row.
KMI_DEV uint2 load4(const LoadConvUp<true>& ld, int k, int m, int jend) {
  const int Hi = ld.Hi, Ho = 2 * Hi, C = ld.C, HW = Ho * Ho;
  if (k >= 9 * C) return make_uint2(0u, 0u);
  const int t = k / C, c = k % C, dy = t / 3 - 1, dx = t % 3 - 1;
  uint16_t v[4] = {0, 0, 0, 0};
  if (m % 4 == 0 && Ho % 4 == 0 && m + 3 < min(jend, ld.B * HW)) {
    const int b = m / HW, r = m % HW, y = r / Ho, x = r % Ho, yy = y + dy;
    if (yy >= 0 && yy < Ho) {
      const uint16_t* row = ld.in + ((long long)b * Hi + (yy >> 1)) * Hi * C + c;
      for (int e = 0; e < 4; ++e) {
        const int xx = x + e + dx;
        if (xx >= 0 && xx < Ho) v[e] = row[(long long)(xx >> 1) * C];
      }
    }
    return pack4(v);
  }
  for (int e = 0; e < 4; ++e) v[e] = m + e < jend ? ld(k, m + e) : uint16_t(0);
  return pack4(v);
}
// Data-gradient operand B(k, n..n+3) = W[(tap, n), o]: tap and o decoded once.
KMI_DEV uint2 load4(const LoadConvWeightT& ld, int k, int n, int jend) {
  if (k >= 36 * ld.Co) return make_uint2(0u, 0u);
  const int o = k % ld.Co, t = (k / ld.Co) % 9;
  const uint16_t* q = ld.w + (long long)t * ld.Ci * ld.Co + o;
  uint16_t v[4];
  for (int e = 0; e < 4; ++e)
    v[e] = (n + e < jend && n + e < ld.Ci) ? q[(long long)(n + e) * ld.Co] : uint16_t(0);
  return pack4(v);
}

// ---- split-K GEMM: P[z][m][n] = Σ_{k in split z} A(m,k) B(k,n) --------------------------------
// 256 threads = 4 waves; block tile 64×64; each wave 32×32 = 2×2 MFMA 16×16×16 tiles.
// MFMA operand layout (lane l): A row l%16, k = 4(l/16)..+3; B col l%16, same k;
// D rows 4(l/16)..+3 of col l%16.
template <class LA, class LB>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, int kchunk, LA la, LB lb,
                                              float* __restrict__ out) {
  __shared__ __align__(8) uint16_t As[64][16 + 4];  // rows of 40 B: 8-byte aligned quads
  __shared__ __align__(8) uint16_t Bs[16][64 + 4];  // rows of 136 B
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  v4f acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  // A tile 64×16: thread → row t/4, 4 consecutive k; B tile 16×64: thread → k row t/16,
  // 4 consecutive n.  The quads of K-step k0+16 are loaded into registers while the MFMAs of
  // K-step k0 run, so the operand latency overlaps the products.
  const int ar = t >> 2, ak = (t & 3) * 4, bk = t >> 4, bn = (t & 15) * 4;
  auto fetch_a = [&](int k0) { return load4(la, m0 + ar, k0 + ak, kend); };
  auto fetch_b = [&](int k0) {
    return k0 + bk < kend ? load4(lb, k0 + bk, n0 + bn, 1 << 30) : make_uint2(0u, 0u);
  };
  uint2 qa = fetch_a(kbeg), qb = fetch_b(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    *reinterpret_cast<uint2*>(&As[ar][ak]) = qa;
    *reinterpret_cast<uint2*>(&Bs[bk][bn]) = qb;
    __syncthreads();
    if (k0 + 16 < kend) {
      qa = fetch_a(k0 + 16);
      qb = fetch_b(k0 + 16);
    }
    const int kq = (lane >> 4) * 4;
    for (int a = 0; a < 2; ++a) {
      v4s af;
      const int r = wm + a * 16 + (lane & 15);
      af[0] = As[r][kq];
      af[1] = As[r][kq + 1];
      af[2] = As[r][kq + 2];
      af[3] = As[r][kq + 3];
      for (int b = 0; b < 2; ++b) {
        v4s bfr;
        const int c = wn + b * 16 + (lane & 15);
        bfr[0] = Bs[kq][c];
        bfr[1] = Bs[kq + 1][c];
        bfr[2] = Bs[kq + 2][c];
        bfr[3] = Bs[kq + 3][c];
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bfr, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float* P = out + (long long)blockIdx.z * M * N;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + a * 16 + (lane >> 4) * 4 + r, n = n0 + wn + b * 16 + (lane & 15);
        if (m < M && n < N) P[(long long)m * N + n] = acc[a][b][r];
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
  if (silu) h = z / (1.f + expf(-z));
  if (H) H[i] = h;
  if (Hb) Hb[i] = f2bf(h);
}

// Σ over split-K partials of products with a long reduction and a small output (weight
// gradients): 16 outputs × 16 split lanes per work-group, lane g summing splits g, g+16, …,
// the lane sums added in lane order (a fixed order: the result is deterministic).
__global__ void __launch_bounds__(256) k_split_sum(const float* __restrict__ P, int splits, long long MN,
                                                    float* __restrict__ out) {
  __shared__ float part[16][17];
  const int o = threadIdx.x & 15, g = threadIdx.x >> 4;
  const long long i = (long long)blockIdx.x * 16 + o;
  float z = 0.f;
  if (i < MN)
    for (int sp = g; sp < splits; sp += 16) z += P[(long long)sp * MN + i];
  part[g][o] = z;
  __syncthreads();
  if (g == 0 && i < MN) {
    float t = 0.f;
    for (int k = 0; k < 16; ++k) t += part[k][o];
    out[i] = t;
  }
}

// dZ = dH ⊙ SiLU'(Z) (or dH when !silu); FP32 and BF16 copies.  nearest-×2 folding of the
// up-sampled data gradient happens inside LoadConvUpGradIn, so dH is already per input pixel.
__global__ void k_act_grad(const float* __restrict__ dH, const float* __restrict__ Z, long long n,
                           int silu, float* __restrict__ dZ, uint16_t* __restrict__ dZb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = dH[i];
  if (silu) {
    const float z = Z[i], s = 1.f / (1.f + expf(-z));
    g = g * (s * (1.f + z * (1.f - s)));
  }
  if (dZ) dZ[i] = g;
  if (dZb) dZb[i] = f2bf(g);
}

// Column sums of X[rows, N] (bias gradients): partial over row chunks, then fixed-order sum.
// Wide rows (N > 128): one thread per column walks the chunk.  Narrow rows: the 256 threads
// of a group split into G = 256 / N row lanes per column (consecutive threads read
// consecutive elements), and the G lane sums of a column are added in lane order.
__global__ void __launch_bounds__(256) k_colsum_partial(const float* __restrict__ X, long long rows,
                                                        int N, int chunk, float* __restrict__ part) {
  const long long r0 = (long long)blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
  if (N > 128) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (long long r = r0; r < r1; ++r) s += X[r * N + n];
    part[(long long)blockIdx.y * N + n] = s;
    return;
  }
  __shared__ float lds[256];
  const int G = 256 / N, t = threadIdx.x, n = t % N, g = t / N;
  float s = 0.f;
  if (g < G)
    for (long long r = r0 + g; r < r1; r += G) s += X[r * N + n];
  lds[t] = s;
  __syncthreads();
  if (t < N) {
    float acc = 0.f;
    for (int k = 0; k < G; ++k) acc += lds[k * N + t];
    part[(long long)blockIdx.y * N + t] = acc;
  }
}
// One group per column: strided partial sums in fixed order, then a fixed-order tree.
__global__ void __launch_bounds__(256) k_colsum_final(const float* __restrict__ part, int nchunk, int N,
                                                      float* __restrict__ out) {
  __shared__ float lds[256];
  const int n = blockIdx.x, t = threadIdx.x;
  float s = 0.f;
  for (int c = t; c < nchunk; c += 256) s += part[(long long)c * N + n];
  lds[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) lds[t] += lds[t + w];
    __syncthreads();
  }
  if (t == 0) out[n] = lds[0];
}

// MSE loss partials and dY = 2 (Y − T) / count.
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             long long n, float* __restrict__ dY,
                                             double* __restrict__ part) {
  __shared__ double lds[4];
  double acc = 0.0;
  const float scale = 2.f / float(n);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    acc += double(d) * double(d);
    dY[i] = scale * d;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (lds[0] + lds[1] + lds[2] + lds[3]) / double(n);
}
// One wave: lane-strided sums in fixed order, then a fixed butterfly.
__global__ void __launch_bounds__(64) k_sum_parts(const double* __restrict__ part, int n, double* __restrict__ out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 64) s += part[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (threadIdx.x == 0) out[0] = s;
}

__global__ void k_to_bf16(const float* __restrict__ x, long long n, uint16_t* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
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

// Gather rows idx[0..B) of the device dataset into the batch buffers.
__global__ void k_gather(const float* __restrict__ X, const float* __restrict__ Y,
                         const int* __restrict__ idx, int B, int din, long long dout,
                         float* __restrict__ xb, float* __restrict__ yb) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)B * dout;
  if (t < total) {
    const int b = int(t / dout);
    yb[t] = Y[(long long)idx[b] * dout + (t % dout)];
  }
  if (t < (long long)B * din) xb[t] = X[(long long)idx[t / din] * din + (t % din)];
}

template <class LA, class LB>
static hipError_t gemm(int M, int N, int K, LA la, LB lb, float* P, int splits, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  int kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 15) / 16 * 16;
  splits = (K + kchunk - 1) / kchunk;
  dim3 g(unsigned((M + 63) / 64), unsigned((N + 63) / 64), unsigned(splits));
  hipLaunchKernelGGL((k_gemm<LA, LB>), g, dim3(256), 0, s, M, N, K, kchunk, la, lb, P);
  return hipGetLastError();
}

static inline dim3 g1(long long n) { return dim3(unsigned((n + 255) / 256)); }

}  // namespace kmi

using namespace kmi;

// ---- the network object -------------------------------------------------------------------
struct kmi_nn {
  kmi_ctx* ctx = nullptr;
  kmi_nn_desc d{};
  int L = 0;  // hw·hw·c0
  int H1 = 0, H2 = 0, H3 = 0;  // spatial sizes after conv 1, 2, 3
  long long nparam = 0;
  // parameter offsets in the flat vector
  long long oW1, ob1, oW2, ob2, oC1, obc1, oC2, obc2, oC3, obc3;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* wb = nullptr;
  int64_t adam_t = 0;
  // activations (max_batch)
  float *x = nullptr, *tgt = nullptr;
  uint16_t* xb = nullptr;
  float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr, *z4 = nullptr, *y = nullptr;
  uint16_t *h1b = nullptr, *h2b = nullptr, *h3b = nullptr, *h4b = nullptr;
  float *dy = nullptr, *dh = nullptr, *dz = nullptr;  // scratch sized for the largest layer
  uint16_t *dyb = nullptr, *dzb = nullptr;
  float* P = nullptr;  // split-K partials
  size_t P_elems = 0;
  float* colpart = nullptr;
  double* lpart = nullptr;
  double* lout = nullptr;
  // device dataset
  float *dsX = nullptr, *dsY = nullptr;
  int dsN = 0;
  int* idx = nullptr;
};

static long long out_elems(const kmi_nn* N, int B) { return (long long)B * N->H3 * N->H3 * N->d.out_c; }

static int nn_alloc(kmi_nn* N) {
  const kmi_nn_desc& d = N->d;
  const int B = d.max_batch;
  N->L = d.latent_hw * d.latent_hw * d.latent_c;
  N->H1 = 2 * d.latent_hw;
  N->H2 = 4 * d.latent_hw;
  N->H3 = 8 * d.latent_hw;
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
  const long long a5 = out_elems(N, B);
  const long long big = std::max({a1, a2, a3, a4, a5});
  // P holds one split of every product (pick_splits adds splits only while they fit): the
  // activations above, the data gradients (B·Hi²·Ci per stage, at most `big`), and the
  // weight gradients of the dense layers (in_dim × hidden, hidden × L) and of the three
  // convolutions (9 Ci × Co).
  const long long wmax = std::max({(long long)d.in_dim * d.hidden, (long long)d.hidden * N->L,
                                   9LL * d.latent_c * d.c1, 9LL * d.c1 * d.c2, 9LL * d.c2 * d.out_c});
  N->P_elems = size_t(std::max({big, wmax, 64LL * 1024 * 1024 / 4}));
  hipError_t e = hipSuccess;
  auto al = [&](void** p, size_t bytes) { if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 16)); };
  al((void**)&N->w, o * 4);
  al((void**)&N->g, o * 4);
  al((void**)&N->m, o * 4);
  al((void**)&N->v, o * 4);
  al((void**)&N->wb, o * 2);
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
  al((void**)&N->dy, a5 * 4);
  al((void**)&N->dyb, a5 * 2);
  al((void**)&N->dh, big * 4);
  al((void**)&N->dz, big * 4);
  al((void**)&N->dzb, big * 2);
  al((void**)&N->P, N->P_elems * 4);
  al((void**)&N->colpart, 1024LL * std::max(N->L, 256) * 4);
  al((void**)&N->lpart, 4096 * 8);
  al((void**)&N->lout, 64);
  al((void**)&N->idx, (size_t)B * 4);
  if (e != hipSuccess) return hip_fail(N->ctx, e, "kmi_nn allocation");
  if ((e = hipMemset(N->m, 0, o * 4)) != hipSuccess || (e = hipMemset(N->v, 0, o * 4)) != hipSuccess)
    return hip_fail(N->ctx, e, "kmi_nn init");
  return KMI_OK;
}

// Splits for a product with a long reduction (K) and a small output, capped by P's size.
static int pick_splits(const kmi_nn* N, long long M, long long Nn, long long K) {
  long long tiles = ((M + 63) / 64) * ((Nn + 63) / 64);
  long long s = std::max<long long>(1, std::min<long long>((2048 + tiles - 1) / tiles, K / 256));
  s = std::min<long long>(s, (long long)N->P_elems / std::max<long long>(1, M * Nn));
  return int(std::max<long long>(1, s));
}

static int colsum(kmi_nn* N, const float* X, long long rows, int n, float* out, hipStream_t s) {
  const int chunk = int(std::max<long long>(256, (rows + 1023) / 1024));
  const int nchunk = int((rows + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_colsum_partial, dim3(n > 128 ? unsigned((n + 255) / 256) : 1u, unsigned(nchunk)),
                     dim3(256), 0, s, X, rows, n, chunk, N->colpart);
  hipLaunchKernelGGL(k_colsum_final, dim3(unsigned(n)), dim3(256), 0, s, N->colpart, nchunk, n, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? KMI_OK : hip_fail(N->ctx, e, "bias gradient");
}

#define NN_CHECK(expr)                                            \
  do {                                                            \
    hipError_t _e = (expr);                                       \
    if (_e != hipSuccess) return hip_fail(N->ctx, _e, #expr);     \
  } while (0)

// Forward pass for batch B from N->x (FP32) into N->y.
static int nn_forward(kmi_nn* N, int B) {
  const kmi_nn_desc& d = N->d;
  hipStream_t s = N->ctx->stream;
  const uint16_t* wb = N->wb;
  const float* w = N->w;
  hipLaunchKernelGGL(k_to_bf16, g1((long long)B * d.in_dim), dim3(256), 0, s, N->x, (long long)B * d.in_dim, N->xb);
  // dense 1
  NN_CHECK(gemm(B, d.hidden, d.in_dim, LoadDense{N->xb, d.in_dim, 1, B, d.in_dim},
                LoadDense{wb + N->oW1, d.hidden, 1, d.in_dim, d.hidden}, N->P, 1, s));
  long long mn = (long long)B * d.hidden;
  hipLaunchKernelGGL(k_epilogue, g1(mn), dim3(256), 0, s, N->P, 1, mn, d.hidden, w + N->ob1, 1, N->z1, (float*)nullptr, N->h1b);
  // dense 2 → latent (NHWC)
  NN_CHECK(gemm(B, N->L, d.hidden, LoadDense{N->h1b, d.hidden, 1, B, d.hidden},
                LoadDense{wb + N->oW2, N->L, 1, d.hidden, N->L}, N->P, 1, s));
  mn = (long long)B * N->L;
  hipLaunchKernelGGL(k_epilogue, g1(mn), dim3(256), 0, s, N->P, 1, mn, N->L, w + N->ob2, 1, N->z2, (float*)nullptr, N->h2b);
  // conv 1..3
  struct Stage { const uint16_t* in; int Hi, Ci, Co; long long ow, ob; float* z; uint16_t* hb; int silu; };
  const Stage st[3] = {{N->h2b, d.latent_hw, d.latent_c, d.c1, N->oC1, N->obc1, N->z3, N->h3b, 1},
                       {N->h3b, N->H1, d.c1, d.c2, N->oC2, N->obc2, N->z4, N->h4b, 1},
                       {N->h4b, N->H2, d.c2, d.out_c, N->oC3, N->obc3, N->y, nullptr, 0}};
  for (const Stage& S : st) {
    const int Ho = 2 * S.Hi;
    const int M = B * Ho * Ho;
    NN_CHECK(gemm(M, S.Co, 9 * S.Ci, LoadConvUp<false>{S.in, B, S.Hi, S.Ci},
                  LoadDense{wb + S.ow, S.Co, 1, 9 * S.Ci, S.Co}, N->P, 1, s));
    mn = (long long)M * S.Co;
    hipLaunchKernelGGL(k_epilogue, g1(mn), dim3(256), 0, s, N->P, 1, mn, S.Co, w + S.ob, S.silu, S.z,
                       (float*)nullptr, S.hb);
  }
  NN_CHECK(hipGetLastError());
  return KMI_OK;
}

// Backward pass from N->dy (FP32 loss gradient) into N->g (all parameter gradients).
static int nn_backward(kmi_nn* N, int B) {
  const kmi_nn_desc& d = N->d;
  hipStream_t s = N->ctx->stream;
  const uint16_t* wb = N->wb;
  float* g = N->g;
  struct Stage { const uint16_t* in; int Hi, Ci, Co; long long ow, ob; const float* zin; int silu_in; };
  // stage k: input activation `in` (BF16, NHWC Hi×Hi×Ci, produced with SiLU from zin)
  const Stage st[3] = {{N->h2b, d.latent_hw, d.latent_c, d.c1, N->oC1, N->obc1, N->z2, 1},
                       {N->h3b, N->H1, d.c1, d.c2, N->oC2, N->obc2, N->z3, 1},
                       {N->h4b, N->H2, d.c2, d.out_c, N->oC3, N->obc3, N->z4, 1}};
  // gradient w.r.t. the stage output pre-activation, FP32 + BF16
  const float* dz = N->dy;
  long long nout = out_elems(N, B);
  hipLaunchKernelGGL(k_to_bf16, g1(nout), dim3(256), 0, s, N->dy, nout, N->dyb);
  const uint16_t* dzb = N->dyb;
  for (int k = 2; k >= 0; --k) {
    const Stage& S = st[k];
    const int Ho = 2 * S.Hi, Mo = B * Ho * Ho;
    // bias gradient
    if (int rc = colsum(N, dz, Mo, S.Co, g + S.ob, s)) return rc;
    // weight gradient dW[(ky,kx,c), o] = Σ_m A(m, (ky,kx,c)) dZ[m, o]
    const int Mw = 9 * S.Ci;
    const int sp = pick_splits(N, Mw, S.Co, Mo);
    NN_CHECK(gemm(Mw, S.Co, Mo, LoadConvUp<true>{S.in, B, S.Hi, S.Ci}, LoadDense{dzb, S.Co, 1, Mo, S.Co}, N->P, sp, s));
    {
      int kchunk = (Mo + sp - 1) / sp;
      kchunk = (kchunk + 15) / 16 * 16;
      const int splits = (Mo + kchunk - 1) / kchunk;
      const long long mn = (long long)Mw * S.Co;
      hipLaunchKernelGGL(k_split_sum, dim3(unsigned((mn + 15) / 16)), dim3(256), 0, s, N->P, splits, mn, g + S.ow);
    }
    // data gradient into the stage input (per input pixel, up-sampling folded in)
    const int Mi = B * S.Hi * S.Hi;
    NN_CHECK(gemm(Mi, S.Ci, 36 * S.Co, LoadConvUpGradIn{dzb, B, S.Hi, S.Co},
                  LoadConvWeightT{wb + S.ow, S.Ci, S.Co}, N->P, 1, s));
    const long long ni = (long long)Mi * S.Ci;
    hipLaunchKernelGGL(k_epilogue, g1(ni), dim3(256), 0, s, N->P, 1, ni, S.Ci, (const float*)nullptr, 0, N->dh,
                       (float*)nullptr, (uint16_t*)nullptr);
    hipLaunchKernelGGL(k_act_grad, g1(ni), dim3(256), 0, s, N->dh, S.zin, ni, S.silu_in, N->dz, N->dzb);
    dz = N->dz;
    dzb = N->dzb;
  }
  // dense 2: dZ2 is in dz/dzb as [B, L]
  if (int rc = colsum(N, dz, B, N->L, g + N->ob2, s)) return rc;
  {
    const int sp = pick_splits(N, d.hidden, N->L, B);
    NN_CHECK(gemm(d.hidden, N->L, B, LoadDense{N->h1b, 1, d.hidden, d.hidden, B}, LoadDense{dzb, N->L, 1, B, N->L},
                  N->P, sp, s));
    int kchunk = ((B + sp - 1) / sp + 15) / 16 * 16;
    const int splits = (B + kchunk - 1) / kchunk;
    const long long mn = (long long)d.hidden * N->L;
    hipLaunchKernelGGL(k_split_sum, dim3(unsigned((mn + 15) / 16)), dim3(256), 0, s, N->P, splits, mn, g + N->oW2);
  }
  // dH1 = dZ2 W2ᵀ (long reduction over the latent: split-K)
  const long long nh = (long long)B * d.hidden;
  {
    const int sp = pick_splits(N, B, d.hidden, N->L);
    NN_CHECK(gemm(B, d.hidden, N->L, LoadDense{dzb, N->L, 1, B, N->L},
                  LoadDense{wb + N->oW2, 1, N->L, N->L, d.hidden}, N->P, sp, s));
    int kchunk = ((N->L + sp - 1) / sp + 15) / 16 * 16;
    const int splits = (N->L + kchunk - 1) / kchunk;
    hipLaunchKernelGGL(k_split_sum, dim3(unsigned((nh + 15) / 16)), dim3(256), 0, s, N->P, splits, nh, N->dh);
  }
  hipLaunchKernelGGL(k_act_grad, g1(nh), dim3(256), 0, s, N->dh, N->z1, nh, 1, N->dz, N->dzb);
  // dense 1
  if (int rc = colsum(N, N->dz, B, d.hidden, g + N->ob1, s)) return rc;
  NN_CHECK(gemm(d.in_dim, d.hidden, B, LoadDense{N->xb, 1, d.in_dim, d.in_dim, B}, LoadDense{N->dzb, d.hidden, 1, B, d.hidden},
                N->P, 1, s));
  const long long n1 = (long long)d.in_dim * d.hidden;
  hipLaunchKernelGGL(k_epilogue, g1(n1), dim3(256), 0, s, N->P, 1, n1, d.hidden, (const float*)nullptr, 0, g + N->oW1,
                     (float*)nullptr, (uint16_t*)nullptr);
  NN_CHECK(hipGetLastError());
  return KMI_OK;
}

static int nn_loss(kmi_nn* N, int B, double* loss) {
  hipStream_t s = N->ctx->stream;
  const long long n = out_elems(N, B);
  const int nb = int(std::min<long long>(4096, (n + 255) / 256));
  hipLaunchKernelGGL(k_mse, dim3(unsigned(nb)), dim3(256), 0, s, N->y, N->tgt, n, N->dy, N->lpart);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, s, N->lpart, nb, N->lout);
  NN_CHECK(hipGetLastError());
  if (loss) {
    NN_CHECK(hipMemcpyAsync(loss, N->lout, 8, hipMemcpyDeviceToHost, s));
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
  void* ptrs[] = {N->w, N->g, N->m, N->v, N->wb, N->x, N->xb, N->tgt, N->z1, N->h1b, N->z2, N->h2b, N->z3,
                  N->h3b, N->z4, N->h4b, N->y, N->dy, N->dyb, N->dh, N->dz, N->dzb, N->P, N->colpart,
                  N->lpart, N->lout, N->idx, N->dsX, N->dsY};
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
  if (N->dsX) (void)hipFree(N->dsX);
  if (N->dsY) (void)hipFree(N->dsY);
  N->dsX = N->dsY = nullptr;
  N->dsN = 0;
  const long long dout = out_elems(N, 1);
  KMI_HIP(c, hipMalloc(&N->dsX, (size_t)n * N->d.in_dim * 4));
  KMI_HIP(c, hipMalloc(&N->dsY, (size_t)n * dout * 4));
  KMI_HIP(c, hipMemcpy(N->dsX, x, (size_t)n * N->d.in_dim * 4, hipMemcpyHostToDevice));
  KMI_HIP(c, hipMemcpy(N->dsY, y, (size_t)n * dout * 4, hipMemcpyHostToDevice));
  N->dsN = n;
  return KMI_OK;
}

// Epochs of Adam over the device dataset in mini-batches of B (a Fisher–Yates shuffle per
// epoch from `seed`; the last partial batch is dropped).  Returns the mean loss of the
// final epoch.
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
  double last = 0.0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < N->dsN; ++i) perm[i] = i;
    for (int i = N->dsN - 1; i > 0; --i) std::swap(perm[i], perm[int(splitmix(s) % uint64_t(i + 1))]);
    const int nb = N->dsN / B;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) {
      KMI_HIP(c, hipMemcpyAsync(N->idx, perm.data() + (size_t)b * B, (size_t)B * 4, hipMemcpyHostToDevice, c->stream));
      hipLaunchKernelGGL(k_gather, g1((long long)B * dout), dim3(256), 0, c->stream, N->dsX, N->dsY, N->idx, B,
                         N->d.in_dim, dout, N->x, N->tgt);
      KMI_HIP(c, hipGetLastError());
      if (int rc = nn_forward(N, B)) return rc;
      double l = 0.0;
      if (int rc = nn_loss(N, B, ep == epochs - 1 ? &l : nullptr)) return rc;
      if (int rc = nn_backward(N, B)) return rc;
      if (int rc = nn_adam(N, lr)) return rc;
      acc += l;
      if (ep != epochs - 1) KMI_HIP(c, hipStreamSynchronize(c->stream));  // perm buffer reuse
    }
    last = acc / nb;
  }
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  if (mean_loss) *mean_loss = last;
  return KMI_OK;
}

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
  int sp;            // splits gemm() runs after rounding the K chunks to whole K-steps
};
static bool gemm_shape(int32_t M, int32_t Nn, int32_t K, int64_t sam, int64_t sak, int64_t sbk, int64_t sbn,
                       int32_t splits, GemmShape& g) {
  if (M < 1 || Nn < 1 || K < 1 || splits < 1 || sam < 0 || sak < 0 || sbk < 0 || sbn < 0) return false;
  g.na = (long long)(M - 1) * sam + (long long)(K - 1) * sak + 1;
  g.nb = (long long)(K - 1) * sbk + (long long)(Nn - 1) * sbn + 1;
  const int kchunk = ((K + splits - 1) / splits + 15) / 16 * 16;  // as gemm() rounds it
  g.splits = splits;
  g.sp = (K + kchunk - 1) / kchunk;
  return true;
}
static hipError_t gemm_bf16_run(hipStream_t s, int32_t M, int32_t Nn, int32_t K, const float* A, int64_t sam,
                                int64_t sak, const float* B, int64_t sbk, int64_t sbn, const GemmShape& g,
                                const GemmBufs& d, float* C) {
  hipError_t e = hipMemcpyAsync(d.dA, A, g.na * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d.dB, B, g.nb * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, g1(g.na), dim3(256), 0, s, d.dA, g.na, d.bA);
    hipLaunchKernelGGL(k_to_bf16, g1(g.nb), dim3(256), 0, s, d.dB, g.nb, d.bB);
    e = gemm(M, Nn, K, LoadDense{d.bA, sam, sak, M, K}, LoadDense{d.bB, sbk, sbn, K, Nn}, d.P, g.splits, s);
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

}  // extern "C"
