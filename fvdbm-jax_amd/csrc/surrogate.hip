// Kernels of the warm-start surrogate: a mesh message-passing network whose
// dense layers run on BF16 MFMA with FP32 accumulation, and the mesh
// aggregation that couples the nodes. No reference counterpart (the
// reference has no learned component); the C ABI is declared in
// include/fvdbm.h (fvdbm_nn_*).
//
// Building blocks (all row-major float32 in HBM; weights are float32 master
// copies, rounded to bf16 when staged in LDS):
//   mlp2   y = silu(x W1 + b1) W2 + b2, x [M,K], W1 [K,64], W2 [64,NO]
//   mlp2 backward (data part): recomputes z = x W1 + b1, then
//          dz = (dy W2^T) * silu'(z), dx = dz W1^T; writes dz and h = silu(z)
//   wgrad  out = A^T B over the rows (A [M,KA], B [M,NB]) as per-chunk
//          partial products plus a fixed-order reduction (deterministic)
//   colsum out = sum over rows, same chunking
//   aggregate: z = [h | sum_e w h_j | sum_e w ex h_j | sum_e w ey h_j]
//          over each node's CSR edge list, and its transpose for dh.
//
// MFMA use: fragments follow the 16x16x16 bf16 layout. For one 64-lane
// wave with i = lane % 16 and g = lane / 16:
//   A (16x16, MxK): lane holds A[i][4g .. 4g+3]
//   B (16x16, KxN): lane holds B[4g .. 4g+3][i]
//   D (16x16, MxN): lane holds D[4g + r][i], r = 0..3
// K is consumed 32 at a time by gfx950's 16x16x32 bf16 MFMA (mfma2: two
// such fragments per operand), 16 at a time only for a 16-wide input.
// The dense layers are computed transposed (hidden units as MFMA rows, data
// rows as MFMA columns): then the D fragment of one layer is, lane by lane,
// exactly the B fragment the next layer needs, so the hidden activations
// never leave registers.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "fvdbm.h"

#pragma clang fp contract(fast)

namespace fvdbm_nn {

constexpr int HID = 64;
constexpr int CHUNK = 1024;  // rows per partial product in wgrad / colsum
constexpr int RGROUP = 32;   // partials summed per thread in the first reduction pass

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two consecutive K = 16 steps, c + A0 B0 + A1 B1, as one gfx950
// 16x16x32 MFMA. Each lane gives that instruction eight k values of A
// and the same eight k positions of B: here its four of (a0, b0) and its
// four of (a1, b1). A and B use one assignment of the 32 k indices to
// (lane group, slot), so the product sums exactly the 32 products of the
// two 16-blocks, in the instruction's own order (FP32 accumulation).
__device__ __forceinline__ f4 mfma2(s4 a0, s4 a1, s4 b0, s4 b1, f4 c) {
  const s8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
  const s8 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c,
                                                  0, 0, 0);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ s4 lds4(const unsigned short *p) {
  return s4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
}

__device__ __forceinline__ s4 bf4(float4 v) {
  return s4{(short)f2bf(v.x), (short)f2bf(v.y), (short)f2bf(v.z), (short)f2bf(v.w)};
}

// ---------------------------------------------------------------------------
// mlp2 forward: one wave per 16-row tile, weights for the whole block in LDS
// ---------------------------------------------------------------------------
template <int K, int NO>
__global__ __launch_bounds__(256) void k_mlp_fwd(const float *__restrict__ x, int M,
                                                 const float *__restrict__ w1,
                                                 const float *__restrict__ b1,
                                                 const float *__restrict__ w2,
                                                 const float *__restrict__ b2,
                                                 float *__restrict__ y) {
  constexpr int KP = K + 4, HP = HID + 4;
  __shared__ unsigned short sW1T[HID * KP];  // [hidden][k]
  __shared__ unsigned short sW2T[NO * HP];   // [out][hidden]
  __shared__ float sb1[HID], sb2[NO];
  for (int e = threadIdx.x; e < K * HID; e += blockDim.x)
    sW1T[(e % HID) * KP + e / HID] = f2bf(w1[e]);
  for (int e = threadIdx.x; e < HID * NO; e += blockDim.x)
    sW2T[(e % NO) * HP + e / NO] = f2bf(w2[e]);
  if (threadIdx.x < HID) sb1[threadIdx.x] = b1[threadIdx.x];
  if (threadIdx.x < NO) sb2[threadIdx.x] = b2[threadIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int ntiles = (M + 15) / 16;
  for (int t = blockIdx.x * 4 + wave; t < ntiles; t += gridDim.x * 4) {
    const int row = t * 16 + i;
    const bool ok = row < M;
    f4 z[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) z[hb] = f4{0.f, 0.f, 0.f, 0.f};
    // the B fragment of x's 16-column block ks: x[row][16 ks + 4g .. +3]
    auto xfrag = [&](int ks) {
      const float4 xv = ok ? *reinterpret_cast<const float4 *>(x + (size_t)row * K + ks * 16 + 4 * g)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
      return bf4(xv);
    };
    if constexpr (K % 32 == 0) {
#pragma unroll
      for (int kp = 0; kp < K / 32; ++kp) {
        const s4 b0 = xfrag(2 * kp), b1 = xfrag(2 * kp + 1);
#pragma unroll
        for (int hb = 0; hb < 4; ++hb) {
          const unsigned short *w = sW1T + (hb * 16 + i) * KP + kp * 32 + 4 * g;
          z[hb] = mfma2(lds4(w), lds4(w + 16), b0, b1, z[hb]);
        }
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < K / 16; ++ks) {
        const s4 b = xfrag(ks);
#pragma unroll
        for (int hb = 0; hb < 4; ++hb)
          z[hb] = mfma(lds4(sW1T + (hb * 16 + i) * KP + ks * 16 + 4 * g), b, z[hb]);
      }
    }
    s4 hfrag[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      float h4[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) h4[r] = silu(z[hb][r] + sb1[hb * 16 + 4 * g + r]);
      hfrag[hb] = bf4(make_float4(h4[0], h4[1], h4[2], h4[3]));
    }
#pragma unroll
    for (int ob = 0; ob < NO / 16; ++ob) {
      f4 acc = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kp = 0; kp < 2; ++kp) {
        const unsigned short *w = sW2T + (ob * 16 + i) * HP + kp * 32 + 4 * g;
        acc = mfma2(lds4(w), lds4(w + 16), hfrag[2 * kp], hfrag[2 * kp + 1], acc);
      }
      if (ok) {
        const int o = ob * 16 + 4 * g;
        *reinterpret_cast<float4 *>(y + (size_t)row * NO + o) =
            make_float4(acc[0] + sb2[o], acc[1] + sb2[o + 1], acc[2] + sb2[o + 2], acc[3] + sb2[o + 3]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// mlp2 backward, data part: dz [M,64], h [M,64] and (optionally) dx [M,K]
// ---------------------------------------------------------------------------
template <int K, int NO>
__global__ __launch_bounds__(256) void k_mlp_bwd(const float *__restrict__ x, int M,
                                                 const float *__restrict__ w1,
                                                 const float *__restrict__ b1,
                                                 const float *__restrict__ w2,
                                                 const float *__restrict__ dy,
                                                 float *__restrict__ dx, float *__restrict__ dz,
                                                 float *__restrict__ hout) {
  constexpr int HP = HID + 4, NP = NO + 4;
  __shared__ unsigned short sW1[K * HP];    // [k][hidden]: z (strided reads) and dx
  __shared__ unsigned short sW2[HID * NP];  // [hidden][out]: dh = dy W2^T
  __shared__ float sb1[HID];
  for (int e = threadIdx.x; e < K * HID; e += blockDim.x) sW1[(e / HID) * HP + e % HID] = f2bf(w1[e]);
  for (int e = threadIdx.x; e < HID * NO; e += blockDim.x) sW2[(e / NO) * NP + e % NO] = f2bf(w2[e]);
  if (threadIdx.x < HID) sb1[threadIdx.x] = b1[threadIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int ntiles = (M + 15) / 16;
  for (int t = blockIdx.x * 4 + wave; t < ntiles; t += gridDim.x * 4) {
    const int row = t * 16 + i;
    const bool ok = row < M;
    f4 z[4], dh[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      z[hb] = f4{0.f, 0.f, 0.f, 0.f};
      dh[hb] = f4{0.f, 0.f, 0.f, 0.f};
    }
    // B fragments of x's and dy's 16-column block ks, and the A fragment
    // W1^T[hidden block hb][k block ks] read down sW1's columns
    auto xfrag = [&](int ks) {
      const float4 xv = ok ? *reinterpret_cast<const float4 *>(x + (size_t)row * K + ks * 16 + 4 * g)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
      return bf4(xv);
    };
    auto dyfrag = [&](int os) {
      const float4 dv = ok ? *reinterpret_cast<const float4 *>(dy + (size_t)row * NO + os * 16 + 4 * g)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
      return bf4(dv);
    };
    auto w1t = [&](int ks, int hb) {
      const unsigned short *p = sW1 + (ks * 16 + 4 * g) * HP + hb * 16 + i;
      return s4{(short)p[0], (short)p[HP], (short)p[2 * HP], (short)p[3 * HP]};
    };
    if constexpr (K % 32 == 0) {
#pragma unroll
      for (int kp = 0; kp < K / 32; ++kp) {
        const s4 b0 = xfrag(2 * kp), b1 = xfrag(2 * kp + 1);
#pragma unroll
        for (int hb = 0; hb < 4; ++hb) z[hb] = mfma2(w1t(2 * kp, hb), w1t(2 * kp + 1, hb), b0, b1, z[hb]);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < K / 16; ++ks) {
        const s4 b = xfrag(ks);
#pragma unroll
        for (int hb = 0; hb < 4; ++hb) z[hb] = mfma(w1t(ks, hb), b, z[hb]);
      }
    }
    // dh^T [hidden][rows] = W2 [hidden][out] . dy^T [out][rows]
    if constexpr (NO % 32 == 0) {
#pragma unroll
      for (int op = 0; op < NO / 32; ++op) {
        const s4 b0 = dyfrag(2 * op), b1 = dyfrag(2 * op + 1);
#pragma unroll
        for (int hb = 0; hb < 4; ++hb) {
          const unsigned short *w = sW2 + (hb * 16 + i) * NP + op * 32 + 4 * g;
          dh[hb] = mfma2(lds4(w), lds4(w + 16), b0, b1, dh[hb]);
        }
      }
    } else {
#pragma unroll
      for (int os = 0; os < NO / 16; ++os) {
        const s4 b = dyfrag(os);
#pragma unroll
        for (int hb = 0; hb < 4; ++hb)
          dh[hb] = mfma(lds4(sW2 + (hb * 16 + i) * NP + os * 16 + 4 * g), b, dh[hb]);
      }
    }
    s4 dzfrag[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      float d4[4], h4[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float zz = z[hb][r] + sb1[hb * 16 + 4 * g + r];
        h4[r] = silu(zz);
        d4[r] = dh[hb][r] * dsilu(zz);
      }
      if (ok) {
        const size_t o = (size_t)row * HID + hb * 16 + 4 * g;
        *reinterpret_cast<float4 *>(dz + o) = make_float4(d4[0], d4[1], d4[2], d4[3]);
        *reinterpret_cast<float4 *>(hout + o) = make_float4(h4[0], h4[1], h4[2], h4[3]);
      }
      dzfrag[hb] = bf4(make_float4(d4[0], d4[1], d4[2], d4[3]));
    }
    if (dx) {
      // dx^T [k][rows] = W1 [k][hidden] . dz^T [hidden][rows]
#pragma unroll
      for (int kb = 0; kb < K / 16; ++kb) {
        f4 acc = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kp = 0; kp < 2; ++kp) {
          const unsigned short *w = sW1 + (kb * 16 + i) * HP + kp * 32 + 4 * g;
          acc = mfma2(lds4(w), lds4(w + 16), dzfrag[2 * kp], dzfrag[2 * kp + 1], acc);
        }
        if (ok)
          *reinterpret_cast<float4 *>(dx + (size_t)row * K + kb * 16 + 4 * g) =
              make_float4(acc[0], acc[1], acc[2], acc[3]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// wgrad: partial[c] = A[rows of chunk c]^T B[rows of chunk c]  ([KA][NB])
// One block per chunk of CHUNK rows, staged 64 rows at a time in LDS as
// bf16, transposed (feature-major, [feature][row]): an MFMA operand
// fragment (one feature, four consecutive rows) is then one 8-byte LDS
// read. The output 16x16 blocks are dealt round-robin to the four waves;
// each wave's accumulators live in registers across the chunk.
// ---------------------------------------------------------------------------
// Stage rows [base, base + R) of X [rows][K] (rows from r1 on read as 0)
// into s as bf16, feature-major ([K][RS]). Work item e takes the 4x4 block
// of features 4c .. 4c+3 and rows 4q .. 4q+3 (q = e % (R/4) fastest, so
// neighbouring lanes write neighbouring 8-byte words of one feature row):
// four 16-byte loads, one per row, and four 8-byte LDS stores, one per
// feature. X rows must be 16-byte aligned (K a multiple of 4).
template <int K, int R, int RS>
__device__ __forceinline__ void stage_transposed(const float *__restrict__ X, int base, int r1,
                                                 unsigned short *__restrict__ s) {
  static_assert(K % 4 == 0 && R % 4 == 0, "4x4 blocks");
  for (int e = threadIdx.x; e < (K / 4) * (R / 4); e += blockDim.x) {
    const int q = e % (R / 4), c = e / (R / 4);
    float4 v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = base + 4 * q + j;
      v[j] = row < r1 ? *reinterpret_cast<const float4 *>(X + (size_t)row * K + 4 * c)
                      : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const float col[4][4] = {{v[0].x, v[1].x, v[2].x, v[3].x},
                             {v[0].y, v[1].y, v[2].y, v[3].y},
                             {v[0].z, v[1].z, v[2].z, v[3].z},
                             {v[0].w, v[1].w, v[2].w, v[3].w}};
#pragma unroll
    for (int f = 0; f < 4; ++f)
      *reinterpret_cast<uint2 *>(s + (4 * c + f) * RS + 4 * q) =
          make_uint2((uint32_t)f2bf(col[f][0]) | ((uint32_t)f2bf(col[f][1]) << 16),
                     (uint32_t)f2bf(col[f][2]) | ((uint32_t)f2bf(col[f][3]) << 16));
  }
}

template <int KA, int NB>
__global__ __launch_bounds__(256) void k_wgrad(const float *__restrict__ A,
                                               const float *__restrict__ B, int M,
                                               float *__restrict__ partial) {
  constexpr int R = 64;          // rows per LDS stage
  constexpr int RP = R + 4;      // row stride of a feature, in bf16 (8-byte aligned, odd in 8-byte units)
  constexpr int NBLK = (KA / 16) * (NB / 16);
  constexpr int PER = (NBLK + 3) / 4;
  __shared__ __align__(16) unsigned short sA[KA * RP];
  __shared__ __align__(16) unsigned short sB[NB * RP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int r0 = blockIdx.x * CHUNK, r1 = min(M, r0 + CHUNK);
  f4 acc[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) acc[p] = f4{0.f, 0.f, 0.f, 0.f};
  for (int base = r0; base < r1; base += R) {
    __syncthreads();
    stage_transposed<KA, R, RP>(A, base, r1, sA);
    stage_transposed<NB, R, RP>(B, base, r1, sB);
    __syncthreads();
#pragma unroll
    for (int mp = 0; mp < R / 32; ++mp) {
      const int m = mp * 32 + 4 * g;  // this lane's four rows of each 16-row slab of the pair
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        const int blk = wave + 4 * p;
        if (blk < NBLK) {
          const int kb = blk / (NB / 16), nb = blk % (NB / 16);
          // A fragments: feature kb*16+i, rows m..m+3 and m+16..m+19; B: feature nb*16+i, same rows
          const unsigned short *pa = sA + (kb * 16 + i) * RP + m;
          const unsigned short *pb = sB + (nb * 16 + i) * RP + m;
          acc[p] = mfma2(*reinterpret_cast<const s4 *>(pa), *reinterpret_cast<const s4 *>(pa + 16),
                         *reinterpret_cast<const s4 *>(pb), *reinterpret_cast<const s4 *>(pb + 16), acc[p]);
        }
      }
    }
  }
  float *out = partial + (size_t)blockIdx.x * KA * NB;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int blk = wave + 4 * p;
    if (blk < NBLK) {
      const int kb = blk / (NB / 16), nb = blk % (NB / 16);
#pragma unroll
      for (int r = 0; r < 4; ++r) out[(size_t)(kb * 16 + 4 * g + r) * NB + nb * 16 + i] = acc[p][r];
    }
  }
}

// partial[c][n] = sum of column n over the rows of chunk c: 256 / N row
// groups per column each sum their rows in order, then the groups are added
// in group order (fixed order, deterministic)
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ A, int M, int N,
                                                float *__restrict__ partial) {
  __shared__ float red[256];
  const int groups = 256 / N;
  const int n = threadIdx.x % N, grp = threadIdx.x / N;
  const int r0 = blockIdx.x * CHUNK, r1 = min(M, r0 + CHUNK);
  float s = 0.f;
  if (grp < groups)
    for (int r = r0 + grp; r < r1; r += groups) s += A[(size_t)r * N + n];
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < N) {
    float t = 0.f;
    for (int g = 0; g < groups; ++g) t += red[g * N + threadIdx.x];
    partial[(size_t)blockIdx.x * N + threadIdx.x] = t;
  }
}

__global__ void k_zero(float *__restrict__ p, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = 0.f;
}

// First pass of the chunk reduction, in place: for group q (blockIdx.y),
// partial[q * RGROUP][e] <- the sum, in chunk order, of partial[c][e] over
// the group's chunks c = q*RGROUP .. Each thread reads only its group's
// chunks and writes its group's first slot, so the pass is race-free.
__global__ void k_reduce_groups(float *__restrict__ partial, int nchunks, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int c0 = blockIdx.y * RGROUP, c1 = min(nchunks, c0 + RGROUP);
  float s = 0.f;
  for (int c = c0; c < c1; ++c) s += partial[(size_t)c * n + e];
  partial[(size_t)c0 * n + e] = s;
}

// Second pass: out[e] = the sum, in group order, of the group sums.
__global__ void k_reduce_final(const float *__restrict__ partial, int nchunks, int n,
                               float *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.f;
  for (int c = 0; c < nchunks; c += RGROUP) s += partial[(size_t)c * n + e];
  out[e] = s;
}

// ---------------------------------------------------------------------------
// aggregation: one wave per node, lane = feature
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_aggregate(const float *__restrict__ h, int n,
                                                   const int32_t *__restrict__ off,
                                                   const int32_t *__restrict__ col,
                                                   const float *__restrict__ w,
                                                   const float *__restrict__ ex,
                                                   const float *__restrict__ ey,
                                                   float *__restrict__ z) {
  const int node = blockIdx.x * 4 + (threadIdx.x >> 6), f = threadIdx.x & 63;
  if (node >= n) return;
  float a0 = 0.f, ax = 0.f, ay = 0.f;
  for (int e = off[node]; e < off[node + 1]; ++e) {
    const float v = h[(size_t)col[e] * HID + f] * w[e];
    a0 += v;
    ax += v * ex[e];
    ay += v * ey[e];
  }
  float *zr = z + (size_t)node * 4 * HID;
  zr[f] = h[(size_t)node * HID + f];
  zr[HID + f] = a0;
  zr[2 * HID + f] = ax;
  zr[3 * HID + f] = ay;
}

// dh[i] = dz[i][0:64] + sum over i's edges e = (i, j) of the transpose
// weights (wt = weight of edge j -> i, (ext, eyt) its direction) times
// dz[j][64:256]
__global__ __launch_bounds__(256) void k_aggregate_bwd(const float *__restrict__ dz, int n,
                                                       const int32_t *__restrict__ off,
                                                       const int32_t *__restrict__ col,
                                                       const float *__restrict__ wt,
                                                       const float *__restrict__ ext,
                                                       const float *__restrict__ eyt,
                                                       float *__restrict__ dh) {
  const int node = blockIdx.x * 4 + (threadIdx.x >> 6), f = threadIdx.x & 63;
  if (node >= n) return;
  float s = dz[(size_t)node * 4 * HID + f];
  for (int e = off[node]; e < off[node + 1]; ++e) {
    const float *dj = dz + (size_t)col[e] * 4 * HID;
    s += wt[e] * (dj[HID + f] + ext[e] * dj[2 * HID + f] + eyt[e] * dj[3 * HID + f]);
  }
  dh[(size_t)node * HID + f] = s;
}

static thread_local std::string g_err;

static int nn_fail(const std::string &m) {
  g_err = m;
  return FVDBM_ERR_INVALID;
}

static int hip_status(const char *what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) return FVDBM_OK;
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return FVDBM_ERR_HIP;
}

static int grid_for(int64_t m) {
  const int64_t tiles = (m + 15) / 16;
  const int64_t blocks = (tiles + 3) / 4;
  return (int)(blocks < 2048 ? (blocks > 0 ? blocks : 1) : 2048);
}

template <int K, int NO>
static void launch_fwd(const float *x, int m, const float *w1, const float *b1, const float *w2,
                       const float *b2, float *y, hipStream_t s) {
  hipLaunchKernelGGL((k_mlp_fwd<K, NO>), dim3(grid_for(m)), dim3(256), 0, s, x, m, w1, b1, w2, b2, y);
}

template <int K, int NO>
static void launch_bwd(const float *x, int m, const float *w1, const float *b1, const float *w2,
                       const float *dy, float *dx, float *dz, float *h, hipStream_t s) {
  hipLaunchKernelGGL((k_mlp_bwd<K, NO>), dim3(grid_for(m)), dim3(256), 0, s, x, m, w1, b1, w2, dy,
                     dx, dz, h);
}

template <int KA, int NB>
static void launch_wgrad(const float *a, const float *b, int m, float *partial, hipStream_t s) {
  const int nch = (m + CHUNK - 1) / CHUNK;
  hipLaunchKernelGGL((k_wgrad<KA, NB>), dim3(nch), dim3(256), 0, s, a, b, m, partial);
}

// out [n] = sum over the nch chunk partials [nch][n] in a fixed order (in
// groups of RGROUP chunks, then the groups); overwrites the partials
static void launch_reduce(float *partial, int nch, int n, float *out, hipStream_t s) {
  const int bx = (n + 255) / 256;
  hipLaunchKernelGGL(k_reduce_groups, dim3(bx, (nch + RGROUP - 1) / RGROUP), dim3(256), 0, s,
                     partial, nch, n);
  hipLaunchKernelGGL(k_reduce_final, dim3(bx), dim3(256), 0, s, partial, nch, n, out);
}

// ---------------------------------------------------------------------------
// Allocation-free weight gradient (fvdbm_nn_wgrad_noalloc): the per-chunk
// partial products live in this file-scope device array instead of a
// caller workspace. Each GPU holds its own copy for the life of the
// process. Its host-side name is not a device address: the address on the
// current device comes from hipGetSymbolAddress, looked up once per device
// and kept in scratch_dev. k_wgrad writes every word of the partials it
// reduces, so nothing needs clearing between calls; calls that overlap on
// one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchFloats = size_t(4) << 20;  // 16 MiB
static __device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static float *scratch_dev[kMaxDevices] = {};

// The scratch array's address on the current device (looked up on a
// device's first call).
static int scratch_on_current_device(float **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    g_err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return FVDBM_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) return nn_fail("device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return FVDBM_ERR_HIP;
    }
    scratch_dev[dev] = static_cast<float *>(p);
  }
  *out = scratch_dev[dev];
  return FVDBM_OK;
}

}  // namespace fvdbm_nn

using namespace fvdbm_nn;

extern "C" {

const char *fvdbm_nn_last_error(void) { return g_err.c_str(); }

int fvdbm_nn_chunk_rows(void) { return CHUNK; }

// the dense kernels move four floats per lane: their row arrays must be
// 16-byte aligned
static bool aligned16(const void *p) { return p == nullptr || ((uintptr_t)p & 15u) == 0; }

#define NN_DISPATCH_K_NO(K_, NO_, CALL) \
  if (k == K_ && n_out == NO_) {        \
    CALL(K_, NO_);                      \
    return hip_status("mlp launch");    \
  }

int fvdbm_nn_mlp_forward(const float *x, int64_t m, int32_t k, const float *w1, const float *b1,
                         const float *w2, const float *b2, int32_t n_out, float *y, void *stream) {
  if (m < 0 || m > INT32_MAX) return nn_fail("row count out of range");
  if (m == 0) return FVDBM_OK;
  if (!x || !w1 || !b1 || !w2 || !b2 || !y) return nn_fail("null pointer");
  if (!aligned16(x) || !aligned16(y)) return nn_fail("x and y must be 16-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int mi = (int)m;
#define CALL_FWD(K_, NO_) launch_fwd<K_, NO_>(x, mi, w1, b1, w2, b2, y, s)
  NN_DISPATCH_K_NO(16, 64, CALL_FWD)
  NN_DISPATCH_K_NO(64, 64, CALL_FWD)
  NN_DISPATCH_K_NO(256, 64, CALL_FWD)
  NN_DISPATCH_K_NO(64, 16, CALL_FWD)
#undef CALL_FWD
  return nn_fail("unsupported (in, out) widths: (16|64|256, 64) or (64, 16)");
}

int fvdbm_nn_mlp_backward(const float *x, int64_t m, int32_t k, const float *w1, const float *b1,
                          const float *w2, int32_t n_out, const float *dy, float *dx, float *dz,
                          float *h, void *stream) {
  if (m < 0 || m > INT32_MAX) return nn_fail("row count out of range");
  if (m == 0) return FVDBM_OK;
  if (!x || !w1 || !b1 || !w2 || !dy || !dz || !h) return nn_fail("null pointer");
  if (!aligned16(x) || !aligned16(dy) || !aligned16(dx) || !aligned16(dz) || !aligned16(h))
    return nn_fail("x, dy, dx, dz and h must be 16-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int mi = (int)m;
#define CALL_BWD(K_, NO_) launch_bwd<K_, NO_>(x, mi, w1, b1, w2, dy, dx, dz, h, s)
  NN_DISPATCH_K_NO(16, 64, CALL_BWD)
  NN_DISPATCH_K_NO(64, 64, CALL_BWD)
  NN_DISPATCH_K_NO(256, 64, CALL_BWD)
  NN_DISPATCH_K_NO(64, 16, CALL_BWD)
#undef CALL_BWD
  return nn_fail("unsupported (in, out) widths: (16|64|256, 64) or (64, 16)");
}

int fvdbm_nn_wgrad(const float *a, const float *b, int64_t m, int32_t ka, int32_t nb, float *partial,
                   float *out, void *stream) {
  if (m < 0 || m > INT32_MAX) return nn_fail("row count out of range");
  if (!a || !b || !partial || !out) return nn_fail("null pointer");
  if (!aligned16(a) || !aligned16(b)) return nn_fail("a and b must be 16-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int mi = (int)m;
  const int nch = (mi + CHUNK - 1) / CHUNK;
  if (mi == 0) {
    hipLaunchKernelGGL(k_zero, dim3((ka * nb + 255) / 256), dim3(256), 0, s, out, ka * nb);
    return hip_status("zero launch");
  }
  if (ka == 16 && nb == 64) launch_wgrad<16, 64>(a, b, mi, partial, s);
  else if (ka == 64 && nb == 64) launch_wgrad<64, 64>(a, b, mi, partial, s);
  else if (ka == 256 && nb == 64) launch_wgrad<256, 64>(a, b, mi, partial, s);
  else if (ka == 64 && nb == 16) launch_wgrad<64, 16>(a, b, mi, partial, s);
  else return nn_fail("unsupported wgrad widths");
  launch_reduce(partial, nch, ka * nb, out, s);
  return hip_status("wgrad launch");
}

int fvdbm_nn_wgrad_noalloc(const float *a, const float *b, int64_t m, int32_t ka, int32_t nb,
                           float *out, void *stream) {
  if (m < 0 || m > INT32_MAX) return nn_fail("row count out of range");
  if (!a || !b || !out) return nn_fail("null pointer");
  if (!aligned16(a) || !aligned16(b)) return nn_fail("a and b must be 16-byte aligned");
  if (!((ka == 16 && nb == 64) || (ka == 64 && nb == 64) || (ka == 256 && nb == 64) ||
        (ka == 64 && nb == 16)))
    return nn_fail("unsupported wgrad widths");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int mi = (int)m;
  const int n = ka * nb;
  const int64_t nch = (m + CHUNK - 1) / CHUNK;
  if ((uint64_t)nch * (uint64_t)n > kScratchFloats)
    return nn_fail("fvdbm_nn_wgrad_noalloc: " + std::to_string(m) + " rows need " +
                   std::to_string(nch * n) + " partial floats, more than the " +
                   std::to_string(kScratchFloats) + " of the static scratch (at most " +
                   std::to_string((int64_t)(kScratchFloats / n) * CHUNK) + " rows for widths (" +
                   std::to_string(ka) + ", " + std::to_string(nb) + "))");
  if (mi == 0) {
    hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, s, out, n);
    return hip_status("zero launch");
  }
  float *partial = nullptr;
  const int rc = scratch_on_current_device(&partial);
  if (rc) return rc;
  if (ka == 16) launch_wgrad<16, 64>(a, b, mi, partial, s);
  else if (ka == 256) launch_wgrad<256, 64>(a, b, mi, partial, s);
  else if (nb == 64) launch_wgrad<64, 64>(a, b, mi, partial, s);
  else launch_wgrad<64, 16>(a, b, mi, partial, s);
  launch_reduce(partial, (int)nch, n, out, s);
  return hip_status("wgrad launch");
}

int fvdbm_nn_colsum(const float *a, int64_t m, int32_t n, float *partial, float *out, void *stream) {
  if (m < 0 || m > INT32_MAX) return nn_fail("row count out of range");
  if (!a || !partial || !out) return nn_fail("null pointer");
  if (n < 1 || n > 256) return nn_fail("column count must be in [1, 256]");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int mi = (int)m;
  const int nch = (mi + CHUNK - 1) / CHUNK;
  if (mi == 0) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, s, out, n);
    return hip_status("zero launch");
  }
  hipLaunchKernelGGL(k_colsum, dim3(nch), dim3(256), 0, s, a, mi, n, partial);
  launch_reduce(partial, nch, n, out, s);
  return hip_status("colsum launch");
}

int fvdbm_nn_aggregate(const float *h, int64_t n_nodes, const int32_t *row_off, const int32_t *col,
                       const float *w, const float *ex, const float *ey, float *z, void *stream) {
  if (n_nodes < 0 || n_nodes > INT32_MAX) return nn_fail("node count out of range");
  if (n_nodes == 0) return FVDBM_OK;
  if (!h || !row_off || !col || !w || !ex || !ey || !z) return nn_fail("null pointer");
  const int n = (int)n_nodes;
  hipLaunchKernelGGL(k_aggregate, dim3((n + 3) / 4), dim3(256), 0, static_cast<hipStream_t>(stream), h,
                     n, row_off, col, w, ex, ey, z);
  return hip_status("aggregate launch");
}

int fvdbm_nn_aggregate_backward(const float *dz, int64_t n_nodes, const int32_t *row_off,
                                const int32_t *col, const float *wt, const float *ext,
                                const float *eyt, float *dh, void *stream) {
  if (n_nodes < 0 || n_nodes > INT32_MAX) return nn_fail("node count out of range");
  if (n_nodes == 0) return FVDBM_OK;
  if (!dz || !row_off || !col || !wt || !ext || !eyt || !dh) return nn_fail("null pointer");
  const int n = (int)n_nodes;
  hipLaunchKernelGGL(k_aggregate_bwd, dim3((n + 3) / 4), dim3(256), 0,
                     static_cast<hipStream_t>(stream), dz, n, row_off, col, wt, ext, eyt, dh);
  return hip_status("aggregate backward launch");
}

}  // extern "C"
