// FVDBM D2Q9 time step for MI355X (gfx950): kernels, device state, C ABI.
//
// One step (reference: Environment.step, src/environment.py:55-65) is one
// fused cell-centric kernel, k_step. Each thread owns one cell:
//   1. loads its 9 pdfs (q-major SoA, coalesced),
//   2. computes rho, u (Dynamics.calc_macro, src/dynamics.py:35-47) and the
//      BGK equilibrium (D2Q9.calc_eq, src/dynamics.py:70-74) in registers,
//   3. for each of its 3 faces in cells.face_indices order gathers the
//      neighbour's 9 pdfs, or builds the ghost pdf from the two face-node
//      pdfs (Faces.calc_ghost, src/containers.py:280-287); velocity/density
//      node pdfs are computed inline from their cells (Nodes.calc_pdfs,
//      src/containers.py:339-390), so no separate node pass or [F,9] flux
//      array ever reaches HBM,
//   4. forms the upwind flux by the sign of ksi.n (calc_upwind_flux,
//      src/containers.py:215-240) or the Lax-Wendroff flux (:242-278), and
//   5. applies f + dt*(1/tau*(feq - f) - sum_k s_k F_k) (src/containers.py:
//      107-121) and stores to the other buffer (double buffering: the
//      kernel never reads what it writes, so it is race-free).
// Arithmetic keeps the reference's operation order with no contraction
// (built with -ffp-contract=off), so results are bitwise reproducible and
// equal to the CPU restatement of the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "errors.h"
#include "fvdbm.h"
#include "layout.h"

namespace fvdbm {

// ----------------------------------------------------------------------------
// errors
// ----------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

int set_error(int code, const std::string &msg) { return fail(code, msg); }

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(FVDBM_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// ----------------------------------------------------------------------------
// lattice constants (src/dynamics.py:54-64)
// ----------------------------------------------------------------------------
constexpr int Q = 9;
__host__ __device__ constexpr int kx(int q) {
  return q == 1 || q == 5 || q == 8 ? 1 : (q == 3 || q == 6 || q == 7 ? -1 : 0);
}
__host__ __device__ constexpr int ky(int q) {
  return q == 2 || q == 5 || q == 6 ? 1 : (q == 4 || q == 7 || q == 8 ? -1 : 0);
}

template <typename T>
struct Lattice {
  T W[Q];
  T C2, C4x2, C2x2;     // C*C, 2*(C*C)*(C*C), 2*C*C with C = 1/sqrt(3)
  T rC2, rC4x2, rC2x2;  // their rounded reciprocals (see div_const)
  T invtau, dt;         // 1/tau rounded from double, delta_t rounded
};

template <typename T>
static Lattice<T> make_lattice(double tau, double dt) {
  Lattice<T> k;
  const double w[Q] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                       1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
  for (int q = 0; q < Q; ++q) k.W[q] = T(w[q]);
  T C = T(1) / std::sqrt(T(3));
  k.C2 = C * C;
  T c4 = k.C2 * k.C2;
  k.C4x2 = T(2) * c4;
  k.C2x2 = T(2) * k.C2;
  k.rC2 = T(1) / k.C2;
  k.rC4x2 = T(1) / k.C4x2;
  k.rC2x2 = T(1) / k.C2x2;
  k.invtau = T(1.0 / tau);
  k.dt = T(dt);
  return k;
}

// ----------------------------------------------------------------------------
// device math: one function per reference formula, same operation order
// ----------------------------------------------------------------------------

// rho = sum_q f_q ; u = (sum_q ksi_q f_q) / rho, sums 0-initialised in q order.
// The terms with ksi component 0 are left out: 0 * f is a zero of either
// sign, and an accumulator that starts at +0 is never -0 (x + y is -0 only
// for two -0 operands), so adding such a zero changes nothing for finite f.
// ksi components of +-1 give +-f exactly.
template <typename T>
__device__ __forceinline__ void macro(const T (&f)[Q], T &rho, T &ux, T &uy) {
  T r = T(0);
#pragma unroll
  for (int q = 0; q < Q; ++q) r += f[q];
  T mx = T(0), my = T(0);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (kx(q) != 0) mx += T(kx(q)) * f[q];
    if (ky(q) != 0) my += T(ky(q)) * f[q];
  }
  rho = r;
  ux = mx / r;
  uy = my / r;
}

// feq_q = (W_q rho) * (((1 + A_q) + B_q) - D) with A = (ksi.u)/C^2,
// B = (ksi.u)^2/(2C^4), D = (u.u)/(2C^2). ksi.u of opposite directions are
// exact negatives and (-a)/b == -(a/b) in round-to-nearest, so only four A
// and four B quotients are distinct; each bracket is still formed per q.
// x / y for one of the three lattice constants y with r = RN(1/y).
// FP32: q0 = RN(x r), rem = x - q0 y (exact with one FMA), q = RN(q0 + rem r).
// An exhaustive CPU check (tests/fastdiv_check.c) shows q equal to the IEEE
// quotient x / y for every x whose biased exponent lies in [25, 251], for
// each of the three constants, and q = +0 for x = +0.
//
// The equilibrium uses it when uu < 1/2 (false for inf and NaN), and the
// IEEE division otherwise. With uu < 1/2 every dividend (ux, uy, c5, c6,
// their squares, uu) is below 4. One of at least 2^-102 is divided exactly;
// a smaller one (zero included) has a quotient below 2^-98 by either
// method, and it is added where the partial bracket is 1 (A: |ksi.u| <
// 2^-102; B: |ksi.u| < 2^-51, so A is absorbed; D: |u| < 2^-51, so every A
// and B is), below half an ulp of 1, so the bracket is the IEEE one either
// way. A zero of either sign in A leaves the bracket unchanged.
__device__ __forceinline__ float fast_div(float x, float y, float r) {
  const float q0 = x * r;
  const float rem = __builtin_fmaf(-q0, y, x);
  return __builtin_fmaf(rem, r, q0);
}

template <typename T>
struct EqTerms {
  T D, A1, A2, A5, A6, B1, B2, B5, B6;
};

template <typename T>
__device__ __forceinline__ EqTerms<T> eq_terms(const Lattice<T> &k, T ux, T uy, T uu, T c5,
                                               T c6) {
  return {uu / k.C2x2,        ux / k.C2,          uy / k.C2,
          c5 / k.C2,          c6 / k.C2,          (ux * ux) / k.C4x2,
          (uy * uy) / k.C4x2, (c5 * c5) / k.C4x2, (c6 * c6) / k.C4x2};
}
template <>
__device__ __forceinline__ EqTerms<float> eq_terms<float>(const Lattice<float> &k, float ux,
                                                          float uy, float uu, float c5, float c6) {
  const float s1 = ux * ux, s2 = uy * uy, s5 = c5 * c5, s6 = c6 * c6;
  if (uu < 0.5f) {
    return {fast_div(uu, k.C2x2, k.rC2x2), fast_div(ux, k.C2, k.rC2),
            fast_div(uy, k.C2, k.rC2),     fast_div(c5, k.C2, k.rC2),
            fast_div(c6, k.C2, k.rC2),     fast_div(s1, k.C4x2, k.rC4x2),
            fast_div(s2, k.C4x2, k.rC4x2), fast_div(s5, k.C4x2, k.rC4x2),
            fast_div(s6, k.C4x2, k.rC4x2)};
  }
  return {uu / k.C2x2, ux / k.C2, uy / k.C2, c5 / k.C2, c6 / k.C2,
          s1 / k.C4x2, s2 / k.C4x2, s5 / k.C4x2, s6 / k.C4x2};
}

template <typename T>
__device__ __forceinline__ void equilibrium(const Lattice<T> &k, T rho, T ux, T uy,
                                            T (&feq)[Q]) {
  T uu = T(0);
  uu += ux * ux;
  uu += uy * uy;
  const T c5 = ux + uy;          // q5; q7 = -c5
  const T c6 = (-ux) + uy;       // q6; q8 = -c6
  const EqTerms<T> t = eq_terms(k, ux, uy, uu, c5, c6);
  const T D = t.D, A1 = t.A1, A2 = t.A2, A5 = t.A5, A6 = t.A6;
  const T B1 = t.B1, B2 = t.B2, B5 = t.B5, B6 = t.B6;
  const T one = T(1);
  // W has three distinct values (W[1..4] and W[5..8] are equal floats)
  const T wr0 = k.W[0] * rho, wr1 = k.W[1] * rho, wr5 = k.W[5] * rho;
  feq[0] = wr0 * (((one + T(0)) + T(0)) - D);
  feq[1] = wr1 * (((one + A1) + B1) - D);
  feq[2] = wr1 * (((one + A2) + B2) - D);
  feq[3] = wr1 * (((one - A1) + B1) - D);
  feq[4] = wr1 * (((one - A2) + B2) - D);
  feq[5] = wr5 * (((one + A5) + B5) - D);
  feq[6] = wr5 * (((one + A6) + B6) - D);
  feq[7] = wr5 * (((one - A5) + B5) - D);
  feq[8] = wr5 * (((one - A6) + B6) - D);
}

// ----------------------------------------------------------------------------
// raw buffer access: 32-bit per-lane byte offsets plus a uniform (SGPR)
// offset; loads past num_records return 0 and stores past it are dropped
// ----------------------------------------------------------------------------
constexpr int32_t kRsrcWord3 = 0x00020000;  // gfx9-family raw buffer descriptor

struct Buf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Buf(const void *p, uint32_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                            kRsrcWord3)) {}
  template <typename T>
  __device__ __forceinline__ T ld(uint32_t voff, uint32_t soff) const;
  template <typename T>
  __device__ __forceinline__ void st(T v, uint32_t voff, uint32_t soff) const;
  __device__ __forceinline__ uint4 ld128(uint32_t voff, uint32_t soff) const {
    auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    return make_uint4(v[0], v[1], v[2], v[3]);
  }
};
template <>
__device__ __forceinline__ float Buf::ld<float>(uint32_t voff, uint32_t soff) const {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double Buf::ld<double>(uint32_t voff, uint32_t soff) const {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return __builtin_bit_cast(double, v);
}
template <>
__device__ __forceinline__ int32_t Buf::ld<int32_t>(uint32_t voff, uint32_t soff) const {
  return (int32_t)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
template <>
__device__ __forceinline__ void Buf::st<float>(float v, uint32_t voff, uint32_t soff) const {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}
template <>
__device__ __forceinline__ void Buf::st<double>(double v, uint32_t voff, uint32_t soff) const {
  using v2u = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, voff, soff, 0);
}

// LDS cell records: a cell's 9 pdfs, read and written one value at a time.
template <typename T>
struct LdsCell;
// Records of Q = 9 values with an odd stride in 32-bit words (9 for FP32;
// FP64 records are 18 words whose even start banks, with each value's
// second word, cover all 32 banks), so the records that 32 lanes read at
// scattered slots spread over all LDS banks. A 12-word stride (16-byte
// aligned records) used only a quarter of the banks: 4-way conflicts.
template <>
struct LdsCell<float> {
  static constexpr int W = 9;  // floats per cell record
  __device__ static __forceinline__ void put(float *base, int slot, const float (&f)[Q]) {
    float *d = base + slot * W;
#pragma unroll
    for (int q = 0; q < Q; ++q) d[q] = f[q];
  }
  __device__ static __forceinline__ void get(const float *base, int slot, float (&f)[Q]) {
    const float *d = base + slot * W;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = d[q];
  }
};
template <>
struct LdsCell<double> {
  static constexpr int W = 9;  // doubles per cell record
  __device__ static __forceinline__ void put(double *base, int slot, const double (&f)[Q]) {
    double *d = base + slot * W;
#pragma unroll
    for (int q = 0; q < Q; ++q) d[q] = f[q];
  }
  __device__ static __forceinline__ void get(const double *base, int slot, double (&f)[Q]) {
    const double *d = base + slot * W;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = d[q];
  }
};

// Stride of a (node, cell) pair's contribution record in the node phase:
// 12 values (rho w, ux w, uy w, 9 neq w) padded to an odd 13 words.
constexpr int kPairW = 13;

// Padded cells from which single handles launch directly by default (see
// create_impl).
constexpr int kDirectLaunchCells = 131072;

// One face of one cell: neighbour code, face normal, sign * length.
// FP32: one 16-byte record; FP64: two 16-byte records (code, pad, nx | ny, sL).
template <typename T>
struct FaceRec {
  int32_t code;
  T nx, ny, sL;
};
template <typename T>
struct GeoLayout;
template <>
struct GeoLayout<float> {
  static constexpr uint32_t bytes = 16;
  __device__ static __forceinline__ FaceRec<float> load(const Buf &b, uint32_t vo, uint32_t so) {
    const uint4 v = b.ld128(vo, so);
    return {(int32_t)v.x, __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w)};
  }
};
template <>
struct GeoLayout<double> {
  static constexpr uint32_t bytes = 32;
  __device__ static __forceinline__ FaceRec<double> load(const Buf &b, uint32_t vo, uint32_t so) {
    const uint4 v0 = b.ld128(vo, so), v1 = b.ld128(vo + 16, so);
    auto d = [](uint32_t lo, uint32_t hi) {
      return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    };
    return {(int32_t)v0.x, d(v0.z, v0.w), d(v1.x, v1.y), d(v1.z, v1.w)};
  }
};

template <typename T>
struct StepArgs {
  const T *fsrc;    // [9][Cp]
  T *fdst;          // [9][Cp]
  const void *geo;  // [3][Cp] FaceRec records (GeoLayout<T>::bytes each)
  const T *__restrict__ lwa;
  const T *__restrict__ lwden2;
  const int32_t *__restrict__ bf_slot;   // [2][nbf] slots in the tile node table
  const int32_t *__restrict__ bf_gnode;  // [2][nbf] node ids
  const T *__restrict__ bf_ratio;        // [nbf]
  const int32_t *__restrict__ tile_node_off;  // [ntiles+1]
  const int32_t *__restrict__ tile_nodes;     // encoded nodes of each tile
  const int32_t *__restrict__ tile_pair_off;  // [ntiles+1]
  const int32_t *__restrict__ entry_pair_off; // [entries+1]
  const int32_t *__restrict__ pair_cell;      // (node, cell) pairs
  const T *__restrict__ pair_w;
  const int32_t *__restrict__ an_type;
  const int32_t *__restrict__ an_cells;  // [nan][K]
  const T *__restrict__ an_w;            // [nan][K]
  const T *__restrict__ an_wsum;
  const T *__restrict__ an_rho;
  const T *__restrict__ an_ux;
  const T *__restrict__ an_uy;
  const T *__restrict__ node_pdf;  // [9][N]
  double *__restrict__ partials;   // [ntiles][2] residual partial sums
  const int32_t *__restrict__ block_tile;  // [grid] tile of each block (XCD-aware), -1 = none
  const int32_t *__restrict__ halo_off;    // [ntiles+1]
  const int32_t *__restrict__ halo_cells;  // out-of-tile neighbours of each tile
  const int32_t *__restrict__ tile_count;  // [ntiles] live cells (first slots of the tile)
  // batched solves: per-case buffer parity, mask and parameters
  const int32_t *__restrict__ tile_case;   // [ntiles]
  const int32_t *__restrict__ case_active; // [ncases]
  const int32_t *__restrict__ case_cur;    // [ncases] buffer holding the case's state
  const T *__restrict__ case_invtau;
  const T *__restrict__ case_dt;
  const T *f0;  // the two pdf buffers (batched solves pick one per case)
  T *f1;
  int batch;
  int bstep;                              // launch index within the current batch interval
  const int32_t *__restrict__ an_case;    // [nan] case of each active node (batched)
  int C, Cp, N, K, nbf, nan;
  int ntiles, xcd_remap, inline_nodes, stage, max_tile_nodes, max_tile_pairs, max_halo;
  Lattice<T> k;
};

// Two-step tiles (k_step2): ring lists, face records, ghost lists and the
// extended node tables (see HostLayout in layout.h). Every face record of
// the two-step kernel names its neighbour by a stage slot: own cells 0..bs-1,
// the tile's ring (S1 \ O, then S2) from bs, then one ghost record per
// ghost face of S1 (the own cells' first, each tile's in boundary-face
// order, then the S1 \ O cells').
template <typename T>
struct PairArgs {
  const int32_t *__restrict__ ring_off;    // [ntiles+1] into ring_cells
  const int32_t *__restrict__ ring_cells;  // S1 \ O then S2, device ids
  const int32_t *__restrict__ ring_n1;     // [ntiles] |S1 \ O|
  const int32_t *__restrict__ rg_off;      // [ntiles] first ring face record of the tile
  const void *rg_geo;                      // [3][nrg_cells] face records of S1 \ O cells
  const void *geo2;                        // [3][Cp] face records of the own cells
  const int32_t *__restrict__ gh_off;      // [ntiles+1] into the ghost lists
  const int32_t *__restrict__ gh_nown;     // [ntiles] ghosts of own cells (listed first)
  const int32_t *__restrict__ gh_s0, *__restrict__ gh_s1;  // extended-table slots of the face nodes
  const int32_t *__restrict__ gh_self;     // stage slot of the cell owning the face
  const T *__restrict__ gh_ratio;          // ghost distance / known distance
  const int32_t *__restrict__ xnode_off, *__restrict__ xnodes, *__restrict__ xentry_pair_off,
      *__restrict__ xpair_off, *__restrict__ xpair_cell;
  const T *__restrict__ xpair_w;
  const int32_t *__restrict__ pair_slot;   // stage slot of each own-table pair
  int nrg_cells;                           // ring cells with face records (all tiles)
  int max_xnodes, max_xpairs;              // LDS partition: table rows, pair records
};

// The 9 pdfs of cell c from a q-major [9][Cp] buffer.
template <typename T>
__device__ __forceinline__ void load_cell(const Buf &f, uint32_t plane, int c, T (&out)[Q]) {
  const uint32_t vo = (uint32_t)c * sizeof(T);
#pragma unroll
  for (int q = 0; q < Q; ++q) out[q] = f.ld<T>(vo, q * plane);
}

// The same through plain global addressing, for cells whose index is known
// to lie in [0, Cp): a two-step tile's own cells and its ring cells (ids the
// layout builder checked). On the 1M-cell channel k_step2 runs 1-2% faster
// this way than with buffer loads at nine per-plane scalar offsets (same
// results bit for bit).
template <typename T>
__device__ __forceinline__ void load_cell_at(const T *__restrict__ f, size_t Cp, int c, T (&out)[Q]) {
#pragma unroll
  for (int q = 0; q < Q; ++q) out[q] = f[(size_t)q * Cp + c];
}

// Nodes.calc_dirichlet_pdf for an active node (src/containers.py:343-390):
// type 1 (velocity): rho <- IDW(cell rho), u fixed; type 2 (density):
// u <- IDW(cell u), rho fixed; pdf = calc_eq(rho, u) + IDW(f_c - feq_c).
// Padding entries (-1) carry weight 0 and contribute +0, as x*0 does.
template <typename T>
__device__ __forceinline__ void active_node(const StepArgs<T> &a, const Buf &fb, uint32_t plane,
                                            int an, T (&out)[Q], T &rho_n, T &ux_n, T &uy_n) {
  const int type = a.an_type[an];
  const int32_t *cells = a.an_cells + (size_t)an * a.K;
  const T *w = a.an_w + (size_t)an * a.K;
  const T wsum = a.an_wsum[an];
  rho_n = a.an_rho[an];
  ux_n = a.an_ux[an];
  uy_n = a.an_uy[an];
  {
    T n0 = T(0), n1 = T(0);
    for (int j = 0; j < a.K; ++j) {
      const int cj = cells[j];
      const T wj = w[j];
      T r = T(0), ux = T(0), uy = T(0);
      if (!(cj < 0 && wj == T(0))) {
        T fc[Q];
        load_cell(fb, plane, cj < 0 ? -2 - cj : cj, fc);
        macro(fc, r, ux, uy);
      }
      if (type == 1) {
        n0 += r * wj;
      } else {
        n0 += ux * wj;
        n1 += uy * wj;
      }
    }
    if (type == 1) {
      rho_n = n0 / wsum;
    } else {
      ux_n = n0 / wsum;
      uy_n = n1 / wsum;
    }
  }
  T acc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) acc[q] = T(0);
  for (int j = 0; j < a.K; ++j) {
    const int cj = cells[j];
    const T wj = w[j];
    if (cj < 0) {
#pragma unroll
      for (int q = 0; q < Q; ++q) acc[q] += T(0) * wj;
    } else {
      T fc[Q], eq[Q], r, ux, uy;
      load_cell(fb, plane, cj, fc);
      macro(fc, r, ux, uy);
      equilibrium(a.k, r, ux, uy, eq);
#pragma unroll
      for (int q = 0; q < Q; ++q) acc[q] += (fc[q] - eq[q]) * wj;
    }
  }
  T eqn[Q];
  equilibrium(a.k, rho_n, ux_n, uy_n, eqn);
#pragma unroll
  for (int q = 0; q < Q; ++q) out[q] = eqn[q] + acc[q] / wsum;
}

// XCD-aware tile order: consecutive workgroup ids go round-robin to the 8
// XCDs, so each XCD gets one contiguous range of tiles (neighbour gathers then
// mostly hit that XCD's own L2); inside a range the tiles with boundary-node
// work come first, so their longer latency overlaps the bulk of the grid.
// block_tile is built on the host; without it the identity order is used.
template <typename T>
__device__ __forceinline__ int tile_of_block(const StepArgs<T> &a) {
  if (!a.xcd_remap) return blockIdx.x;
  return a.block_tile[blockIdx.x];
}

template <typename T>
__device__ __forceinline__ void block_reduce_store(T v0, T v1, double *out) {
  double d0 = double(v0), d1 = double(v1);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    d0 += __shfl_down(d0, off, 64);
    d1 += __shfl_down(d1, off, 64);
  }
  __shared__ double s0[16], s1[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s0[wave] = d0;
    s1[wave] = d1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    double t0 = 0.0, t1 = 0.0;
    for (int w = 0; w < nw; ++w) {
      t0 += s0[w];
      t1 += s1[w];
    }
    out[0] = t0;
    out[1] = t1;
  }
}

// One time step for the cells of one tile.
// Phase 1 (tiles that own ghost faces only): the distinct face nodes those
// ghosts read are evaluated once per tile into LDS, [9][n] (Nodes.calc_pdfs
// for active nodes; the stored pdf for type-0 nodes). Nothing of the cell
// step is live yet, so the node IDW does not raise the step's register use.
// Phase 2: one cell per thread (see the file header).
// SCHEME 0: first-order upwind; 1: Lax-Wendroff. RESID: also accumulate the
// residual partial sums of |u^{n+1}-u^n|^2 and |u^{n+1}|^2 for this tile.
template <typename T, int SCHEME, bool RESID>
__global__ void __launch_bounds__(1024) k_step(StepArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *sh = reinterpret_cast<T *>(smem_raw);
  const int tile = tile_of_block(a);
  if (tile < 0 || tile >= a.ntiles) return;  // uniform per block
  const T *fsrc = a.fsrc;
  T *fdst = a.fdst;
  T invtau = a.k.invtau, dt = a.k.dt;
  if (a.batch) {  // the tile's case: skip it when frozen, else its parity and parameters
    const int cs = a.tile_case[tile];
    if (!a.case_active[cs]) return;  // uniform per block
    // case_cur holds the parity at the start of the interval; the launch
    // index within it (each launch swaps the buffers once) gives the
    // current one (advanced after the interval)
    const int cur = (a.case_cur[cs] + a.bstep) & 1;
    fsrc = cur ? a.f1 : a.f0;
    fdst = cur ? const_cast<T *>(a.f0) : a.f1;
    invtau = a.case_invtau[cs];
    dt = a.case_dt[cs];
  }
  const uint32_t plane = (uint32_t)a.Cp * sizeof(T);  // bytes between q planes
  const Buf fin(fsrc, Q * plane);
  const bool live = (int)threadIdx.x < a.tile_count[tile];
  const int t0 = a.tile_node_off[tile];
  const int ntn = a.tile_node_off[tile + 1] - t0;
  const bool nodes_here = a.inline_nodes && ntn > 0;  // uniform per block
  const int bs = blockDim.x;
  if (nodes_here) {
    // 1a: one thread per (node, cell) pair: that cell's weighted IDW
    //     contributions [r w, ux w, uy w, (f_q - feq_q) w] into LDS
    const int pb = a.tile_pair_off[tile], np = a.tile_pair_off[tile + 1] - pb;
    T *con = sh + Q * a.max_tile_nodes;
    for (int p = threadIdx.x; p < np; p += bs) {
      const int cj = a.pair_cell[pb + p];
      const T wj = a.pair_w[pb + p];
      T fc[Q], r, ux, uy;
      load_cell(fin, plane, cj < 0 ? -2 - cj : cj, fc);
      macro(fc, r, ux, uy);
      T *o = con + kPairW * p;
      o[0] = r * wj;
      o[1] = ux * wj;
      o[2] = uy * wj;
      if (cj >= 0) {
        T eq[Q];
        equilibrium(a.k, r, ux, uy, eq);
#pragma unroll
        for (int q = 0; q < Q; ++q) o[3 + q] = (fc[q] - eq[q]) * wj;
      } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) o[3 + q] = T(0) * wj;  // padding: neq is 0
      }
    }
    __syncthreads();
    // 1b: one thread per table entry: the IDW sums in nodes.cells_index
    //     order (0-initialised, as the reference's jnp.sum), then the
    //     Dirichlet pdf feq(rho, u) + IDW(neq) (src/containers.py:353-361)
    for (int i = threadIdx.x; i < ntn; i += bs) {
      const int e = a.tile_nodes[t0 + i];
      T out[Q];
      if (e >= 0) {
        const int type = a.an_type[e];
        const int p0 = a.entry_pair_off[t0 + i] - pb, p1 = a.entry_pair_off[t0 + i + 1] - pb;
        T n0 = T(0), n1 = T(0), acc[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] = T(0);
        for (int p = p0; p < p1; ++p) {
          const T *o = con + kPairW * p;
          if (type == 1) {
            n0 += o[0];
          } else {
            n0 += o[1];
            n1 += o[2];
          }
#pragma unroll
          for (int q = 0; q < Q; ++q) acc[q] += o[3 + q];
        }
        const T wsum = a.an_wsum[e];
        T rho_n = a.an_rho[e], ux_n = a.an_ux[e], uy_n = a.an_uy[e];
        if (type == 1) {
          rho_n = n0 / wsum;
        } else {
          ux_n = n0 / wsum;
          uy_n = n1 / wsum;
        }
        T eqn[Q];
        equilibrium(a.k, rho_n, ux_n, uy_n, eqn);
#pragma unroll
        for (int q = 0; q < Q; ++q) out[q] = eqn[q] + acc[q] / wsum;
      } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) out[q] = a.node_pdf[(size_t)q * a.N + (-1 - e)];
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) sh[q * ntn + i] = out[q];
    }
  }

  // Stage: the tile's own pdfs (registers and LDS) and its halo (the
  // out-of-tile neighbours, LDS only) as [9][bs + max_halo]; the face
  // records then name each neighbour by its LDS slot, so no gather of
  // phase 2 touches L2/HBM.
  const int base = tile * bs;
  const int c = base + threadIdx.x;
  // 16-byte aligned start of the stage (after node table and contributions)
  const int pre = (Q * a.max_tile_nodes + kPairW * a.max_tile_pairs + 7) & ~7;
  T *shf = static_cast<T *>(__builtin_assume_aligned(sh + pre, 16));
  T f[Q];
  FaceRec<T> fr[3] = {};
  if (live) {
    load_cell(fin, plane, c, f);
    // face records are loaded with the pdfs, before the barrier, so both
    // share one memory round trip
    const Buf geo(a.geo, 3 * (uint32_t)a.Cp * GeoLayout<T>::bytes);
    const uint32_t gplane = (uint32_t)a.Cp * GeoLayout<T>::bytes;
    const uint32_t gvo = (uint32_t)c * GeoLayout<T>::bytes;
#pragma unroll
    for (int kf = 0; kf < 3; ++kf) fr[kf] = GeoLayout<T>::load(geo, gvo, kf * gplane);
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = T(0);
  }
  if (a.stage) {
    LdsCell<T>::put(shf, threadIdx.x, f);
    const int h0 = a.halo_off[tile], nh = a.halo_off[tile + 1] - h0;
    for (int i = threadIdx.x; i < nh; i += bs) {
      T g[Q];
      load_cell(fin, plane, a.halo_cells[h0 + i], g);
      LdsCell<T>::put(shf, bs + i, g);
    }
  }
  if (nodes_here || a.stage) __syncthreads();

  T res_d = T(0), res_n = T(0);
  if (live) {
    T acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = T(0);
#pragma unroll
    for (int kf = 0; kf < 3; ++kf) {
      const int32_t code = fr[kf].code;
      const T nx = fr[kf].nx, ny = fr[kf].ny, sL = fr[kf].sL;
      T la = T(0), den2 = T(1);
      if constexpr (SCHEME == 1) {
        const size_t at = (size_t)kf * a.Cp + c;
        la = a.lwa[at];
        den2 = a.lwden2[at];
      }
      const bool slot1 = (code & kSlot1Bit) != 0;
      T o[Q];
      if (code >= 0) {
        const int nbr = code & kIndexMask;  // LDS slot when staged, else cell id
        if (a.stage) {
          LdsCell<T>::get(shf, nbr, o);
        } else {
          load_cell(fin, plane, nbr, o);
        }
      } else {
        // ghost (Faces.calc_ghost + utils.extrap_pdf):
        // fbar = mean(f_p, f_r); g = fbar + (fbar - f_self) * (d_ghost / d_known)
        const int bf = code & kIndexMask;
        const T ratio = a.bf_ratio[bf];
        if (a.inline_nodes) {
          const int s0 = a.bf_slot[bf], s1 = a.bf_slot[a.nbf + bf];
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const T fbar = (sh[q * ntn + s0] + sh[q * ntn + s1]) / T(2);
            o[q] = fbar + (fbar - f[q]) * ratio;
          }
        } else {
          const int g0 = a.bf_gnode[bf], g1 = a.bf_gnode[a.nbf + bf];
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const T fbar =
                (a.node_pdf[(size_t)q * a.N + g0] + a.node_pdf[(size_t)q * a.N + g1]) / T(2);
            o[q] = fbar + (fbar - f[q]) * ratio;
          }
        }
      }
      if constexpr (SCHEME == 0) {
        // ksi.n for q = 1..8. Opposite directions give exact negatives, so
        // two sums cover all eight; the value can differ from the literal
        // 0-initialised dot product only in the sign of a zero, which makes
        // only the sign of a zero flux differ, and an accumulator started at
        // +0 absorbs a zero of either sign identically. q = 0 has ksi = 0:
        // its flux is a zero and is never added.
        const T p5 = nx + ny, p6 = ny - nx;
        const T phi[Q] = {T(0), nx, ny, -nx, -ny, p5, p6, -p5, -p6};
#pragma unroll
        for (int q = 1; q < Q; ++q) {
          // upwind: slot 0 when ksi.n >= 0, else slot 1
          const bool take_slot0 = phi[q] >= T(0);
          const T up = (take_slot0 != slot1) ? f[q] : o[q];
          acc[q] += (up * phi[q]) * sL;
        }
      } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          T phi = T(0);
          phi += T(kx(q)) * nx;
          phi += T(ky(q)) * ny;
          const T fu = slot1 ? o[q] : f[q];
          const T fd = slot1 ? f[q] : o[q];
          const T interp = la - (phi * dt) / den2;
          const T fc = fu + (fd - fu) * interp;
          acc[q] += (fc * phi) * sL;
        }
      }
    }
    T rho, ux, uy, feq[Q];
    macro(f, rho, ux, uy);
    equilibrium(a.k, rho, ux, uy, feq);
    const Buf fout(fdst, Q * plane);
    const uint32_t vo = (uint32_t)c * sizeof(T);
    T fn[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      fn[q] = f[q] + dt * (invtau * (feq[q] - f[q]) - acc[q]);
      fout.st<T>(fn[q], vo, q * plane);
    }
    if constexpr (RESID) {
      T r2, vx, vy;
      macro(fn, r2, vx, vy);
      const T dx = vx - ux, dy = vy - uy;
      res_d = dx * dx + dy * dy;
      res_n = vx * vx + vy * vy;
    }
  }
  if constexpr (RESID) block_reduce_store(res_d, res_n, a.partials + 2 * (size_t)tile);
}

// Ghost records of one step (Faces.calc_ghost + utils.extrap_pdf,
// src/containers.py:280-287, utils/utils.py:153-154): for ghost face i of
// the list [g0, g0 + n), fbar = mean of its two node pdfs in the LDS node
// table (rows of length nrow) and g = fbar + (fbar - f_self) * ratio, with
// f_self the owning cell's record in the stage; g goes to stage slot
// slot0 + i. One lane per ghost face, so the cell updates read every face
// neighbour as a stage record and a wave never runs the extrapolation for
// a few of its lanes.
template <typename T>
__device__ __forceinline__ void ghost_records(T *shf, const T *tab, int nrow,
                                              const int32_t *__restrict__ s0v,
                                              const int32_t *__restrict__ s1v,
                                              const int32_t *__restrict__ selfv,
                                              const T *__restrict__ ratiov, int g0, int n,
                                              int slot0) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int s0 = s0v[g0 + i], s1 = s1v[g0 + i];
    const T ratio = ratiov[g0 + i];
    T fs[Q], g[Q];
    LdsCell<T>::get(shf, selfv[g0 + i], fs);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const T fbar = (tab[q * nrow + s0] + tab[q * nrow + s1]) / T(2);
      g[q] = fbar + (fbar - fs[q]) * ratio;
    }
    LdsCell<T>::put(shf, slot0 + i, g);
  }
}

// One upwind cell update with every face neighbour, ghost records
// included, in the LDS stage (code = stage slot | slot-1 bit): the
// arithmetic of k_step<T, 0, ..> phase 2, operation for operation.
template <typename T>
__device__ __forceinline__ void cell_update_staged(const Lattice<T> &k, const T (&f)[Q],
                                                   const FaceRec<T> (&fr)[3], const T *shf,
                                                   T invtau, T dt, T (&fn)[Q]) {
  T acc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) acc[q] = T(0);
#pragma unroll
  for (int kf = 0; kf < 3; ++kf) {
    const int32_t code = fr[kf].code;
    const T nx = fr[kf].nx, ny = fr[kf].ny, sL = fr[kf].sL;
    const bool slot1 = (code & kSlot1Bit) != 0;
    T o[Q];
    LdsCell<T>::get(shf, code & kIndexMask, o);
    const T p5 = nx + ny, p6 = ny - nx;
    const T phi[Q] = {T(0), nx, ny, -nx, -ny, p5, p6, -p5, -p6};
#pragma unroll
    for (int q = 1; q < Q; ++q) {
      const bool take_slot0 = phi[q] >= T(0);
      const T up = (take_slot0 != slot1) ? f[q] : o[q];
      acc[q] += (up * phi[q]) * sL;
    }
  }
  T rho, ux, uy, feq[Q];
  macro(f, rho, ux, uy);
  equilibrium(k, rho, ux, uy, feq);
#pragma unroll
  for (int q = 0; q < Q; ++q) fn[q] = f[q] + dt * (invtau * (feq[q] - f[q]) - acc[q]);
}

// Node values of table entries [0, nent) (rows of length nrow) from their
// (node, cell) pairs [pb, pb + np) (entry_pair_off absolute), the pair
// cells read from global memory (FROM_LDS false: pair_cell device ids,
// -2 - id for a weighted padding entry) or from the LDS stage (pair_slot):
// k_step's phase 1 arithmetic. Ends with a barrier.
template <typename T, bool FROM_LDS>
__device__ __forceinline__ void node_table(const StepArgs<T> &a, T *tab, int nrow, T *con,
                                           const int32_t *__restrict__ entries, int nent,
                                           const int32_t *__restrict__ epo, int pb, int np,
                                           const int32_t *__restrict__ pcell,
                                           const int32_t *__restrict__ pslot,
                                           const T *__restrict__ pw, const Buf &fin,
                                           uint32_t plane, const T *shf) {
  const int bs = blockDim.x;
  for (int p = threadIdx.x; p < np; p += bs) {
    const int cj = pcell[pb + p];
    const T wj = pw[pb + p];
    T fc[Q], r, ux, uy;
    if constexpr (FROM_LDS) {
      LdsCell<T>::get(shf, pslot[pb + p], fc);
    } else {
      load_cell(fin, plane, cj < 0 ? -2 - cj : cj, fc);
    }
    macro(fc, r, ux, uy);
    T *o = con + kPairW * p;
    o[0] = r * wj;
    o[1] = ux * wj;
    o[2] = uy * wj;
    if (cj >= 0) {
      T eq[Q];
      equilibrium(a.k, r, ux, uy, eq);
#pragma unroll
      for (int q = 0; q < Q; ++q) o[3 + q] = (fc[q] - eq[q]) * wj;
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) o[3 + q] = T(0) * wj;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nent; i += bs) {
    const int e = entries[i];
    T out[Q];
    if (e >= 0) {
      const int type = a.an_type[e];
      const int p0 = epo[i] - pb, p1 = epo[i + 1] - pb;
      T n0 = T(0), n1 = T(0), acc[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) acc[q] = T(0);
      for (int p = p0; p < p1; ++p) {
        const T *o = con + kPairW * p;
        if (type == 1) {
          n0 += o[0];
        } else {
          n0 += o[1];
          n1 += o[2];
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] += o[3 + q];
      }
      const T wsum = a.an_wsum[e];
      T rho_n = a.an_rho[e], ux_n = a.an_ux[e], uy_n = a.an_uy[e];
      if (type == 1) {
        rho_n = n0 / wsum;
      } else {
        ux_n = n0 / wsum;
        uy_n = n1 / wsum;
      }
      T eqn[Q];
      equilibrium(a.k, rho_n, ux_n, uy_n, eqn);
#pragma unroll
      for (int q = 0; q < Q; ++q) out[q] = eqn[q] + acc[q] / wsum;
    } else {
#pragma unroll
      for (int q = 0; q < Q; ++q) out[q] = a.node_pdf[(size_t)q * a.N + (-1 - e)];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) tab[q * nrow + i] = out[q];
  }
  __syncthreads();
}

// Two steps of one tile (upwind, single handle, staged, inline nodes):
// step n for S1 = the tile and its ring from the inputs of S1 and S2 (the
// extended node table from global memory), the results kept in the LDS
// stage, then the own node table from those results and step n + 1 for
// the tile's cells, written to fdst. The intermediate state never goes to
// HBM; every value equals what two launches of k_step compute. Batched
// handles: a.bstep is the launch index within the interval (as in k_step),
// and a frozen case's tiles return at once.
template <typename T>
__global__ void __launch_bounds__(1024) k_step2(StepArgs<T> a, PairArgs<T> p) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *sh = reinterpret_cast<T *>(smem_raw);
  const int tile = tile_of_block(a);
  if (tile < 0 || tile >= a.ntiles) return;  // uniform per block
  const T *fsrc = a.fsrc;
  T *fdst = a.fdst;
  T invtau = a.k.invtau, dt = a.k.dt;
  if (a.batch) {  // ring cells and node-fan cells belong to the tile's case
    const int cs = a.tile_case[tile];
    if (!a.case_active[cs]) return;  // uniform per block
    const int cur = (a.case_cur[cs] + a.bstep) & 1;
    fsrc = cur ? a.f1 : a.f0;
    fdst = cur ? const_cast<T *>(a.f0) : a.f1;
    invtau = a.case_invtau[cs];
    dt = a.case_dt[cs];
  }
  const uint32_t plane = (uint32_t)a.Cp * sizeof(T);
  const Buf fin(fsrc, Q * plane);
  const int bs = blockDim.x;
  const bool live = (int)threadIdx.x < a.tile_count[tile];
  const int x0 = p.xnode_off[tile], nx = p.xnode_off[tile + 1] - x0;
  const int t0 = a.tile_node_off[tile], ntn = a.tile_node_off[tile + 1] - t0;
  T *con = sh + Q * p.max_xnodes;
  const int pre = (Q * p.max_xnodes + kPairW * p.max_xpairs + 7) & ~7;
  T *shf = static_cast<T *>(__builtin_assume_aligned(sh + pre, 16));
  // step-n node values of every ghost face of S1, from global memory
  if (nx > 0) {
    const int pb = p.xpair_off[tile];
    node_table<T, false>(a, sh, nx, con, p.xnodes + x0, nx, p.xentry_pair_off + x0, pb,
                         p.xpair_off[tile + 1] - pb, p.xpair_cell, nullptr, p.xpair_w, fin, plane,
                         shf);
  }
  // stage: own cells, then S1 \ O and S2
  const int c = tile * bs + threadIdx.x;
  T f[Q];
  FaceRec<T> fr[3] = {};
  if (live) {
    load_cell_at(fsrc, (size_t)a.Cp, c, f);
    const Buf geo(p.geo2, 3 * (uint32_t)a.Cp * GeoLayout<T>::bytes);
    const uint32_t gplane = (uint32_t)a.Cp * GeoLayout<T>::bytes;
    const uint32_t gvo = (uint32_t)c * GeoLayout<T>::bytes;
#pragma unroll
    for (int kf = 0; kf < 3; ++kf) fr[kf] = GeoLayout<T>::load(geo, gvo, kf * gplane);
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = T(0);
  }
  LdsCell<T>::put(shf, threadIdx.x, f);
  const int r0 = p.ring_off[tile], nr = p.ring_off[tile + 1] - r0, n1 = p.ring_n1[tile];
  // ghost records: the own cells' ghost faces, then the S1 \ O cells', in
  // stage slots after the ring
  const int g0 = p.gh_off[tile], ng = p.gh_off[tile + 1] - g0, ngo = p.gh_nown[tile];
  for (int i = threadIdx.x; i < nr; i += bs) {
    T g[Q];
    load_cell_at(fsrc, (size_t)a.Cp, p.ring_cells[r0 + i], g);
    LdsCell<T>::put(shf, bs + i, g);
  }
  // the ring cell of this thread (n1 <= bs) and its face records
  const bool ring_here = (int)threadIdx.x < n1;
  FaceRec<T> rr[3] = {};
  if (ring_here) {
    const Buf rgeo(p.rg_geo, 3 * (uint32_t)p.nrg_cells * GeoLayout<T>::bytes);
    const uint32_t rplane = (uint32_t)p.nrg_cells * GeoLayout<T>::bytes;
    const uint32_t rvo = (uint32_t)(p.rg_off[tile] + (int)threadIdx.x) * GeoLayout<T>::bytes;
#pragma unroll
    for (int kf = 0; kf < 3; ++kf) rr[kf] = GeoLayout<T>::load(rgeo, rvo, kf * rplane);
  }
  __syncthreads();
  if (ng > 0) {  // uniform per block
    ghost_records(shf, sh, nx, p.gh_s0, p.gh_s1, p.gh_self, p.gh_ratio, g0, ng, bs + nr);
    __syncthreads();
  }
  // step n for the own cell and the ring cell of this thread
  T fo[Q], fq[Q];
  if (live) cell_update_staged(a.k, f, fr, shf, invtau, dt, fo);
  if (ring_here) {
    T g[Q];
    LdsCell<T>::get(shf, bs + threadIdx.x, g);
    cell_update_staged(a.k, g, rr, shf, invtau, dt, fq);
  }
  __syncthreads();  // every step-n input has been read
  if (live) LdsCell<T>::put(shf, threadIdx.x, fo);
  if (ring_here) LdsCell<T>::put(shf, bs + threadIdx.x, fq);
  __syncthreads();
  // step-(n+1) node values of the own ghost faces, from the stage, and the
  // own cells' ghost records from them (the first ngo of the list, same
  // slots)
  if (ntn > 0) {
    const int pb = a.tile_pair_off[tile];
    node_table<T, true>(a, sh, nx, con, a.tile_nodes + t0, ntn, a.entry_pair_off + t0, pb,
                        a.tile_pair_off[tile + 1] - pb, a.pair_cell, p.pair_slot, a.pair_w, fin,
                        plane, shf);
  }
  if (ngo > 0) {  // uniform per block
    ghost_records(shf, sh, nx, p.gh_s0, p.gh_s1, p.gh_self, p.gh_ratio, g0, ngo, bs + nr);
    __syncthreads();
  }
  if (live) {
    T fn[Q];
    cell_update_staged(a.k, fo, fr, shf, invtau, dt, fn);
#pragma unroll
    for (int q = 0; q < Q; ++q) fdst[(size_t)q * a.Cp + c] = fn[q];
  }
}

// Active nodes from the pdf in fsrc: node pdf (type 1/2), node rho (type 1)
// and node vel (type 2) written to the node arrays (reference node order).
template <typename T>
__global__ void k_nodes(StepArgs<T> a, const int32_t *__restrict__ an_gid, T *__restrict__ npdf,
                        T *__restrict__ nrho, T *__restrict__ nvel) {
  const int an = blockIdx.x * blockDim.x + threadIdx.x;
  if (an >= a.nan) return;
  const uint32_t plane = (uint32_t)a.Cp * sizeof(T);
  const T *src = a.fsrc;
  if (a.batch) src = ((a.case_cur[a.an_case[an]] + a.bstep) & 1) ? a.f1 : a.f0;
  const Buf fin(src, Q * plane);
  T out[Q], r, ux, uy;
  active_node(a, fin, plane, an, out, r, ux, uy);
  const int g = an_gid[an];
#pragma unroll
  for (int q = 0; q < Q; ++q) npdf[(size_t)q * a.N + g] = out[q];
  if (a.an_type[an] == 1) {
    nrho[g] = r;
  } else {
    nvel[g] = ux;
    nvel[(size_t)a.N + g] = uy;
  }
}

// Cell moments and equilibrium of fsrc (cells.rho / vel / pdf_eq), device order.
template <typename T>
__global__ void k_cell_moments(StepArgs<T> a, T *__restrict__ rho, T *__restrict__ vel,
                               T *__restrict__ feq) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.C) return;
  const uint32_t plane = (uint32_t)a.Cp * sizeof(T);
  const Buf fin(a.fsrc, Q * plane);
  T f[Q], eq[Q], r, ux, uy;
  load_cell(fin, plane, c, f);
  macro(f, r, ux, uy);
  equilibrium(a.k, r, ux, uy, eq);
  rho[c] = r;
  vel[c] = ux;
  vel[(size_t)a.Cp + c] = uy;
#pragma unroll
  for (int q = 0; q < Q; ++q) feq[(size_t)q * a.Cp + c] = eq[q];
}

// Face flux (faces.pdf) in reference face order, from fsrc and the node pdf
// array (active nodes already refreshed by k_nodes). Follows
// calc_upwind_flux / calc_lax_wendroff_flux face by face, ghosts on either
// slot, index -1 of a missing known cell read as the last reference cell.
struct FaceArrays {
  const int32_t *c0, *c1, *n0, *n1;
  const void *d0, *d1, *nx, *ny, *L;
  int F, last_cell;
};

template <typename T, int SCHEME>
__global__ void k_face_flux(StepArgs<T> a, FaceArrays fa, T *__restrict__ flux) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= fa.F) return;
  const uint32_t plane = (uint32_t)a.Cp * sizeof(T);
  const Buf fin(a.fsrc, Q * plane);
  const int ia = fa.c0[j], ib = fa.c1[j];
  const T d0 = static_cast<const T *>(fa.d0)[j], d1 = static_cast<const T *>(fa.d1)[j];
  const T nx = static_cast<const T *>(fa.nx)[j], ny = static_cast<const T *>(fa.ny)[j];
  const T L = static_cast<const T *>(fa.L)[j];
  T fl[Q], fr[Q];
  for (int side = 0; side < 2; ++side) {
    const int idx = side == 0 ? ia : ib;
    T tmp[Q];
    if (idx != -1) {
      load_cell(fin, plane, idx, tmp);
    } else {
      int known = side == 0 ? ib : ia;
      if (known < 0) known = fa.last_cell;
      const T gd = side == 0 ? d0 : d1, kd = side == 0 ? d1 : d0;
      const T ratio = gd / kd;
      T fk[Q];
      load_cell(fin, plane, known, fk);
      const int p = fa.n0[j], r = fa.n1[j];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const T fbar = (a.node_pdf[(size_t)q * a.N + p] + a.node_pdf[(size_t)q * a.N + r]) / T(2);
        tmp[q] = fbar + (fbar - fk[q]) * ratio;
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) (side == 0 ? fl : fr)[q] = tmp[q];
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    T phi = T(0);
    phi += T(kx(q)) * nx;
    phi += T(ky(q)) * ny;
    T v;
    if constexpr (SCHEME == 0) {
      v = ((phi >= T(0)) ? fl[q] : fr[q]) * phi;
    } else {
      const T s = d0 + d1;
      const T interp = (d0 / s) - (phi * a.k.dt) / (T(2) * s);
      v = (fl[q] + (fr[q] - fl[q]) * interp) * phi;
    }
    flux[(size_t)q * fa.F + j] = v * L;
  }
}

// Deterministic sum of the per-tile residual partials (fixed order).
__global__ void k_residual_finalize(const double *__restrict__ partials, int ntiles,
                                    double *__restrict__ out) {
  __shared__ double s0[256], s1[256];
  double t0 = 0.0, t1 = 0.0;
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) {
    t0 += partials[2 * (size_t)i];
    t1 += partials[2 * (size_t)i + 1];
  }
  s0[threadIdx.x] = t0;
  s1[threadIdx.x] = t1;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s0[threadIdx.x] += s0[threadIdx.x + s];
      s1[threadIdx.x] += s1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s0[0];
    out[1] = s1[0];
  }
}

// Plain device copy (16 B per lane per iteration, grid-stride): the
// known-size reference for converting FETCH_SIZE / WRITE_SIZE counters and
// for the attainable copy bandwidth on this device.
__global__ void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Batched solves, after an interval of n steps taken by `swaps` launches
// (launch index 0..swaps-1; a two-step launch swaps the buffers once):
// active cases advance their step count by n and their parity by swaps.
__global__ void k_batch_advance(int ncases, const int32_t *__restrict__ active,
                                int32_t *__restrict__ cur, int64_t *__restrict__ steps, int n,
                                int swaps) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncases || !active[k]) return;
  cur[k] ^= (swaps & 1);
  steps[k] += n;
}

// Batched solves, after a checked interval of `every` steps: per case the
// residual of its last step
// from its tiles' partial sums (fixed order), then the case stops when it is
// converged (r < tol), non-finite, or (growth > 0) r > growth * best r.
// One block; n_active = number of cases still running.
enum { kRunning = 0, kConverged = 1, kDiverged = 2, kGrowing = 3, kStopped = 4 };
__global__ void k_batch_check(int ncases, const int32_t *__restrict__ case_tile_off,
                              const double *__restrict__ partials, double tol, double growth,
                              int32_t *__restrict__ active, int32_t *__restrict__ status,
                              double *__restrict__ resid, double *__restrict__ best,
                              int32_t *__restrict__ n_active, int32_t *__restrict__ cur,
                              int64_t *__restrict__ steps, int every, int swaps) {
  __shared__ int cnt[256];
  int mine = 0;
  for (int k = threadIdx.x; k < ncases; k += blockDim.x) {
    if (active[k]) {
      // the interval's steps: parity and count, as k_batch_advance
      cur[k] ^= (swaps & 1);
      steps[k] += every;
      double d = 0.0, n = 0.0;
      for (int t = case_tile_off[k]; t < case_tile_off[k + 1]; ++t) {
        d += partials[2 * (size_t)t];
        n += partials[2 * (size_t)t + 1];
      }
      // non-finite sums (a NaN or Inf pdf anywhere in the case) give a NaN
      // residual, so the case stops as diverged rather than passing r < tol
      const double r = !(isfinite(d) && isfinite(n)) ? NAN
                       : n > 0.0                     ? sqrt(d) / sqrt(n)
                       : d > 0.0                     ? INFINITY
                                                     : 0.0;
      resid[k] = r;
      if (!isfinite(r)) {
        status[k] = kDiverged;
        active[k] = 0;
      } else if (r < tol) {
        status[k] = kConverged;
        active[k] = 0;
      } else if (growth > 0.0 && r > growth * best[k]) {
        status[k] = kGrowing;
        active[k] = 0;
      }
      if (r < best[k]) best[k] = r;
    }
    mine += active[k] ? 1 : 0;
  }
  cnt[threadIdx.x] = mine;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) cnt[threadIdx.x] += cnt[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_active = cnt[0];
}

template <typename T>
__global__ void k_fill_uniform(T *__restrict__ f, int Cp, int C, Lattice<T> k) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Cp) return;
#pragma unroll
  for (int q = 0; q < Q; ++q) f[(size_t)q * Cp + c] = c < C ? k.W[q] : T(0);
}

// Moments of the current pdf per device slot; batched handles read each
// case's current buffer. Same arithmetic as Dynamics.calc_macro.
template <typename T>
__global__ void k_slot_macros(const T *__restrict__ f0, const T *__restrict__ f1, int Cp, int cur,
                              int block, const int32_t *__restrict__ tile_case,
                              const int32_t *__restrict__ case_cur, T *__restrict__ rho,
                              T *__restrict__ ux, T *__restrict__ uy) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Cp) return;
  if (tile_case) cur = case_cur[tile_case[c / block]];
  const T *f = cur ? f1 : f0;
  T v[Q], r, x, y;
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = f[(size_t)q * Cp + c];
  macro(v, r, x, y);
  rho[c] = r;
  ux[c] = x;
  uy[c] = y;
}

// Inverse-distance weighting at every node (utils.weighted_avg,
// utils/utils.py:34-45, applied to all nodes rather than boundary nodes
// only): w = 1/d, 0 where negative (padding); sum(x w) / sum(w), sums from
// 0 in list order. Padding entries add +0 and are skipped.
template <typename T>
__global__ void k_node_idw(int N, int K, const int32_t *__restrict__ nc, const T *__restrict__ nd,
                           const T *__restrict__ rho, const T *__restrict__ ux,
                           const T *__restrict__ uy, T *__restrict__ nrho, T *__restrict__ nux,
                           T *__restrict__ nuy) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  T ws = T(0), sr = T(0), sx = T(0), sy = T(0);
  for (int k = 0; k < K; ++k) {
    const int c = nc[(size_t)n * K + k];
    T w = T(1) / nd[(size_t)n * K + k];
    if (w < T(0)) w = T(0);
    ws += w;
    if (c >= 0) {
      sr += rho[c] * w;
      sx += ux[c] * w;
      sy += uy[c] * w;
    }
  }
  nrho[n] = sr / ws;
  nux[n] = sx / ws;
  nuy[n] = sy / ws;
}

// Node -> cell: inverse-distance weighting of a cell's vertices (weights
// 1/d, d the vertex-to-centroid distance, vertices in ascending node id).
// Slots without vertices (batch padding) get rho = 1, u = 0.
template <typename T>
__global__ void k_cell_from_nodes(int Cp, const int32_t *__restrict__ off,
                                  const int32_t *__restrict__ node, const T *__restrict__ d,
                                  const T *__restrict__ nrho, const T *__restrict__ nux,
                                  const T *__restrict__ nuy, T *__restrict__ rho,
                                  T *__restrict__ ux, T *__restrict__ uy) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Cp) return;
  if (off[c] == off[c + 1]) {
    rho[c] = T(1);
    ux[c] = T(0);
    uy[c] = T(0);
    return;
  }
  T ws = T(0), sr = T(0), sx = T(0), sy = T(0);
  for (int e = off[c]; e < off[c + 1]; ++e) {
    const int n = node[e];
    const T w = T(1) / d[e];
    ws += w;
    sr += nrho[n] * w;
    sx += nux[n] * w;
    sy += nuy[n] * w;
  }
  rho[c] = sr / ws;
  ux[c] = sx / ws;
  uy[c] = sy / ws;
}

// pdf = feq(rho, u) per slot (D2Q9.calc_eq arithmetic), q-major.
template <typename T>
__global__ void k_equilibrium_fill(int Cp, Lattice<T> k, const T *__restrict__ rho,
                                   const T *__restrict__ ux, const T *__restrict__ uy,
                                   T *__restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Cp) return;
  T eq[Q];
  equilibrium(k, rho[c], ux[c], uy[c], eq);
#pragma unroll
  for (int q = 0; q < Q; ++q) out[(size_t)q * Cp + c] = eq[q];
}

// ----------------------------------------------------------------------------
// handle
// ----------------------------------------------------------------------------
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct Handle {
  int precision = FVDBM_FP32, scheme = FVDBM_UPWIND, variant = 0, block = 256;
  int graph_steps = 32, reorder = 1;
  double tau = 1.0, dt = 1.0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<DevBuf> bufs;
  int64_t device_bytes = 0;
  // layout scalars
  int C = 0, Cp = 0, F = 0, N = 0, K = 0, nbc = 0, nbf = 0, nan = 0, bandwidth = 0;
  int last_cell_new = 0, ntiles = 0;
  std::vector<int32_t> new_to_old;
  std::vector<int32_t> an_gid_host, an_type_host;
  // device pointers (typed as void to share between precisions)
  void *f[2] = {nullptr, nullptr};
  void *geo = nullptr, *lwa = nullptr, *lwden2 = nullptr;
  void *bf_slot = nullptr, *bf_gnode = nullptr, *bf_ratio = nullptr;
  void *tile_node_off = nullptr, *tile_nodes = nullptr;
  void *tile_pair_off = nullptr, *entry_pair_off = nullptr, *pair_cell = nullptr, *pair_w = nullptr;
  int max_tile_nodes = 0, max_tile_pairs = 0;
  void *block_tile = nullptr, *halo_off = nullptr, *halo_cells = nullptr;
  int max_halo = 0;
  void *an_gid = nullptr, *an_type = nullptr, *an_cells = nullptr, *an_w = nullptr,
       *an_wsum = nullptr, *an_rho = nullptr, *an_ux = nullptr, *an_uy = nullptr;
  void *node_pdf = nullptr, *node_rho = nullptr, *node_vel = nullptr;
  void *fc0 = nullptr, *fc1 = nullptr, *fn0 = nullptr, *fn1 = nullptr, *fd0 = nullptr,
       *fd1 = nullptr, *fnx = nullptr, *fny = nullptr, *fL = nullptr;
  void *partials = nullptr, *resid = nullptr;
  void *obs_rho = nullptr, *obs_vel = nullptr, *obs_feq = nullptr, *obs_flux = nullptr;
  void *tile_count = nullptr;
  // two-step tiles (k_step2): used unless variant bit 8 is set, for single
  // upwind handles whose rings and LDS fit
  bool pairs = false;
  size_t lds2 = 0;
  // LDS classes of the two-step launch: with npair_class = 2, tiles whose
  // LDS fits pair_lds[0] run in a first launch sized for it (so LDS does not
  // limit its occupancy) and the others in a second one; 0 = one launch over
  // block_tile with lds2
  int npair_class = 0;
  // with two classes the second (few tiles, larger LDS) runs concurrently on
  // stream2, forked from and joined back into the handle's stream (evf, evj)
  hipStream_t stream2 = nullptr;
  hipEvent_t evf = nullptr, evj = nullptr;
  void *pair_bt[2] = {nullptr, nullptr};
  int pair_grid[2] = {0, 0};
  size_t pair_lds[2] = {0, 0};
  int pair_xnodes[2] = {0, 0}, pair_xpairs[2] = {0, 0};
  int nrg_cells = 0, max_xnodes = 0, max_xpairs = 0;
  void *ring_off = nullptr, *ring_cells = nullptr, *ring_n1 = nullptr, *rg_off = nullptr,
       *rg_geo = nullptr, *geo2 = nullptr, *gh_off = nullptr, *gh_nown = nullptr,
       *gh_s0 = nullptr, *gh_s1 = nullptr, *gh_self = nullptr, *gh_ratio = nullptr,
       *xnode_off = nullptr,
       *xnodes = nullptr, *xentry_pair_off = nullptr, *xpair_off = nullptr,
       *xpair_cell = nullptr, *xpair_w = nullptr, *pair_slot = nullptr;
  // captured runs of two-step launches, keyed by (starting parity, launches)
  std::map<std::pair<int, int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs2;
  // two-step handles: after a run that ended in a two-step launch
  // (prev_two_back), f[cur ^ 1] holds the pdf two steps back; the pdf at the
  // start of the last step is one k_step from it, computed into fmid only
  // when an observable is read (mid_ready)
  void *fmid = nullptr;
  bool prev_two_back = false, mid_ready = false;
  // all-node interpolation: node -> (slot, d) lists [N][K], cell -> vertex
  // CSR in slot order, and lazily allocated work arrays
  void *all_nc = nullptr, *all_nd = nullptr, *cn_off = nullptr, *cn_node = nullptr, *cn_d = nullptr;
  void *w_rho = nullptr, *w_ux = nullptr, *w_uy = nullptr, *w_nrho = nullptr, *w_nux = nullptr,
       *w_nuy = nullptr, *w_pdf = nullptr;
  // batched solves
  int batch = 0, ncases = 1;
  int bstep = 0;             // launch index within the batch interval being launched
  void *an_case = nullptr;   // [nan] case of each active node
  hipGraphExec_t sexec = nullptr;  // batched: graph_steps steps + advance
  hipGraph_t sgraph = nullptr;
  std::vector<int32_t> case_tile_off_host;
  std::vector<int64_t> case_cell_off;  // reference cell range of each case
  void *tile_case = nullptr, *case_tile_off = nullptr, *case_active = nullptr,
       *case_cur = nullptr, *case_steps = nullptr, *case_status = nullptr,
       *case_resid = nullptr, *case_best = nullptr, *case_invtau = nullptr,
       *case_dt = nullptr, *n_active = nullptr;
  std::vector<double> case_tau_host, case_dt_host;
  hipGraphExec_t bexec = nullptr;
  hipGraph_t bgraph = nullptr;
  int bgraph_every = 0;
  double bgraph_tol = 0.0, bgraph_growth = 0.0;
  int cur = 0;           // buffer holding the current pdf
  int64_t steps = 0;     // steps taken since the last init/set of the cell pdf
  bool has_prev = false; // the pdf at the start of the last step is known (f[cur^1], or see fmid)
  bool observed = false; // observable buffers are up to date
  // graphs: one per starting buffer parity
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};

  ~Handle() {
    // steps are queued asynchronously: let them finish before their graphs,
    // buffers and streams go
    if (stream2) (void)hipStreamSynchronize(stream2);
    if (stream) (void)hipStreamSynchronize(stream);
    for (int i = 0; i < 2; ++i) {
      if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
      if (graph[i]) (void)hipGraphDestroy(graph[i]);
    }
    for (auto &kv : graphs2) {
      if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
      if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
    }
    if (bexec) (void)hipGraphExecDestroy(bexec);
    if (sexec) (void)hipGraphExecDestroy(sexec);
    if (sgraph) (void)hipGraphDestroy(sgraph);
    if (bgraph) (void)hipGraphDestroy(bgraph);
    for (auto &b : bufs)
      if (b.p) (void)hipFree(b.p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (evf) (void)hipEventDestroy(evf);
    if (evj) (void)hipEventDestroy(evj);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
  }

  size_t esize() const { return precision == FVDBM_FP64 ? 8 : 4; }

  int alloc(void **out, size_t bytes) {
    if (bytes == 0) bytes = 256;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
      return fail(FVDBM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    bufs.push_back({p, bytes});
    device_bytes += (int64_t)bytes;
    *out = p;
    return FVDBM_OK;
  }

  template <typename V>
  int upload(void **out, const std::vector<V> &v) {
    int rc = alloc(out, v.size() * sizeof(V));
    if (rc) return rc;
    if (!v.empty()) HIP_TRY(hipMemcpy(*out, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    return FVDBM_OK;
  }
};

template <typename T>
static StepArgs<T> make_args(const Handle &h, int src) {
  StepArgs<T> a{};
  a.fsrc = static_cast<const T *>(h.f[src]);
  a.fdst = static_cast<T *>(h.f[src ^ 1]);
  a.geo = h.geo;
  a.lwa = static_cast<const T *>(h.lwa);
  a.lwden2 = static_cast<const T *>(h.lwden2);
  a.bf_slot = static_cast<const int32_t *>(h.bf_slot);
  a.bf_gnode = static_cast<const int32_t *>(h.bf_gnode);
  a.tile_node_off = static_cast<const int32_t *>(h.tile_node_off);
  a.tile_nodes = static_cast<const int32_t *>(h.tile_nodes);
  a.tile_pair_off = static_cast<const int32_t *>(h.tile_pair_off);
  a.entry_pair_off = static_cast<const int32_t *>(h.entry_pair_off);
  a.pair_cell = static_cast<const int32_t *>(h.pair_cell);
  a.pair_w = static_cast<const T *>(h.pair_w);
  a.bf_ratio = static_cast<const T *>(h.bf_ratio);
  a.an_type = static_cast<const int32_t *>(h.an_type);
  a.an_cells = static_cast<const int32_t *>(h.an_cells);
  a.an_w = static_cast<const T *>(h.an_w);
  a.an_wsum = static_cast<const T *>(h.an_wsum);
  a.an_rho = static_cast<const T *>(h.an_rho);
  a.an_ux = static_cast<const T *>(h.an_ux);
  a.an_uy = static_cast<const T *>(h.an_uy);
  a.node_pdf = static_cast<const T *>(h.node_pdf);
  a.partials = static_cast<double *>(h.partials);
  a.block_tile = static_cast<const int32_t *>(h.block_tile);
  a.tile_count = static_cast<const int32_t *>(h.tile_count);
  a.batch = h.batch;
  a.bstep = h.bstep;
  a.an_case = static_cast<const int32_t *>(h.an_case);
  a.tile_case = static_cast<const int32_t *>(h.tile_case);
  a.case_active = static_cast<const int32_t *>(h.case_active);
  a.case_cur = static_cast<const int32_t *>(h.case_cur);
  a.case_invtau = static_cast<const T *>(h.case_invtau);
  a.case_dt = static_cast<const T *>(h.case_dt);
  a.f0 = static_cast<const T *>(h.f[0]);
  a.f1 = static_cast<T *>(h.f[1]);
  a.halo_off = static_cast<const int32_t *>(h.halo_off);
  a.halo_cells = static_cast<const int32_t *>(h.halo_cells);
  a.max_halo = h.max_halo;
  a.C = h.C;
  a.Cp = h.Cp;
  a.N = h.N;
  a.K = h.K;
  a.nbf = h.nbf;
  a.nan = h.nan;
  a.ntiles = h.ntiles;
  a.xcd_remap = (h.variant & 1) ? 0 : 1;
  a.inline_nodes = (h.variant & 2) ? 0 : 1;
  a.stage = (h.variant & 4) ? 0 : 1;
  a.max_tile_nodes = h.max_tile_nodes;
  a.max_tile_pairs = h.max_tile_pairs;
  a.k = make_lattice<T>(h.tau, h.dt);
  return a;
}

// dynamic LDS of k_step: node table, pair contributions, pdf stage, face table
static size_t lds_bytes(const Handle &h, size_t es) {
  const bool stage = (h.variant & 4) == 0;
  size_t n = ((size_t)Q * h.max_tile_nodes + kPairW * (size_t)h.max_tile_pairs + 7) & ~(size_t)7;
  if (stage) n += (size_t)9 * (h.block + h.max_halo);
  return n * es;
}

static int grid_for(const Handle &h, int remap) {
  return remap ? ((h.ntiles + 7) / 8) * 8 : h.ntiles;
}

// Launch one step reading f[src] and writing f[src ^ 1] (or dst); appends
// to the stream.
template <typename T>
static int launch_step(Handle &h, int src, bool resid, void *dst = nullptr) {
  StepArgs<T> a = make_args<T>(h, src);
  if (dst) a.fdst = static_cast<T *>(dst);
  if (!a.inline_nodes && h.nan > 0) {
    // two-kernel variant: refresh active node pdfs first
    const int nb = (h.nan + 255) / 256;
    hipLaunchKernelGGL((k_nodes<T>), dim3(nb), dim3(256), 0, h.stream, a,
                       static_cast<const int32_t *>(h.an_gid), static_cast<T *>(h.node_pdf),
                       static_cast<T *>(h.node_rho), static_cast<T *>(h.node_vel));
  }
  const dim3 grid(grid_for(h, a.xcd_remap)), block(h.block);
  const size_t lds = lds_bytes(h, sizeof(T));
  const int key = (h.scheme == FVDBM_UPWIND ? 0 : 2) + (resid ? 1 : 0);
  switch (key) {
    case 0: hipLaunchKernelGGL((k_step<T, 0, false>), grid, block, lds, h.stream, a); break;
    case 1: hipLaunchKernelGGL((k_step<T, 0, true>), grid, block, lds, h.stream, a); break;
    case 2: hipLaunchKernelGGL((k_step<T, 1, false>), grid, block, lds, h.stream, a); break;
    default: hipLaunchKernelGGL((k_step<T, 1, true>), grid, block, lds, h.stream, a); break;
  }
  HIP_TRY(hipGetLastError());
  return FVDBM_OK;
}

static int launch_step_any(Handle &h, int src, bool resid) {
  return h.precision == FVDBM_FP64 ? launch_step<double>(h, src, resid)
                                   : launch_step<float>(h, src, resid);
}

template <typename T>
static int launch_pair_typed(Handle &h, int src) {
  StepArgs<T> a = make_args<T>(h, src);
  PairArgs<T> p{};
  p.ring_off = static_cast<const int32_t *>(h.ring_off);
  p.ring_cells = static_cast<const int32_t *>(h.ring_cells);
  p.ring_n1 = static_cast<const int32_t *>(h.ring_n1);
  p.rg_off = static_cast<const int32_t *>(h.rg_off);
  p.rg_geo = h.rg_geo;
  p.geo2 = h.geo2;
  p.gh_off = static_cast<const int32_t *>(h.gh_off);
  p.gh_nown = static_cast<const int32_t *>(h.gh_nown);
  p.gh_s0 = static_cast<const int32_t *>(h.gh_s0);
  p.gh_s1 = static_cast<const int32_t *>(h.gh_s1);
  p.gh_self = static_cast<const int32_t *>(h.gh_self);
  p.gh_ratio = static_cast<const T *>(h.gh_ratio);
  p.xnode_off = static_cast<const int32_t *>(h.xnode_off);
  p.xnodes = static_cast<const int32_t *>(h.xnodes);
  p.xentry_pair_off = static_cast<const int32_t *>(h.xentry_pair_off);
  p.xpair_off = static_cast<const int32_t *>(h.xpair_off);
  p.xpair_cell = static_cast<const int32_t *>(h.xpair_cell);
  p.xpair_w = static_cast<const T *>(h.xpair_w);
  p.pair_slot = static_cast<const int32_t *>(h.pair_slot);
  p.nrg_cells = h.nrg_cells;
  p.max_xnodes = h.max_xnodes;
  p.max_xpairs = h.max_xpairs;
  if (h.npair_class == 0) {
    hipLaunchKernelGGL((k_step2<T>), dim3(grid_for(h, a.xcd_remap)), dim3(h.block), h.lds2,
                       h.stream, a, p);
    HIP_TRY(hipGetLastError());
    return FVDBM_OK;
  }
  // the classes partition the tiles, and a tile reads only fsrc and writes
  // only its own cells of fdst: two launches are one two-step launch, and
  // they may run at the same time. The heavy class goes first, on stream2,
  // so its few long tiles overlap the light class instead of trailing it.
  if (h.stream2) {
    HIP_TRY(hipEventRecord(h.evf, h.stream));
    HIP_TRY(hipStreamWaitEvent(h.stream2, h.evf, 0));
  }
  for (int k = h.npair_class - 1; k >= 0; --k) {
    a.block_tile = static_cast<const int32_t *>(h.pair_bt[k]);
    p.max_xnodes = h.pair_xnodes[k];
    p.max_xpairs = h.pair_xpairs[k];
    hipStream_t s = (k == 1 && h.stream2) ? h.stream2 : h.stream;
    hipLaunchKernelGGL((k_step2<T>), dim3(h.pair_grid[k]), dim3(h.block), h.pair_lds[k], s, a, p);
    HIP_TRY(hipGetLastError());
  }
  if (h.stream2) {
    HIP_TRY(hipEventRecord(h.evj, h.stream2));
    HIP_TRY(hipStreamWaitEvent(h.stream, h.evj, 0));
  }
  return FVDBM_OK;
}

// Two steps reading f[src] and writing f[src ^ 1].
static int launch_pair(Handle &h, int src) {
  return h.precision == FVDBM_FP64 ? launch_pair_typed<double>(h, src)
                                   : launch_pair_typed<float>(h, src);
}

static int launch_batch_advance(Handle &h, int n, int swaps) {
  hipLaunchKernelGGL(k_batch_advance, dim3((h.ncases + 255) / 256), dim3(256), 0, h.stream,
                     h.ncases, static_cast<const int32_t *>(h.case_active),
                     static_cast<int32_t *>(h.case_cur), static_cast<int64_t *>(h.case_steps), n,
                     swaps);
  HIP_TRY(hipGetLastError());
  return FVDBM_OK;
}

// Batched: the n steps of an interval (the residual on the last when
// asked), as two-step launches where the handle has them. Launch i runs
// with h.bstep = i and swaps the buffers once; *swaps = launches made.
static int launch_batch_steps(Handle &h, int n, bool resid_last, int *swaps) {
  int rc = FVDBM_OK;
  int j = 0, i = 0;
  const int plain_end = n - (resid_last ? 1 : 0);
  if (h.pairs)
    for (; j + 2 <= plain_end && rc == FVDBM_OK; j += 2, ++i) {
      h.bstep = i;
      rc = launch_pair(h, 0);
    }
  for (; j < n && rc == FVDBM_OK; ++j, ++i) {
    h.bstep = i;
    rc = launch_step_any(h, 0, resid_last && j == n - 1);
  }
  h.bstep = 0;
  *swaps = i;
  return rc;
}

// Batched: an interval of n steps (the residual on the last when asked)
// followed by the parity/step advance of active cases.
static int launch_batch_interval(Handle &h, int n, bool resid_last) {
  int swaps = 0;
  const int rc = launch_batch_steps(h, n, resid_last, &swaps);
  return rc ? rc : launch_batch_advance(h, n, swaps);
}

static int launch_batch_check(Handle &h, double tol, double growth, int every, int swaps) {
  hipLaunchKernelGGL(k_batch_check, dim3(1), dim3(256), 0, h.stream, h.ncases,
                     static_cast<const int32_t *>(h.case_tile_off),
                     static_cast<const double *>(h.partials), tol, growth,
                     static_cast<int32_t *>(h.case_active), static_cast<int32_t *>(h.case_status),
                     static_cast<double *>(h.case_resid), static_cast<double *>(h.case_best),
                     static_cast<int32_t *>(h.n_active), static_cast<int32_t *>(h.case_cur),
                     static_cast<int64_t *>(h.case_steps), every, swaps);
  HIP_TRY(hipGetLastError());
  return FVDBM_OK;
}

// One check interval of a batched run as a graph: (every - 1) plain steps,
// then a residual step and the per-case convergence check.
static int ensure_batch_graph(Handle &h, int every, double tol, double growth) {
  if (h.bexec && h.bgraph_every == every && h.bgraph_tol == tol && h.bgraph_growth == growth)
    return FVDBM_OK;
  if (h.bexec) {
    (void)hipGraphExecDestroy(h.bexec);
    h.bexec = nullptr;
  }
  if (h.bgraph) {
    (void)hipGraphDestroy(h.bgraph);
    h.bgraph = nullptr;
  }
  HIP_TRY(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
  int swaps = 0;
  int rc = launch_batch_steps(h, every, true, &swaps);
  if (!rc) rc = launch_batch_check(h, tol, growth, every, swaps);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h.stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(FVDBM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  h.bgraph = g;
  HIP_TRY(hipGraphInstantiate(&h.bexec, g, nullptr, nullptr, 0));
  h.bgraph_every = every;
  h.bgraph_tol = tol;
  h.bgraph_growth = growth;
  return FVDBM_OK;
}

static int ensure_graph(Handle &h, int parity) {
  if (h.gexec[parity]) return FVDBM_OK;
  HIP_TRY(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
  int src = parity;
  int rc = FVDBM_OK;
  for (int s = 0; s < h.graph_steps && rc == FVDBM_OK; ++s) {
    rc = launch_step_any(h, src, false);
    src ^= 1;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h.stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(FVDBM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  h.graph[parity] = g;
  HIP_TRY(hipGraphInstantiate(&h.gexec[parity], g, nullptr, nullptr, 0));
  return FVDBM_OK;
}

// Two-step tiles: a captured run of `launches` pair launches starting from
// `parity` (cached per (parity, launches)); *out is its executable graph.
static int ensure_graph2(Handle &h, int parity, int launches, hipGraphExec_t *out) {
  const auto key = std::make_pair(parity, launches);
  auto it = h.graphs2.find(key);
  if (it != h.graphs2.end()) {
    *out = it->second.second;
    return FVDBM_OK;
  }
  HIP_TRY(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
  int src = parity;
  int rc = FVDBM_OK;
  for (int s = 0; s < launches && rc == FVDBM_OK; ++s) {
    rc = launch_pair(h, src);
    src ^= 1;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h.stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(FVDBM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(FVDBM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  h.graphs2[key] = std::make_pair(g, x);
  *out = x;
  return FVDBM_OK;
}

// `launches` pair launches from f[h.cur]: captured runs of graph_steps / 2
// launches, then one captured run for the rest (a single launch goes
// direct). With dry, nothing is launched and only the graphs such a run
// would replay are built (so a timed run captures nothing).
static int run_pair_launches(Handle &h, int64_t launches, bool dry) {
  int cur = h.cur;
  const int per = h.graph_steps >= 2 ? h.graph_steps / 2 : 0;
  hipGraphExec_t x = nullptr;
  while (per > 0 && launches > 0) {
    const int nl = (int)std::min<int64_t>(launches, per);
    if (nl < 2) break;
    int rc = ensure_graph2(h, cur, nl, &x);
    if (rc) return rc;
    if (!dry) {
      HIP_TRY(hipGraphLaunch(x, h.stream));
      h.steps += 2 * (int64_t)nl;
    }
    cur ^= nl & 1;
    launches -= nl;
  }
  for (; launches > 0; --launches) {
    if (!dry) {
      int rc = launch_pair(h, cur);
      if (rc) return rc;
      h.steps += 2;
    }
    cur ^= 1;
  }
  if (!dry) h.cur = cur;
  return FVDBM_OK;
}

// Batched: a graph of graph_steps steps and the advance, replayed.
static int ensure_batch_step_graph(Handle &h) {
  if (h.sexec) return FVDBM_OK;
  HIP_TRY(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
  int rc = launch_batch_interval(h, h.graph_steps, false);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h.stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(FVDBM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  h.sgraph = g;
  HIP_TRY(hipGraphInstantiate(&h.sexec, g, nullptr, nullptr, 0));
  return FVDBM_OK;
}

static int run_batch_steps(Handle &h, int64_t n, bool resid_last) {
  int64_t rest = n;
  if (h.graph_steps > 0) {
    while (rest - (resid_last ? 1 : 0) >= h.graph_steps) {
      int rc = ensure_batch_step_graph(h);
      if (rc) return rc;
      HIP_TRY(hipGraphLaunch(h.sexec, h.stream));
      rest -= h.graph_steps;
    }
  }
  while (rest > 0) {
    const int chunk = (int)std::min<int64_t>(rest, 1024);
    int rc = launch_batch_interval(h, chunk, resid_last && chunk == rest);
    if (rc) return rc;
    rest -= chunk;
  }
  h.steps += n;
  h.has_prev = true;
  h.observed = false;
  return FVDBM_OK;
}

// n steps (the residual on the last when asked). With dry (no residual),
// only the graphs such a run replays are built.
static int run_steps(Handle &h, int64_t n, bool resid_last, bool dry = false) {
  if (n <= 0) return FVDBM_OK;
  if (h.batch) return dry ? FVDBM_OK : run_batch_steps(h, n, resid_last);
  if (h.pairs && !resid_last && n >= 2) {
    // two-step launches only; an odd count starts with one single step.
    // f[cur ^ 1] then holds the pdf two steps back (see prev_two_back)
    const int cur0 = h.cur;
    if (n & 1) {
      if (!dry) {
        int rc = launch_step_any(h, h.cur, false);
        if (rc) return rc;
        ++h.steps;
      }
      h.cur ^= 1;
      --n;
    }
    int rc = run_pair_launches(h, n / 2, dry);
    if (dry) {
      h.cur = cur0;
      return rc;
    }
    if (rc) return rc;
    h.prev_two_back = true;
    h.mid_ready = false;
    h.has_prev = true;
    h.observed = false;
    return FVDBM_OK;
  }
  if (dry) return FVDBM_OK;
  if (h.pairs && n >= 3) {
    // residual runs: pairs of steps for all but the last one or two, so the
    // last (residual) step is a single launch and f[cur ^ 1] keeps the pdf
    // at its start
    const int64_t launches = (n - 1) / 2;
    int rc = run_pair_launches(h, launches, false);
    if (rc) return rc;
    n -= 2 * launches;  // the remaining single steps, the last included
  }
  int64_t plain = resid_last ? n - 1 : n;
  if (h.graph_steps > 0 && !h.batch) {
    while (plain >= h.graph_steps) {
      int rc = ensure_graph(h, h.cur);
      if (rc) return rc;
      HIP_TRY(hipGraphLaunch(h.gexec[h.cur], h.stream));
      plain -= h.graph_steps;  // graph_steps is even: parity unchanged
      h.steps += h.graph_steps;
    }
  }
  for (; plain > 0; --plain) {
    int rc = launch_step_any(h, h.cur, false);
    if (rc) return rc;
    h.cur ^= 1;
    ++h.steps;
  }
  if (resid_last) {
    int rc = launch_step_any(h, h.cur, true);
    if (rc) return rc;
    h.cur ^= 1;
    ++h.steps;
  }
  h.prev_two_back = false;
  h.has_prev = true;
  h.observed = false;
  return FVDBM_OK;
}

// LDS of a two-step launch over a set of tiles: the kernel partitions it
// with the set's largest node table (rows), largest pair count (extended or
// own table) and largest ring plus ghost records: [9][rows] table, pair
// records rounded to 8 words, then the stage of block + ring + ghost cell
// records.
struct PairLds {
  int rows = 0, pairs = 0, ring = 0;
  template <typename T>
  void add(const HostLayout<T> &L, int t, int ghosts) {
    rows = std::max(rows, L.xnode_off[t + 1] - L.xnode_off[t]);
    pairs = std::max(pairs, std::max(L.xpair_off[t + 1] - L.xpair_off[t],
                                     L.tile_pair_off[t + 1] - L.tile_pair_off[t]));
    ring = std::max(ring, L.ring_off[t + 1] - L.ring_off[t] + ghosts);
  }
  size_t bytes(int block, size_t es) const {
    const size_t pre = ((size_t)Q * rows + kPairW * (size_t)pairs + 7) & ~(size_t)7;
    return (pre + (size_t)9 * (block + ring)) * es;
  }
};

// LDS classes of the two-step launch. One launch sizes its LDS for the
// largest tile; where that LDS, not the registers, limits the resident
// blocks per CU (boundary-heavy meshes: a few tiles with large node tables),
// the tiles that fit the register-limited share of the CU's LDS run in a
// launch of their own with that smaller size, the rest in a second launch.
// Each class keeps the XCD-aware order (boundary tiles first per range).
// Without occupancy information the single launch stays.
template <typename T>
static int build_pair_classes(Handle &h, const HostLayout<T> &L,
                              const std::vector<int32_t> &ghosts) {
  int dev = 0, lds_cu = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(k_step2<T>),
                                                   h.block, 0) != hipSuccess ||
      nb <= 0 || lds_cu <= 0) {
    (void)hipGetLastError();  // leave no error for the next launch check to report
    return FVDBM_OK;
  }
  const size_t es = sizeof(T);
  const size_t budget = ((size_t)lds_cu / (size_t)nb) & ~(size_t)1023;
  if (h.lds2 <= budget) return FVDBM_OK;  // LDS does not limit the occupancy
  std::vector<size_t> need(L.ntiles);
  for (int t = 0; t < L.ntiles; ++t) {
    PairLds one;
    one.add(L, t, ghosts[t]);
    need[t] = one.bytes(h.block, es);
  }
  // the light class: tiles up to a threshold, lowered until the class's
  // LDS (from its component maxima) fits the budget
  size_t thr = budget;
  PairLds cls[2];
  int n_light = 0;
  for (;;) {
    cls[0] = PairLds();
    cls[1] = PairLds();
    n_light = 0;
    for (int t = 0; t < L.ntiles; ++t) {
      const int k = need[t] <= thr ? 0 : 1;
      cls[k].add(L, t, ghosts[t]);
      n_light += k == 0;
    }
    if (cls[0].bytes(h.block, es) <= budget || thr < 1024) break;
    thr -= 512;
  }
  if (n_light == 0 || n_light == L.ntiles || cls[0].bytes(h.block, es) > budget ||
      cls[1].bytes(h.block, es) > 65536)
    return FVDBM_OK;
  const int per = (L.ntiles + 7) / 8;
  for (int k = 0; k < 2; ++k) {
    std::vector<int32_t> order[8];
    size_t longest = 0;
    for (int x = 0; x < 8; ++x) {
      const int lo = std::min(x * per, L.ntiles), hi = std::min((x + 1) * per, L.ntiles);
      for (int pass = 0; pass < 2; ++pass)  // tiles with boundary nodes first
        for (int t = lo; t < hi; ++t) {
          const bool nodes = L.tile_node_off[t + 1] > L.tile_node_off[t];
          if (nodes == (pass == 0) && (need[t] <= thr) == (k == 0)) order[x].push_back(t);
        }
      longest = std::max(longest, order[x].size());
    }
    std::vector<int32_t> bt(8 * longest, -1);
    for (int x = 0; x < 8; ++x)
      for (size_t i = 0; i < order[x].size(); ++i) bt[i * 8 + x] = order[x][i];
    const int rc = h.upload(&h.pair_bt[k], bt);
    if (rc) return rc;
    h.pair_grid[k] = (int)bt.size();
    h.pair_lds[k] = cls[k].bytes(h.block, es);
    h.pair_xnodes[k] = cls[k].rows;
    h.pair_xpairs[k] = cls[k].pairs;
  }
  HIP_TRY(hipStreamCreateWithFlags(&h.stream2, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&h.evf, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&h.evj, hipEventDisableTiming));
  h.npair_class = 2;
  return FVDBM_OK;
}

template <typename T>
static int create_typed(const fvdbm_mesh_desc &m, Handle &h, const int32_t *cell_case = nullptr,
                        int ncases = 1) {
  HostLayout<T> L;
  std::string err;
  if (!build_layout<T>(m, h.reorder, h.block, L, err, cell_case, ncases))
    return fail(FVDBM_ERR_INVALID, err);
  h.C = L.C;
  h.Cp = L.Cp;
  h.F = L.F;
  h.N = L.N;
  h.K = L.K;
  h.nbc = L.n_bcells;
  h.nbf = L.nbf;
  h.nan = L.nan;
  h.bandwidth = L.bandwidth;
  h.last_cell_new = L.last_cell_new;
  // raw buffer descriptors address at most 2^32 - 1 bytes
  if ((uint64_t)Q * L.Cp * sizeof(T) >= (1ull << 32) ||
      (uint64_t)3 * L.Cp * GeoLayout<T>::bytes >= (1ull << 32))
    return fail(FVDBM_ERR_INVALID, "mesh too large for one handle (pdf array must be < 4 GiB)");
  h.ntiles = L.ntiles;
  h.max_tile_nodes = L.max_tile_nodes;
  h.max_tile_pairs = L.max_tile_pairs;
  // LDS per block is capped at 64 KiB: node tables beyond 40 KiB use the
  // two-kernel variant, and the pdf staging is dropped when it does not fit
  if (((size_t)Q * L.max_tile_nodes + kPairW * (size_t)L.max_tile_pairs) * sizeof(T) > 40960) {
    h.variant |= 2;
    h.max_tile_nodes = 0;
    h.max_tile_pairs = 0;
  }
  h.max_halo = L.max_halo;
  if (lds_bytes(h, sizeof(T)) > 65536) h.variant |= 4;  // the stage does not fit
  const bool staged = (h.variant & 4) == 0;
  h.new_to_old = L.new_to_old;
  h.an_gid_host = L.an_gid;
  h.an_type_host = L.an_type;
  const size_t es = sizeof(T);
  int rc = 0;
  for (int i = 0; i < 2 && !rc; ++i) rc = h.alloc(&h.f[i], (size_t)Q * h.Cp * es);
  if (!rc) {
    // interleave (code, nx, ny, sL) per cell face into GeoLayout<T> records
    const size_t rec = GeoLayout<T>::bytes, n = 3 * (size_t)L.Cp;
    std::vector<unsigned char> g(n * rec, 0);
    const std::vector<int32_t> &codes = staged ? L.nb_slot : L.nb;
    for (size_t i = 0; i < n; ++i) {
      unsigned char *d = g.data() + i * rec;
      std::memcpy(d, &codes[i], 4);
      const size_t o = sizeof(T) == 4 ? 4 : 8;
      std::memcpy(d + o, &L.nx[i], sizeof(T));
      std::memcpy(d + o + sizeof(T), &L.ny[i], sizeof(T));
      std::memcpy(d + o + 2 * sizeof(T), &L.sL[i], sizeof(T));
    }
    rc = h.upload(&h.geo, g);
  }
  if (!rc && h.scheme == FVDBM_LAX_WENDROFF) rc = h.upload(&h.lwa, L.lwa);
  if (!rc && h.scheme == FVDBM_LAX_WENDROFF) rc = h.upload(&h.lwden2, L.lwden2);
  if (!rc) rc = h.upload(&h.bf_slot, L.bf_slot);
  if (!rc) rc = h.upload(&h.bf_gnode, L.bf_gnode);
  if (!rc) rc = h.upload(&h.tile_node_off, L.tile_node_off);
  if (!rc) rc = h.upload(&h.tile_nodes, L.tile_nodes);
  if (!rc) rc = h.upload(&h.tile_pair_off, L.tile_pair_off);
  if (!rc) rc = h.upload(&h.entry_pair_off, L.entry_pair_off);
  if (!rc) rc = h.upload(&h.pair_cell, L.pair_cell);
  if (!rc) rc = h.upload(&h.pair_w, L.pair_w);
  if (!rc) rc = h.upload(&h.tile_count, L.tile_count);
  if (!rc && cell_case) {
    h.batch = 1;
    h.ncases = L.ncases;
    h.case_tile_off_host = L.case_tile_off;
    rc = h.upload(&h.tile_case, L.tile_case);
    if (!rc) rc = h.upload(&h.case_tile_off, L.case_tile_off);
    std::vector<int32_t> ones(h.ncases, 1), zeros(h.ncases, 0);
    std::vector<int64_t> z64(h.ncases, 0);
    std::vector<double> zd(h.ncases, 0.0), inf(h.ncases, INFINITY);
    std::vector<T> it(h.ncases), dtv(h.ncases);
    for (int k = 0; k < h.ncases; ++k) {
      it[k] = T(1.0 / h.case_tau_host[k]);  // 1/tau in double, then rounded
      dtv[k] = T(h.case_dt_host[k]);
    }
    if (!rc) rc = h.upload(&h.case_active, ones);
    if (!rc) rc = h.upload(&h.case_cur, zeros);
    if (!rc) rc = h.upload(&h.case_steps, z64);
    if (!rc) rc = h.upload(&h.case_status, zeros);
    if (!rc) rc = h.upload(&h.case_resid, zd);
    if (!rc) rc = h.upload(&h.case_best, inf);
    if (!rc) rc = h.upload(&h.case_invtau, it);
    if (!rc) rc = h.upload(&h.case_dt, dtv);
    if (!rc) {
      // each active node's case: that of its first listed cell
      std::vector<int32_t> anc(std::max<size_t>(L.an_gid.size(), 1), 0);
      for (size_t i = 0; i < L.an_gid.size(); ++i) {
        const int c0 = m.node_cells[(size_t)L.an_gid[i] * m.k_max];
        anc[i] = c0 >= 0 ? cell_case[c0] : 0;
      }
      rc = h.upload(&h.an_case, anc);
    }
    if (!rc) rc = h.alloc(&h.n_active, sizeof(int32_t));
  }
  if (!rc) rc = h.upload(&h.halo_off, L.halo_off);
  if (!rc) rc = h.upload(&h.halo_cells, L.halo_cells);
  if (!rc) {
    // block -> tile for the XCD-aware launch: block b runs on XCD b % 8 and
    // takes the (b / 8)-th tile of that XCD's contiguous range, boundary
    // tiles of the range first
    const int per = (L.ntiles + 7) / 8;
    std::vector<int32_t> bt(8 * (size_t)per, -1);
    for (int x = 0; x < 8; ++x) {
      const int lo = std::min(x * per, L.ntiles), hi = std::min((x + 1) * per, L.ntiles);
      std::vector<int32_t> order;
      for (int t = lo; t < hi; ++t)
        if (L.tile_node_off[t + 1] > L.tile_node_off[t]) order.push_back(t);
      for (int t = lo; t < hi; ++t)
        if (L.tile_node_off[t + 1] == L.tile_node_off[t]) order.push_back(t);
      for (size_t i = 0; i < order.size(); ++i) bt[i * 8 + x] = order[i];
    }
    rc = h.upload(&h.block_tile, bt);
  }
  if (!rc) rc = h.upload(&h.bf_ratio, L.bf_ratio);
  if (!rc) rc = h.upload(&h.an_gid, L.an_gid);
  if (!rc) rc = h.upload(&h.an_type, L.an_type);
  if (!rc) rc = h.upload(&h.an_cells, L.an_cells);
  if (!rc) rc = h.upload(&h.an_w, L.an_w);
  if (!rc) rc = h.upload(&h.an_wsum, L.an_wsum);
  if (!rc) rc = h.upload(&h.an_rho, L.an_rho);
  if (!rc) rc = h.upload(&h.an_ux, L.an_ux);
  if (!rc) rc = h.upload(&h.an_uy, L.an_uy);
  if (!rc) rc = h.alloc(&h.node_pdf, (size_t)Q * h.N * es);
  if (!rc) rc = h.alloc(&h.node_rho, (size_t)h.N * es);
  if (!rc) rc = h.alloc(&h.node_vel, (size_t)2 * h.N * es);
  if (!rc) rc = h.upload(&h.fc0, L.fc0);
  if (!rc) rc = h.upload(&h.fc1, L.fc1);
  if (!rc) rc = h.upload(&h.fn0, L.fn0);
  if (!rc) rc = h.upload(&h.fn1, L.fn1);
  if (!rc) rc = h.upload(&h.fd0, L.fd0);
  if (!rc) rc = h.upload(&h.fd1, L.fd1);
  if (!rc) rc = h.upload(&h.fnx, L.fnx);
  if (!rc) rc = h.upload(&h.fny, L.fny);
  if (!rc) rc = h.upload(&h.fL, L.fL);
  if (!rc) rc = h.alloc(&h.partials, (size_t)2 * h.ntiles * sizeof(double));
  if (!rc) rc = h.alloc(&h.resid, 2 * sizeof(double));
  if (!rc) {
    // node lists with cells as device slots (all nodes, for the all-node
    // interpolation), and each cell's vertices in ascending node id
    const int K = m.k_max;
    const T *nd = static_cast<const T *>(m.node_dists);
    std::vector<int32_t> nc((size_t)h.N * K);
    std::vector<T> ndv((size_t)h.N * K);
    std::vector<int32_t> cnt((size_t)h.Cp + 1, 0);
    for (size_t i = 0; i < nc.size(); ++i) {
      const int c = m.node_cells[i];
      nc[i] = c >= 0 ? L.old_to_new[c] : -1;
      ndv[i] = nd[i];
      if (c >= 0) cnt[(size_t)nc[i] + 1] += 1;
    }
    for (int c = 0; c < h.Cp; ++c) cnt[c + 1] += cnt[c];
    std::vector<int32_t> cn_node(cnt[h.Cp]), fill(cnt.begin(), cnt.end() - 1);
    std::vector<T> cn_d(cnt[h.Cp]);
    for (int n = 0; n < h.N; ++n)
      for (int k = 0; k < K; ++k) {
        const int s = nc[(size_t)n * K + k];
        if (s < 0) continue;
        cn_node[fill[s]] = n;
        cn_d[fill[s]++] = ndv[(size_t)n * K + k];
      }
    rc = h.upload(&h.all_nc, nc);
    if (!rc) rc = h.upload(&h.all_nd, ndv);
    if (!rc) rc = h.upload(&h.cn_off, cnt);
    if (!rc) rc = h.upload(&h.cn_node, cn_node);
    if (!rc) rc = h.upload(&h.cn_d, cn_d);
  }
  if (!rc && !(h.variant & 8) && h.scheme == FVDBM_UPWIND && !(h.variant & 6) && L.pair_ok) {
    // two-step tiles, when their LDS fits a block. Ghost lists: per tile the
    // ghost faces of the own cells (device order, faces 0..2: ascending
    // boundary-face index), then those of the S1 \ O cells (ring order);
    // ghost j of tile t is staged at slot block + |ring of t| + j, and the
    // two-step face records name it there instead of by its ghost bit.
    const int bs = h.block;
    const size_t Cp = (size_t)L.Cp;
    std::vector<int32_t> gh_off(L.ntiles + 1, 0), gh_nown(L.ntiles, 0), gh_s0, gh_s1, gh_self;
    std::vector<T> gh_ratio;
    std::vector<int32_t> code2(L.nb_slot), rgc(L.rg_code), gh_count(L.ntiles, 0);
    int max_stage = 0;  // largest ring + ghost records of a tile
    for (int t = 0, rc0 = 0; t < L.ntiles; ++t) {
      const int nr = L.ring_off[t + 1] - L.ring_off[t], n1 = L.ring_n1[t];
      gh_off[t] = (int32_t)gh_s0.size();
      int j = 0;
      for (int i = t * bs; i < t * bs + L.tile_count[t]; ++i)
        for (int k = 0; k < 3; ++k) {
          const size_t at = (size_t)k * Cp + i;
          const int32_t code = L.nb_slot[at];
          if (code >= 0) continue;
          const int bf = code & kIndexMask;
          gh_s0.push_back(L.bf_slot[bf]);
          gh_s1.push_back(L.bf_slot[(size_t)L.nbf + bf]);
          gh_ratio.push_back(L.bf_ratio[bf]);
          gh_self.push_back(i - t * bs);
          code2[at] = (code & kSlot1Bit) | (bs + nr + j);
          ++j;
        }
      gh_nown[t] = j;
      for (int r = 0; r < n1; ++r, ++rc0)
        for (int k = 0; k < 3; ++k) {
          const size_t si = 3 * (size_t)rc0 + k;
          const int32_t code = L.rg_code[si];
          if (code >= 0) continue;
          const int b = code & kIndexMask;
          gh_s0.push_back(L.rgb_slot[b]);
          gh_s1.push_back(L.rgb_slot[(size_t)L.nrg + b]);
          gh_ratio.push_back(L.rgb_ratio[b]);
          gh_self.push_back(bs + r);
          rgc[si] = (code & kSlot1Bit) | (bs + nr + j);
          ++j;
        }
      gh_count[t] = j;
      max_stage = std::max(max_stage, nr + j);
    }
    gh_off[L.ntiles] = (int32_t)gh_s0.size();
    const size_t pre2 = ((size_t)Q * L.max_xnodes + kPairW * (size_t)L.max_xpairs + 7) & ~(size_t)7;
    const size_t lds2 = (pre2 + (size_t)9 * (bs + max_stage)) * es;
    if (lds2 <= 65536) {
      const int nrgc = (int)(L.rg_code.size() / 3);
      const size_t rec = GeoLayout<T>::bytes;
      const size_t o = sizeof(T) == 4 ? 4 : 8;
      std::vector<unsigned char> g(3 * (size_t)std::max(nrgc, 1) * rec, 0);
      for (int r = 0; r < nrgc; ++r)
        for (int k = 0; k < 3; ++k) {
          unsigned char *d = g.data() + ((size_t)k * nrgc + r) * rec;
          const size_t si = 3 * (size_t)r + k;
          std::memcpy(d, &rgc[si], 4);
          std::memcpy(d + o, &L.rg_nx[si], sizeof(T));
          std::memcpy(d + o + sizeof(T), &L.rg_ny[si], sizeof(T));
          std::memcpy(d + o + 2 * sizeof(T), &L.rg_sL[si], sizeof(T));
        }
      std::vector<unsigned char> g2(3 * Cp * rec, 0);
      for (size_t i = 0; i < 3 * Cp; ++i) {
        unsigned char *d = g2.data() + i * rec;
        std::memcpy(d, &code2[i], 4);
        std::memcpy(d + o, &L.nx[i], sizeof(T));
        std::memcpy(d + o + sizeof(T), &L.ny[i], sizeof(T));
        std::memcpy(d + o + 2 * sizeof(T), &L.sL[i], sizeof(T));
      }
      std::vector<int32_t> rg_off(L.ntiles, 0);
      for (int t = 1; t < L.ntiles; ++t) rg_off[t] = rg_off[t - 1] + L.ring_n1[t - 1];
      rc = h.upload(&h.rg_geo, g);
      if (!rc) rc = h.upload(&h.geo2, g2);
      if (!rc) rc = h.upload(&h.rg_off, rg_off);
      if (!rc) rc = h.upload(&h.ring_off, L.ring_off);
      if (!rc) rc = h.upload(&h.ring_cells, L.ring_cells);
      if (!rc) rc = h.upload(&h.ring_n1, L.ring_n1);
      if (!rc) rc = h.upload(&h.gh_off, gh_off);
      if (!rc) rc = h.upload(&h.gh_nown, gh_nown);
      if (!rc) rc = h.upload(&h.gh_s0, gh_s0);
      if (!rc) rc = h.upload(&h.gh_s1, gh_s1);
      if (!rc) rc = h.upload(&h.gh_self, gh_self);
      if (!rc) rc = h.upload(&h.gh_ratio, gh_ratio);
      if (!rc) rc = h.upload(&h.xnode_off, L.xnode_off);
      if (!rc) rc = h.upload(&h.xnodes, L.xnodes);
      if (!rc) rc = h.upload(&h.xentry_pair_off, L.xentry_pair_off);
      if (!rc) rc = h.upload(&h.xpair_off, L.xpair_off);
      if (!rc) rc = h.upload(&h.xpair_cell, L.xpair_cell);
      if (!rc) rc = h.upload(&h.xpair_w, L.xpair_w);
      if (!rc) rc = h.upload(&h.pair_slot, L.pair_slot);
      if (!rc) rc = h.alloc(&h.fmid, (size_t)Q * h.Cp * es);
      if (!rc) {
        h.pairs = true;
        h.lds2 = lds2;
        h.nrg_cells = nrgc;
        h.max_xnodes = L.max_xnodes;
        h.max_xpairs = L.max_xpairs;
      }
      if (!rc && !(h.variant & 1)) rc = build_pair_classes<T>(h, L, gh_count);
    }
  }
  if (!rc && !h.pairs) h.variant |= 8;  // one step per launch (two-step tiles off or unavailable)
  if (rc) return rc;
  // node rho/vel hold the boundary-condition values (reference node order)
  std::vector<T> rho(h.N), vel(2 * (size_t)h.N);
  const T *mr = static_cast<const T *>(m.node_rho);
  const T *mv = static_cast<const T *>(m.node_vel);
  for (int i = 0; i < h.N; ++i) {
    rho[i] = mr[i];
    vel[i] = mv[2 * (size_t)i];
    vel[(size_t)h.N + i] = mv[2 * (size_t)i + 1];
  }
  HIP_TRY(hipMemcpy(h.node_rho, rho.data(), rho.size() * es, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h.node_vel, vel.data(), vel.size() * es, hipMemcpyHostToDevice));
  return FVDBM_OK;
}

template <typename T>
static int init_uniform_typed(Handle &h) {
  Lattice<T> k = make_lattice<T>(h.tau, h.dt);
  // every slot, holes included (cells of later cases sit past slot C)
  for (int i = 0; i < 2; ++i)
    hipLaunchKernelGGL((k_fill_uniform<T>), dim3((h.Cp + 255) / 256), dim3(256), 0, h.stream,
                       static_cast<T *>(h.f[i]), h.Cp, h.Cp, k);
  // node pdf <- W (both type 0 and active nodes)
  hipLaunchKernelGGL((k_fill_uniform<T>), dim3((h.N + 255) / 256), dim3(256), 0, h.stream,
                     static_cast<T *>(h.node_pdf), h.N, h.N, k);
  HIP_TRY(hipGetLastError());
  if (h.batch) {
    std::vector<int32_t> ones(h.ncases, 1), zeros(h.ncases, 0);
    std::vector<int64_t> z64(h.ncases, 0);
    std::vector<double> inf(h.ncases, INFINITY);
    HIP_TRY(hipMemcpy(h.case_active, ones.data(), 4 * (size_t)h.ncases, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.case_cur, zeros.data(), 4 * (size_t)h.ncases, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.case_status, zeros.data(), 4 * (size_t)h.ncases, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.case_steps, z64.data(), 8 * (size_t)h.ncases, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.case_best, inf.data(), 8 * (size_t)h.ncases, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipStreamSynchronize(h.stream));
  h.cur = 0;
  h.steps = 0;
  h.has_prev = false;
  h.prev_two_back = false;
  h.observed = false;
  return FVDBM_OK;
}

// After a run that ended in a two-step launch: the pdf at the start of the
// last step, one step from f[cur ^ 1] (bitwise the two-step launch's
// intermediate state), into fmid. Done before anything the step reads
// (boundary-node values) changes, and at most once per run.
template <typename T>
static int materialize_mid(Handle &h) {
  if (!h.has_prev || !h.prev_two_back || h.mid_ready) return FVDBM_OK;
  const int rc = launch_step<T>(h, h.cur ^ 1, false, h.fmid);
  if (rc) return rc;
  h.mid_ready = true;
  return FVDBM_OK;
}

template <typename T>
static int observe_typed(Handle &h) {
  if (h.observed || !h.has_prev) return FVDBM_OK;
  int rc = 0;
  if (!h.obs_rho) rc = h.alloc(&h.obs_rho, (size_t)h.Cp * sizeof(T));
  if (!rc && !h.obs_vel) rc = h.alloc(&h.obs_vel, (size_t)2 * h.Cp * sizeof(T));
  if (!rc && !h.obs_feq) rc = h.alloc(&h.obs_feq, (size_t)Q * h.Cp * sizeof(T));
  if (!rc && !h.obs_flux) rc = h.alloc(&h.obs_flux, (size_t)Q * h.F * sizeof(T));
  if (rc) return rc;
  rc = materialize_mid<T>(h);
  if (rc) return rc;
  StepArgs<T> a = make_args<T>(h, h.cur ^ 1);  // the pdf at the start of the last step
  if (h.prev_two_back) a.fsrc = static_cast<const T *>(h.fmid);
  if (h.nan > 0)
    hipLaunchKernelGGL((k_nodes<T>), dim3((h.nan + 255) / 256), dim3(256), 0, h.stream, a,
                       static_cast<const int32_t *>(h.an_gid), static_cast<T *>(h.node_pdf),
                       static_cast<T *>(h.node_rho), static_cast<T *>(h.node_vel));
  hipLaunchKernelGGL((k_cell_moments<T>), dim3((h.C + 255) / 256), dim3(256), 0, h.stream, a,
                     static_cast<T *>(h.obs_rho), static_cast<T *>(h.obs_vel),
                     static_cast<T *>(h.obs_feq));
  FaceArrays fa;
  fa.c0 = static_cast<const int32_t *>(h.fc0);
  fa.c1 = static_cast<const int32_t *>(h.fc1);
  fa.n0 = static_cast<const int32_t *>(h.fn0);
  fa.n1 = static_cast<const int32_t *>(h.fn1);
  fa.d0 = h.fd0;
  fa.d1 = h.fd1;
  fa.nx = h.fnx;
  fa.ny = h.fny;
  fa.L = h.fL;
  fa.F = h.F;
  fa.last_cell = h.last_cell_new;
  if (h.scheme == FVDBM_UPWIND)
    hipLaunchKernelGGL((k_face_flux<T, 0>), dim3((h.F + 255) / 256), dim3(256), 0, h.stream, a,
                       fa, static_cast<T *>(h.obs_flux));
  else
    hipLaunchKernelGGL((k_face_flux<T, 1>), dim3((h.F + 255) / 256), dim3(256), 0, h.stream, a,
                       fa, static_cast<T *>(h.obs_flux));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h.stream));
  h.observed = true;
  return FVDBM_OK;
}

template <typename T>
static int get_typed(Handle &h, int field, void *dst) {
  if (h.batch && field != FVDBM_CELL_PDF)
    return fail(FVDBM_ERR_INVALID, "a batched handle exposes FVDBM_CELL_PDF only");
  T *out = static_cast<T *>(dst);
  const size_t es = sizeof(T);
  HIP_TRY(hipStreamSynchronize(h.stream));
  auto cell_q_major = [&](const void *dev, int comps, size_t stride) -> int {
    std::vector<T> tmp((size_t)comps * stride);
    hipError_t e = hipMemcpy(tmp.data(), dev, tmp.size() * es, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(FVDBM_ERR_HIP, hipGetErrorString(e));
    for (int i = 0; i < h.Cp; ++i) {
      const int o = h.new_to_old[i];
      if (o < 0) continue;
      for (int q = 0; q < comps; ++q) out[(size_t)o * comps + q] = tmp[(size_t)q * stride + i];
    }
    return FVDBM_OK;
  };
  auto node_q_major = [&](const void *dev, int comps) -> int {
    std::vector<T> tmp((size_t)comps * h.N);
    hipError_t e = hipMemcpy(tmp.data(), dev, tmp.size() * es, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(FVDBM_ERR_HIP, hipGetErrorString(e));
    for (int i = 0; i < h.N; ++i)
      for (int q = 0; q < comps; ++q) out[(size_t)i * comps + q] = tmp[(size_t)q * h.N + i];
    return FVDBM_OK;
  };
  switch (field) {
    case FVDBM_CELL_PDF:
      if (h.batch) {  // each case's current buffer
        std::vector<int32_t> cur(h.ncases);
        HIP_TRY(hipMemcpy(cur.data(), h.case_cur, cur.size() * 4, hipMemcpyDeviceToHost));
        std::vector<T> t0((size_t)Q * h.Cp), t1((size_t)Q * h.Cp);
        HIP_TRY(hipMemcpy(t0.data(), h.f[0], t0.size() * es, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(t1.data(), h.f[1], t1.size() * es, hipMemcpyDeviceToHost));
        for (int k = 0; k < h.ncases; ++k) {
          const std::vector<T> &src = cur[k] ? t1 : t0;
          for (int i = h.case_tile_off_host[k] * h.block; i < h.case_tile_off_host[k + 1] * h.block;
               ++i) {
            const int o = h.new_to_old[i];
            if (o < 0) continue;
            for (int q = 0; q < Q; ++q) out[(size_t)o * Q + q] = src[(size_t)q * h.Cp + i];
          }
        }
        return FVDBM_OK;
      }
      return cell_q_major(h.f[h.cur], Q, h.Cp);
    case FVDBM_CELL_RHO:
    case FVDBM_CELL_VEL:
    case FVDBM_CELL_PDF_EQ:
      if (!h.has_prev) {
        // Cells.__init__: zeros until the first step (src/containers.py:56-58)
        size_t n = (size_t)h.C * (field == FVDBM_CELL_RHO ? 1 : field == FVDBM_CELL_VEL ? 2 : Q);
        std::fill(out, out + n, T(0));
        return FVDBM_OK;
      } else {
        int rc = observe_typed<T>(h);
        if (rc) return rc;
        if (field == FVDBM_CELL_RHO) return cell_q_major(h.obs_rho, 1, h.Cp);
        if (field == FVDBM_CELL_VEL) return cell_q_major(h.obs_vel, 2, h.Cp);
        return cell_q_major(h.obs_feq, Q, h.Cp);
      }
    case FVDBM_FACE_FLUX:
      if (!h.has_prev) {
        Lattice<T> k = make_lattice<T>(h.tau, h.dt);
        for (int j = 0; j < h.F; ++j)
          for (int q = 0; q < Q; ++q) out[(size_t)j * Q + q] = k.W[q];
        return FVDBM_OK;
      } else {
        int rc = observe_typed<T>(h);
        if (rc) return rc;
        std::vector<T> tmp((size_t)Q * h.F);
        HIP_TRY(hipMemcpy(tmp.data(), h.obs_flux, tmp.size() * es, hipMemcpyDeviceToHost));
        for (int j = 0; j < h.F; ++j)
          for (int q = 0; q < Q; ++q) out[(size_t)j * Q + q] = tmp[(size_t)q * h.F + j];
        return FVDBM_OK;
      }
    case FVDBM_NODE_PDF:
    case FVDBM_NODE_RHO:
    case FVDBM_NODE_VEL: {
      int rc = observe_typed<T>(h);
      if (rc) return rc;
      if (field == FVDBM_NODE_PDF) return node_q_major(h.node_pdf, Q);
      if (field == FVDBM_NODE_RHO) {
        HIP_TRY(hipMemcpy(out, h.node_rho, (size_t)h.N * es, hipMemcpyDeviceToHost));
        return FVDBM_OK;
      }
      return node_q_major(h.node_vel, 2);
    }
    default:
      return fail(FVDBM_ERR_INVALID, "unknown field");
  }
}

template <typename T>
static int set_typed(Handle &h, int field, const void *src) {
  if (h.batch && field != FVDBM_CELL_PDF)
    return fail(FVDBM_ERR_INVALID, "a batched handle exposes FVDBM_CELL_PDF only");
  const T *in = static_cast<const T *>(src);
  const size_t es = sizeof(T);
  if (field == FVDBM_NODE_PDF || field == FVDBM_NODE_VEL || field == FVDBM_NODE_RHO) {
    // the start-of-step state is computed with the values the step used
    const int rc = materialize_mid<T>(h);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(h.stream));
  if (field == FVDBM_CELL_PDF) {
    std::vector<T> tmp((size_t)Q * h.Cp, T(0));
    for (int i = 0; i < h.Cp; ++i) {
      const int o = h.new_to_old[i];
      if (o < 0) continue;
      for (int q = 0; q < Q; ++q) tmp[(size_t)q * h.Cp + i] = in[(size_t)o * Q + q];
    }
    if (h.batch) {  // into each case's current buffer
      std::vector<int32_t> cur(h.ncases);
      HIP_TRY(hipMemcpy(cur.data(), h.case_cur, cur.size() * 4, hipMemcpyDeviceToHost));
      for (int k = 0; k < h.ncases; ++k) {
        const size_t i0 = (size_t)h.case_tile_off_host[k] * h.block;
        const size_t n = (size_t)(h.case_tile_off_host[k + 1] - h.case_tile_off_host[k]) * h.block;
        for (int q = 0; q < Q; ++q)
          HIP_TRY(hipMemcpy(static_cast<T *>(h.f[cur[k]]) + (size_t)q * h.Cp + i0,
                            tmp.data() + (size_t)q * h.Cp + i0, n * es, hipMemcpyHostToDevice));
      }
      return FVDBM_OK;
    }
    // write the current buffer only: f[cur^1] keeps the pdf at the start of
    // the last step (or two steps back), which the derived fields still describe
    HIP_TRY(hipMemcpy(h.f[h.cur], tmp.data(), tmp.size() * es, hipMemcpyHostToDevice));
    return FVDBM_OK;
  }
  if (field == FVDBM_NODE_PDF || field == FVDBM_NODE_VEL) {
    const int comps = field == FVDBM_NODE_PDF ? Q : 2;
    std::vector<T> tmp((size_t)comps * h.N);
    for (int i = 0; i < h.N; ++i)
      for (int q = 0; q < comps; ++q) tmp[(size_t)q * h.N + i] = in[(size_t)i * comps + q];
    HIP_TRY(hipMemcpy(field == FVDBM_NODE_PDF ? h.node_pdf : h.node_vel, tmp.data(),
                      tmp.size() * es, hipMemcpyHostToDevice));
    if (field == FVDBM_NODE_VEL) {
      // velocity-node values feed the step: refresh the active-node table
      std::vector<T> ux(h.nan), uy(h.nan);
      for (int a = 0; a < h.nan; ++a) {
        ux[a] = in[2 * (size_t)h.an_gid_host[a]];
        uy[a] = in[2 * (size_t)h.an_gid_host[a] + 1];
      }
      if (h.nan) {
        HIP_TRY(hipMemcpy(h.an_ux, ux.data(), ux.size() * es, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(h.an_uy, uy.data(), uy.size() * es, hipMemcpyHostToDevice));
      }
    }
    h.observed = false;
    return FVDBM_OK;
  }
  if (field == FVDBM_NODE_RHO) {
    HIP_TRY(hipMemcpy(h.node_rho, in, (size_t)h.N * es, hipMemcpyHostToDevice));
    std::vector<T> r(h.nan);
    for (int a = 0; a < h.nan; ++a) r[a] = in[h.an_gid_host[a]];
    if (h.nan) HIP_TRY(hipMemcpy(h.an_rho, r.data(), r.size() * es, hipMemcpyHostToDevice));
    h.observed = false;
    return FVDBM_OK;
  }
  return fail(FVDBM_ERR_INVALID, "field cannot be set");
}

// ----------------------------------------------------------------------------
// all-node interpolation and equilibrium initialisation
// ----------------------------------------------------------------------------
template <typename T>
static int ensure_work(Handle &h) {
  if (h.w_pdf) return FVDBM_OK;
  const size_t es = sizeof(T);
  int rc = h.alloc(&h.w_rho, (size_t)h.Cp * es);
  if (!rc) rc = h.alloc(&h.w_ux, (size_t)h.Cp * es);
  if (!rc) rc = h.alloc(&h.w_uy, (size_t)h.Cp * es);
  if (!rc) rc = h.alloc(&h.w_nrho, (size_t)h.N * es);
  if (!rc) rc = h.alloc(&h.w_nux, (size_t)h.N * es);
  if (!rc) rc = h.alloc(&h.w_nuy, (size_t)h.N * es);
  if (!rc) rc = h.alloc(&h.w_pdf, (size_t)Q * h.Cp * es);
  return rc;
}

template <typename T>
static int launch_slot_macros(Handle &h) {
  hipLaunchKernelGGL((k_slot_macros<T>), dim3((h.Cp + 255) / 256), dim3(256), 0, h.stream,
                     static_cast<const T *>(h.f[0]), static_cast<const T *>(h.f[1]), h.Cp, h.cur,
                     h.block, h.batch ? static_cast<const int32_t *>(h.tile_case) : nullptr,
                     static_cast<const int32_t *>(h.case_cur), static_cast<T *>(h.w_rho),
                     static_cast<T *>(h.w_ux), static_cast<T *>(h.w_uy));
  HIP_TRY(hipGetLastError());
  return FVDBM_OK;
}

// rho [N] and u [N][2] at every node from the current cell pdf
template <typename T>
static int node_fields_typed(Handle &h, void *rho_out, void *vel_out) {
  int rc = ensure_work<T>(h);
  if (!rc) rc = launch_slot_macros<T>(h);
  if (rc) return rc;
  hipLaunchKernelGGL((k_node_idw<T>), dim3((h.N + 255) / 256), dim3(256), 0, h.stream, h.N, h.K,
                     static_cast<const int32_t *>(h.all_nc), static_cast<const T *>(h.all_nd),
                     static_cast<const T *>(h.w_rho), static_cast<const T *>(h.w_ux),
                     static_cast<const T *>(h.w_uy), static_cast<T *>(h.w_nrho),
                     static_cast<T *>(h.w_nux), static_cast<T *>(h.w_nuy));
  HIP_TRY(hipGetLastError());
  std::vector<T> r(h.N), x(h.N), y(h.N);
  HIP_TRY(hipMemcpyAsync(r.data(), h.w_nrho, h.N * sizeof(T), hipMemcpyDeviceToHost, h.stream));
  HIP_TRY(hipMemcpyAsync(x.data(), h.w_nux, h.N * sizeof(T), hipMemcpyDeviceToHost, h.stream));
  HIP_TRY(hipMemcpyAsync(y.data(), h.w_nuy, h.N * sizeof(T), hipMemcpyDeviceToHost, h.stream));
  HIP_TRY(hipStreamSynchronize(h.stream));
  T *ro = static_cast<T *>(rho_out), *vo = static_cast<T *>(vel_out);
  for (int n = 0; n < h.N; ++n) {
    if (ro) ro[n] = r[n];
    if (vo) {
      vo[2 * (size_t)n] = x[n];
      vo[2 * (size_t)n + 1] = y[n];
    }
  }
  return FVDBM_OK;
}

// rho [C] and u [C][2] of the current cell pdf (reference cell order)
template <typename T>
static int cell_fields_typed(Handle &h, void *rho_out, void *vel_out) {
  int rc = ensure_work<T>(h);
  if (!rc) rc = launch_slot_macros<T>(h);
  if (rc) return rc;
  std::vector<T> r(h.Cp), x(h.Cp), y(h.Cp);
  HIP_TRY(hipMemcpyAsync(r.data(), h.w_rho, h.Cp * sizeof(T), hipMemcpyDeviceToHost, h.stream));
  HIP_TRY(hipMemcpyAsync(x.data(), h.w_ux, h.Cp * sizeof(T), hipMemcpyDeviceToHost, h.stream));
  HIP_TRY(hipMemcpyAsync(y.data(), h.w_uy, h.Cp * sizeof(T), hipMemcpyDeviceToHost, h.stream));
  HIP_TRY(hipStreamSynchronize(h.stream));
  T *ro = static_cast<T *>(rho_out), *vo = static_cast<T *>(vel_out);
  for (int i = 0; i < h.Cp; ++i) {
    const int o = h.new_to_old[i];
    if (o < 0) continue;
    if (ro) ro[o] = r[i];
    if (vo) {
      vo[2 * (size_t)o] = x[i];
      vo[2 * (size_t)o + 1] = y[i];
    }
  }
  return FVDBM_OK;
}

// The q-major pdf in w_pdf becomes the current cell pdf: of the handle, or
// of every case of a batch (into each case's current buffer).
template <typename T>
static int install_work_pdf(Handle &h) {
  const size_t es = sizeof(T);
  if (!h.batch) {
    HIP_TRY(hipMemcpyAsync(h.f[h.cur], h.w_pdf, (size_t)Q * h.Cp * es, hipMemcpyDeviceToDevice,
                           h.stream));
  } else {
    std::vector<int32_t> cur(h.ncases);
    HIP_TRY(hipMemcpyAsync(cur.data(), h.case_cur, cur.size() * 4, hipMemcpyDeviceToHost, h.stream));
    HIP_TRY(hipStreamSynchronize(h.stream));
    for (int k = 0; k < h.ncases; ++k) {
      const size_t i0 = (size_t)h.case_tile_off_host[k] * h.block;
      const size_t n = (size_t)(h.case_tile_off_host[k + 1] - h.case_tile_off_host[k]) * h.block;
      for (int q = 0; q < Q; ++q)
        HIP_TRY(hipMemcpyAsync(static_cast<T *>(h.f[cur[k]]) + (size_t)q * h.Cp + i0,
                               static_cast<const T *>(h.w_pdf) + (size_t)q * h.Cp + i0, n * es,
                               hipMemcpyDeviceToDevice, h.stream));
    }
  }
  HIP_TRY(hipStreamSynchronize(h.stream));
  h.observed = false;
  return FVDBM_OK;
}

// cell pdf <- feq(rho, u): from cell values (reference cell order) or, with
// at_nodes, from node values interpolated to the cells
template <typename T>
static int set_equilibrium_typed(Handle &h, const void *rho_in, const void *vel_in, bool at_nodes) {
  int rc = ensure_work<T>(h);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(h.stream));
  const T *ri = static_cast<const T *>(rho_in), *vi = static_cast<const T *>(vel_in);
  if (!at_nodes) {
    std::vector<T> r(h.Cp, T(1)), x(h.Cp, T(0)), y(h.Cp, T(0));
    for (int i = 0; i < h.Cp; ++i) {
      const int o = h.new_to_old[i];
      if (o < 0) continue;
      r[i] = ri[o];
      x[i] = vi[2 * (size_t)o];
      y[i] = vi[2 * (size_t)o + 1];
    }
    HIP_TRY(hipMemcpy(h.w_rho, r.data(), r.size() * sizeof(T), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.w_ux, x.data(), x.size() * sizeof(T), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.w_uy, y.data(), y.size() * sizeof(T), hipMemcpyHostToDevice));
  } else {
    std::vector<T> x(h.N), y(h.N);
    for (int n = 0; n < h.N; ++n) {
      x[n] = vi[2 * (size_t)n];
      y[n] = vi[2 * (size_t)n + 1];
    }
    HIP_TRY(hipMemcpy(h.w_nrho, ri, h.N * sizeof(T), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.w_nux, x.data(), x.size() * sizeof(T), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.w_nuy, y.data(), y.size() * sizeof(T), hipMemcpyHostToDevice));
    hipLaunchKernelGGL((k_cell_from_nodes<T>), dim3((h.Cp + 255) / 256), dim3(256), 0, h.stream,
                       h.Cp, static_cast<const int32_t *>(h.cn_off),
                       static_cast<const int32_t *>(h.cn_node), static_cast<const T *>(h.cn_d),
                       static_cast<const T *>(h.w_nrho), static_cast<const T *>(h.w_nux),
                       static_cast<const T *>(h.w_nuy), static_cast<T *>(h.w_rho),
                       static_cast<T *>(h.w_ux), static_cast<T *>(h.w_uy));
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL((k_equilibrium_fill<T>), dim3((h.Cp + 255) / 256), dim3(256), 0, h.stream,
                     h.Cp, make_lattice<T>(h.tau, h.dt), static_cast<const T *>(h.w_rho),
                     static_cast<const T *>(h.w_ux), static_cast<const T *>(h.w_uy),
                     static_cast<T *>(h.w_pdf));
  HIP_TRY(hipGetLastError());
  return install_work_pdf<T>(h);
}

}  // namespace fvdbm

// ============================================================================
// C ABI
// ============================================================================
using fvdbm::Handle;
using fvdbm::fail;

extern "C" {

const char *fvdbm_last_error(void) { return fvdbm::g_err.c_str(); }

int fvdbm_abi_version(void) { return FVDBM_ABI_VERSION; }

int fvdbm_device_count(int32_t *count) {
  if (!count) return fail(FVDBM_ERR_INVALID, "null count");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return FVDBM_OK;
}

int fvdbm_variant_count(int32_t *count) {
  if (!count) return fail(FVDBM_ERR_INVALID, "null count");
  // bit 0: no XCD remap; bit 1: separate node kernel; bit 2: no LDS staging;
  // bit 3: one step per launch (no two-step tiles)
  *count = 16;
  return FVDBM_OK;
}

static int create_impl(const fvdbm_mesh_desc *mesh, const fvdbm_params *p,
                       const fvdbm_batch_desc *batch, void **out) {
  if (!mesh || !p || !out) return fail(FVDBM_ERR_INVALID, "null argument");
  *out = nullptr;
  if (p->precision != FVDBM_FP32 && p->precision != FVDBM_FP64)
    return fail(FVDBM_ERR_INVALID, "precision must be FVDBM_FP32 or FVDBM_FP64");
  if (p->scheme != FVDBM_UPWIND && p->scheme != FVDBM_LAX_WENDROFF)
    return fail(FVDBM_ERR_INVALID, "Unknown flux scheme");
  if (!(p->tau > 0.0) || !std::isfinite(p->delta_t))
    return fail(FVDBM_ERR_INVALID, "tau must be > 0 and delta_t finite");
  int block = p->block ? p->block : 256;
  if (block < 64 || block > 1024 || block % 64)
    return fail(FVDBM_ERR_INVALID, "block must be a multiple of 64 in [64, 1024]");
  if (p->variant < 0 || p->variant > 15) return fail(FVDBM_ERR_INVALID, "variant out of range");
  int gs = p->graph_steps == 0 ? 32 : p->graph_steps;
  if (gs > 0 && (gs % 2)) ++gs;  // even, so a replay preserves buffer parity
  std::unique_ptr<Handle> h(new Handle());
  h->precision = p->precision;
  h->scheme = p->scheme;
  h->variant = p->variant;
  h->block = block;
  h->graph_steps = gs;
  h->reorder = p->reorder;
  h->tau = p->tau;
  h->dt = p->delta_t;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(FVDBM_ERR_HIP, "no HIP device available");
  const int32_t *cell_case = nullptr;
  int ncases = 1;
  if (batch) {
    if (batch->n_cases < 1 || !batch->cell_case || !batch->case_tau || !batch->case_dt)
      return fail(FVDBM_ERR_INVALID, "batch descriptor needs n_cases >= 1 and its arrays");
    ncases = batch->n_cases;
    cell_case = batch->cell_case;
    for (int k = 0; k < ncases; ++k) {
      if (!(batch->case_tau[k] > 0.0) || !std::isfinite(batch->case_dt[k]))
        return fail(FVDBM_ERR_INVALID, "case tau must be > 0 and dt finite");
      h->case_tau_host.push_back(batch->case_tau[k]);
      h->case_dt_host.push_back(batch->case_dt[k]);
    }
    h->graph_steps = -1;
  }
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&h->ev0));
  HIP_TRY(hipEventCreate(&h->ev1));
  int rc = p->precision == FVDBM_FP64
               ? fvdbm::create_typed<double>(*mesh, *h, cell_case, ncases)
               : fvdbm::create_typed<float>(*mesh, *h, cell_case, ncases);
  if (rc) return rc;
  // Default launch mode of a single handle: below kDirectLaunchCells a
  // launch takes a few microseconds, about what the host needs to queue
  // one, so runs replay captured graphs; from there on the host queues the
  // next launch while one runs, and plain launches leave no gap that a
  // graph replay would close (20-step runs were 2-6% faster at 0.25M-1M
  // cells, equal at 20k-100k; same results).
  if (!batch && p->graph_steps == 0 && h->Cp >= fvdbm::kDirectLaunchCells) h->graph_steps = -1;
  rc = p->precision == FVDBM_FP64 ? fvdbm::init_uniform_typed<double>(*h)
                                  : fvdbm::init_uniform_typed<float>(*h);
  if (rc) return rc;
  *out = h.release();
  return FVDBM_OK;
}

int fvdbm_create(const fvdbm_mesh_desc *mesh, const fvdbm_params *p, void **out) {
  return create_impl(mesh, p, nullptr, out);
}

int fvdbm_create_batch(const fvdbm_mesh_desc *mesh, const fvdbm_params *p,
                       const fvdbm_batch_desc *batch, void **out) {
  if (!batch) return fail(FVDBM_ERR_INVALID, "null batch descriptor");
  return create_impl(mesh, p, batch, out);
}

int fvdbm_batch_run(void *handle, int64_t max_steps, int32_t check_every, double tol,
                    double growth, int32_t *n_active, int64_t *steps_run) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !n_active || !steps_run) return fail(FVDBM_ERR_INVALID, "null argument");
  if (!h->batch) return fail(FVDBM_ERR_STATE, "not a batched handle");
  if (check_every < 1 || max_steps < 0) return fail(FVDBM_ERR_INVALID, "bad check interval");
  int rc = FVDBM_OK;
  if (max_steps >= check_every) rc = fvdbm::ensure_batch_graph(*h, check_every, tol, growth);
  if (rc) return rc;
  int32_t na = h->ncases;
  int64_t done = 0;
  while (done + check_every <= max_steps && na > 0) {
    HIP_TRY(hipGraphLaunch(h->bexec, h->stream));
    done += check_every;
    HIP_TRY(hipMemcpyAsync(&na, h->n_active, sizeof(na), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  if (done < max_steps && na > 0) {
    // a last, shorter interval (direct launches) so that no case runs past
    // max_steps; its residual check is the same as a full interval's
    const int rest = (int)(max_steps - done);
    int swaps = 0;
    rc = fvdbm::launch_batch_steps(*h, rest, true, &swaps);
    if (!rc) rc = fvdbm::launch_batch_check(*h, tol, growth, rest, swaps);
    if (rc) return rc;
    done += rest;
    HIP_TRY(hipMemcpyAsync(&na, h->n_active, sizeof(na), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  *n_active = na;
  *steps_run = done;
  return FVDBM_OK;
}

int fvdbm_batch_status(void *handle, int64_t *steps, double *residual, int32_t *status,
                       int32_t *active) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  if (!h->batch) return fail(FVDBM_ERR_STATE, "not a batched handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  const size_t n = h->ncases;
  if (steps) HIP_TRY(hipMemcpy(steps, h->case_steps, 8 * n, hipMemcpyDeviceToHost));
  if (residual) HIP_TRY(hipMemcpy(residual, h->case_resid, 8 * n, hipMemcpyDeviceToHost));
  if (status) HIP_TRY(hipMemcpy(status, h->case_status, 4 * n, hipMemcpyDeviceToHost));
  if (active) HIP_TRY(hipMemcpy(active, h->case_active, 4 * n, hipMemcpyDeviceToHost));
  return FVDBM_OK;
}

int fvdbm_batch_set_active(void *handle, const int32_t *active, int32_t reset_steps) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !active) return fail(FVDBM_ERR_INVALID, "null argument");
  if (!h->batch) return fail(FVDBM_ERR_STATE, "not a batched handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  const size_t n = h->ncases;
  std::vector<int32_t> act(active, active + n), st(n);
  std::vector<double> best(n);
  std::vector<int64_t> steps(n);
  HIP_TRY(hipMemcpy(st.data(), h->case_status, 4 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(best.data(), h->case_best, 8 * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(steps.data(), h->case_steps, 8 * n, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < n; ++k) {
    act[k] = act[k] ? 1 : 0;
    if (act[k]) {  // (re)started: running, no residual history
      st[k] = fvdbm::kRunning;
      best[k] = INFINITY;
      if (reset_steps) steps[k] = 0;
    } else if (st[k] == fvdbm::kRunning) {
      st[k] = fvdbm::kStopped;
    }
  }
  HIP_TRY(hipMemcpy(h->case_active, act.data(), 4 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->case_status, st.data(), 4 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->case_best, best.data(), 8 * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->case_steps, steps.data(), 8 * n, hipMemcpyHostToDevice));
  return FVDBM_OK;
}

int fvdbm_destroy(void *handle) {
  delete static_cast<Handle *>(handle);
  return FVDBM_OK;
}

int fvdbm_info_get(void *handle, fvdbm_info *o) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !o) return fail(FVDBM_ERR_INVALID, "null argument");
  o->n_cells = h->C;
  o->n_faces = h->F;
  o->n_nodes = h->N;
  o->k_max = h->K;
  o->n_cells_padded = h->Cp;
  o->n_boundary_cells = h->nbc;
  o->n_boundary_faces = h->nbf;
  o->n_active_nodes = h->nan;
  o->bandwidth = h->bandwidth;
  o->precision = h->precision;
  o->scheme = h->scheme;
  o->variant = h->variant;
  o->block = h->block;
  o->graph_steps = h->graph_steps;
  o->steps_done = h->steps;
  o->device_bytes = h->device_bytes;
  o->n_cases = h->ncases;
  o->launch_classes = h->pairs ? std::max(h->npair_class, 1) : 0;
  return FVDBM_OK;
}

int fvdbm_init_uniform(void *handle) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  return h->precision == FVDBM_FP64 ? fvdbm::init_uniform_typed<double>(*h)
                                    : fvdbm::init_uniform_typed<float>(*h);
}

int fvdbm_node_fields(void *handle, void *node_rho, void *node_vel) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || (!node_rho && !node_vel)) return fail(FVDBM_ERR_INVALID, "null argument");
  return h->precision == FVDBM_FP64 ? fvdbm::node_fields_typed<double>(*h, node_rho, node_vel)
                                    : fvdbm::node_fields_typed<float>(*h, node_rho, node_vel);
}

int fvdbm_cell_fields(void *handle, void *cell_rho, void *cell_vel) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || (!cell_rho && !cell_vel)) return fail(FVDBM_ERR_INVALID, "null argument");
  return h->precision == FVDBM_FP64 ? fvdbm::cell_fields_typed<double>(*h, cell_rho, cell_vel)
                                    : fvdbm::cell_fields_typed<float>(*h, cell_rho, cell_vel);
}

int fvdbm_set_equilibrium(void *handle, const void *cell_rho, const void *cell_vel) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !cell_rho || !cell_vel) return fail(FVDBM_ERR_INVALID, "null argument");
  return h->precision == FVDBM_FP64
             ? fvdbm::set_equilibrium_typed<double>(*h, cell_rho, cell_vel, false)
             : fvdbm::set_equilibrium_typed<float>(*h, cell_rho, cell_vel, false);
}

int fvdbm_set_equilibrium_nodes(void *handle, const void *node_rho, const void *node_vel) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !node_rho || !node_vel) return fail(FVDBM_ERR_INVALID, "null argument");
  return h->precision == FVDBM_FP64
             ? fvdbm::set_equilibrium_typed<double>(*h, node_rho, node_vel, true)
             : fvdbm::set_equilibrium_typed<float>(*h, node_rho, node_vel, true);
}

int fvdbm_set_field(void *handle, int32_t field, const void *src) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !src) return fail(FVDBM_ERR_INVALID, "null argument");
  return h->precision == FVDBM_FP64 ? fvdbm::set_typed<double>(*h, field, src)
                                    : fvdbm::set_typed<float>(*h, field, src);
}

int fvdbm_get_field(void *handle, int32_t field, void *dst) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !dst) return fail(FVDBM_ERR_INVALID, "null argument");
  return h->precision == FVDBM_FP64 ? fvdbm::get_typed<double>(*h, field, dst)
                                    : fvdbm::get_typed<float>(*h, field, dst);
}

int fvdbm_step(void *handle, int64_t nsteps) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  if (nsteps < 0) return fail(FVDBM_ERR_INVALID, "nsteps < 0");
  return fvdbm::run_steps(*h, nsteps, false);
}

int fvdbm_step_residual(void *handle, int64_t nsteps, double *residual) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !residual) return fail(FVDBM_ERR_INVALID, "null argument");
  if (nsteps < 1) return fail(FVDBM_ERR_INVALID, "nsteps must be >= 1");
  int rc = fvdbm::run_steps(*h, nsteps, true);
  if (rc) return rc;
  hipLaunchKernelGGL(fvdbm::k_residual_finalize, dim3(1), dim3(256), 0, h->stream,
                     static_cast<const double *>(h->partials), h->ntiles,
                     static_cast<double *>(h->resid));
  HIP_TRY(hipGetLastError());
  double r[2];
  HIP_TRY(hipMemcpyAsync(r, h->resid, sizeof(r), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  *residual = r[1] > 0.0 ? std::sqrt(r[0]) / std::sqrt(r[1]) : (r[0] > 0.0 ? INFINITY : 0.0);
  if (!std::isfinite(r[0]) || !std::isfinite(r[1])) *residual = NAN;
  return FVDBM_OK;
}

int fvdbm_prepare(void *handle, int64_t nsteps) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  if (nsteps < 0) return fail(FVDBM_ERR_INVALID, "nsteps < 0");
  if (h->batch) return h->graph_steps > 0 ? fvdbm::ensure_batch_step_graph(*h) : FVDBM_OK;
  // from either buffer parity: the steps before the run decide which
  const int cur = h->cur;
  int rc = FVDBM_OK;
  for (int par = 0; par < 2 && rc == FVDBM_OK; ++par) {
    h->cur = par;
    rc = h->pairs ? fvdbm::run_steps(*h, nsteps, false, true)
                  : (h->graph_steps > 0 && nsteps >= h->graph_steps ? fvdbm::ensure_graph(*h, par)
                                                                     : FVDBM_OK);
  }
  h->cur = cur;
  return rc;
}

int fvdbm_step_timed(void *handle, int64_t nsteps, double *ms) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !ms) return fail(FVDBM_ERR_INVALID, "null argument");
  // build graphs outside the timed region
  {
    // the graphs this run replays, captured outside the timed region
    int rc = FVDBM_OK;
    if (h->graph_steps > 0 && nsteps >= h->graph_steps)
      rc = h->batch ? fvdbm::ensure_batch_step_graph(*h)
                    : h->pairs ? FVDBM_OK : fvdbm::ensure_graph(*h, h->cur);
    if (!rc && h->pairs && !h->batch) rc = fvdbm::run_steps(*h, nsteps, false, true);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  int rc = fvdbm::run_steps(*h, nsteps, false);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, h->ev0, h->ev1));
  *ms = t;
  return FVDBM_OK;
}

int fvdbm_synchronize(void *handle) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  return FVDBM_OK;
}

int fvdbm_copy_probe(int64_t bytes, int32_t reps, double *ms_per_copy) {
  if (bytes < 16 || reps < 1 || !ms_per_copy) return fail(FVDBM_ERR_INVALID, "bad copy probe arguments");
  const size_t n = (size_t)bytes / 16;
  void *src = nullptr, *dst = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  };
  hipError_t e = hipMalloc(&src, n * 16);
  if (e == hipSuccess) e = hipMalloc(&dst, n * 16);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(src, 0x3c, n * 16, st);
  const int grid = 256 * 8 * 4;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fvdbm::k_copy_probe, dim3(grid), dim3(256), 0, st,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
    e = hipEventRecord(e0, st);
  }
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    hipLaunchKernelGGL(fvdbm::k_copy_probe, dim3(grid), dim3(256), 0, st,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  cleanup();
  if (e != hipSuccess) return fail(FVDBM_ERR_HIP, std::string("copy probe: ") + hipGetErrorString(e));
  *ms_per_copy = t / reps;
  return FVDBM_OK;
}

int fvdbm_get_permutation(void *handle, int32_t *new_to_old) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !new_to_old) return fail(FVDBM_ERR_INVALID, "null argument");
  size_t j = 0;
  for (int32_t o : h->new_to_old)
    if (o >= 0) new_to_old[j++] = o;
  return FVDBM_OK;
}

}  // extern "C"
