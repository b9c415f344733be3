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
#include <memory>
#include <string>
#include <vector>

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
  T C2, C4x2, C2x2;  // C*C, 2*(C*C)*(C*C), 2*C*C with C = 1/sqrt(3)
  T invtau, dt;      // 1/tau rounded from double, delta_t rounded
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
  k.invtau = T(1.0 / tau);
  k.dt = T(dt);
  return k;
}

// ----------------------------------------------------------------------------
// device math: one function per reference formula, same operation order
// ----------------------------------------------------------------------------

// rho = sum_q f_q ; u = (sum_q ksi_q f_q) / rho, sums 0-initialised in q order
template <typename T>
__device__ __forceinline__ void macro(const T (&f)[Q], T &rho, T &ux, T &uy) {
  T r = T(0);
#pragma unroll
  for (int q = 0; q < Q; ++q) r += f[q];
  T mx = T(0), my = T(0);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    mx += T(kx(q)) * f[q];
    my += T(ky(q)) * f[q];
  }
  rho = r;
  ux = mx / r;
  uy = my / r;
}

// feq_q = (W_q rho) * (((1 + A_q) + B_q) - D) with A = (ksi.u)/C^2,
// B = (ksi.u)^2/(2C^4), D = (u.u)/(2C^2). ksi.u of opposite directions are
// exact negatives and (-a)/b == -(a/b) in round-to-nearest, so only four A
// and four B quotients are distinct; each bracket is still formed per q.
template <typename T>
__device__ __forceinline__ void equilibrium(const Lattice<T> &k, T rho, T ux, T uy,
                                            T (&feq)[Q]) {
  T uu = T(0);
  uu += ux * ux;
  uu += uy * uy;
  const T D = uu / k.C2x2;
  const T c5 = ux + uy;          // q5; q7 = -c5
  const T c6 = (-ux) + uy;       // q6; q8 = -c6
  const T A1 = ux / k.C2, A2 = uy / k.C2, A5 = c5 / k.C2, A6 = c6 / k.C2;
  const T B1 = (ux * ux) / k.C4x2, B2 = (uy * uy) / k.C4x2;
  const T B5 = (c5 * c5) / k.C4x2, B6 = (c6 * c6) / k.C4x2;
  const T one = T(1);
  feq[0] = (k.W[0] * rho) * (((one + T(0)) + T(0)) - D);
  feq[1] = (k.W[1] * rho) * (((one + A1) + B1) - D);
  feq[2] = (k.W[2] * rho) * (((one + A2) + B2) - D);
  feq[3] = (k.W[3] * rho) * (((one - A1) + B1) - D);
  feq[4] = (k.W[4] * rho) * (((one - A2) + B2) - D);
  feq[5] = (k.W[5] * rho) * (((one + A5) + B5) - D);
  feq[6] = (k.W[6] * rho) * (((one + A6) + B6) - D);
  feq[7] = (k.W[7] * rho) * (((one - A5) + B5) - D);
  feq[8] = (k.W[8] * rho) * (((one - A6) + B6) - D);
}

template <typename T>
struct StepArgs {
  const T *__restrict__ fsrc;  // [9][Cp]
  T *__restrict__ fdst;        // [9][Cp]
  const int32_t *__restrict__ nb;  // [3][Cp]
  const T *__restrict__ nx;
  const T *__restrict__ ny;
  const T *__restrict__ sL;
  const T *__restrict__ lwa;
  const T *__restrict__ lwden2;
  const int32_t *__restrict__ bf_slot;   // [2][nbf] slots in the tile node table
  const int32_t *__restrict__ bf_gnode;  // [2][nbf] node ids
  const T *__restrict__ bf_ratio;        // [nbf]
  const int32_t *__restrict__ tile_node_off;  // [ntiles+1]
  const int32_t *__restrict__ tile_nodes;     // encoded nodes of each tile
  const int32_t *__restrict__ an_type;
  const int32_t *__restrict__ an_cells;  // [nan][K]
  const T *__restrict__ an_w;            // [nan][K]
  const T *__restrict__ an_wsum;
  const T *__restrict__ an_rho;
  const T *__restrict__ an_ux;
  const T *__restrict__ an_uy;
  const T *__restrict__ node_pdf;  // [9][N]
  double *__restrict__ partials;   // [ntiles][2] residual partial sums
  int C, Cp, N, K, nbf, nan;
  int ntiles, xcd_remap, inline_nodes;
  Lattice<T> k;
};

template <typename T>
__device__ __forceinline__ void load_cell(const T *__restrict__ f, int Cp, int c, T (&out)[Q]) {
#pragma unroll
  for (int q = 0; q < Q; ++q) out[q] = f[(size_t)q * Cp + c];
}

// Nodes.calc_dirichlet_pdf for an active node (src/containers.py:343-390):
// type 1 (velocity): rho <- IDW(cell rho), u fixed; type 2 (density):
// u <- IDW(cell u), rho fixed; pdf = calc_eq(rho, u) + IDW(f_c - feq_c).
// Padding entries (-1) carry weight 0 and contribute +0, as x*0 does.
template <typename T>
__device__ __forceinline__ void active_node(const StepArgs<T> &a, int an, T (&out)[Q], T &rho_n,
                                            T &ux_n, T &uy_n) {
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
        load_cell(a.fsrc, a.Cp, cj < 0 ? -2 - cj : cj, fc);
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
      load_cell(a.fsrc, a.Cp, cj, fc);
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
// XCDs, so give each XCD one contiguous range of tiles; neighbour gathers
// then mostly hit that XCD's own L2.
__device__ __forceinline__ int tile_of_block(int b, int nblocks, int remap) {
  if (!remap) return b;
  const int per = nblocks >> 3;  // grid is a multiple of 8 when remapping
  return (b & 7) * per + (b >> 3);
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
  extern __shared__ unsigned char smem_raw[];
  T *sh = reinterpret_cast<T *>(smem_raw);
  const int tile = tile_of_block(blockIdx.x, gridDim.x, a.xcd_remap);
  if (tile >= a.ntiles) return;  // uniform per block
  const int t0 = a.tile_node_off[tile];
  const int ntn = a.tile_node_off[tile + 1] - t0;
  if (a.inline_nodes && ntn > 0) {  // uniform per block
    for (int i = threadIdx.x; i < ntn; i += blockDim.x) {
      const int e = a.tile_nodes[t0 + i];
      T out[Q];
      if (e >= 0) {
        T r, ux, uy;
        active_node(a, e, out, r, ux, uy);
      } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) out[q] = a.node_pdf[(size_t)q * a.N + (-1 - e)];
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) sh[q * ntn + i] = out[q];
    }
    __syncthreads();
  }

  const int c = tile * blockDim.x + threadIdx.x;
  T res_d = T(0), res_n = T(0);
  if (c < a.C) {
    T f[Q], acc[Q];
    load_cell(a.fsrc, a.Cp, c, f);
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = T(0);
#pragma unroll
    for (int kf = 0; kf < 3; ++kf) {
      const size_t at = (size_t)kf * a.Cp + c;
      const int32_t code = a.nb[at];
      const T nx = a.nx[at], ny = a.ny[at], sL = a.sL[at];
      const bool slot1 = (code & kSlot1Bit) != 0;
      T o[Q];
      if (code >= 0) {
        load_cell(a.fsrc, a.Cp, code & kIndexMask, o);
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
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          T phi = T(0);
          phi += T(kx(q)) * nx;
          phi += T(ky(q)) * ny;
          // upwind: slot 0 when ksi.n >= 0, else slot 1
          const bool take_slot0 = phi >= T(0);
          const T up = (take_slot0 != slot1) ? f[q] : o[q];
          acc[q] += (up * phi) * sL;
        }
      } else {
        const T la = a.lwa[at], den2 = a.lwden2[at];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          T phi = T(0);
          phi += T(kx(q)) * nx;
          phi += T(ky(q)) * ny;
          const T fu = slot1 ? o[q] : f[q];
          const T fd = slot1 ? f[q] : o[q];
          const T interp = la - (phi * a.k.dt) / den2;
          const T fc = fu + (fd - fu) * interp;
          acc[q] += (fc * phi) * sL;
        }
      }
    }
    T rho, ux, uy, feq[Q];
    macro(f, rho, ux, uy);
    equilibrium(a.k, rho, ux, uy, feq);
    T fn[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      fn[q] = f[q] + a.k.dt * (a.k.invtau * (feq[q] - f[q]) - acc[q]);
      a.fdst[(size_t)q * a.Cp + c] = fn[q];
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

// Active nodes from the pdf in fsrc: node pdf (type 1/2), node rho (type 1)
// and node vel (type 2) written to the node arrays (reference node order).
template <typename T>
__global__ void k_nodes(StepArgs<T> a, const int32_t *__restrict__ an_gid, T *__restrict__ npdf,
                        T *__restrict__ nrho, T *__restrict__ nvel) {
  const int an = blockIdx.x * blockDim.x + threadIdx.x;
  if (an >= a.nan) return;
  T out[Q], r, ux, uy;
  active_node(a, an, out, r, ux, uy);
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
  T f[Q], eq[Q], r, ux, uy;
  load_cell(a.fsrc, a.Cp, c, f);
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
  const int ia = fa.c0[j], ib = fa.c1[j];
  const T d0 = static_cast<const T *>(fa.d0)[j], d1 = static_cast<const T *>(fa.d1)[j];
  const T nx = static_cast<const T *>(fa.nx)[j], ny = static_cast<const T *>(fa.ny)[j];
  const T L = static_cast<const T *>(fa.L)[j];
  T fl[Q], fr[Q];
  for (int side = 0; side < 2; ++side) {
    const int idx = side == 0 ? ia : ib;
    T tmp[Q];
    if (idx != -1) {
      load_cell(a.fsrc, a.Cp, idx, tmp);
    } else {
      int known = side == 0 ? ib : ia;
      if (known < 0) known = fa.last_cell;
      const T gd = side == 0 ? d0 : d1, kd = side == 0 ? d1 : d0;
      const T ratio = gd / kd;
      T fk[Q];
      load_cell(a.fsrc, a.Cp, known, fk);
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

template <typename T>
__global__ void k_fill_uniform(T *__restrict__ f, int Cp, int C, Lattice<T> k) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Cp) return;
#pragma unroll
  for (int q = 0; q < Q; ++q) f[(size_t)q * Cp + c] = c < C ? k.W[q] : T(0);
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
  void *nb = nullptr, *nx = nullptr, *ny = nullptr, *sL = nullptr, *lwa = nullptr,
       *lwden2 = nullptr;
  void *bf_slot = nullptr, *bf_gnode = nullptr, *bf_ratio = nullptr;
  void *tile_node_off = nullptr, *tile_nodes = nullptr;
  int max_tile_nodes = 0;
  void *an_gid = nullptr, *an_type = nullptr, *an_cells = nullptr, *an_w = nullptr,
       *an_wsum = nullptr, *an_rho = nullptr, *an_ux = nullptr, *an_uy = nullptr;
  void *node_pdf = nullptr, *node_rho = nullptr, *node_vel = nullptr;
  void *fc0 = nullptr, *fc1 = nullptr, *fn0 = nullptr, *fn1 = nullptr, *fd0 = nullptr,
       *fd1 = nullptr, *fnx = nullptr, *fny = nullptr, *fL = nullptr;
  void *partials = nullptr, *resid = nullptr;
  void *obs_rho = nullptr, *obs_vel = nullptr, *obs_feq = nullptr, *obs_flux = nullptr;
  int cur = 0;           // buffer holding the current pdf
  int64_t steps = 0;     // steps taken since the last init/set of the cell pdf
  bool has_prev = false; // f[cur^1] holds the pdf at the start of the last step
  bool observed = false; // observable buffers are up to date
  // graphs: one per starting buffer parity
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};

  ~Handle() {
    for (int i = 0; i < 2; ++i) {
      if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
      if (graph[i]) (void)hipGraphDestroy(graph[i]);
    }
    for (auto &b : bufs)
      if (b.p) (void)hipFree(b.p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
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
  a.nb = static_cast<const int32_t *>(h.nb);
  a.nx = static_cast<const T *>(h.nx);
  a.ny = static_cast<const T *>(h.ny);
  a.sL = static_cast<const T *>(h.sL);
  a.lwa = static_cast<const T *>(h.lwa);
  a.lwden2 = static_cast<const T *>(h.lwden2);
  a.bf_slot = static_cast<const int32_t *>(h.bf_slot);
  a.bf_gnode = static_cast<const int32_t *>(h.bf_gnode);
  a.tile_node_off = static_cast<const int32_t *>(h.tile_node_off);
  a.tile_nodes = static_cast<const int32_t *>(h.tile_nodes);
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
  a.C = h.C;
  a.Cp = h.Cp;
  a.N = h.N;
  a.K = h.K;
  a.nbf = h.nbf;
  a.nan = h.nan;
  a.ntiles = h.ntiles;
  a.xcd_remap = (h.variant & 1) ? 0 : 1;
  a.inline_nodes = (h.variant & 2) ? 0 : 1;
  a.k = make_lattice<T>(h.tau, h.dt);
  return a;
}

static int grid_for(const Handle &h, int remap) {
  return remap ? ((h.ntiles + 7) / 8) * 8 : h.ntiles;
}

// Launch one step reading f[src]; appends to the stream.
template <typename T>
static int launch_step(Handle &h, int src, bool resid) {
  StepArgs<T> a = make_args<T>(h, src);
  if (!a.inline_nodes && h.nan > 0) {
    // two-kernel variant: refresh active node pdfs first
    const int nb = (h.nan + 255) / 256;
    hipLaunchKernelGGL((k_nodes<T>), dim3(nb), dim3(256), 0, h.stream, a,
                       static_cast<const int32_t *>(h.an_gid), static_cast<T *>(h.node_pdf),
                       static_cast<T *>(h.node_rho), static_cast<T *>(h.node_vel));
  }
  const dim3 grid(grid_for(h, a.xcd_remap)), block(h.block);
  const size_t lds = a.inline_nodes ? (size_t)Q * h.max_tile_nodes * sizeof(T) : 0;
  if (h.scheme == FVDBM_UPWIND) {
    if (resid)
      hipLaunchKernelGGL((k_step<T, 0, true>), grid, block, lds, h.stream, a);
    else
      hipLaunchKernelGGL((k_step<T, 0, false>), grid, block, lds, h.stream, a);
  } else {
    if (resid)
      hipLaunchKernelGGL((k_step<T, 1, true>), grid, block, lds, h.stream, a);
    else
      hipLaunchKernelGGL((k_step<T, 1, false>), grid, block, lds, h.stream, a);
  }
  HIP_TRY(hipGetLastError());
  return FVDBM_OK;
}

static int launch_step_any(Handle &h, int src, bool resid) {
  return h.precision == FVDBM_FP64 ? launch_step<double>(h, src, resid)
                                   : launch_step<float>(h, src, resid);
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

static int run_steps(Handle &h, int64_t n, bool resid_last) {
  if (n <= 0) return FVDBM_OK;
  int64_t plain = resid_last ? n - 1 : n;
  if (h.graph_steps > 0) {
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
  h.has_prev = true;
  h.observed = false;
  return FVDBM_OK;
}

template <typename T>
static int create_typed(const fvdbm_mesh_desc &m, Handle &h) {
  HostLayout<T> L;
  std::string err;
  if (!build_layout<T>(m, h.reorder, h.block, L, err)) return fail(FVDBM_ERR_INVALID, err);
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
  h.ntiles = L.ntiles;
  h.max_tile_nodes = L.max_tile_nodes;
  // node tables beyond 64 KiB of LDS per block: use the two-kernel variant
  if ((size_t)Q * L.max_tile_nodes * sizeof(T) > 65536) h.variant |= 2;
  h.new_to_old = L.new_to_old;
  h.an_gid_host = L.an_gid;
  h.an_type_host = L.an_type;
  const size_t es = sizeof(T);
  int rc = 0;
  for (int i = 0; i < 2 && !rc; ++i) rc = h.alloc(&h.f[i], (size_t)Q * h.Cp * es);
  if (!rc) rc = h.upload(&h.nb, L.nb);
  if (!rc) rc = h.upload(&h.nx, L.nx);
  if (!rc) rc = h.upload(&h.ny, L.ny);
  if (!rc) rc = h.upload(&h.sL, L.sL);
  if (!rc && h.scheme == FVDBM_LAX_WENDROFF) rc = h.upload(&h.lwa, L.lwa);
  if (!rc && h.scheme == FVDBM_LAX_WENDROFF) rc = h.upload(&h.lwden2, L.lwden2);
  if (!rc) rc = h.upload(&h.bf_slot, L.bf_slot);
  if (!rc) rc = h.upload(&h.bf_gnode, L.bf_gnode);
  if (!rc) rc = h.upload(&h.tile_node_off, L.tile_node_off);
  if (!rc) rc = h.upload(&h.tile_nodes, L.tile_nodes);
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
  for (int i = 0; i < 2; ++i)
    hipLaunchKernelGGL((k_fill_uniform<T>), dim3((h.Cp + 255) / 256), dim3(256), 0, h.stream,
                       static_cast<T *>(h.f[i]), h.Cp, h.C, k);
  // node pdf <- W (both type 0 and active nodes)
  hipLaunchKernelGGL((k_fill_uniform<T>), dim3((h.N + 255) / 256), dim3(256), 0, h.stream,
                     static_cast<T *>(h.node_pdf), h.N, h.N, k);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h.stream));
  h.cur = 0;
  h.steps = 0;
  h.has_prev = false;
  h.observed = false;
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
  StepArgs<T> a = make_args<T>(h, h.cur ^ 1);  // the pdf at the start of the last step
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
  T *out = static_cast<T *>(dst);
  const size_t es = sizeof(T);
  HIP_TRY(hipStreamSynchronize(h.stream));
  auto cell_q_major = [&](const void *dev, int comps, size_t stride) -> int {
    std::vector<T> tmp((size_t)comps * stride);
    hipError_t e = hipMemcpy(tmp.data(), dev, tmp.size() * es, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(FVDBM_ERR_HIP, hipGetErrorString(e));
    for (int i = 0; i < h.C; ++i) {
      const int o = h.new_to_old[i];
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
  const T *in = static_cast<const T *>(src);
  const size_t es = sizeof(T);
  HIP_TRY(hipStreamSynchronize(h.stream));
  if (field == FVDBM_CELL_PDF) {
    std::vector<T> tmp((size_t)Q * h.Cp, T(0));
    for (int i = 0; i < h.C; ++i) {
      const int o = h.new_to_old[i];
      for (int q = 0; q < Q; ++q) tmp[(size_t)q * h.Cp + i] = in[(size_t)o * Q + q];
    }
    // write the current buffer only: f[cur^1] keeps the pdf at the start of
    // the last step, which the derived cell fields still describe
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
  *count = 4;  // bit 0: no XCD remap; bit 1: separate node kernel
  return FVDBM_OK;
}

int fvdbm_create(const fvdbm_mesh_desc *mesh, const fvdbm_params *p, void **out) {
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
  if (p->variant < 0 || p->variant > 3) return fail(FVDBM_ERR_INVALID, "variant out of range");
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
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&h->ev0));
  HIP_TRY(hipEventCreate(&h->ev1));
  int rc = p->precision == FVDBM_FP64 ? fvdbm::create_typed<double>(*mesh, *h)
                                      : fvdbm::create_typed<float>(*mesh, *h);
  if (rc) return rc;
  rc = p->precision == FVDBM_FP64 ? fvdbm::init_uniform_typed<double>(*h)
                                  : fvdbm::init_uniform_typed<float>(*h);
  if (rc) return rc;
  *out = h.release();
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
  return FVDBM_OK;
}

int fvdbm_init_uniform(void *handle) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  return h->precision == FVDBM_FP64 ? fvdbm::init_uniform_typed<double>(*h)
                                    : fvdbm::init_uniform_typed<float>(*h);
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

int fvdbm_step_timed(void *handle, int64_t nsteps, double *ms) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !ms) return fail(FVDBM_ERR_INVALID, "null argument");
  // build graphs outside the timed region
  if (h->graph_steps > 0 && nsteps >= h->graph_steps) {
    int rc = fvdbm::ensure_graph(*h, h->cur);
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

int fvdbm_get_permutation(void *handle, int32_t *new_to_old) {
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !new_to_old) return fail(FVDBM_ERR_INVALID, "null argument");
  std::copy(h->new_to_old.begin(), h->new_to_old.end(), new_to_old);
  return FVDBM_OK;
}

}  // extern "C"
