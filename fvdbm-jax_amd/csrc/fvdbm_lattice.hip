// FVDBM step for any supported lattice: D2Q13 (src/dynamics.py:77-102), and
// D2Q9 through the same code as a cross-check of the tuned D2Q9 kernels in
// fvdbm_step.hip. One step is one kernel: one thread per cell for the face
// fluxes and the BGK update (src/containers.py:107-121, 191-287), a tile
// with ghost faces first evaluating the boundary nodes they read
// (Nodes.calc_pdfs, src/containers.py:339-390) into LDS. Same reference operation order and
// no contraction (-ffp-contract=off) as the tuned kernels, so D2Q9 results
// equal theirs bit for bit and D2Q13 results equal the C oracle's.
//
// Layout (device cell order and every per-cell record) comes from the same
// host builder (layout.cpp); pdfs are q-major [Q][Cp], node pdfs [Q][N].
// C ABI: fvdbm_lattice_* in include/fvdbm.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "errors.h"
#include "fvdbm.h"
#include "layout.h"

namespace fvdbm_lattice {

using fvdbm::HostLayout;
using fvdbm::kIndexMask;
using fvdbm::kSlot1Bit;

static int fail(int code, const std::string &m) { return fvdbm::set_error(code, m); }

#define LHIP_TRY(expr)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(FVDBM_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));     \
  } while (0)

// Velocities (src/dynamics.py:54-62 and 80-96): D2Q13 is D2Q9 plus
// q = 9..12 = (2,0), (0,2), (-2,0), (0,-2).
__host__ __device__ constexpr int kx(int q) {
  return q == 1 || q == 5 || q == 8 ? 1
         : q == 3 || q == 6 || q == 7 ? -1
         : q == 9                     ? 2
         : q == 11                    ? -2
                                      : 0;
}
__host__ __device__ constexpr int ky(int q) {
  return q == 2 || q == 5 || q == 6 ? 1
         : q == 4 || q == 7 || q == 8 ? -1
         : q == 10                    ? 2
         : q == 12                    ? -2
                                      : 0;
}

// W rounded from the Python doubles; C = 1/sqrt(3) (D2Q9) or 1/sqrt(2)
// (D2Q13) in the working precision; powers of C as lax.integer_pow forms
// them: C^2 = C*C, C^4 = C^2*C^2, C^6 = C^2*C^4.
template <typename T, int QN>
struct Consts {
  T W[QN];
  T C2, C4x2, C2x2, C6x2;      // C^2, 2 C^4, 2 C^2, 2 C^6
  T rC2, rC4x2, rC2x2, rC6x2;  // their rounded reciprocals (FP32 fast division)
  T invtau, dt;                // 1/tau rounded from double, delta_t rounded
};

template <typename T, int QN>
static Consts<T, QN> make_consts(double tau, double dt) {
  static const double w9[9] = {4.0 / 9,  1.0 / 9,  1.0 / 9,  1.0 / 9, 1.0 / 9,
                               1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
  static const double w13[13] = {3.0 / 8,  1.0 / 12, 1.0 / 12, 1.0 / 12, 1.0 / 12,
                                 1.0 / 16, 1.0 / 16, 1.0 / 16, 1.0 / 16, 1.0 / 96,
                                 1.0 / 96, 1.0 / 96, 1.0 / 96};
  Consts<T, QN> k;
  for (int q = 0; q < QN; ++q) k.W[q] = T(QN == 13 ? w13[q] : w9[q]);
  const T C = T(1) / std::sqrt(T(QN == 13 ? 2 : 3));
  k.C2 = C * C;
  const T c4 = k.C2 * k.C2;
  const T c6 = k.C2 * c4;
  k.C4x2 = T(2) * c4;
  k.C2x2 = T(2) * k.C2;
  k.C6x2 = T(2) * c6;
  k.rC2 = T(1) / k.C2;
  k.rC4x2 = T(1) / k.C4x2;
  k.rC2x2 = T(1) / k.C2x2;
  k.rC6x2 = T(1) / k.C6x2;
  k.invtau = T(1.0 / tau);
  k.dt = T(dt);
  return k;
}

// rho = sum_q f_q; u = (sum_q ksi_q f_q) / rho (0-initialised, q order)
template <typename T, int QN>
__device__ __forceinline__ void macro(const T (&f)[QN], T &rho, T &ux, T &uy) {
  T r = T(0);
#pragma unroll
  for (int q = 0; q < QN; ++q) r += f[q];
  // terms with a zero ksi component add a zero to an accumulator that is
  // never -0 (it starts at +0): left out, exact for finite f
  T mx = T(0), my = T(0);
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    if (kx(q) != 0) mx += T(kx(q)) * f[q];
    if (ky(q) != 0) my += T(ky(q)) * f[q];
  }
  rho = r;
  ux = mx / r;
  uy = my / r;
}

// velocities in pairs of opposites
constexpr int kPos[6] = {1, 2, 5, 6, 9, 10};   // and their opposites:
constexpr int kNeg[6] = {3, 4, 7, 8, 11, 12};  // ksi_neg = -ksi_pos

// x / y for y one of the lattice constants (C^2, 2C^4, 2C^2, 2C^6) with
// r = RN(1/y), in FP32: q0 = RN(x r), rem = x - q0 y (exact with one FMA),
// q = RN(q0 + rem r). An exhaustive CPU check over every FP32 x whose
// biased exponent lies in [25, 251] shows q equal to the IEEE quotient x / y
// for each constant of D2Q9 and of D2Q13 (tests/fastdiv_check.c); for
// x = +0 it returns +0, as the IEEE division does.
__device__ __forceinline__ float fast_div(float x, float y, float r) {
  const float q0 = x * r;
  const float rem = __builtin_fmaf(-q0, y, x);
  return __builtin_fmaf(rem, r, q0);
}

// (W rho) * ((((1 + cu/C^2) + cu^2/(2C^4)) - uu/(2C^2)) [D2Q13: + cu^3/(2C^6)
// - 3 cu uu/(2C^4)]), cu = ksi.u and uu = u.u as 0-initialised sums,
// cu^3 = cu*(cu*cu) (src/dynamics.py:70-74, 101-102).
// Opposite velocities give exactly negated cu (round-to-nearest is
// sign-symmetric), hence negated cu/C^2, cu^3/(2C^6) and 3 cu uu/(2C^4) and
// the same cu^2/(2C^4): each quotient is formed once per pair of opposite
// directions and each bracket is still summed in the reference's order for
// every q. A literal per-q evaluation can differ only in the sign of a zero
// cu, which leaves every bracket unchanged.
// q = 0: cu0 = (0 + 0 ux) + 0 uy is +0 (or NaN), and dividing +0 or NaN by
// a constant returns it unchanged, so its bracket uses cu0 in place of each
// quotient: the literal values without the four divisions.
//
// FP32 with uu < 1/2 (false for inf and NaN) divides with fast_div, and the
// brackets equal the IEEE ones: every dividend (cu, cu^2, cu^3, uu,
// 3 cu uu; none is -0, the literal sums never are) is below 4. A dividend
// of at least 2^-102 is divided exactly. A smaller one (zero included) has
// a quotient below 2^-98 by either method, and it is added where the
// partial bracket is at least 1/4 in magnitude, below half its ulp, so it
// is absorbed either way: A, B (|cu| < 2^-51) and D (|u| < 2^-51) where the
// partial sum is 1; E (|cu| < 2^-34, so A and B are absorbed) to
// 1 - D >= 1/2 - 2^-24; G to that same bracket, or, when |cu| >= 2^-34, to
// 1 (then uu < 2^-69, so |u| < 2^-34 and A, B, D, E are all below 2^-26).
template <bool FAST, typename T, int QN>
__device__ __forceinline__ void equilibrium_div(const Consts<T, QN> &k, T rho, T ux, T uy, T uu,
                                                T (&out)[QN]) {
  constexpr int kPairs = QN == 13 ? 6 : 4;
  auto div = [](T x, T y, T r) {
    if constexpr (FAST)
      return fast_div(x, y, r);
    else
      return x / y;
  };
  const T D = div(uu, k.C2x2, k.rC2x2);
  {
    T c0 = T(0);
    c0 += T(0) * ux;
    c0 += T(0) * uy;
    T br = ((T(1) + c0) + c0 * c0) - D;
    if constexpr (QN == 13) br = (br + c0 * (c0 * c0)) - (T(3) * c0) * uu;
    out[0] = (k.W[0] * rho) * br;
  }
#pragma unroll
  for (int i = 0; i < kPairs; ++i) {
    const int qp = kPos[i], qn = kNeg[i];
    T cu = T(0);
    cu += T(kx(qp)) * ux;
    cu += T(ky(qp)) * uy;
    const T A = div(cu, k.C2, k.rC2), B = div(cu * cu, k.C4x2, k.rC4x2);
    T bp = ((T(1) + A) + B) - D;
    T bn = ((T(1) + (-A)) + B) - D;
    if constexpr (QN == 13) {
      const T E = div(cu * (cu * cu), k.C6x2, k.rC6x2);
      const T G = div((T(3) * cu) * uu, k.C4x2, k.rC4x2);
      bp = (bp + E) - G;
      bn = (bn + (-E)) - (-G);
    }
    out[qp] = (k.W[qp] * rho) * bp;
    out[qn] = (k.W[qn] * rho) * bn;
  }
}

template <typename T, int QN>
__device__ __forceinline__ void equilibrium(const Consts<T, QN> &k, T rho, T ux, T uy,
                                            T (&out)[QN]) {
  T uu = T(0);
  uu += ux * ux;
  uu += uy * uy;
  if constexpr (std::is_same<T, float>::value) {
    if (uu < 0.5f) {
      equilibrium_div<true>(k, rho, ux, uy, uu, out);
      return;
    }
  }
  equilibrium_div<false>(k, rho, ux, uy, uu, out);
}

// Upwind ksi.n for q >= 1: opposite velocities give exact negatives and
// ksi = (+-2, 0), (0, +-2) gives 2n exactly, so these equal the
// 0-initialised dot products (src/containers.py:235) up to the sign of a
// zero; a zero of either sign selects the same upwind side (phi >= 0) and
// makes a zero flux, which an accumulator started at +0 absorbs alike.
// q = 0 (ksi = 0) has a zero flux, and the callers skip it.
template <typename T, int QN>
__device__ __forceinline__ void upwind_phi(T nx, T ny, T (&phi)[QN]) {
  const T p5 = nx + ny, p6 = ny - nx;
  phi[0] = T(0);
  phi[1] = nx;
  phi[2] = ny;
  phi[3] = -nx;
  phi[4] = -ny;
  phi[5] = p5;
  phi[6] = p6;
  phi[7] = -p5;
  phi[8] = -p6;
  if constexpr (QN == 13) {
    const T a = nx + nx, b = ny + ny;
    phi[9] = a;
    phi[10] = b;
    phi[11] = -a;
    phi[12] = -b;
  }
}

template <typename T>
struct Args {
  const T *fsrc;
  T *fdst;
  int C, Cp, N, K, nbf, nan;
  const int32_t *code;  // [3][Cp] neighbour device id | slot-1 bit, or ghost bit | record
  const T *nx, *ny, *sL, *lwa, *lwden2;  // [3][Cp]
  const int32_t *bf_gnode;               // [2][nbf] face-node ids of each ghost face
  const T *bf_ratio;                     // [nbf] d_ghost / d_known
  T *node_pdf, *node_rho, *node_vel;     // [Q][N], [N], [2][N]
  const int32_t *an_gid, *an_type, *an_cells;
  const T *an_w, *an_wsum, *an_rho, *an_ux, *an_uy;
  const int32_t *block_tile;  // [grid] tile (block of cells) of each workgroup, -1 = none
  // per-tile node tables: entries of tile t are [tile_node_off[t],
  // tile_node_off[t+1]) of tile_nodes (>= 0 active node, < 0: -1 - node id
  // with a stored pdf); bf_slot [2][nbf] = a ghost face's two entries
  const int32_t *tile_node_off, *tile_nodes, *bf_slot;
};

template <typename T, int QN>
__device__ __forceinline__ void load_cell(const T *f, size_t Cp, int c, T (&out)[QN]) {
#pragma unroll
  for (int q = 0; q < QN; ++q) out[q] = f[(size_t)q * Cp + c];
}

// Active node `an` from fsrc: type 1 (velocity) rho <- IDW(cell rho), type
// 2 (density) u <- IDW(cell u); pdf = calc_eq(rho, u) + IDW(f_c - feq_c)
// (src/containers.py:339-390). Padding entries (weight 0) contribute +0,
// as x*0 does in the reference. Sixteen lanes per node: lane j computes the
// weighted contributions of the node's j-th listed cell (loads and
// arithmetic in parallel), then every lane of the group adds them up in
// list order through shuffles, so each sum is the reference's
// 0-initialised ordered sum. Called by every lane of the wave in uniform
// control flow (the shuffles); `active` false gives zeros.
constexpr int kNodeLanes = 16;

template <typename T, int QN>
__device__ __forceinline__ void group_node(const Args<T> &a, const Consts<T, QN> &k, int an,
                                           bool active, T (&out)[QN], T &rho_n, T &ux_n,
                                           T &uy_n) {
  const int lane = threadIdx.x & (kNodeLanes - 1);
  const int base = (threadIdx.x & 63) & ~(kNodeLanes - 1);  // the group's first lane in the wave
  const int type = active ? a.an_type[an] : 1;
  const int32_t *cells = a.an_cells + (size_t)(active ? an : 0) * a.K;
  const T *w = a.an_w + (size_t)(active ? an : 0) * a.K;
  T n0 = T(0), n1 = T(0), acc[QN];
#pragma unroll
  for (int q = 0; q < QN; ++q) acc[q] = T(0);
  for (int j0 = 0; j0 < a.K; j0 += kNodeLanes) {
    const int j = j0 + lane;
    T vr = T(0), vx = T(0), vy = T(0), vq[QN];
#pragma unroll
    for (int q = 0; q < QN; ++q) vq[q] = T(0);
    if (active && j < a.K) {
      const int cj = cells[j];
      const T wj = w[j];
      T r = T(0), ux = T(0), uy = T(0), fc[QN];
      if (!(cj < 0 && wj == T(0))) {
        load_cell<T, QN>(a.fsrc, a.Cp, cj < 0 ? -2 - cj : cj, fc);
        macro<T, QN>(fc, r, ux, uy);
      }
      vr = r * wj;
      vx = ux * wj;
      vy = uy * wj;
      if (cj < 0) {
#pragma unroll
        for (int q = 0; q < QN; ++q) vq[q] = T(0) * wj;
      } else {
        T eq[QN];
        equilibrium<T, QN>(k, r, ux, uy, eq);
#pragma unroll
        for (int q = 0; q < QN; ++q) vq[q] = (fc[q] - eq[q]) * wj;
      }
    }
    const int nj = a.K - j0 < kNodeLanes ? a.K - j0 : kNodeLanes;
    for (int t = 0; t < nj; ++t) {
      const int src = base + t;
      const T sr = __shfl(vr, src), sx = __shfl(vx, src), sy = __shfl(vy, src);
      if (type == 1) {
        n0 += sr;
      } else {
        n0 += sx;
        n1 += sy;
      }
#pragma unroll
      for (int q = 0; q < QN; ++q) acc[q] += __shfl(vq[q], src);
    }
  }
  const T wsum = active ? a.an_wsum[an] : T(1);
  rho_n = active ? a.an_rho[an] : T(1);
  ux_n = active ? a.an_ux[an] : T(0);
  uy_n = active ? a.an_uy[an] : T(0);
  if (type == 1) {
    rho_n = n0 / wsum;
  } else {
    ux_n = n0 / wsum;
    uy_n = n1 / wsum;
  }
  T eqn[QN];
  equilibrium<T, QN>(k, rho_n, ux_n, uy_n, eqn);
#pragma unroll
  for (int q = 0; q < QN; ++q) out[q] = eqn[q] + acc[q] / wsum;
}

// Every active node from fsrc into the node arrays (the observables; the
// step computes the values it needs per tile). 16 nodes per 256 threads.
template <typename T, int QN>
__global__ void __launch_bounds__(256) k_lat_nodes(Args<T> a, Consts<T, QN> k) {
  const int an = blockIdx.x * (blockDim.x / kNodeLanes) + (int)(threadIdx.x / kNodeLanes);
  const bool active = an < a.nan;  // uniform per group; no return before the shuffles
  T out[QN], rho_n, ux_n, uy_n;
  group_node<T, QN>(a, k, an, active, out, rho_n, ux_n, uy_n);
  if (!active || (threadIdx.x & (kNodeLanes - 1)) != 0) return;
  const int g = a.an_gid[an];
#pragma unroll
  for (int q = 0; q < QN; ++q) a.node_pdf[(size_t)q * a.N + g] = out[q];
  if (a.an_type[an] == 1) {
    a.node_rho[g] = rho_n;
  } else {
    a.node_vel[g] = ux_n;
    a.node_vel[(size_t)a.N + g] = uy_n;
  }
}

// One cell per thread, tiles of blockDim.x consecutive cells taken in the
// XCD-aware order of block_tile (workgroups go round-robin to the 8 XCDs;
// each XCD gets one contiguous range of tiles, so neighbour gathers mostly
// hit its own L2). A tile with ghost faces first evaluates the distinct
// face nodes they read into LDS (active nodes by group_node, stored pdfs
// of the others). The three faces in cells.face_indices order (ghost
// g = fbar + (fbar - f_self) * ratio from the two face-node pdfs), upwind
// by the sign of ksi.n (SCHEME 0) or Lax-Wendroff with slot 0 upstream
// (SCHEME 1), flux (pdf*phi)*(s*L) summed in face order, then
// f + dt*(1/tau*(feq - f) - sum) into fdst.
template <typename T, int QN, int SCHEME>
__global__ void __launch_bounds__(1024) k_lat_step(Args<T> a, Consts<T, QN> k) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *tab = reinterpret_cast<T *>(smem_raw);  // [QN][ntn] node pdfs of the tile's ghost faces
  const int tile = a.block_tile[blockIdx.x];
  if (tile < 0) return;  // uniform per block
  const int t0 = a.tile_node_off[tile], ntn = a.tile_node_off[tile + 1] - t0;
  if (ntn > 0) {  // uniform per block: the node values this tile's ghost faces read
    const int groups = blockDim.x / kNodeLanes, group = threadIdx.x / kNodeLanes;
    const bool lead = (threadIdx.x & (kNodeLanes - 1)) == 0;
    for (int i0 = 0; i0 < ntn; i0 += groups) {
      const int i = i0 + group;
      const int e = i < ntn ? a.tile_nodes[t0 + i] : -1;
      T out[QN], r, ux, uy;
      group_node<T, QN>(a, k, e, i < ntn && e >= 0, out, r, ux, uy);
      if (lead && i < ntn) {
#pragma unroll
        for (int q = 0; q < QN; ++q)
          tab[q * ntn + i] = e >= 0 ? out[q] : a.node_pdf[(size_t)q * a.N + (-1 - e)];
      }
    }
    __syncthreads();
  }
  const int c = tile * blockDim.x + threadIdx.x;
  if (c >= a.C) return;  // after the only barrier
  const size_t Cp = a.Cp;
  T f[QN], acc[QN];
  load_cell<T, QN>(a.fsrc, Cp, c, f);
#pragma unroll
  for (int q = 0; q < QN; ++q) acc[q] = T(0);
  for (int kf = 0; kf < 3; ++kf) {
    const size_t at = (size_t)kf * Cp + c;
    const int32_t code = a.code[at];
    const T nx = a.nx[at], ny = a.ny[at], sL = a.sL[at];
    const bool slot1 = (code & kSlot1Bit) != 0;
    T o[QN];
    if (code >= 0) {
      load_cell<T, QN>(a.fsrc, Cp, code & kIndexMask, o);
    } else {
      const int bf = code & kIndexMask;
      const T ratio = a.bf_ratio[bf];
      const int s0 = a.bf_slot[bf], s1 = a.bf_slot[a.nbf + bf];
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        const T fbar = (tab[q * ntn + s0] + tab[q * ntn + s1]) / T(2);
        o[q] = fbar + (fbar - f[q]) * ratio;
      }
    }
    T la = T(0), den2 = T(1);
    if constexpr (SCHEME == 1) {
      la = a.lwa[at];
      den2 = a.lwden2[at];
    }
    if constexpr (SCHEME == 0) {
      T phi[QN];
      upwind_phi<T, QN>(nx, ny, phi);
#pragma unroll
      for (int q = 1; q < QN; ++q) {
        const bool take_slot0 = phi[q] >= T(0);
        const T up = (take_slot0 != slot1) ? f[q] : o[q];
        acc[q] += (up * phi[q]) * sL;
      }
    } else {
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        T phi = T(0);
        phi += T(kx(q)) * nx;
        phi += T(ky(q)) * ny;
        const T fu = slot1 ? o[q] : f[q];
        const T fd = slot1 ? f[q] : o[q];
        const T interp = la - (phi * k.dt) / den2;
        const T fcv = fu + (fd - fu) * interp;
        acc[q] += (fcv * phi) * sL;
      }
    }
  }
  T rho, ux, uy, eq[QN];
  macro<T, QN>(f, rho, ux, uy);
  equilibrium<T, QN>(k, rho, ux, uy, eq);
#pragma unroll
  for (int q = 0; q < QN; ++q)
    a.fdst[(size_t)q * Cp + c] = f[q] + k.dt * (k.invtau * (eq[q] - f[q]) - acc[q]);
}

// ----------------------------------------------------------------------------
// two-step tiles (upwind): one launch advances a tile by two steps
// ----------------------------------------------------------------------------
// The ring lists, ring face records and extended node tables come from the
// host layout builder (the tuned D2Q9 kernel uses the same ones): with O a
// tile's cells, S1 = O, its face neighbours and every cell a node of an O
// ghost face reads, S2 = the face neighbours of S1 outside it. Step n runs
// for S1 from the inputs of S1 and S2 staged in LDS, the results stay in
// LDS, then step n + 1 runs for O. Every value equals what two k_lat_step
// launches compute.
// Every face code of the two-step kernel is a stage slot (| slot-1 bit):
// own cells 0..bs-1, the tile's ring from bs, then one ghost record per
// ghost face of S1 (the own cells' first, in boundary-face order, then the
// S1 \ O cells').
template <typename T>
struct PArgs {
  const int32_t *code2;                    // [3][Cp] own face codes with stage slots
  const int32_t *ring_off, *ring_cells, *ring_n1, *rg_off;
  const int32_t *rg_code;                  // [3][nrgc] face records of the S1 \ O cells
  const T *rg_nx, *rg_ny, *rg_sL;
  const int32_t *gh_off, *gh_nown;         // [ntiles+1] ghost lists; [ntiles] own ghosts (first)
  const int32_t *gh_s0, *gh_s1, *gh_self;  // table slots of the face nodes; owning cell's slot
  const T *gh_ratio;                       // ghost distance / known distance
  const int32_t *xnode_off, *xnodes, *xentry_pair_off, *xpair_off, *xpair_cell;
  const T *xpair_w;
  const int32_t *entry_pair_off, *tile_pair_off, *pair_cell, *pair_slot;
  const T *pair_w;
  int nrgc, max_xnodes, max_xpairs;
};

// Node table entries [0, nent) into tab (rows of length nrow) from their
// (node, cell) pairs [pb, pb + np) (offsets epo absolute): one thread per
// pair puts the pair's weighted contributions [r w, ux w, uy w, (f - feq) w]
// (its cell read from fsrc, or from the stage slot pslot when FROM_LDS)
// into an LDS record of PairRec<QN>::W words, then one thread per entry
// adds its pairs' records in pair order: group_node's arithmetic (pairs of
// weight-0 padding are left out: they add +0). Ends with a barrier.
template <int QN>
struct PairRec {
  static constexpr int W = 3 + QN + ((3 + QN) % 2 == 0 ? 1 : 0);  // odd stride: spread LDS banks
};

template <typename T, int QN, bool FROM_LDS>
__device__ __forceinline__ void lat_table_rec(const Args<T> &a, const Consts<T, QN> &k, T *tab,
                                              int nrow, T *con,
                                              const int32_t *__restrict__ entries, int nent,
                                              const int32_t *__restrict__ epo, int pb, int np,
                                              const int32_t *__restrict__ pcell,
                                              const int32_t *__restrict__ pslot,
                                              const T *__restrict__ pw, const T *stage) {
  constexpr int RW = PairRec<QN>::W;
  const int bs = blockDim.x;
  for (int pi = threadIdx.x; pi < np; pi += bs) {
    const int cj = pcell[pb + pi];
    const T wj = pw[pb + pi];
    T fc[QN], r, ux, uy;
    if constexpr (FROM_LDS) {
      const int sl = pslot[pb + pi];
#pragma unroll
      for (int q = 0; q < QN; ++q) fc[q] = stage[sl * QN + q];
    } else {
      load_cell<T, QN>(a.fsrc, a.Cp, cj < 0 ? -2 - cj : cj, fc);
    }
    macro<T, QN>(fc, r, ux, uy);
    T *o = con + RW * pi;
    o[0] = r * wj;
    o[1] = ux * wj;
    o[2] = uy * wj;
    if (cj >= 0) {
      T eq[QN];
      equilibrium<T, QN>(k, r, ux, uy, eq);
#pragma unroll
      for (int q = 0; q < QN; ++q) o[3 + q] = (fc[q] - eq[q]) * wj;
    } else {
#pragma unroll
      for (int q = 0; q < QN; ++q) o[3 + q] = T(0) * wj;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nent; i += bs) {
    const int e = entries[i];
    T out[QN];
    if (e >= 0) {
      const int type = a.an_type[e];
      const int p0 = epo[i] - pb, p1 = epo[i + 1] - pb;
      T n0 = T(0), n1 = T(0), acc[QN];
#pragma unroll
      for (int q = 0; q < QN; ++q) acc[q] = T(0);
      for (int pi = p0; pi < p1; ++pi) {
        const T *o = con + RW * pi;
        if (type == 1) {
          n0 += o[0];
        } else {
          n0 += o[1];
          n1 += o[2];
        }
#pragma unroll
        for (int q = 0; q < QN; ++q) acc[q] += o[3 + q];
      }
      const T wsum = a.an_wsum[e];
      T rho_n = a.an_rho[e], ux_n = a.an_ux[e], uy_n = a.an_uy[e];
      if (type == 1) {
        rho_n = n0 / wsum;
      } else {
        ux_n = n0 / wsum;
        uy_n = n1 / wsum;
      }
      T eqn[QN];
      equilibrium<T, QN>(k, rho_n, ux_n, uy_n, eqn);
#pragma unroll
      for (int q = 0; q < QN; ++q) out[q] = eqn[q] + acc[q] / wsum;
    } else {
#pragma unroll
      for (int q = 0; q < QN; ++q) out[q] = a.node_pdf[(size_t)q * a.N + (-1 - e)];
    }
#pragma unroll
    for (int q = 0; q < QN; ++q) tab[q * nrow + i] = out[q];
  }
  __syncthreads();
}

// Ghost records of one step (Faces.calc_ghost + utils.extrap_pdf,
// src/containers.py:280-287, utils/utils.py:153-154): ghost face i of
// [g0, g0 + n) gets fbar = mean of its two node rows in tab (rows of
// length nrow) and g = fbar + (fbar - f_self) * ratio, with f_self the
// owning cell's stage record, stored at stage slot slot0 + i. One lane per
// ghost face, so the cell updates read every face neighbour from the stage.
template <typename T, int QN>
__device__ __forceinline__ void lat_ghost_records(T *stage, const T *tab, int nrow,
                                                  const PArgs<T> &p, int g0, int n, int slot0) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int s0 = p.gh_s0[g0 + i], s1 = p.gh_s1[g0 + i], self = p.gh_self[g0 + i];
    const T ratio = p.gh_ratio[g0 + i];
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      const T fbar = (tab[q * nrow + s0] + tab[q * nrow + s1]) / T(2);
      stage[(slot0 + i) * QN + q] = fbar + (fbar - stage[self * QN + q]) * ratio;
    }
  }
}

// One upwind update of a cell whose face neighbours, ghost records
// included, are in the stage ([slot][QN]): k_lat_step's arithmetic.
template <typename T, int QN>
__device__ __forceinline__ void lat_update_staged(const Consts<T, QN> &k, const T (&f)[QN],
                                                  const int32_t (&code)[3], const T (&nx)[3],
                                                  const T (&ny)[3], const T (&sL)[3],
                                                  const T *stage, T (&fn)[QN]) {
  T acc[QN];
#pragma unroll
  for (int q = 0; q < QN; ++q) acc[q] = T(0);
#pragma unroll
  for (int kf = 0; kf < 3; ++kf) {
    const bool slot1 = (code[kf] & kSlot1Bit) != 0;
    T o[QN];
    {
      const int sl = code[kf] & kIndexMask;
#pragma unroll
      for (int q = 0; q < QN; ++q) o[q] = stage[sl * QN + q];
    }
    T phi[QN];
    upwind_phi<T, QN>(nx[kf], ny[kf], phi);
#pragma unroll
    for (int q = 1; q < QN; ++q) {
      const bool take_slot0 = phi[q] >= T(0);
      const T up = (take_slot0 != slot1) ? f[q] : o[q];
      acc[q] += (up * phi[q]) * sL[kf];
    }
  }
  T rho, ux, uy, eq[QN];
  macro<T, QN>(f, rho, ux, uy);
  equilibrium<T, QN>(k, rho, ux, uy, eq);
#pragma unroll
  for (int q = 0; q < QN; ++q) fn[q] = f[q] + k.dt * (k.invtau * (eq[q] - f[q]) - acc[q]);
}

template <typename T, int QN>
__global__ void __launch_bounds__(1024) k_lat_step2(Args<T> a, PArgs<T> p, Consts<T, QN> k) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *tab = reinterpret_cast<T *>(smem_raw);              // [QN][nx] extended node table
  T *con = tab + QN * p.max_xnodes;                      // pair records of a node table
  T *stage = con + PairRec<QN>::W * p.max_xpairs;        // [slot][QN]: O, then S1 \ O and S2
  const int tile = a.block_tile[blockIdx.x];
  if (tile < 0) return;  // uniform per block
  const int bs = blockDim.x;
  const size_t Cp = a.Cp;
  const int c = tile * bs + threadIdx.x;
  const bool live = c < a.C;
  const int x0 = p.xnode_off[tile], nx = p.xnode_off[tile + 1] - x0;
  const int t0 = a.tile_node_off[tile], ntn = a.tile_node_off[tile + 1] - t0;
  // step-n node values of every ghost face of S1, from global memory
  if (nx > 0) {
    const int pb = p.xpair_off[tile];
    lat_table_rec<T, QN, false>(a, k, tab, nx, con, p.xnodes + x0, nx, p.xentry_pair_off + x0, pb,
                                p.xpair_off[tile + 1] - pb, p.xpair_cell, nullptr, p.xpair_w, stage);
  }
  // stage: own cells, then S1 \ O and S2; face records in registers
  T f[QN];
  int32_t cd[3] = {0, 0, 0};
  T fnx[3] = {}, fny[3] = {}, fsl[3] = {};
  if (live) {
    load_cell<T, QN>(a.fsrc, Cp, c, f);
#pragma unroll
    for (int kf = 0; kf < 3; ++kf) {
      const size_t at = (size_t)kf * Cp + c;
      cd[kf] = p.code2[at];
      fnx[kf] = a.nx[at];
      fny[kf] = a.ny[at];
      fsl[kf] = a.sL[at];
    }
  } else {
#pragma unroll
    for (int q = 0; q < QN; ++q) f[q] = T(0);
  }
#pragma unroll
  for (int q = 0; q < QN; ++q) stage[threadIdx.x * QN + q] = f[q];
  const int r0 = p.ring_off[tile], nr = p.ring_off[tile + 1] - r0, n1 = p.ring_n1[tile];
  for (int i = threadIdx.x; i < nr; i += bs) {
    T g[QN];
    load_cell<T, QN>(a.fsrc, Cp, p.ring_cells[r0 + i], g);
#pragma unroll
    for (int q = 0; q < QN; ++q) stage[(bs + i) * QN + q] = g[q];
  }
  const int g0 = p.gh_off[tile], ng = p.gh_off[tile + 1] - g0, ngo = p.gh_nown[tile];
  const bool ring_here = (int)threadIdx.x < n1;
  int32_t rc[3] = {0, 0, 0};
  T rnx[3] = {}, rny[3] = {}, rsl[3] = {};
  if (ring_here) {
    const int rr = p.rg_off[tile] + threadIdx.x;
#pragma unroll
    for (int kf = 0; kf < 3; ++kf) {
      const size_t at = (size_t)kf * p.nrgc + rr;
      rc[kf] = p.rg_code[at];
      rnx[kf] = p.rg_nx[at];
      rny[kf] = p.rg_ny[at];
      rsl[kf] = p.rg_sL[at];
    }
  }
  __syncthreads();
  if (ng > 0) {  // uniform per block
    lat_ghost_records<T, QN>(stage, tab, nx, p, g0, ng, bs + nr);
    __syncthreads();
  }
  // step n for the own cell and the ring cell of this thread
  T fo[QN], fq[QN];
  if (live) lat_update_staged<T, QN>(k, f, cd, fnx, fny, fsl, stage, fo);
  if (ring_here) {
    T g[QN];
#pragma unroll
    for (int q = 0; q < QN; ++q) g[q] = stage[(bs + threadIdx.x) * QN + q];
    lat_update_staged<T, QN>(k, g, rc, rnx, rny, rsl, stage, fq);
  }
  __syncthreads();  // every step-n input has been read
  if (live) {
#pragma unroll
    for (int q = 0; q < QN; ++q) stage[threadIdx.x * QN + q] = fo[q];
  }
  if (ring_here) {
#pragma unroll
    for (int q = 0; q < QN; ++q) stage[(bs + threadIdx.x) * QN + q] = fq[q];
  }
  __syncthreads();
  // step-(n+1) node values of the own ghost faces, from the stage
  if (ntn > 0) {
    const int pb = p.tile_pair_off[tile];
    lat_table_rec<T, QN, true>(a, k, tab, nx, con, a.tile_nodes + t0, ntn, p.entry_pair_off + t0,
                               pb, p.tile_pair_off[tile + 1] - pb, p.pair_cell, p.pair_slot,
                               p.pair_w, stage);
  }
  if (ngo > 0) {  // the own cells' ghost records from the step-(n+1) table
    lat_ghost_records<T, QN>(stage, tab, nx, p, g0, ngo, bs + nr);
    __syncthreads();
  }
  if (live) {
    T fn[QN];
    lat_update_staged<T, QN>(k, fo, cd, fnx, fny, fsl, stage, fn);
#pragma unroll
    for (int q = 0; q < QN; ++q) a.fdst[(size_t)q * Cp + c] = fn[q];
  }
}

// cells.rho / vel / pdf_eq of fsrc (device order)
template <typename T, int QN>
__global__ void k_lat_moments(Args<T> a, Consts<T, QN> k, T *rho, T *vel, T *feq) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.C) return;
  T f[QN], eq[QN], r, ux, uy;
  load_cell<T, QN>(a.fsrc, a.Cp, c, f);
  macro<T, QN>(f, r, ux, uy);
  equilibrium<T, QN>(k, r, ux, uy, eq);
  rho[c] = r;
  vel[c] = ux;
  vel[(size_t)a.Cp + c] = uy;
#pragma unroll
  for (int q = 0; q < QN; ++q) feq[(size_t)q * a.Cp + c] = eq[q];
}

// faces.pdf in reference face order from fsrc and the node pdfs
// (calc_upwind_flux / calc_lax_wendroff_flux face by face, ghosts on
// either slot, index -1 of a missing known cell read as the last cell)
struct FaceArrays {
  const int32_t *c0, *c1, *n0, *n1;
  const void *d0, *d1, *nx, *ny, *L;
  int F, last_cell;
};

template <typename T, int QN, int SCHEME>
__global__ void k_lat_flux(Args<T> a, Consts<T, QN> k, FaceArrays fa, T *flux) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= fa.F) return;
  const int ia = fa.c0[j], ib = fa.c1[j];
  const T d0 = static_cast<const T *>(fa.d0)[j], d1 = static_cast<const T *>(fa.d1)[j];
  const T nx = static_cast<const T *>(fa.nx)[j], ny = static_cast<const T *>(fa.ny)[j];
  const T L = static_cast<const T *>(fa.L)[j];
  T fl[QN], fr[QN];
  for (int side = 0; side < 2; ++side) {
    const int idx = side == 0 ? ia : ib;
    T tmp[QN];
    if (idx != -1) {
      load_cell<T, QN>(a.fsrc, a.Cp, idx, tmp);
    } else {
      int known = side == 0 ? ib : ia;
      if (known < 0) known = fa.last_cell;
      const T gd = side == 0 ? d0 : d1, kd = side == 0 ? d1 : d0;
      const T ratio = gd / kd;
      T fk[QN];
      load_cell<T, QN>(a.fsrc, a.Cp, known, fk);
      const int p = fa.n0[j], r = fa.n1[j];
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        const T fbar = (a.node_pdf[(size_t)q * a.N + p] + a.node_pdf[(size_t)q * a.N + r]) / T(2);
        tmp[q] = fbar + (fbar - fk[q]) * ratio;
      }
    }
#pragma unroll
    for (int q = 0; q < QN; ++q) (side == 0 ? fl : fr)[q] = tmp[q];
  }
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    T phi = T(0);
    phi += T(kx(q)) * nx;
    phi += T(ky(q)) * ny;
    T v;
    if constexpr (SCHEME == 0) {
      v = ((phi >= T(0)) ? fl[q] : fr[q]) * phi;
    } else {
      const T s = d0 + d1;
      const T interp = (d0 / s) - (phi * k.dt) / (T(2) * s);
      v = (fl[q] + (fr[q] - fl[q]) * interp) * phi;
    }
    flux[(size_t)q * fa.F + j] = v * L;
  }
}

// rows [0, n) of a q-major [Q][stride] array <- W
template <typename T, int QN>
__global__ void k_lat_fill(T *f, int n, int stride, Consts<T, QN> k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int q = 0; q < QN; ++q) f[(size_t)q * stride + i] = k.W[q];
}

// ----------------------------------------------------------------------------
// handle
// ----------------------------------------------------------------------------
struct Handle {
  int q = 13, precision = FVDBM_FP32, scheme = FVDBM_UPWIND, block = 256, graph_steps = 32;
  double tau = 1.0, dt = 1.0;
  int C = 0, Cp = 0, F = 0, N = 0, K = 0, nbf = 0, nan = 0, nbc = 0, bandwidth = 0,
      last_cell_new = 0;
  std::vector<int32_t> new_to_old, an_gid_host;
  void *f[2] = {nullptr, nullptr};
  void *code = nullptr, *nx = nullptr, *ny = nullptr, *sL = nullptr, *lwa = nullptr,
       *lwden2 = nullptr, *bf_gnode = nullptr, *bf_ratio = nullptr;
  void *node_pdf = nullptr, *node_rho = nullptr, *node_vel = nullptr;
  void *an_gid = nullptr, *an_type = nullptr, *an_cells = nullptr, *an_w = nullptr,
       *an_wsum = nullptr, *an_rho = nullptr, *an_ux = nullptr, *an_uy = nullptr;
  void *fc0 = nullptr, *fc1 = nullptr, *fn0 = nullptr, *fn1 = nullptr, *fd0 = nullptr,
       *fd1 = nullptr, *fnx = nullptr, *fny = nullptr, *fL = nullptr;
  void *obs_rho = nullptr, *obs_vel = nullptr, *obs_feq = nullptr, *obs_flux = nullptr;
  void *block_tile = nullptr, *tile_node_off = nullptr, *tile_nodes = nullptr, *bf_slot = nullptr;
  int grid = 1, max_tile_nodes = 0;
  // two-step tiles (upwind, when the layout's rings fit and the LDS does)
  bool pairs = false, two_step = true;  // two_step: wanted (params.variant bit 8 clear)
  size_t lds2 = 0;
  int nrgc = 0, max_xnodes = 0, max_xpairs = 0;
  void *code2 = nullptr, *ring_off = nullptr, *ring_cells = nullptr, *ring_n1 = nullptr,
       *rg_off = nullptr, *rg_code = nullptr, *rg_nx = nullptr, *rg_ny = nullptr, *rg_sL = nullptr,
       *gh_off = nullptr, *gh_nown = nullptr, *gh_s0 = nullptr, *gh_s1 = nullptr,
       *gh_self = nullptr, *gh_ratio = nullptr, *xnode_off = nullptr, *xnodes = nullptr,
       *xentry_pair_off = nullptr, *xpair_off = nullptr, *xpair_cell = nullptr, *xpair_w = nullptr,
       *entry_pair_off = nullptr, *tile_pair_off = nullptr, *pair_cell = nullptr,
       *pair_slot = nullptr, *pair_w = nullptr;
  // captured runs of two-step launches, keyed by (starting parity, launches)
  std::map<std::pair<int, int>, std::pair<hipGraph_t, hipGraphExec_t>> graphs2;
  // two-step handles: after a run that ended in a two-step launch
  // (prev_two_back), f[cur ^ 1] holds the pdf two steps back; the pdf at the
  // start of the last step is one step from it, computed into fmid only
  // when an observable is read (mid_ready)
  void *fmid = nullptr;
  bool prev_two_back = false, mid_ready = false;
  int cur = 0;
  bool has_prev = false, observed = false;
  int64_t steps = 0, device_bytes = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<void *> bufs;

  ~Handle() {
    // steps are queued asynchronously: let them finish before their graphs,
    // buffers and stream go
    if (stream) (void)hipStreamSynchronize(stream);
    for (int i = 0; i < 2; ++i) {
      if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
      if (graph[i]) (void)hipGraphDestroy(graph[i]);
    }
    for (auto &kv : graphs2) {
      if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
      if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
    }
    for (void *b : bufs) (void)hipFree(b);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int alloc(void **out, size_t bytes) {
    if (bytes == 0) bytes = 256;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(FVDBM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    bufs.push_back(p);
    device_bytes += (int64_t)bytes;
    *out = p;
    return FVDBM_OK;
  }
  template <typename V>
  int upload(void **out, const std::vector<V> &v) {
    int rc = alloc(out, v.size() * sizeof(V));
    if (rc) return rc;
    if (!v.empty()) LHIP_TRY(hipMemcpy(*out, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    return FVDBM_OK;
  }
};

template <typename T>
static Args<T> make_args(const Handle &h, int src) {
  Args<T> a;
  a.fsrc = static_cast<const T *>(h.f[src]);
  a.fdst = static_cast<T *>(h.f[src ^ 1]);
  a.C = h.C;
  a.Cp = h.Cp;
  a.N = h.N;
  a.K = h.K;
  a.nbf = h.nbf;
  a.nan = h.nan;
  a.code = static_cast<const int32_t *>(h.code);
  a.nx = static_cast<const T *>(h.nx);
  a.ny = static_cast<const T *>(h.ny);
  a.sL = static_cast<const T *>(h.sL);
  a.lwa = static_cast<const T *>(h.lwa);
  a.lwden2 = static_cast<const T *>(h.lwden2);
  a.bf_gnode = static_cast<const int32_t *>(h.bf_gnode);
  a.bf_ratio = static_cast<const T *>(h.bf_ratio);
  a.node_pdf = static_cast<T *>(h.node_pdf);
  a.node_rho = static_cast<T *>(h.node_rho);
  a.node_vel = static_cast<T *>(h.node_vel);
  a.an_gid = static_cast<const int32_t *>(h.an_gid);
  a.an_type = static_cast<const int32_t *>(h.an_type);
  a.an_cells = static_cast<const int32_t *>(h.an_cells);
  a.an_w = static_cast<const T *>(h.an_w);
  a.an_wsum = static_cast<const T *>(h.an_wsum);
  a.an_rho = static_cast<const T *>(h.an_rho);
  a.an_ux = static_cast<const T *>(h.an_ux);
  a.an_uy = static_cast<const T *>(h.an_uy);
  a.block_tile = static_cast<const int32_t *>(h.block_tile);
  a.tile_node_off = static_cast<const int32_t *>(h.tile_node_off);
  a.tile_nodes = static_cast<const int32_t *>(h.tile_nodes);
  a.bf_slot = static_cast<const int32_t *>(h.bf_slot);
  return a;
}

static inline dim3 grid1(int n) { return dim3((unsigned)std::max(1, (n + 255) / 256)); }

// one step reading f[src]: the active nodes, then the cells
template <typename T, int QN>
static int launch_step(Handle &h, int src, void *dst = nullptr) {
  Args<T> a = make_args<T>(h, src);
  if (dst) a.fdst = static_cast<T *>(dst);
  const Consts<T, QN> k = make_consts<T, QN>(h.tau, h.dt);
  const dim3 grid((unsigned)h.grid);
  const size_t lds = (size_t)QN * h.max_tile_nodes * sizeof(T);
  if (h.scheme == FVDBM_UPWIND)
    hipLaunchKernelGGL((k_lat_step<T, QN, 0>), grid, dim3(h.block), lds, h.stream, a, k);
  else
    hipLaunchKernelGGL((k_lat_step<T, QN, 1>), grid, dim3(h.block), lds, h.stream, a, k);
  LHIP_TRY(hipGetLastError());
  return FVDBM_OK;
}

template <typename F>
static int dispatch(const Handle &h, F &&fn) {
  if (h.precision == FVDBM_FP64) return h.q == 13 ? fn((double *)nullptr, std::integral_constant<int, 13>())
                                                   : fn((double *)nullptr, std::integral_constant<int, 9>());
  return h.q == 13 ? fn((float *)nullptr, std::integral_constant<int, 13>())
                   : fn((float *)nullptr, std::integral_constant<int, 9>());
}

static int step_any(Handle &h, int src) {
  return dispatch(h, [&](auto tp, auto qn) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_step<T, decltype(qn)::value>(h, src);
  });
}

// two steps reading f[src] and writing f[src ^ 1]
template <typename T, int QN>
static int launch_pair(Handle &h, int src) {
  const Args<T> a = make_args<T>(h, src);
  const Consts<T, QN> k = make_consts<T, QN>(h.tau, h.dt);
  PArgs<T> p;
  p.code2 = static_cast<const int32_t *>(h.code2);
  p.ring_off = static_cast<const int32_t *>(h.ring_off);
  p.ring_cells = static_cast<const int32_t *>(h.ring_cells);
  p.ring_n1 = static_cast<const int32_t *>(h.ring_n1);
  p.rg_off = static_cast<const int32_t *>(h.rg_off);
  p.rg_code = static_cast<const int32_t *>(h.rg_code);
  p.rg_nx = static_cast<const T *>(h.rg_nx);
  p.rg_ny = static_cast<const T *>(h.rg_ny);
  p.rg_sL = static_cast<const T *>(h.rg_sL);
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
  p.entry_pair_off = static_cast<const int32_t *>(h.entry_pair_off);
  p.tile_pair_off = static_cast<const int32_t *>(h.tile_pair_off);
  p.pair_cell = static_cast<const int32_t *>(h.pair_cell);
  p.pair_slot = static_cast<const int32_t *>(h.pair_slot);
  p.pair_w = static_cast<const T *>(h.pair_w);
  p.nrgc = h.nrgc;
  p.max_xnodes = h.max_xnodes;
  p.max_xpairs = h.max_xpairs;
  hipLaunchKernelGGL((k_lat_step2<T, QN>), dim3((unsigned)h.grid), dim3(h.block), h.lds2, h.stream,
                     a, p, k);
  LHIP_TRY(hipGetLastError());
  return FVDBM_OK;
}

static int pair_any(Handle &h, int src) {
  return dispatch(h, [&](auto tp, auto qn) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return launch_pair<T, decltype(qn)::value>(h, src);
  });
}

// a captured run of `launches` two-step launches starting from `parity`
// (cached per (parity, launches)); *out is its executable graph
static int ensure_graph2(Handle &h, int parity, int launches, hipGraphExec_t *out) {
  const auto key = std::make_pair(parity, launches);
  auto it = h.graphs2.find(key);
  if (it != h.graphs2.end()) {
    *out = it->second.second;
    return FVDBM_OK;
  }
  LHIP_TRY(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
  int src = parity, rc = FVDBM_OK;
  for (int s = 0; s < launches && rc == FVDBM_OK; ++s) {
    rc = pair_any(h, src);
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

static int ensure_graph(Handle &h, int parity) {
  if (h.gexec[parity]) return FVDBM_OK;
  LHIP_TRY(hipStreamBeginCapture(h.stream, hipStreamCaptureModeThreadLocal));
  int src = parity, rc = FVDBM_OK;
  for (int s = 0; s < h.graph_steps && rc == FVDBM_OK; ++s) {
    rc = step_any(h, src);
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
  LHIP_TRY(hipGraphInstantiate(&h.gexec[parity], g, nullptr, nullptr, 0));
  return FVDBM_OK;
}

// n steps. Upwind handles with two-step tiles run two-step launches only
// (an odd count starts with one single step): captured runs of
// graph_steps / 2 launches, then one captured run for the rest. With dry,
// nothing is launched and only the graphs such a run replays are built.
static int run_steps(Handle &h, int64_t n, bool dry = false) {
  if (n <= 0) return FVDBM_OK;
  if (h.pairs && n >= 2) {
    int cur = h.cur;
    if (n & 1) {
      if (!dry) {
        int rc = step_any(h, cur);
        if (rc) return rc;
        ++h.steps;
      }
      cur ^= 1;
      --n;
    }
    int64_t launches = n / 2;
    const int per = h.graph_steps >= 2 ? h.graph_steps / 2 : 0;
    hipGraphExec_t x = nullptr;
    while (per > 0 && launches > 0) {
      const int nl = (int)std::min<int64_t>(launches, per);
      if (nl < 2) break;
      int rc = ensure_graph2(h, cur, nl, &x);
      if (rc) return rc;
      if (!dry) {
        LHIP_TRY(hipGraphLaunch(x, h.stream));
        h.steps += 2 * (int64_t)nl;
      }
      cur ^= nl & 1;
      launches -= nl;
    }
    for (; launches > 0; --launches) {
      if (!dry) {
        int rc = pair_any(h, cur);
        if (rc) return rc;
        h.steps += 2;
      }
      cur ^= 1;
    }
    if (dry) return FVDBM_OK;
    // f[cur ^ 1] holds the pdf two steps back (see prev_two_back)
    h.cur = cur;
    h.prev_two_back = true;
    h.mid_ready = false;
    h.has_prev = true;
    h.observed = false;
    return FVDBM_OK;
  }
  if (dry) return h.graph_steps > 0 && n >= h.graph_steps ? ensure_graph(h, h.cur) : FVDBM_OK;
  if (h.graph_steps > 0) {
    while (n >= h.graph_steps) {
      int rc = ensure_graph(h, h.cur);
      if (rc) return rc;
      LHIP_TRY(hipGraphLaunch(h.gexec[h.cur], h.stream));
      n -= h.graph_steps;  // graph_steps is even: parity unchanged
      h.steps += h.graph_steps;
    }
  }
  for (; n > 0; --n) {
    int rc = step_any(h, h.cur);
    if (rc) return rc;
    h.cur ^= 1;
    ++h.steps;
  }
  h.prev_two_back = false;
  h.has_prev = true;
  h.observed = false;
  return FVDBM_OK;
}

template <typename T, int QN>
static int init_uniform(Handle &h) {
  const Consts<T, QN> k = make_consts<T, QN>(h.tau, h.dt);
  for (int i = 0; i < 2; ++i)
    hipLaunchKernelGGL((k_lat_fill<T, QN>), grid1(h.Cp), dim3(256), 0, h.stream,
                       static_cast<T *>(h.f[i]), h.Cp, h.Cp, k);
  hipLaunchKernelGGL((k_lat_fill<T, QN>), grid1(h.N), dim3(256), 0, h.stream,
                     static_cast<T *>(h.node_pdf), h.N, h.N, k);
  LHIP_TRY(hipGetLastError());
  LHIP_TRY(hipStreamSynchronize(h.stream));
  h.cur = 0;
  h.steps = 0;
  h.has_prev = false;
  h.prev_two_back = false;
  h.observed = false;
  return FVDBM_OK;
}

template <typename T, int QN>
static int create_typed(const fvdbm_mesh_desc &m, Handle &h, int reorder) {
  HostLayout<T> L;
  std::string err;
  if (!fvdbm::build_layout<T>(m, reorder, h.block, L, err)) return fail(FVDBM_ERR_INVALID, err);
  h.C = L.C;
  h.Cp = L.Cp;
  h.F = L.F;
  h.N = L.N;
  h.K = L.K;
  h.nbf = L.nbf;
  h.nan = L.nan;
  h.nbc = L.n_bcells;
  h.bandwidth = L.bandwidth;
  h.last_cell_new = L.last_cell_new;
  h.new_to_old = L.new_to_old;
  h.an_gid_host = L.an_gid;
  const size_t es = sizeof(T);
  int rc = FVDBM_OK;
  for (int i = 0; i < 2 && !rc; ++i) rc = h.alloc(&h.f[i], (size_t)QN * h.Cp * es);
  if (!rc) rc = h.upload(&h.code, L.nb);
  if (!rc) rc = h.upload(&h.nx, L.nx);
  if (!rc) rc = h.upload(&h.ny, L.ny);
  if (!rc) rc = h.upload(&h.sL, L.sL);
  if (!rc && h.scheme == FVDBM_LAX_WENDROFF) rc = h.upload(&h.lwa, L.lwa);
  if (!rc && h.scheme == FVDBM_LAX_WENDROFF) rc = h.upload(&h.lwden2, L.lwden2);
  if (!rc) rc = h.upload(&h.bf_gnode, L.bf_gnode);
  if (!rc) rc = h.upload(&h.bf_ratio, L.bf_ratio);
  if (!rc) rc = h.alloc(&h.node_pdf, (size_t)QN * h.N * es);
  if (!rc) rc = h.alloc(&h.node_rho, (size_t)h.N * es);
  if (!rc) rc = h.alloc(&h.node_vel, (size_t)2 * h.N * es);
  if (!rc) rc = h.upload(&h.an_gid, L.an_gid);
  if (!rc) rc = h.upload(&h.an_type, L.an_type);
  if (!rc) rc = h.upload(&h.an_cells, L.an_cells);
  if (!rc) rc = h.upload(&h.an_w, L.an_w);
  if (!rc) rc = h.upload(&h.an_wsum, L.an_wsum);
  if (!rc) rc = h.upload(&h.an_rho, L.an_rho);
  if (!rc) rc = h.upload(&h.an_ux, L.an_ux);
  if (!rc) rc = h.upload(&h.an_uy, L.an_uy);
  if (!rc) rc = h.upload(&h.fc0, L.fc0);
  if (!rc) rc = h.upload(&h.fc1, L.fc1);
  if (!rc) rc = h.upload(&h.fn0, L.fn0);
  if (!rc) rc = h.upload(&h.fn1, L.fn1);
  if (!rc) rc = h.upload(&h.fd0, L.fd0);
  if (!rc) rc = h.upload(&h.fd1, L.fd1);
  if (!rc) rc = h.upload(&h.fnx, L.fnx);
  if (!rc) rc = h.upload(&h.fny, L.fny);
  if (!rc) rc = h.upload(&h.fL, L.fL);
  if (!rc) rc = h.upload(&h.tile_node_off, L.tile_node_off);
  if (!rc) rc = h.upload(&h.tile_nodes, L.tile_nodes);
  if (!rc) rc = h.upload(&h.bf_slot, L.bf_slot);
  h.max_tile_nodes = L.max_tile_nodes;
  if ((size_t)QN * h.max_tile_nodes * sizeof(T) > 65536)
    return fail(FVDBM_ERR_INVALID, "a tile's boundary-node table exceeds the LDS of a block");
  if (!rc) {
    // workgroup -> tile: workgroup b runs on XCD b % 8 and takes the
    // (b / 8)-th tile of that XCD's contiguous range, tiles with boundary
    // nodes first
    const int per = (L.ntiles + 7) / 8;
    std::vector<int32_t> bt(8 * (size_t)std::max(per, 1), -1);
    for (int x = 0; x < 8; ++x) {
      const int lo = std::min(x * per, L.ntiles), hi = std::min((x + 1) * per, L.ntiles);
      std::vector<int32_t> order;
      for (int t = lo; t < hi; ++t)
        if (L.tile_node_off[t + 1] > L.tile_node_off[t]) order.push_back(t);
      for (int t = lo; t < hi; ++t)
        if (L.tile_node_off[t + 1] == L.tile_node_off[t]) order.push_back(t);
      for (size_t i = 0; i < order.size(); ++i) bt[i * 8 + x] = order[i];
    }
    h.grid = (int)bt.size();
    rc = h.upload(&h.block_tile, bt);
  }
  if (rc) return rc;
  // two-step tiles: upwind, rings within a tile's threads, LDS of a block.
  // Ghost lists: per tile the ghost faces of the own cells (device order,
  // faces 0..2: ascending boundary-face index), then those of the S1 \ O
  // cells (ring order); ghost j of tile t is staged at slot block + |ring
  // of t| + j, and the two-step face codes name it there.
  const int bs = h.block;
  const size_t Cp = (size_t)L.Cp;
  const int nrgc = (int)(L.rg_code.size() / 3);
  std::vector<int32_t> gh_off(L.ntiles + 1, 0), gh_nown(L.ntiles, 0), gh_s0, gh_s1, gh_self;
  std::vector<T> gh_ratio;
  std::vector<int32_t> code2(L.nb_slot), rgc(L.rg_code);
  int max_stage = 0;
  for (int t = 0, r0 = 0; t < L.ntiles; ++t) {
    const int nr = L.ring_off[t + 1] - L.ring_off[t], n1 = L.ring_n1[t];
    gh_off[t] = (int32_t)gh_s0.size();
    int j = 0;
    for (int i = t * bs; i < t * bs + L.tile_count[t]; ++i)
      for (int kf = 0; kf < 3; ++kf) {
        const size_t at = (size_t)kf * Cp + i;
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
    for (int r = 0; r < n1; ++r, ++r0)
      for (int kf = 0; kf < 3; ++kf) {
        const size_t si = 3 * (size_t)r0 + kf;
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
    max_stage = std::max(max_stage, nr + j);
  }
  gh_off[L.ntiles] = (int32_t)gh_s0.size();
  const size_t lds2 = ((size_t)QN * L.max_xnodes + (size_t)PairRec<QN>::W * L.max_xpairs +
                       (size_t)QN * (bs + max_stage)) * es;
  if (h.two_step && h.scheme == FVDBM_UPWIND && L.pair_ok && lds2 <= 65536) {
    std::vector<int32_t> rcode(3 * (size_t)std::max(nrgc, 1), 0);
    std::vector<T> rnx(rcode.size(), T(0)), rny(rcode.size(), T(0)), rsl(rcode.size(), T(0));
    for (int r = 0; r < nrgc; ++r)
      for (int kf = 0; kf < 3; ++kf) {
        const size_t d = (size_t)kf * nrgc + r, src = 3 * (size_t)r + kf;
        rcode[d] = rgc[src];
        rnx[d] = L.rg_nx[src];
        rny[d] = L.rg_ny[src];
        rsl[d] = L.rg_sL[src];
      }
    std::vector<int32_t> rg_off(L.ntiles, 0);
    for (int t = 1; t < L.ntiles; ++t) rg_off[t] = rg_off[t - 1] + L.ring_n1[t - 1];
    rc = h.upload(&h.code2, code2);
    if (!rc) rc = h.upload(&h.ring_off, L.ring_off);
    if (!rc) rc = h.upload(&h.ring_cells, L.ring_cells);
    if (!rc) rc = h.upload(&h.ring_n1, L.ring_n1);
    if (!rc) rc = h.upload(&h.rg_off, rg_off);
    if (!rc) rc = h.upload(&h.rg_code, rcode);
    if (!rc) rc = h.upload(&h.rg_nx, rnx);
    if (!rc) rc = h.upload(&h.rg_ny, rny);
    if (!rc) rc = h.upload(&h.rg_sL, rsl);
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
    if (!rc) rc = h.upload(&h.entry_pair_off, L.entry_pair_off);
    if (!rc) rc = h.upload(&h.tile_pair_off, L.tile_pair_off);
    if (!rc) rc = h.upload(&h.pair_cell, L.pair_cell);
    if (!rc) rc = h.upload(&h.pair_slot, L.pair_slot);
    if (!rc) rc = h.upload(&h.pair_w, L.pair_w);
    if (!rc) rc = h.alloc(&h.fmid, (size_t)QN * h.Cp * sizeof(T));
    if (rc) return rc;
    h.pairs = true;
    h.lds2 = lds2;
    h.nrgc = nrgc;
    h.max_xnodes = L.max_xnodes;
    h.max_xpairs = L.max_xpairs;
  }
  // node rho/vel hold the boundary-condition values (reference node order)
  std::vector<T> rho(h.N), vel(2 * (size_t)h.N);
  const T *mr = static_cast<const T *>(m.node_rho);
  const T *mv = static_cast<const T *>(m.node_vel);
  for (int i = 0; i < h.N; ++i) {
    rho[i] = mr ? mr[i] : T(0);
    vel[i] = mv ? mv[2 * (size_t)i] : T(0);
    vel[(size_t)h.N + i] = mv ? mv[2 * (size_t)i + 1] : T(0);
  }
  LHIP_TRY(hipMemcpy(h.node_rho, rho.data(), rho.size() * es, hipMemcpyHostToDevice));
  LHIP_TRY(hipMemcpy(h.node_vel, vel.data(), vel.size() * es, hipMemcpyHostToDevice));
  return init_uniform<T, QN>(h);
}

// After a run that ended in a two-step launch: the pdf at the start of the
// last step, one step from f[cur ^ 1] (bitwise the two-step launch's
// intermediate state), into fmid. Done before anything the step reads
// (boundary-node values) changes, and at most once per run.
template <typename T, int QN>
static int materialize_mid(Handle &h) {
  if (!h.has_prev || !h.prev_two_back || h.mid_ready) return FVDBM_OK;
  const int rc = launch_step<T, QN>(h, h.cur ^ 1, h.fmid);
  if (rc) return rc;
  h.mid_ready = true;
  return FVDBM_OK;
}

// derived fields of the pdf at the start of the last step
template <typename T, int QN>
static int observe(Handle &h) {
  if (h.observed || !h.has_prev) return FVDBM_OK;
  int rc = FVDBM_OK;
  if (!h.obs_rho) rc = h.alloc(&h.obs_rho, (size_t)h.Cp * sizeof(T));
  if (!rc && !h.obs_vel) rc = h.alloc(&h.obs_vel, (size_t)2 * h.Cp * sizeof(T));
  if (!rc && !h.obs_feq) rc = h.alloc(&h.obs_feq, (size_t)QN * h.Cp * sizeof(T));
  if (!rc && !h.obs_flux) rc = h.alloc(&h.obs_flux, (size_t)QN * h.F * sizeof(T));
  if (rc) return rc;
  rc = materialize_mid<T, QN>(h);
  if (rc) return rc;
  Args<T> a = make_args<T>(h, h.cur ^ 1);
  if (h.prev_two_back) a.fsrc = static_cast<const T *>(h.fmid);
  const Consts<T, QN> k = make_consts<T, QN>(h.tau, h.dt);
  if (h.nan > 0)  // node values at the start of the last step; 16 nodes per block
    hipLaunchKernelGGL((k_lat_nodes<T, QN>), dim3((unsigned)((h.nan + 15) / 16)), dim3(256), 0,
                       h.stream, a, k);
  hipLaunchKernelGGL((k_lat_moments<T, QN>), grid1(h.C), dim3(256), 0, h.stream, a, k,
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
    hipLaunchKernelGGL((k_lat_flux<T, QN, 0>), grid1(h.F), dim3(256), 0, h.stream, a, k, fa,
                       static_cast<T *>(h.obs_flux));
  else
    hipLaunchKernelGGL((k_lat_flux<T, QN, 1>), grid1(h.F), dim3(256), 0, h.stream, a, k, fa,
                       static_cast<T *>(h.obs_flux));
  LHIP_TRY(hipGetLastError());
  LHIP_TRY(hipStreamSynchronize(h.stream));
  h.observed = true;
  return FVDBM_OK;
}

template <typename T, int QN>
static int get_typed(Handle &h, int field, void *dst) {
  T *out = static_cast<T *>(dst);
  const size_t es = sizeof(T);
  LHIP_TRY(hipStreamSynchronize(h.stream));
  auto cells = [&](const void *dev, int comps) -> int {
    std::vector<T> tmp((size_t)comps * h.Cp);
    LHIP_TRY(hipMemcpy(tmp.data(), dev, tmp.size() * es, hipMemcpyDeviceToHost));
    for (int i = 0; i < h.Cp; ++i) {
      const int o = h.new_to_old[i];
      if (o < 0) continue;
      for (int q = 0; q < comps; ++q) out[(size_t)o * comps + q] = tmp[(size_t)q * h.Cp + i];
    }
    return FVDBM_OK;
  };
  auto nodes = [&](const void *dev, int comps) -> int {
    std::vector<T> tmp((size_t)comps * h.N);
    LHIP_TRY(hipMemcpy(tmp.data(), dev, tmp.size() * es, hipMemcpyDeviceToHost));
    for (int i = 0; i < h.N; ++i)
      for (int q = 0; q < comps; ++q) out[(size_t)i * comps + q] = tmp[(size_t)q * h.N + i];
    return FVDBM_OK;
  };
  switch (field) {
    case FVDBM_CELL_PDF:
      return cells(h.f[h.cur], QN);
    case FVDBM_CELL_RHO:
    case FVDBM_CELL_VEL:
    case FVDBM_CELL_PDF_EQ: {
      if (!h.has_prev) {  // Cells.__init__: zeros until the first step
        const size_t n = (size_t)h.C * (field == FVDBM_CELL_RHO ? 1 : field == FVDBM_CELL_VEL ? 2 : QN);
        std::fill(out, out + n, T(0));
        return FVDBM_OK;
      }
      int rc = observe<T, QN>(h);
      if (rc) return rc;
      if (field == FVDBM_CELL_RHO) return cells(h.obs_rho, 1);
      if (field == FVDBM_CELL_VEL) return cells(h.obs_vel, 2);
      return cells(h.obs_feq, QN);
    }
    case FVDBM_FACE_FLUX: {
      if (!h.has_prev) {  // Faces.__init__: W rows
        const Consts<T, QN> k = make_consts<T, QN>(h.tau, h.dt);
        for (int j = 0; j < h.F; ++j)
          for (int q = 0; q < QN; ++q) out[(size_t)j * QN + q] = k.W[q];
        return FVDBM_OK;
      }
      int rc = observe<T, QN>(h);
      if (rc) return rc;
      std::vector<T> tmp((size_t)QN * h.F);
      LHIP_TRY(hipMemcpy(tmp.data(), h.obs_flux, tmp.size() * es, hipMemcpyDeviceToHost));
      for (int j = 0; j < h.F; ++j)
        for (int q = 0; q < QN; ++q) out[(size_t)j * QN + q] = tmp[(size_t)q * h.F + j];
      return FVDBM_OK;
    }
    case FVDBM_NODE_PDF:
    case FVDBM_NODE_RHO:
    case FVDBM_NODE_VEL: {
      int rc = observe<T, QN>(h);
      if (rc) return rc;
      if (field == FVDBM_NODE_PDF) return nodes(h.node_pdf, QN);
      if (field == FVDBM_NODE_VEL) return nodes(h.node_vel, 2);
      LHIP_TRY(hipMemcpy(out, h.node_rho, (size_t)h.N * es, hipMemcpyDeviceToHost));
      return FVDBM_OK;
    }
    default:
      return fail(FVDBM_ERR_INVALID, "unknown field");
  }
}

template <typename T, int QN>
static int set_typed(Handle &h, int field, const void *src) {
  const T *in = static_cast<const T *>(src);
  const size_t es = sizeof(T);
  if (field == FVDBM_NODE_PDF || field == FVDBM_NODE_VEL || field == FVDBM_NODE_RHO) {
    // the start-of-step state is computed with the values the step used
    const int rc = materialize_mid<T, QN>(h);
    if (rc) return rc;
  }
  LHIP_TRY(hipStreamSynchronize(h.stream));
  if (field == FVDBM_CELL_PDF) {
    // the current buffer only: the other keeps the start of the last step
    std::vector<T> tmp((size_t)QN * h.Cp, T(0));
    for (int i = 0; i < h.Cp; ++i) {
      const int o = h.new_to_old[i];
      if (o < 0) continue;
      for (int q = 0; q < QN; ++q) tmp[(size_t)q * h.Cp + i] = in[(size_t)o * QN + q];
    }
    LHIP_TRY(hipMemcpy(h.f[h.cur], tmp.data(), tmp.size() * es, hipMemcpyHostToDevice));
    return FVDBM_OK;
  }
  if (field == FVDBM_NODE_PDF || field == FVDBM_NODE_VEL) {
    const int comps = field == FVDBM_NODE_PDF ? QN : 2;
    std::vector<T> tmp((size_t)comps * h.N);
    for (int i = 0; i < h.N; ++i)
      for (int q = 0; q < comps; ++q) tmp[(size_t)q * h.N + i] = in[(size_t)i * comps + q];
    LHIP_TRY(hipMemcpy(field == FVDBM_NODE_PDF ? h.node_pdf : h.node_vel, tmp.data(),
                       tmp.size() * es, hipMemcpyHostToDevice));
    if (field == FVDBM_NODE_VEL && h.nan) {  // velocity-node values feed the step
      std::vector<T> ux(h.nan), uy(h.nan);
      for (int a = 0; a < h.nan; ++a) {
        ux[a] = in[2 * (size_t)h.an_gid_host[a]];
        uy[a] = in[2 * (size_t)h.an_gid_host[a] + 1];
      }
      LHIP_TRY(hipMemcpy(h.an_ux, ux.data(), ux.size() * es, hipMemcpyHostToDevice));
      LHIP_TRY(hipMemcpy(h.an_uy, uy.data(), uy.size() * es, hipMemcpyHostToDevice));
    }
    h.observed = false;
    return FVDBM_OK;
  }
  if (field == FVDBM_NODE_RHO) {
    LHIP_TRY(hipMemcpy(h.node_rho, in, (size_t)h.N * es, hipMemcpyHostToDevice));
    if (h.nan) {  // density-node values feed the step
      std::vector<T> r(h.nan);
      for (int a = 0; a < h.nan; ++a) r[a] = in[h.an_gid_host[a]];
      LHIP_TRY(hipMemcpy(h.an_rho, r.data(), r.size() * es, hipMemcpyHostToDevice));
    }
    h.observed = false;
    return FVDBM_OK;
  }
  return fail(FVDBM_ERR_INVALID, "field cannot be written (derived from the pdf)");
}

}  // namespace fvdbm_lattice

using fvdbm_lattice::Handle;

extern "C" {

int fvdbm_lattice_create(const fvdbm_mesh_desc *mesh, const fvdbm_params *p, int32_t quivers,
                         void **out) {
  using namespace fvdbm_lattice;
  if (!mesh || !p || !out) return fail(FVDBM_ERR_INVALID, "null argument");
  *out = nullptr;
  if (quivers != 9 && quivers != 13) return fail(FVDBM_ERR_INVALID, "quivers must be 9 or 13");
  if (p->precision != FVDBM_FP32 && p->precision != FVDBM_FP64)
    return fail(FVDBM_ERR_INVALID, "precision must be FVDBM_FP32 or FVDBM_FP64");
  if (p->scheme != FVDBM_UPWIND && p->scheme != FVDBM_LAX_WENDROFF)
    return fail(FVDBM_ERR_INVALID, "Unknown flux scheme");
  if (!(p->tau > 0.0) || !std::isfinite(p->delta_t))
    return fail(FVDBM_ERR_INVALID, "tau must be > 0 and delta_t finite");
  const int block = p->block ? p->block : 256;
  if (block < 64 || block > 1024 || block % 64)
    return fail(FVDBM_ERR_INVALID, "block must be a multiple of 64 in [64, 1024]");
  int gs = p->graph_steps == 0 ? 32 : p->graph_steps;
  if (gs > 0 && (gs % 2)) ++gs;  // even, so a replay preserves buffer parity
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(FVDBM_ERR_HIP, "no HIP device available");
  std::unique_ptr<Handle> h(new Handle());
  h->q = quivers;
  h->precision = p->precision;
  h->scheme = p->scheme;
  h->block = block;
  h->graph_steps = gs;
  h->two_step = (p->variant & 8) == 0;
  h->tau = p->tau;
  h->dt = p->delta_t;
  LHIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  LHIP_TRY(hipEventCreate(&h->ev0));
  LHIP_TRY(hipEventCreate(&h->ev1));
  const int reorder = p->reorder;
  int rc = dispatch(*h, [&](auto tp, auto qn) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return create_typed<T, decltype(qn)::value>(*mesh, *h, reorder);
  });
  if (rc) return rc;
  // default launch mode as for the tuned handle: plain launches from 128k
  // padded cells, where a launch outlasts the host's queueing of the next
  if (p->graph_steps == 0 && h->Cp >= 131072) h->graph_steps = -1;
  *out = h.release();
  return FVDBM_OK;
}

int fvdbm_lattice_destroy(void *handle) {
  delete static_cast<Handle *>(handle);
  return FVDBM_OK;
}

int fvdbm_lattice_init_uniform(void *handle) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  return dispatch(*h, [&](auto tp, auto qn) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return init_uniform<T, decltype(qn)::value>(*h);
  });
}

int fvdbm_lattice_step(void *handle, int64_t nsteps) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  if (nsteps < 0) return fail(FVDBM_ERR_INVALID, "nsteps < 0");
  return run_steps(*h, nsteps);
}

int fvdbm_lattice_prepare(void *handle, int64_t nsteps) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  if (nsteps < 0) return fail(FVDBM_ERR_INVALID, "nsteps < 0");
  const int cur = h->cur;  // from either buffer parity
  int rc = FVDBM_OK;
  for (int par = 0; par < 2 && rc == FVDBM_OK; ++par) {
    h->cur = par;
    rc = run_steps(*h, nsteps, true);
  }
  h->cur = cur;
  return rc;
}

int fvdbm_lattice_step_timed(void *handle, int64_t nsteps, double *ms) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !ms) return fail(FVDBM_ERR_INVALID, "null argument");
  if (nsteps < 0) return fail(FVDBM_ERR_INVALID, "nsteps < 0");
  {
    int rc = run_steps(*h, nsteps, true);  // the graphs this run replays, outside the timed region
    if (rc) return rc;
  }
  LHIP_TRY(hipEventRecord(h->ev0, h->stream));
  int rc = run_steps(*h, nsteps);
  if (rc) return rc;
  LHIP_TRY(hipEventRecord(h->ev1, h->stream));
  LHIP_TRY(hipEventSynchronize(h->ev1));
  float t = 0.f;
  LHIP_TRY(hipEventElapsedTime(&t, h->ev0, h->ev1));
  *ms = t;
  return FVDBM_OK;
}

int fvdbm_lattice_synchronize(void *handle) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h) return fail(FVDBM_ERR_INVALID, "null handle");
  LHIP_TRY(hipStreamSynchronize(h->stream));
  return FVDBM_OK;
}

int fvdbm_lattice_get_field(void *handle, int32_t field, void *dst) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !dst) return fail(FVDBM_ERR_INVALID, "null argument");
  return dispatch(*h, [&](auto tp, auto qn) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return get_typed<T, decltype(qn)::value>(*h, field, dst);
  });
}

int fvdbm_lattice_set_field(void *handle, int32_t field, const void *src) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !src) return fail(FVDBM_ERR_INVALID, "null argument");
  return dispatch(*h, [&](auto tp, auto qn) {
    using T = std::remove_pointer_t<decltype(tp)>;
    return set_typed<T, decltype(qn)::value>(*h, field, src);
  });
}

int fvdbm_lattice_info_get(void *handle, fvdbm_info *o) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !o) return fail(FVDBM_ERR_INVALID, "null argument");
  std::memset(o, 0, sizeof(*o));
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
  o->variant = 4 | (h->pairs ? 0 : 8);  // no one-step staging; two-step tiles or not
  o->block = h->block;
  o->graph_steps = h->graph_steps;
  o->steps_done = h->steps;
  o->device_bytes = h->device_bytes;
  o->n_cases = 1;
  o->launch_classes = h->pairs ? 1 : 0;
  return FVDBM_OK;
}

int fvdbm_lattice_get_permutation(void *handle, int32_t *new_to_old) {
  using namespace fvdbm_lattice;
  Handle *h = static_cast<Handle *>(handle);
  if (!h || !new_to_old) return fail(FVDBM_ERR_INVALID, "null argument");
  for (int i = 0; i < h->C; ++i) new_to_old[i] = h->new_to_old[i];
  return FVDBM_OK;
}

}  // extern "C"
