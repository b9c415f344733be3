/*
 * fvdbm_oracle.c — CPU restatement of the reference FVDBM step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it. The product step runs in the HIP library and fails loudly when
 * that library is missing.
 *
 * What it restates (reference = MADALERAU/FVDBM-JAX):
 *   Environment.step            src/environment.py:55-65 (phase order)
 *   Cells.calc_macros/calc_eqs  src/containers.py:93-105
 *   Nodes.calc_pdfs             src/containers.py:339-408
 *   Faces.calc_upwind_flux      src/containers.py:215-240, calc_ghost :280-287
 *   Cells.calc_pdfs/calc_pdf    src/containers.py:107-121
 *   D2Q9 constants/calc_eq      src/dynamics.py:35-74
 *   extrapolate/weighted_avg    utils/utils.py:34-60, extrap_pdf :153-154
 *
 * Arithmetic conventions (the reference runs JAX without x64, i.e. FP32):
 *   - every jnp.sum / jnp.dot is a 0-initialised sum in index order;
 *   - no fused multiply-add (built with -ffp-contract=off);
 *   - divisions are divisions (never reciprocal multiplies);
 *   - C = 1/sqrt(3) in the working precision, C^2 = C*C,
 *     C^4 = (C*C)*(C*C) (lax.integer_pow by squaring);
 *   - index -1 in a gather wraps to the last row (JAX index normalisation).
 * XLA's exact reduction order cannot be observed here (jax is absent), so the
 * FP32 agreement with the JAX path is argued from these conventions; the
 * oracle itself is pinned by the invariant, closed-form, symmetry and planted-
 * mistake tests under tests/.
 *
 * Layouts are the reference container layouts, row-major:
 *   cell pdf [C][Q], face flux [F][Q], node pdf [N][Q], vel [.][2].
 *
 * Two lattices: D2Q9 (oracle_step_f32/_f64, Q = 9) and D2Q13
 * (oracle_step_q13_f32/_f64; src/dynamics.py:77-102: the D2Q9 velocities
 * plus (+-2,0), (0,+-2), W = 3/8, 1/12, 1/16, 1/96, C = 1/sqrt(2) and the
 * cubic equilibrium terms). The same step body is compiled for each.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the element loops (1 without OpenMP); n < 1 leaves it. */
int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n >= 1) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

#define Q 9

/* src/dynamics.py:54-62 */
static const int KSI[Q][2] = {{0, 0},  {1, 0},  {0, 1},   {-1, 0}, {0, -1},
                              {1, 1},  {-1, 1}, {-1, -1}, {1, -1}};

/* Planted mistakes, used only by tests that show the oracle checks catch a
 * plausible error. 0 = faithful. */
enum {
  MUT_NONE = 0,
  MUT_DOWNWIND = 1,         /* upwind choice reversed */
  MUT_NO_GHOST = 2,         /* ghost replaced by the known cell */
  MUT_NO_SIGN = 3,          /* face sign ignored in the cell sum */
  MUT_TRANSPOSE_KSI = 4,    /* ksi_x and ksi_y swapped */
  MUT_DROP_QUADRATIC = 5,   /* (ksi.u)^2 term of feq dropped */
  MUT_IDW_UNIFORM = 6,      /* IDW weights replaced by 1 */
  MUT_IDW_UNNORMALISED = 7, /* IDW sum not divided by sum of weights */
  MUT_NO_COLLISION = 8,     /* BGK term dropped */
  MUT_LW_D1_FOR_D0 = 9,     /* Lax-Wendroff: d1 / (d0 + d1) in place of d0 / (d0 + d1) */
  MUT_LW_DT_SIGN = 10,      /* Lax-Wendroff: + (phi dt) / (2 (d0 + d1)) */
  MUT_LW_SWAP_UD = 11,      /* Lax-Wendroff: upwind and downwind slots exchanged */
};

typedef struct {
  int32_t n_cells, n_faces, n_nodes, k_max;
  const int32_t *cell_faces; /* [C][3]  cells.face_indices */
  const int32_t *cell_signs; /* [C][3]  cells.face_normals (+-1) */
  const int32_t *face_cells; /* [F][2]  faces.stencil_cells_index */
  const void *face_dists;    /* [F][2]  faces.stencil_dists */
  const int32_t *face_nodes; /* [F][2]  faces.nodes_index */
  const void *face_n;        /* [F][2]  faces.n */
  const void *face_L;        /* [F]     faces.L */
  const int32_t *node_cells; /* [N][K]  nodes.cells_index */
  const void *node_dists;    /* [N][K]  nodes.cell_dists */
  const int32_t *node_type;  /* [N]     nodes.type */
} oracle_mesh;

typedef struct {
  void *cell_pdf; /* [C][9] in/out */
  void *cell_rho; /* [C]    out: moments of the pdf at the start of the step */
  void *cell_vel; /* [C][2] out */
  void *cell_feq; /* [C][9] out */
  void *face_pdf; /* [F][9] out: the flux (faces.pdf) */
  void *node_pdf; /* [N][9] in/out */
  void *node_rho; /* [N]    in/out */
  void *node_vel; /* [N][2] in/out */
} oracle_state;

struct consts_f32 { float W[Q], C, C2, C4x2, C2x2; };
struct consts_f64 { double W[Q], C, C2, C4x2, C2x2; };
struct consts_q13_f32 { float W[13], C, C2, C4x2, C2x2, C6x2; };
struct consts_q13_f64 { double W[13], C, C2, C4x2, C2x2, C6x2; };

/* W = jnp.array([4/9, 1/9 x4, 1/36 x4]) rounds the Python doubles;
 * C = 1/jnp.sqrt(3) (src/dynamics.py:63-64). */
static void init_consts_f32(struct consts_f32 *k) {
  const double w[Q] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                       1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
  for (int q = 0; q < Q; ++q) k->W[q] = (float)w[q];
  k->C = 1.0f / sqrtf(3.0f);
  k->C2 = k->C * k->C;
  float c4 = k->C2 * k->C2;
  k->C4x2 = 2.0f * c4;
  k->C2x2 = 2.0f * k->C2;
}
static void init_consts_f64(struct consts_f64 *k) {
  const double w[Q] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                       1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
  for (int q = 0; q < Q; ++q) k->W[q] = w[q];
  k->C = 1.0 / sqrt(3.0);
  k->C2 = k->C * k->C;
  double c4 = k->C2 * k->C2;
  k->C4x2 = 2.0 * c4;
  k->C2x2 = 2.0 * k->C2;
}

#define REAL float
#define SFX f32
#include "oracle_step.inc"
#undef REAL
#undef SFX

#define REAL double
#define SFX f64
#include "oracle_step.inc"
#undef REAL
#undef SFX

/* D2Q13 (src/dynamics.py:77-102). W = jnp.array([3/8, 1/12 x4, 1/16 x4,
 * 1/96 x4]) rounds the Python doubles; C = 1/jnp.sqrt(2); powers of C as
 * lax.integer_pow computes them (by squaring): C^2 = C*C, C^4 = C^2*C^2,
 * C^6 = C^2*C^4. */
static const int KSI13[13][2] = {{0, 0},  {1, 0},  {0, 1},   {-1, 0}, {0, -1},
                                 {1, 1},  {-1, 1}, {-1, -1}, {1, -1}, {2, 0},
                                 {0, 2},  {-2, 0}, {0, -2}};
static const double W13[13] = {3.0 / 8,  1.0 / 12, 1.0 / 12, 1.0 / 12, 1.0 / 12,
                               1.0 / 16, 1.0 / 16, 1.0 / 16, 1.0 / 16, 1.0 / 96,
                               1.0 / 96, 1.0 / 96, 1.0 / 96};
static void init_consts_q13_f32(struct consts_q13_f32 *k) {
  for (int q = 0; q < 13; ++q) k->W[q] = (float)W13[q];
  k->C = 1.0f / sqrtf(2.0f);
  k->C2 = k->C * k->C;
  float c4 = k->C2 * k->C2;
  float c6 = k->C2 * c4;
  k->C4x2 = 2.0f * c4;
  k->C2x2 = 2.0f * k->C2;
  k->C6x2 = 2.0f * c6;
}
static void init_consts_q13_f64(struct consts_q13_f64 *k) {
  for (int q = 0; q < 13; ++q) k->W[q] = W13[q];
  k->C = 1.0 / sqrt(2.0);
  k->C2 = k->C * k->C;
  double c4 = k->C2 * k->C2;
  double c6 = k->C2 * c4;
  k->C4x2 = 2.0 * c4;
  k->C2x2 = 2.0 * k->C2;
  k->C6x2 = 2.0 * c6;
}

#undef Q
#define Q 13
#define KSI KSI13
#define REAL float
#define SFX q13_f32
#include "oracle_step.inc"
#undef REAL
#undef SFX
#define REAL double
#define SFX q13_f64
#include "oracle_step.inc"
#undef REAL
#undef SFX

int oracle_calc_eq_q13_f32(int n, const float *rho, const float *vel, float *out) {
  struct consts_q13_f32 k;
  init_consts_q13_f32(&k);
  for (int i = 0; i < n; ++i)
    calc_eq_q13_f32(&k, rho[i], vel[2 * i], vel[2 * i + 1], out + (size_t)i * 13, 0);
  return 0;
}
int oracle_calc_eq_q13_f64(int n, const double *rho, const double *vel, double *out) {
  struct consts_q13_f64 k;
  init_consts_q13_f64(&k);
  for (int i = 0; i < n; ++i)
    calc_eq_q13_f64(&k, rho[i], vel[2 * i], vel[2 * i + 1], out + (size_t)i * 13, 0);
  return 0;
}
int oracle_calc_macro_q13_f32(int n, const float *f, float *rho, float *vel) {
  for (int i = 0; i < n; ++i)
    calc_macro_q13_f32(f + (size_t)i * 13, rho + i, vel + 2 * i, vel + 2 * i + 1, 0);
  return 0;
}
#undef KSI
#undef Q
#define Q 9

/* Equilibrium of (rho, u) for a batch of points, used by tests for the
 * closed-form moment checks and for warm-start initialisation checks. */
int oracle_calc_eq_f32(int n, const float *rho, const float *vel, float *out) {
  struct consts_f32 k;
  init_consts_f32(&k);
  for (int i = 0; i < n; ++i)
    calc_eq_f32(&k, rho[i], vel[2 * i], vel[2 * i + 1], out + (size_t)i * Q, 0);
  return 0;
}
int oracle_calc_eq_f64(int n, const double *rho, const double *vel, double *out) {
  struct consts_f64 k;
  init_consts_f64(&k);
  for (int i = 0; i < n; ++i)
    calc_eq_f64(&k, rho[i], vel[2 * i], vel[2 * i + 1], out + (size_t)i * Q, 0);
  return 0;
}
int oracle_calc_macro_f32(int n, const float *f, float *rho, float *vel) {
  for (int i = 0; i < n; ++i)
    calc_macro_f32(f + (size_t)i * Q, rho + i, vel + 2 * i, vel + 2 * i + 1, 0);
  return 0;
}
