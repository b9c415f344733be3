/*
 * fvdbm.h — C ABI of the MI355X FVDBM solver library (libfvdbm.so).
 *
 * The reference (MADALERAU/FVDBM-JAX) has no FFI: its boundary is the Python
 * object API. Each entry point below replaces one piece of that API and is
 * bound from Python with ctypes (fvdbm-jax_amd/_lib.py):
 *
 *   fvdbm_create       Environment(cells, faces, nodes) + Environment.init()
 *                      (src/environment.py:15-18, 38-42; Cells/Faces/Nodes.init
 *                      src/containers.py:64-69, 157-163, 309-315) — uploads the
 *                      static arrays Mesher.to_env produced (src/mesher.py:610-693)
 *   fvdbm_init_uniform Container.__init__ pdf = calc_eq(1, 0) (src/containers.py:29)
 *   fvdbm_set_field    assignment to a dynamic container attribute (pytree child)
 *   fvdbm_get_field    reading cells.pdf / rho / vel / pdf_eq, faces.pdf,
 *                      nodes.pdf / rho / vel (src/containers.py:53-62,146-155,299-307)
 *   fvdbm_step         env = env.step()  (src/environment.py:55-65), n times
 *   fvdbm_destroy      end of the Environment's lifetime
 *
 * Conventions: every entry point returns 0 on success and a non-zero
 * FVDBM_ERR_* code otherwise; fvdbm_last_error() gives a message for the
 * calling thread. A handle owns its device buffers and one HIP stream; it is
 * not thread-safe, separate handles are independent. Arrays are host
 * pointers in the reference's row-major layout and in the handle's precision
 * (float for FVDBM_FP32, double for FVDBM_FP64) unless stated otherwise.
 */
#ifndef FVDBM_H
#define FVDBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVDBM_ABI_VERSION 1

enum {
  FVDBM_OK = 0,
  FVDBM_ERR_INVALID = 1, /* bad argument or mesh (ValueError in Python) */
  FVDBM_ERR_HIP = 2,     /* HIP runtime error */
  FVDBM_ERR_ALLOC = 3,   /* device or host allocation failed */
  FVDBM_ERR_STATE = 4,   /* call not valid in the handle's state */
};

enum { FVDBM_FP32 = 0, FVDBM_FP64 = 1 };
enum { FVDBM_UPWIND = 0, FVDBM_LAX_WENDROFF = 1 };

/* Observable fields, each in reference order and layout. */
enum {
  FVDBM_CELL_PDF = 0,    /* [C][9]  cells.pdf (current) */
  FVDBM_CELL_RHO = 1,    /* [C]     cells.rho    } moments of the pdf at the */
  FVDBM_CELL_VEL = 2,    /* [C][2]  cells.vel    } start of the last step, as */
  FVDBM_CELL_PDF_EQ = 3, /* [C][9]  cells.pdf_eq } the reference leaves them */
  FVDBM_FACE_FLUX = 4,   /* [F][9]  faces.pdf: flux of the last step */
  FVDBM_NODE_PDF = 5,    /* [N][9]  nodes.pdf */
  FVDBM_NODE_RHO = 6,    /* [N]     nodes.rho */
  FVDBM_NODE_VEL = 7,    /* [N][2]  nodes.vel */
};

/* Static mesh, exactly the arrays Mesher.to_env hands to the containers
 * (src/mesher.py:632-691) plus the boundary-condition arrays that
 * set_vel_node / set_rho_node fill (src/mesher.py:697-742). Float arrays are
 * in the handle's precision. */
typedef struct fvdbm_mesh_desc {
  int32_t n_cells, n_faces, n_nodes, k_max;
  const int32_t *cell_faces; /* [C][3] cells.face_indices */
  const int32_t *cell_signs; /* [C][3] cells.face_normals (+-1) */
  const int32_t *face_cells; /* [F][2] faces.stencil_cells_index (-1 = ghost) */
  const void *face_dists;    /* [F][2] faces.stencil_dists */
  const int32_t *face_nodes; /* [F][2] faces.nodes_index */
  const void *face_n;        /* [F][2] faces.n */
  const void *face_L;        /* [F]    faces.L */
  const int32_t *node_cells; /* [N][K] nodes.cells_index (-1 padded) */
  const void *node_dists;    /* [N][K] nodes.cell_dists (-1 padded) */
  const int32_t *node_type;  /* [N]    nodes.type: 0 none, 1 velocity, 2 density */
  const void *node_rho;      /* [N]    nodes.rho  (density-node values) */
  const void *node_vel;      /* [N][2] nodes.vel  (velocity-node values) */
  /* Optional (may be NULL): cell centroids [C][2] (Mesher.cell_centers),
   * used only to order cells along a Hilbert curve (reorder = 3). */
  const double *cell_xy;
} fvdbm_mesh_desc;

/* D2Q9(tau, delta_t) (src/dynamics.py:65-68) and layout/launch options. */
typedef struct fvdbm_params {
  double tau;
  double delta_t;
  int32_t precision;  /* FVDBM_FP32 | FVDBM_FP64 */
  int32_t scheme;     /* FVDBM_UPWIND | FVDBM_LAX_WENDROFF (Faces.flux_scheme) */
  int32_t reorder;    /* device cell order: 0 reference order, 1 reverse Cuthill-McKee
                         with ghost-face cells first, 2 RCM, 3 Hilbert curve of the
                         centroids with ghost-face cells first, 4 Hilbert curve
                         (3 and 4 need cell_xy) */
  int32_t variant;    /* step kernel variant bits, 0 = default: 1 no XCD-aware tile order,
                         2 boundary nodes in their own kernel, 4 no LDS staging,
                         8 one step per launch (by default single upwind handles
                         advance two steps per launch where the tiles fit; the
                         handle's info reports bit 8 when they do not) */
  int32_t block;      /* threads per block (64..1024, multiple of 64), 0 = default */
  int32_t graph_steps;/* steps per captured hipGraph, <0 = no graphs, 0 = default: 32 below
                         131072 padded cells, plain launches from there (single handles;
                         fvdbm_info reports the mode chosen) */
} fvdbm_params;

typedef struct fvdbm_info {
  int32_t n_cells, n_faces, n_nodes, k_max;
  int32_t n_cells_padded;
  int32_t n_boundary_cells; /* cells with a ghost face (stored first) */
  int32_t n_boundary_faces;
  int32_t n_active_nodes;   /* nodes of type 1 or 2 */
  int32_t bandwidth;        /* max |i - j| over neighbouring cells, device order */
  int32_t precision, scheme, variant, block, graph_steps;
  int64_t steps_done;
  int64_t device_bytes;     /* device memory owned by the handle */
  int32_t n_cases;          /* 1, or the case count of a batched handle */
  int32_t launch_classes;   /* launches per two-step launch: 1, or 2 when tiles are split by
                               their LDS size; 0 without two-step tiles */
} fvdbm_info;

/* Batched solves (no counterpart in the reference, which runs one mesh per
 * Environment): the mesh descriptor holds several independent meshes
 * concatenated (no face joins two cases), cell_case gives each cell's case,
 * cases contiguous and in order. Each case has its own tau and delta_t, its
 * own step count and a convergence mask; one kernel launch steps all active
 * cases. */
typedef struct fvdbm_batch_desc {
  int32_t n_cases;
  const int32_t *cell_case; /* [C] case id of each cell, nondecreasing */
  const double *case_tau;   /* [n_cases] */
  const double *case_dt;    /* [n_cases] */
} fvdbm_batch_desc;

/* Per-case status values of fvdbm_batch_status. */
enum { FVDBM_RUNNING = 0, FVDBM_CONVERGED = 1, FVDBM_DIVERGED = 2, FVDBM_GROWING = 3,
       FVDBM_STOPPED = 4 };

const char *fvdbm_last_error(void);
int fvdbm_abi_version(void);
/* Number of HIP devices visible (0 without a GPU; never an error). */
int fvdbm_device_count(int32_t *count);
/* Number of step-kernel variants the autotuner may choose from. */
int fvdbm_variant_count(int32_t *count);

int fvdbm_create(const fvdbm_mesh_desc *mesh, const fvdbm_params *params,
                 void **handle);
int fvdbm_destroy(void *handle);
int fvdbm_info_get(void *handle, fvdbm_info *out);

/* Cell and node pdfs <- W; moments, pdf_eq zeroed; flux <- W
 * (the state right after Environment.init()). */
int fvdbm_init_uniform(void *handle);
/* Write a field (FVDBM_CELL_PDF, FVDBM_NODE_PDF, FVDBM_NODE_RHO,
 * FVDBM_NODE_VEL) from host memory in reference order. */
int fvdbm_set_field(void *handle, int32_t field, const void *src);
/* Read any FVDBM_* field into host memory in reference order. */
int fvdbm_get_field(void *handle, int32_t field, void *dst);

/* Advance nsteps steps (asynchronous on the handle's stream). */
int fvdbm_step(void *handle, int64_t nsteps);
/* Advance nsteps steps and return the residual of the last one,
 * R = ||u^{n+1} - u^n||_2 / ||u^{n+1}||_2 over all cells (synchronous). */
int fvdbm_step_residual(void *handle, int64_t nsteps, double *residual);
/* Capture the hipGraphs a later run of nsteps steps replays (from either
 * buffer parity), on the host: nothing runs and the state is unchanged, so
 * that run (fvdbm_step or fvdbm_step_timed) launches without capturing. */
int fvdbm_prepare(void *handle, int64_t nsteps);
/* Advance nsteps steps between two HIP events on the handle's stream and
 * return the elapsed milliseconds (synchronous). Work already queued on
 * the stream runs before the first event. */
int fvdbm_step_timed(void *handle, int64_t nsteps, double *ms);
int fvdbm_synchronize(void *handle);

/* ------------------------------------------------------------------------
 * Any-lattice handle: D2Q13 (src/dynamics.py:77-102; quivers = 13) and D2Q9
 * (quivers = 9, the same plain kernels, a cross-check of the tuned path).
 * Replaces Environment.init/step for an Environment built on D2Q13
 * (src/environment.py:38-65). Mesh descriptor and params as for
 * fvdbm_create (params.variant: bit 8 = one step per launch, other bits
 * ignored; by default upwind handles advance two steps per launch where the
 * tiles fit, and info reports variant 4, or 12 for one step per launch);
 * fields as for
 * fvdbm_get_field / fvdbm_set_field with Q = quivers pdf components.
 * Errors through fvdbm_last_error(). No batched handles, residual or
 * warm-start entry points for this handle.
 * ------------------------------------------------------------------------ */
int fvdbm_lattice_create(const fvdbm_mesh_desc *mesh, const fvdbm_params *params,
                         int32_t quivers, void **handle);
int fvdbm_lattice_destroy(void *handle);
int fvdbm_lattice_init_uniform(void *handle);
int fvdbm_lattice_step(void *handle, int64_t nsteps);
int fvdbm_lattice_prepare(void *handle, int64_t nsteps);
int fvdbm_lattice_step_timed(void *handle, int64_t nsteps, double *ms);
int fvdbm_lattice_synchronize(void *handle);
int fvdbm_lattice_get_field(void *handle, int32_t field, void *dst);
int fvdbm_lattice_set_field(void *handle, int32_t field, const void *src);
int fvdbm_lattice_info_get(void *handle, fvdbm_info *out);
int fvdbm_lattice_get_permutation(void *handle, int32_t *new_to_old);

/* Measurement aid (no reference counterpart): copies `bytes` between two
 * fresh device buffers `reps` times with a plain 16-byte-per-lane kernel
 * (k_copy_probe) and returns milliseconds per copy. Used to scale the
 * FETCH_SIZE / WRITE_SIZE counters and as the attainable-bandwidth check. */
int fvdbm_copy_probe(int64_t bytes, int32_t reps, double *ms_per_copy);

/* Batched handle (see fvdbm_batch_desc). Only FVDBM_CELL_PDF can be read
 * or written (each case's current buffer); every case starts at W. */
int fvdbm_create_batch(const fvdbm_mesh_desc *mesh, const fvdbm_params *params,
                       const fvdbm_batch_desc *batch, void **handle);
/* Run active cases in intervals of check_every steps (one captured graph
 * per interval) until none is active or max_steps have run; when
 * max_steps is not a multiple of check_every the last interval is the
 * shorter remainder, so no case runs more than max_steps. After each
 * interval's last step a case stops as CONVERGED when its residual
 * ||u^{n+1}-u^n||/||u^{n+1}|| is below tol, DIVERGED when it is not finite,
 * GROWING when growth > 0 and it exceeds growth times its smallest residual
 * so far. Returns the number of active cases and the steps run. */
int fvdbm_batch_run(void *handle, int64_t max_steps, int32_t check_every, double tol,
                    double growth, int32_t *n_active, int64_t *steps_run);
/* Per-case steps taken, last residual, status (FVDBM_RUNNING ...), active
 * flag; any pointer may be NULL. */
int fvdbm_batch_status(void *handle, int64_t *steps, double *residual, int32_t *status,
                       int32_t *active);
/* Set the active mask; a case switched on restarts its residual history
 * (and its step count when reset_steps). */
int fvdbm_batch_set_active(void *handle, const int32_t *active, int32_t reset_steps);

/* Interpolation and equilibrium initialisation (single and batched
 * handles; arrays in reference order, batches concatenated case by case,
 * element type per the handle's precision).
 *
 * fvdbm_node_fields: rho [N] and u [N][2] at every node, the reference's
 * inverse-distance weighting (utils.weighted_avg, utils/utils.py:34-45; used
 * there for boundary nodes only, Nodes.calc_density / calc_velocity,
 * src/containers.py:364-374) of the moments of the current cell pdf.
 * Either output may be NULL. */
int fvdbm_node_fields(void *handle, void *node_rho, void *node_vel);
/* rho [C] and u [C][2] of the current cell pdf (Dynamics.calc_macro,
 * src/dynamics.py:35-47); either output may be NULL. */
int fvdbm_cell_fields(void *handle, void *cell_rho, void *cell_vel);
/* Cell pdf <- feq(rho, u) (D2Q9.calc_eq, src/dynamics.py:70-74), the
 * equilibrium initialisation Cells.__init__ uses for rho = 1, u = 0
 * (src/containers.py:24-30) with given fields. */
int fvdbm_set_equilibrium(void *handle, const void *cell_rho, const void *cell_vel);
/* As fvdbm_set_equilibrium from node values: each cell takes the
 * inverse-distance weighting (weights 1/d, d the node's distance to the
 * cell centre as in the node lists) of its three vertices. */
int fvdbm_set_equilibrium_nodes(void *handle, const void *node_rho, const void *node_vel);

/* Device cell order: new_to_old[i] = reference cell id of device cell i. */
int fvdbm_get_permutation(void *handle, int32_t *new_to_old);

/* Host-only (no device needed): the cell order fvdbm_create would use for
 * this mesh with `reorder` (see fvdbm_params), and stats = {boundary cells,
 * boundary faces, active nodes, bandwidth}. Returns FVDBM_ERR_INVALID for an
 * inconsistent mesh. */
int fvdbm_layout_order(const fvdbm_mesh_desc *mesh, int32_t precision, int32_t reorder,
                       int32_t *new_to_old, int32_t *stats);

/* ------------------------------------------------------------------------
 * Warm-start surrogate kernels (no reference counterpart: the reference has
 * no learned component). Device pointers to row-major float32 arrays; the
 * stream is a hipStream_t (NULL = the default stream). Dense layers use
 * bf16 MFMA operands with float32 accumulation; weights are float32 master
 * copies rounded to bf16 on load. Hidden width 64. The row arrays of the
 * dense kernels (x, y, dy, dx, dz, h; a and b of the weight gradient) must
 * be 16-byte aligned: a misaligned one fails with FVDBM_ERR_INVALID.
 * ------------------------------------------------------------------------ */
const char *fvdbm_nn_last_error(void);
/* Rows per partial product of fvdbm_nn_wgrad / fvdbm_nn_colsum: their
 * `partial` workspace holds ceil(m / rows) * (output size) floats. */
int fvdbm_nn_chunk_rows(void);
/* y [m][n_out] = silu(x [m][k] w1 [k][64] + b1) w2 [64][n_out] + b2;
 * (k, n_out) in {(16,64), (64,64), (256,64), (64,16)}. */
int fvdbm_nn_mlp_forward(const float *x, int64_t m, int32_t k, const float *w1, const float *b1,
                         const float *w2, const float *b2, int32_t n_out, float *y, void *stream);
/* Data part of the backward pass of fvdbm_nn_mlp_forward: with
 * z = x w1 + b1 recomputed, dz [m][64] = (dy w2^T) * silu'(z),
 * h [m][64] = silu(z) and, unless dx is NULL, dx [m][k] = dz w1^T. */
int fvdbm_nn_mlp_backward(const float *x, int64_t m, int32_t k, const float *w1, const float *b1,
                          const float *w2, int32_t n_out, const float *dy, float *dx, float *dz,
                          float *h, void *stream);
/* out [ka][nb] = a [m][ka]^T b [m][nb]; (ka, nb) as for the layers.
 * Deterministic: per-chunk partial products, summed in chunk order. */
int fvdbm_nn_wgrad(const float *a, const float *b, int64_t m, int32_t ka, int32_t nb, float *partial,
                   float *out, void *stream);
/* As fvdbm_nn_wgrad, with no workspace argument and no allocation. No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc). The chunk partials go to a static 16 MiB device
 * array (one per GPU, held until the process ends), so
 * ceil(m / fvdbm_nn_chunk_rows()) * ka * nb must not exceed 4194304: at
 * most 262144 rows for (256, 64), 1048576 for (64, 64) and 4194304 for
 * (16, 64) and (64, 16); larger m fails with FVDBM_ERR_INVALID. Calls that
 * overlap on one GPU share the array and race. Results equal
 * fvdbm_nn_wgrad's bit for bit (the same kernels). */
int fvdbm_nn_wgrad_noalloc(const float *a, const float *b, int64_t m, int32_t ka, int32_t nb,
                           float *out, void *stream);
/* out [n] = column sums of a [m][n] (n <= 256), deterministic as above. */
int fvdbm_nn_colsum(const float *a, int64_t m, int32_t n, float *partial, float *out, void *stream);
/* Mesh aggregation over a CSR edge list (row_off [n+1], col [E]):
 * z [n][256] = [h_i | sum_e w_e h_j | sum_e w_e ex_e h_j | sum_e w_e ey_e h_j],
 * h [n][64]. */
int fvdbm_nn_aggregate(const float *h, int64_t n_nodes, const int32_t *row_off, const int32_t *col,
                       const float *w, const float *ex, const float *ey, float *z, void *stream);
/* Its transpose: dh_i = dz_i[0:64] + sum over i's edges e = (i, j) of
 * wt_e (dz_j[64:128] + ext_e dz_j[128:192] + eyt_e dz_j[192:256]), where
 * (wt, ext, eyt) are the weight and direction of the reverse edge j -> i. */
int fvdbm_nn_aggregate_backward(const float *dz, int64_t n_nodes, const int32_t *row_off,
                                const int32_t *col, const float *wt, const float *ext,
                                const float *eyt, float *dh, void *stream);

/* Input graph of one porous case, on the host (host pointers; thread-safe,
 * no device work): mesh nodes xy [n_mesh][2] and mesh edges [n_edges][2],
 * plus a virtual gx x gy grid over [0,1] x [0,height] (nodes n_mesh ..,
 * row-major) and a global node (last), n = n_mesh + gx*gy + 1 nodes. Edges,
 * each stored both ways: mesh edges (unit direction), mesh node <-> its grid
 * cell (direction 0), grid node <-> its 8 neighbours (unit direction between
 * centres), grid node <-> global node (direction 0). CSR sorted by (source,
 * target): row_off [n+1], col [E], w = 1/deg(source), ex, ey, and the
 * reverse edge's wt = 1/deg(target), ext, eyt, with
 * E = fvdbm_nn_graph_edges(n_mesh, n_edges, gx, gy) (-1 for bad sizes).
 * Errors: FVDBM_ERR_INVALID, message in fvdbm_nn_graph_error(). */
int64_t fvdbm_nn_graph_edges(int64_t n_mesh, int64_t n_edges, int32_t gx, int32_t gy);
const char *fvdbm_nn_graph_error(void);
/* Node inputs of one porous case, on the host, rows in the node order
 * above, x [n][width] (width >= 16): columns 0-8 = (x, y, signed distance
 * to the nearest disc, marker == marker_codes[0..3], u_in, tau), each as
 * (float(v) - shift[c]) / scale[c]; virtual nodes take the cell centre
 * (the global node (0.5, 0.25)), y = 0.25 and distance shift[2] when gy = 1,
 * and flag column 9 (grid) or 10 (global); columns 11-14 = the channel's
 * cross-section at the node's x on a section_grid-point grid (open
 * fraction a, 1/a, share of the total resistance sum 1/a^3 downstream, log
 * of that total), as (v - sshift) / sscale; other columns 0. discs
 * [n_discs][3] = (cx, cy, r). */
int fvdbm_nn_case_features(const double *xy, int64_t n_mesh, const int32_t *markers,
                           const int32_t *marker_codes, const double *discs, int32_t n_discs,
                           double u_in, double tau, int32_t gx, int32_t gy, double height,
                           int32_t section_grid, const float *shift, const float *scale,
                           const float *sshift, const float *sscale, int32_t width, float *x);
int fvdbm_nn_case_graph(const double *xy, int64_t n_mesh, const int32_t *edges, int64_t n_edges,
                        int32_t gx, int32_t gy, double height, int32_t *row_off, int32_t *col,
                        float *w, float *ex, float *ey, float *wt, float *ext, float *eyt);

#ifdef __cplusplus
}
#endif

#endif /* FVDBM_H */
