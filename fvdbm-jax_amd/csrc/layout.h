// Host-side construction of the device layout: cell renumbering (reverse
// Cuthill-McKee, cells with a ghost face first), per cell-face structure of
// arrays, boundary-face and active-node records. Pure C++ (no HIP calls), so
// it is also exercised by the CPU tests through fvdbm_layout_* helpers.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fvdbm.h"

namespace fvdbm {

// Bits of a cell-face code (nb array):
//   bit 31 set      : ghost face; bits 0..29 = boundary-face record index
//   bit 31 clear    : interior face; bits 0..29 = neighbour cell (device id)
//   bit 30 set      : this cell sits in stencil slot 1 of the face
constexpr int32_t kGhostBit = int32_t(0x80000000u);
constexpr int32_t kSlot1Bit = 0x40000000;
constexpr int32_t kIndexMask = 0x3fffffff;

template <typename T>
struct HostLayout {
  // C cells occupy Cp device slots; a slot whose new_to_old is -1 is a hole
  // (batched layouts start every case on a tile boundary). The live cells of
  // a tile are its first tile_count[t] slots.
  int C = 0, Cp = 0, F = 0, N = 0, K = 0;
  int ncases = 1;
  std::vector<int32_t> tile_count, tile_case, case_tile_off;
  int n_bcells = 0;   // cells with >= 1 ghost face, stored first when reorder==1
  int nbf = 0;        // boundary-face records
  int nan = 0;        // active (type 1 / 2) nodes
  int bandwidth = 0;
  int last_cell_new = 0;  // device id of reference cell C-1 (JAX's index -1)
  std::vector<int32_t> new_to_old, old_to_new;

  // per cell-face, [3][Cp]
  std::vector<int32_t> nb;
  // same codes with the neighbour given as its slot in the tile's LDS stage
  // (own cells 0..tile-1, then the tile's halo cells)
  std::vector<int32_t> nb_slot;
  // halo of each tile: distinct out-of-tile neighbour cells, ascending
  int max_halo = 0;
  std::vector<int32_t> halo_off, halo_cells;

  std::vector<T> nx, ny, sL;      // face normal, sign * L
  std::vector<T> lwa, lwden2;     // Lax-Wendroff: d0/(d0+d1), 2*(d0+d1)

  // boundary-face records: [2][nbf] endpoint slots in the owning tile's
  // node table, [2][nbf] endpoint node ids, [nbf] ghost/known distance ratio
  std::vector<int32_t> bf_slot, bf_gnode;
  std::vector<T> bf_ratio;

  // per-tile node tables (tile = `block` consecutive device cells): the
  // distinct face nodes the tile's ghost faces read, CSR over tiles.
  // entry >= 0: active-node index; < 0: -1 - node id (stored pdf)
  int ntiles = 0, max_tile_nodes = 0;
  std::vector<int32_t> tile_node_off, tile_nodes;

  // (node, cell) pairs of the active nodes in each tile table, in table
  // order and, per node, in nodes.cells_index order; padding entries with
  // weight 0 are left out (they add +0). pair_cell: device cell id, or
  // -2 - (wrapped id) for a padding entry with a non-zero weight.
  int max_tile_pairs = 0;
  std::vector<int32_t> tile_pair_off;   // [ntiles+1]
  std::vector<int32_t> entry_pair_off;  // [tile_nodes.size()+1], absolute
  std::vector<int32_t> pair_cell;
  std::vector<T> pair_w;

  // active nodes
  std::vector<int32_t> an_gid, an_type;
  std::vector<int32_t> an_cells;  // [nan][K]: device id, or -2 - (wrapped id) for padding
  std::vector<T> an_w;            // [nan][K]  1/d, negatives -> 0
  std::vector<T> an_wsum;         // [nan]     0-initialised ordered sum of an_w
  std::vector<T> an_rho, an_ux, an_uy;

  // Two-step tiles (one launch advances a tile by two steps). With O the
  // tile's cells: S1 = O, its face neighbours (the halo, same order) and
  // every cell a node of an O ghost face reads; S2 = the face neighbours of
  // S1 outside it. Step n is computed for S1 from the inputs of S1 and S2,
  // step n + 1 for O from those results. Per tile, ring = S1 \ O then S2
  // (device ids), staged after the tile's own slots; n1 = |S1 \ O|.
  bool pair_ok = false;       // every tile fits (n1 <= tile, stage within bounds)
  int max_ring = 0;           // max |S1 \ O| + |S2|
  std::vector<int32_t> ring_off, ring_cells, ring_n1;
  // face records of the S1 \ O cells ([3] per ring cell, in ring order):
  // code = stage slot of the neighbour | slot-1 bit, or ghost bit | index
  // into the ring-ghost records; normal and sign * length as for own cells
  std::vector<int32_t> rg_code;
  std::vector<T> rg_nx, rg_ny, rg_sL;
  // ring-ghost records: [2][nrg] endpoint slots in the tile's extended node
  // table, [nrg] ghost/known distance ratio
  int nrg = 0;
  std::vector<int32_t> rgb_slot;
  std::vector<T> rgb_ratio;
  // extended node table of each tile: its own table (same slots) followed
  // by the nodes of the ring cells' ghost faces; pairs as for the own table
  int max_xnodes = 0, max_xpairs = 0;
  std::vector<int32_t> xnode_off, xnodes, xentry_pair_off, xpair_off, xpair_cell;
  std::vector<T> xpair_w;
  // stage slot of each (node, cell) pair of the own tables (parallel to
  // pair_cell), for the node values at the start of the second step
  std::vector<int32_t> pair_slot;

  // faces in reference order (flux observable)
  std::vector<int32_t> fc0, fc1, fn0, fn1;
  std::vector<T> fd0, fd1, fnx, fny, fL;
};

// `tile` is the number of cells per thread block; the cell count is padded
// to a multiple of it. With cell_case (case id of every cell, ncases cases)
// the cells are grouped by case and each case starts on a tile boundary.
template <typename T>
bool build_layout(const fvdbm_mesh_desc &m, int reorder, int tile,
                  HostLayout<T> &out, std::string &err,
                  const int32_t *cell_case = nullptr, int ncases = 1);

// Cells sorted by the Hilbert index of their centroids ([n][2]).
std::vector<int32_t> hilbert_order(int n, const double *xy);

// Reverse Cuthill-McKee over an undirected graph in CSR form.
// Returns perm with perm[new] = old.
std::vector<int32_t> reverse_cuthill_mckee(int n, const std::vector<int32_t> &xadj,
                                           const std::vector<int32_t> &adj);

}  // namespace fvdbm
