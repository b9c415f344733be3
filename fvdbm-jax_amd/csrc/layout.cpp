// Host-side device-layout construction (see layout.h).
#include "layout.h"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace fvdbm {

namespace {

// BFS over the unvisited component of `s`; returns the last level and the
// eccentricity of s. `stamp`/`mark` avoid clearing a per-call array.
int bfs_last_level(int s, const std::vector<int32_t> &xadj,
                   const std::vector<int32_t> &adj, const std::vector<char> &placed,
                   std::vector<int32_t> &mark, int32_t stamp,
                   std::vector<int32_t> &queue, std::vector<int32_t> &last) {
  queue.clear();
  queue.push_back(s);
  mark[s] = stamp;
  int depth = 0;
  size_t level_start = 0, level_end = 1;
  for (;;) {
    for (size_t h = level_start; h < level_end; ++h) {
      int v = queue[h];
      for (int e = xadj[v]; e < xadj[v + 1]; ++e) {
        int u = adj[e];
        if (placed[u] || mark[u] == stamp) continue;
        mark[u] = stamp;
        queue.push_back(u);
      }
    }
    if (queue.size() == level_end) break;
    level_start = level_end;
    level_end = queue.size();
    ++depth;
  }
  last.assign(queue.begin() + level_start, queue.begin() + level_end);
  return depth;
}

// Index of (x, y) along a Hilbert curve over a 2^order x 2^order grid.
uint64_t hilbert_index(uint32_t x, uint32_t y, int order) {
  uint64_t d = 0;
  for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += (uint64_t)s * s * ((3u * rx) ^ ry);
    if (ry == 0) {  // rotate the quadrant
      if (rx == 1) {
        x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
        y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

}  // namespace

std::vector<int32_t> hilbert_order(int n, const double *xy) {
  double x0 = xy[0], x1 = xy[0], y0 = xy[1], y1 = xy[1];
  for (int i = 0; i < n; ++i) {
    x0 = std::min(x0, xy[2 * i]);
    x1 = std::max(x1, xy[2 * i]);
    y0 = std::min(y0, xy[2 * i + 1]);
    y1 = std::max(y1, xy[2 * i + 1]);
  }
  const int order = 20;
  const double span = std::max(x1 - x0, y1 - y0);
  const double scale = span > 0 ? ((1u << order) - 1) / span : 0.0;
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  for (int i = 0; i < n; ++i) {
    const uint32_t gx = (uint32_t)((xy[2 * i] - x0) * scale);
    const uint32_t gy = (uint32_t)((xy[2 * i + 1] - y0) * scale);
    key[i] = {hilbert_index(gx, gy, order), i};
  }
  std::sort(key.begin(), key.end());
  std::vector<int32_t> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = key[i].second;
  return perm;
}

std::vector<int32_t> reverse_cuthill_mckee(int n, const std::vector<int32_t> &xadj,
                                           const std::vector<int32_t> &adj) {
  std::vector<int32_t> deg(n);
  for (int i = 0; i < n; ++i) deg[i] = xadj[i + 1] - xadj[i];
  std::vector<char> placed(n, 0);
  std::vector<int32_t> mark(n, -1), queue, last, order, nbrs;
  order.reserve(n);
  int32_t stamp = 0;
  for (int s0 = 0; s0 < n; ++s0) {
    if (placed[s0]) continue;
    // George-Liu pseudo-peripheral node: restart from the lowest-degree node
    // of the last BFS level while the eccentricity grows.
    int root = s0;
    int ecc = bfs_last_level(root, xadj, adj, placed, mark, stamp++, queue, last);
    for (int it = 0; it < 16; ++it) {
      int best = last[0];
      for (int v : last)
        if (deg[v] < deg[best] || (deg[v] == deg[best] && v < best)) best = v;
      int e2 = bfs_last_level(best, xadj, adj, placed, mark, stamp++, queue, last);
      if (e2 <= ecc) break;
      ecc = e2;
      root = best;
    }
    // Cuthill-McKee: BFS, unplaced neighbours in increasing (degree, id).
    size_t head = order.size();
    order.push_back(root);
    placed[root] = 1;
    while (head < order.size()) {
      int v = order[head++];
      nbrs.clear();
      for (int e = xadj[v]; e < xadj[v + 1]; ++e) {
        int u = adj[e];
        if (!placed[u]) {
          placed[u] = 1;
          nbrs.push_back(u);
        }
      }
      std::sort(nbrs.begin(), nbrs.end(), [&](int a, int b) {
        return deg[a] < deg[b] || (deg[a] == deg[b] && a < b);
      });
      order.insert(order.end(), nbrs.begin(), nbrs.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Two-step tiles (see HostLayout): ring lists, ring face records, extended
// node tables and stage slots of the own-table pairs. Leaves pair_ok false
// when some tile's ring does not fit the one-ring-cell-per-thread pass.
template <typename T>
static void build_pairs(const fvdbm_mesh_desc &m, HostLayout<T> &L, int pad,
                        const std::vector<int32_t> &other, const std::vector<int32_t> &slot) {
  const T *fd = static_cast<const T *>(m.face_dists);
  const T *fn = static_cast<const T *>(m.face_n);
  const T *fL = static_cast<const T *>(m.face_L);
  const int K = L.K;
  const size_t Cp = (size_t)L.Cp;
  L.pair_ok = true;
  L.max_ring = 0;
  L.ring_off.assign(L.ntiles + 1, 0);
  L.ring_cells.clear();
  L.ring_n1.assign(L.ntiles, 0);
  L.rg_code.clear();
  L.rg_nx.clear();
  L.rg_ny.clear();
  L.rg_sL.clear();
  L.rgb_slot.clear();
  L.rgb_ratio.clear();
  std::vector<int32_t> rgb0, rgb1;
  L.xnode_off.assign(L.ntiles + 1, 0);
  L.xnodes.clear();
  L.xentry_pair_off.clear();
  L.xpair_off.assign(L.ntiles + 1, 0);
  L.xpair_cell.clear();
  L.xpair_w.clear();
  L.pair_slot.assign(L.pair_cell.size(), -1);
  L.max_xnodes = 0;
  L.max_xpairs = 0;
  std::vector<int32_t> node_an(L.N, -1);
  for (int a = 0; a < L.nan; ++a) node_an[L.an_gid[a]] = a;
  std::vector<int32_t> stage_of(Cp, -1);  // device cell -> stage slot in the current tile
  std::vector<int32_t> xslot_of(L.N, -1);  // node -> slot in the current extended table
  std::vector<int32_t> ring, s2, touched_nodes;
  auto pair_cell_id = [&](int32_t cj) { return cj >= 0 ? cj : -2 - cj; };
  for (int t = 0; t < L.ntiles; ++t) {
    const int lo = t * pad, hi = lo + L.tile_count[t];
    for (int i = lo; i < hi; ++i) stage_of[i] = i - lo;
    // S1 \ O: the halo (ascending, as staged by the one-step kernel), then
    // the other cells the own node table's pairs read (ascending)
    ring.assign(L.halo_cells.begin() + L.halo_off[t], L.halo_cells.begin() + L.halo_off[t + 1]);
    for (size_t r = 0; r < ring.size(); ++r) stage_of[ring[r]] = pad + (int)r;
    std::vector<int32_t> f1;
    for (int p = L.tile_pair_off[t]; p < L.tile_pair_off[t + 1]; ++p) {
      const int c = pair_cell_id(L.pair_cell[p]);
      if (stage_of[c] < 0) f1.push_back(c);
    }
    std::sort(f1.begin(), f1.end());
    f1.erase(std::unique(f1.begin(), f1.end()), f1.end());
    for (int c : f1) {
      stage_of[c] = pad + (int)ring.size();
      ring.push_back(c);
    }
    const int n1 = (int)ring.size();
    // S2: face neighbours of S1 \ O outside S1
    s2.clear();
    for (int r = 0; r < n1; ++r) {
      const int c = L.new_to_old[ring[r]];
      for (int k = 0; k < 3; ++k) {
        const int o = other[3 * c + k];
        if (o < 0) continue;
        const int j = L.old_to_new[o];
        if (stage_of[j] < 0) s2.push_back(j);
      }
    }
    std::sort(s2.begin(), s2.end());
    s2.erase(std::unique(s2.begin(), s2.end()), s2.end());
    for (size_t r = 0; r < s2.size(); ++r) stage_of[s2[r]] = pad + n1 + (int)r;
    if (n1 > pad) L.pair_ok = false;
    L.ring_n1[t] = n1;
    L.ring_off[t] = (int32_t)L.ring_cells.size();
    L.ring_cells.insert(L.ring_cells.end(), ring.begin(), ring.end());
    L.ring_cells.insert(L.ring_cells.end(), s2.begin(), s2.end());
    L.max_ring = std::max(L.max_ring, n1 + (int)s2.size());
    // extended node table: the own table, then the ring cells' ghost nodes
    L.xnode_off[t] = (int32_t)L.xnodes.size();
    touched_nodes.clear();
    for (int e = L.tile_node_off[t]; e < L.tile_node_off[t + 1]; ++e) {
      const int v = L.tile_nodes[e];
      const int node = v >= 0 ? L.an_gid[v] : -1 - v;
      xslot_of[node] = e - L.tile_node_off[t];
      touched_nodes.push_back(node);
      L.xnodes.push_back(v);
    }
    auto xslot = [&](int node) {
      if (xslot_of[node] < 0) {
        xslot_of[node] = (int32_t)(L.xnodes.size() - L.xnode_off[t]);
        L.xnodes.push_back(node_an[node] >= 0 ? node_an[node] : -1 - node);
        touched_nodes.push_back(node);
      }
      return xslot_of[node];
    };
    // ring face records
    for (int r = 0; r < n1; ++r) {
      const int c = L.new_to_old[ring[r]];
      for (int k = 0; k < 3; ++k) {
        const int fid = m.cell_faces[3 * c + k];
        const int sl = slot[3 * c + k], o = other[3 * c + k];
        int32_t code;
        if (o >= 0) {
          code = stage_of[L.old_to_new[o]];
        } else {
          const T d0 = fd[2 * (size_t)fid], d1 = fd[2 * (size_t)fid + 1];
          code = kGhostBit | (int32_t)L.rgb_ratio.size();
          L.rgb_ratio.push_back(sl == 0 ? d1 / d0 : d0 / d1);
          rgb0.push_back(xslot(m.face_nodes[2 * (size_t)fid]));
          rgb1.push_back(xslot(m.face_nodes[2 * (size_t)fid + 1]));
        }
        if (sl == 1) code |= kSlot1Bit;
        L.rg_code.push_back(code);
        L.rg_nx.push_back(fn[2 * (size_t)fid]);
        L.rg_ny.push_back(fn[2 * (size_t)fid + 1]);
        L.rg_sL.push_back(T(m.cell_signs[3 * c + k]) * fL[fid]);
      }
    }
    // pairs of the extended table (own entries: the own pairs)
    L.xpair_off[t] = (int32_t)L.xpair_cell.size();
    for (size_t e = L.xnode_off[t]; e < L.xnodes.size(); ++e) {
      L.xentry_pair_off.push_back((int32_t)L.xpair_cell.size());
      const int an = L.xnodes[e];
      if (an < 0) continue;
      for (int k = 0; k < K; ++k) {
        const int32_t cj = L.an_cells[(size_t)an * K + k];
        const T w = L.an_w[(size_t)an * K + k];
        if (cj < 0 && w == T(0)) continue;
        L.xpair_cell.push_back(cj);
        L.xpair_w.push_back(w);
      }
    }
    L.max_xnodes = std::max(L.max_xnodes, (int)L.xnodes.size() - L.xnode_off[t]);
    L.max_xpairs = std::max(L.max_xpairs, (int)L.xpair_cell.size() - L.xpair_off[t]);
    // stage slots of the own pairs
    for (int p = L.tile_pair_off[t]; p < L.tile_pair_off[t + 1]; ++p)
      L.pair_slot[p] = stage_of[pair_cell_id(L.pair_cell[p])];
    // reset the per-tile maps
    for (int i = lo; i < hi; ++i) stage_of[i] = -1;
    for (int c : ring) stage_of[c] = -1;
    for (int c : s2) stage_of[c] = -1;
    for (int nd_ : touched_nodes) xslot_of[nd_] = -1;
  }
  L.ring_off[L.ntiles] = (int32_t)L.ring_cells.size();
  L.xnode_off[L.ntiles] = (int32_t)L.xnodes.size();
  L.xpair_off[L.ntiles] = (int32_t)L.xpair_cell.size();
  L.xentry_pair_off.push_back((int32_t)L.xpair_cell.size());
  L.nrg = (int)L.rgb_ratio.size();
  L.rgb_slot.resize(2 * (size_t)L.nrg);
  std::copy(rgb0.begin(), rgb0.end(), L.rgb_slot.begin());
  std::copy(rgb1.begin(), rgb1.end(), L.rgb_slot.begin() + L.nrg);
  // defensive range checks of everything k_step2 indexes with these tables
  for (int t = 0, rc = 0; t < L.ntiles && L.pair_ok; ++t) {
    const int nr = L.ring_off[t + 1] - L.ring_off[t], n1 = L.ring_n1[t];
    const int nx = L.xnode_off[t + 1] - L.xnode_off[t];
    for (int r = 0; r < n1; ++r, ++rc)
      for (int k = 0; k < 3; ++k) {
        const int32_t code = L.rg_code[3 * (size_t)rc + k];
        if (code >= 0) {
          if ((code & kIndexMask) >= pad + nr) L.pair_ok = false;
        } else {
          const int b = code & kIndexMask;
          if (b >= L.nrg || L.rgb_slot[b] >= nx || L.rgb_slot[L.nrg + b] >= nx) L.pair_ok = false;
        }
      }
    for (int q = L.tile_pair_off[t]; q < L.tile_pair_off[t + 1]; ++q)
      if (L.pair_slot[q] < 0 || L.pair_slot[q] >= pad + n1) L.pair_ok = false;
  }
}

template <typename T>
bool build_layout(const fvdbm_mesh_desc &m, int reorder, int tile,
                  HostLayout<T> &L, std::string &err, const int32_t *cell_case, int ncases) {
  const int C = m.n_cells, F = m.n_faces, N = m.n_nodes, K = m.k_max;
  if (C <= 0 || F <= 0 || N <= 0 || K < 0) {
    err = "mesh must have cells, faces and nodes";
    return false;
  }
  if (!m.cell_faces || !m.cell_signs || !m.face_cells || !m.face_dists ||
      !m.face_nodes || !m.face_n || !m.face_L || !m.node_type ||
      (K > 0 && (!m.node_cells || !m.node_dists)) || !m.node_rho || !m.node_vel) {
    err = "mesh descriptor has a null array";
    return false;
  }
  if (C > kIndexMask || F > kIndexMask || N > kIndexMask) {
    err = "mesh too large for 30-bit indices";
    return false;
  }
  const T *fd = static_cast<const T *>(m.face_dists);
  const T *fn = static_cast<const T *>(m.face_n);
  const T *fL = static_cast<const T *>(m.face_L);
  const T *nd = static_cast<const T *>(m.node_dists);
  const T *nrho = static_cast<const T *>(m.node_rho);
  const T *nvel = static_cast<const T *>(m.node_vel);

  L.C = C;
  L.F = F;
  L.N = N;
  L.K = K;

  // --- per cell-face: slot of the cell in the face stencil and the other side
  std::vector<int32_t> other(3 * (size_t)C), slot(3 * (size_t)C);
  std::vector<char> has_ghost(C, 0);
  for (int c = 0; c < C; ++c) {
    for (int k = 0; k < 3; ++k) {
      int fid = m.cell_faces[3 * c + k];
      if (fid < 0 || fid >= F) {
        err = "cells.face_indices out of range at cell " + std::to_string(c);
        return false;
      }
      int s = m.cell_signs[3 * c + k];
      if (s < -1 || s > 1) {
        err = "cells.face_normals must hold signs in {-1,0,1}";
        return false;
      }
      int a = m.face_cells[2 * fid], b = m.face_cells[2 * fid + 1];
      if (a < -1 || a >= C || b < -1 || b >= C) {
        err = "faces.stencil_cells_index out of range at face " + std::to_string(fid);
        return false;
      }
      int sl, o;
      if (a == c && b != c) {
        sl = 0;
        o = b;
      } else if (b == c && a != c) {
        sl = 1;
        o = a;
      } else {
        err = "cell " + std::to_string(c) + " is not in the stencil of its face " +
              std::to_string(fid);
        return false;
      }
      slot[3 * c + k] = sl;
      other[3 * c + k] = o;
      if (o == -1) {
        has_ghost[c] = 1;
        int p = m.face_nodes[2 * fid], r = m.face_nodes[2 * fid + 1];
        if (p < 0 || p >= N || r < 0 || r >= N) {
          err = "faces.nodes_index out of range at face " + std::to_string(fid);
          return false;
        }
      }
    }
  }

  // --- cell order
  std::vector<int32_t> perm(C);
  if (reorder < 0 || reorder > 4) {
    err = "reorder must be 0, 1, 2, 3 or 4";
    return false;
  }
  if (reorder >= 3 && !m.cell_xy) {
    err = "reorder 3/4 (Hilbert order) needs cell centroids (cell_xy)";
    return false;
  }
  if (reorder == 0) {
    std::iota(perm.begin(), perm.end(), 0);
  } else if (reorder >= 3) {
    perm = hilbert_order(C, m.cell_xy);
    if (reorder == 3)
      std::stable_partition(perm.begin(), perm.end(),
                            [&](int32_t c) { return has_ghost[c] != 0; });
  } else {
    std::vector<int32_t> xadj(C + 1, 0), adj;
    adj.reserve(3 * (size_t)C);
    for (int c = 0; c < C; ++c) {
      for (int k = 0; k < 3; ++k)
        if (other[3 * c + k] >= 0) adj.push_back(other[3 * c + k]);
      xadj[c + 1] = (int32_t)adj.size();
    }
    perm = reverse_cuthill_mckee(C, xadj, adj);
    if (reorder == 1) {
      std::stable_partition(perm.begin(), perm.end(),
                            [&](int32_t c) { return has_ghost[c] != 0; });
    }
  }
  const int pad = tile > 0 ? tile : 1;
  // device slots: cases (if any) grouped and tile-aligned, holes = -1
  if (cell_case) {
    if (ncases < 1) {
      err = "ncases must be >= 1";
      return false;
    }
    for (int c = 0; c < C; ++c)
      if (cell_case[c] < 0 || cell_case[c] >= ncases) {
        err = "cell case id out of range at cell " + std::to_string(c);
        return false;
      }
    for (int c = 0; c < C; ++c)
      for (int k = 0; k < 3; ++k)
        if (other[3 * c + k] >= 0 && cell_case[other[3 * c + k]] != cell_case[c]) {
          err = "a face joins cells of different cases";
          return false;
        }
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int32_t a, int32_t b) { return cell_case[a] < cell_case[b]; });
  }
  L.ncases = cell_case ? ncases : 1;
  L.new_to_old.clear();
  L.case_tile_off.assign(L.ncases + 1, 0);
  {
    size_t i = 0;
    for (int k = 0; k < L.ncases; ++k) {
      L.case_tile_off[k] = (int32_t)(L.new_to_old.size() / pad);
      while (i < perm.size() && (!cell_case || cell_case[perm[i]] == k))
        L.new_to_old.push_back(perm[i++]);
      while (L.new_to_old.size() % pad) L.new_to_old.push_back(-1);
    }
    L.case_tile_off[L.ncases] = (int32_t)(L.new_to_old.size() / pad);
  }
  L.Cp = (int)L.new_to_old.size();
  const size_t Cp = (size_t)L.Cp;
  L.old_to_new.assign(C, 0);
  for (size_t i = 0; i < Cp; ++i)
    if (L.new_to_old[i] >= 0) L.old_to_new[L.new_to_old[i]] = (int32_t)i;
  L.n_bcells = 0;
  for (int c = 0; c < C; ++c) L.n_bcells += has_ghost[c];
  L.last_cell_new = L.old_to_new[C - 1];
  L.tile_count.assign(Cp / pad, 0);
  L.tile_case.assign(Cp / pad, 0);
  for (size_t i = 0; i < Cp; ++i)
    if (L.new_to_old[i] >= 0) ++L.tile_count[i / pad];
  for (int k = 0; k < L.ncases; ++k)
    for (int t = L.case_tile_off[k]; t < L.case_tile_off[k + 1]; ++t) L.tile_case[t] = k;

  // --- active nodes
  std::vector<int32_t> node_an(N, -1);
  for (int i = 0; i < N; ++i) {
    int t = m.node_type[i];
    if (t == 1 || t == 2) {
      node_an[i] = (int32_t)L.an_gid.size();
      L.an_gid.push_back(i);
      L.an_type.push_back(t);
    } else if (t != 0) {
      err = "nodes.type must be 0, 1 or 2";
      return false;
    }
  }
  L.nan = (int)L.an_gid.size();
  L.an_cells.assign((size_t)L.nan * K, 0);
  L.an_w.assign((size_t)L.nan * K, T(0));
  L.an_wsum.assign(L.nan, T(0));
  L.an_rho.resize(L.nan);
  L.an_ux.resize(L.nan);
  L.an_uy.resize(L.nan);
  for (int a = 0; a < L.nan; ++a) {
    int i = L.an_gid[a];
    T wsum = T(0);
    for (int k = 0; k < K; ++k) {
      int ci = m.node_cells[(size_t)i * K + k];
      T d = nd[(size_t)i * K + k];
      // utils.extrapolate: w = 1/d, negative weights -> 0 (utils/utils.py:58-59)
      T w = T(1) / d;
      if (w < T(0)) w = T(0);
      if (ci < -1 || ci >= C) {
        err = "nodes.cells_index out of range at node " + std::to_string(i);
        return false;
      }
      L.an_cells[(size_t)a * K + k] = ci >= 0 ? L.old_to_new[ci] : -2 - L.last_cell_new;
      L.an_w[(size_t)a * K + k] = w;
      wsum += w;
    }
    L.an_wsum[a] = wsum;
    L.an_rho[a] = nrho[i];
    L.an_ux[a] = nvel[2 * (size_t)i];
    L.an_uy[a] = nvel[2 * (size_t)i + 1];
  }

  // --- per cell-face arrays in device order
  L.nb.assign(3 * Cp, 0);
  L.nx.assign(3 * Cp, T(0));
  L.ny.assign(3 * Cp, T(0));
  L.sL.assign(3 * Cp, T(0));
  L.lwa.assign(3 * Cp, T(0));
  L.lwden2.assign(3 * Cp, T(1));
  L.bf_ratio.clear();
  std::vector<int32_t> bf0, bf1, bs0, bs1;
  L.ntiles = L.Cp / pad;
  L.tile_node_off.assign(L.ntiles + 1, 0);
  L.tile_nodes.clear();
  std::vector<int32_t> slot_of(N, -1);  // node -> slot in the current tile table
  std::vector<int32_t> touched;
  int cur_tile = -1;
  auto tile_slot = [&](int node) {
    if (slot_of[node] < 0) {
      slot_of[node] = (int32_t)(L.tile_nodes.size() - L.tile_node_off[cur_tile]);
      L.tile_nodes.push_back(node_an[node] >= 0 ? node_an[node] : -1 - node);
      touched.push_back(node);
    }
    return slot_of[node];
  };
  int bw = 0;
  for (int i = 0; i < (int)Cp; ++i) {
    const int t = i / pad;
    while (cur_tile < t) {  // close the previous tile's node table
      for (int nd_ : touched) slot_of[nd_] = -1;
      touched.clear();
      ++cur_tile;
      L.tile_node_off[cur_tile] = (int32_t)L.tile_nodes.size();
    }
    const int c = L.new_to_old[i];
    if (c < 0) continue;  // hole
    for (int k = 0; k < 3; ++k) {
      int fid = m.cell_faces[3 * c + k];
      int sl = slot[3 * c + k], o = other[3 * c + k];
      size_t at = (size_t)k * Cp + i;
      T d0 = fd[2 * (size_t)fid], d1 = fd[2 * (size_t)fid + 1];
      L.nx[at] = fn[2 * (size_t)fid];
      L.ny[at] = fn[2 * (size_t)fid + 1];
      L.sL[at] = T(m.cell_signs[3 * c + k]) * fL[fid];
      T s = d0 + d1;
      L.lwa[at] = d0 / s;
      L.lwden2[at] = T(2) * s;
      int32_t code;
      if (o >= 0) {
        int j = L.old_to_new[o];
        code = j;
        bw = std::max(bw, std::abs(j - i));
      } else {
        int bf = (int)L.bf_ratio.size();
        // ghost distance / known distance: (d1/d0) when the cell is slot 0
        L.bf_ratio.push_back(sl == 0 ? d1 / d0 : d0 / d1);
        int p = m.face_nodes[2 * (size_t)fid], r = m.face_nodes[2 * (size_t)fid + 1];
        bf0.push_back(p);
        bf1.push_back(r);
        bs0.push_back(tile_slot(p));
        bs1.push_back(tile_slot(r));
        code = kGhostBit | bf;
      }
      if (sl == 1) code |= kSlot1Bit;
      L.nb[at] = code;
    }
  }
  while (cur_tile < L.ntiles - 1) {
    ++cur_tile;
    L.tile_node_off[cur_tile] = (int32_t)L.tile_nodes.size();
  }
  L.tile_node_off[L.ntiles] = (int32_t)L.tile_nodes.size();
  L.max_tile_nodes = 0;
  for (int t = 0; t < L.ntiles; ++t)
    L.max_tile_nodes = std::max(L.max_tile_nodes, L.tile_node_off[t + 1] - L.tile_node_off[t]);
  // (node, cell) pairs per tile table entry
  L.tile_pair_off.assign(L.ntiles + 1, 0);
  L.entry_pair_off.assign(L.tile_nodes.size() + 1, 0);
  L.pair_cell.clear();
  L.pair_w.clear();
  L.max_tile_pairs = 0;
  for (int t = 0; t < L.ntiles; ++t) {
    L.tile_pair_off[t] = (int32_t)L.pair_cell.size();
    for (int e = L.tile_node_off[t]; e < L.tile_node_off[t + 1]; ++e) {
      L.entry_pair_off[e] = (int32_t)L.pair_cell.size();
      const int an = L.tile_nodes[e];
      if (an < 0) continue;
      for (int k = 0; k < K; ++k) {
        const int32_t cj = L.an_cells[(size_t)an * K + k];
        const T w = L.an_w[(size_t)an * K + k];
        if (cj < 0 && w == T(0)) continue;
        L.pair_cell.push_back(cj);
        L.pair_w.push_back(w);
      }
    }
    L.max_tile_pairs =
        std::max(L.max_tile_pairs, (int)L.pair_cell.size() - L.tile_pair_off[t]);
  }
  L.tile_pair_off[L.ntiles] = (int32_t)L.pair_cell.size();
  L.entry_pair_off[L.tile_nodes.size()] = (int32_t)L.pair_cell.size();
  // halos and LDS-slot codes
  L.halo_off.assign(L.ntiles + 1, 0);
  L.halo_cells.clear();
  L.nb_slot = L.nb;
  L.max_halo = 0;
  {
    std::vector<int32_t> h;
    for (int t = 0; t < L.ntiles; ++t) {
      L.halo_off[t] = (int32_t)L.halo_cells.size();
      const int lo = t * pad, hi = t * pad + L.tile_count[t];
      h.clear();
      for (int k = 0; k < 3; ++k)
        for (int i = lo; i < hi; ++i) {
          const int32_t code = L.nb[(size_t)k * Cp + i];
          if (code < 0) continue;
          const int j = code & kIndexMask;
          if (j < lo || j >= t * pad + pad) h.push_back(j);
        }
      std::sort(h.begin(), h.end());
      h.erase(std::unique(h.begin(), h.end()), h.end());
      L.halo_cells.insert(L.halo_cells.end(), h.begin(), h.end());
      L.max_halo = std::max(L.max_halo, (int)h.size());
      for (int k = 0; k < 3; ++k)
        for (int i = lo; i < hi; ++i) {
          int32_t &code = L.nb_slot[(size_t)k * Cp + i];
          if (code < 0) continue;
          const int j = code & kIndexMask;
          int slot;
          if (j >= t * pad && j < t * pad + pad) {
            slot = j - t * pad;
          } else {
            slot = pad + (int)(std::lower_bound(h.begin(), h.end(), j) - h.begin());
          }
          code = (code & ~kIndexMask) | slot;
        }
    }
    L.halo_off[L.ntiles] = (int32_t)L.halo_cells.size();
  }
  L.nbf = (int)L.bf_ratio.size();
  L.bf_gnode.resize(2 * (size_t)L.nbf);
  L.bf_slot.resize(2 * (size_t)L.nbf);
  std::copy(bf0.begin(), bf0.end(), L.bf_gnode.begin());
  std::copy(bf1.begin(), bf1.end(), L.bf_gnode.begin() + L.nbf);
  std::copy(bs0.begin(), bs0.end(), L.bf_slot.begin());
  std::copy(bs1.begin(), bs1.end(), L.bf_slot.begin() + L.nbf);
  L.bandwidth = bw;

  build_pairs(m, L, pad, other, slot);

  // --- faces in reference order (for the flux observable)
  L.fc0.resize(F);
  L.fc1.resize(F);
  L.fn0.resize(F);
  L.fn1.resize(F);
  L.fd0.resize(F);
  L.fd1.resize(F);
  L.fnx.resize(F);
  L.fny.resize(F);
  L.fL.resize(F);
  for (int j = 0; j < F; ++j) {
    int a = m.face_cells[2 * j], b = m.face_cells[2 * j + 1];
    if (a < -1 || a >= C || b < -1 || b >= C) {
      err = "faces.stencil_cells_index out of range at face " + std::to_string(j);
      return false;
    }
    int p = m.face_nodes[2 * j], r = m.face_nodes[2 * j + 1];
    if (p < 0 || p >= N || r < 0 || r >= N) {
      err = "faces.nodes_index out of range at face " + std::to_string(j);
      return false;
    }
    L.fc0[j] = a >= 0 ? L.old_to_new[a] : -1;
    L.fc1[j] = b >= 0 ? L.old_to_new[b] : -1;
    L.fn0[j] = p;
    L.fn1[j] = r;
    L.fd0[j] = fd[2 * (size_t)j];
    L.fd1[j] = fd[2 * (size_t)j + 1];
    L.fnx[j] = fn[2 * (size_t)j];
    L.fny[j] = fn[2 * (size_t)j + 1];
    L.fL[j] = fL[j];
  }
  return true;
}

template bool build_layout<float>(const fvdbm_mesh_desc &, int, int, HostLayout<float> &,
                                  std::string &, const int32_t *, int);
template bool build_layout<double>(const fvdbm_mesh_desc &, int, int, HostLayout<double> &,
                                   std::string &, const int32_t *, int);

}  // namespace fvdbm

// Host-only entry point (no device needed): the device cell order and layout
// statistics the library would use for this mesh.
extern "C" int fvdbm_layout_order(const fvdbm_mesh_desc *mesh, int32_t precision,
                                  int32_t reorder, int32_t *new_to_old, int32_t *stats) {
  if (!mesh || !new_to_old || !stats) return FVDBM_ERR_INVALID;
  std::string err;
  if (precision == FVDBM_FP64) {
    fvdbm::HostLayout<double> L;
    if (!fvdbm::build_layout<double>(*mesh, reorder, 1, L, err)) return FVDBM_ERR_INVALID;
    std::copy(L.new_to_old.begin(), L.new_to_old.end(), new_to_old);
    stats[0] = L.n_bcells; stats[1] = L.nbf; stats[2] = L.nan; stats[3] = L.bandwidth;
  } else {
    fvdbm::HostLayout<float> L;
    if (!fvdbm::build_layout<float>(*mesh, reorder, 1, L, err)) return FVDBM_ERR_INVALID;
    std::copy(L.new_to_old.begin(), L.new_to_old.end(), new_to_old);
    stats[0] = L.n_bcells; stats[1] = L.nbf; stats[2] = L.nan; stats[3] = L.bandwidth;
  }
  return FVDBM_OK;
}
