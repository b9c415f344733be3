// Inputs of the warm-start surrogate for one porous case, built on the
// host: the node features and the input graph (the mesh's node graph plus a
// virtual grid and a global node, in the CSR form the aggregation kernels
// read). No reference counterpart (the reference has no learned component);
// declared in include/fvdbm.h. Both functions are plain host code without
// shared state, so threads build several cases at once.
//
// Nodes: the n_mesh mesh nodes, then the gx * gy grid nodes (row-major,
// index j * gx + i, cell [i/gx, (i+1)/gx] x [j/gy, (j+1)/gy] * height), then
// the global node. Undirected edges, each stored both ways with opposite
// unit directions:
//   * every mesh edge, direction unit(p_v - p_u);
//   * every mesh node and the grid node of its cell, direction 0;
//   * every grid node and its neighbours at (+1,0), (+1,+1), (0,+1), (-1,+1)
//     (so each of the 8 neighbours once), direction unit(centre offset);
//   * every grid node and the global node, direction 0.
// CSR rows are sorted by (source, target). Per edge: w = 1/deg(source),
// (ex, ey); for the backward pass the reverse edge's weight wt =
// 1/deg(target) and direction (ext, eyt). Every division, hypot and cast
// is the one the NumPy construction performs (float64, then float32), so
// the arrays are identical to it.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "fvdbm.h"

namespace {

thread_local std::string g_graph_err;

int64_t grid_pairs(int64_t gx, int64_t gy) {
  return (gx - 1) * gy + (gx - 1) * (gy - 1) + gx * (gy - 1) + (gx - 1) * (gy - 1);
}

// min over discs (cx, cy, r) of |p - c| - r; +inf without discs
double signed_distance(double px, double py, const double *discs, int nd) {
  double d = INFINITY;
  for (int k = 0; k < nd; ++k) {
    const double v = std::hypot(px - discs[3 * k], py - discs[3 * k + 1]) - discs[3 * k + 2];
    d = std::min(d, v);
  }
  return d;
}

float scaled(double v, float shift, float scale) { return ((float)v - shift) / scale; }

}  // namespace

extern "C" {

int64_t fvdbm_nn_graph_edges(int64_t n_mesh, int64_t n_edges, int32_t gx, int32_t gy) {
  if (n_mesh < 0 || n_edges < 0 || gx < 1 || gy < 1) return -1;
  const int64_t s = (int64_t)gx * gy;
  return 2 * (n_edges + n_mesh + grid_pairs(gx, gy) + s);
}

const char *fvdbm_nn_graph_error(void) { return g_graph_err.c_str(); }

int fvdbm_nn_case_features(const double *xy, int64_t n_mesh, const int32_t *markers,
                           const int32_t *marker_codes, const double *discs, int32_t n_discs,
                           double u_in, double tau, int32_t gx, int32_t gy, double height,
                           int32_t section_grid, const float *shift, const float *scale,
                           const float *sshift, const float *sscale, int32_t width, float *x) {
  if (n_mesh < 0 || gx < 1 || gy < 1 || n_discs < 0 || section_grid < 1 || width < 16 ||
      !(height > 0.0)) {
    g_graph_err = "bad sizes";
    return FVDBM_ERR_INVALID;
  }
  if ((n_mesh > 0 && (!xy || !markers)) || !marker_codes || (n_discs > 0 && !discs) || !shift ||
      !scale || !sshift || !sscale || !x) {
    g_graph_err = "null pointer";
    return FVDBM_ERR_INVALID;
  }
  const int64_t S = (int64_t)gx * gy, n = n_mesh + S + 1;
  std::fill(x, x + n * width, 0.0f);
  // cross-section of the channel on a grid of x: open fraction a, the
  // hydraulic resistance 1/a^3, its total and the share downstream of x
  const int G = section_grid;
  std::vector<double> a(G), down(G);
  double total = 0.0;
  for (int i = 0; i < G; ++i) {
    const double xs = (i + 0.5) / G;
    double blocked = 0.0;
    for (int k = 0; k < n_discs; ++k) {
      const double dxs = xs - discs[3 * k], r = discs[3 * k + 2];
      const double d2 = r * r - dxs * dxs;
      blocked += d2 > 0.0 ? 2.0 * std::sqrt(d2) : 0.0;
    }
    a[i] = std::min(std::max(1.0 - blocked / height, 0.05), 1.0);
    total += 1.0 / std::pow(a[i], 3.0);
  }
  total /= G;
  double run = 0.0;
  for (int i = G - 1; i >= 0; --i) {
    run += 1.0 / std::pow(a[i], 3.0);
    down[i] = run / G / total;
  }
  const float ltot = (float)std::log(total);
  auto section = [&](int64_t row, double px) {
    int64_t i = (int64_t)(px * G);
    i = std::min<int64_t>(std::max<int64_t>(i, 0), G - 1);
    float *o = x + row * width + 11;
    o[0] = ((float)a[i] - sshift[0]) / sscale[0];
    o[1] = ((float)(1.0 / a[i]) - sshift[1]) / sscale[1];
    o[2] = ((float)down[i] - sshift[2]) / sscale[2];
    o[3] = (ltot - sshift[3]) / sscale[3];
  };
  for (int64_t i = 0; i < n_mesh; ++i) {
    float *o = x + i * width;
    const double px = xy[2 * i], py = xy[2 * i + 1];
    o[0] = scaled(px, shift[0], scale[0]);
    o[1] = scaled(py, shift[1], scale[1]);
    o[2] = scaled(signed_distance(px, py, discs, n_discs), shift[2], scale[2]);
    for (int j = 0; j < 4; ++j)
      o[3 + j] = scaled(markers[i] == marker_codes[j] ? 1.0 : 0.0, shift[3 + j], scale[3 + j]);
    o[7] = scaled(u_in, shift[7], scale[7]);
    o[8] = scaled(tau, shift[8], scale[8]);
    section(i, px);
  }
  // virtual nodes: grid cell centres (y and distance only for a 2D grid),
  // then the global node at (0.5, 0.25)
  for (int64_t s = 0; s <= S; ++s) {
    float *o = x + (n_mesh + s) * width;
    double vx = 0.5, vy = 0.25, vd = shift[2];
    if (s < S) {
      const int64_t i = s % gx, j = s / gx;
      vx = (i + 0.5) / gx;
      if (gy > 1) {
        vy = (j + 0.5) / gy * height;
        vd = signed_distance(vx, vy, discs, n_discs);
      }
    }
    o[0] = scaled(vx, shift[0], scale[0]);
    o[1] = scaled(vy, shift[1], scale[1]);
    o[2] = scaled(vd, shift[2], scale[2]);
    for (int j = 0; j < 4; ++j) o[3 + j] = scaled(0.0, shift[3 + j], scale[3 + j]);
    o[7] = scaled(u_in, shift[7], scale[7]);
    o[8] = scaled(tau, shift[8], scale[8]);
    o[s < S ? 9 : 10] = 1.0f;
    section(n_mesh + s, vx);
  }
  return FVDBM_OK;
}

int fvdbm_nn_case_graph(const double *xy, int64_t n_mesh, const int32_t *edges, int64_t n_edges,
                        int32_t gx, int32_t gy, double height, int32_t *row_off, int32_t *col,
                        float *w, float *ex, float *ey, float *wt, float *ext, float *eyt) {
  const int64_t ne = fvdbm_nn_graph_edges(n_mesh, n_edges, gx, gy);
  if (ne < 0 || !(height > 0.0)) {
    g_graph_err = "bad sizes";
    return FVDBM_ERR_INVALID;
  }
  const int64_t S = (int64_t)gx * gy, n = n_mesh + S + 1;
  if (n > INT32_MAX || ne > INT32_MAX) {
    g_graph_err = "graph too large";
    return FVDBM_ERR_INVALID;
  }
  if ((n_mesh > 0 && !xy) || (n_edges > 0 && !edges) || !row_off || !col || !w || !ex || !ey ||
      !wt || !ext || !eyt) {
    g_graph_err = "null pointer";
    return FVDBM_ERR_INVALID;
  }
  // directed edges in pairs: 2k = (u -> v, d), 2k + 1 = (v -> u, -d)
  std::vector<int32_t> src(ne), dst(ne);
  std::vector<double> dx(ne), dy(ne);
  int64_t k = 0;
  auto add = [&](int64_t u, int64_t v, double ux, double uy) {
    src[k] = (int32_t)u;
    dst[k] = (int32_t)v;
    dx[k] = ux;
    dy[k] = uy;
    src[k + 1] = (int32_t)v;
    dst[k + 1] = (int32_t)u;
    dx[k + 1] = -ux;
    dy[k + 1] = -uy;
    k += 2;
  };
  for (int64_t e = 0; e < n_edges; ++e) {
    const int64_t a = edges[2 * e], b = edges[2 * e + 1];
    if (a < 0 || b < 0 || a >= n_mesh || b >= n_mesh || a == b) {
      g_graph_err = "mesh edge " + std::to_string(e) + " out of range";
      return FVDBM_ERR_INVALID;
    }
    const double ddx = xy[2 * b] - xy[2 * a], ddy = xy[2 * b + 1] - xy[2 * a + 1];
    const double L = std::hypot(ddx, ddy);
    add(a, b, ddx / L, ddy / L);
  }
  for (int64_t i = 0; i < n_mesh; ++i) {
    int64_t ci = (int64_t)(xy[2 * i] * gx), cj = (int64_t)(xy[2 * i + 1] / height * gy);
    ci = std::min<int64_t>(ci, gx - 1);
    cj = std::min<int64_t>(cj, gy - 1);
    if (ci < 0 || cj < 0) {
      g_graph_err = "mesh node " + std::to_string(i) + " outside the grid";
      return FVDBM_ERR_INVALID;
    }
    add(i, n_mesh + cj * gx + ci, 0.0, 0.0);
  }
  const int dirs[4][2] = {{1, 0}, {1, 1}, {0, 1}, {-1, 1}};
  for (const auto &d : dirs) {
    const double ox = (double)d[0] / gx, oy = (double)d[1] / gy * height;
    const double L = std::hypot(ox, oy);
    for (int64_t j = 0; j < gy; ++j)
      for (int64_t i = 0; i < gx; ++i) {
        const int64_t i2 = i + d[0], j2 = j + d[1];
        if (i2 < 0 || i2 >= gx || j2 < 0 || j2 >= gy) continue;
        add(n_mesh + j * gx + i, n_mesh + j2 * gx + i2, ox / L, oy / L);
      }
  }
  for (int64_t s = 0; s < S; ++s) add(n_mesh + s, n_mesh + S, 0.0, 0.0);
  if (k != ne) {
    g_graph_err = "edge count mismatch";
    return FVDBM_ERR_INVALID;
  }
  // rows by source (counting sort), each row by target
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t e = 0; e < ne; ++e) ++off[src[e] + 1];
  for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
  std::vector<int64_t> fill(off.begin(), off.end() - 1);
  std::vector<int32_t> order(ne);
  for (int64_t e = 0; e < ne; ++e) order[fill[src[e]]++] = (int32_t)e;
  for (int64_t i = 0; i < n; ++i) {
    int32_t *r = order.data() + off[i];
    const int64_t len = off[i + 1] - off[i];
    if (len <= 32) {  // insertion sort (stable) for the short rows
      for (int64_t p = 1; p < len; ++p) {
        const int32_t v = r[p];
        int64_t q = p;
        for (; q > 0 && dst[r[q - 1]] > dst[v]; --q) r[q] = r[q - 1];
        r[q] = v;
      }
    } else {
      std::stable_sort(r, r + len, [&](int32_t a, int32_t b) { return dst[a] < dst[b]; });
    }
  }
  for (int64_t i = 0; i <= n; ++i) row_off[i] = (int32_t)off[i];
  for (int64_t p = 0; p < ne; ++p) {
    const int32_t e = order[p];
    const int64_t ds = off[src[e] + 1] - off[src[e]], dd = off[dst[e] + 1] - off[dst[e]];
    const int32_t r = e ^ 1;  // the reverse edge
    col[p] = dst[e];
    w[p] = (float)(1.0 / (double)ds);
    ex[p] = (float)dx[e];
    ey[p] = (float)dy[e];
    wt[p] = (float)(1.0 / (double)dd);
    ext[p] = (float)dx[r];
    eyt[p] = (float)dy[r];
  }
  return FVDBM_OK;
}

}  // extern "C"
