// graph.cpp — .adj loading, side swap, vertex ordering and CSR build.
//
// Semantics follow CSRBiGraph::Read (src/BiGraph.cu:13-278):
//   * line i of the file is V vertex i; its tokens are U ids; |U| = max id + 1
//     (src/BiGraph.cu:27-37);
//   * uvtrans 1 swaps the sides when |U| < |V|, 2 when |U| > |V|
//     (src/BiGraph.cu:66-73, 217-227);
//   * order 1 / 2 sorts V rows by (degree asc / desc, then lexicographic row)
//     (src/BiGraph.cu:74-92, 228-239); in fast mode every order != 0 other
//     than 1 sorts descending, as the reference's fast path does
//     (src/BiGraph.cu:236-237); order 3 outside fast mode is the unilateral
//     (2-hop degeneracy) order (src/BiGraph.cu:93-210);
//   * U lists are rebuilt from the reordered rows, so every list is ascending
//     (src/BiGraph.cu:240-246);
//   * the 2-hop statistics are computed only outside fast mode
//     (src/BiGraph.cu:39-65, 212-215).
// Deliberate differences: rows are sorted and de-duplicated (the reference
// assumes sorted, duplicate-free lines from preprocess/data_convert.cpp:20-42),
// a missing file is an error instead of an empty graph (Q10), and original ids
// of both sides are kept for hashing.
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "graph.hpp"

namespace gmbe {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

double now_s() {
  using namespace std::chrono;
  return duration_cast<duration<double>>(steady_clock::now().time_since_epoch())
      .count();
}

// Transpose rows (nRows lists over ids < nCols) into column lists.  Columns
// come out ascending because rows are visited in ascending order.
static void transpose(uint32_t nRows, uint32_t nCols, const std::vector<int64_t>& off,
                      const std::vector<uint32_t>& adj, std::vector<int64_t>& toff,
                      std::vector<uint32_t>& tadj) {
  toff.assign((size_t)nCols + 1, 0);
  for (uint64_t e = 0; e < adj.size(); ++e) toff[(size_t)adj[e] + 1]++;
  for (uint32_t c = 0; c < nCols; ++c) toff[c + 1] += toff[c];
  tadj.resize(adj.size());
  std::vector<int64_t> pos(toff.begin(), toff.end() - 1);
  for (uint32_t r = 0; r < nRows; ++r)
    for (int64_t e = off[r]; e < off[r + 1]; ++e) tadj[pos[adj[e]]++] = r;
}

// max over rows r of |union of cols(c) over c in row r| - 1, i.e. the number
// of distinct 2-hop neighbours including r itself minus one
// (src/BiGraph.cu:47-65).
static int64_t max_two_hop(uint32_t nRows, const std::vector<int64_t>& off,
                           const std::vector<uint32_t>& adj,
                           const std::vector<int64_t>& toff,
                           const std::vector<uint32_t>& tadj) {
  std::vector<uint32_t> mark(nRows, UINT32_MAX);
  int64_t best = 0;
  for (uint32_t r = 0; r < nRows; ++r) {
    int64_t n = 0;
    for (int64_t e = off[r]; e < off[r + 1]; ++e) {
      uint32_t c = adj[e];
      for (int64_t f = toff[c]; f < toff[c + 1]; ++f) {
        uint32_t r2 = tadj[f];
        if (mark[r2] != r) {
          mark[r2] = r;
          ++n;
        }
      }
    }
    best = std::max(best, n - 1);
  }
  return best;
}

// Unilateral order: repeatedly take the vertex with the fewest remaining
// 2-hop neighbours (bucket queue, Batagelj-Zaversnik style), the procedure of
// src/BiGraph.cu:93-203.  Returns perm[new] = old.
static std::vector<uint32_t> unilateral_order(uint32_t nV, const std::vector<int64_t>& off,
                                              const std::vector<uint32_t>& adj,
                                              const std::vector<int64_t>& toff,
                                              const std::vector<uint32_t>& tadj) {
  std::vector<int64_t> key(nV, 0);
  std::vector<uint32_t> mark(nV, UINT32_MAX);
  for (uint32_t v = 0; v < nV; ++v) {
    int64_t n = 0;
    for (int64_t e = off[v]; e < off[v + 1]; ++e)
      for (int64_t f = toff[adj[e]]; f < toff[adj[e] + 1]; ++f) {
        uint32_t w = tadj[f];
        if (w != v && mark[w] != v) {
          mark[w] = v;
          ++n;
        }
      }
    key[v] = n;
  }
  int64_t kmax = 0;
  for (auto k : key) kmax = std::max(kmax, k);
  std::vector<int64_t> bucket_start(kmax + 2, 0);
  for (auto k : key) bucket_start[k + 1]++;
  for (int64_t k = 0; k <= kmax; ++k) bucket_start[k + 1] += bucket_start[k];
  std::vector<uint32_t> order(nV), where(nV);
  {
    std::vector<int64_t> fill(bucket_start.begin(), bucket_start.end() - 1);
    for (uint32_t v = 0; v < nV; ++v) {
      where[v] = (uint32_t)fill[key[v]];
      order[where[v]] = v;
      fill[key[v]]++;
    }
  }
  std::fill(mark.begin(), mark.end(), UINT32_MAX);
  for (uint32_t i = 0; i < nV; ++i) {
    uint32_t v = order[i];
    for (int64_t e = off[v]; e < off[v + 1]; ++e)
      for (int64_t f = toff[adj[e]]; f < toff[adj[e] + 1]; ++f) {
        uint32_t w = tadj[f];
        if (mark[w] == v) continue;
        mark[w] = v;
        if (key[w] > key[v]) {
          // move w to the front of its bucket, then shrink its key by one
          int64_t kw = key[w];
          uint32_t front = (uint32_t)bucket_start[kw];
          uint32_t other = order[front];
          if (other != w) {
            order[where[w]] = other;
            where[other] = where[w];
            order[front] = w;
            where[w] = front;
          }
          bucket_start[kw]++;
          key[w]--;
        }
      }
  }
  return order;
}

int build_graph(uint32_t nU, uint32_t nV, const int64_t* off, const uint32_t* adj,
                const std::vector<uint32_t>& u_orig_in,
                const std::vector<uint32_t>& v_orig_in, int in_swapped,
                const gmbe_option& opt, gmbe_graph* g) {
  if (opt.uvtrans < 0 || opt.uvtrans > 2 || opt.order < 0 || opt.order > 3)
    return set_error(GMBE_ERR_ARG, "invalid option uvtrans=%d order=%d",
                     opt.uvtrans, opt.order);
  // rows: sorted, de-duplicated
  std::vector<int64_t> roff((size_t)nV + 1, 0);
  std::vector<uint32_t> radj;
  radj.reserve(off[nV] - off[0]);
  for (uint32_t v = 0; v < nV; ++v) {
    size_t s = radj.size();
    if (off[v + 1] < off[v]) return set_error(GMBE_ERR_ARG, "row %u: negative length", v);
    for (int64_t e = off[v]; e < off[v + 1]; ++e) {
      if (adj[e] >= nU)
        return set_error(GMBE_ERR_ARG, "row %u: id %u >= |U| %u", v, adj[e], nU);
      radj.push_back(adj[e]);
    }
    std::sort(radj.begin() + s, radj.end());
    radj.erase(std::unique(radj.begin() + s, radj.end()), radj.end());
    roff[v + 1] = (int64_t)radj.size();
  }
  std::vector<int64_t> coff;
  std::vector<uint32_t> cadj;
  transpose(nV, nU, roff, radj, coff, cadj);

  std::vector<uint32_t> u_orig = u_orig_in, v_orig = v_orig_in;
  int64_t dU = 0, dV = 0, d2U = -1, d2V = -1;
  for (uint32_t v = 0; v < nV; ++v) dV = std::max(dV, roff[v + 1] - roff[v]);
  for (uint32_t u = 0; u < nU; ++u) dU = std::max(dU, coff[u + 1] - coff[u]);
  if (!opt.fast_mode) {
    d2U = max_two_hop(nU, coff, cadj, roff, radj);
    d2V = max_two_hop(nV, roff, radj, coff, cadj);
  }
  int swapped = in_swapped;
  if ((opt.uvtrans == 1 && nU < nV) || (opt.uvtrans == 2 && nU > nV)) {
    std::swap(nU, nV);
    std::swap(roff, coff);
    std::swap(radj, cadj);
    std::swap(u_orig, v_orig);
    std::swap(dU, dV);
    std::swap(d2U, d2V);
    swapped ^= 1;
  }
  // perm[new] = old
  std::vector<uint32_t> perm(nV);
  std::iota(perm.begin(), perm.end(), 0u);
  bool sort_rows = opt.fast_mode ? opt.order != 0 : (opt.order == 1 || opt.order == 2);
  if (sort_rows) {
    bool asc = opt.order == 1;
    std::sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
      int64_t la = roff[a + 1] - roff[a], lb = roff[b + 1] - roff[b];
      if (la != lb) return asc ? la < lb : la > lb;
      const uint32_t* pa = radj.data() + roff[a];
      const uint32_t* pb = radj.data() + roff[b];
      for (int64_t i = 0; i < la; ++i)
        if (pa[i] != pb[i]) return pa[i] < pb[i];
      return a < b;  // identical rows: keep input order
    });
  } else if (!opt.fast_mode && opt.order == 3) {
    perm = unilateral_order(nV, roff, radj, coff, cadj);
  }
  g->nU = nU;
  g->nV = nV;
  g->E = radj.size();
  g->v_off.assign((size_t)nV + 1, 0);
  g->v_adj.resize(radj.size());
  g->v_orig.resize(nV);
  for (uint32_t i = 0; i < nV; ++i) {
    uint32_t o = perm[i];
    int64_t len = roff[o + 1] - roff[o];
    std::copy(radj.begin() + roff[o], radj.begin() + roff[o + 1],
              g->v_adj.begin() + g->v_off[i]);
    g->v_off[i + 1] = g->v_off[i] + len;
    g->v_orig[i] = v_orig[o];
  }
  g->u_orig = u_orig;
  transpose(nV, nU, g->v_off, g->v_adj, g->u_off, g->u_adj);
  g->swapped = swapped;
  g->dU = dU;
  g->dV = dV;
  g->d2U = d2U;
  g->d2V = d2V;
  g->opt = opt;
  return GMBE_OK;
}

void hash_keys(const gmbe_graph& g, std::vector<uint64_t>& kU, std::vector<uint64_t>& kV) {
  // internal U is the original U side unless swapped
  int sideU = g.swapped ? 1 : 0, sideV = g.swapped ? 0 : 1;
  kU.resize(g.nU);
  kV.resize(g.nV);
  for (uint32_t u = 0; u < g.nU; ++u) kU[u] = side_key(sideU, g.u_orig[u]);
  for (uint32_t v = 0; v < g.nV; ++v) kV[v] = side_key(sideV, g.v_orig[v]);
}

}  // namespace gmbe

using namespace gmbe;

extern "C" {

void gmbe_option_default(gmbe_option* opt) {
  if (!opt) return;
  opt->uvtrans = 0;
  opt->order = 1;
  opt->fast_mode = 0;
}

static std::vector<uint32_t> iota_u32(uint32_t n) {
  std::vector<uint32_t> r(n);
  std::iota(r.begin(), r.end(), 0u);
  return r;
}

static int gmbe_graph_read_adj_impl(const char* path, const gmbe_option* opt, gmbe_graph** out) {
  if (!path || !out) return set_error(GMBE_ERR_ARG, "null argument");
  *out = nullptr;
  FILE* fp = fopen(path, "rb");
  if (!fp) return set_error(GMBE_ERR_IO, "can't open file %s: %s", path, strerror(errno));
  std::vector<char> buf;
  {
    char chunk[1 << 16];
    size_t n;
    while ((n = fread(chunk, 1, sizeof(chunk), fp)) > 0) buf.insert(buf.end(), chunk, chunk + n);
  }
  fclose(fp);
  std::vector<int64_t> off{0};
  std::vector<uint32_t> adj;
  uint64_t nU = 0;
  size_t i = 0, n = buf.size();
  uint64_t line = 0;
  while (i < n) {
    // one line
    while (i < n && buf[i] != '\n') {
      char c = buf[i];
      if (c >= '0' && c <= '9') {
        uint64_t x = 0;
        while (i < n && buf[i] >= '0' && buf[i] <= '9') {
          x = x * 10 + (uint64_t)(buf[i] - '0');
          if (x > 0xFFFFFFFEull)
            return set_error(GMBE_ERR_FORMAT, "%s:%llu: id out of range", path,
                             (unsigned long long)line + 1);
          ++i;
        }
        adj.push_back((uint32_t)x);
        nU = std::max<uint64_t>(nU, x + 1);
      } else if (c == ' ' || c == '\t' || c == '\r') {
        ++i;
      } else {
        return set_error(GMBE_ERR_FORMAT, "%s:%llu: unexpected character 0x%02x", path,
                         (unsigned long long)line + 1, (unsigned)(unsigned char)c);
      }
    }
    off.push_back((int64_t)adj.size());
    ++line;
    if (i < n) ++i;  // skip '\n'
  }
  if (off.size() - 1 > 0xFFFFFFFEull) return set_error(GMBE_ERR_CAPACITY, "too many lines");
  uint32_t nV = (uint32_t)(off.size() - 1);
  gmbe_option o;
  if (opt) o = *opt; else gmbe_option_default(&o);
  gmbe_graph* g = new gmbe_graph();
  int rc = build_graph((uint32_t)nU, nV, off.data(), adj.data(), iota_u32((uint32_t)nU),
                       iota_u32(nV), 0, o, g);
  if (rc != GMBE_OK) {
    delete g;
    return rc;
  }
  g->name = path;
  *out = g;
  return GMBE_OK;
}

int gmbe_graph_read_adj(const char* path, const gmbe_option* opt, gmbe_graph** out) {
  return gmbe::guarded("gmbe_graph_read_adj", [&] {
    return gmbe_graph_read_adj_impl(path, opt, out);
  });
}

static int gmbe_graph_from_csr_impl(uint32_t nU, uint32_t nV, const int64_t* v_off,
                                    const uint32_t* v_adj, const gmbe_option* opt,
                                    gmbe_graph** out) {
  if (!v_off || (!v_adj && nV && v_off[nV] > v_off[0]) || !out)
    return set_error(GMBE_ERR_ARG, "null argument");
  if (nU == UINT32_MAX || nV == UINT32_MAX)
    return set_error(GMBE_ERR_CAPACITY, "vertex count out of range");
  *out = nullptr;
  gmbe_option o;
  if (opt) o = *opt; else gmbe_option_default(&o);
  gmbe_graph* g = new gmbe_graph();
  int rc = build_graph(nU, nV, v_off, v_adj, iota_u32(nU), iota_u32(nV), 0, o, g);
  if (rc != GMBE_OK) {
    delete g;
    return rc;
  }
  g->name = "<csr>";
  *out = g;
  return GMBE_OK;
}

int gmbe_graph_from_csr(uint32_t nU, uint32_t nV, const int64_t* v_off, const uint32_t* v_adj,
                        const gmbe_option* opt, gmbe_graph** out) {
  return gmbe::guarded("gmbe_graph_from_csr", [&] {
    return gmbe_graph_from_csr_impl(nU, nV, v_off, v_adj, opt, out);
  });
}

static int gmbe_graph_derive_impl(const gmbe_graph* src, const gmbe_option* opt, gmbe_graph** out) {
  if (!src || !out) return set_error(GMBE_ERR_ARG, "null argument");
  *out = nullptr;
  gmbe_option o;
  if (opt) o = *opt; else gmbe_option_default(&o);
  gmbe_graph* g = new gmbe_graph();
  int rc = build_graph(src->nU, src->nV, src->v_off.data(), src->v_adj.data(), src->u_orig,
                       src->v_orig, src->swapped, o, g);
  if (rc != GMBE_OK) {
    delete g;
    return rc;
  }
  g->name = src->name;
  *out = g;
  return GMBE_OK;
}

int gmbe_graph_derive(const gmbe_graph* src, const gmbe_option* opt, gmbe_graph** out) {
  return gmbe::guarded("gmbe_graph_derive", [&] { return gmbe_graph_derive_impl(src, opt, out); });
}

static int gmbe_graph_stats_impl(const gmbe_graph* g, int64_t out[8]) {
  if (!g || !out) return set_error(GMBE_ERR_ARG, "null argument");
  out[0] = g->nU;
  out[1] = g->nV;
  out[2] = (int64_t)g->E;
  out[3] = g->dU;
  out[4] = g->dV;
  out[5] = g->d2U;
  out[6] = g->d2V;
  out[7] = g->swapped;
  return GMBE_OK;
}

int gmbe_graph_stats(const gmbe_graph* g, int64_t out[8]) {
  return gmbe::guarded("gmbe_graph_stats", [&] { return gmbe_graph_stats_impl(g, out); });
}

static int gmbe_graph_csr_impl(const gmbe_graph* g, const int64_t** v_off, const uint32_t** v_adj,
                               const int64_t** u_off, const uint32_t** u_adj, const uint32_t** u_orig,
                               const uint32_t** v_orig) {
  if (!g) return set_error(GMBE_ERR_ARG, "null graph");
  if (v_off) *v_off = g->v_off.data();
  if (v_adj) *v_adj = g->v_adj.data();
  if (u_off) *u_off = g->u_off.data();
  if (u_adj) *u_adj = g->u_adj.data();
  if (u_orig) *u_orig = g->u_orig.data();
  if (v_orig) *v_orig = g->v_orig.data();
  return GMBE_OK;
}

int gmbe_graph_csr(const gmbe_graph* g, const int64_t** v_off, const uint32_t** v_adj,
                   const int64_t** u_off, const uint32_t** u_adj, const uint32_t** u_orig,
                   const uint32_t** v_orig) {
  return gmbe::guarded("gmbe_graph_csr", [&] {
    return gmbe_graph_csr_impl(g, v_off, v_adj, u_off, u_adj, u_orig, v_orig);
  });
}

static int gmbe_graph_write_adj_impl(const gmbe_graph* g, const char* path) {
  if (!g || !path) return set_error(GMBE_ERR_ARG, "null argument");
  FILE* fp = fopen(path, "wb");
  if (!fp) return set_error(GMBE_ERR_IO, "can't open %s for writing", path);
  std::string line;
  for (uint32_t v = 0; v < g->nV; ++v) {
    line.clear();
    for (int64_t e = g->v_off[v]; e < g->v_off[v + 1]; ++e) {
      if (e > g->v_off[v]) line.push_back(' ');
      line += std::to_string(g->v_adj[e]);
    }
    line.push_back('\n');
    if (fwrite(line.data(), 1, line.size(), fp) != line.size()) {
      fclose(fp);
      return set_error(GMBE_ERR_IO, "write failed: %s", path);
    }
  }
  fclose(fp);
  return GMBE_OK;
}

int gmbe_graph_write_adj(const gmbe_graph* g, const char* path) {
  return gmbe::guarded("gmbe_graph_write_adj", [&] { return gmbe_graph_write_adj_impl(g, path); });
}

static int gmbe_graph_write_edges_impl(const gmbe_graph* g, const char* path) {
  if (!g || !path) return set_error(GMBE_ERR_ARG, "null argument");
  // rows are the original V side, columns the original U side, in original ids
  const bool sw = g->swapped != 0;
  const uint32_t nrows = sw ? g->nU : g->nV, ncols = sw ? g->nV : g->nU;
  const int64_t* off = sw ? g->u_off.data() : g->v_off.data();
  const uint32_t* adj = sw ? g->u_adj.data() : g->v_adj.data();
  const std::vector<uint32_t>& row_orig = sw ? g->u_orig : g->v_orig;
  const std::vector<uint32_t>& col_orig = sw ? g->v_orig : g->u_orig;
  std::vector<uint32_t> rows(nrows);
  for (uint32_t i = 0; i < nrows; ++i) rows[i] = i;
  std::sort(rows.begin(), rows.end(),
            [&](uint32_t a, uint32_t b) { return row_orig[a] < row_orig[b]; });
  FILE* fp = fopen(path, "wb");
  if (!fp) return set_error(GMBE_ERR_IO, "can't open %s for writing", path);
  std::string buf = std::to_string(nrows) + " " + std::to_string(ncols) + "\n";
  std::vector<uint32_t> cols;
  bool ok = true;
  for (uint32_t i : rows) {
    cols.clear();
    for (int64_t e = off[i]; e < off[i + 1]; ++e) cols.push_back(col_orig[adj[e]]);
    std::sort(cols.begin(), cols.end());
    const std::string r = std::to_string(row_orig[i]) + " ";
    for (uint32_t c : cols) {
      buf += r;
      buf += std::to_string(c);
      buf.push_back('\n');
    }
    if (buf.size() > (1u << 20)) {
      ok = ok && fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
      buf.clear();
    }
  }
  ok = ok && fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
  fclose(fp);
  if (!ok) return set_error(GMBE_ERR_IO, "write failed: %s", path);
  return GMBE_OK;
}

int gmbe_graph_write_edges(const gmbe_graph* g, const char* path) {
  return gmbe::guarded("gmbe_graph_write_edges", [&] {
    return gmbe_graph_write_edges_impl(g, path);
  });
}

void gmbe_graph_destroy(gmbe_graph* g) { delete g; }

uint64_t gmbe_hash_biclique(const uint32_t* A, size_t nA, const uint32_t* B, size_t nB) {
  uint64_t SA = 0, SB = 0;
  for (size_t i = 0; i < nA; ++i) SA += side_key(0, A[i]);
  for (size_t i = 0; i < nB; ++i) SB += side_key(1, B[i]);
  return mix64(SA ^ rotl64(SB, 31));
}

const char* gmbe_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
