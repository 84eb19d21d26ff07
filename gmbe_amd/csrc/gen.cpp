// gen.cpp — seeded synthetic bipartite workloads.
//
// The reference downloads real graphs (preprocess/prepare_dataset.sh:6-37);
// with no network the benchmark graphs are generated here instead.  All
// randomness comes from xoshiro256** (Blackman & Vigna 2018) seeded through
// splitmix64, so a (params, seed) pair always yields the same graph on every
// platform.  Chung-Lu endpoints are drawn proportionally to
// w_i ~ (i+1)^(-1/(gamma-1)), scaled so that each side sums to m and capped at
// `cap`; every vertex first receives one edge to a weight-proportional
// partner so that no vertex is isolated (isolated enumeration-side vertices
// are undefined behaviour in the reference, src/IterFinderGpu.cu:682-697).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "graph.hpp"

namespace gmbe {
namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint32_t below(uint32_t n) {  // uniform in [0, n), n > 0 (Lemire)
    return (uint32_t)(((unsigned __int128)next() * n) >> 64);
  }
};

// Open-addressing set of (u, v) pairs.
struct EdgeSet {
  std::vector<uint64_t> slot;
  uint64_t mask, n = 0;
  explicit EdgeSet(uint64_t expected) {
    uint64_t cap = 16;
    while (cap < expected * 2 + 16) cap <<= 1;
    slot.assign(cap, ~0ull);
    mask = cap - 1;
  }
  static uint64_t h(uint64_t k) { return mix64(k); }
  bool insert(uint64_t k) {
    uint64_t i = h(k) & mask;
    while (slot[i] != ~0ull) {
      if (slot[i] == k) return false;
      i = (i + 1) & mask;
    }
    slot[i] = k;
    ++n;
    return true;
  }
  bool contains(uint64_t k) const {
    uint64_t i = h(k) & mask;
    while (slot[i] != ~0ull) {
      if (slot[i] == k) return true;
      i = (i + 1) & mask;
    }
    return false;
  }
  bool erase(uint64_t k) {  // backward-shift deletion
    uint64_t i = h(k) & mask;
    while (slot[i] != k) {
      if (slot[i] == ~0ull) return false;
      i = (i + 1) & mask;
    }
    uint64_t j = i;
    for (;;) {
      j = (j + 1) & mask;
      if (slot[j] == ~0ull) break;
      uint64_t home = h(slot[j]) & mask;
      // move slot[j] into the hole at i if its home is not in (i, j]
      bool in_range = (i <= j) ? (home > i && home <= j) : (home > i || home <= j);
      if (!in_range) {
        slot[i] = slot[j];
        i = j;
      }
    }
    slot[i] = ~0ull;
    --n;
    return true;
  }
};

struct Sampler {  // draws index i with probability w_i / sum(w)
  std::vector<double> cum;
  explicit Sampler(const std::vector<double>& w) : cum(w.size()) {
    double s = 0;
    for (size_t i = 0; i < w.size(); ++i) cum[i] = (s += w[i]);
  }
  uint32_t draw(Rng& r) const {
    double x = r.uniform() * cum.back();
    size_t i = std::upper_bound(cum.begin(), cum.end(), x) - cum.begin();
    return (uint32_t)std::min(i, cum.size() - 1);
  }
};

std::vector<double> powerlaw_weights(uint32_t n, double gamma, double total, double cap) {
  std::vector<double> w(n);
  double ex = -1.0 / (gamma - 1.0), s = 0;
  for (uint32_t i = 0; i < n; ++i) s += (w[i] = std::pow((double)i + 1.0, ex));
  double scale = total / s;
  for (auto& x : w) x = std::min(x * scale, cap);
  return w;
}

struct Builder {
  uint32_t nU, nV, cap;
  EdgeSet set;
  std::vector<uint32_t> degU, degV;
  std::vector<uint64_t> edges;  // (u << 32) | v, insertion order
  Builder(uint32_t nU_, uint32_t nV_, uint64_t m, uint32_t cap_)
      : nU(nU_), nV(nV_), cap(cap_), set(m), degU(nU_, 0), degV(nV_, 0) {
    edges.reserve(m);
  }
  bool add(uint32_t u, uint32_t v) {
    if (degU[u] >= cap || degV[v] >= cap) return false;
    uint64_t k = ((uint64_t)u << 32) | v;
    if (!set.insert(k)) return false;
    edges.push_back(k);
    degU[u]++;
    degV[v]++;
    return true;
  }
};

// Draw partners until one is accepted (bounded attempts; falls back to a
// uniform partner, then gives up).
template <class DrawU, class DrawV>
int fill_min_degree(Builder& b, Rng& r, DrawU du, DrawV dv) {
  for (uint32_t v = 0; v < b.nV; ++v) {
    if (b.degV[v]) continue;
    bool ok = false;
    for (int t = 0; t < 1000 && !ok; ++t) ok = b.add(du(r), v);
    for (int t = 0; t < 100000 && !ok; ++t) ok = b.add(r.below(b.nU), v);
    if (!ok) return set_error(GMBE_ERR_ARG, "generator: cannot give V vertex %u an edge", v);
  }
  for (uint32_t u = 0; u < b.nU; ++u) {
    if (b.degU[u]) continue;
    bool ok = false;
    for (int t = 0; t < 1000 && !ok; ++t) ok = b.add(u, dv(r));
    for (int t = 0; t < 100000 && !ok; ++t) ok = b.add(u, r.below(b.nV));
    if (!ok) return set_error(GMBE_ERR_ARG, "generator: cannot give U vertex %u an edge", u);
  }
  return GMBE_OK;
}

int finish(Builder& b, const std::vector<uint64_t>& edges, gmbe_graph** out) {
  // file orientation: rows = V, items = U
  std::vector<int64_t> off((size_t)b.nV + 1, 0);
  for (uint64_t k : edges) off[(k & 0xffffffffu) + 1]++;
  for (uint32_t v = 0; v < b.nV; ++v) off[v + 1] += off[v];
  std::vector<uint32_t> adj(edges.size());
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  for (uint64_t k : edges) adj[pos[k & 0xffffffffu]++] = (uint32_t)(k >> 32);
  gmbe_option o{0, 0, 1};
  return gmbe_graph_from_csr(b.nU, b.nV, off.data(), adj.data(), &o, out);
}

int generate(const gmbe_gen_params& p, gmbe_graph** out) {
  if (p.nU == 0 || p.nV == 0) return set_error(GMBE_ERR_ARG, "generator: empty side");
  uint64_t full = (uint64_t)p.nU * p.nV;
  if (p.m > full / 2 + 1 && p.kind != 1)
    return set_error(GMBE_ERR_ARG, "generator: m=%llu too dense for Chung-Lu",
                     (unsigned long long)p.m);
  if (p.m > full) return set_error(GMBE_ERR_ARG, "generator: m exceeds |U||V|");
  if (p.m < std::max(p.nU, p.nV))
    return set_error(GMBE_ERR_ARG, "generator: m < max(|U|,|V|) leaves isolated vertices");
  Rng r(p.seed);
  uint32_t cap = p.cap ? p.cap : UINT32_MAX;
  Builder b(p.nU, p.nV, p.m + (uint64_t)p.blocks * p.block_max * p.block_max, cap);
  if (p.kind == 1) {
    int rc = fill_min_degree(b, r, [&](Rng& rr) { return rr.below(p.nU); },
                             [&](Rng& rr) { return rr.below(p.nV); });
    if (rc) return rc;
    uint64_t attempts = 0;
    while (b.edges.size() < p.m) {
      b.add(r.below(p.nU), r.below(p.nV));
      if (++attempts > p.m * 1000) return set_error(GMBE_ERR_ARG, "generator: stuck");
    }
  } else if (p.kind == 2 || p.kind == 3) {
    Sampler su(powerlaw_weights(p.nU, p.gamma_u, (double)p.m, (double)cap));
    Sampler sv(powerlaw_weights(p.nV, p.gamma_v, (double)p.m, (double)cap));
    int rc = fill_min_degree(b, r, [&](Rng& rr) { return su.draw(rr); },
                             [&](Rng& rr) { return sv.draw(rr); });
    if (rc) return rc;
    uint64_t attempts = 0;
    while (b.edges.size() < p.m) {
      uint32_t u = su.draw(r), v = sv.draw(r);
      b.add(u, v);
      if (++attempts > p.m * 200)
        return set_error(GMBE_ERR_ARG, "generator: cannot reach %llu distinct edges",
                         (unsigned long long)p.m);
    }
  } else {
    return set_error(GMBE_ERR_ARG, "generator: unknown kind %d", p.kind);
  }
  std::vector<uint64_t> edges = b.edges;
  if (p.kind == 3 && p.blocks) {
    if (p.block_min == 0 || p.block_max < p.block_min || p.block_max > p.nU ||
        p.block_max > p.nV)
      return set_error(GMBE_ERR_ARG, "generator: bad block size range");
    // planted blocks: the cap does not apply to planted edges
    std::vector<uint64_t> removed;
    std::vector<uint32_t> us, vs;
    for (uint32_t k = 0; k < p.blocks; ++k) {
      uint32_t span = p.block_max - p.block_min + 1;
      uint32_t a = p.block_min + r.below(span), bb = p.block_min + r.below(span);
      auto pick = [&](uint32_t n, uint32_t cnt, std::vector<uint32_t>& outv) {
        outv.clear();
        while (outv.size() < cnt) {
          uint32_t x = r.below(n);
          if (std::find(outv.begin(), outv.end(), x) == outv.end()) outv.push_back(x);
        }
      };
      pick(p.nU, a, us);
      pick(p.nV, bb, vs);
      std::vector<uint64_t> block;
      for (uint32_t u : us)
        for (uint32_t v : vs) {
          uint64_t key = ((uint64_t)u << 32) | v;
          block.push_back(key);
          if (b.set.insert(key)) {
            edges.push_back(key);
            b.degU[u]++;
            b.degV[v]++;
          }
        }
      uint64_t drop = (uint64_t)std::llround(p.drop * (double)a * (double)bb);
      // partial Fisher-Yates over the block's pairs
      for (uint64_t i = 0; i < drop && i < block.size(); ++i) {
        uint64_t j = i + r.below((uint32_t)(block.size() - i));
        std::swap(block[i], block[j]);
        uint64_t key = block[i];
        uint32_t u = (uint32_t)(key >> 32), v = (uint32_t)(key & 0xffffffffu);
        // never isolate a vertex
        if (b.set.contains(key) && b.degU[u] > 1 && b.degV[v] > 1) {
          b.set.erase(key);
          b.degU[u]--;
          b.degV[v]--;
          removed.push_back(key);
        }
      }
    }
    if (!removed.empty()) {
      std::vector<uint64_t> kept;
      kept.reserve(edges.size());
      for (uint64_t k : edges)
        if (b.set.contains(k)) kept.push_back(k);
      edges.swap(kept);
    }
  }
  return finish(b, edges, out);
}

}  // namespace
}  // namespace gmbe

using namespace gmbe;

extern "C" {

static int gmbe_generate_impl(const gmbe_gen_params* p, gmbe_graph** out) {
  if (!p || !out) return set_error(GMBE_ERR_ARG, "null argument");
  *out = nullptr;
  return generate(*p, out);
}

int gmbe_generate(const gmbe_gen_params* p, gmbe_graph** out) {
  return gmbe::guarded("gmbe_generate", [&] { return gmbe_generate_impl(p, out); });
}

static int gmbe_generate_small_impl(uint64_t seed, uint32_t lo, uint32_t hi, double dlo, double dhi,
                                    double gamma, gmbe_graph** out) {
  if (!out || lo == 0 || hi < lo || dlo <= 0 || dhi < dlo)
    return set_error(GMBE_ERR_ARG, "bad small-graph parameters");
  Rng r(seed);
  double l0 = std::log((double)lo), l1 = std::log((double)hi);
  uint32_t nU = (uint32_t)std::llround(std::exp(l0 + (l1 - l0) * r.uniform()));
  uint32_t nV = (uint32_t)std::llround(std::exp(l0 + (l1 - l0) * r.uniform()));
  double d = dlo + (dhi - dlo) * r.uniform();
  uint64_t m = (uint64_t)std::llround(d * nU);
  uint64_t full = (uint64_t)nU * nV;
  m = std::max<uint64_t>(m, std::max(nU, nV));
  m = std::min<uint64_t>(m, full / 2);
  m = std::max<uint64_t>(m, std::max(nU, nV));
  gmbe_gen_params p{};
  p.kind = 2;
  p.nU = nU;
  p.nV = nV;
  p.m = m;
  p.gamma_u = p.gamma_v = gamma;
  p.cap = 0;
  p.seed = r.next();
  return gmbe_generate(&p, out);
}

int gmbe_generate_small(uint64_t seed, uint32_t lo, uint32_t hi, double dlo, double dhi,
                        double gamma, gmbe_graph** out) {
  return gmbe::guarded("gmbe_generate_small", [&] {
    return gmbe_generate_small_impl(seed, lo, hi, dlo, dhi, gamma, out);
  });
}

static int gmbe_ego_subgraph_impl(const gmbe_graph* g, uint32_t u, uint64_t max_edges,
                                  gmbe_graph** out) {
  // g is in file orientation: rows = V (original V ids), items = U.
  if (!g || !out) return set_error(GMBE_ERR_ARG, "null argument");
  if (g->swapped) return set_error(GMBE_ERR_ARG, "ego: graph must be in file orientation");
  if (u >= g->nU) return set_error(GMBE_ERR_ARG, "ego: vertex %u out of range", u);
  *out = nullptr;
  std::vector<uint32_t> vs(g->u_adj.begin() + g->u_off[u], g->u_adj.begin() + g->u_off[u + 1]);
  // V' in ascending id, whole rows until the edge cap
  std::vector<uint32_t> keepV;
  uint64_t e = 0;
  for (uint32_t v : vs) {
    uint64_t d = (uint64_t)(g->v_off[v + 1] - g->v_off[v]);
    if (e + d > max_edges) break;
    e += d;
    keepV.push_back(v);
  }
  std::vector<uint32_t> Us;
  for (uint32_t v : keepV)
    Us.insert(Us.end(), g->v_adj.begin() + g->v_off[v], g->v_adj.begin() + g->v_off[v + 1]);
  std::sort(Us.begin(), Us.end());
  Us.erase(std::unique(Us.begin(), Us.end()), Us.end());
  std::vector<int64_t> off{0};
  std::vector<uint32_t> adj;
  adj.reserve(e);
  for (uint32_t v : keepV) {
    for (int64_t k = g->v_off[v]; k < g->v_off[v + 1]; ++k)
      adj.push_back((uint32_t)(std::lower_bound(Us.begin(), Us.end(), g->v_adj[k]) - Us.begin()));
    off.push_back((int64_t)adj.size());
  }
  gmbe_option o{0, 0, 1};
  int rc = gmbe_graph_from_csr((uint32_t)Us.size(), (uint32_t)keepV.size(), off.data(),
                               adj.data(), &o, out);
  if (rc == GMBE_OK) {
    // keep the ids of the resident graph as original ids
    for (size_t i = 0; i < Us.size(); ++i) (*out)->u_orig[i] = g->u_orig[Us[i]];
    for (size_t i = 0; i < keepV.size(); ++i) (*out)->v_orig[i] = g->v_orig[keepV[i]];
  }
  return rc;
}

int gmbe_ego_subgraph(const gmbe_graph* g, uint32_t u, uint64_t max_edges, gmbe_graph** out) {
  return gmbe::guarded("gmbe_ego_subgraph", [&] {
    return gmbe_ego_subgraph_impl(g, u, max_edges, out);
  });
}

}  // extern "C"
