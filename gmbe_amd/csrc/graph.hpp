// graph.hpp — host-side bipartite graph of the engine (internal C++ header).
//
// Internal orientation follows the reference's CSRBiGraph (src/BiGraph.h:23-38):
// V is the enumerated side (rows of the .adj file after the optional side
// swap), U the other one.  Ids are uint32 and offsets int64 (the reference
// uses int32 for both).  Unlike the reference, the internal->original id map
// of both sides is kept so that results can be stated in original ids.
#pragma once

#include <cstdint>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <system_error>
#include <vector>

#include "../../include/gmbe.h"

struct gmbe_graph {
  uint32_t nU = 0, nV = 0;
  uint64_t E = 0;
  std::vector<int64_t> v_off;   // nV + 1   (row_offset)
  std::vector<uint32_t> v_adj;  // E        (column_indices), ascending per row
  std::vector<int64_t> u_off;   // nU + 1   (rev_row_offset)
  std::vector<uint32_t> u_adj;  // E        (rev_column_indices), ascending
  std::vector<uint32_t> u_orig; // internal U id -> original id
  std::vector<uint32_t> v_orig; // internal V id -> original id
  int swapped = 0;              // 1: internal U is the original V side
  int64_t dU = 0, dV = 0, d2U = -1, d2V = -1;
  gmbe_option opt{0, 1, 0};
  std::string name;
};

namespace gmbe {

// Thread-local last error (gmbe_last_error).
int set_error(int code, const char* fmt, ...);

// Runs the body of a C entry point.  A C++ exception (a host allocation, a
// thread that cannot start) does not unwind into the C caller: it becomes an
// error code, with its message behind gmbe_last_error().
template <class F>
int guarded(const char* fn, F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return set_error(GMBE_ERR_NOMEM, "%s: host allocation failed", fn);
  } catch (const std::length_error& e) {
    return set_error(GMBE_ERR_CAPACITY, "%s: %s", fn, e.what());
  } catch (const std::system_error& e) {
    return set_error(GMBE_ERR_NOMEM, "%s: %s", fn, e.what());
  } catch (const std::exception& e) {
    return set_error(GMBE_ERR_INTERNAL, "%s: %s", fn, e.what());
  } catch (...) {
    return set_error(GMBE_ERR_INTERNAL, "%s: unknown exception", fn);
  }
}

// Build a working graph from file-orientation rows.  rows[v] lists U ids.
// v_orig_of_row / u_orig give the original ids of the input rows / U ids
// (identity for a fresh file).  in_swapped tells whether the input rows are
// already the original U side (when deriving from a derived graph).
int build_graph(uint32_t nU, uint32_t nV, const int64_t* off, const uint32_t* adj,
                const std::vector<uint32_t>& u_orig_in,
                const std::vector<uint32_t>& v_orig_in, int in_swapped,
                const gmbe_option& opt, gmbe_graph* g);

// Per-vertex hash keys of the internal sides, already folded with the
// original side (see gmbe_hash_biclique).
void hash_keys(const gmbe_graph& g, std::vector<uint64_t>& kU,
               std::vector<uint64_t>& kV);

static inline uint64_t mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t rotl64(uint64_t x, int r) {
  return (x << r) | (x >> (64 - r));
}
static inline uint64_t side_key(int side, uint32_t id) {
  return mix64(((uint64_t)id << 1) | (uint64_t)side);
}
// Combine the internal-L and internal-R key sums into the biclique hash.
static inline uint64_t biclique_hash(uint64_t sumL, uint64_t sumR, int swapped) {
  // A = original-U side, B = original-V side
  uint64_t SA = swapped ? sumR : sumL;
  uint64_t SB = swapped ? sumL : sumR;
  return mix64(SA ^ rotl64(SB, 31));
}

double now_s();

}  // namespace gmbe
