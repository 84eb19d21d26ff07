// batch.hip — batched form: many small graphs in one persistent launch.
#include <hip/hip_runtime.h>

#include "device_graph.hpp"
#include "graph.hpp"

using namespace gmbe;

extern "C" {

int gmbe_batch_create(const gmbe_graph* const* graphs, int n, int bound_height, int bound_size,
                      gmbe_batch** out) {
  (void)graphs; (void)n; (void)bound_height; (void)bound_size;
  if (out) *out = nullptr;
  return set_error(GMBE_ERR_UNSUPPORTED, "batched form not built yet");
}
int gmbe_batch_set(gmbe_batch* b, const char* key, int64_t value) {
  (void)b; (void)key; (void)value;
  return set_error(GMBE_ERR_UNSUPPORTED, "batched form not built yet");
}
int gmbe_batch_execute(gmbe_batch* b, gmbe_result* per_graph, gmbe_result* total) {
  (void)b; (void)per_graph; (void)total;
  return set_error(GMBE_ERR_UNSUPPORTED, "batched form not built yet");
}
void gmbe_batch_destroy(gmbe_batch* b) { (void)b; }

int gmbe_serve_create(const gmbe_graph* resident, int slots, uint64_t mem_budget,
                      int bound_height, int bound_size, gmbe_server** out) {
  (void)resident; (void)slots; (void)mem_budget; (void)bound_height; (void)bound_size;
  if (out) *out = nullptr;
  return set_error(GMBE_ERR_UNSUPPORTED, "serving loop not built yet");
}
int gmbe_serve_submit_graph(gmbe_server* s, const gmbe_graph* g, uint64_t* ticket) {
  (void)s; (void)g; (void)ticket;
  return set_error(GMBE_ERR_UNSUPPORTED, "serving loop not built yet");
}
int gmbe_serve_submit_ego(gmbe_server* s, uint32_t u, uint64_t max_edges, uint64_t* ticket) {
  (void)s; (void)u; (void)max_edges; (void)ticket;
  return set_error(GMBE_ERR_UNSUPPORTED, "serving loop not built yet");
}
int gmbe_serve_poll(gmbe_server* s, gmbe_serve_done* out, int max, int wait, int* n_out) {
  (void)s; (void)out; (void)max; (void)wait;
  if (n_out) *n_out = 0;
  return set_error(GMBE_ERR_UNSUPPORTED, "serving loop not built yet");
}
void gmbe_serve_destroy(gmbe_server* s) { (void)s; }

}  // extern "C"
