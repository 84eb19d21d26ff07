// main.cpp — mbe_gpu, the command line of the reference (src/main.cu:76-149):
//   -i graph.adj  -s sel  -t uvtrans  -o order  -f (fast mode)  -x ngpus  -p
//   -m bound_height  -n bound_size
// with the same stdout lines.  -m / -n are parsed here (the reference lists
// them in its README but leaves them out of the getopt string, Q1).
#include <signal.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gmbe.hpp"

static void FinderTest(BicliqueFinder* finder = nullptr, char* fn = nullptr) {
  if (finder == nullptr) return;
  finder->Execute();
  finder->PrintResult(fn);
  delete finder;
}

static void on_term(int) {
  printf("Total processing time is more than 7200s\n");
  fflush(stdout);
  _exit(0);
}

static double now_s() {
  using namespace std::chrono;
  return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char* argv[]) {
  signal(SIGTERM, on_term);
  int ngpus = 1, op, sel = 0, bound_height = 20, bound_size = 1500;
  Option opt;
  char graph_name[4096] = "db/Writers.adj";
  while ((op = getopt(argc, argv, "h:x:i:s:t:o:fpm:n:")) != -1) {
    switch (op) {
      case 'i':
        snprintf(graph_name, sizeof(graph_name), "%s", optarg);
        break;
      case 't': opt.uvtrans = atoi(optarg); break;
      case 's': sel = atoi(optarg); break;
      case 'o': opt.order = atoi(optarg); break;
      case 'f': opt.fast_mode = 1; break;
      case 'm': bound_height = atoi(optarg); break;
      case 'n': bound_size = atoi(optarg); break;
      case 'x': ngpus = atoi(optarg); break;
      case 'p': print_exit_times() = true; break;  // per-block exit times after the run
      default: break;
    }
  }
  double start_t = now_s();
  CSRBiGraph* graph = new CSRBiGraph();
  graph->Read(graph_name, opt);
  char dev[256] = "";
  int cus = 0;
  gmbe_device_info(dev, sizeof(dev), &cus);
  printf("Max blocks:%u Warps per block:%u Max1d:0x%x Max2d:0x%x\n", (unsigned)cus, 4u, 0x400u,
         (unsigned)graph->V_size);
  printf("trans:%d sel:%d order:%d fast_mode:%d\n", opt.uvtrans, sel, opt.order, opt.fast_mode);
  printf("Graph load time: %lf s\n", now_s() - start_t);
  switch (sel) {
    case 0: FinderTest(new IterFinderGpu(graph)); break;
    case 1: FinderTest(new IterFinderGpu2(graph)); break;
    case 2: FinderTest(new IterFinderGpu6(graph, bound_height, bound_size)); break;
    case 3:
      if (ngpus != 1) {
        fprintf(stderr, "GPUassert: -s 3 -x %d: this build runs on one GPU\n", ngpus);
        return 8;
      }
      FinderTest(new IterFinderGpu6(graph, bound_height, bound_size));
      break;
    default: break;
  }
  delete graph;
  return 0;
}
