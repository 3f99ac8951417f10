// gala_main -- drop-in for the reference CLI (LinXi-lx/GALA-mg src/gala_main.cu:8-57):
//   gala_main -f <graph.bin> [-w] [-t f32] [-o communities.txt] [-d device]
// Prints the reference's per-iteration / per-round lines (README.md:57-89):
//   vertex number / load success / round headers / Iteration:i Q:... / timers /
//   number of communities / final line / execution time / elapsed time.
// min_modularity is fixed at 1e-6 as in gala_main.cu:48.  -w selects int32
// weights (the reference's flag, which its loader ignores, graph.cpp:15,200);
// -t f32 selects float32 weights with FP64 accumulation; -o writes the composed
// final community of every input vertex, one per line.
#include <sys/time.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gala.h"

static double now_ms() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec * 1e3 + tv.tv_usec * 1e-3;
}

struct Printer {
  gala_ctx* ctx;
  int last_round = -1;
};

static void on_iter(void* user, int round, int iteration, double q, int64_t, const int32_t*, int32_t) {
  Printer* p = (Printer*)user;
  if (round != p->last_round) {
    std::printf("===============round:%d===============\n", round);
    p->last_round = round;
  }
  std::printf("Iteration:%d Q:%f\n", iteration, q);
}

static void on_round(void* user, int round, double q, int32_t, const int32_t*, int32_t C, const int64_t*,
                     const int32_t*, const int64_t*, const double*) {
  Printer* p = (Printer*)user;
  gala_result r;
  gala_partial_result(p->ctx, &r);
  double t = r.round_ms[round], dm = r.round_decide_ms[round], up = r.round_update_ms[round];
  std::printf("time without data init = %fms\n", t);
  std::printf("decideandmove time = %fms weight updating time = %fms remaining time = %fms\n", dm, up,
              t - dm - up);
  if (round == 0) {
    std::printf("louvain time in the first round = %fms\n", t);
    std::printf("build time in the first round = %fms\n", r.round_build_ms[0]);
  }
  std::printf("number of communities:%d modularity:%f\n", C, q);
}

int main(int argc, char** argv) {
  double t0 = now_ms();
  std::string file, out_file;
  int wmode = GALA_W_NONE, device = 0;
  int opt;
  while ((opt = getopt(argc, argv, "f:wt:o:d:")) != -1) {
    switch (opt) {
      case 'f': file = optarg; break;
      case 'w': wmode = GALA_W_I32; break;
      case 't':
        if (std::strcmp(optarg, "f32") == 0) wmode = GALA_W_F32;
        else if (std::strcmp(optarg, "i32") == 0) wmode = GALA_W_I32;
        else { std::fprintf(stderr, "unknown weight type %s\n", optarg); return 1; }
        break;
      case 'o': out_file = optarg; break;
      case 'd': device = std::atoi(optarg); break;
      default:
        std::fprintf(stderr, "Usage: %s -f <graph.bin> [-w] [-t f32|i32] [-o comm.txt] [-d dev]\n", argv[0]);
        return 1;
    }
  }
  if (file.empty()) {
    std::fprintf(stderr, "Usage: %s -f <graph.bin> [-w] [-t f32|i32] [-o comm.txt] [-d dev]\n", argv[0]);
    return 1;
  }
  int32_t n = 0;
  int64_t m = 0;
  int rc = gala_read_bin(file.c_str(), wmode, &n, &m, nullptr, nullptr, nullptr);
  if (rc) { std::fprintf(stderr, "cannot read %s: %s\n", file.c_str(), gala_strerror(rc)); return 2; }
  std::printf("vertex number:%d edge number:%lld\n", n, (long long)(m / 2));
  gala_opts o;
  gala_default_opts(&o);
  gala_ctx* ctx = nullptr;
  if ((rc = gala_create(&ctx, device, &o))) { std::fprintf(stderr, "gala_create: %s\n", gala_strerror(rc)); return 3; }
  if ((rc = gala_load_bin(ctx, file.c_str(), wmode))) {
    std::fprintf(stderr, "load: %s\n", gala_strerror(rc));
    gala_destroy(ctx);
    return 2;
  }
  std::printf("load success\n");
  Printer p{ctx};
  gala_result r;
  rc = gala_run(ctx, on_iter, on_round, &p, &r);
  if (rc) {
    std::fprintf(stderr, "run: %s\n", gala_strerror(rc));
    gala_destroy(ctx);
    return 4;
  }
  std::printf("=====================================\n");
  std::printf("final number of communities:%d -> %d final modularity:%f\n", n, r.n_communities, r.modularity);
  std::printf("execution time without data transfer = %fms\n", r.t_total_ms);
  if (!out_file.empty()) {
    std::vector<int32_t> comm(n);
    gala_get_communities(ctx, comm.data());
    FILE* f = std::fopen(out_file.c_str(), "w");
    if (!f) { std::fprintf(stderr, "cannot write %s\n", out_file.c_str()); gala_destroy(ctx); return 2; }
    for (int32_t v = 0; v < n; v++) std::fprintf(f, "%d\n", comm[v]);
    std::fclose(f);
  }
  gala_destroy(ctx);
  std::printf("elapsed time = %fms\n", now_ms() - t0);
  return 0;
}
