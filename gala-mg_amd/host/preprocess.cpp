// preprocess -- drop-in for the reference format producer
// (LinXi-lx/GALA-mg src/preprocess.cpp:7-47):
//   preprocess -f <edges.txt> [-w] [-o out_dir] [-g gpu]  ->  <edges.txt>.bin  (or out_dir/<name>.bin)
// Same ingest semantics (graph.cpp:18-151) through gala_ingest_text, or with
// -g through gala_ingest_text_device on that GPU (the same output); unlike the
// reference, weighted graphs get every weight written and input self-loops do
// not truncate the neighbour array.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/gala.h"

int main(int argc, char** argv) {
  std::string input, out_dir;
  int wmode = GALA_W_NONE;
  int gpu = -1;
  int opt;
  while ((opt = getopt(argc, argv, "f:wo:g:")) != -1) {
    switch (opt) {
      case 'f': input = optarg; break;
      case 'w': wmode = GALA_W_I32; break;
      case 'o': out_dir = optarg; break;
      case 'g': gpu = std::atoi(optarg); break;
      default:
        std::fprintf(stderr, "Usage: %s -f <filename> [-w -o <output_dirname> -g <gpu>]\n", argv[0]);
        return 1;
    }
  }
  if (input.empty()) {
    std::printf("Input file not specified. Usage: %s -f <filename> [-w -o <output_dirname>]\n", argv[0]);
    return 1;
  }
  std::string output;
  if (!out_dir.empty()) {
    size_t k = input.find_last_of('/');
    output = out_dir + "/" + (k == std::string::npos ? input : input.substr(k + 1)) + ".bin";
  } else {
    output = input + ".bin";
  }
  std::printf("%s\n", output.c_str());
  int32_t n = 0;
  int64_t m = 0;
  gala_ctx* ctx = nullptr;
  if (gpu >= 0) {
    int rc = gala_create(&ctx, gpu, nullptr);
    if (rc) { std::fprintf(stderr, "gpu %d: %s\n", gpu, gala_strerror(rc)); return 2; }
  }
  // the two-call protocol, on the host or on the GPU
  auto ingest = [&](int64_t* e, int32_t* nb, void* w) {
    return ctx ? gala_ingest_text_device(ctx, input.c_str(), wmode, &n, &m, e, nb, w)
               : gala_ingest_text(input.c_str(), wmode, &n, &m, e, nb, w);
  };
  int rc = ingest(nullptr, nullptr, nullptr);
  if (rc) { std::fprintf(stderr, "ingest failed: %s\n", gala_strerror(rc)); gala_destroy(ctx); return 2; }
  std::vector<int64_t> end(n);
  std::vector<int32_t> nbr(m), w(wmode != GALA_W_NONE ? m : 0);
  rc = ingest(end.data(), nbr.data(), w.empty() ? nullptr : w.data());
  gala_destroy(ctx);
  if (rc) { std::fprintf(stderr, "ingest failed: %s\n", gala_strerror(rc)); return 2; }
  std::printf("vertex number:%d edge number:%lld\n", n, (long long)(m / 2));
  rc = gala_write_bin(output.c_str(), n, end.data(), nbr.data(), w.empty() ? nullptr : w.data(), wmode);
  if (rc) { std::fprintf(stderr, "write failed: %s\n", gala_strerror(rc)); return 2; }
  std::printf("store successfully. \n");
  return 0;
}
