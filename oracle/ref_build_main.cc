// Test infrastructure only: command-line front end of oracle/ref_driver.cc.
// Builds an HNSW index with the reference sources from a raw little-endian
// float32 matrix file and saves it in the reference's optimized format
// (hnsw.cc:748-806), as nmslib.init(...).addDataPointBatch/createIndex/
// saveIndex would.
//   ref_build <space> <data.f32> <n> <dim> <params k=v,k=v> <out.index>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <cstdint>

extern "C" {
const char* ref_last_error();
void* ref_create(const char* space, const float* data, int64_t n, int dim);
int ref_create_index(void* h, const char* params);
int ref_save_index(void* h, const char* path);
void ref_free(void* h);
}

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s space data.f32 n dim params out\n", argv[0]);
    return 2;
  }
  int64_t n = std::atoll(argv[3]);
  int dim = std::atoi(argv[4]);
  std::vector<float> data((size_t)n * dim);
  FILE* f = std::fopen(argv[2], "rb");
  if (!f || std::fread(data.data(), sizeof(float), data.size(), f) != data.size()) {
    std::fprintf(stderr, "cannot read %s\n", argv[2]);
    return 1;
  }
  std::fclose(f);
  void* h = ref_create(argv[1], data.data(), n, dim);
  if (!h) { std::fprintf(stderr, "%s\n", ref_last_error()); return 1; }
  std::vector<float>().swap(data);
  if (ref_create_index(h, argv[5])) { std::fprintf(stderr, "%s\n", ref_last_error()); return 1; }
  if (ref_save_index(h, argv[6])) { std::fprintf(stderr, "%s\n", ref_last_error()); return 1; }
  ref_free(h);
  std::printf("saved %s\n", argv[6]);
  return 0;
}
