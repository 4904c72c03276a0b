// test_inplace: the reference driver's command line over libeithot.so.
//   test_inplace d1 .. dn p1 .. pn elem_bytes n_sub alpha
// Same argv (1-based perm), stdout lines and appended log files as
// ITNT/src/test_inplace.cpp:14-89,142-177: pinned host buffer initialised
// with v(i) = i, H2D, in-place transpose, D2H, closed-form check
// (tensor_util.cpp:112-142), timings with and without the PCIe copies.
// EITHOT_MEMORY_TYPE=managed selects the reference's other mode
// (padding_managed_transpose, ITNT/src/test_inplace.cpp:91-139): one managed
// allocation of 1.03x the tensor, initialised and checked by the host in
// place, transposed by the device with no explicit copies.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/eithot.h"

#define HIPCHECK(x)                                                          \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 1;                                                              \
    }                                                                        \
  } while (0)

static void put(unsigned char* p, uint64_t i, int elem, uint64_t v) { memcpy(p + i * elem, &v, elem); }
static uint64_t get(const unsigned char* p, uint64_t i, int elem) {
  uint64_t v = 0;
  memcpy(&v, p + i * elem, elem);
  return v;
}

// Closed-form check: the element that started at linear index s must be at
// its destination, and with v(i) = i it holds s (truncated to elem bytes).
// Prints the reference's lines (ITNT/src/tensor_util.cpp:112-142).
static bool verify_inplace(const unsigned char* h, const std::vector<int64_t>& dims,
                           const std::vector<int>& perm, int rank, int elem, uint64_t vol) {
  printf("Inplace transpose verify_inplace\n");
  std::vector<uint64_t> ostride(rank);
  uint64_t s = 1;
  for (int i = 0; i < rank; ++i) {
    ostride[perm[i] - 1] = s;
    s *= (uint64_t)dims[perm[i] - 1];
  }
  const uint64_t mask = elem == 8 ? ~0ull : ((1ull << (8 * elem)) - 1);
  unsigned nthr = std::max(1u, std::thread::hardware_concurrency());
  std::vector<int> ok(nthr, 1);
  std::vector<std::thread> th;
  for (unsigned t0 = 0; t0 < nthr; ++t0)
    th.emplace_back([&, t0] {
      for (uint64_t src = vol * t0 / nthr; src < vol * (t0 + 1) / nthr; ++src) {
        uint64_t rem = src, dst = 0;
        for (int i = 0; i < rank; ++i) {
          dst += (rem % (uint64_t)dims[i]) * ostride[i];
          rem /= (uint64_t)dims[i];
        }
        if (get(h, dst, elem) != (src & mask)) {
          ok[t0] = 0;
          return;
        }
      }
    });
  for (auto& x : th) x.join();
  bool all = true;
  for (int x : ok) all = all && x;
  printf(all ? "Correct transpose\n" : "Incorrect transpose\n");
  return all;
}

static void append_logs(double t_ms, double bytes) {
  FILE* f = fopen("inplace_bench.txt", "a+");
  if (f) {
    fprintf(f, "%.5f\n", t_ms);
    fclose(f);
  }
  f = fopen("inplace_bench_throughput.txt", "a+");
  if (f) {
    fprintf(f, "%.5f\n", bytes * 2 / 1e6 / t_ms);
    fclose(f);
  }
}

// The managed-memory mode (ITNT/src/test_inplace.cpp:91-139).
static int managed_transpose(const std::vector<int64_t>& dims, const std::vector<int>& perm,
                             int rank, int elem, int n_sub, double alpha, uint64_t vol,
                             size_t bytes) {
  const size_t alloc = (size_t)((1 + 0.03) * (double)bytes);
  unsigned char* d = nullptr;
  HIPCHECK(hipMallocManaged((void**)&d, alloc));
  for (uint64_t i = 0; i < vol; ++i) put(d, i, elem, i);
  long long perm_int = 0;
  for (int i = 0; i < rank; ++i) perm_int = perm_int * 10 + perm[i];
  printf("%dD Inplace %lld transpose", rank, perm_int);
  printf(", Data size = %.5f GB\n", (double)alloc / 1e9);
  printf("number of sub_tensors = %d\n", n_sub);
  hipEvent_t start, stop;
  HIPCHECK(hipEventCreate(&start));
  HIPCHECK(hipEventCreate(&stop));
  HIPCHECK(hipEventRecord(start, 0));
  const int rc = eithot_transpose(d, dims.data(), perm.data(), rank, (size_t)elem, n_sub, alpha);
  if (rc != EITHOT_OK) {
    fprintf(stderr, "eithot_transpose failed (%d): %s\n", rc, eithot_last_error());
    return 1;
  }
  HIPCHECK(hipDeviceSynchronize());
  HIPCHECK(hipEventRecord(stop, 0));
  HIPCHECK(hipEventSynchronize(stop));
  float t = 0;
  HIPCHECK(hipEventElapsedTime(&t, start, stop));
  const bool all = verify_inplace(d, dims, perm, rank, elem, vol);
  printf("Execution Time: %.5fms\n", t);
  append_logs(t, (double)alloc);
  HIPCHECK(hipFree(d));
  return all ? 0 : 3;
}

int main(int argc, char** argv) {
  const int rank = (argc - 1) / 2 - 1;
  if (rank < 1 || argc != 2 * (rank + 1) + 2) {
    fprintf(stderr, "usage: %s d1..dn p1..pn elem_bytes n_sub alpha\n", argv[0]);
    return 2;
  }
  std::vector<int64_t> dims(rank);
  std::vector<int> perm(rank);
  int k = 1;
  for (int i = 0; i < rank; ++i) dims[i] = atoll(argv[k++]);
  for (int i = 0; i < rank; ++i) perm[i] = atoi(argv[k++]);
  const int elem = atoi(argv[k++]);
  const int n_sub = atoi(argv[k++]);
  const double alpha = atof(argv[k++]);
  uint64_t vol = 1;
  for (int64_t d : dims) vol *= (uint64_t)d;
  const size_t bytes = vol * (size_t)elem;
  const char* mt = getenv("EITHOT_MEMORY_TYPE");
  if (mt && !strcmp(mt, "managed"))
    return managed_transpose(dims, perm, rank, elem, n_sub, alpha, vol, bytes);
  long long perm_int = 0;
  for (int i = 0; i < rank; ++i) perm_int = perm_int * 10 + perm[i];
  printf("%dD Inplace %lld transpose\n", rank, perm_int);
  printf("number of sub_tensors = %d\n", n_sub);
  printf("Dims =");
  for (int i = 0; i < rank; ++i) printf(" %lld", (long long)dims[i]);
  printf(", Data size = %.5f GB\n", (double)bytes / 1e9);

  unsigned char* h = nullptr;
  void* d = nullptr;
  HIPCHECK(hipHostMalloc((void**)&h, bytes, hipHostMallocDefault));
  HIPCHECK(hipMalloc(&d, bytes));
  for (uint64_t i = 0; i < vol; ++i) put(h, i, elem, i);
  hipEvent_t start, h2d, trans, d2h, stop;
  HIPCHECK(hipEventCreate(&start));
  HIPCHECK(hipEventCreate(&h2d));
  HIPCHECK(hipEventCreate(&trans));
  HIPCHECK(hipEventCreate(&d2h));
  HIPCHECK(hipEventCreate(&stop));
  HIPCHECK(hipEventRecord(start, 0));
  HIPCHECK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
  HIPCHECK(hipEventRecord(h2d, 0));
  const int rc = eithot_transpose(d, dims.data(), perm.data(), rank, (size_t)elem, n_sub, alpha);
  HIPCHECK(hipEventRecord(trans, 0));
  if (rc != EITHOT_OK) {
    fprintf(stderr, "eithot_transpose failed (%d): %s\n", rc, eithot_last_error());
    return 1;
  }
  HIPCHECK(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
  HIPCHECK(hipEventRecord(d2h, 0));
  HIPCHECK(hipDeviceSynchronize());
  HIPCHECK(hipEventRecord(stop, 0));
  HIPCHECK(hipEventSynchronize(stop));
  float t = 0, th2d = 0, td2h = 0;
  HIPCHECK(hipEventElapsedTime(&th2d, start, h2d));
  HIPCHECK(hipEventElapsedTime(&td2h, trans, d2h));
  HIPCHECK(hipEventElapsedTime(&t, start, stop));

  const bool all = verify_inplace(h, dims, perm, rank, elem, vol);
  printf("Execution Time without h2d2h: %.5fms\n", t - th2d - td2h);
  printf("Execution Time: %.5fms\n", t);
  printf("Peak extra device memory = %zu B (%.5f of tensor bytes)\n", eithot_last_peak_extra_bytes(),
         (double)eithot_last_peak_extra_bytes() / (double)bytes);
  append_logs(t, (double)bytes);
  (void)hipFree(d);
  (void)hipHostFree(h);
  return all ? 0 : 3;
}
