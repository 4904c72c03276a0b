// Nd_estimation: the reference planner's command line over libeithot.so.
//   Nd_estimation d1 .. dn p1 .. pn elem_bytes n_sub alpha
// Prints the same lines as Nd_memory_estimation/src/test_inplace.cpp:18-147
// and appends "dims perm elem nb alpha" to proper_input.txt.
// EITHOT_ESTIMATOR=native (default) plans for this build's kernels;
// EITHOT_ESTIMATOR=reference evaluates the reference's own size model on the
// CUDA device profile in EITHOT_REF_PROFILE="smem,sms,threads_per_sm,pad_occ"
// (default 49152,82,1536,1, an RTX 3090).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/eithot.h"

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
  (void)atoi(argv[k++]);
  double alpha = atof(argv[k++]);
  const char* mode = getenv("EITHOT_ESTIMATOR");
  int nb = 1;
  if (mode && !strcmp(mode, "reference")) {
    eithot_ref_profile p = {49152, 82, 1536, 1};
    const char* pr = getenv("EITHOT_REF_PROFILE");
    if (pr) {
      long long a = 0;
      int b = 0, c = 0, d = 0;
      if (sscanf(pr, "%lld,%d,%d,%d", &a, &b, &c, &d) == 4) p = {a, b, c, d};
    }
    double init = 0, nr = 0, fin = 0, fr = 0, aout = 0;
    int flags = 0;
    if (eithot_estimate_reference(dims.data(), perm.data(), rank, (size_t)elem, alpha, &p, &init,
                                  &nr, &fin, &fr, &nb, &aout, &flags) != EITHOT_OK) {
      fprintf(stderr, "estimate failed: %s\n", eithot_last_error());
      return 1;
    }
    printf("init Extra Memory size = %.5f MB\n", init / 1e6);
    if (flags & 1) printf("new ratio = %f\n", nr);
    if (flags & 2)
      printf("Final Extra Memory Size = %.5f MB, take %.5f percent input data size\n", fin / 1e6, fr);
    alpha = aout;
  } else {
    int nsub = 1;
    double need = 0;
    size_t extra = 0;
    if (eithot_estimate(dims.data(), perm.data(), rank, (size_t)elem, alpha, &nsub, &need, &extra) !=
        EITHOT_OK) {
      fprintf(stderr, "estimate failed: %s\n", eithot_last_error());
      return 1;
    }
    nb = nsub;
    printf("init Extra Memory size = %.5f MB\n", (double)extra / 1e6);
    if (extra > 0) {
      printf("new ratio = %f\n", alpha);
      printf("Final Extra Memory Size = %.5f MB, take %.5f percent input data size\n",
             (double)extra / 1e6, need);
    }
    alpha = std::max(alpha, std::ceil(need * 100.0) / 100.0);
  }
  printf("\t");
  for (int i = 0; i < rank; ++i) printf("%lld ", (long long)dims[i]);
  for (int i = 0; i < rank; ++i) printf("%d ", perm[i]);
  printf("%d %d %.2f\n", elem, nb, alpha);
  FILE* f = fopen("proper_input.txt", "a+");
  if (f) {
    for (int i = 0; i < rank; ++i) fprintf(f, "%lld ", (long long)dims[i]);
    for (int i = 0; i < rank; ++i) fprintf(f, "%d ", perm[i]);
    fprintf(f, "%d %d %.2f\n", elem, nb, alpha);
    fclose(f);
  }
  return 0;
}
