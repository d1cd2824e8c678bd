// tensorsvm-predict: scores a LIBSVM-format test file with a model file written by
// tensorsvm-train, on the GPU (the role of the external svm-predict / thundersvm-predict
// in the reference's README.md:27-37).  Usage and output follow svm-predict:
//   tensorsvm-predict [-q] test_file model_file output_file
// writes one predicted label per line and prints "Accuracy = x% (hit/total) (classification)".
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/tsvm.h"

int main(int argc, char* argv[]) {
  int quiet = 0, a = 1;
  if (a < argc && !strcmp(argv[a], "-q")) {
    quiet = 1;
    a++;
  }
  if (argc - a != 3) {
    printf("Usage: tensorsvm-predict [-q] test_file model_file output_file\n");
    return 1;
  }
  const char *test_path = argv[a], *model_path = argv[a + 1], *out_path = argv[a + 2];
  tsvm_model* m = nullptr;
  tsvm_status st = tsvm_read_model(model_path, &m);
  if (st != TSVM_OK) {
    fprintf(stderr, "can't open model file %s (status %d)\n", model_path, st);
    return 1;
  }
  float *lab = nullptr, *inst = nullptr;
  int64_t n = 0, d = 0;
  if (tsvm_read_libsvm(test_path, 0, &lab, &inst, &n, &d) != TSVM_OK) {
    fprintf(stderr, "can't open input file %s\n", test_path);
    tsvm_model_free(m);
    return 1;
  }
  tsvm_ctx* ctx = nullptr;
  if (tsvm_create(0, &ctx) != TSVM_OK) {
    fprintf(stderr, "no GPU device\n");
    return 1;
  }
  float* pred = (float*)malloc(sizeof(float) * (n > 0 ? n : 1));
  st = tsvm_predict(ctx, m, inst, n, d, nullptr, pred);
  if (st != TSVM_OK) {
    fprintf(stderr, "prediction failed: %s\n", tsvm_last_error(ctx));
    return 1;
  }
  FILE* f = fopen(out_path, "w");
  if (!f) {
    fprintf(stderr, "can't open output file %s\n", out_path);
    return 1;
  }
  long hit = 0;
  for (int64_t i = 0; i < n; i++) {
    fprintf(f, "%g\n", pred[i]);
    if (pred[i] == lab[i]) hit++;
  }
  fclose(f);
  if (!quiet)
    printf("Accuracy = %g%% (%ld/%ld) (classification)\n", n ? 100.0 * hit / n : 0.0, hit, (long)n);
  free(pred);
  tsvm_free(lab);
  tsvm_free(inst);
  tsvm_model_free(m);
  tsvm_destroy(ctx);
  return 0;
}
