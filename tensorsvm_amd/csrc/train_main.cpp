// tensorsvm-train: the drop-in command line of the reference trainer
// (main/parsecmd/help_msg, tensorsvm-train.cu:858-1156), running the hot path through
// libtsvm's C-ABI on one GPU.
//
// Kept options: -k rank, -c C, -g gamma, -t 0|2, -s 0|1, -N n, -scale256, -analysis,
// -test file, -tensorcore (= -precision bf16), positional training file and model file
// (default: training file + ".model").  Defaults as the reference's globals (66-69):
// K=50, C=1, g=0, t=0.  Added: -precision fp32|bf16|split|fp8, -stop ref|gap, -tol x, -refine n,
// -maxit n, -q (quiet), -json file (machine-readable summary of the run).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/tsvm.h"

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void help_msg() {
  printf("Usage: tensorsvm-train [options] training_set_file [model_file]\n");
  printf(
      "options:\n"
      "-t kernel_type : 0 -- linear: u'*v ; 2 -- radial basis function: exp(-gamma*|u-v|^2)\n"
      "-k rank : rank of the kernel approximation (default 50)\n"
      "-g gamma : set gamma in kernel function (default 0)\n"
      "-c cost : set the parameter C of C-SVC (default 1)\n"
      "-s solver : 0 -- approx interior point method (default); 1 -- projected gradient descent\n"
      "-N n_samples : only consider the first n_samples in training\n"
      "-test file : held-out LIBSVM file; prints the prediction accuracy\n"
      "-scale256 : divide every feature value by 256\n"
      "-analysis : compute fnorm(K-U*U'), the error of the low rank approximation (RBF)\n"
      "-precision fp32|bf16|split|fp8 : arithmetic of the kernel products (default fp32;\n"
      "    split: bf16 with the factor's sketch in two fp16 terms; fp8: bf16 with an fp8\n"
      "    tier in the solver's capacitance)\n"
      "-tensorcore : same as -precision bf16\n"
      "-stop ref|gap : IPM stop rule (ref: mu,|r_x|,|r_y| < 1e-7; gap: relative gap <= -tol)\n"
      "-tol x : tolerance of -stop gap (default 1e-5)\n"
      "-refine n : refinement sweeps of the Newton steps: n > 0 for each of the affine and the\n"
      "    corrector's steps, -n for the combined step, started from the affine one (the first\n"
      "    sweep adds the corrector's step; default -1)\n"
      "-maxit n : IPM iteration limit (default 100)\n"
      "-init ref|feasible : IPM initial point (feasible: zero initial dual residual)\n"
      "-q : quiet\n"
      "-json file : write a JSON summary of the run\n");
}

void map_test_labels(std::vector<float>& lab, int pos, int neg) {
  for (auto& v : lab) {
    if (v == (float)pos) v = 1;
    else if (v == (float)neg) v = -1;
  }
}

}  // namespace

int main(int argc, char* argv[]) {
  tsvm_train_params prm;
  tsvm_train_default_params(&prm);
  prm.kernel_type = 0;  // the reference's default T=0 (66-69)
  prm.verbose = 1;
  const char* train_path = nullptr;
  const char* test_path = nullptr;
  const char* json_path = nullptr;
  std::string model_path;
  long NN = 0;
  int solver = 0, scale256 = 0, analysis = 0;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        fprintf(stderr, "missing value for %s\n", argv[i]);
        exit(1);
      }
      return argv[++i];
    };
    if (a == "-c") prm.C = atof(next());
    else if (a == "-g") prm.gamma = atof(next());
    else if (a == "-test") test_path = next();
    else if (a == "-t") {
      prm.kernel_type = atoi(next());
      if (prm.kernel_type != 0 && prm.kernel_type != 2) {
        printf("Error: unimplemented -t %d\n", prm.kernel_type);
        exit(1);
      }
    } else if (a == "-N") NN = atol(next());
    else if (a == "-k") prm.k = atol(next());
    else if (a == "-s") solver = atoi(next());
    else if (a == "-analysis") analysis = 1;
    else if (a == "-tensorcore") prm.precision = TSVM_PREC_BF16;
    else if (a == "-scale256") scale256 = 1;
    else if (a == "-precision") {
      std::string p = next();
      if (p == "fp32") prm.precision = TSVM_PREC_FP32;
      else if (p == "bf16") prm.precision = TSVM_PREC_BF16;
      else if (p == "split") prm.precision = TSVM_PREC_FP16X2;
      else if (p == "fp8") prm.precision = TSVM_PREC_FP8;
      else {
        fprintf(stderr, "unknown precision %s\n", p.c_str());
        exit(1);
      }
    } else if (a == "-stop") {
      std::string p = next();
      if (p == "gap") prm.stop_mode = TSVM_STOP_GAP;
      else if (p == "ref") prm.stop_mode = TSVM_STOP_REF;
      else {
        fprintf(stderr, "unknown stop rule %s\n", p.c_str());
        exit(1);
      }
    } else if (a == "-tol") prm.tol = atof(next());
    else if (a == "-refine") prm.refine = atoi(next());
    else if (a == "-maxit") prm.maxit = atoi(next());
    else if (a == "-init") {
      std::string p = next();
      if (p == "feasible") prm.init = 1;
      else if (p == "ref") prm.init = 0;
      else {
        fprintf(stderr, "unknown initial point %s\n", p.c_str());
        exit(1);
      }
    }
    else if (a == "-q") prm.verbose = 0;
    else if (a == "-json") json_path = next();
    else if (!train_path) train_path = argv[i];
    else model_path = argv[i];
  }
  if (!train_path) {
    help_msg();
    return 1;
  }
  if (model_path.empty()) model_path = std::string(train_path) + ".model";
  if (prm.kernel_type == 2 && (prm.k <= 0 || prm.k > TSVM_MAX_RANK)) {
    fprintf(stderr, "Error: -k %ld outside 1..%d\n", (long)prm.k, TSVM_MAX_RANK);
    return 1;
  }
  const bool v = prm.verbose != 0;

  clock_t before = clock();
  float *labels = nullptr, *inst = nullptr;
  int64_t n = 0, d = 0;
  tsvm_status st = tsvm_read_libsvm(train_path, scale256, &labels, &inst, &n, &d);
  if (st != TSVM_OK) {
    fprintf(stderr, st == TSVM_ERR_IO ? "Can't open %s\n" : "Wrong input format in %s\n", train_path);
    return 1;
  }
  if (v) printf("Dataset size %ld #features %ld\n", (long)n, (long)d);
  if (prm.precision != TSVM_PREC_FP32 && v) printf("\x1b[31mUsing BF16 MFMA\x1b[39m\n");
  if (NN != 0 && NN < n) {
    if (v) printf("Truncating the input file to first %ld instances\n", NN);
    n = NN;
  }
  int pos, neg;
  int64_t bad = tsvm_map_labels(labels, n, &pos, &neg);
  if (bad) printf("Error: %ld labels outside {%d, %d}\n", (long)bad, pos, neg);
  if (v) {
    printf("found labels: %d(+1) %d(-1)\n", pos, neg);
    printf("Reading files took %.3f seconds\n", 1.0 * (clock() - before) / CLOCKS_PER_SEC);
  }
  tsvm_ctx* ctx = nullptr;
  st = tsvm_create(0, &ctx);
  if (st != TSVM_OK) {
    fprintf(stderr, "no GPU device (status %d)\n", st);
    return 1;
  }
  if (prm.kernel_type == 2 && solver == 1) {
    printf("unimplemented pgd\n");
    tsvm_destroy(ctx);
    return 0;
  }
  std::vector<double> alpha(n), w(d);
  prm.analysis = analysis && prm.kernel_type == 2;
  tsvm_train_info info;
  const double t0 = now();
  st = tsvm_train(ctx, inst, labels, n, d, &prm, nullptr, alpha.data(), w.data(), &info);
  const double t_train = now() - t0;
  if (st != TSVM_OK) {
    fprintf(stderr, "training failed: %s\n", tsvm_last_error(ctx));
    tsvm_destroy(ctx);
    return 1;
  }
  if (v) printf("Writemodel ");
  if (prm.kernel_type == 0)
    st = tsvm_write_model_linear(model_path.c_str(), w.data(), d, info.b, pos, neg);
  else
    st = tsvm_write_model_rbf(model_path.c_str(), alpha.data(), labels, inst, n, d, prm.C,
                              prm.gamma, info.b, pos, neg);
  if (st != TSVM_OK) {
    fprintf(stderr, "Can't open %s\n", model_path.c_str());
    tsvm_destroy(ctx);
    return 1;
  }
  if (v) printf("done\n");
  double acc = -1;
  if (test_path) {
    float *tl = nullptr, *tx = nullptr;
    int64_t tn = 0, td = 0;
    if (tsvm_read_libsvm(test_path, scale256, &tl, &tx, &tn, &td) != TSVM_OK) {
      fprintf(stderr, "Can't read %s\n", test_path);
    } else {
      std::vector<float> T((size_t)(tn * d), 0.0f);
      for (int64_t i = 0; i < tn; i++)
        for (int64_t t = 0; t < d && t < td; t++) T[i * d + t] = tx[i * td + t];
      std::vector<float> lab(tl, tl + tn);
      map_test_labels(lab, pos, neg);
      std::vector<double> f(tn);
      long hit = 0;
      if (prm.kernel_type == 2) {
        tsvm_decision_function(ctx, inst, labels, alpha.data(), n, d, prm.gamma, info.b, T.data(), tn,
                               f.data());
      } else {
        for (int64_t i = 0; i < tn; i++) {
          double s = info.b;
          for (int64_t t = 0; t < d; t++) s += w[t] * T[i * d + t];
          f[i] = s;
        }
      }
      for (int64_t i = 0; i < tn; i++)
        if (f[i] * lab[i] > 0) hit++;
      acc = tn ? 1.0 * hit / tn : 0.0;
      printf("prediction accuracy %.3f (%ld/%ld)\n", acc, hit, (long)tn);
      tsvm_free(tl);
      tsvm_free(tx);
    }
  }
  if (json_path) {
    FILE* jf = fopen(json_path, "w");
    if (jf) {
      fprintf(jf,
              "{\"n\": %ld, \"d\": %ld, \"k\": %ld, \"C\": %g, \"gamma\": %g, \"precision\": \"%s\", "
              "\"train_s\": %.6f, \"lra_s\": %.6f, \"ipm_s\": %.6f, \"iters\": %d, \"reason\": %d, "
              "\"pobj\": %.12e, \"dobj\": %.12e, \"gap_rel\": %.6e, \"b\": %.12e, \"nsv\": %ld, "
              "\"nfree\": %ld, \"accuracy\": %.6f}\n",
              (long)n, (long)d, (long)prm.k, prm.C, prm.gamma,
              prm.precision == TSVM_PREC_BF16 ? "bf16" : (prm.precision == TSVM_PREC_FP16X2 ? "split" : (prm.precision == TSVM_PREC_FP8 ? "fp8" : "fp32")),
              t_train, info.lra.t_total,
              info.ipm.t_total, info.ipm.iters, info.ipm.reason, info.ipm.pobj, info.ipm.dobj,
              info.ipm.gap_rel, info.b, (long)info.nsv, (long)info.nfree, acc);
      fclose(jf);
    }
  }
  tsvm_destroy(ctx);
  tsvm_free(labels);
  tsvm_free(inst);
  return 0;
}
