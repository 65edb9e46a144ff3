// acdc_hip_lasso_path -- the Lasso lambda-path experiment on the gfx950 path.
//
// Counterpart of cpp/src/experiments/lasso_console_solver.cpp:18-171: a
// dense CSV matrix A and a CSV b (-A, -b, the reference's two options),
// lambda from lambdaHigh down to lambdaLow in `pieces` geometric steps
// (lambda *= (low/high)^(1/pieces) while lambda >= low, in float as the
// reference steps it), each lambda warm-started from the previous x and run
// for `rounds` rounds of `multiply` x n asynchronous updates (the
// reference's N = 20 rounds of 100 n), with L_j = 1 / ||a_j||^2 (the
// reference's h_Li, sigma = 1).  Per round it prints and logs the
// reference's row "threads,n,m,sigma,time,iterations,objective,nnz" (time:
// seconds inside the rounds; iterations: epochs of n), to
// <out>/lasso.log, and after each lambda writes x, one value per line with
// 15 significant digits, to <out>/lasso_x_<lambda>.log.
// The reference recomputes the residual in double every third round; here
// the FP32 path refreshes every `refresh` epochs (or FP64 throughout, -p 0).
// Options beyond the reference's (defaults = its constants):
//   -H <lambda high, 100>  -L <lambda low, 1e-5>  -N <pieces, 20>
//   -R <rounds, 20>  -m <multiply, 100>  -t <workers, 4 = its omp_threads>
//   -o <output directory, results>  -p <0 fp64 | 1 fp32+refresh (default)>
//   -s <seed>  -g <relative-gap target: stop a lambda's rounds early, 0 = never>
#include <getopt.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <ctime>
#include <vector>

#include "../../include/acdc_hip.h"
#include "../host/acdc/file_reader.h"

namespace {
void check(acdc_status st, const char *what) {
  if (st != ACDC_OK) throw std::runtime_error(std::string(what) + ": " + acdc_last_error());
}
double now_s() {
  timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}
}  // namespace

int main(int argc, char *argv[]) {
  std::string matrixA, vectorB, outDir = "results";
  float lambdaHigh = 100, lambdaLow = 0.00001f;
  int pieces = 20, rounds = 20, multiply = 100, tau = 4, precision = ACDC_PREC_FP32_REFRESH;
  uint64_t seed = 1;
  double gapTarget = 0;
  int c;
  while ((c = getopt(argc, argv, "A:b:H:L:N:R:m:t:o:p:s:g:h")) != -1) {
    switch (c) {
      case 'A': matrixA = optarg; break;
      case 'b': vectorB = optarg; break;
      case 'H': lambdaHigh = (float)std::atof(optarg); break;
      case 'L': lambdaLow = (float)std::atof(optarg); break;
      case 'N': pieces = std::atoi(optarg); break;
      case 'R': rounds = std::atoi(optarg); break;
      case 'm': multiply = std::atoi(optarg); break;
      case 't': tau = std::atoi(optarg); break;
      case 'o': outDir = optarg; break;
      case 'p': precision = std::atoi(optarg); break;
      case 's': seed = std::strtoull(optarg, nullptr, 10); break;
      case 'g': gapTarget = std::atof(optarg); break;
      default:
        std::cerr << "usage: acdc_hip_lasso_path -A <matrix.csv> -b <b.csv> [-H high] [-L low]\n"
                     "       [-N pieces] [-R rounds] [-m multiply] [-t workers] [-o outdir]\n"
                     "       [-p precision] [-s seed] [-g gap target]\n";
        return c == 'h' ? 0 : 2;
    }
  }
  if (matrixA.empty() || vectorB.empty() || pieces < 1 || rounds < 1 || multiply < 1 ||
      !(lambdaHigh > 0) || !(lambdaLow > 0)) {
    std::cerr << "acdc_hip_lasso_path: -A and -b are required; pieces, rounds, multiply and "
                 "both lambdas must be positive\n";
    return 2;
  }
  try {
    ProblemData<int, double> inst;
    InputOuputHelper::loadDenseCSV(matrixA, vectorB, inst);
    std::cout << "Loaded data with m=" << inst.m << " and n=" << inst.n << std::endl;
    acdc_problem *p = nullptr;
    check(acdc_problem_create_csc(inst.m, inst.n, (int64_t)inst.A_csc_values.size(),
                                  inst.A_csc_col_ptr.data(), inst.A_csc_row_idx.data(),
                                  inst.A_csc_values.data(), inst.b.data(), &p),
          "acdc_problem_create_csc");
    acdc_settings st{};
    st.lambda = lambdaHigh;
    st.tau = tau;
    st.seed = seed;
    st.beta_mode = ACDC_BETA_NONE;  // L_j = 1 / ||a_j||^2, the reference's h_Li
    st.precision = precision;
    st.refresh_period = precision == ACDC_PREC_FP32_REFRESH ? 5 : 0;
    st.step_mode = ACDC_STEP_ASYNC;
    acdc_solver *s = nullptr;
    check(acdc_solver_create(p, ACDC_LOSS_L1_SQUARE, &st, &s), "acdc_solver_create");
    std::ofstream log((outDir + "/lasso.log").c_str());
    if (!log.is_open()) throw std::runtime_error("cannot write " + outDir + "/lasso.log");
    const double sigma = 1.0;
    double totalRunningTime = 0, iterations = 0;
    acdc_report r;
    check(acdc_report_get(s, &r), "acdc_report_get");
    log << std::setprecision(16) << tau << "," << inst.n << "," << inst.m << "," << sigma << ","
        << totalRunningTime << "," << iterations << "," << r.primal << std::endl;
    std::vector<double> x((size_t)inst.n);
    const float delta = std::pow(lambdaLow / lambdaHigh, 1 / (pieces + 0.0f));
    for (float lambda = lambdaHigh; lambda >= lambdaLow; lambda = lambda * delta) {
      std::cout << "lambda " << lambda << std::endl;
      check(acdc_set_lambda(s, lambda), "acdc_set_lambda");
      for (int round = 0; round < rounds; ++round) {
        const double t0 = now_s();
        check(acdc_run_updates(s, (uint64_t)multiply * (uint64_t)inst.n), "acdc_run_updates");
        check(acdc_synchronize(s), "acdc_synchronize");
        totalRunningTime += now_s() - t0;
        iterations += multiply;
        check(acdc_report_get(s, &r), "acdc_report_get");
        std::ostringstream row;
        row << tau << "," << inst.n << "," << inst.m << "," << sigma << "," << totalRunningTime
            << "," << iterations << "," << std::setprecision(15) << r.primal << ","
            << (long long)r.nnz_x;
        std::cout << row.str() << std::endl;
        log << std::setprecision(16) << tau << "," << inst.n << "," << inst.m << "," << sigma
            << "," << totalRunningTime << "," << iterations << "," << std::setprecision(15)
            << r.primal << "," << (long long)r.nnz_x << std::endl;
        if (gapTarget > 0 && r.rel_gap <= gapTarget) break;
      }
      check(acdc_get_x(s, x.data(), inst.n), "acdc_get_x");
      std::stringstream name;
      name << outDir << "/lasso_x_" << lambda << ".log";
      std::ofstream out(name.str().c_str());
      if (!out.is_open()) throw std::runtime_error("cannot write " + name.str());
      for (double v : x) out << std::setprecision(15) << v << "\n";
    }
    acdc_solver_destroy(s);
    acdc_problem_destroy(p);
  } catch (const std::exception &e) {
    std::cerr << "acdc_hip_lasso_path: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
