// acdc_hip_solver -- console front end of the gfx950 PCDM path.
//
// Counterpart of cpp/src/frontends/solvers/MultiCoreSolver.cpp:18-115 with
// the getopt letters of cpp/src/helpers/option_console_parser.h:102-114:
//   -A <prefix|file>  text CSC prefix (losses 0/10) or LIBSVM file (loss 2)
//   -b <file>         right-hand side b (losses 0/10)
//   -l <lambda>       regularisation
//   -t <tau>          coordinates per synchronous step (totalThreads)
//   -f <loss>         0 L1-square (Lasso), 2 SVM dual hinge, 10 NNLS
//   -I <n>            updates per bulk (innerIterations)
//   -c <n>            number of bulks (the reference stores -c into a bool)
//   -i <0|1>          show initial/intermediate/last objective values
//   -r <file>         CSV log file
//   -T <seconds>      stop once the bulks' pure computation time reaches it
//                     (the reference parses -T and ignores it; here it sets
//                     settings.total_execution_time)
// and the HIP knobs (no reference counterpart):
//   -s <seed>  -p <0 fp64 | 1 fp32+refresh>  -R <refresh epochs>  -L <lanes>
//   -P <0 auto | 1 per-step launches | 2 persistent grid | 3 per step, streamed scatter>
//   -M <0 sync | 1 async (default, the multicore executor's workers)>
//   -B <-1 reference rule (default) | 0 omega | 2 none | 3 spectral>
//   -S <0|1>   spectral ESO beta (the power method of mSDCA.cpp:188-286),
//              the same as -B 3
// In the reference, -I falls through into -M (option_console_parser.h:89-93);
// here each flag sets only its own field.  omega and sigma are computed as in
// MultiCoreSolver.cpp:24-38, :55-62, :108-110 and handed to the executor in
// instance.sigma; under the reference rule the Lasso and NNLS ignore it
// (MulticoreSquareLoss.h:27-40) and the SVM dual uses it.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../host/acdc/HipEngineExecutor.h"
#include "../host/acdc/Solver.h"
#include "../host/acdc/file_reader.h"

static void usage() {
  std::cerr << "usage: acdc_hip_solver -A <data> [-b <b file>] -l <lambda> [-t tau] [-f loss]\n"
               "                       [-I updates per bulk] [-c bulks] [-i 0|1] [-r log]\n"
               "                       [-T seconds of computation]\n"
               "                       [-s seed] [-p precision] [-R refresh epochs] [-L lanes]\n"
               "                       [-P launch mode] [-M step mode: 0 sync, 1 async]\n"
               "                       [-B beta rule: -1 reference, 0 omega, 2 none, 3 spectral]\n"
               "                       [-S 1: spectral ESO beta]\n";
}

int main(int argc, char *argv[]) {
  OptimizationSettings settings;
  settings.verbose = true;
  std::string matrixA, vectorB, logPath;
  double lambda = 0;
  int c;
  while ((c = getopt(argc, argv, "A:b:l:t:f:I:c:i:r:T:s:p:R:L:P:M:B:S:h")) != -1) {
    switch (c) {
      case 'A': matrixA = optarg; break;
      case 'b': vectorB = optarg; break;
      case 'l': lambda = std::atof(optarg); break;
      case 't': settings.totalThreads = std::atoi(optarg); break;
      case 'f': settings.lossFunction = std::atoi(optarg); break;
      case 'I': settings.innerIterations = std::strtoul(optarg, nullptr, 10); break;
      case 'c': settings.iters_bulkIterations_count = std::strtoul(optarg, nullptr, 10); break;
      case 'i': {
        const bool v = std::atoi(optarg) != 0;
        settings.showIntermediateObjectiveValue = v;
        settings.showInitialObjectiveValue = v;
        settings.showLastObjectiveValue = v;
        break;
      }
      case 'r': logPath = optarg; break;
      case 'T': settings.total_execution_time = std::atof(optarg); break;
      case 's': settings.hipSeed = std::strtoull(optarg, nullptr, 10); break;
      case 'p': settings.hipPrecision = std::atoi(optarg); break;
      case 'R': settings.hipRefreshPeriod = std::atoi(optarg); break;
      case 'L': settings.hipLanesPerCoord = std::atoi(optarg); break;
      case 'P': settings.hipLaunchMode = std::atoi(optarg); break;
      case 'M': settings.hipStepMode = std::atoi(optarg); break;
      case 'B': settings.hipBetaMode = std::atoi(optarg); break;
      case 'S':
        if (std::atoi(optarg) != 0) settings.hipBetaMode = ACDC_BETA_SPECTRAL;
        break;
      default: usage(); return c == 'h' ? 0 : 2;
    }
  }
  if (matrixA.empty()) {
    usage();
    return 2;
  }
  std::ofstream logFile;
  if (!logPath.empty()) {
    logFile.open(logPath.c_str());
    settings.logToFile = true;
    settings.logFile = &logFile;
  }
  try {
    ProblemData<int, double> instance;
    if (settings.lossFunction == ACDC_LOSS_SVM_DUAL_HINGE) {
      InputOuputHelper::loadLibSVMData(matrixA, instance);
      instance.lambda = lambda;
    } else {
      if (vectorB.empty()) throw std::runtime_error("-b <b file> is required for losses 0 and 10");
      InputOuputHelper::loadCSCData(matrixA, vectorB, lambda, instance);
    }
    std::cout << "Data loaded with size: " << instance.m << " x " << instance.n << std::endl;
    // omega: the most entries of one row (CSC) or one feature (SVM rows),
    // counted from 1 (MultiCoreSolver.cpp:24-38, :55-62)
    {
      const std::vector<int> &ix = settings.lossFunction == ACDC_LOSS_SVM_DUAL_HINGE
                                       ? instance.A_csr_col_idx
                                       : instance.A_csc_row_idx;
      std::vector<int> counts((size_t)instance.m, 0);
      double omega = 1;
      for (int i : ix)
        if (++counts[(size_t)i] > omega) omega = counts[(size_t)i];
      std::cout << "Omega: " << omega << std::endl;
      instance.sigma = 1 + (omega - 1) * (settings.totalThreads - 1) / (instance.n - 1.0);
    }
    HipEngineExecutor<int, double> executor(instance, &settings);
    std::cout << "ESO beta: " << instance.sigma << std::endl;
    Solver<int, double> solver(executor);
    solver.runSolver();
  } catch (const std::exception &e) {
    std::cerr << "acdc_hip_solver: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
