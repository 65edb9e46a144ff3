// Host-side checks of the C++ mirror that need no GPU: the text CSC and
// LIBSVM loaders, the CSV log rows.  Exit code 0 = pass.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "../../ac-dc_amd/host/acdc/file_reader.h"
#include "../../ac-dc_amd/host/acdc/losses.h"

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  {  // text CSC round trip; the last value has no trailing newline
    std::ofstream(dir + "/a_values.txt") << "1.5 -2\n0.25";
    std::ofstream(dir + "/a_colPtr.txt") << "0\n2\n3\n";
    std::ofstream(dir + "/a_rowIdx.txt") << "0 2 1";
    std::ofstream(dir + "/b.txt") << "1\n2\n3\n";
    ProblemData<int, double> p;
    InputOuputHelper::loadCSCData(dir + "/a", dir + "/b.txt", 0.5, p);
    CHECK(p.m == 3 && p.n == 2 && p.lambda == 0.5);
    CHECK(p.A_csc_values.size() == 3 && p.A_csc_values[2] == 0.25);
    CHECK(p.A_csc_row_idx[1] == 2 && p.A_csc_col_ptr[2] == 3);
    InputOuputHelper::saveCSCData(dir + "/c", dir + "/cb.txt", p);
    ProblemData<int, double> q;
    InputOuputHelper::loadCSCData(dir + "/c", dir + "/cb.txt", 0.5, q);
    CHECK(q.A_csc_values == p.A_csc_values && q.b == p.b);
    bool threw = false;
    try {
      InputOuputHelper::loadCSCData(dir + "/missing", dir + "/b.txt", 0.5, q);
    } catch (const std::exception &) {
      threw = true;
    }
    CHECK(threw);
  }
  {  // text column pointers + raw row ids and values of one rank
    const int rows[3] = {0, 2, 1};
    const double vals[3] = {1.5, -2.0, 0.25};
    std::ofstream(dir + "/m_colptr_3.bin") << "0 2 3\n";
    std::ofstream(dir + "/m_rowid_3.bin", std::ios::binary)
        .write(reinterpret_cast<const char *>(rows), sizeof(rows));
    std::ofstream(dir + "/m_values_3.bin", std::ios::binary)
        .write(reinterpret_cast<const char *>(vals), sizeof(vals));
    std::ofstream(dir + "/mb.txt") << "1\n2\n3\n";
    ProblemData<int, double> p;
    InputOuputHelper::loadAsciiAndBinaryCSC(dir + "/m", std::string(".bin"), dir + "/mb.txt", 0.5,
                                            2, 3, p);
    CHECK(p.m == 3 && p.n == 2 && p.lambda == 0.5);
    CHECK(p.A_csc_col_ptr[2] == 3 && p.A_csc_row_idx[1] == 2 && p.A_csc_values[2] == 0.25);
    bool threw = false;
    try {  // a short values file
      std::ofstream(dir + "/m_values_3.bin", std::ios::binary)
          .write(reinterpret_cast<const char *>(vals), sizeof(double));
      InputOuputHelper::loadAsciiAndBinaryCSC(dir + "/m", std::string(".bin"), dir + "/mb.txt",
                                              0.5, 2, 3, p);
    } catch (const std::exception &) {
      threw = true;
    }
    CHECK(threw);
  }
  {  // LIBSVM: 1-based features, label 0 -> -1
    std::ofstream(dir + "/svm.txt") << "+1 1:0.5 3:2\n0 2:1\n-1 4:-1.5 1:3\n";
    ProblemData<int, double> p;
    InputOuputHelper::loadLibSVMData(dir + "/svm.txt", p);
    CHECK(p.n == 3 && p.m == 4);
    CHECK(p.b[0] == 1 && p.b[1] == -1 && p.b[2] == -1);
    CHECK(p.A_csr_row_ptr.size() == 4 && p.A_csr_row_ptr[3] == 5);
    CHECK(p.A_csr_col_idx[0] == 0 && p.A_csr_col_idx[1] == 2 && p.A_csr_col_idx[3] == 3);
    CHECK(p.A_csr_values[4] == 3.0);
  }
  {  // CSV rows keep the reference's column order
    acdc_report r{};
    r.primal = 2.5;
    r.dual = 2.0;
    r.gap = 0.5;
    r.accuracy = 0.75;
    OptimizationStatistics st;
    st.elapsedIterations = 7;
    st.elapsedPureComputationTime = 1.0;
    st.elapsed_time = 2.0;
    CHECK(acdc_losses::log_row(ACDC_LOSS_L1_SQUARE, r, st) == "7,2.5,1,2,0.5");
    CHECK(acdc_losses::log_row(ACDC_LOSS_SVM_DUAL_HINGE, r, st) == "7,0.5,2.5,2,0.75,1,2,0");
    CHECK(acdc_losses::log_row(ACDC_LOSS_SVM_DUAL_HINGE, r, st, 0.5) == "7,0.5,2.5,2,0.75,1,2,0.5");
    CHECK(acdc_losses::log_headers(ACDC_LOSS_SVM_DUAL_HINGE).rfind("Elapsed iterations, Current duality gap", 0) == 0);
  }
  std::printf("host mirror loaders: ok\n");
  return 0;
}
