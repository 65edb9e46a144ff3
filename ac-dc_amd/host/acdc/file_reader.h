// file_reader.h -- the on-disk formats of the path.
//
// Text CSC (InputOuputHelper::loadCSCData, cpp/src/utils/file_reader.h:108-137):
//   <prefix>_values.txt, <prefix>_colPtr.txt, <prefix>_rowIdx.txt and a b file,
//   whitespace-separated; n = |colPtr| - 1, m = |b|.  Unlike the reference's
//   loadVectorData (:33-48), a final value without a trailing newline is kept,
//   and a missing file is an error rather than an empty vector.
// Text + binary CSC of one rank (loadANSCIAndBinaryData, file_reader.h:50-106):
//   <prefix>_colptr_<rank><suffix> holds n+1 whitespace-separated column
//   pointers as text, <prefix>_rowid_<rank><suffix> the row indices as raw
//   L, <prefix>_values_<rank><suffix> the values as raw D (host byte order);
//   b from a text file.  Short or missing files are errors.
// Dense CSV (readCSVFile, parse_data_size_from_CSV_file and
//   parse_data_from_CSV_file, cpp/src/utils/file_reader.h:138-194): m lines of
//   n comma-separated numbers, n read from the first line; stored
//   column-major (entry (row, col) at row + col m).  Unlike the reference, a
//   short or malformed line is an error.
// LIBSVM rows (loadDistributedSparseSVMRowData,
//   cpp/src/utils/distributed_instances_loader.h:68-193): "label idx:val ..."
//   per sample, 1-based feature indices shifted to 0-based, label 0 read as -1;
//   m = largest feature index, n = samples.  Unlike the reference, labels >= 100
//   and malformed tokens are errors instead of silently skipped rows.
#pragma once
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "structures.h"

namespace InputOuputHelper {

template <typename G>
void loadVectorData(const std::string &fileLocation, std::vector<G> &vector) {
  std::ifstream f(fileLocation.c_str());
  if (!f.is_open()) throw std::runtime_error("file not found: " + fileLocation);
  vector.clear();
  G val;
  while (f >> val) vector.push_back(val);
  if (!f.eof()) throw std::runtime_error("malformed number in " + fileLocation);
}

template <typename L, typename D>
void loadCSCData(const std::string &matrixPrefix, const std::string &vectorBFileName,
                 double lambda, ProblemData<L, D> &instance) {
  loadVectorData(vectorBFileName, instance.b);
  loadVectorData(matrixPrefix + "_values.txt", instance.A_csc_values);
  loadVectorData(matrixPrefix + "_colPtr.txt", instance.A_csc_col_ptr);
  loadVectorData(matrixPrefix + "_rowIdx.txt", instance.A_csc_row_idx);
  if (instance.A_csc_col_ptr.empty()) throw std::runtime_error("empty column pointer file");
  instance.n = (L)(instance.A_csc_col_ptr.size() - 1);
  instance.m = (L)instance.b.size();
  instance.lambda = (D)lambda;
}

template <typename L, typename D>
void loadAsciiAndBinaryCSC(const std::string &matrixPrefix, const std::string &matrixSuffix,
                           const std::string &vectorBFileName, double lambda, L local_n,
                           int rank, ProblemData<L, D> &instance) {
  loadVectorData(vectorBFileName, instance.b);
  const std::string tag = "_" + std::to_string(rank) + matrixSuffix;
  {
    const std::string path = matrixPrefix + "_colptr" + tag;
    std::ifstream f(path.c_str());
    if (!f.is_open()) throw std::runtime_error("file not found: " + path);
    instance.A_csc_col_ptr.resize((size_t)local_n + 1);
    for (auto &e : instance.A_csc_col_ptr)
      if (!(f >> e)) throw std::runtime_error("short column pointer file: " + path);
  }
  const size_t nnz = (size_t)instance.A_csc_col_ptr[(size_t)local_n];
  auto raw = [&](const std::string &name, auto &v) {
    const std::string path = matrixPrefix + name + tag;
    std::ifstream f(path.c_str(), std::ios::in | std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("file not found: " + path);
    v.resize(nnz);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(nnz * sizeof(v[0])));
    if ((size_t)f.gcount() != nnz * sizeof(v[0])) throw std::runtime_error("short binary file: " + path);
  };
  raw("_rowid", instance.A_csc_row_idx);
  raw("_values", instance.A_csc_values);
  instance.n = local_n;
  instance.m = (L)instance.b.size();
  instance.lambda = (D)lambda;
}

template <typename L, typename D>
void saveCSCData(const std::string &matrixPrefix, const std::string &vectorBFileName,
                 const ProblemData<L, D> &instance) {
  auto dump = [](const std::string &path, const auto &v) {
    std::ofstream f(path.c_str());
    if (!f.is_open()) throw std::runtime_error("cannot write " + path);
    f.precision(17);
    for (const auto &e : v) f << e << "\n";
  };
  dump(vectorBFileName, instance.b);
  dump(matrixPrefix + "_values.txt", instance.A_csc_values);
  dump(matrixPrefix + "_colPtr.txt", instance.A_csc_col_ptr);
  dump(matrixPrefix + "_rowIdx.txt", instance.A_csc_row_idx);
}

template <typename L, typename D>
void loadLibSVMData(const std::string &file, ProblemData<L, D> &part) {
  std::ifstream f(file.c_str());
  if (!f.is_open()) throw std::runtime_error("file not found: " + file);
  part.A_csr_row_ptr.assign(1, 0);
  part.A_csr_col_idx.clear();
  part.A_csr_values.clear();
  part.b.clear();
  long long maxFeature = 0;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string tok;
    if (!(ls >> tok)) continue;  // blank line
    int label = 0;
    if (std::sscanf(tok.c_str(), "%d", &label) != 1 || label >= 100)
      throw std::runtime_error("bad label '" + tok + "' in " + file);
    part.b.push_back(label == 0 ? (D)-1 : (D)(label > 0 ? 1 : -1));
    while (ls >> tok) {
      long long idx = 0;
      double v = 0;
      if (std::sscanf(tok.c_str(), "%lld:%lf", &idx, &v) != 2 || idx < 1)
        throw std::runtime_error("bad feature '" + tok + "' in " + file);
      part.A_csr_col_idx.push_back((L)(idx - 1));
      part.A_csr_values.push_back((D)v);
      if (idx > maxFeature) maxFeature = idx;
    }
    part.A_csr_row_ptr.push_back((L)part.A_csr_col_idx.size());
  }
  part.n = (L)part.b.size();
  part.m = (L)maxFeature;
}

// Dense CSV matrix into column-major `data` (m rows, n columns).
template <typename F>
void readCSVFile(std::vector<F> &data, unsigned int &ld, unsigned int &m, unsigned int &n,
                 const std::string &file) {
  std::ifstream f(file.c_str());
  if (!f.is_open()) throw std::runtime_error("file not found: " + file);
  std::vector<std::vector<F>> rows;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line == "\r") continue;
    std::vector<F> row;
    std::stringstream ls(line);
    std::string cell;
    while (std::getline(ls, cell, ',')) {
      std::stringstream cs(cell);
      double v = 0;
      if (!(cs >> v)) throw std::runtime_error("malformed number '" + cell + "' in " + file);
      row.push_back((F)v);
    }
    if (!rows.empty() && row.size() != rows[0].size())
      throw std::runtime_error("rows of different lengths in " + file);
    rows.push_back(std::move(row));
  }
  m = (unsigned int)rows.size();
  n = m ? (unsigned int)rows[0].size() : 0u;
  ld = m;
  data.assign((size_t)m * n, (F)0);
  for (unsigned int r = 0; r < m; ++r)
    for (unsigned int c = 0; c < n; ++c) data[r + (size_t)c * m] = rows[r][c];
}

// A dense CSV matrix as a CSC instance (its nonzeros) and a CSV b, as the
// reference's lambda-path experiment loads them (lasso_console_solver.cpp:76-112).
template <typename L, typename D>
void loadDenseCSV(const std::string &matrixFile, const std::string &bFile,
                  ProblemData<L, D> &instance) {
  std::vector<D> a;
  unsigned int ld = 0, m = 0, n = 0;
  readCSVFile(a, ld, m, n, matrixFile);
  instance.m = (L)m;
  instance.n = (L)n;
  instance.A_csc_col_ptr.assign(1, 0);
  instance.A_csc_row_idx.clear();
  instance.A_csc_values.clear();
  for (unsigned int c = 0; c < n; ++c) {
    for (unsigned int r = 0; r < m; ++r) {
      const D v = a[r + (size_t)c * m];
      if (v != 0) {
        instance.A_csc_row_idx.push_back((L)r);
        instance.A_csc_values.push_back(v);
      }
    }
    instance.A_csc_col_ptr.push_back((L)instance.A_csc_row_idx.size());
  }
  std::vector<D> b;
  unsigned int bm = 0, bn = 0;
  readCSVFile(b, ld, bm, bn, bFile);
  if (bm != m || bn != 1) throw std::runtime_error("b must be an m x 1 CSV matching A: " + bFile);
  instance.b = b;
}

}  // namespace InputOuputHelper
