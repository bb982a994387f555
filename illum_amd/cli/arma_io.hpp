// Armadillo binary file formats, as the reference writes them through
// arma::mat::save / arma::sp_mat::save (arma_util.hpp:10-13, illum_cli.cpp:281-283,
// 474-507) and as src/py/arma.py:4-35 reads the dense kind:
//   dense : "ARMA_MAT_BIN_FN008\n" "<rows> <cols>\n" then rows*cols FP64, column-major
//   sparse: "ARMA_SPM_BIN_FN008\n" "<rows> <cols> <nnz>\n" then nnz FP64 values,
//           nnz uint64 row indices, cols+1 uint64 column pointers (CSC)
// Plain raw-ASCII matrices (sun position files, illum_cli.cpp:329) are read by
// load_raw_ascii.
#pragma once

#include <cstdint>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace illum_cli {

inline void save_dense(const std::string &path, const std::vector<double> &colmajor, int64_t rows,
                       int64_t cols) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  f << "ARMA_MAT_BIN_FN008\n" << rows << ' ' << cols << '\n';
  f.write(reinterpret_cast<const char *>(colmajor.data()), (std::streamsize)(8 * rows * cols));
}

inline std::vector<double> load_dense(const std::string &path, int64_t *rows, int64_t *cols) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot read " + path);
  std::string head;
  std::getline(f, head);
  if (head != "ARMA_MAT_BIN_FN008") throw std::runtime_error(path + ": not an FP64 Armadillo matrix");
  f >> *rows >> *cols;
  f.get();
  std::vector<double> v((size_t)(*rows * *cols));
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(8 * v.size()));
  if (!f) throw std::runtime_error(path + ": truncated");
  return v;
}

inline void save_sparse(const std::string &path, int64_t rows, int64_t cols,
                        const std::vector<int64_t> &colptr, const std::vector<int32_t> &rowidx,
                        const std::vector<double> &val) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  const int64_t nnz = (int64_t)val.size();
  f << "ARMA_SPM_BIN_FN008\n" << rows << ' ' << cols << ' ' << nnz << '\n';
  f.write(reinterpret_cast<const char *>(val.data()), (std::streamsize)(8 * nnz));
  std::vector<uint64_t> ri(rowidx.begin(), rowidx.end());
  std::vector<uint64_t> cp(colptr.begin(), colptr.end());
  f.write(reinterpret_cast<const char *>(ri.data()), (std::streamsize)(8 * nnz));
  f.write(reinterpret_cast<const char *>(cp.data()), (std::streamsize)(8 * cp.size()));
}

inline void load_sparse(const std::string &path, int64_t *rows, int64_t *cols,
                        std::vector<int64_t> &colptr, std::vector<int32_t> &rowidx,
                        std::vector<double> &val) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot read " + path);
  std::string head;
  std::getline(f, head);
  if (head != "ARMA_SPM_BIN_FN008") throw std::runtime_error(path + ": not an FP64 Armadillo sparse matrix");
  int64_t nnz;
  f >> *rows >> *cols >> nnz;
  f.get();
  val.resize(nnz);
  std::vector<uint64_t> ri(nnz), cp(*cols + 1);
  f.read(reinterpret_cast<char *>(val.data()), (std::streamsize)(8 * nnz));
  f.read(reinterpret_cast<char *>(ri.data()), (std::streamsize)(8 * nnz));
  f.read(reinterpret_cast<char *>(cp.data()), (std::streamsize)(8 * cp.size()));
  if (!f) throw std::runtime_error(path + ": truncated");
  rowidx.assign(ri.begin(), ri.end());
  colptr.assign(cp.begin(), cp.end());
}

// Whitespace-separated numbers, one matrix row per line.
inline std::vector<std::vector<double>> load_raw_ascii(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot read " + path);
  std::vector<std::vector<double>> rows;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::vector<double> r;
    double x;
    while (ss >> x) r.push_back(x);
    if (!r.empty()) rows.push_back(r);
  }
  return rows;
}

}  // namespace illum_cli
