// io.cpp -- the reference's matrix file formats (host code).
//
// Binary matrix (Utils/matrixIO.h:18-35, Utils/matrixIO.inl:4-110): int32 rows,
// int32 cols, then rows*cols elements row-major, native byte order, element
// type chosen by the caller (the reference's phase maps are float32).
// Complex text (Utils/Utils.cpp:80-166 and 196-281 read, 168-193 and 282-307
// write): a first line "rows,cols", then one "re,im" line per element in
// row-major order.
#include "../../include/glsu.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace glsu {
void set_error(const char* msg);
}

static int io_fail(int code, const std::string& msg) {
  glsu::set_error(msg.c_str());
  return code;
}

extern "C" int glsu_read_matrix_bin(const char* path, int elem_size, int* rows, int* cols,
                                    void* data, size_t capacity) {
  if (!path || !rows || !cols || elem_size <= 0) return io_fail(GLSU_ERR_ARG, "invalid argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(GLSU_ERR_ARG, std::string("can't open input matrix file: ") + path);
  int hdr[2];
  if (std::fread(hdr, sizeof(int), 2, f) != 2 || hdr[0] < 0 || hdr[1] < 0) {
    std::fclose(f);
    return io_fail(GLSU_ERR_ARG, std::string("error reading matrix header: ") + path);
  }
  *rows = hdr[0];
  *cols = hdr[1];
  if (!data) {  // size query
    std::fclose(f);
    return GLSU_OK;
  }
  const size_t n = (size_t)hdr[0] * hdr[1];
  if (n * (size_t)elem_size > capacity) {
    std::fclose(f);
    return io_fail(GLSU_ERR_SIZE, "buffer too small for the matrix");
  }
  const size_t got = std::fread(data, (size_t)elem_size, n, f);
  std::fclose(f);
  if (got < n)
    return io_fail(GLSU_ERR_ARG, std::string("matrix data shorter than its header says: ") + path);
  return GLSU_OK;
}

extern "C" int glsu_write_matrix_bin(const char* path, int elem_size, int rows, int cols,
                                     const void* data) {
  if (!path || !data || elem_size <= 0 || rows < 0 || cols < 0)
    return io_fail(GLSU_ERR_ARG, "invalid argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return io_fail(GLSU_ERR_ARG, std::string("can't open output file: ") + path);
  const int hdr[2] = {rows, cols};
  const size_t n = (size_t)rows * cols;
  const bool ok = std::fwrite(hdr, sizeof(int), 2, f) == 2 &&
                  std::fwrite(data, (size_t)elem_size, n, f) == n;
  std::fclose(f);
  return ok ? GLSU_OK : io_fail(GLSU_ERR_ARG, std::string("error writing ") + path);
}

extern "C" int glsu_read_complex_text(const char* path, int* rows, int* cols, double* data,
                                      size_t capacity) {
  if (!path || !rows || !cols) return io_fail(GLSU_ERR_ARG, "invalid argument");
  std::ifstream in(path);
  if (!in) return io_fail(GLSU_ERR_ARG, std::string("can't open ") + path);
  std::string line;
  if (!std::getline(in, line)) return io_fail(GLSU_ERR_ARG, "empty file");
  {
    std::vector<long> size;
    std::stringstream ls(line);
    std::string cell;
    while (std::getline(ls, cell, ',')) size.push_back(std::strtol(cell.c_str(), nullptr, 10));
    if (size.size() != 2 || size[0] < 0 || size[1] < 0)
      return io_fail(GLSU_ERR_ARG, "first line must be rows,cols");
    *rows = (int)size[0];
    *cols = (int)size[1];
  }
  if (!data) return GLSU_OK;  // size query
  const size_t n = (size_t)*rows * *cols;
  if (2 * n > capacity) return io_fail(GLSU_ERR_SIZE, "buffer too small for the matrix");
  size_t i = 0;
  while (i < n && std::getline(in, line)) {
    std::stringstream ls(line);
    std::string re, im;
    if (!std::getline(ls, re, ',') || !std::getline(ls, im, ','))
      return io_fail(GLSU_ERR_ARG, "element lines must be re,im");
    data[2 * i] = std::strtod(re.c_str(), nullptr);
    data[2 * i + 1] = std::strtod(im.c_str(), nullptr);
    ++i;
  }
  if (i < n) return io_fail(GLSU_ERR_ARG, "fewer elements than the header says");
  return GLSU_OK;
}

extern "C" int glsu_write_complex_text(const char* path, int rows, int cols, const double* data,
                                       int precision) {
  if (!path || !data || rows < 0 || cols < 0) return io_fail(GLSU_ERR_ARG, "invalid argument");
  std::ofstream out(path);
  if (!out) return io_fail(GLSU_ERR_ARG, std::string("can't open output file: ") + path);
  if (precision > 0) out.precision(precision);  // the reference uses the stream default (6)
  out << rows << "," << cols << "\n";
  const size_t n = (size_t)rows * cols;
  for (size_t i = 0; i < n; ++i) out << data[2 * i] << "," << data[2 * i + 1] << "\n";
  return out ? GLSU_OK : io_fail(GLSU_ERR_ARG, std::string("error writing ") + path);
}
