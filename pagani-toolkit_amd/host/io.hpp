// io.hpp -- file formats and directory conventions of the PAGANI command-line
// tools, kept byte-compatible with the reference:
//   .nii  NIfTI-1 single file, 352-byte header read as shorts
//         (dims at [21..24], datatype [35], bitpix [36]), data from byte 352
//         (U/CorMat/CUCorMat/CUCORMAT/CorMat.cpp:138-235)
//   .csr  int32 N+1, int32 R[N+1], int32 nnz, int32 C[nnz]
//         [, int32 nnz, float32 V[nnz]] (U CorMat.cpp:337-342, W CorMat.cpp:340-346)
//   .nm   int32 N, float32[N] (U/BC/CUBC/CUBC_main.cpp:94-100)
//   BNA_time_log  appended "tool\t<file>GPU\tkernel time\t<s>s" lines
//         (U/Degree/Degree/Degree.cpp:94)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pgio {

struct Csr {
    int32_t n = 0;
    std::vector<int32_t> row, col;
    std::vector<float> val;  // empty for binary graphs
};

// Files in `dir` whose extension is exactly `ext` (".nii", ".csr"), sorted,
// skipping `exclude` (the reference skips mask.nii).
std::vector<std::string> list_dir(const std::string &dir, const std::string &ext,
                                  const std::string &exclude = "");
std::string join(const std::string &dir, const std::string &name);
std::string strip_ext(const std::string &path);
bool make_dir(const std::string &dir);

// NIfTI mask: values converted to float (uint8 or float32 storage).  Returns
// false with a message for other datatypes (U CorMat.cpp:150-167).  hdr, when
// given, receives the 352-byte header as 176 shorts.
bool read_mask(const std::string &path, std::vector<float> &mask, int dims[3], std::string &err,
               short *hdr = nullptr);
// 4-D series gathered through the mask (mask >= thr), time-major T x N as the
// reference's BOLD[l*N+i]; float64 input is narrowed to float32.
bool read_bold(const std::string &path, const std::vector<float> &mask, float thr, int &n, int &t,
               std::vector<float> &bold, std::string &err);

bool read_csr(const std::string &path, bool weighted, Csr &g, std::string &err);
bool write_csr(const std::string &path, const int32_t *row, int32_t n, const int32_t *col,
               int64_t nnz, const float *val);
bool write_nm(const std::string &path, const float *x, int32_t n);
bool read_nm(const std::string &path, std::vector<float> &x);
void time_log(const std::string &line);

}  // namespace pgio
