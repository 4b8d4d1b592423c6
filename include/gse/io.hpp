// io.hpp -- the text formats and error metric of the reference's driver
// program (cpu_gse.cpp), for callers of include/gse/gse.hpp.
//
//   loadArray<T>(width, numLine, filename, array)  cpu_gse.cpp:152-171
//       numLine lines of `width` whitespace-separated values (xyzq files:
//       width 4, x y z q per line; exclusion files: width 1)
//   saveForce<T>(filename, numCoord, fx, fy, fz)    cpu_gse.cpp:173-183
//       one "fx fy fz" line per particle, default stream formatting
//   readForce<T>(filename, numCoord, fx, fy, fz)    cpu_gse.cpp:185-195
//   saveMatrix<T>(filename, L, sizeX, sizeY)        cpu_gse.cpp:197-211
//   Erms(n, refx, refy, refz, x, y, z)              cpu_gse.cpp:131-146
//
// The formats are kept so that files written by either program read in the
// other.  Failures throw gse_amd::Error where the reference prints and
// calls exit(1).
#ifndef GSE_AMD_IO_HPP
#define GSE_AMD_IO_HPP

#include <cmath>
#include <fstream>
#include <string>

#include "gse.hpp"

namespace gse_amd {

template <typename T>
void loadArray(const int width, const int numLine, const char* filename, T* array) {
  std::ifstream file(filename);
  if (!file.is_open()) throw Error(GSE_ERROR_INVALID_VALUE, std::string("cannot open ") + filename);
  for (int i = 0; i < numLine; i++)
    for (int k = 0; k < width; k++)
      if (!(file >> array[i * width + k]))
        throw Error(GSE_ERROR_INVALID_VALUE, std::string("error reading ") + filename + " at line " +
                                                 std::to_string(i + 1));
}

template <typename T>
void saveForce(const char* filename, const int numCoord, const T* forceX, const T* forceY,
               const T* forceZ) {
  std::ofstream file(filename);
  if (!file.is_open()) throw Error(GSE_ERROR_INVALID_VALUE, std::string("cannot open ") + filename);
  for (int i = 0; i < numCoord; i++)
    file << forceX[i] << " " << forceY[i] << " " << forceZ[i] << std::endl;
}

template <typename T>
void readForce(const char* filename, const int numCoord, T* forceX, T* forceY, T* forceZ) {
  std::ifstream file(filename);
  if (!file.is_open()) throw Error(GSE_ERROR_INVALID_VALUE, std::string("cannot open ") + filename);
  for (int i = 0; i < numCoord; i++)
    if (!(file >> forceX[i] >> forceY[i] >> forceZ[i]))
      throw Error(GSE_ERROR_INVALID_VALUE, std::string("error reading ") + filename);
}

template <typename T>
void saveMatrix(const char* filename, const T* L, const int sizeX, const int sizeY) {
  std::ofstream file(filename);
  if (!file.is_open()) throw Error(GSE_ERROR_INVALID_VALUE, std::string("cannot open ") + filename);
  for (int y = 0; y < sizeY; y++) {
    for (int x = 0; x < sizeX; x++) file << L[x + y * sizeX] << " ";
    file << std::endl;
  }
}

// relative force RMS error sqrt(sum |F - F_ref|^2 / sum |F_ref|^2)
inline double Erms(const int n, const double* refx, const double* refy, const double* refz,
                   const double* x, const double* y, const double* z) {
  double rms = 0.0, ref = 0.0;
  for (int i = 0; i < n; i++) {
    const double dx = x[i] - refx[i], dy = y[i] - refy[i], dz = z[i] - refz[i];
    rms += dx * dx + dy * dy + dz * dz;
    ref += refx[i] * refx[i] + refy[i] * refy[i] + refz[i] * refz[i];
  }
  return std::sqrt(rms / ref);
}

}  // namespace gse_amd

#endif  // GSE_AMD_IO_HPP
