// gse.hpp -- C++ drop-in classes for the GSE k-space path of ap-hynninen/GSE,
// running on one MI355X through the C-ABI of include/gse_hip.h.
//
// The classes keep the reference's names, template parameters, constructor
// arguments and method signatures, so a caller switches by including this
// header and naming namespace gse_amd:
//   CpuGrid<T>              CpuGrid.h:17-106, Grid.h:12-53
//   CpuGaussCharge<AT,CT>   CpuGaussCharge.h:11-49
//   CpuConvolution<T>       CpuConvolution.h:9-38
//   CpuGreensFunc<T>        CpuGreensFunc.h:9-19 (GSE: CpuGreensFuncGSE.h,
//                           G2: CpuGreensFuncG2.h)
//   CpuFFTSolver<T>         CpuFFTSolver.h:9-43
//   CpuGSE<AT,CT>, CpuGSEk<AT,CT>   CpuGSE.h:18-68, CpuGSEk.h:18-37
//   xyzq_t<T>               TypesGSE.h:20-22
// and the reference's GPU classes, over device memory:
//   CudaGrid<T>             CudaGrid.h:14-42
//   CudaGaussCharge<AT,CT>  CudaGaussCharge.h:11-33
//   CudaConvolution<T>      CudaConvolution.h:9-37
// and the comparators the reference's driver runs beside GSE:
//   CpuEwaldDirect          CpuEwaldDirect.h:10-44 (host arrays)
//   CpuEwaldRecip           CpuEwaldRecip.h:12-46 (host arrays)
//   CudaPMERecip<AT,CT>     cuda/CudaPMERecip.h:36-182 (device arrays)
//
// Memory model.  As in the reference, the Cpu* classes work on HOST memory
// owned by the caller (CpuGrid allocates with new[]).  Each call
// copies its inputs to device buffers that the objects keep between calls,
// runs the HIP kernels and copies its outputs back; the calls therefore
// include PCIe transfers and return with the results in host memory.
// The Cuda* classes work on DEVICE memory as the reference's do: CudaGrid
// owns a device array, CudaGaussCharge reads device particles, and results
// stay on the device until CpuGrid::copy(const CudaGrid&) brings them back.
// Their calls are ordered on the null stream and return when the device work
// is done, like the reference's synchronous CUDA calls.  Callers that manage
// their own streams call include/gse_hip.h directly with device pointers.
//
// Differences from the reference, all on error paths or outside the path:
//   * failures throw gse_amd::Error (status + message) instead of assert()
//     or exit(1); FP32 CpuFFTSolver, which exits in the reference
//     (CpuFFTSolver.cpp:75-95), runs;
//   * getRhoS()/getPhi() return host copies refreshed from the device at each
//     call (writes into them are not seen by later device steps);
//   * CpuGrid::save throws when the file cannot be opened.
#ifndef GSE_AMD_GSE_HPP
#define GSE_AMD_GSE_HPP

#include <cmath>
#include <cstddef>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gse_hip.h"

namespace gse_amd {

class Error : public std::runtime_error {
 public:
  Error(gse_status_t s, const std::string& msg)
      : std::runtime_error(std::string("gse: ") + gse_status_string(s) + ": " + msg), status(s) {}
  gse_status_t status;
};

inline void check(gse_status_t s) {
  if (s != GSE_SUCCESS) throw Error(s, gse_last_error());
}

template <typename T>
struct dtype_of;
template <>
struct dtype_of<float> {
  static constexpr gse_dtype_t value = GSE_FLOAT32;
};
template <>
struct dtype_of<double> {
  static constexpr gse_dtype_t value = GSE_FLOAT64;
};

// TypesGSE.h:20-22
template <typename T>
struct xyzq_t {
  T x, y, z, q;
};

const double pi_dbl = 3.14159265358979323846;

// Growable device allocation owned by one object.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { (void)gse_device_free(p_); }
  void* get(size_t bytes) {
    if (bytes > n_ || !p_) {
      check(gse_device_free(p_));
      p_ = nullptr;
      n_ = 0;
      check(gse_device_malloc(&p_, bytes ? bytes : 1));
      n_ = bytes;
    }
    return p_;
  }
  void* ptr() const { return p_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Grid.h:12-53
template <typename T>
class Grid {
 protected:
  const int sizeX;
  const int sizeY;
  const int sizeZ;
  const int size;
  T* data;

 public:
  Grid(const int sizeX, const int sizeY, const int sizeZ)
      : sizeX(sizeX), sizeY(sizeY), sizeZ(sizeZ), size(sizeX * sizeY * sizeZ), data(nullptr) {}
  T* getDataPointer() { return data; }
  const T* getDataPointer() const { return data; }
  int getPos(const int x, const int y, const int z) const {
    const int xt = (x + sizeX) % sizeX;
    const int yt = (y + sizeY) % sizeY;
    const int zt = (z + sizeZ) % sizeZ;
    return xt + (yt + zt * sizeY) * sizeX;
  }
  int getSizeX() const { return sizeX; }
  int getSizeY() const { return sizeY; }
  int getSizeZ() const { return sizeZ; }
  int getSize() const { return size; }
};

template <typename T>
class CudaGrid;

// CpuGrid.h:17-106: host grid, x fastest.  The device mirror is a staging
// buffer used by the classes below.
template <typename T>
class CpuGrid : public Grid<T> {
 public:
  CpuGrid(const int sizeX, const int sizeY, const int sizeZ) : Grid<T>(sizeX, sizeY, sizeZ) {
    this->data = new T[this->size]();
  }
  ~CpuGrid() { delete[] this->data; }
  CpuGrid(const CpuGrid&) = delete;
  CpuGrid& operator=(const CpuGrid&) = delete;

  void clear() {
    for (int i = 0; i < this->size; i++) this->data[i] = (T)0;
  }
  void add(const Grid<T>& grid) {
    const T* g = grid.getDataPointer();
    for (int i = 0; i < this->size; i++) this->data[i] += g[i];
  }
  template <typename T2>
  void sub(const Grid<T2>& grid) {
    const T2* g = grid.getDataPointer();
    for (int i = 0; i < this->size; i++) this->data[i] -= (T)g[i];
  }
  template <typename T2>
  void copy(const CpuGrid<T2>& grid) {
    const T2* g = grid.getDataPointer();
    for (int i = 0; i < this->size; i++) this->data[i] = (T)g[i];
  }
  // CpuGrid.h:47-50: device -> host, synchronous
  void copy(const CudaGrid<T>& d_grid);
  void scale(const T val) {
    for (int i = 0; i < this->size; i++) this->data[i] *= val;
  }
  T sum() const {
    T r = (T)0;
    for (int i = 0; i < this->size; i++) r += this->data[i];
    return r;
  }
  T maxAbsValue() const {
    T m = (T)0;
    for (int i = 0; i < this->size; i++) {
      const T v = std::fabs(this->data[i]);
      m = m < v ? v : m;
    }
    return m;
  }
  T getDataValue(const int x, const int y, const int z) const {
    return this->data[this->getPos(x, y, z)];
  }
  void save(const char* filename) const {
    std::ofstream file(filename);
    if (!file.is_open())
      throw Error(GSE_ERROR_INVALID_VALUE, std::string("cannot open ") + filename);
    for (int z = 0; z < this->sizeZ; z++)
      for (int y = 0; y < this->sizeY; y++)
        for (int x = 0; x < this->sizeX; x++) file << getDataValue(x, y, z) << std::endl;
  }

  // --- device staging (not in the reference) ---
  size_t bytes() const { return sizeof(T) * (size_t)this->size; }
  void* deviceBuffer() const { return dev_.get(bytes()); }
  const void* toDevice() const {
    void* d = dev_.get(bytes());
    check(gse_memcpy_to_device(d, this->data, bytes()));
    return d;
  }
  void fromDevice() { check(gse_memcpy_to_host(this->data, dev_.ptr(), bytes())); }

 private:
  mutable DeviceBuffer dev_;
};

namespace detail {
template <typename CT>
const void* upload_xyzq(DeviceBuffer& buf, int n, const xyzq_t<CT>* xyzq) {
  const size_t bytes = sizeof(xyzq_t<CT>) * (size_t)(n > 0 ? n : 0);
  void* d = buf.get(bytes);
  if (n > 0) check(gse_memcpy_to_device(d, xyzq, bytes));
  return d;
}
template <typename AT>
void download(AT* dst, const void* src, int n) {
  if (n > 0) check(gse_memcpy_to_host(dst, src, sizeof(AT) * (size_t)n));
}
template <typename G>
void check_dims(const G& g, int X, int Y, int Z) {
  if (g.getSizeX() != X || g.getSizeY() != Y || g.getSizeZ() != Z)
    throw Error(GSE_ERROR_INVALID_VALUE, "grid size does not match the object's grid");
}
}  // namespace detail

// CpuGaussCharge.h:11-49
template <typename AT, typename CT>
class CpuGaussCharge {
 public:
  CpuGaussCharge(const int sizeX, const int sizeY, const int sizeZ)
      : sizeX(sizeX), sizeY(sizeY), sizeZ(sizeZ) {
    check(gse_gauss_charge_create(&h_, dtype_of<AT>::value, dtype_of<CT>::value, sizeX, sizeY,
                                  sizeZ));
  }
  ~CpuGaussCharge() { (void)gse_gauss_charge_destroy(h_); }
  CpuGaussCharge(const CpuGaussCharge&) = delete;
  CpuGaussCharge& operator=(const CpuGaussCharge&) = delete;

  int calcSupportSize(const double rcut, const double boxx, const double boxy, const double boxz) {
    int n = 0;
    check(gse_gauss_charge_calc_support_size(h_, rcut, boxx, boxy, boxz, &n));
    return n;
  }

  void calcElectricFieldOnGrid(const double boxx, const double boxy, const double boxz,
                               const CpuGrid<CT>& phi, CpuGrid<CT>& Ex, CpuGrid<CT>& Ey,
                               CpuGrid<CT>& Ez) {
    grids_ok(phi), grids_ok(Ex), grids_ok(Ey), grids_ok(Ez);
    check(gse_gauss_charge_calc_electric_field_on_grid(h_, boxx, boxy, boxz, phi.toDevice(),
                                                       Ex.deviceBuffer(), Ey.deviceBuffer(),
                                                       Ez.deviceBuffer()));
    Ex.fromDevice();
    Ey.fromDevice();
    Ez.fromDevice();
  }

  void interpolateForce(const double sigma, const double rcut, const int numCoord,
                        const xyzq_t<CT>* xyzq, const double boxx, const double boxy,
                        const double boxz, const CpuGrid<CT>& phi, AT* forceX, AT* forceY,
                        AT* forceZ) {
    grids_ok(phi);
    const void* dx = detail::upload_xyzq(xyzq_, numCoord, xyzq);
    const size_t fb = sizeof(AT) * (size_t)(numCoord > 0 ? numCoord : 0);
    void* f = force_.get(3 * fb);
    char* fc = (char*)f;
    check(gse_gauss_charge_interpolate_force(h_, sigma, rcut, numCoord, dx, boxx, boxy, boxz,
                                             phi.toDevice(), fc, fc + fb, fc + 2 * fb));
    detail::download(forceX, fc, numCoord);
    detail::download(forceY, fc + fb, numCoord);
    detail::download(forceZ, fc + 2 * fb, numCoord);
  }

  void calcDipoleSum(const double boxx, const double boxy, const double boxz,
                     const CpuGrid<CT>& rho, AT& dipSumX, AT& dipSumY, AT& dipSumZ) {
    grids_ok(rho);
    double d[3];
    check(gse_gauss_charge_calc_dipole_sum(h_, boxx, boxy, boxz, rho.toDevice(), d));
    dipSumX = (AT)d[0];
    dipSumY = (AT)d[1];
    dipSumZ = (AT)d[2];
  }

  void spreadChargeToGrid(const double sigma, const double rcut, const int numCoord,
                          const xyzq_t<CT>* xyzq, const double boxx, const double boxy,
                          const double boxz, CpuGrid<CT>& rho, const bool normalize = false) {
    grids_ok(rho);
    const void* dx = detail::upload_xyzq(xyzq_, numCoord, xyzq);
    check(gse_gauss_charge_spread_charge_to_grid(h_, sigma, rcut, numCoord, dx, boxx, boxy, boxz,
                                                 rho.deviceBuffer(), normalize ? 1 : 0));
    rho.fromDevice();
  }

  void spreadChargeOnGrid(const double sigma, const double rcut, const double boxx,
                          const double boxy, const double boxz, const CpuGrid<CT>& rho,
                          CpuGrid<CT>& rhoS) {
    grids_ok(rho), grids_ok(rhoS);
    check(gse_gauss_charge_spread_charge_on_grid(h_, sigma, rcut, boxx, boxy, boxz,
                                                 rho.toDevice(), rhoS.deviceBuffer()));
    rhoS.fromDevice();
  }

  gse_gauss_charge_t handle() const { return h_; }

 private:
  void grids_ok(const Grid<CT>& g) const { detail::check_dims(g, sizeX, sizeY, sizeZ); }
  const int sizeX, sizeY, sizeZ;
  gse_gauss_charge_t h_ = nullptr;
  DeviceBuffer xyzq_, force_;
};

// CpuConvolution.h:9-38
template <typename T>
class CpuConvolution {
 public:
  CpuConvolution(const int sizeX, const int sizeY, const int sizeZ)
      : sizeX(sizeX), sizeY(sizeY), sizeZ(sizeZ) {
    check(gse_convolution_create(&h_, dtype_of<T>::value, sizeX, sizeY, sizeZ));
  }
  ~CpuConvolution() { (void)gse_convolution_destroy(h_); }
  CpuConvolution(const CpuConvolution&) = delete;
  CpuConvolution& operator=(const CpuConvolution&) = delete;

  void setupKernel(const int kernelRadiusX, const int kernelRadiusY, const int kernelRadiusZ,
                   const T* h_kernelX, const T* h_kernelY, const T* h_kernelZ) {
    check(gse_convolution_setup_kernel(h_, kernelRadiusX, kernelRadiusY, kernelRadiusZ,
                                       h_kernelX, h_kernelY, h_kernelZ));
  }
  // X pass only, as CpuConvolution.cpp:116-151
  void run(const CpuGrid<T>& data, CpuGrid<T>& dataXYZ) {
    detail::check_dims(data, sizeX, sizeY, sizeZ);
    detail::check_dims(dataXYZ, sizeX, sizeY, sizeZ);
    check(gse_convolution_run(h_, data.toDevice(), dataXYZ.deviceBuffer()));
    dataXYZ.fromDevice();
  }
  // all three passes (not in the reference class)
  void runXYZ(const CpuGrid<T>& data, CpuGrid<T>& dataXYZ) {
    detail::check_dims(data, sizeX, sizeY, sizeZ);
    detail::check_dims(dataXYZ, sizeX, sizeY, sizeZ);
    check(gse_convolution_run_xyz(h_, data.toDevice(), dataXYZ.deviceBuffer()));
    dataXYZ.fromDevice();
  }

 private:
  const int sizeX, sizeY, sizeZ;
  gse_convolution_t h_ = nullptr;
};

// CpuGreensFunc.h:9-19.  mul() works on a host half spectrum of
// ((nfftX/2+1)*2, nfftY, nfftZ) values, [z][y][x<=X/2] interleaved complex.
// A caller may derive its own kernel; CpuFFTSolver then calls its mul() on
// the host between the device transforms.
template <typename T>
class CpuGreensFunc {
 public:
  CpuGreensFunc() {}
  virtual ~CpuGreensFunc() {}
  virtual void mul(const int nfftX, const int nfftY, const int nfftZ, const double boxx,
                   const double boxy, const double boxz, const double sigma,
                   CpuGrid<T>& rho) = 0;
  // device kind of the built-in kernels, -1 for caller-defined ones
  virtual int deviceKind() const { return -1; }
};

namespace detail {
template <typename T>
void greens_mul(gse_greens_kind_t kind, int X, int Y, int Z, double bx, double by, double bz,
                double sigma, CpuGrid<T>& rho) {
  const void* d = rho.toDevice();
  check(gse_greens_func_mul(dtype_of<T>::value, kind, X, Y, Z, bx, by, bz, sigma,
                            const_cast<void*>(d), nullptr));
  rho.fromDevice();
}
}  // namespace detail

// CpuGreensFuncGSE.h:9-19 (CpuGreensFuncGSE.cpp:9-56)
template <typename T>
class CpuGreensFuncGSE : public CpuGreensFunc<T> {
 public:
  void mul(const int nfftX, const int nfftY, const int nfftZ, const double boxx,
           const double boxy, const double boxz, const double sigma, CpuGrid<T>& rho) override {
    detail::greens_mul(GSE_GREENS_GSE, nfftX, nfftY, nfftZ, boxx, boxy, boxz, sigma, rho);
  }
  int deviceKind() const override { return GSE_GREENS_GSE; }
};

// CpuGreensFuncG2.h (CpuGreensFuncG2.cpp:9-60)
template <typename T>
class CpuGreensFuncG2 : public CpuGreensFunc<T> {
 public:
  void mul(const int nfftX, const int nfftY, const int nfftZ, const double boxx,
           const double boxy, const double boxz, const double sigma, CpuGrid<T>& rho) override {
    detail::greens_mul(GSE_GREENS_G2, nfftX, nfftY, nfftZ, boxx, boxy, boxz, sigma, rho);
  }
  int deviceKind() const override { return GSE_GREENS_G2; }
};

// CpuFFTSolver.h:9-43.  Built-in Green's functions run fused on the device
// (r2c -> k-pass -> c2r); a caller-defined one is applied on the host.
template <typename T>
class CpuFFTSolver {
 public:
  CpuFFTSolver(CpuGreensFunc<T>& greensFunc, const int sizeX, const int sizeY, const int sizeZ)
      : sizeX(sizeX), sizeY(sizeY), sizeZ(sizeZ), greensFunc(greensFunc),
        rhoK((sizeX / 2 + 1) * 2, sizeY, sizeZ) {
    const int k = greensFunc.deviceKind();
    check(gse_fft_solver_create(&h_, dtype_of<T>::value,
                                k < 0 ? GSE_GREENS_GSE : (gse_greens_kind_t)k, sizeX, sizeY,
                                sizeZ));
  }
  ~CpuFFTSolver() { (void)gse_fft_solver_destroy(h_); }
  CpuFFTSolver(const CpuFFTSolver&) = delete;
  CpuFFTSolver& operator=(const CpuFFTSolver&) = delete;

  void run(CpuGrid<T>& phi, const CpuGrid<T>& rho, const double sigma, const double boxx,
           const double boxy, const double boxz) {
    detail::check_dims(phi, sizeX, sizeY, sizeZ);
    detail::check_dims(rho, sizeX, sizeY, sizeZ);
    if (greensFunc.deviceKind() >= 0) {
      check(gse_fft_solver_run(h_, phi.deviceBuffer(), rho.toDevice(), sigma, boxx, boxy, boxz));
    } else {
      check(gse_fft_solver_r2c(h_, rho.toDevice(), rhoK.deviceBuffer()));
      rhoK.fromDevice();
      greensFunc.mul(sizeX, sizeY, sizeZ, boxx, boxy, boxz, sigma, rhoK);
      check(gse_fft_solver_c2r(h_, const_cast<void*>(rhoK.toDevice()), phi.deviceBuffer()));
    }
    phi.fromDevice();
  }

 private:
  const int sizeX, sizeY, sizeZ;
  CpuGreensFunc<T>& greensFunc;
  CpuGrid<T> rhoK;
  gse_fft_solver_t h_ = nullptr;
};

// CudaGrid.h:14-42: a grid in device memory (x fastest), allocated by the
// constructor and freed by the destructor.
template <typename T>
class CudaGrid : public Grid<T> {
 public:
  CudaGrid(const int sizeX, const int sizeY, const int sizeZ) : Grid<T>(sizeX, sizeY, sizeZ) {
    void* p = nullptr;
    check(gse_device_malloc(&p, bytes()));
    this->data = (T*)p;
  }
  ~CudaGrid() { (void)gse_device_free(this->data); }
  CudaGrid(const CudaGrid&) = delete;
  CudaGrid& operator=(const CudaGrid&) = delete;

  // CudaGrid.h:28-30
  void clear() { check(gse_device_clear(this->data, bytes())); }
  // CudaGrid.h:32-35: host -> device, synchronous
  void copy(const CpuGrid<T>& h_grid) {
    detail::check_dims(h_grid, this->sizeX, this->sizeY, this->sizeZ);
    check(gse_memcpy_to_device(this->data, h_grid.getDataPointer(), bytes()));
  }
  // CudaGrid.h:37-40: device -> device, synchronous
  void copy(const CudaGrid<T>& d_grid) {
    detail::check_dims(d_grid, this->sizeX, this->sizeY, this->sizeZ);
    check(gse_memcpy_device_to_device(this->data, d_grid.getDataPointer(), bytes()));
  }
  size_t bytes() const { return sizeof(T) * (size_t)this->size; }
};

template <typename T>
void CpuGrid<T>::copy(const CudaGrid<T>& d_grid) {
  detail::check_dims(d_grid, this->sizeX, this->sizeY, this->sizeZ);
  check(gse_memcpy_to_host(this->data, d_grid.getDataPointer(), bytes()));
}

// CudaGaussCharge.h:11-33: sigma1 spread of DEVICE particles into a device
// grid (CudaGaussCharge.cu:255-301).  The reference spreads with float
// atomics; this spread is the fixed-point one of CpuGaussCharge, so the grid
// is bit-identical run to run and equals the CPU spread to rounding.
template <typename AT, typename CT>
class CudaGaussCharge {
 public:
  CudaGaussCharge(const int sizeX, const int sizeY, const int sizeZ)
      : sizeX(sizeX), sizeY(sizeY), sizeZ(sizeZ) {
    check(gse_gauss_charge_create(&h_, dtype_of<AT>::value, dtype_of<CT>::value, sizeX, sizeY,
                                  sizeZ));
  }
  ~CudaGaussCharge() { (void)gse_gauss_charge_destroy(h_); }
  CudaGaussCharge(const CudaGaussCharge&) = delete;
  CudaGaussCharge& operator=(const CudaGaussCharge&) = delete;

  // CudaGaussCharge.h:29-32; xyzq is a device array of numCoord records
  void spreadChargeToGrid(const double sigma, const double rcut, const int numCoord,
                          const xyzq_t<CT>* xyzq, const double boxx, const double boxy,
                          const double boxz, CudaGrid<CT>& rho) {
    detail::check_dims(rho, sizeX, sizeY, sizeZ);
    check(gse_gauss_charge_spread_charge_to_grid(h_, sigma, rcut, numCoord, xyzq, boxx, boxy,
                                                 boxz, rho.getDataPointer(), 0));
    check(gse_device_synchronize());
  }

  gse_gauss_charge_t handle() const { return h_; }

 private:
  const int sizeX, sizeY, sizeZ;
  gse_gauss_charge_t h_ = nullptr;
};

// CudaConvolution.h:9-37: separable convolution of device grids.  run() is
// the X pass followed by the copy into dataXYZ, as CudaConvolution.cu:115-150
// does; runXYZ() (not in the reference class) runs all three passes.
template <typename T>
class CudaConvolution {
 public:
  CudaConvolution(const int sizeX, const int sizeY, const int sizeZ)
      : sizeX(sizeX), sizeY(sizeY), sizeZ(sizeZ) {
    check(gse_convolution_create(&h_, dtype_of<T>::value, sizeX, sizeY, sizeZ));
  }
  ~CudaConvolution() { (void)gse_convolution_destroy(h_); }
  CudaConvolution(const CudaConvolution&) = delete;
  CudaConvolution& operator=(const CudaConvolution&) = delete;

  // CudaConvolution.h:33-34: the taps are HOST arrays of 2R+1 values
  void setupKernel(const int kernelRadiusX, const int kernelRadiusY, const int kernelRadiusZ,
                   const T* h_kernelX, const T* h_kernelY, const T* h_kernelZ) {
    check(gse_convolution_setup_kernel(h_, kernelRadiusX, kernelRadiusY, kernelRadiusZ,
                                       h_kernelX, h_kernelY, h_kernelZ));
  }
  // CudaConvolution.h:35
  void run(const CudaGrid<T>& data, CudaGrid<T>& dataXYZ) {
    detail::check_dims(data, sizeX, sizeY, sizeZ);
    detail::check_dims(dataXYZ, sizeX, sizeY, sizeZ);
    check(gse_convolution_run(h_, data.getDataPointer(), dataXYZ.getDataPointer()));
    check(gse_device_synchronize());
  }
  void runXYZ(const CudaGrid<T>& data, CudaGrid<T>& dataXYZ) {
    detail::check_dims(data, sizeX, sizeY, sizeZ);
    detail::check_dims(dataXYZ, sizeX, sizeY, sizeZ);
    check(gse_convolution_run_xyz(h_, data.getDataPointer(), dataXYZ.getDataPointer()));
    check(gse_device_synchronize());
  }

 private:
  const int sizeX, sizeY, sizeZ;
  gse_convolution_t h_ = nullptr;
};

// CpuEwaldDirect.h:10-44: direct-space erfc sum within rcut plus the self
// energy, FP64.  Host particle and force arrays, as the reference; the
// forces it returns are dE/dr (CpuEwaldDirect.cpp:107-117).
class CpuEwaldDirect {
 public:
  CpuEwaldDirect(const double sigma, const double rcut, const double boxx, const double boxy,
                 const double boxz, const double ccelec = 332.0716) {
    check(gse_ewald_direct_create(&h_, sigma, rcut, boxx, boxy, boxz, ccelec));
  }
  ~CpuEwaldDirect() { (void)gse_ewald_direct_destroy(h_); }
  CpuEwaldDirect(const CpuEwaldDirect&) = delete;
  CpuEwaldDirect& operator=(const CpuEwaldDirect&) = delete;

  // CpuEwaldDirect.h:39
  void setupExclusions(const int n_iblo14_in, const int* iblo14_in, const int* inb14_in) {
    check(gse_ewald_direct_setup_exclusions(h_, n_iblo14_in, iblo14_in, inb14_in));
  }
  // CpuEwaldDirect.h:41-42
  double calcForceEnergy(const int n, const xyzq_t<double>* xyzq, double* fx, double* fy,
                         double* fz) {
    const void* dx = detail::upload_xyzq(xyzq_, n, xyzq);
    const size_t fb = sizeof(double) * (size_t)(n > 0 ? n : 0);
    char* f = (char*)force_.get(3 * fb + sizeof(double));
    double* e = (double*)(f + 3 * fb);
    check(gse_ewald_direct_calc_force_energy(h_, n, dx, (double*)f, (double*)(f + fb),
                                             (double*)(f + 2 * fb), e));
    detail::download(fx, f, n);
    detail::download(fy, f + fb, n);
    detail::download(fz, f + 2 * fb, n);
    double eh = 0.0;
    check(gse_memcpy_to_host(&eh, e, sizeof(double)));
    return eh;
  }

 private:
  gse_ewald_direct_t h_ = nullptr;
  DeviceBuffer xyzq_, force_;
};

// CpuEwaldRecip.h:12-46: the exact Ewald k-space sum, FP64, host arrays;
// the forces are -dE/dr (CpuEwaldRecip.cpp:125-139).
class CpuEwaldRecip {
 public:
  CpuEwaldRecip(const double sigma, const int kcut, const double boxx_in, const double boxy_in,
                const double boxz_in, const double ccelec = 332.0716) {
    check(gse_ewald_recip_create(&h_, sigma, kcut, boxx_in, boxy_in, boxz_in, ccelec));
  }
  ~CpuEwaldRecip() { (void)gse_ewald_recip_destroy(h_); }
  CpuEwaldRecip(const CpuEwaldRecip&) = delete;
  CpuEwaldRecip& operator=(const CpuEwaldRecip&) = delete;

  // CpuEwaldRecip.h:35-36
  double calcForceEnergy(const int n, const xyzq_t<double>* xyzq, double* fx, double* fy,
                         double* fz) {
    const void* dx = detail::upload_xyzq(xyzq_, n, xyzq);
    const size_t fb = sizeof(double) * (size_t)(n > 0 ? n : 0);
    char* f = (char*)force_.get(3 * fb + sizeof(double));
    double* e = (double*)(f + 3 * fb);
    check(gse_ewald_recip_calc_force_energy(h_, n, dx, (double*)f, (double*)(f + fb),
                                            (double*)(f + 2 * fb), e));
    detail::download(fx, f, n);
    detail::download(fy, f + fb, n);
    detail::download(fz, f + 2 * fb, n);
    double eh = 0.0;
    check(gse_memcpy_to_host(&eh, e, sizeof(double)));
    return eh;
  }

 private:
  gse_ewald_recip_t h_ = nullptr;
  DeviceBuffer xyzq_, force_;
};

// cuda/CudaPMERecip.h:36-182: SPME reciprocal space on DEVICE particles
// (xyzq_t<CT>, the reference's float4) and device force arrays, one GPU,
// BOX FFT.  The energies and the virial that the reference adds into its
// CudaEnergyVirial registry are kept here and read with getEnergy(),
// getSelfEnergy() and getVirial().
template <typename AT, typename CT>
class CudaPMERecip {
 public:
  CudaPMERecip(const int nfftx, const int nffty, const int nfftz, const int order) {
    check(gse_pme_recip_create(&h_, dtype_of<AT>::value, dtype_of<CT>::value, nfftx, nffty,
                               nfftz, order));
    ev_.get(sizeof(double) * 11);
  }
  ~CudaPMERecip() { (void)gse_pme_recip_destroy(h_); }
  CudaPMERecip(const CudaPMERecip&) = delete;
  CudaPMERecip& operator=(const CudaPMERecip&) = delete;

  // CudaPMERecip.h:155
  void spread_charge(const xyzq_t<CT>* xyzq, const int ncoord, const double* recip) {
    check(gse_pme_recip_spread_charge(h_, xyzq, ncoord, recip));
  }
  // CudaPMERecip.h:173-174
  void r2c_fft() { check(gse_pme_recip_r2c_fft(h_)); }
  void c2r_fft() { check(gse_pme_recip_c2r_fft(h_)); }
  // CudaPMERecip.h:157-158
  void scalar_sum(const double* recip, const double kappa, const bool calc_energy,
                  const bool calc_virial) {
    double* ev = (double*)ev_.ptr();
    check(gse_pme_recip_scalar_sum(h_, recip, kappa, calc_energy, calc_virial, ev, ev + 1));
  }
  // CudaPMERecip.h:160
  void calc_self_energy(const xyzq_t<CT>* xyzq, const int ncoord, const double kappa) {
    check(gse_pme_recip_calc_self_energy(h_, xyzq, ncoord, kappa, (double*)ev_.ptr() + 10));
  }
  // CudaPMERecip.h:165-167: force[i], force[i+stride], force[i+2*stride] = dE/dr_i
  void gather_force(const xyzq_t<CT>* xyzq, const int ncoord, const double* recip,
                    const int stride, AT* force) {
    check(gse_pme_recip_gather_force(h_, xyzq, ncoord, recip, stride, force));
  }
  double getEnergy() const { return read(0); }
  double getSelfEnergy() const { return read(10); }
  void getVirial(double* virial) const {
    check(gse_memcpy_to_host(virial, (const double*)ev_.ptr() + 1, 9 * sizeof(double)));
  }
  gse_pme_recip_t handle() const { return h_; }

 private:
  double read(int k) const {
    double v = 0.0;
    check(gse_memcpy_to_host(&v, (const double*)ev_.ptr() + k, sizeof(double)));
    return v;
  }
  gse_pme_recip_t h_ = nullptr;
  mutable DeviceBuffer ev_;
};

// CpuGSE.h:18-68.  The device handle holds rhoS and phiM.
template <typename AT, typename CT>
class CpuGSE {
 public:
  CpuGSE(const int sizeX, const int sizeY, const int sizeZ, const double sigma,
         const double kappa, const double lambdaSigma, const double lambdaSigma1,
         const double boxx, const double boxy, const double boxz, const double ccelec = 332.0716)
      : ccelec(ccelec), boxx(boxx), boxy(boxy), boxz(boxz), sigma(sigma), kappa(kappa),
        lambdaSigma(lambdaSigma), lambdaSigma1(lambdaSigma1), rhoS(sizeX, sizeY, sizeZ),
        phiM(sizeX, sizeY, sizeZ) {
    check(gse_gsek_create(&h_, dtype_of<AT>::value, dtype_of<CT>::value, sizeX, sizeY, sizeZ,
                          sigma, kappa, lambdaSigma, lambdaSigma1, boxx, boxy, boxz, ccelec));
    check(gse_gsek_get_params(h_, &sigma1, &sigma2, nullptr));
  }
  virtual ~CpuGSE() { (void)gse_gsek_destroy(h_); }
  CpuGSE(const CpuGSE&) = delete;
  CpuGSE& operator=(const CpuGSE&) = delete;

  void printInfo() {
    std::printf("sigma = %lf kappa = %lf lambdaSigma = %lf lambdaSigma1 = %lf\n", sigma, kappa,
                lambdaSigma, lambdaSigma1);
    std::printf("sigma1 = %lf sigma2 = %lf\n", sigma1, sigma2);
    std::printf("grid %d x %d x %d, box %lf %lf %lf\n", rhoS.getSizeX(), rhoS.getSizeY(),
                rhoS.getSizeZ(), boxx, boxy, boxz);
  }
  void setBoxSize(const double bx, const double by, const double bz) {
    check(gse_gsek_set_box_size(h_, bx, by, bz));
    boxx = bx;
    boxy = by;
    boxz = bz;
  }
  virtual void spreadCharge1(const int numCoord, const xyzq_t<CT>* xyzq) = 0;
  virtual void solvePoisson() = 0;
  virtual AT calculateEnergy() = 0;
  // CpuGSE.cpp:70-80: sphere-support force times ccelec
  void interpolateForce(const int numCoord, const xyzq_t<CT>* xyzq, AT* forceX, AT* forceY,
                        AT* forceZ) {
    const void* dx = detail::upload_xyzq(xyzq_, numCoord, xyzq);
    const size_t fb = sizeof(AT) * (size_t)(numCoord > 0 ? numCoord : 0);
    char* f = (char*)force_.get(3 * fb);
    check(gse_gsek_interpolate_force(h_, numCoord, dx, f, f + fb, f + 2 * fb));
    detail::download(forceX, f, numCoord);
    detail::download(forceY, f + fb, numCoord);
    detail::download(forceZ, f + 2 * fb, numCoord);
  }
  CpuGrid<CT>& getRhoS() {
    check(gse_gsek_copy_grid(h_, 0, rhoS.deviceBuffer()));
    rhoS.fromDevice();
    return rhoS;
  }
  CpuGrid<CT>& getPhi() {
    check(gse_gsek_copy_grid(h_, 1, phiM.deviceBuffer()));
    phiM.fromDevice();
    return phiM;
  }
  double getSigma1() { return sigma1; }
  double getSigma2() { return sigma2; }
  double getLambdaSigma1() { return lambdaSigma1; }
  gse_gsek_t handle() const { return h_; }

 protected:
  const double ccelec;
  double boxx, boxy, boxz;
  double sigma, kappa, lambdaSigma, lambdaSigma1, sigma1 = 0, sigma2 = 0;
  CpuGrid<CT> rhoS;
  CpuGrid<CT> phiM;
  gse_gsek_t h_ = nullptr;
  DeviceBuffer xyzq_, force_;
};

// CpuGSEk.h:18-37
template <typename AT, typename CT>
class CpuGSEk : public CpuGSE<AT, CT> {
 public:
  CpuGSEk(const int sizeX, const int sizeY, const int sizeZ, const double sigma,
          const double kappa, const double lambdaSigma, const double lambdaSigma1,
          const double boxx, const double boxy, const double boxz, const double ccelec = 332.0716)
      : CpuGSE<AT, CT>(sizeX, sizeY, sizeZ, sigma, kappa, lambdaSigma, lambdaSigma1, boxx, boxy,
                       boxz, ccelec) {}
  // CpuGSEk.cpp:30-34
  void spreadCharge1(const int numCoord, const xyzq_t<CT>* xyzq) override {
    const void* dx = detail::upload_xyzq(this->xyzq_, numCoord, xyzq);
    check(gse_gsek_spread_charge1(this->h_, numCoord, dx));
  }
  // CpuGSEk.cpp:36-38
  void spreadCharge2(CpuGrid<CT>& rhoOut) {
    detail::check_dims(rhoOut, this->rhoS.getSizeX(), this->rhoS.getSizeY(), this->rhoS.getSizeZ());
    check(gse_gsek_spread_charge2(this->h_, rhoOut.deviceBuffer()));
    rhoOut.fromDevice();
  }
  // CpuGSEk.cpp:40-43
  void solvePoisson() override { check(gse_gsek_solve_poisson(this->h_)); }
  // CpuGSEk.cpp:56-58 -> CpuGSE.cpp:86-104
  AT calculateEnergy() override {
    double e = 0.0;
    check(gse_gsek_calculate_energy(this->h_, &e));
    return (AT)e;
  }
};

}  // namespace gse_amd

#endif  // GSE_AMD_GSE_HPP
