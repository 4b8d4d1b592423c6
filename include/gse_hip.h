/*
 * gse_hip.h -- C-ABI of the MI355X (gfx950) GSE reciprocal-space library.
 *
 * This is the boundary between callers (the C++ drop-in classes in
 * include/gse/, Python via ctypes, any FFI) and the HIP kernels.  It mirrors
 * the reference's class API for the GSE k-space path of ap-hynninen/GSE;
 * each entry point names the reference declaration it replaces.
 *
 * Conventions
 *   - Plain C types only; every call returns gse_status_t and never exits or
 *     prints.  gse_last_error() gives the calling thread's last message.
 *   - Arrays passed to the gse_* calls are DEVICE pointers unless the
 *     parameter says "host".  Calls are asynchronous on the handle's HIP
 *     stream (gse_*_set_stream; default: the null stream), except those
 *     documented as synchronising.
 *   - Particles are xyzq records {x, y, z, q} of the handle's compute type CT
 *     (TypesGSE.h:20-22).  Grids are x-fastest, data[x + (y + z*Y)*X]
 *     (Grid.h:32-44).  Forces and potentials have the accumulation type AT.
 *     Half spectra hold (X/2+1)*Y*Z interleaved complex values.
 *   - The supported (AT, CT) pairs are the reference's instantiations
 *     (CpuGaussCharge.cpp:524-526): (f32,f32), (f64,f32), (f64,f64).
 */
#ifndef GSE_HIP_H
#define GSE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSE_API __attribute__((visibility("default")))

typedef enum gse_status {
  GSE_SUCCESS = 0,
  GSE_ERROR_INVALID_VALUE = 1, /* bad size, null pointer, dtype pair, parameter */
  GSE_ERROR_OUT_OF_MEMORY = 2,
  GSE_ERROR_HIP = 3,           /* HIP runtime or kernel launch failure */
  GSE_ERROR_FFT = 4,           /* rocFFT failure */
  GSE_ERROR_BAD_PARTICLES = 5, /* non-finite or out-of-range particles seen */
  GSE_ERROR_NOT_SUPPORTED = 6
} gse_status_t;

typedef enum gse_dtype { GSE_FLOAT32 = 0, GSE_FLOAT64 = 1 } gse_dtype_t;

/* Green's function plug-ins (CpuGreensFunc.h:13-19):
 *   GSE: 4 pi/(XYZ) exp(-sigma^2 k^2/2)/k^2, k=0 -> 0 (CpuGreensFuncGSE.cpp:9-56);
 *        sigma = 0 gives the bare Coulomb kernel.
 *   G2 : exp(-sigma^2 k^2/2)/(XYZ) (CpuGreensFuncG2.cpp:9-60).
 *   HERMITIAN: see below. */
typedef enum gse_greens_kind {
  GSE_GREENS_GSE = 0,
  GSE_GREENS_G2 = 1,
  /* the HERMITIAN (Mehrstellen) finite-difference Poisson operator of
   * CpuMultiGridSolver.cpp:237-251 solved exactly: phi = A^-1 B (-4 pi rho)
   * with zero mean (sigma unused) -- the discrete solution the reference's
   * red-black Gauss-Seidel sweeps converge to */
  GSE_GREENS_HERMITIAN = 2
} gse_greens_kind_t;

GSE_API const char* gse_status_string(gse_status_t status);
GSE_API const char* gse_last_error(void);
GSE_API int gse_version(void);

/* Bin edge (grid cells) and LDS budget (bytes) used to tile spread and
 * gather; 0 keeps the built-in choice.  Process-wide. */
GSE_API gse_status_t gse_set_tiling(int bin_edge, int lds_budget_bytes);

/* Named launch-shape knobs, process-wide (the autotuner's interface):
 *   bin_edge      preferred bin edge in grid cells            (1..64, 8)
 *   brick_budget  LDS bytes allowed for an int64 brick        (64 KiB)
 *   spread_block  threads per spread workgroup, 0 = automatic (0,256,512,1024)
 *   gather_block  threads per gather workgroup, 0 = automatic (0,256,512)
 *   spread_stage  LDS bytes for staged particles in the spread (16 KiB)
 *   gather_stage  LDS bytes for staged particles in the gather, 0 = automatic
 *                 (24 KiB for FP32 grids, 16 KiB for FP64) (0)
 *   spread_lds32  FP32 grids: 32-bit accumulators, per-bin scale (0/1, 1)
 *   spread_cols   column-owner register spread (0/1, 0); 0 = LDS-atomic brick
 *   spread_flush  1: each bin stores its brick and one pass sums the bricks
 *                 covering each grid point; 0: bricks added with atomics (0/1, 1)
 *   reduce_tile   that sum by one workgroup per 8^3 bin where the grid allows (0/1, 1)
 *   reduce_nt     brick loads of that FP32 sum: 0 ordinary, 1 non-temporal (0/1, 0)
 *   reduce_tile2  that sum over 64-bit bricks with two x points per thread (0/1, 1)
 *   bin_agg       binning claims bin slots once per workgroup and bin, with
 *                 256 x this many particles per workgroup (grids of at most
 *                 16384 bins); 0 = one claim per particle (0,1,2,4,8; 4)
 *   fused_kz      FFT solves of handles created afterwards run 2D x-y
 *                 transforms around one kernel that does the z transform,
 *                 the G(k) multiply with its energy/virial sums and the
 *                 inverse z transform (grid sizes whose Z has no prime
 *                 factor above 8); 0 = 3D transforms and a separate k-pass (0/1, 1)
 *   gather_group  lanes per charge in the FP32 gather of 13-point supports:
 *                 32 (two charges per wave) or 16 (four) (32)
 *   zpass_cols    spectrum columns per workgroup of the fused z pass, for
 *                 handles created afterwards (4, 8, 16; 8)
 *   zpass4        the fused z pass runs as a four-step transform in
 *                 registers (two LDS exchanges per direction) for the column
 *                 lengths it is built for (Z = 80, 216), for handles created
 *                 afterwards; 0 = the Stockham form for every Z (0/1, 1)
 *   bin_coarse_block  threads per workgroup of the first (super-bin) pass of
 *                 the binning of grids above 16384 bins (256, 512, 1024; 1024)
 *   bin_order     the spread's and the gather's workgroups take the bins of
 *                 each XCD's share by descending charge count, so the longest
 *                 start first; the counts are those of the last binning of
 *                 the same shape, sorted by eight spare workgroups of its
 *                 gather (results do not depend on the order); 0 = bin
 *                 order (0/1, 1)
 *   ypass4        for Y = 216 and Y = 80 the x-y transforms
 *                 around the fused z pass run as 1D x rocFFT transforms and
 *                 a four-step y pass in registers (rocFFT's 2D plan takes
 *                 two kernels at 216), for handles created afterwards (0/1,
 *                 1)
 *   zpass4_cols   spectrum columns per workgroup of the four-step z pass;
 *                 0 = 32 for FP32 with Z > 128, else 16 (0, 16, 32; 0)
 *   xy_plane      the x-y transforms around the fused z pass run as one
 *                 workgroup per z plane in LDS (lengths with radices 2..7,
 *                 a plane within 160 KiB of LDS), for handles created
 *                 afterwards; 0 = rocFFT 2D transforms (0/1, 0: rocFFT's
 *                 2D kernels measured faster at 64^2 to 128^2 planes)
 *   spread_mma    FP32 spreads with int32 bricks and a 13-point support run
 *                 on the matrix cores (int8 limb products summed exactly per
 *                 bin, the brick rounded once); 0 = LDS integer atomics (0/1, 0) */
GSE_API gse_status_t gse_set_tuning_param(const char* name, int value);
GSE_API gse_status_t gse_get_tuning_param(const char* name, int* value);

/* Device memory for callers that hold host buffers (the C++ drop-in classes
 * of include/gse/gse.hpp use these; no reference counterpart: the reference
 * CPU classes own host memory, CpuGrid.h:21-24).  The copies are synchronous
 * and order after all work on the null stream. */
GSE_API gse_status_t gse_device_malloc(void** ptr, size_t bytes);
GSE_API gse_status_t gse_device_free(void* ptr);
GSE_API gse_status_t gse_memcpy_to_device(void* dst, const void* src, size_t bytes);
GSE_API gse_status_t gse_memcpy_to_host(void* dst, const void* src, size_t bytes);
/* The device-side helpers of CudaGrid (CudaGrid.h:28-40): zero `bytes` at a
 * device pointer, device -> device copy (both synchronous, null stream), and
 * a wait for all device work, which the reference's CUDA classes do after
 * their kernels (CudaConvolution.cu:149). */
GSE_API gse_status_t gse_memcpy_device_to_device(void* dst, const void* src, size_t bytes);
GSE_API gse_status_t gse_device_clear(void* ptr, size_t bytes);
GSE_API gse_status_t gse_device_synchronize(void);

/* ------------------------------------------------------------------------
 * CpuGaussCharge<AT,CT>  (CpuGaussCharge.h:11-49)
 * ---------------------------------------------------------------------- */
typedef struct gse_gauss_charge_s* gse_gauss_charge_t;

/* CpuGaussCharge.h:30  CpuGaussCharge(sizeX, sizeY, sizeZ) */
GSE_API gse_status_t gse_gauss_charge_create(gse_gauss_charge_t* out, gse_dtype_t at,
                                             gse_dtype_t ct, int sizeX, int sizeY, int sizeZ);
/* CpuGaussCharge.h:31  ~CpuGaussCharge() */
GSE_API gse_status_t gse_gauss_charge_destroy(gse_gauss_charge_t h);
GSE_API gse_status_t gse_gauss_charge_set_stream(gse_gauss_charge_t h, void* hip_stream);

/* CpuGaussCharge.h:32-33  calcSupportSize (host only, no device work) */
GSE_API gse_status_t gse_gauss_charge_calc_support_size(gse_gauss_charge_t h, double rcut,
                                                        double boxx, double boxy, double boxz,
                                                        int* nsup);

/* CpuGaussCharge.h:42-45  spreadChargeToGrid(sigma, rcut, numCoord, xyzq,
 * boxx, boxy, boxz, rho, normalize).  rho is overwritten (the reference
 * clears it first).  Accumulation is int64 fixed point: the grid is
 * bit-identical run to run. */
GSE_API gse_status_t gse_gauss_charge_spread_charge_to_grid(gse_gauss_charge_t h, double sigma,
                                                            double rcut, int numCoord,
                                                            const void* xyzq, double boxx,
                                                            double boxy, double boxz, void* rho,
                                                            int normalize);

/* No counterpart in the reference: spreadChargeToGrid for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * bit-identical result as gse_gauss_charge_spread_charge_to_grid, but every
 * buffer beyond xyzq and rho lives in one static device array of
 * GSE_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends); the call
 * makes no hipMalloc/hipFree and no host allocation.  It needs about
 * 8*X*Y*Z (int64 grid) + 4*n + slots*(8 + 4*sizeof(CT)) + the bricks
 * (8^3 bins: (8+2n_sup)^3 x 4 or 8 bytes per bin) bytes, e.g. 39 MB for
 * 48,000 FP32 charges on an 80^3 grid; a call needing more returns
 * GSE_ERROR_INVALID_VALUE with the size in gse_last_error().  All calls on
 * one GPU share that GPU's array: calls that overlap in time on one GPU
 * (other streams, other threads) race and must be ordered by the caller. */
#define GSE_NOALLOC_SCRATCH_BYTES (64u << 20)
GSE_API gse_status_t gse_gauss_charge_spread_charge_to_grid_noalloc(
    gse_gauss_charge_t h, double sigma, double rcut, int numCoord, const void* xyzq, double boxx,
    double boxy, double boxz, void* rho, int normalize);

/* CpuGaussCharge.h:46-48  spreadChargeOnGrid(sigma, rcut, box, rho, rhoS):
 * separable X->Y->Z periodic Gaussian convolution with radius ceil(rcut/h)
 * and the (2 pi sigma^2)^-3/2 h^3 prefactor on the Z taps. */
GSE_API gse_status_t gse_gauss_charge_spread_charge_on_grid(gse_gauss_charge_t h, double sigma,
                                                            double rcut, double boxx,
                                                            double boxy, double boxz,
                                                            const void* rho, void* rhoS);

/* CpuGaussCharge.h:36-39  interpolateForce(sigma, rcut, numCoord, xyzq,
 * box, phi, forceX, forceY, forceZ): sphere support r^2 < rcut^2, no
 * ccelec factor (CpuGaussCharge.cpp:105-183). */
GSE_API gse_status_t gse_gauss_charge_interpolate_force(gse_gauss_charge_t h, double sigma,
                                                        double rcut, int numCoord,
                                                        const void* xyzq, double boxx,
                                                        double boxy, double boxz,
                                                        const void* phi, void* forceX,
                                                        void* forceY, void* forceZ);

/* CpuGaussCharge.h:34-35  calcElectricFieldOnGrid */
GSE_API gse_status_t gse_gauss_charge_calc_electric_field_on_grid(gse_gauss_charge_t h,
                                                                  double boxx, double boxy,
                                                                  double boxz, const void* phi,
                                                                  void* Ex, void* Ey, void* Ez);

/* CpuGaussCharge.h:40-41  calcDipoleSum; dipSum is a host array of 3
 * doubles.  Synchronises the stream. */
GSE_API gse_status_t gse_gauss_charge_calc_dipole_sum(gse_gauss_charge_t h, double boxx,
                                                      double boxy, double boxz,
                                                      const void* rho, double* dipSum);

/* Build-defined gather over one cube walk (no reference counterpart; it
 * extends interpolateForce).  Outputs may be NULL; `scale` multiplies all. */
#define GSE_GATHER_SPHERE_FORCE 1 /* reference force, sphere support */
#define GSE_GATHER_CUBE_FORCE 2   /* exact -dE/dr, cube support */
#define GSE_GATHER_POTENTIAL 4    /* phi_i = (2 pi s^2)^-3/2 h^3 sum_cube w phi */
#define GSE_GATHER_VIRIAL 8       /* sum_i q_i prefF sum_cube w phi d_a d_b */
typedef struct gse_gather_desc {
  int what;         /* mask of GSE_GATHER_* */
  double scale;     /* e.g. ccelec */
  void* forceX;     /* sphere force (AT) */
  void* forceY;
  void* forceZ;
  void* cubeX;      /* cube force (AT) */
  void* cubeY;
  void* cubeZ;
  void* potential;  /* AT */
  double* virial;   /* device, 9 doubles (row-major 3x3), overwritten */
} gse_gather_desc_t;
GSE_API gse_status_t gse_gauss_charge_gather(gse_gauss_charge_t h, double sigma, double rcut,
                                             int numCoord, const void* xyzq, double boxx,
                                             double boxy, double boxz, const void* phi,
                                             const gse_gather_desc_t* desc);

/* Number of particles the last spread/gather rejected (non-finite or
 * |x/h| >= 1e9).  Synchronises the stream. */
GSE_API gse_status_t gse_gauss_charge_invalid_count(gse_gauss_charge_t h, int* count);

/* Bin layout of the handle's last spread or gather and the kernel that
 * converted its fixed-point sums (build-defined diagnostic, no reference
 * counterpart).  info receives GSE_LAYOUT_FIELDS host ints:
 *   [0..2] bin edge Bx, By, Bz (cells)   [3] number of bins
 *   [4..6] brick edge Lx, Ly, Lz          [7] 1: one bin, global atomics
 *   [8] conversion kernel of the last spread: -1 none yet, 0 atomic flush,
 *       1 generic brick reduce, 2 tiled 32-bit, 3 tiled 32-bit with
 *       non-temporal loads, 4 tiled 64-bit two points per thread, 5 tiled
 *       one point per thread
 *   [9] particle slots per bin (the rest goes to the overflow list)
 *   [10] k-space form of the handle's FFT solve (gse_fft_solver_get_form):
 *        -1 no solver (gauss-charge handles)
 *   [11] kernel of the last binned spread: -1 none yet, 0 LDS integer
 *        atomics, 1 matrix cores (tuning knob spread_mma) */
#define GSE_LAYOUT_FIELDS 12
GSE_API gse_status_t gse_gauss_charge_get_layout(gse_gauss_charge_t h, int* info);

/* ------------------------------------------------------------------------
 * CpuConvolution<T>  (CpuConvolution.h:9-38)
 * ---------------------------------------------------------------------- */
typedef struct gse_convolution_s* gse_convolution_t;

/* CpuConvolution.h:33 */
GSE_API gse_status_t gse_convolution_create(gse_convolution_t* out, gse_dtype_t t, int sizeX,
                                            int sizeY, int sizeZ);
/* CpuConvolution.h:34 */
GSE_API gse_status_t gse_convolution_destroy(gse_convolution_t h);
GSE_API gse_status_t gse_convolution_set_stream(gse_convolution_t h, void* hip_stream);
/* CpuConvolution.h:35-36  setupKernel(Rx, Ry, Rz, h_kernelX, h_kernelY,
 * h_kernelZ): the taps are HOST arrays of 2R+1 values of type T. */
GSE_API gse_status_t gse_convolution_setup_kernel(gse_convolution_t h, int kernelRadiusX,
                                                  int kernelRadiusY, int kernelRadiusZ,
                                                  const void* h_kernelX, const void* h_kernelY,
                                                  const void* h_kernelZ);
/* CpuConvolution.h:37  run(data, dataXYZ): the reference convolves along X
 * only and copies the result (CpuConvolution.cpp:134-150); so does this. */
GSE_API gse_status_t gse_convolution_run(gse_convolution_t h, const void* data, void* dataXYZ);
/* The full separable X->Y->Z convolution with the three kernels. */
GSE_API gse_status_t gse_convolution_run_xyz(gse_convolution_t h, const void* data,
                                             void* dataXYZ);

/* ------------------------------------------------------------------------
 * CpuFFTSolver<T> with a CpuGreensFunc<T>  (CpuFFTSolver.h:9-43)
 * ---------------------------------------------------------------------- */
typedef struct gse_fft_solver_s* gse_fft_solver_t;

/* CpuFFTSolver.h:39  CpuFFTSolver(greensFunc, sizeX, sizeY, sizeZ); the
 * Green's function is chosen by kind. */
GSE_API gse_status_t gse_fft_solver_create(gse_fft_solver_t* out, gse_dtype_t t,
                                           gse_greens_kind_t kind, int sizeX, int sizeY,
                                           int sizeZ);
/* CpuFFTSolver.h:40 */
GSE_API gse_status_t gse_fft_solver_destroy(gse_fft_solver_t h);
GSE_API gse_status_t gse_fft_solver_set_stream(gse_fft_solver_t h, void* hip_stream);
/* CpuFFTSolver.h:41-42  run(phi, rho, sigma, boxx, boxy, boxz):
 * phi = c2r(G . r2c(rho)), both transforms unnormalised (FFTW semantics). */
GSE_API gse_status_t gse_fft_solver_run(gse_fft_solver_t h, void* phi, const void* rho,
                                        double sigma, double boxx, double boxy, double boxz);
/* The two transforms of CpuFFTSolver.cpp:88 and :94 on their own, for a
 * caller that applies its own Green's function between them.  c2r may
 * overwrite rhoK. */
GSE_API gse_status_t gse_fft_solver_r2c(gse_fft_solver_t h, const void* rho, void* rhoK);
GSE_API gse_status_t gse_fft_solver_c2r(gse_fft_solver_t h, void* rhoK, void* phi);
/* Build-defined diagnostic (no reference counterpart): *form = 1 when run()
 * takes 2D x-y transforms around the fused z pass (z transform, G(k)
 * multiply and energy/virial sums, inverse z transform in one kernel), 2
 * when that z pass is the four-step register form, 3 and 4 for those two
 * when the x-y transforms are the plane kernels instead of rocFFT 2D
 * transforms, 5 and 6 when they are 1D x rocFFT transforms and the
 * four-step y pass, 0 when run() takes 3D transforms and a separate k-pass.
 * The form is fixed at create time by the fused_kz, zpass4, xy_plane and
 * ypass4 tuning knobs and the factors of the grid sizes. */
GSE_API gse_status_t gse_fft_solver_get_form(gse_fft_solver_t h, int* form);

/* CpuGreensFunc.h:16-18  mul(nfftX, nfftY, nfftZ, box, sigma, rhoK) */
GSE_API gse_status_t gse_greens_func_mul(gse_dtype_t t, gse_greens_kind_t kind, int nfftX,
                                         int nfftY, int nfftZ, double boxx, double boxy,
                                         double boxz, double sigma, void* rhoK,
                                         void* hip_stream);

/* ------------------------------------------------------------------------
 * CpuGSE / CpuGSEk<AT,CT>  (CpuGSE.h:18-68, CpuGSEk.h:18-37)
 * ---------------------------------------------------------------------- */
typedef struct gse_gsek_s* gse_gsek_t;

/* CpuGSEk.h:28-31  CpuGSEk(sizeX, sizeY, sizeZ, sigma, kappa, lambdaSigma,
 * lambdaSigma1, boxx, boxy, boxz, ccelec); asserts of CpuGSE.cpp:33-44
 * (0 <= kappa <= 0.5) become GSE_ERROR_INVALID_VALUE. */
GSE_API gse_status_t gse_gsek_create(gse_gsek_t* out, gse_dtype_t at, gse_dtype_t ct,
                                     int sizeX, int sizeY, int sizeZ, double sigma,
                                     double kappa, double lambdaSigma, double lambdaSigma1,
                                     double boxx, double boxy, double boxz, double ccelec);
/* CpuGSEk.h:32 */
GSE_API gse_status_t gse_gsek_destroy(gse_gsek_t h);
GSE_API gse_status_t gse_gsek_set_stream(gse_gsek_t h, void* hip_stream);
/* CpuGSE.h:58  setBoxSize */
GSE_API gse_status_t gse_gsek_set_box_size(gse_gsek_t h, double boxx, double boxy, double boxz);
/* CpuGSE.h:65-67  getSigma1, getSigma2, getLambdaSigma1 (host) */
GSE_API gse_status_t gse_gsek_get_params(gse_gsek_t h, double* sigma1, double* sigma2,
                                         double* lambdaSigma1);
/* CpuGSEk.h:33  spreadCharge1(numCoord, xyzq): sigma1 spread into rhoS */
GSE_API gse_status_t gse_gsek_spread_charge1(gse_gsek_t h, int numCoord, const void* xyzq);
/* CpuGSEk.h:34  spreadCharge2(rhoOut): sigma2 on-grid spread of rhoS */
GSE_API gse_status_t gse_gsek_spread_charge2(gse_gsek_t h, void* rhoOut);
/* CpuGSEk.h:35  solvePoisson(): phiM = FFT solve of rhoS with G_GSE(sigma2) */
GSE_API gse_status_t gse_gsek_solve_poisson(gse_gsek_t h);
/* CpuGSEk.h:36  calculateEnergy(): 0.5*ccelec*h^3*sum(rhoS*phiM)
 * (CpuGSE.cpp:86-104).  Synchronises; host result. */
GSE_API gse_status_t gse_gsek_calculate_energy(gse_gsek_t h, double* energy);
/* CpuGSE.h:61  interpolateForce(numCoord, xyzq, forceX, forceY, forceZ):
 * sphere-support force times ccelec (CpuGSE.cpp:70-80). */
GSE_API gse_status_t gse_gsek_interpolate_force(gse_gsek_t h, int numCoord, const void* xyzq,
                                                void* forceX, void* forceY, void* forceZ);
/* CpuGSE.h:63-64  getRhoS / getPhi: device pointers owned by the handle */
GSE_API gse_status_t gse_gsek_get_rho_s(gse_gsek_t h, void** rhoS);
GSE_API gse_status_t gse_gsek_get_phi(gse_gsek_t h, void** phi);
/* As gse_gauss_charge_get_layout, for the handle's own spread and gather. */
GSE_API gse_status_t gse_gsek_get_layout(gse_gsek_t h, int* info);
/* Asynchronous device copy of rhoS (which = 0) or phiM (which = 1) into dst. */
GSE_API gse_status_t gse_gsek_copy_grid(gse_gsek_t h, int which, void* dst);

/* Build-defined fused step (no reference counterpart).
 * mode 0 (K): spread sigma1 -> r2c -> G_GSE(sigma2) k-pass with FP64
 *   energy/virial -> c2r -> gather; reproduces CpuGSEk.
 * mode 1 (R): spread sigma1 -> on-grid sigma2 convolution (rcut2) -> r2c ->
 *   G_GSE(0) k-pass -> c2r -> energy sum(rho1*phi) -> gather; the
 *   composition of CpuGaussCharge + CpuFFTSolver as CpuGSEr drives it.
 * Outputs may be NULL.  energy[0] (device) receives E; virial (device, 9
 * doubles, mode K only) receives W_ab = -dE/d eps_ab. */
typedef struct gse_step_desc {
  int mode;
  void* forceX;     /* reference sphere force x ccelec (AT) */
  void* forceY;
  void* forceZ;
  void* potential;  /* phi_i = dE/dq_i (AT) */
  void* cubeX;      /* exact force -dE/dr_i (AT) */
  void* cubeY;
  void* cubeZ;
  double* energy;   /* device, 1 double */
  double* virial;   /* device, 9 doubles */
} gse_step_desc_t;
GSE_API gse_status_t gse_gsek_step(gse_gsek_t h, int numCoord, const void* xyzq,
                                   const gse_step_desc_t* desc);

/* Build-defined adjoint sweep (mode K): recomputes spread -> k-pass ->
 * gather from xyzq (no grid is kept from the forward) and writes, scaled
 * by the upstream gradient *grad_energy (device scalar; NULL = 1):
 *   dEdq[i] = dE/dq_i, dEdx/y/z[i] = dE/dr_i (cube support, exact),
 * and virial (device, 9 doubles, unscaled) = -dE/d eps.  Any output NULL. */
GSE_API gse_status_t gse_gsek_adjoint(gse_gsek_t h, int numCoord, const void* xyzq,
                                      const double* grad_energy, void* dEdq, void* dEdx,
                                      void* dEdy, void* dEdz, double* virial);

/* Stage timing of gse_gsek_step: with profiling on, every step records HIP
 * events on the handle's stream around its stages
 *   0 bin, 1 spread, 2 fixed-point -> grid conversion, 3 r2c FFT,
 *   4 k-pass, 5 c2r FFT, 6 gather
 * (mode R adds its convolution to stage 3).  gse_gsek_stage_times
 * synchronises, writes the mean milliseconds per stage over the recorded
 * steps into ms[0..6], the number of steps into *steps, and clears them. */
#define GSE_NUM_STAGES 7
GSE_API gse_status_t gse_gsek_set_profiling(gse_gsek_t h, int enable);
GSE_API gse_status_t gse_gsek_stage_times(gse_gsek_t h, double* ms, int* steps);

/* Closed-form corrections in FP64 (device out[2]): out[0] self energy
 * -ccelec sum q^2 / sqrt(2 pi sigma^2) (CpuEwaldDirect.cpp:124-128),
 * out[1] neutralising-background energy -ccelec pi Q^2 sigma^2 / V. */
GSE_API gse_status_t gse_gsek_corrections(gse_gsek_t h, int numCoord, const void* xyzq,
                                          double* out);

/* ------------------------------------------------------------------------
 * CpuGSEr<AT,CT>  (CpuGSEr.h:18-37): real-space GSE.  sigma1 spread into
 * rho, on-grid sigma2 convolution into rhoS, the Poisson equation of the
 * HERMITIAN stencil for -4 pi rhoS solved exactly in k-space (the reference
 * relaxes 200 red-black Gauss-Seidel sweeps towards the same solution; the
 * exact solve fixes the free constant of phi to zero mean), energy
 * 1/2 ccelec h^3 sum rho phi, forces ccelec * interpolateForce(sigma1, phi).
 * ---------------------------------------------------------------------- */
typedef struct gse_gser_s* gse_gser_t;
/* CpuGSEr.h:28-31  CpuGSEr(sizeX, sizeY, sizeZ, sigma, kappa, lambdaSigma,
 * lambdaSigma1, boxx, boxy, boxz, ccelec) */
GSE_API gse_status_t gse_gser_create(gse_gser_t* out, gse_dtype_t at, gse_dtype_t ct, int sizeX,
                                     int sizeY, int sizeZ, double sigma, double kappa,
                                     double lambdaSigma, double lambdaSigma1, double boxx,
                                     double boxy, double boxz, double ccelec);
/* CpuGSEr.h:32 */
GSE_API gse_status_t gse_gser_destroy(gse_gser_t h);
GSE_API gse_status_t gse_gser_set_stream(gse_gser_t h, void* hip_stream);
/* CpuGSEr.h:33  spreadCharge1(numCoord, xyzq): sigma1 spread into rho */
GSE_API gse_status_t gse_gser_spread_charge1(gse_gser_t h, int numCoord, const void* xyzq);
/* CpuGSEr.h:34  spreadCharge2(): sigma2 on-grid convolution rho -> rhoS */
GSE_API gse_status_t gse_gser_spread_charge2(gse_gser_t h);
/* CpuGSEr.h:35  solvePoisson(): phiM from -4 pi rhoS (HERMITIAN stencil) */
GSE_API gse_status_t gse_gser_solve_poisson(gse_gser_t h);
/* CpuGSEr.h:36  calculateEnergy(): 1/2 ccelec h^3 sum rho phiM
 * (CpuGSEr.cpp:157-160).  Synchronises; host result. */
GSE_API gse_status_t gse_gser_calculate_energy(gse_gser_t h, double* energy);
/* CpuGSE.h:61  interpolateForce: sphere force x ccelec on phiM */
GSE_API gse_status_t gse_gser_interpolate_force(gse_gser_t h, int numCoord, const void* xyzq,
                                                void* forceX, void* forceY, void* forceZ);
/* device copy of rho (which = 0), rhoS (1) or phiM (2) into dst */
GSE_API gse_status_t gse_gser_copy_grid(gse_gser_t h, int which, void* dst);
/* Build-defined fused step: spreadCharge1, spreadCharge2, solvePoisson,
 * energy (device) and the requested per-particle outputs (no virial). */
GSE_API gse_status_t gse_gser_step(gse_gser_t h, int numCoord, const void* xyzq,
                                   const gse_step_desc_t* desc);

/* ------------------------------------------------------------------------
 * CpuGSEu<AT,CT>  (CpuGSEu.h:18-42): the u-series GSE.  sigma1 spread into
 * rhoS; phiM = sum_j (2 pi s_j^2)^(3/2)/s_j * (rhoS convolved with the
 * on-grid Gaussian of s_j = sqrt(sigma^2 b^(2j) - 2 sigma1^2)), j < N, added
 * from j = N-1 down; energy 2 ln(b) 1/2 ccelec h^3 sum rhoS phiM; forces
 * 2 ln(b) ccelec interpolateForce(sigma1, phiM).  The reference writes
 * rhoG%d.txt at every solve; this library writes no files.
 * ---------------------------------------------------------------------- */
typedef struct gse_gseu_s* gse_gseu_t;
/* CpuGSEu.h:31-35  CpuGSEu(N, b, sizeX, sizeY, sizeZ, sigma, kappa,
 * lambdaSigma, lambdaSigma1, boxx, boxy, boxz, ccelec); N in [1, 64], b > 1 */
GSE_API gse_status_t gse_gseu_create(gse_gseu_t* out, gse_dtype_t at, gse_dtype_t ct, int N,
                                     double b, int sizeX, int sizeY, int sizeZ, double sigma,
                                     double kappa, double lambdaSigma, double lambdaSigma1,
                                     double boxx, double boxy, double boxz, double ccelec);
/* CpuGSEu.h:36 */
GSE_API gse_status_t gse_gseu_destroy(gse_gseu_t h);
GSE_API gse_status_t gse_gseu_set_stream(gse_gseu_t h, void* hip_stream);
/* CpuGSEu.h:39  spreadCharge1(numCoord, xyzq) */
GSE_API gse_status_t gse_gseu_spread_charge1(gse_gseu_t h, int numCoord, const void* xyzq);
/* CpuGSEu.h:40  solvePoisson() */
GSE_API gse_status_t gse_gseu_solve_poisson(gse_gseu_t h);
/* CpuGSEu.h:41  calculateEnergy().  Synchronises; host result. */
GSE_API gse_status_t gse_gseu_calculate_energy(gse_gseu_t h, double* energy);
/* CpuGSEu.h:38  interpolateForce(numCoord, xyzq, forceX, forceY, forceZ) */
GSE_API gse_status_t gse_gseu_interpolate_force(gse_gseu_t h, int numCoord, const void* xyzq,
                                                void* forceX, void* forceY, void* forceZ);
/* device copy of rhoS (which = 0) or phiM (1) into dst */
GSE_API gse_status_t gse_gseu_copy_grid(gse_gseu_t h, int which, void* dst);
/* Build-defined fused step: spreadCharge1, solvePoisson, energy (device)
 * and the requested per-particle outputs, all scaled by 2 ln(b) (no virial). */
GSE_API gse_status_t gse_gseu_step(gse_gseu_t h, int numCoord, const void* xyzq,
                                   const gse_step_desc_t* desc);

/* ------------------------------------------------------------------------
 * Batched independent systems (build-defined; no reference counterpart:
 * the reference runs one system per CpuGSEk object)
 *
 * One launch per stage covers all systems: spread and gather walk a table of
 * (system, bin) items with a fixed grid of workgroups, the FFTs run as one
 * batched rocFFT transform per distinct grid size, the k-pass and the
 * per-system energies are single launches.  gse_batch_step_graph captures
 * that sequence once in a HIP graph and replays it.
 * ---------------------------------------------------------------------- */
typedef struct gse_batch_s* gse_batch_t;

/* n systems; for system s: sizes[3s..3s+2] grid, boxes[3s..3s+2] box,
 * counts[s] particles.  All systems share sigma, kappa, lambdas, ccelec and
 * the (AT,CT) pair.  Host arrays.  Systems whose support does not fit the
 * LDS path give GSE_ERROR_NOT_SUPPORTED. */
GSE_API gse_status_t gse_batch_create(gse_batch_t* out, gse_dtype_t at, gse_dtype_t ct,
                                      int nsys, const int* sizes, const double* boxes,
                                      const int* counts, double sigma, double kappa,
                                      double lambdaSigma, double lambdaSigma1, double ccelec);
GSE_API gse_status_t gse_batch_destroy(gse_batch_t h);
GSE_API gse_status_t gse_batch_set_stream(gse_batch_t h, void* hip_stream);
/* xyzq: all systems' particles concatenated in system order (device).
 * forceX/Y/Z: AT arrays of the same concatenated length, the reference
 * sphere-support force times ccelec (all NULL: no gather); potential
 * (phi_i, AT) may be NULL; energies: device, nsys doubles (may be NULL),
 * E_s = 1/2 ccelec h^3 sum rho phi of system s, as CpuGSEk computes it. */
GSE_API gse_status_t gse_batch_step(gse_batch_t h, const void* xyzq, void* forceX, void* forceY,
                                    void* forceZ, void* potential, double* energies);
/* As gse_batch_step, replayed from a HIP graph captured at the first call
 * (and again whenever a pointer differs from the captured ones).  The
 * capture waits for the handle's stream once; replays are ordered on it. */
GSE_API gse_status_t gse_batch_step_graph(gse_batch_t h, const void* xyzq, void* forceX,
                                          void* forceY, void* forceZ, void* potential,
                                          double* energies);
/* Launch shape of the batch: bin edge in cells (1..64), threads per spread
 * and gather workgroup (256, 512; 0 = default).  Re-lays out the batch. */
GSE_API gse_status_t gse_batch_set_config(gse_batch_t h, int bin_edge, int spread_block,
                                          int gather_block);
/* config[0..2] = bin edge, spread block, gather block in use (host). */
GSE_API gse_status_t gse_batch_get_config(gse_batch_t h, int* config);
/* Autotuner: times `reps` batched steps on xyzq for each candidate shape
 * (bin edge 6, 8, 10 x spread block 256, 512 x gather block 256, 512),
 * keeps the fastest and writes its mean ms per step to *best_ms (host, may
 * be NULL).  Synchronises. */
GSE_API gse_status_t gse_batch_autotune(gse_batch_t h, const void* xyzq, int reps,
                                        double* best_ms);

/* ------------------------------------------------------------------------
 * SPME reciprocal space, the comparator cpu_gse/gpu_gse run beside GSE
 * (CudaPMERecip<AT,CT,CT2>, cuda/CudaPMERecip.h:36-182; BOX FFT on one GPU).
 * The charge grid accumulates in int64 fixed point (bit-identical for any
 * particle order) and is converted to CT; FFTs are rocFFT; energy and virial
 * are FP64 sums.  ccelec is the reference's 332.0716 (CudaPMERecip.h:25).
 * recip: HOST array of 9 doubles, the reciprocal box (orthorhombic only:
 * off-diagonal entries must be zero, as CudaPMERecip.cu:3028-3031 requires).
 * ---------------------------------------------------------------------- */
typedef struct gse_pme_recip_s* gse_pme_recip_t;

/* CudaPMERecip.h:143-145  CudaPMERecip(nfftx, nffty, nfftz, order, BOX, 1, 0, ...);
 * order 4, 6 or 8; every dimension at least the order. */
GSE_API gse_status_t gse_pme_recip_create(gse_pme_recip_t* out, gse_dtype_t at, gse_dtype_t ct,
                                          int nfftx, int nffty, int nfftz, int order);
GSE_API gse_status_t gse_pme_recip_destroy(gse_pme_recip_t h);
/* CudaPMERecip.h:152 */
GSE_API gse_status_t gse_pme_recip_set_stream(gse_pme_recip_t h, void* hip_stream);
/* CudaPMERecip.h:155  spread_charge(xyzq, ncoord, recip): B-spline spread
 * of device particles {x, y, z, q} (CT) into the handle's grid */
GSE_API gse_status_t gse_pme_recip_spread_charge(gse_pme_recip_t h, const void* xyzq, int ncoord,
                                                 const double* recip);
/* CudaPMERecip.h:173-174  r2c_fft(), c2r_fft() on the handle's grid */
GSE_API gse_status_t gse_pme_recip_r2c_fft(gse_pme_recip_t h);
GSE_API gse_status_t gse_pme_recip_c2r_fft(gse_pme_recip_t h);
/* CudaPMERecip.h:157-158  scalar_sum(recip, kappa, calc_energy, calc_virial):
 * multiplies the spectrum by the SPME kernel and, when asked, writes the
 * reciprocal energy (device, 1 double) and virial (device, 9 doubles,
 * row-major) -- overwritten, where the reference adds into its registry. */
GSE_API gse_status_t gse_pme_recip_scalar_sum(gse_pme_recip_t h, const double* recip,
                                              double kappa, int calc_energy, int calc_virial,
                                              double* energy, double* virial);
/* CudaPMERecip.h:165-167  gather_force(xyzq, ncoord, recip, stride, force):
 * force[i], force[i+stride], force[i+2*stride] (AT) receive dE/dr_i, the
 * reference's sign (its driver negates them, gpu_gse.cu:145-149). */
GSE_API gse_status_t gse_pme_recip_gather_force(gse_pme_recip_t h, const void* xyzq, int ncoord,
                                                const double* recip, int stride, void* force);
/* CudaPMERecip.h:160  calc_self_energy: -kappa ccelec/sqrt(pi) sum q^2
 * written to energy (device, 1 double) */
GSE_API gse_status_t gse_pme_recip_calc_self_energy(gse_pme_recip_t h, const void* xyzq,
                                                    int ncoord, double kappa, double* energy);
/* Build-defined fused step as gpu_gse.cu:61-73 sequences it for an
 * orthorhombic box: spread -> r2c -> scalar_sum -> c2r -> gather, with the
 * forces -dE/dr in fx, fy, fz (AT; all NULL: no gather); energy and virial
 * (device) may be NULL. */
GSE_API gse_status_t gse_pme_recip_step(gse_pme_recip_t h, const void* xyzq, int ncoord,
                                        double boxx, double boxy, double boxz, double kappa,
                                        double* energy, double* virial, void* fx, void* fy,
                                        void* fz);
/* Device pointer of the handle's CT grid: the charge grid after a spread,
 * the convolved grid after c2r_fft. */
GSE_API gse_status_t gse_pme_recip_get_grid(gse_pme_recip_t h, void** grid);

/* ------------------------------------------------------------------------
 * Ewald direct space + self energy (CpuEwaldDirect, CpuEwaldDirect.h:10-44),
 * FP64 as the reference.  Pairs within rcut (minimum image) are found
 * through cells of edge >= rcut; each thread sums its own particle's terms
 * in a fixed order, so results are deterministic.
 * ---------------------------------------------------------------------- */
typedef struct gse_ewald_direct_s* gse_ewald_direct_t;

/* CpuEwaldDirect.h:34-36  CpuEwaldDirect(sigma, rcut, boxx, boxy, boxz, ccelec) */
GSE_API gse_status_t gse_ewald_direct_create(gse_ewald_direct_t* out, double sigma, double rcut,
                                             double boxx, double boxy, double boxz,
                                             double ccelec);
/* CpuEwaldDirect.h:37 */
GSE_API gse_status_t gse_ewald_direct_destroy(gse_ewald_direct_t h);
GSE_API gse_status_t gse_ewald_direct_set_stream(gse_ewald_direct_t h, void* hip_stream);
/* CpuEwaldDirect.h:39  setupExclusions(n_iblo14, iblo14, inb14): HOST
 * arrays in the reference's format (iblo14[j] running end of atom j's list
 * in inb14, 1-based atom numbers, sign ignored); n_iblo14 = 0 clears them. */
GSE_API gse_status_t gse_ewald_direct_setup_exclusions(gse_ewald_direct_t h, int n_iblo14,
                                                       const int* iblo14, const int* inb14);
/* CpuEwaldDirect.h:41-42  calcForceEnergy(n, xyzq, fx, fy, fz): device
 * xyzq (double), fx/fy/fz (double) receive dE/dr_i (the reference's sign:
 * CpuEwaldDirect.cpp:107-117), energy (device, 1 double) receives
 * E_direct + E_self. */
GSE_API gse_status_t gse_ewald_direct_calc_force_energy(gse_ewald_direct_t h, int n,
                                                        const void* xyzq, double* fx, double* fy,
                                                        double* fz, double* energy);

/* ------------------------------------------------------------------------
 * Ewald reciprocal sum (CpuEwaldRecip, CpuEwaldRecip.h:12-46), FP64: the
 * exact k-space sum over 0 < |m| <= kcut the reference's driver compares
 * every grid method with.  One thread per k-vector forms S(k) (particles in
 * order), one thread per particle sums its force over the k-vectors in the
 * reference's loop order.
 * ---------------------------------------------------------------------- */
typedef struct gse_ewald_recip_s* gse_ewald_recip_t;

/* CpuEwaldRecip.h:29-31  CpuEwaldRecip(sigma, kcut, boxx, boxy, boxz, ccelec) */
GSE_API gse_status_t gse_ewald_recip_create(gse_ewald_recip_t* out, double sigma, int kcut,
                                            double boxx, double boxy, double boxz,
                                            double ccelec);
/* CpuEwaldRecip.h:32 */
GSE_API gse_status_t gse_ewald_recip_destroy(gse_ewald_recip_t h);
GSE_API gse_status_t gse_ewald_recip_set_stream(gse_ewald_recip_t h, void* hip_stream);
/* CpuEwaldRecip.h:35-36  calcForceEnergy(n, xyzq, fx, fy, fz): device xyzq
 * (double); fx/fy/fz (device double, all NULL: energy only) receive the
 * forces -dE/dr_i (CpuEwaldRecip.cpp:125-139); energy (device, 1 double). */
GSE_API gse_status_t gse_ewald_recip_calc_force_energy(gse_ewald_recip_t h, int n,
                                                       const void* xyzq, double* fx, double* fy,
                                                       double* fz, double* energy);

#ifdef __cplusplus
}
#endif
#endif /* GSE_HIP_H */
