// msvd.hpp -- C++ adapter over libofrr.so that keeps the reference's container
// and solver interfaces (namespace msvd): Vector<T> / Matrix<T> / Location
// (CPP/containers/vector.hpp:21-214, matrix.hpp:15-189,
// CPP/core/memory/location.hpp:10-13), MSVDStatus
// (CPP/core/utils/error_handling.hpp:24-62), the BLAS-like calls of
// CPP/linalg/blas/mvops.hpp, the Hessenberg process of
// CPP/linalg/factorization/hessenberg.hpp:22-31 and the solver stubs of
// CPP/linalg/solvers/eig/{rr_eig,si_eig,arnoldi_eig}.hpp and
// CPP/linalg/solvers/svd/{si_svd,rr_svd}.hpp.
//
// Header only; host C++ with the HIP runtime API.  Link with -lofrr
// -lamdhip64.  With clang (hipcc / amdclang++) the reference's half-precision
// element type __half (and __hip_bfloat16) is available; g++ sees the FP64 /
// FP32 subset.  Device buffers come from hipMalloc; results are complete
// when a call returns (every call synchronises the library's context, as the
// reference's synchronous copies do).  Argument and location errors throw
// std::runtime_error where the reference throws; other failures return the
// MSVDStatus code of the library.
#ifndef OFRR_MSVD_HPP
#define OFRR_MSVD_HPP

#include <hip/hip_runtime_api.h>
#if defined(__clang__)
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#define OFRR_MSVD_HALF 1
#endif

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../ofrr.h"

namespace msvd {

enum class Location { kHOST = 0, kDEVICE };

// status codes of the reference (same numeric values as the C-ABI returns)
enum class MSVDStatus {
   kSuccess = 0,
   kErrorInvalidArgument = 1000,
   kErrorInvalidOperation = 1001,
   kErrorOutOfMemory = 1002,
   kErrorNotImplemented = 1003,
   kErrorInternalError = 1004,
   kErrorNaN = 1005,
   kErrorInf = 1006,
   kErrorBLASInvalidLocation = 2000,
   kErrorBLASMixedPrecisionCPU = 2001,
   kErrorBLASHalfPrecisionCPU = 2002,
   kErrorBLASUnsupportedType = 2003,
   kErrorBLASCuBLASFailure = 2004,
   kErrorBLASInvalidSize = 2005,
   kErrorFactorizationFailed = 3000,
   kErrorSingularMatrix = 3001,
   kErrorSolverDivergence = 4000,
   kErrorSolverMaxIterations = 4001,
   kErrorFileNotFound = 5000,
   kErrorInvalidFormat = 5001,
   kErrorCUDAGeneral = 6000,
   kErrorCUDAMemory = 6001,
   kErrorCUDALaunch = 6002,
   kErrorCUDASync = 6003
};

namespace detail {

// one library context per host thread (replaces the CUDAHandler singleton,
// CPP/core/utils/cuda_handler.cu:9-45)
inline ofrr_ctx *ctx() {
   struct Holder {
      ofrr_ctx *c = nullptr;
      ~Holder() { if (c) ofrr_ctx_destroy(c); }
   };
   static thread_local Holder h;
   if (!h.c) {
      if (ofrr_ctx_create(0, &h.c) != 0) throw std::runtime_error("msvd: no HIP device");
      ofrr_ctx_set_stream(h.c, nullptr);  // the null stream, like the reference
   }
   return h.c;
}

inline void hip_check(hipError_t e, const char *what) {
   if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline MSVDStatus done(int st) {
   if (st == 0) st = ofrr_ctx_sync(ctx());
   return static_cast<MSVDStatus>(st);
}

template <typename T> struct dtype;
template <> struct dtype<double> { static constexpr int value = OFRR_F64; };
template <> struct dtype<float> { static constexpr int value = OFRR_F32; };
#ifdef OFRR_MSVD_HALF
template <> struct dtype<__half> { static constexpr int value = OFRR_F16; };
template <> struct dtype<__hip_bfloat16> { static constexpr int value = OFRR_BF16; };
#endif

// element <-> double on the host
template <typename T> inline T to_elem(double x) { return (T)x; }
template <typename T> inline double from_elem(T x) { return (double)x; }
#ifdef OFRR_MSVD_HALF
template <> inline __half to_elem<__half>(double x) { return __float2half((float)x); }
template <> inline double from_elem<__half>(__half x) { return (double)__half2float(x); }
template <> inline __hip_bfloat16 to_elem<__hip_bfloat16>(double x) { return __float2bfloat16((float)x); }
template <> inline double from_elem<__hip_bfloat16>(__hip_bfloat16 x) { return (double)__bfloat162float(x); }
#endif

inline double eps_of(int dt) {
   return dt == OFRR_F64 ? 2.220446049250313e-16 : dt == OFRR_F32 ? 1.1920928955078125e-07 : 9.765625e-04;
}

}  // namespace detail

// ---------------------------------------------------------------------------
// Vector<T>: owning or non-owning buffer on the host or the device
template <typename T> class Vector {
   T *_data_ptr = nullptr;
   Location _location = Location::kDEVICE;
   size_t _length = 0;
   bool _owns_data = false;

   void alloc(size_t n, Location loc) {
      _length = n;
      _location = loc;
      _owns_data = true;
      if (n == 0) return;
      if (loc == Location::kDEVICE) {
         detail::hip_check(hipMalloc(&_data_ptr, n * sizeof(T)), "Vector: hipMalloc");
         detail::hip_check(hipMemset(_data_ptr, 0, n * sizeof(T)), "Vector: hipMemset");
      } else {
         _data_ptr = static_cast<T *>(std::calloc(n, sizeof(T)));
         if (!_data_ptr) throw std::runtime_error("Vector: out of host memory");
      }
   }

 public:
   Vector() = default;
   explicit Vector(size_t size, Location loc = Location::kDEVICE) { alloc(size, loc); }
   // view of external memory (never freed)
   Vector(T *external_data, size_t size, Location loc)
       : _data_ptr(external_data), _location(loc), _length(size), _owns_data(false) {}
   Vector(const Vector &o) {
      alloc(o._length, o._location);
      if (_length)
         detail::hip_check(hipMemcpy(_data_ptr, o._data_ptr, _length * sizeof(T), hipMemcpyDefault),
                           "Vector: copy");
   }
   Vector(Vector &&o) noexcept { swap(o); }
   Vector &operator=(const Vector &o) {
      if (this != &o) { Vector t(o); swap(t); }
      return *this;
   }
   Vector &operator=(Vector &&o) noexcept {
      if (this != &o) { Free(); swap(o); }
      return *this;
   }
   ~Vector() { Free(); }
   void swap(Vector &o) noexcept {
      std::swap(_data_ptr, o._data_ptr);
      std::swap(_location, o._location);
      std::swap(_length, o._length);
      std::swap(_owns_data, o._owns_data);
   }
   void Free() {
      if (_owns_data && _data_ptr) {
         if (_location == Location::kDEVICE) (void)hipFree(_data_ptr);
         else std::free(_data_ptr);
      }
      _data_ptr = nullptr;
      _length = 0;
      _owns_data = false;
   }
   void to_device() { move_to(Location::kDEVICE); }
   void to_host() { move_to(Location::kHOST); }
   void move_to(Location loc) {
      if (loc == _location) return;
      Vector t(_length, loc);
      if (_length)
         detail::hip_check(hipMemcpy(t._data_ptr, _data_ptr, _length * sizeof(T), hipMemcpyDefault),
                           "Vector: transfer");
      *this = std::move(t);
   }
   void fill(T val) {
      std::vector<T> h(_length, val);
      if (_length)
         detail::hip_check(hipMemcpy(_data_ptr, h.data(), _length * sizeof(T), hipMemcpyDefault), "fill");
   }
   // uniform entries in [-1, 1) (vector.hpp:127-133) from the library's
   // Philox4x32-10 stream (seed): 2 u - 1 with u ~ U[0, 1), on the device
   void fill_random(unsigned long long seed = 815) {
      if (_location == Location::kDEVICE) {
         ofrr_ctx *cx = detail::ctx();
         int st = ofrr_fill_randu(cx, detail::dtype<T>::value, (int64_t)_length, _data_ptr, 1, seed, 0, 0);
         if (!st) st = ofrr_axpb(cx, detail::dtype<T>::value, (int64_t)_length, 1, _data_ptr, 1, 1, 2.0, -1.0);
         if (detail::done(st) != MSVDStatus::kSuccess) throw std::runtime_error("fill_random failed");
      } else {
         Vector d(_length, Location::kDEVICE);
         d.fill_random(seed);
         detail::hip_check(hipMemcpy(_data_ptr, d.data(), _length * sizeof(T), hipMemcpyDefault), "fill_random");
      }
   }
   // every element times alpha (vector.hpp:140)
   void scale(T alpha) {
      if (_length == 0) return;
      if (_location == Location::kDEVICE) {
         int st = ofrr_axpb(detail::ctx(), detail::dtype<T>::value, (int64_t)_length, 1, _data_ptr, 1, 1,
                            detail::from_elem<T>(alpha), 0.0);
         if (detail::done(st) != MSVDStatus::kSuccess) throw std::runtime_error("scale failed");
      } else {
         const double a = detail::from_elem<T>(alpha);
         for (size_t i = 0; i < _length; ++i) _data_ptr[i] = detail::to_elem<T>(a * detail::from_elem<T>(_data_ptr[i]));
      }
   }
   // a new vector with the same contents in another element type, same
   // location (vector.hpp:148; on the device without a host round trip)
   template <typename U> Vector<U> cast() const {
      Vector<U> out(_length, _location);
      if (_length == 0) return out;
      if (_location == Location::kDEVICE) {
         int st = ofrr_convert(detail::ctx(), (int64_t)_length, 1, _data_ptr, detail::dtype<T>::value, 1, 1,
                               out.data(), detail::dtype<U>::value, 1, 1);
         if (detail::done(st) != MSVDStatus::kSuccess) throw std::runtime_error("cast failed");
      } else {
         for (size_t i = 0; i < _length; ++i) out[i] = detail::to_elem<U>(detail::from_elem<T>(_data_ptr[i]));
      }
      return out;
   }
   T *data() const { return _data_ptr; }
   T &operator[](size_t i) { return _data_ptr[i]; }
   const T &operator[](size_t i) const { return _data_ptr[i]; }
   Location location() const { return _location; }
   size_t length() const { return _length; }
   bool owns_data() const { return _owns_data; }
};

// ---------------------------------------------------------------------------
// Matrix<T>: column-major, element (i, j) at data[i + j * ld]
template <typename T> class Matrix {
   Vector<T> _data;
   size_t _num_rows = 0, _num_cols = 0, _ld = 0;

 public:
   Matrix() = default;
   Matrix(size_t rows, size_t cols, Location loc = Location::kDEVICE)
       : _data(rows * cols, loc), _num_rows(rows), _num_cols(cols), _ld(rows) {}
   Matrix(T *external_data, size_t rows, size_t cols, size_t ld, Location loc)
       : _data(external_data, ld * cols, loc), _num_rows(rows), _num_cols(cols), _ld(ld) {}
   T &operator()(size_t i, size_t j) { return _data[i + j * _ld]; }
   const T &operator()(size_t i, size_t j) const { return _data[i + j * _ld]; }
   void to_host() { _data.to_host(); }
   void to_device() { _data.to_device(); }
   void fill(T val) { _data.fill(val); }
   void fill_random(unsigned long long seed = 815) { _data.fill_random(seed); }
   T *data() { return _data.data(); }
   const T *data() const { return _data.data(); }
   size_t rows() const { return _num_rows; }
   size_t cols() const { return _num_cols; }
   size_t ld() const { return _ld; }
   Location location() const { return _data.location(); }
   // a view of rows [r0, r0 + nr) x columns [c0, c0 + nc) (matrix.hpp:116),
   // on the host or the device: it shares this matrix's memory and ld
   Matrix<T> submatrix(size_t start_row, size_t start_col, size_t num_rows, size_t num_cols) {
      if (start_row + num_rows > _num_rows || start_col + num_cols > _num_cols)
         throw std::runtime_error("Submatrix dimensions exceed matrix bounds");
      return Matrix<T>(_data.data() + start_row + start_col * _ld, num_rows, num_cols, _ld, location());
   }
   // every element times alpha (matrix.hpp:145)
   void scale(T alpha) {
      if (_num_rows == 0 || _num_cols == 0) return;
      if (location() == Location::kDEVICE) {
         int st = ofrr_axpb(detail::ctx(), detail::dtype<T>::value, (int64_t)_num_rows, (int64_t)_num_cols,
                            _data.data(), 1, (int64_t)_ld, detail::from_elem<T>(alpha), 0.0);
         if (detail::done(st) != MSVDStatus::kSuccess) throw std::runtime_error("scale failed");
      } else {
         const double a = detail::from_elem<T>(alpha);
         for (size_t j = 0; j < _num_cols; ++j)
            for (size_t i = 0; i < _num_rows; ++i)
               (*this)(i, j) = detail::to_elem<T>(a * detail::from_elem<T>((*this)(i, j)));
      }
   }
   // a new matrix (ld = rows) with the same contents in another element type,
   // same location (matrix.hpp:152; on the device without a host round trip)
   template <typename U> Matrix<U> cast() const {
      Matrix<U> out(_num_rows, _num_cols, location());
      if (_num_rows == 0 || _num_cols == 0) return out;
      if (location() == Location::kDEVICE) {
         int st = ofrr_convert(detail::ctx(), (int64_t)_num_rows, (int64_t)_num_cols, _data.data(),
                               detail::dtype<T>::value, 1, (int64_t)_ld, out.data(), detail::dtype<U>::value, 1,
                               (int64_t)_num_rows);
         if (detail::done(st) != MSVDStatus::kSuccess) throw std::runtime_error("cast failed");
      } else {
         for (size_t j = 0; j < _num_cols; ++j)
            for (size_t i = 0; i < _num_rows; ++i)
               out(i, j) = detail::to_elem<U>(detail::from_elem<T>((*this)(i, j)));
      }
      return out;
   }
};

namespace detail {
// a device copy of a matrix (the matrix itself when already on the device)
template <typename T> struct DevView {
   Matrix<T> copy;
   const T *p = nullptr;
   size_t ld = 0;
   explicit DevView(const Matrix<T> &A) {
      if (A.location() == Location::kDEVICE) {
         p = A.data();
         ld = A.ld();
      } else {
         copy = Matrix<T>(A.rows(), A.cols(), Location::kDEVICE);
         hip_check(hipMemcpy2D(copy.data(), A.rows() * sizeof(T), A.data(), A.ld() * sizeof(T),
                               A.rows() * sizeof(T), A.cols(), hipMemcpyHostToDevice),
                   "matrix upload");
         p = copy.data();
         ld = A.rows();
      }
   }
};
template <typename T> void store(Matrix<T> &dst, const double *src_dev, size_t rows, size_t cols, size_t lds) {
   std::vector<double> h(rows * cols);
   hip_check(hipMemcpy2D(h.data(), rows * sizeof(double), src_dev, lds * sizeof(double),
                         rows * sizeof(double), cols, hipMemcpyDeviceToHost),
             "result download");
   std::vector<T> t(h.size());
   for (size_t i = 0; i < h.size(); ++i) t[i] = to_elem<T>(h[i]);
   hip_check(hipMemcpy2D(dst.data(), dst.ld() * sizeof(T), t.data(), rows * sizeof(T), rows * sizeof(T),
                         cols, hipMemcpyDefault),
             "result store");
}
template <typename T> void set_diag(Matrix<T> &D, const std::vector<double> &v) {
   std::vector<T> h(D.ld() * D.cols(), to_elem<T>(0.0));
   for (size_t i = 0; i < std::min(v.size(), std::min(D.rows(), D.cols())); ++i) h[i + i * D.ld()] = to_elem<T>(v[i]);
   hip_check(hipMemcpy(D.data(), h.data(), h.size() * sizeof(T), hipMemcpyDefault), "diag store");
}
}  // namespace detail

// ---------------------------------------------------------------------------
// mvops (CPP/linalg/blas/mvops.hpp)

// 1-based index of the first max |x_i| and its signed value (mvops.hpp:98-99);
// location kDEVICE: result and result_val refer to device memory
template <typename T, typename T_O = T>
MSVDStatus iamax(const Vector<T> &x, int &result, T_O &result_val, Location location = Location::kHOST) {
   if (x.location() != Location::kDEVICE) return MSVDStatus::kErrorBLASInvalidLocation;
   Vector<int32_t> idx(1, Location::kDEVICE);
   Vector<double> val(1, Location::kDEVICE);
   int st = ofrr_iamax(detail::ctx(), detail::dtype<T>::value, (int64_t)x.length(), x.data(), 1, idx.data(),
                       val.data());
   MSVDStatus s = detail::done(st);
   if (s != MSVDStatus::kSuccess) return s;
   int32_t hi;
   double hv;
   detail::hip_check(hipMemcpy(&hi, idx.data(), 4, hipMemcpyDeviceToHost), "iamax");
   detail::hip_check(hipMemcpy(&hv, val.data(), 8, hipMemcpyDeviceToHost), "iamax");
   const T_O tv = detail::to_elem<T_O>(hv);
   if (location == Location::kHOST) {
      result = hi;
      result_val = tv;
   } else {
      detail::hip_check(hipMemcpy(&result, &hi, sizeof(int), hipMemcpyHostToDevice), "iamax");
      detail::hip_check(hipMemcpy(&result_val, &tv, sizeof(T_O), hipMemcpyHostToDevice), "iamax");
   }
   return MSVDStatus::kSuccess;
}

// 0-based per-column argmax |A(:,j)| (mvops.hpp:113-114)
template <typename T>
MSVDStatus matrix_iamax(const Matrix<T> &A, int *results, Location location = Location::kHOST) {
   detail::DevView<T> a(A);
   Vector<int32_t> idx(A.cols(), Location::kDEVICE);
   int st = ofrr_matrix_iamax(detail::ctx(), detail::dtype<T>::value, (int64_t)A.rows(), (int)A.cols(), a.p, 1,
                              (int64_t)a.ld, idx.data());
   MSVDStatus s = detail::done(st);
   if (s != MSVDStatus::kSuccess) return s;
   detail::hip_check(hipMemcpy(results, idx.data(), A.cols() * 4,
                               location == Location::kHOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice),
                     "matrix_iamax");
   return MSVDStatus::kSuccess;
}

// NaN / Inf scan (mvops.hpp:159-174)
template <typename T> MSVDStatus check_special_values(const Matrix<T> &A) {
   detail::DevView<T> a(A);
   return static_cast<MSVDStatus>(ofrr_check_special(detail::ctx(), detail::dtype<T>::value, (int64_t)A.rows(),
                                                     (int)A.cols(), a.p, 1, (int64_t)a.ld));
}
template <typename T> MSVDStatus check_special_values(const Vector<T> &x) {
   Matrix<T> v(x.data(), x.length(), 1, x.length(), x.location());
   return check_special_values(v);
}

// C = alpha op(A) op(B) + beta C (mvops.hpp:79-82, mvops.cu:423-605): any
// transposes and sizes.  The reference's checks and their messages come first
// (mvops.cu:433-449, asserted by mvops_test.cpp:734-769); host matrices need
// one element type throughout, as the reference's CBLAS path does, and are
// multiplied on the device (copied there and back).  Device instantiations:
// the reference's <double,double,double>, <float,float,float>,
// <__half,__half,float>, <__half,__half,__half>, and BF16 and mixed output
// types beside them (FP16 compute accumulates in FP32).  Tall-skinny
// projection products A^T B (k >= 4096) run on the FP64-accumulated Gram
// kernel; everything else on the tiled general product.
template <typename T_I, typename T_O, typename T_COMPUTE>
MSVDStatus gemm(bool transA, bool transB, const T_COMPUTE &alpha, const Matrix<T_I> &A, const Matrix<T_I> &B,
                const T_COMPUTE &beta, Matrix<T_O> &C) {
   const size_t m = transA ? A.cols() : A.rows(), k = transA ? A.rows() : A.cols();
   const size_t n = transB ? B.rows() : B.cols(), k2 = transB ? B.cols() : B.rows();
   if (A.location() != B.location() || A.location() != C.location())
      throw std::runtime_error("All matrices must be in the same location (CPU or GPU)");
   if (k != k2) throw std::runtime_error("Matrix dimensions do not match for multiplication");
   if (C.rows() != m || C.cols() != n) throw std::runtime_error("Output matrix C has wrong dimensions");
   const bool host = A.location() == Location::kHOST;
   if (host) {
      if (!std::is_same<T_I, T_O>::value || !std::is_same<T_I, T_COMPUTE>::value)
         throw std::runtime_error("For CPU GEMM, input, output, and compute types must be the same");
      if (!std::is_same<T_I, double>::value && !std::is_same<T_I, float>::value)
         throw std::runtime_error("Unsupported type combination for CPU GEMM");
   }
   if (m == 0 || n == 0) return MSVDStatus::kSuccess;
   ofrr_ctx *cx = detail::ctx();
   const int dt = detail::dtype<T_I>::value, ct = detail::dtype<T_O>::value;
   const int comp = detail::dtype<T_COMPUTE>::value;
   const double a = detail::from_elem<T_COMPUTE>(alpha), bt = detail::from_elem<T_COMPUTE>(beta);
   detail::DevView<T_I> da(A), db(B);
   // C on the device: itself, or a device copy of a host C
   Matrix<T_O> cdev;
   T_O *cp = C.data();
   size_t ldc = C.ld();
   if (host) {
      cdev = Matrix<T_O>(m, n, Location::kDEVICE);
      detail::hip_check(hipMemcpy2D(cdev.data(), m * sizeof(T_O), C.data(), C.ld() * sizeof(T_O), m * sizeof(T_O),
                                    n, hipMemcpyHostToDevice),
                        "gemm");
      cp = cdev.data();
      ldc = m;
   }
   int st;
   if (transA && !transB && k >= 4096) {
      // V^T W: FP64 accumulation with a fixed-order split over the rows
      Matrix<double> G(m, n, Location::kDEVICE);
      st = (bt != 0.0) ? ofrr_convert(cx, (int64_t)m, (int64_t)n, cp, ct, 1, (int64_t)ldc, G.data(), OFRR_F64, 1,
                                      (int64_t)m)
                       : 0;
      if (!st)
         st = ofrr_gram(cx, (int64_t)k, (int)m, (int)n, da.p, dt, 1, (int64_t)da.ld, db.p, dt, 1, (int64_t)db.ld, a,
                        bt, G.data(), (int)m);
      if (!st)
         st = ofrr_convert(cx, (int64_t)m, (int64_t)n, G.data(), OFRR_F64, 1, (int64_t)m, cp, ct, 1, (int64_t)ldc);
   } else {
      st = ofrr_gemm(cx, transA ? 1 : 0, transB ? 1 : 0, (int64_t)m, (int64_t)n, (int64_t)k, a, da.p, dt,
                     (int64_t)da.ld, db.p, dt, (int64_t)db.ld, bt, cp, ct, (int64_t)ldc, comp);
   }
   if (st == (int)MSVDStatus::kErrorBLASUnsupportedType) return MSVDStatus::kErrorBLASCuBLASFailure;
   MSVDStatus s = detail::done(st);
   if (s != MSVDStatus::kSuccess) return s;
   if (host)
      detail::hip_check(hipMemcpy2D(C.data(), C.ld() * sizeof(T_O), cp, m * sizeof(T_O), m * sizeof(T_O), n,
                                    hipMemcpyDeviceToHost),
                        "gemm");
   return MSVDStatus::kSuccess;
}

// symmetric eigenproblems (mvops.hpp:184-197): A (upper triangle used) is
// overwritten with the eigenvectors, W receives the ascending eigenvalues;
// lwork == -1 is a workspace query (work[0] = 1: the GPU needs none from the caller)
template <typename T> MSVDStatus syev(Matrix<T> &A, Vector<T> &W, T *work, int lwork) {
   if (lwork == -1) { if (work) work[0] = T(1); return MSVDStatus::kSuccess; }
   const size_t n = A.rows();
   if (A.cols() != n || W.length() < n) return MSVDStatus::kErrorBLASInvalidSize;
   std::vector<T> h(A.ld() * n);
   detail::hip_check(hipMemcpy(h.data(), A.data(), h.size() * sizeof(T), hipMemcpyDefault), "syev");
   std::vector<double> s(n * n);
   for (size_t j = 0; j < n; ++j)
      for (size_t i = 0; i <= j; ++i) s[i + j * n] = s[j + i * n] = detail::from_elem<T>(h[i + j * A.ld()]);
   Matrix<double> dA(n, n), dV(n, n);
   Vector<double> dw(n);
   detail::hip_check(hipMemcpy(dA.data(), s.data(), s.size() * 8, hipMemcpyHostToDevice), "syev");
   MSVDStatus st = detail::done(ofrr_syev_batched(detail::ctx(), 1, (int)n, dA.data(), (int)n, 0, dw.data(), 0,
                                                  dV.data(), (int)n, 0));
   if (st != MSVDStatus::kSuccess) return st;
   detail::store(A, dV.data(), n, n, n);
   Matrix<T> wm(W.data(), n, 1, n, W.location());
   detail::store(wm, dw.data(), n, 1, n);
   return MSVDStatus::kSuccess;
}

template <typename T> MSVDStatus sygv(Matrix<T> &A, Matrix<T> &B, Vector<T> &W, T *work, int lwork) {
   if (lwork == -1) { if (work) work[0] = T(1); return MSVDStatus::kSuccess; }
   const size_t n = A.rows();
   if (A.cols() != n || B.rows() != n || B.cols() != n || W.length() < n) return MSVDStatus::kErrorBLASInvalidSize;
   auto sym = [n](const Matrix<T> &M) {
      std::vector<T> h(M.ld() * n);
      detail::hip_check(hipMemcpy(h.data(), M.data(), h.size() * sizeof(T), hipMemcpyDefault), "sygv");
      std::vector<double> s(n * n);
      for (size_t j = 0; j < n; ++j)
         for (size_t i = 0; i <= j; ++i) s[i + j * n] = s[j + i * n] = detail::from_elem<T>(h[i + j * M.ld()]);
      return s;
   };
   const std::vector<double> sa = sym(A), sb = sym(B);
   Matrix<double> dA(n, n), dB(n, n), dV(n, n);
   Vector<double> dw(n);
   detail::hip_check(hipMemcpy(dA.data(), sa.data(), sa.size() * 8, hipMemcpyHostToDevice), "sygv");
   detail::hip_check(hipMemcpy(dB.data(), sb.data(), sb.size() * 8, hipMemcpyHostToDevice), "sygv");
   MSVDStatus st = detail::done(ofrr_sygv(detail::ctx(), (int)n, dA.data(), (int)n, dB.data(), (int)n, dw.data(),
                                          dV.data(), (int)n));
   if (st != MSVDStatus::kSuccess) return st;
   detail::store(A, dV.data(), n, n, n);
   Matrix<T> wm(W.data(), n, 1, n, W.location());
   detail::store(wm, dw.data(), n, 1, n);
   return MSVDStatus::kSuccess;
}

// ---------------------------------------------------------------------------
// Hessenberg process (hessenberg.hpp:22-108): A and Q on the device, R on the
// host; returns the skip flags (1 = dropped column).  hessenberg and
// hessenberg_v2 leave a dropped column out of later eliminations (v1 / v2 of
// hessenberg.cu); hessenberg_v3 eliminates with it (hessenberg.cu:768-779).
// The arithmetic is FP64 for an FP64 Q and FP32 otherwise (the reference's
// <__half, __half, float> instantiation computes in FP32 as well).
namespace detail {
inline double default_select_tol(int dt) { return eps_of(dt); }
template <typename T_I, typename T_O>
std::vector<int> hessenberg_run(int variant, const char *name, const Matrix<T_I> &A, Matrix<T_O> &Q,
                                Matrix<T_O> &R, double select_tol) {
   if (A.location() != Location::kDEVICE) throw std::runtime_error("Input matrix A must be on device");
   if (Q.location() != Location::kDEVICE) throw std::runtime_error("Output matrix Q must be on device");
   if (R.location() != Location::kHOST) throw std::runtime_error("Output matrix R must be on host");
   const size_t m = A.rows(), n = A.cols();
   if (Q.rows() != m || Q.cols() != n) throw std::runtime_error("Output matrix Q has incorrect dimensions");
   if (R.rows() != n || R.cols() != n) throw std::runtime_error("Output matrix R has incorrect dimensions");
   ofrr_ctx *cx = ctx();
   int st = ofrr_convert(cx, (int64_t)m, (int64_t)n, A.data(), dtype<T_I>::value, 1, (int64_t)A.ld(), Q.data(),
                         dtype<T_O>::value, 1, (int64_t)Q.ld());
   Matrix<double> dR(n, n);
   Vector<int32_t> perm(n), skip(n);
   if (!st)
      st = (variant == 1 ? ofrr_hessenberg_v3 : ofrr_hessenberg)(cx, dtype<T_O>::value, (int64_t)m, (int)n,
                                                                Q.data(), 1, (int64_t)Q.ld(), select_tol, dR.data(),
                                                                (int)n, perm.data(), skip.data());
   if (done(st) != MSVDStatus::kSuccess) throw std::runtime_error(std::string(name) + ": " + ofrr_ctx_last_error(cx));
   store(R, dR.data(), n, n, n);
   std::vector<int32_t> hs(n);
   hip_check(hipMemcpy(hs.data(), skip.data(), n * 4, hipMemcpyDeviceToHost), name);
   return std::vector<int>(hs.begin(), hs.end());
}
}  // namespace detail

template <typename T_I, typename T_O, typename T_COMPUTE>
std::vector<int> hessenberg(const Matrix<T_I> &A, Matrix<T_O> &Q, Matrix<T_O> &R,
                            T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                            T_O orth_tol = detail::to_elem<T_O>(0.0), T_O reorth_tol = detail::to_elem<T_O>(0.0)) {
   (void)orth_tol;
   (void)reorth_tol;
   return detail::hessenberg_run(0, "hessenberg", A, Q, R, detail::from_elem<T_O>(select_tol));
}

template <typename T_I, typename T_O, typename T_COMPUTE>
std::vector<int> hessenberg_v2(const Matrix<T_I> &A, Matrix<T_O> &Q, Matrix<T_O> &R,
                               T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                               T_O orth_tol = detail::to_elem<T_O>(0.0), T_O reorth_tol = detail::to_elem<T_O>(0.0)) {
   (void)orth_tol;
   (void)reorth_tol;
   return detail::hessenberg_run(0, "hessenberg_v2", A, Q, R, detail::from_elem<T_O>(select_tol));
}

template <typename T_I, typename T_O, typename T_COMPUTE>
std::vector<int> hessenberg_v3(const Matrix<T_I> &A, Matrix<T_O> &Q, Matrix<T_O> &R,
                               T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                               T_O orth_tol = detail::to_elem<T_O>(0.0), T_O reorth_tol = detail::to_elem<T_O>(0.0)) {
   (void)orth_tol;
   (void)reorth_tol;
   return detail::hessenberg_run(1, "hessenberg_v3", A, Q, R, detail::from_elem<T_O>(select_tol));
}

// hessenberg_v3_step (hessenberg.hpp:93-104): column i of Q (device, m x n)
// against its finished columns 0..i-1, then its pivot.  perm_indices_d
// (device, n ints) holds the 1-based argmax rows of the finished columns and
// receives column i's; scale_factor_matrix (device, n x n, T_COMPUTE) receives
// R(0..i, i) in its column i; skip_d[i] (device) and *skip_h (host) the drop
// flag.  R and constant_perm_val_pointer are unused, as in the reference (its
// R is filled from scale_factor_matrix after the last column).  Returns 1 when
// the column was dropped.
template <typename T_O, typename T_COMPUTE>
int hessenberg_v3_step(Matrix<T_O> &Q, Matrix<T_O> &R, int *skip_h, int *skip_d, size_t i, int *perm_indices_d,
                       Matrix<T_COMPUTE> &scale_factor_matrix, T_COMPUTE *constant_perm_val_pointer,
                       T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value))) {
   (void)R;
   (void)constant_perm_val_pointer;
   const size_t m = Q.rows(), n = Q.cols();
   if (Q.location() != Location::kDEVICE) throw std::runtime_error("Output matrix Q must be on device");
   if (i >= n || scale_factor_matrix.rows() < n || scale_factor_matrix.cols() < n)
      throw std::runtime_error("hessenberg_v3_step: bad column or scale factor matrix");
   ofrr_ctx *cx = detail::ctx();
   Matrix<double> dR(n, n);
   // the finished coefficients are only read at the pivot rows, so the
   // column's R arrives complete in dR(:, i)
   int32_t dropped = 0;
   int st = ofrr_hessenberg_step(cx, 1, detail::dtype<T_O>::value, (int64_t)m, (int)n, Q.data(), 1, (int64_t)Q.ld(),
                                 (int)i, detail::from_elem<T_O>(select_tol), dR.data(), (int)n, perm_indices_d,
                                 skip_d, &dropped);
   if (!st)
      st = ofrr_convert(cx, (int64_t)i + 1, 1, dR.data() + i * n, OFRR_F64, 1, (int64_t)n,
                        &scale_factor_matrix(0, i), detail::dtype<T_COMPUTE>::value, 1,
                        (int64_t)scale_factor_matrix.ld());
   if (detail::done(st) != MSVDStatus::kSuccess)
      throw std::runtime_error(std::string("hessenberg_v3_step: ") + ofrr_ctx_last_error(cx));
   if (skip_h) *skip_h = dropped;
   return dropped ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Gram-Schmidt QR (qr.hpp:9-141): A and Q on the device, R on the host; returns
// the skip flags (1 = dropped column).  orth_tol and reorth_tol act in mgs
// only; the others take them as placeholders, as the reference does.
namespace detail {
template <typename T_I, typename T_O>
std::vector<int> gs_qr(int variant, const char *name, const Matrix<T_I> &A, Matrix<T_O> &Q,
                       Matrix<T_O> &R, double select_tol, double orth_tol, double reorth_tol) {
   if (A.location() != Location::kDEVICE) throw std::runtime_error("Input matrix A must be on device");
   if (Q.location() != Location::kDEVICE) throw std::runtime_error("Output matrix Q must be on device");
   if (R.location() != Location::kHOST) throw std::runtime_error("Output matrix R must be on host");
   const size_t m = A.rows(), n = A.cols();
   if (Q.rows() != m || Q.cols() != n) throw std::runtime_error("Output matrix Q has incorrect dimensions");
   if (R.rows() != n || R.cols() != n) throw std::runtime_error("Output matrix R has incorrect dimensions");
   ofrr_ctx *cx = ctx();
   int st = ofrr_convert(cx, (int64_t)m, (int64_t)n, A.data(), dtype<T_I>::value, 1, (int64_t)A.ld(), Q.data(),
                         dtype<T_O>::value, 1, (int64_t)Q.ld());
   Matrix<double> dR(n, n);
   Vector<int32_t> skip(n);
   if (!st)
      st = ofrr_gs_qr(cx, variant, dtype<T_O>::value, (int64_t)m, (int)n, Q.data(), 1, (int64_t)Q.ld(),
                      select_tol, orth_tol, reorth_tol, dR.data(), (int)n, skip.data());
   if (done(st) != MSVDStatus::kSuccess)
      throw std::runtime_error(std::string(name) + ": " + ofrr_ctx_last_error(cx));
   store(R, dR.data(), n, n, n);
   std::vector<int32_t> hs(n);
   hip_check(hipMemcpy(hs.data(), skip.data(), n * 4, hipMemcpyDeviceToHost), name);
   return std::vector<int>(hs.begin(), hs.end());
}

// column i of Q against its final columns 0..i-1 (the *_step functions); on
// return R_dev holds R(0..i, i) in FP64 and skip[i] the drop flag
template <typename T_O>
void gs_qr_step(int variant, const char *name, Matrix<T_O> &Q, std::vector<int> &skip, size_t i,
                double select_tol, double orth_tol, double reorth_tol, Matrix<double> &R_dev) {
   const size_t m = Q.rows(), n = Q.cols();
   if (Q.location() != Location::kDEVICE) throw std::runtime_error("Output matrix Q must be on device");
   if (i >= n || skip.size() < n) throw std::runtime_error(std::string(name) + ": bad column");
   std::vector<int32_t> hs(skip.begin(), skip.begin() + n);
   Vector<int32_t> dskip(n);
   hip_check(hipMemcpy(dskip.data(), hs.data(), n * 4, hipMemcpyHostToDevice), name);
   int32_t dropped = 0;
   int st = ofrr_gs_qr_step(ctx(), variant, dtype<T_O>::value, (int64_t)m, (int)n, Q.data(), 1, (int64_t)Q.ld(),
                            (int)i, select_tol, orth_tol, reorth_tol, R_dev.data(), (int)n, dskip.data(),
                            &dropped);
   if (st != 0) throw std::runtime_error(std::string(name) + ": " + ofrr_ctx_last_error(ctx()));
   skip[i] = dropped;
}
}  // namespace detail

template <typename T_I, typename T_O, typename T_COMPUTE>
std::vector<int> mgs(const Matrix<T_I> &A, Matrix<T_O> &Q, Matrix<T_O> &R,
                     T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                     T_O orth_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                     T_O reorth_tol = detail::to_elem<T_O>(0.70710678118654752)) {
   return detail::gs_qr(OFRR_QR_MGS, "mgs", A, Q, R, detail::from_elem<T_O>(select_tol), detail::from_elem<T_O>(orth_tol), detail::from_elem<T_O>(reorth_tol));
}

template <typename T_I, typename T_O, typename T_COMPUTE>
std::vector<int> mgs_v2(const Matrix<T_I> &A, Matrix<T_O> &Q, Matrix<T_O> &R,
                        T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                        T_O orth_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                        T_O reorth_tol = detail::to_elem<T_O>(0.70710678118654752)) {
   (void)reorth_tol;
   return detail::gs_qr(OFRR_QR_MGS_V2, "mgs_v2", A, Q, R, detail::from_elem<T_O>(select_tol), detail::from_elem<T_O>(orth_tol), -1.0);
}

template <typename T_I, typename T_O, typename T_COMPUTE>
std::vector<int> cgs(const Matrix<T_I> &A, Matrix<T_O> &Q, Matrix<T_O> &R,
                     T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                     T_O orth_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                     T_O reorth_tol = detail::to_elem<T_O>(0.70710678118654752)) {
   (void)reorth_tol;
   return detail::gs_qr(OFRR_QR_CGS, "cgs", A, Q, R, detail::from_elem<T_O>(select_tol), detail::from_elem<T_O>(orth_tol), -1.0);
}

template <typename T_I, typename T_O, typename T_COMPUTE>
std::vector<int> cgs2(const Matrix<T_I> &A, Matrix<T_O> &Q, Matrix<T_O> &R,
                      T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                      T_O orth_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
                      T_O reorth_tol = detail::to_elem<T_O>(0.70710678118654752)) {
   (void)reorth_tol;
   return detail::gs_qr(OFRR_QR_CGS2, "cgs2", A, Q, R, detail::from_elem<T_O>(select_tol), detail::from_elem<T_O>(orth_tol), -1.0);
}

// mgs_step (qr.hpp:52-75): R on the host; R(j, i) for the kept j < i and R(i, i)
// of an accepted column are written, other entries are left as they are
template <typename T_O, typename T_COMPUTE>
int mgs_step(Matrix<T_O> &Q, Matrix<T_O> &R, std::vector<int> &skip, size_t i,
             T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
             T_O orth_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value)),
             T_O reorth_tol = detail::to_elem<T_O>(0.70710678118654752)) {
   if (R.location() != Location::kHOST) throw std::runtime_error("Output matrix R must be on host");
   const size_t n = Q.cols();
   Matrix<double> dR(n, n);
   detail::gs_qr_step(OFRR_QR_MGS, "mgs_step", Q, skip, i, detail::from_elem<T_O>(select_tol), detail::from_elem<T_O>(orth_tol),
                      (double)reorth_tol, dR);
   std::vector<double> col(i + 1);
   detail::hip_check(hipMemcpy(col.data(), dR.data() + i * n, (i + 1) * 8, hipMemcpyDeviceToHost), "mgs_step");
   for (size_t j = 0; j < i; ++j)
      if (!skip[j]) R(j, i) = detail::to_elem<T_O>(col[j]);
   if (!skip[i]) R(i, i) = detail::to_elem<T_O>(col[i]);
   return skip[i];
}

// cgs_step / cgs2_step (qr.hpp:100-107, 133-141): R_device (device, n x n)
// receives R(0..i-1, i), r_diag[i] (host) the norm of an accepted column
template <typename T_O, typename T_COMPUTE>
int cgs_step(Matrix<T_O> &Q, Matrix<T_O> &R_device, Vector<T_O> &r_diag, std::vector<int> &skip, size_t i,
             T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value))) {
   const size_t n = Q.cols();
   Matrix<double> dR(n, n);
   detail::gs_qr_step(OFRR_QR_CGS, "cgs_step", Q, skip, i, detail::from_elem<T_O>(select_tol),
                      detail::eps_of(detail::dtype<T_O>::value), -1.0, dR);
   if (i > 0) {
      int st = ofrr_convert(detail::ctx(), (int64_t)i, 1, dR.data() + i * n, OFRR_F64, 1, (int64_t)n,
                            &R_device(0, i), detail::dtype<T_O>::value, 1, (int64_t)R_device.ld());
      if (detail::done(st) != MSVDStatus::kSuccess) throw std::runtime_error("cgs_step: R store");
   }
   double d = 0.0;
   detail::hip_check(hipMemcpy(&d, dR.data() + i * n + i, 8, hipMemcpyDeviceToHost), "cgs_step");
   if (!skip[i]) r_diag[i] = detail::to_elem<T_O>(d);
   return skip[i];
}

template <typename T_O, typename T_COMPUTE>
int cgs2_step(Matrix<T_O> &Q, Matrix<T_O> &R_device, Vector<T_O> &r_correction, Vector<T_O> &r_diag,
              std::vector<int> &skip, size_t i,
              T_O select_tol = detail::to_elem<T_O>(detail::eps_of(detail::dtype<T_O>::value))) {
   (void)r_correction;
   const size_t n = Q.cols();
   Matrix<double> dR(n, n);
   detail::gs_qr_step(OFRR_QR_CGS2, "cgs2_step", Q, skip, i, detail::from_elem<T_O>(select_tol),
                      detail::eps_of(detail::dtype<T_O>::value), -1.0, dR);
   if (i > 0) {
      int st = ofrr_convert(detail::ctx(), (int64_t)i, 1, dR.data() + i * n, OFRR_F64, 1, (int64_t)n,
                            &R_device(0, i), detail::dtype<T_O>::value, 1, (int64_t)R_device.ld());
      if (detail::done(st) != MSVDStatus::kSuccess) throw std::runtime_error("cgs2_step: R store");
   }
   double d = 0.0;
   detail::hip_check(hipMemcpy(&d, dR.data() + i * n + i, 8, hipMemcpyDeviceToHost), "cgs2_step");
   if (!skip[i]) r_diag[i] = detail::to_elem<T_O>(d);
   return skip[i];
}

// ---------------------------------------------------------------------------
// Rayleigh-Ritz (rr_eig.hpp:21-35)

// ofrr_eig: A = V^T A V-like, B = V^T V-like (m x m); V (m x k) <- the Ritz
// coefficients of the k largest pairs, D (k x k) <- diag(theta), descending;
// select_tol = the truncation tolerance of the Gram eigendecomposition
template <typename T_I, typename T_O, typename T_COMPUTE>
MSVDStatus ofrr_eig(const Matrix<T_I> &A, const Matrix<T_I> &B, Matrix<T_O> &V, Matrix<T_O> &D,
                    T_COMPUTE select_tol = detail::to_elem<T_COMPUTE>(1e-10)) {
   const size_t m = A.rows(), k = std::min<size_t>(V.cols(), m);
   if (A.cols() != m || B.rows() != m || B.cols() != m || V.rows() != m) return MSVDStatus::kErrorBLASInvalidSize;
   Matrix<double> dA(m, m), dB(m, m), Y(m, k);
   Vector<double> th(k);
   ofrr_ctx *cx = detail::ctx();
   detail::DevView<T_I> a(A), b(B);
   int st = ofrr_convert(cx, (int64_t)m, (int64_t)m, a.p, detail::dtype<T_I>::value, 1, (int64_t)a.ld, dA.data(),
                         OFRR_F64, 1, (int64_t)m);
   if (!st)
      st = ofrr_convert(cx, (int64_t)m, (int64_t)m, b.p, detail::dtype<T_I>::value, 1, (int64_t)b.ld, dB.data(),
                        OFRR_F64, 1, (int64_t)m);
   if (!st)
      st = ofrr_rr_truncated(cx, (int)m, dB.data(), (int)m, dA.data(), (int)m, detail::from_elem<T_COMPUTE>(select_tol),
                             (int)k, th.data(), Y.data(), (int)m, nullptr);
   MSVDStatus s = detail::done(st);
   if (s != MSVDStatus::kSuccess) return s;
   detail::store(V, Y.data(), m, k, m);
   std::vector<double> ht(k);
   detail::hip_check(hipMemcpy(ht.data(), th.data(), k * 8, hipMemcpyDeviceToHost), "ofrr_eig");
   detail::set_diag(D, ht);
   return MSVDStatus::kSuccess;
}

// rr_eig: Rayleigh-Ritz of a dense symmetric A on the basis V (n x m, m <= 384):
// V <- V Y, D <- diag(theta) (descending); generalised form eig(V^T A V, V^T V)
// or the truncated OFRR form (select_tol).  Products in FP64 accumulation.
template <typename T_I, typename T_O, typename T_COMPUTE>
MSVDStatus rr_eig(const Matrix<T_I> &A, Matrix<T_O> &V, Matrix<T_O> &D, bool use_generalized = false,
                  T_COMPUTE select_tol = detail::to_elem<T_COMPUTE>(1e-10)) {
   const size_t n = A.rows(), m = V.cols();
   if (A.cols() != n || V.rows() != n || m == 0 || m > 384) return MSVDStatus::kErrorBLASInvalidSize;
   ofrr_ctx *cx = detail::ctx();
   detail::DevView<T_I> a(A);
   detail::DevView<T_O> v(V);
   Matrix<double> A64(n, n), V64(n, m), AV(n, m), G(m, m), H(m, m), Y(m, m), Vn(n, m);
   Vector<double> th(m);
   const double tol = detail::from_elem<T_COMPUTE>(select_tol);
   int st = ofrr_convert(cx, (int64_t)n, (int64_t)n, a.p, detail::dtype<T_I>::value, 1, (int64_t)a.ld, A64.data(),
                         OFRR_F64, 1, (int64_t)n);
   if (!st)
      st = ofrr_convert(cx, (int64_t)n, (int64_t)m, v.p, detail::dtype<T_O>::value, 1, (int64_t)v.ld, V64.data(),
                        OFRR_F64, 1, (int64_t)n);
   if (!st)
      st = ofrr_gemm(cx, 0, 0, (int64_t)n, (int64_t)m, (int64_t)n, 1.0, A64.data(), OFRR_F64, (int64_t)n,
                     V64.data(), OFRR_F64, (int64_t)n, 0.0, AV.data(), OFRR_F64, (int64_t)n, OFRR_F64);
   if (!st)
      st = ofrr_gram(cx, (int64_t)n, (int)m, (int)m, V64.data(), OFRR_F64, 1, (int64_t)n, V64.data(), OFRR_F64, 1,
                     (int64_t)n, 1.0, 0.0, G.data(), (int)m);
   if (!st)
      st = ofrr_gram(cx, (int64_t)n, (int)m, (int)m, V64.data(), OFRR_F64, 1, (int64_t)n, AV.data(), OFRR_F64, 1,
                     (int64_t)n, 1.0, 0.0, H.data(), (int)m);
   if (!st)
      st = use_generalized
               ? ofrr_rr_generalized(cx, (int)m, G.data(), (int)m, H.data(), (int)m, tol, th.data(), Y.data(), (int)m)
               : ofrr_rr_truncated(cx, (int)m, G.data(), (int)m, H.data(), (int)m, tol, (int)m, th.data(), Y.data(),
                                   (int)m, nullptr);
   if (!st)
      st = ofrr_gemm(cx, 0, 0, (int64_t)n, (int64_t)m, (int64_t)m, 1.0, V64.data(), OFRR_F64, (int64_t)n, Y.data(),
                     OFRR_F64, (int64_t)m, 0.0, Vn.data(), OFRR_F64, (int64_t)n, OFRR_F64);
   if (!st)
      st = ofrr_convert(cx, (int64_t)n, (int64_t)m, Vn.data(), OFRR_F64, 1, (int64_t)n, V.data(),
                        detail::dtype<T_O>::value, 1, (int64_t)V.ld());
   MSVDStatus s = detail::done(st);
   if (s != MSVDStatus::kSuccess) return s;
   std::vector<double> ht(m);
   detail::hip_check(hipMemcpy(ht.data(), th.data(), m * 8, hipMemcpyDeviceToHost), "rr_eig");
   detail::set_diag(D, ht);
   return MSVDStatus::kSuccess;
}

// ---------------------------------------------------------------------------
// drivers over a dense symmetric A (si_eig.hpp:24-27, arnoldi_eig.hpp:25-28)

namespace detail {
// growth-block format of the drivers: the output element type's own for
// 16-bit outputs (a mixed-precision basis), FP64 otherwise
template <typename T_O> constexpr int growth_format() {
   return dtype<T_O>::value == OFRR_F16 ? OFRR_F16 : dtype<T_O>::value == OFRR_BF16 ? OFRR_BF16 : OFRR_F64;
}
}  // namespace detail

// block subspace iteration: filtered OFRR over the dense A.  V (n x k, k <= 128)
// start block in / Ritz block out; D (k x k) <- diag(theta), descending.  The
// growth blocks are FP16 / BF16 for a 16-bit T_O (the mixed-precision path:
// FP32 accumulation, FP64 Ritz block, Gram and Rayleigh-Ritz), FP64 otherwise.
// The leading half of the block is converged to tol (the rest are guard
// vectors of the filter); filter degree 12 per iteration, max_iter iterations.
template <typename T_I, typename T_O, typename T_COMPUTE>
MSVDStatus si_eig(const Matrix<T_I> &A, Matrix<T_O> &V, Matrix<T_O> &D, int max_iter = 100,
                  T_COMPUTE tol = detail::to_elem<T_COMPUTE>(1e-10), bool use_generalized = false,
                  T_COMPUTE select_tol = detail::to_elem<T_COMPUTE>(1e-10)) {
   (void)use_generalized;
   const size_t n = A.rows(), k = V.cols();
   if (A.cols() != n || V.rows() != n || k < 2 || k > 128) return MSVDStatus::kErrorBLASInvalidSize;
   ofrr_ctx *cx = detail::ctx();
   detail::DevView<T_I> a(A);
   detail::DevView<T_O> v(V);
   Matrix<double> A64(n, n), X(k, n);
   int st = ofrr_convert(cx, (int64_t)n, (int64_t)n, a.p, detail::dtype<T_I>::value, 1, (int64_t)a.ld, A64.data(),
                         OFRR_F64, 1, (int64_t)n);
   if (!st)
      st = ofrr_convert(cx, (int64_t)n, (int64_t)k, v.p, detail::dtype<T_O>::value, 1, (int64_t)v.ld, X.data(),
                        OFRR_F64, (int64_t)k, 1);
   if (st) return static_cast<MSVDStatus>(st);
   ofrr_filter_opts o{};
   o.k = (int)k;
   o.nwant = (int)std::max<size_t>(1, k / 2);
   o.nd = 1;
   o.degree = 12;
   o.maxiter = max_iter;
   o.basis_dtype = detail::growth_format<T_O>();
   o.tol = detail::from_elem<T_COMPUTE>(tol);
   o.tau = detail::from_elem<T_COMPUTE>(select_tol);
   std::vector<double> th(k), res(k);
   st = ofrr_filtered_eig_dense(cx, (int64_t)n, A64.data(), (int64_t)n, &o, X.data(), th.data(), res.data(),
                                nullptr);
   if (st != 0 && st != 4001) return static_cast<MSVDStatus>(st);
   const int st2 = ofrr_convert(cx, (int64_t)n, (int64_t)k, X.data(), OFRR_F64, (int64_t)k, 1, V.data(),
                                detail::dtype<T_O>::value, 1, (int64_t)V.ld());
   if (detail::done(st2) != MSVDStatus::kSuccess) return static_cast<MSVDStatus>(st2);
   detail::set_diag(D, th);
   return static_cast<MSVDStatus>(st);
}

// Arnoldi-Hessenberg (arnoldi_hess_eig.m): basis dimension V.cols(), the
// wanted pairs V.cols() / 3, restarts up to max_iter; A is taken as CSR of
// its nonzeros.  V (n x kdim) out: the Ritz vectors in its first columns;
// D (k x k) <- diag(theta) descending
template <typename T_I, typename T_O, typename T_COMPUTE>
MSVDStatus arnoldi_eig(const Matrix<T_I> &A, Matrix<T_O> &V, Matrix<T_O> &D, int max_iter = 100,
                       T_COMPUTE tol = detail::to_elem<T_COMPUTE>(1e-10), bool use_generalized = false,
                  T_COMPUTE select_tol = detail::to_elem<T_COMPUTE>(1e-10)) {
   (void)tol;
   (void)use_generalized;
   const size_t n = A.rows();
   const int kdim = (int)std::min<size_t>(V.cols(), 255);
   const int k = std::max(1, kdim / 3);
   if (A.cols() != n || V.rows() != n || kdim < 2) return MSVDStatus::kErrorBLASInvalidSize;
   std::vector<T_I> h(A.ld() * n);
   detail::hip_check(hipMemcpy(h.data(), A.data(), h.size() * sizeof(T_I), hipMemcpyDefault), "arnoldi_eig");
   std::vector<int64_t> rp(n + 1, 0);
   std::vector<int32_t> ci;
   std::vector<double> va;
   for (size_t i = 0; i < n; ++i) {
      for (size_t j = 0; j < n; ++j) {
         const double x = detail::from_elem<T_I>(h[i + j * A.ld()]);
         if (x != 0.0) { ci.push_back((int32_t)j); va.push_back(x); }
      }
      rp[i + 1] = (int64_t)ci.size();
   }
   Vector<int64_t> drp(n + 1);
   Vector<int32_t> dci(std::max<size_t>(1, ci.size()));
   Vector<double> dva(std::max<size_t>(1, va.size())), v0(n);
   Matrix<double> X(n, k);
   detail::hip_check(hipMemcpy(drp.data(), rp.data(), rp.size() * 8, hipMemcpyHostToDevice), "arnoldi_eig");
   detail::hip_check(hipMemcpy(dci.data(), ci.data(), ci.size() * 4, hipMemcpyHostToDevice), "arnoldi_eig");
   detail::hip_check(hipMemcpy(dva.data(), va.data(), va.size() * 8, hipMemcpyHostToDevice), "arnoldi_eig");
   v0.fill_random(0);
   std::vector<double> th(k);
   int built = 0;
   int st = ofrr_arnoldi_hess_eig(detail::ctx(), (int64_t)n, (int64_t)va.size(), drp.data(), dci.data(), dva.data(),
                                  k, kdim, max_iter, detail::from_elem<T_COMPUTE>(select_tol), 0.0, v0.data(), th.data(), X.data(),
                                  (int64_t)n, &built);
   MSVDStatus s = detail::done(st);
   if (s != MSVDStatus::kSuccess) return s;
   detail::store(V, X.data(), n, (size_t)k, n);
   detail::set_diag(D, th);
   return MSVDStatus::kSuccess;
}

// ---------------------------------------------------------------------------
// truncated SVD (si_svd.hpp, rr_svd.hpp) of a dense FP32 A (M x N)

// si_svd: filtered OFRR on A^T A (BF16 growth products); U (M x k), S (k x k)
// <- diag(sigma) descending, V (N x k, k <= 128) start block in / right
// singular vectors out; the largest k/2 triplets to tol
template <typename T_I, typename T_O, typename T_COMPUTE>
MSVDStatus si_svd(const Matrix<T_I> &A, Matrix<T_O> &U, Matrix<T_O> &S, Matrix<T_O> &V, int max_iter = 100,
                  T_COMPUTE tol = detail::to_elem<T_COMPUTE>(1e-10), bool use_generalized = false,
                  T_COMPUTE select_tol = detail::to_elem<T_COMPUTE>(1e-10)) {
   (void)use_generalized;
   const size_t M = A.rows(), N = A.cols(), k = V.cols();
   if (V.rows() != N || U.rows() != M || U.cols() < k) return MSVDStatus::kErrorBLASInvalidSize;
   ofrr_ctx *cx = detail::ctx();
   detail::DevView<T_I> a(A);
   detail::DevView<T_O> v(V);
   Matrix<float> A32(M, N);
   Matrix<double> X(k, N), Ud(k, M);
   int st = ofrr_convert(cx, (int64_t)M, (int64_t)N, a.p, detail::dtype<T_I>::value, 1, (int64_t)a.ld, A32.data(),
                         OFRR_F32, 1, (int64_t)M);
   if (!st)
      st = ofrr_convert(cx, (int64_t)N, (int64_t)k, v.p, detail::dtype<T_O>::value, 1, (int64_t)v.ld, X.data(),
                        OFRR_F64, (int64_t)k, 1);
   if (st) return static_cast<MSVDStatus>(st);
   ofrr_filter_opts o{(int)k, (int)std::max<size_t>(1, k / 2), 1, 8, max_iter, OFRR_BF16,
                      detail::from_elem<T_COMPUTE>(tol), detail::from_elem<T_COMPUTE>(select_tol), 0.0, 0.0};
   std::vector<double> sg(k), res(k);
   st = ofrr_svd_filtered(cx, (int64_t)M, (int64_t)N, A32.data(), (int64_t)M, &o, X.data(), Ud.data(), sg.data(),
                          res.data(), nullptr);
   if (st != 0 && st != 4001) return static_cast<MSVDStatus>(st);
   int st2 = ofrr_convert(cx, (int64_t)N, (int64_t)k, X.data(), OFRR_F64, (int64_t)k, 1, V.data(),
                          detail::dtype<T_O>::value, 1, (int64_t)V.ld());
   if (!st2)
      st2 = ofrr_convert(cx, (int64_t)M, (int64_t)k, Ud.data(), OFRR_F64, (int64_t)k, 1, U.data(),
                         detail::dtype<T_O>::value, 1, (int64_t)U.ld());
   if (detail::done(st2) != MSVDStatus::kSuccess) return static_cast<MSVDStatus>(st2);
   detail::set_diag(S, sg);
   return static_cast<MSVDStatus>(st);
}

// rr_svd: Rayleigh-Ritz for singular triplets on the right basis V (N x k):
// truncated RR of the pencil (V^T A^T A V, V^T V) in FP64; V <- V Y,
// S <- diag(sigma) descending, U <- A V / sigma
template <typename T_I, typename T_O, typename T_COMPUTE>
MSVDStatus rr_svd(const Matrix<T_I> &A, Matrix<T_O> &U, Matrix<T_O> &S, Matrix<T_O> &V,
                  bool use_generalized = false, T_COMPUTE select_tol = detail::to_elem<T_COMPUTE>(1e-10)) {
   (void)use_generalized;
   const size_t M = A.rows(), N = A.cols(), k = V.cols();
   if (V.rows() != N || U.rows() != M || U.cols() < k || k > 128) return MSVDStatus::kErrorBLASInvalidSize;
   ofrr_ctx *cx = detail::ctx();
   detail::DevView<T_I> a(A);
   detail::DevView<T_O> v(V);
   const int nc = (int)((k + 15) / 16 * 16);
   Matrix<double> Arm(N, M), Vr(nc, N), AV(nc, M), G(k, k), H(k, k), Y(k, k), Vn(nc, N), Un(nc, M);
   Vector<double> th(k);
   // A row-major in FP64 (row stride N): P of A V
   int st = ofrr_convert(cx, (int64_t)M, (int64_t)N, a.p, detail::dtype<T_I>::value, 1, (int64_t)a.ld, Arm.data(),
                         OFRR_F64, (int64_t)N, 1);
   if (!st)
      st = ofrr_convert(cx, (int64_t)N, (int64_t)k, v.p, detail::dtype<T_O>::value, 1, (int64_t)v.ld, Vr.data(),
                        OFRR_F64, nc, 1);
   if (!st)
      st = ofrr_dense_gemm(cx, (int64_t)M, (int64_t)N, nc, Arm.data(), OFRR_F64, (int64_t)N, Vr.data(), OFRR_F64,
                           nc, 1.0, 0.0, AV.data(), OFRR_F64, nc);
   if (!st)
      st = ofrr_gram(cx, (int64_t)N, (int)k, (int)k, Vr.data(), OFRR_F64, nc, 1, Vr.data(), OFRR_F64, nc, 1, 1.0,
                     0.0, G.data(), (int)k);
   if (!st)
      st = ofrr_gram(cx, (int64_t)M, (int)k, (int)k, AV.data(), OFRR_F64, nc, 1, AV.data(), OFRR_F64, nc, 1, 1.0,
                     0.0, H.data(), (int)k);
   if (!st)
      st = ofrr_rr_truncated(cx, (int)k, G.data(), (int)k, H.data(), (int)k, detail::from_elem<T_COMPUTE>(select_tol), (int)k,
                             th.data(), Y.data(), (int)k, nullptr);
   if (!st) {
      const void *vb[1] = {Vr.data()}, *ub[1] = {AV.data()};
      const int dt[1] = {OFRR_F64}, kb[1] = {(int)k}, ldc[1] = {(int)k};
      const int64_t rs[1] = {nc}, cs[1] = {1};
      const double *cb[1] = {Y.data()};
      st = ofrr_basis_update(cx, (int64_t)N, (int)k, 1, vb, dt, kb, rs, cs, cb, ldc, Vn.data(), OFRR_F64, nc, 1);
      if (!st)
         st = ofrr_basis_update(cx, (int64_t)M, (int)k, 1, ub, dt, kb, rs, cs, cb, ldc, Un.data(), OFRR_F64, nc,
                                1);
   }
   MSVDStatus s = detail::done(st);
   if (s != MSVDStatus::kSuccess) return s;
   std::vector<double> ht(k), sg(k);
   detail::hip_check(hipMemcpy(ht.data(), th.data(), k * 8, hipMemcpyDeviceToHost), "rr_svd");
   for (size_t j = 0; j < k; ++j) sg[j] = std::sqrt(std::max(0.0, ht[j]));
   // V <- V Y; U <- (A V Y) / sigma
   std::vector<double> hv(nc * N), hu(nc * M);
   detail::hip_check(hipMemcpy(hv.data(), Vn.data(), hv.size() * 8, hipMemcpyDeviceToHost), "rr_svd");
   detail::hip_check(hipMemcpy(hu.data(), Un.data(), hu.size() * 8, hipMemcpyDeviceToHost), "rr_svd");
   std::vector<T_O> tv(V.ld() * k), tu(U.ld() * U.cols(), detail::to_elem<T_O>(0.0));
   for (size_t j = 0; j < k; ++j) {
      for (size_t i = 0; i < N; ++i) tv[i + j * V.ld()] = detail::to_elem<T_O>(hv[i * nc + j]);
      for (size_t i = 0; i < M; ++i) tu[i + j * U.ld()] = detail::to_elem<T_O>(sg[j] > 0 ? hu[i * nc + j] / sg[j] : 0.0);
   }
   detail::hip_check(hipMemcpy(V.data(), tv.data(), tv.size() * sizeof(T_O), hipMemcpyDefault), "rr_svd");
   detail::hip_check(hipMemcpy(U.data(), tu.data(), tu.size() * sizeof(T_O), hipMemcpyDefault), "rr_svd");
   detail::set_diag(S, sg);
   return MSVDStatus::kSuccess;
}

// ---- matrix input: MatrixReader / CSRMatrix of CPP/io/matrix_reader.hpp:7-45,
// reading the binary CSR files through ofrr_csr_file_read; a missing file
// throws std::runtime_error with the reference's message
// (matrix_reader.cpp:15-33).
class CSRMatrix {
  public:
   std::vector<int> row_ptr;
   std::vector<int> col_ind;
   std::vector<double> values;
   int rows;
   int cols;
   CSRMatrix() : rows(0), cols(0) {}
};

class MatrixReader {
  public:
   static CSRMatrix read_csr_matrix(const std::string &prefix) {
      int64_t rows = 0, cols = 0, nnz = 0;
      if (ofrr_csr_file_info(prefix.c_str(), &rows, &cols, &nnz) != 0)
         throw std::runtime_error(ofrr_io_last_error());
      std::vector<int64_t> rp((size_t)rows + 1);
      CSRMatrix m;
      m.col_ind.resize((size_t)nnz);
      m.values.resize((size_t)nnz);
      if (ofrr_csr_file_read(prefix.c_str(), rp.data(), m.col_ind.data(), m.values.data()) != 0)
         throw std::runtime_error(ofrr_io_last_error());
      m.row_ptr.assign(rp.begin(), rp.end());
      m.rows = (int)rows;
      m.cols = (int)cols;
      return m;
   }
   static std::vector<double> read_singular_values(const std::string &prefix) {
      int64_t count = 0;
      if (ofrr_singular_values_file(prefix.c_str(), nullptr, 0, &count) != 0) return {};
      std::vector<double> v((size_t)count);
      (void)ofrr_singular_values_file(prefix.c_str(), v.data(), count, &count);
      return v;
   }
};

}  // namespace msvd

#endif  // OFRR_MSVD_HPP
